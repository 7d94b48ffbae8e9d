// hip_matrix.h -- CSR matrix on the device with SpMV through the gfx950 library.
//
// Interface of csr::matrix / csr::gpu_matrix (source/common/csr/matrix.h:86-140,
// source/common/csr/gpu_matrix.h:119-131, 322-375): init(nnz, size_col,
// size_row), set(data, col_ind, row_ptr) from host arrays, init_set, get_dim,
// axpy(alpha, x, beta, y, transpose) computing y := alpha*A*x + beta*y.
// Zero-based int32 ordinals like the reference's default.  Unlike the reference
// there is no per-call allocation; rows can be evaluated as compensated dot
// products (use_compensated_rows), and a fused residual r := b - A x is offered.
// Transposed products (axpy and the triangular solves with transpose = true) run
// the ordinary kernels on A^T, formed once on the device (hpb_csr_transpose_*) on
// the first transposed call and kept until the matrix changes.
// triangular_solve_lower / _upper (gpu_matrix.h:142-320) solve with the lower
// triangle and a unit diagonal / the upper triangle and the stored diagonal; the
// level analysis of each triangle is made on first use and kept until the
// pattern changes (the reference redoes it on every call).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../hip_vector_operations.h"

namespace csr {

template <typename T>
class hip_matrix {
public:
    struct dimensions_t {
        int columns = 0;
        int rows = 0;
    };
    using scalar_type = T;
    using vector_type = T*;

    explicit hip_matrix(hpb_handle_t h) : h_(h) {}
    hip_matrix(const hip_matrix&) = delete;
    hip_matrix& operator=(const hip_matrix&) = delete;
    ~hip_matrix()
    {
        drop_analyses_();
        release_();
    }

    void set_dim(int matrix_columns, int matrix_rows)
    {
        dim_.columns = matrix_columns;
        dim_.rows = matrix_rows;
    }
    dimensions_t get_dim() const { return dim_; }

    // sizes as in the reference: nnz, size_col = length of col_ind (= nnz), size_row = rows + 1
    void init(int64_t nnz, int64_t size_col, int64_t size_row)
    {
        if (initialized_) return;
        nnz_ = nnz;
        size_col_ = size_col;
        size_row_ = size_row;
        hpb::check(hpb_malloc((void**)&data_, sizeof(T) * (size_t)nnz_), "csr::hip_matrix::init");
        hpb::check(hpb_malloc((void**)&col_ind_, sizeof(int) * (size_t)size_col_), "csr::hip_matrix::init");
        hpb::check(hpb_malloc((void**)&row_ptr_, sizeof(int) * (size_t)size_row_), "csr::hip_matrix::init");
        if (dim_.rows == 0) set_dim((int)(size_row_ - 1), (int)(size_row_ - 1));
        initialized_ = true;
    }
    void set(const T* data, const int* col_ind, const int* row_ptr)
    {
        if (!initialized_) throw std::runtime_error("csr::matrix: trying to set an uninitialized matrix.");
        drop_analyses_();
        hpb::check(hpb_memcpy_h2d(h_, data_, data, sizeof(T) * (size_t)nnz_), "csr::hip_matrix::set");
        hpb::check(hpb_memcpy_h2d(h_, col_ind_, col_ind, sizeof(int) * (size_t)size_col_), "csr::hip_matrix::set");
        hpb::check(hpb_memcpy_h2d(h_, row_ptr_, row_ptr, sizeof(int) * (size_t)size_row_), "csr::hip_matrix::set");
    }
    void init_set(int64_t nnz, int64_t size_col, int64_t size_row, const T* data, const int* col_ind,
                  const int* row_ptr, int matrix_columns, int matrix_rows)
    {
        set_dim(matrix_columns, matrix_rows);
        init(nnz, size_col, size_row);
        set(data, col_ind, row_ptr);
    }
    // adopt device arrays (e.g. from the library's stencil builders); own = false
    // leaves freeing them to the caller (views sharing the index arrays)
    void adopt_device(int rows, int cols, int64_t nnz, int* row_ptr, int* col_ind, T* data, bool own = true)
    {
        release_();
        drop_analyses_();
        set_dim(cols, rows);
        nnz_ = nnz;
        size_col_ = nnz;
        size_row_ = (int64_t)rows + 1;
        row_ptr_ = row_ptr;
        col_ind_ = col_ind;
        data_ = data;
        own_ = own;
        initialized_ = true;
    }

    void use_compensated_rows(bool v) { comp_ = v; }
    bool compensated_rows() const { return comp_; }

    // y := alpha*A*x + beta*y
    // y := alpha*A^T*x + beta*y with transpose (x of length rows, y of length columns)
    void axpy(const T& alpha, const vector_type& x, const T& beta, vector_type& y, bool transpose = false) const
    {
        auto fn = comp_ ? hpb::abi<T>::csr_spmvc : hpb::abi<T>::csr_spmv;
        if (transpose) {
            transpose_();
            hpb::check(fn(h_, dim_.columns, dim_.rows, nnz_, t_row_ptr_, t_col_ind_, t_data_, alpha, x, beta, y),
                       "csr::hip_matrix::axpy (transposed)");
            return;
        }
        hpb::check(fn(h_, dim_.rows, dim_.columns, nnz_, row_ptr_, col_ind_, data_, alpha, x, beta, y),
                   "csr::hip_matrix::axpy");
    }

    // L x = alpha b: strictly lower triangle, unit diagonal (transpose: L^T x = alpha b)
    void triangular_solve_lower(const vector_type& b, vector_type& x, const T& alpha = T(1.0),
                                bool transpose = false) const
    {
        trsv_(0, b, x, alpha, transpose);
    }
    // U x = alpha b: upper triangle with the stored diagonal (transpose: U^T x = alpha b)
    void triangular_solve_upper(const vector_type& b, vector_type& x, const T& alpha = T(1.0),
                                bool transpose = false) const
    {
        trsv_(1, b, x, alpha, transpose);
    }
    // in-place ILU(0) of the values (pattern kept): L strictly below, U on and above
    void ilu0_in_place()
    {
        drop_transpose_();
        hpb::check(hpb::abi<T>::csr_ilu0(h_, analysis_(0), data_), "csr::hip_matrix::ilu0");
    }

    int rows() const { return dim_.rows; }
    int64_t nnz() const { return nnz_; }
    const int* row_ptr() const { return row_ptr_; }
    const int* col_ind() const { return col_ind_; }
    const T* data() const { return data_; }
    hpb_handle_t handle() const { return h_; }

private:
    hpb_trsv_t analysis_(int uplo) const
    {
        hpb_trsv_t& a = uplo == 0 ? lower_ : upper_;
        if (!a)
            hpb::check(hpb_csr_trsv_analysis(h_, uplo, dim_.rows, nnz_, row_ptr_, col_ind_, &a),
                       "csr::hip_matrix: triangular analysis");
        return a;
    }
    // on A^T: L^T is the strictly upper triangle of A^T with a unit diagonal (kind 3),
    // U^T the lower triangle of A^T with the stored diagonal (kind 2)
    hpb_trsv_t analysis_t_(int uplo) const
    {
        transpose_();
        hpb_trsv_t& a = uplo == 0 ? t_lower_ : t_upper_;
        if (!a)
            hpb::check(hpb_csr_trsv_analysis(h_, uplo == 0 ? 3 : 2, dim_.columns, nnz_, t_row_ptr_, t_col_ind_, &a),
                       "csr::hip_matrix: transposed triangular analysis");
        return a;
    }
    void trsv_(int uplo, const vector_type& b, vector_type& x, const T& alpha, bool transpose) const
    {
        if (transpose) {
            hpb::check(hpb::abi<T>::csr_trsv(h_, analysis_t_(uplo), t_data_, alpha, b, x, comp_ ? 1 : 0),
                       "csr::hip_matrix: transposed triangular solve");
            return;
        }
        hpb::check(hpb::abi<T>::csr_trsv(h_, analysis_(uplo), data_, alpha, b, x, comp_ ? 1 : 0),
                   "csr::hip_matrix: triangular solve");
    }
    void transpose_() const
    {
        if (t_row_ptr_) return;
        hpb::check(hpb_malloc((void**)&t_row_ptr_, sizeof(int) * ((size_t)dim_.columns + 1)), "csr::hip_matrix: A^T");
        hpb::check(hpb_malloc((void**)&t_col_ind_, sizeof(int) * (size_t)nnz_), "csr::hip_matrix: A^T");
        hpb::check(hpb_malloc((void**)&t_data_, sizeof(T) * (size_t)nnz_), "csr::hip_matrix: A^T");
        hpb::check(hpb::abi<T>::csr_transpose(h_, dim_.rows, dim_.columns, nnz_, row_ptr_, col_ind_, data_, t_row_ptr_,
                                              t_col_ind_, t_data_),
                   "csr::hip_matrix: A^T");
    }
    void drop_analyses_()
    {
        hpb_trsv_destroy(lower_);
        hpb_trsv_destroy(upper_);
        lower_ = upper_ = nullptr;
        drop_transpose_();
    }
    // A^T holds a copy of the values: dropped whenever they change
    void drop_transpose_() const
    {
        hpb_trsv_destroy(t_lower_);
        hpb_trsv_destroy(t_upper_);
        t_lower_ = t_upper_ = nullptr;
        if (t_row_ptr_) hpb_free(t_row_ptr_);
        if (t_col_ind_) hpb_free(t_col_ind_);
        if (t_data_) hpb_free(t_data_);
        t_row_ptr_ = nullptr;
        t_col_ind_ = nullptr;
        t_data_ = nullptr;
    }
    void release_()
    {
        if (own_) {
            if (data_) hpb_free(data_);
            if (col_ind_) hpb_free(col_ind_);
            if (row_ptr_) hpb_free(row_ptr_);
        }
        own_ = true;
        data_ = nullptr;
        col_ind_ = nullptr;
        row_ptr_ = nullptr;
        initialized_ = false;
    }
    hpb_handle_t h_;
    dimensions_t dim_;
    int64_t nnz_ = 0, size_col_ = 0, size_row_ = 0;
    T* data_ = nullptr;
    int* col_ind_ = nullptr;
    int* row_ptr_ = nullptr;
    bool initialized_ = false;
    bool comp_ = false;
    bool own_ = true;
    mutable hpb_trsv_t lower_ = nullptr, upper_ = nullptr, t_lower_ = nullptr, t_upper_ = nullptr;
    mutable int* t_row_ptr_ = nullptr;   // A^T, made on the first transposed call
    mutable int* t_col_ind_ = nullptr;
    mutable T* t_data_ = nullptr;
};

}  // namespace csr
