// matrix_market_reader.h -- MatrixMarket file -> csr::hip_matrix.
//
// Interface of csr::matrix_market_reader (source/common/csr/matrix_market_reader.h:48-246):
// read_file(name), get_matrix_dim(rows, cols), set_csr_matrix(Matrix*).  The
// parsing is the library's hpb_mm_read_csr (same file semantics).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "hpb.h"

namespace csr {

template <typename T>
class matrix_market_reader {
public:
    matrix_market_reader() = default;
    ~matrix_market_reader() { release_(); }
    matrix_market_reader(const matrix_market_reader&) = delete;
    matrix_market_reader& operator=(const matrix_market_reader&) = delete;

    void read_file(const std::string& file_name)
    {
        release_();
        if (hpb_mm_read_csr(file_name.c_str(), &rows_, &cols_, &nnz_, &rp_, &ci_, &va_) != 0)
            throw std::runtime_error(std::string("csr::matrix_market_reader::read_file: ") + hpb_last_error());
    }
    void get_matrix_dim(int& size_rows, int& size_cols) const
    {
        if (!rp_) throw std::runtime_error("csr::matrix_market_reader::get_matrix_dim: call read_file first");
        size_rows = rows_;
        size_cols = cols_;
    }
    int64_t nnz() const { return nnz_; }
    const int* row_ptr() const { return rp_; }
    const int* col_ind() const { return ci_; }
    const double* values() const { return va_; }

    // same as the reference: init(nnz, size_col, size_row), set_dim, set(data, col, row)
    template <class Matrix>
    void set_csr_matrix(Matrix* mat) const
    {
        if (!rp_) throw std::runtime_error("csr::matrix_market_reader::set_csr_matrix: call read_file first");
        mat->set_dim(cols_, rows_);
        mat->init(nnz_, nnz_, (int64_t)rows_ + 1);
        if constexpr (sizeof(T) == sizeof(double)) {
            mat->set(reinterpret_cast<const T*>(va_), ci_, rp_);
        } else {
            std::vector<T> v(va_, va_ + nnz_);
            mat->set(v.data(), ci_, rp_);
        }
    }

private:
    void release_()
    {
        hpb_mm_free(rp_);
        hpb_mm_free(ci_);
        hpb_mm_free(va_);
        rp_ = ci_ = nullptr;
        va_ = nullptr;
    }
    int rows_ = 0, cols_ = 0;
    int64_t nnz_ = 0;
    int* rp_ = nullptr;
    int* ci_ = nullptr;
    double* va_ = nullptr;
};

}  // namespace csr
