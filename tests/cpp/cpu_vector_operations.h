// Test double: a host implementation of the VectorOperations concept, so that the
// solver templates run on the CPU (the reference does the same with
// cpu_vector_operations, source/common/cpu_vector_operations.h:32-220, and the
// inline test double of numerical_algos/lin_solvers/tests/cpu_1st_call_alloc_test.cpp:41-125).
// High precision uses the sequential Ogita-Rump-Oishi Dot2 (dot_product.h:108-121).
#pragma once
#include <cmath>
#include <cstddef>

template <typename T>
struct cpu_vector_operations {
    typedef T scalar_type;
    typedef T* vector_type;
    typedef T norm_type;
    typedef T Tsc;
    bool location = false;
    size_t sz_;
    mutable bool high_ = false;
    bool provide_multi_ = true;

    explicit cpu_vector_operations(size_t sz) : sz_(sz) {}
    void use_high_precision() const { high_ = true; }
    void use_standard_precision() const { high_ = false; }
    bool get_current_precision() const { return high_; }
    void init_vector(vector_type& x) const { x = nullptr; }
    void free_vector(vector_type& x) const { delete[] x; x = nullptr; }
    void start_use_vector(vector_type& x) const { if (!x) x = new T[sz_](); }
    void stop_use_vector(vector_type&) const {}
    size_t get_vector_size() const { return sz_; }
    bool check_is_valid_number(const vector_type& x) const
    {
        for (size_t i = 0; i < sz_; ++i)
            if (!std::isfinite(x[i])) return false;
        return true;
    }
    T scalar_prod(const vector_type& x, const vector_type& y) const
    {
        if (!high_) {
            T s = 0;
            for (size_t j = 0; j < sz_; ++j) s += x[j] * y[j];
            return s;
        }
        T s = 0, c = 0;
        for (size_t j = 0; j < sz_; ++j) {
            T p = x[j] * y[j];
            T pi = std::fma(x[j], y[j], -p);
            T sn = s + p, bs = sn - s, as = sn - bs;
            T t = (p - bs) + (s - as);
            s = sn;
            c = c + pi + t;
        }
        return s + c;
    }
    T norm(const vector_type& x) const { return std::sqrt(scalar_prod(x, x)); }
    void assign_scalar(T a, vector_type& x) const { for (size_t i = 0; i < sz_; ++i) x[i] = a; }
    void scale(T a, vector_type& x) const { for (size_t i = 0; i < sz_; ++i) x[i] *= a; }
    void assign(const vector_type& x, vector_type& y) const { for (size_t i = 0; i < sz_; ++i) y[i] = x[i]; }
    void assign_mul(T a, const vector_type& x, vector_type& y) const { for (size_t i = 0; i < sz_; ++i) y[i] = a * x[i]; }
    void add_mul(T a, const vector_type& x, vector_type& y) const { for (size_t i = 0; i < sz_; ++i) y[i] += a * x[i]; }
    void add_mul(T a, const vector_type& x, T b, vector_type& y) const
    {
        for (size_t i = 0; i < sz_; ++i) y[i] = a * x[i] + b * y[i];
    }
    // z = a x + b y
    void assign_mul(T a, const vector_type& x, T b, const vector_type& y, vector_type& z) const
    {
        for (size_t i = 0; i < sz_; ++i) z[i] = a * x[i] + b * y[i];
    }
    // z = a x + b y + c z
    void add_mul(T a, const vector_type& x, T b, const vector_type& y, T c, vector_type& z) const
    {
        for (size_t i = 0; i < sz_; ++i) z[i] = a * x[i] + b * y[i] + c * z[i];
    }
};

// same operations plus the fused Gram-Schmidt pair of the concept extension
template <typename T>
struct cpu_vector_operations_multi : cpu_vector_operations<T> {
    using base = cpu_vector_operations<T>;
    using typename base::vector_type;
    explicit cpu_vector_operations_multi(size_t sz) : base(sz) {}
    void multi_scalar_prod(int k, const vector_type* V, const vector_type& w, T* out) const
    {
        for (int j = 0; j < k; ++j) out[j] = this->scalar_prod(V[j], w);
    }
    void multi_add_mul(int k, const T* a, const vector_type* V, vector_type& y) const
    {
        for (size_t i = 0; i < this->sz_; ++i) {
            T s = y[i];
            for (int j = 0; j < k; ++j) s += a[j] * V[j][i];
            y[i] = s;
        }
    }
};
