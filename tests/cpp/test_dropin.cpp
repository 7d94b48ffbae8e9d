// Drop-in C++ layer on the GPU: one call of every member of the VectorOperations
// concept (hpb::hip_vector_operations), of the reduction objects
// (hip_reduction_ogita, hip_reduction), the exact dot (dot_product_exact), the
// vector-pair generator (generate_vector_pair) and the CSR matrix (csr::hip_matrix),
// for FP64 and FP32.  Interfaces: source/common/gpu_vector_operations.h:86-538,
// source/common/ogita/gpu_reduction_ogita.h:36-96, source/common/gpu_reduction.h:36-90,
// source/high_prec/dot_product_gmp.hpp:38-164, source/generate_vector_pair.hpp:40-216,
// source/common/csr/gpu_matrix.h:119-375.
//
// usage: test_dropin <in_dir> <out_dir>
//   in_dir holds x.f64 y.f64 z.f64 w.f64 (n doubles each, n from the file size) and
//   the CSR of a small square matrix: rp.i32 ci.i32 val.f64.
//   out_dir receives one raw little-endian file per result, <tag>_<name>.<f64|f32|i64>;
//   the pytest side (tests/test_gpu_dropin.py) checks them against the oracle.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "csr/hip_matrix.h"
#include "generate_vector_pair.h"
#include "high_prec/dot_product_exact.h"
#include "hip_reduction.h"
#include "hip_reduction_ogita.h"
#include "hip_vector_operations.h"

static std::string g_out;

template <typename V>
static std::vector<V> load(const std::string& path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    const size_t bytes = (size_t)f.tellg();
    f.seekg(0);
    std::vector<V> v(bytes / sizeof(V));
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)bytes);
    return v;
}

template <typename V> const char* ext();
template <> const char* ext<double>() { return "f64"; }
template <> const char* ext<float>() { return "f32"; }
template <> const char* ext<int64_t>() { return "i64"; }

template <typename V>
static void save(const std::string& name, const V* p, size_t n)
{
    std::ofstream f(g_out + "/" + name + "." + ext<V>(), std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(V)));
}
template <typename V>
static void save1(const std::string& name, V v)
{
    save<V>(name, &v, 1);
}

template <typename T>
struct runner {
    using vo_t = hpb::hip_vector_operations<T>;
    using T_vec = T*;
    hpb_handle_t h;
    size_t n;
    std::string tag;
    std::vector<T> X, Y, Z, W;

    runner(hpb_handle_t h_, const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& z,
           const std::vector<double>& w, const char* tag_)
        : h(h_), n(x.size()), tag(tag_), X(x.begin(), x.end()), Y(y.begin(), y.end()), Z(z.begin(), z.end()),
          W(w.begin(), w.end())
    {
    }

    std::string nm(const std::string& s) const { return tag + "_" + s; }

    void dev(const vo_t& vo, T_vec& d, const std::vector<T>& host) const
    {
        vo.init_vector(d);
        vo.start_use_vector(d);
        std::vector<T> tmp(host);
        T* p = tmp.data();
        vo.set(p, d);
    }
    void out(const vo_t& vo, const std::string& name, const T_vec& d) const
    {
        std::vector<T> hbuf(n);
        T* p = hbuf.data();
        vo.get(d, p);
        save<T>(nm(name), hbuf.data(), n);
    }

    void vector_operations()
    {
        vo_t vo(n, h);
        T_vec x, y, z, w, t;
        dev(vo, x, X);
        dev(vo, y, Y);
        dev(vo, z, Z);
        dev(vo, w, W);
        vo.init_vector(t);
        vo.start_use_vector(t);
        save1<int64_t>(nm("size"), (int64_t)vo.get_vector_size());
        save1<int64_t>(nm("l2_size"), (int64_t)vo.get_l2_size());
        save1<int64_t>(nm("device_location"), vo.device_location() ? 1 : 0);

        // reductions in both precisions (gpu_vector_operations.h:129-150, 255-403)
        for (int hp = 0; hp < 2; ++hp) {
            if (hp) vo.use_high_precision();
            else vo.use_standard_precision();
            const std::string p = hp ? "hp_" : "sp_";
            save1<int64_t>(nm(p + "precision"), vo.get_current_precision() ? 1 : 0);
            save1<T>(nm(p + "scalar_prod"), vo.scalar_prod(x, y));
            T r;
            vo.scalar_prod(x, y, &r);
            save1<T>(nm(p + "scalar_prod_ptr"), r);
            save1<T>(nm(p + "absolute_sum"), vo.absolute_sum(x));
            vo.absolute_sum(x, &r);
            save1<T>(nm(p + "absolute_sum_ptr"), r);
            save1<T>(nm(p + "norm"), vo.norm(x));
            vo.norm(x, &r);
            save1<T>(nm(p + "norm_ptr"), r);
            save1<T>(nm(p + "norm_l2"), vo.norm_l2(x));
            save1<T>(nm(p + "norm_rank1"), vo.norm_rank1(x, T(3.25)));
            save1<T>(nm(p + "norm_rank1_l2"), vo.norm_rank1_l2(x, T(3.25)));
            T* d4;
            vo.init_vector(d4);
            vo.start_use_vector(d4);
            vo.scalar_prod_dev(x, y, d4);
            vo.norm_dev(x, d4 + 4);
            std::vector<T> hb(n);
            T* hp4 = hb.data();
            vo.get(d4, hp4);
            save<T>(nm(p + "dev8"), hb.data(), 8);
            vo.free_vector(d4);
            vo.assign(y, t);
            save1<T>(nm(p + "normalize"), vo.normalize(t));
            out(vo, p + "normalized", t);
            std::vector<T_vec> V = {x, y, z};
            T md[3];
            vo.multi_scalar_prod(3, V.data(), w, md);
            save<T>(nm(p + "multi_scalar_prod"), md, 3);
            const T coef[3] = {T(0.5), T(-1.25), T(2.0)};
            vo.assign(w, t);
            vo.multi_add_mul(3, coef, V.data(), t);
            out(vo, p + "multi_add_mul", t);
        }
        vo.use_standard_precision();
        save1<int64_t>(nm("valid_before"), vo.check_is_valid_number(x) ? 1 : 0);

        // elementwise (gpu_vector_operations.h:404-493)
        vo.assign_scalar(T(1.5), t);
        out(vo, "assign_scalar", t);
        vo.assign(x, t);
        vo.add_mul_scalar(T(0.75), T(-2.0), t);
        out(vo, "add_mul_scalar", t);
        vo.assign(x, t);
        vo.scale(T(-3.5), t);
        out(vo, "scale", t);
        vo.assign(x, t);
        out(vo, "assign", t);
        T_vec s1, s2;
        dev(vo, s1, X);
        dev(vo, s2, Y);
        vo.swap(s1, s2);
        out(vo, "swap_a", s1);
        out(vo, "swap_b", s2);
        vo.make_abs_copy(x, t);
        out(vo, "make_abs_copy", t);
        vo.assign(y, t);
        vo.make_abs(t);
        out(vo, "make_abs", t);
        vo.assign_mul(T(1.25), x, t);
        out(vo, "assign_mul1", t);
        vo.assign(y, t);
        vo.add_mul(T(-0.5), x, t);
        out(vo, "add_mul1", t);
        vo.assign_mul(T(1.25), x, T(-0.75), y, t);
        out(vo, "assign_mul2", t);
        vo.assign(y, t);
        vo.add_mul(T(1.25), x, T(-0.75), t);
        out(vo, "add_mul2", t);
        vo.assign(z, t);
        vo.add_mul(T(1.25), x, T(-0.75), y, T(0.5), t);
        out(vo, "add_mul3", t);
        vo.mul_pointwise(T(1.5), x, T(-2.0), y, t);
        out(vo, "mul_pointwise3", t);
        vo.assign(x, t);
        vo.mul_pointwise(t, T(-2.0), y);
        out(vo, "mul_pointwise2", t);
        vo.div_pointwise(T(1.5), x, T(-2.0), z, t);
        out(vo, "div_pointwise3", t);
        vo.assign(x, t);
        vo.div_pointwise(t, T(-2.0), z);
        out(vo, "div_pointwise2", t);
        vo.assign(z, t);
        vo.add_mul(T(1.25), x, T(-0.75), y, T(0.5), w, T(2.0), t);
        out(vo, "add_mul4", t);
        vo.assign_mul(T(1.25), x, T(-0.75), y, T(0.5), z, T(2.0), w, t);
        out(vo, "assign_mul4", t);
        vo.assign_mul_compensated(T(1.25), x, T(-0.75), y, t);
        out(vo, "assign_mul_compensated", t);

        // single entries, validity, random fill
        vo.assign(x, t);
        vo.set_value_at_point(T(42.0), 7, t);
        vo.set_value_at_point(T(-1.0), 2, t, 5);
        save1<T>(nm("get_value_at_point"), vo.get_value_at_point(7, t));
        out(vo, "set_value_at_point", t);
        vo.set_value_at_point(T(NAN), n - 1, t);
        save1<int64_t>(nm("valid_after_nan"), vo.check_is_valid_number(t) ? 1 : 0);
        vo.assign_random(t, (uint64_t)1234, (uint32_t)3);
        out(vo, "assign_random01", t);
        vo.assign_random(t, T(-1.0), T(1.0), (uint64_t)1234, (uint32_t)3);
        out(vo, "assign_random_ab", t);

        // host staging: view, set(x), get_buffer, debug_view
        T* v = vo.view(x);
        save<T>(nm("view"), v, n);
        T* buf = vo.get_buffer();
        for (size_t i = 0; i < n; ++i) buf[i] = T(2) * buf[i];
        vo.set(t);
        out(vo, "set_from_buffer", t);
        const std::string dv = g_out + "/" + nm("debug_view.txt");
        vo.debug_view(x, dv);

        // rank-1 lifetime calls
        T_vec r1;
        vo.init_vector_rank1(r1);
        vo.start_use_vector_rank1(r1);
        vo.stop_use_vector(r1);
        vo.free_vector(r1);
        save1<int64_t>(nm("rank1_freed"), r1 == nullptr ? 1 : 0);
        for (T_vec* p : {&x, &y, &z, &w, &t, &s1, &s2}) {
            vo.stop_use_vector(*p);
            vo.free_vector(*p);
        }
    }

    void reduction_objects()
    {
        vo_t vo(n, h);
        T_vec x, y;
        dev(vo, x, X);
        dev(vo, y, Y);
        hpb::hip_reduction_ogita<T, T*> ro(n);
        save1<T>(nm("ogita_sum"), ro.sum(x));
        save1<T>(nm("ogita_asum"), ro.asum(x));
        save1<T>(nm("ogita_dot"), ro.dot(x, y));
        save1<T>(nm("ogita_dot_s"), ro.last_s());
        save1<T>(nm("ogita_dot_c"), ro.last_c());
        save1<T>(nm("ogita_norm"), ro.norm(x));
        hpb::hip_reduction<T, T*> rn(n, h);
        auto mm = rn.min_max(x);
        save1<T>(nm("naive_min"), mm.first);
        save1<T>(nm("naive_max"), mm.second);
        save1<T>(nm("naive_sum"), rn.sum(x));
        save1<T>(nm("naive_sum_debug"), rn.sum_debug(x));
        save1<T>(nm("naive_asum"), rn.asum(x));
        save1<T>(nm("naive_dot"), rn.dot(x, y));
        // exact dot of host arrays (use_cpu, the default) and of device arrays (use_gpu)
        hpb::dot_product_exact<T, T*> ex(512);
        ex.set_arrays(n, X.data(), Y.data());
        save1<T>(nm("exact_dot_host"), ex.dot_exact());
        save1<double>(nm("exact_hi"), ex.exact_hi());
        save1<double>(nm("exact_lo"), ex.exact_lo());
        const T approx = ro.dot(x, y);
        save1<T>(nm("err_abs"), ex.get_error(approx));
        save1<T>(nm("err_rel"), ex.get_error_relative(approx));
        save1<T>(nm("err_abs_T"), ex.get_error_T(approx));
        save1<T>(nm("err_rel_T"), ex.get_error_relative_T(approx));
        save1<T>(nm("sum_two"), ex.sum_two(T(1), T(1e-30)));
        hpb::dot_product_exact<T, T*> exd(512, h);
        exd.use_gpu(n);
        exd.set_arrays(n, x, y);
        save1<T>(nm("exact_dot_dev"), exd.dot_exact());
        vo.free_vector(x);
        vo.free_vector(y);
    }

    void generator()
    {
        vo_t vo(n, h);
        hpb::hip_vector_operations<double> vod(n, h);
        hpb::dot_product_exact<double, double*> ex;
        hpb::generate_vector_pair<vo_t, hpb::hip_vector_operations<double>, hpb::dot_product_exact<double, double*>>
            gen(&vo, &vod, &ex);
        gen.dot_exact_cuda();
        gen.set_seed(5);
        T_vec x, y;
        vo.init_vector(x);
        vo.start_use_vector(x);
        vo.init_vector(y);
        vo.start_use_vector(y);
        const T kappa = gen.generate(x, y, T(std::is_same<T, double>::value ? 1e16 : 1e6));
        save1<T>(nm("gen_kappa"), kappa);
        save1<int64_t>(nm("gen_seed"), (int64_t)gen.last_seed());
        save1<int64_t>(nm("gen_tail"), (int64_t)gen.last_tail_length());
        save1<double>(nm("gen_exact_hi"), gen.exact_hi());
        save1<double>(nm("gen_exact_lo"), gen.exact_lo());
        save<int64_t>(nm("gen_limbs"), reinterpret_cast<const int64_t*>(gen.exact_limbs()), 4);
        save1<T>(nm("gen_kappa_again"), gen.condition_number_max(x, y));
        out(vo, "gen_x", x);
        out(vo, "gen_y", y);
        const T kappa2 = gen.generate(x, y, T(std::is_same<T, double>::value ? 1e16 : 1e6));
        save1<T>(nm("gen2_kappa"), kappa2);
        save1<int64_t>(nm("gen2_seed"), (int64_t)gen.last_seed());
        vo.free_vector(x);
        vo.free_vector(y);
    }

    void matrix(const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<double>& val64)
    {
        const int rows = (int)rp.size() - 1;
        const int64_t nnz = (int64_t)ci.size();
        std::vector<T> val(val64.begin(), val64.end());
        csr::hip_matrix<T> A(h);
        A.init_set(nnz, nnz, rows + 1, val.data(), ci.data(), rp.data(), rows, rows);
        save1<int64_t>(nm("mat_rows"), A.get_dim().rows);
        save1<int64_t>(nm("mat_cols"), A.get_dim().columns);
        hpb::hip_vector_operations<T> vo((size_t)rows, h);
        std::vector<T> xs(X.begin(), X.begin() + rows), ys(Y.begin(), Y.begin() + rows);
        T_vec x, y, t;
        vo.init_vector(x);
        vo.start_use_vector(x);
        T* px = xs.data();
        vo.set(px, x);
        vo.init_vector(y);
        vo.start_use_vector(y);
        T* py = ys.data();
        vo.set(py, y);
        vo.init_vector(t);
        vo.start_use_vector(t);
        auto o = [&](const std::string& name, const T_vec& d) {
            std::vector<T> hb((size_t)rows);
            T* p = hb.data();
            vo.get(d, p);
            save<T>(nm(name), hb.data(), (size_t)rows);
        };
        vo.assign(y, t);
        A.axpy(T(1.5), x, T(-0.5), t);
        o("mat_axpy", t);
        vo.assign(y, t);
        A.axpy(T(1.5), x, T(-0.5), t, true);
        o("mat_axpy_t", t);
        A.triangular_solve_lower(y, t, T(2.0));
        o("mat_lower", t);
        A.triangular_solve_upper(y, t, T(2.0));
        o("mat_upper", t);
        A.triangular_solve_lower(y, t, T(2.0), true);
        o("mat_lower_t", t);
        A.triangular_solve_upper(y, t, T(2.0), true);
        o("mat_upper_t", t);
        vo.free_vector(x);
        vo.free_vector(y);
        vo.free_vector(t);
    }
};

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <in_dir> <out_dir>\n", argv[0]);
        return 2;
    }
    const std::string in = argv[1];
    g_out = argv[2];
    try {
        auto x = load<double>(in + "/x.f64"), y = load<double>(in + "/y.f64"), z = load<double>(in + "/z.f64"),
             w = load<double>(in + "/w.f64");
        auto rp = load<int>(in + "/rp.i32"), ci = load<int>(in + "/ci.i32");
        auto val = load<double>(in + "/val.f64");
        hpb_handle_t h;
        hpb::check(hpb_create(&h, nullptr), "hpb_create");
        {
            runner<double> r(h, x, y, z, w, "d");
            std::printf("f64 vector operations\n");
            std::fflush(stdout);
            r.vector_operations();
            std::printf("f64 reduction objects\n");
            std::fflush(stdout);
            r.reduction_objects();
            std::printf("f64 generator\n");
            std::fflush(stdout);
            r.generator();
            std::printf("f64 matrix\n");
            std::fflush(stdout);
            r.matrix(rp, ci, val);
        }
        {
            runner<float> r(h, x, y, z, w, "s");
            std::printf("f32 vector operations\n");
            std::fflush(stdout);
            r.vector_operations();
            std::printf("f32 reduction objects\n");
            std::fflush(stdout);
            r.reduction_objects();
            std::printf("f32 generator\n");
            std::fflush(stdout);
            r.generator();
            std::printf("f32 matrix\n");
            std::fflush(stdout);
            r.matrix(rp, ci, val);
        }
        hpb_destroy(h);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "test_dropin: %s\n", e.what());
        return 1;
    }
    std::printf("done\n");
    return 0;
}
