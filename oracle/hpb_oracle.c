/*
 * hpb_oracle.c -- CPU oracle for the compensated BLAS1 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, on the host, the algorithms of
 * evstigneevnm/high_precision_blas1 that the GPU library re-implements.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only as the checker.  The GPU library never links or calls it.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (see oracle/build_oracle.py).
 * -ffp-contract=off is essential: every a*b+c below must round twice, exactly as
 * written, so that TwoSum/TwoProd stay error-free and the restated orders hold.
 *
 * Contents (each with the reference file:line it follows):
 *   two_sum / two_prod ........ source/common/dot_product.h:50-66
 *   dot_naive / dot2 ........... source/common/dot_product.h:96-121
 *   sum_naive / kahan / sum2 ... source/sums.hpp:52-66,110-176
 *   threaded dot2 / sum2 ....... source/common/threaded_reduction.h:128-178,220-246,
 *                                307-399,418-435 (merge in part order, see below)
 *   reference GPU tree order ... source/common/ogita/gpu_reduction_ogita_impl.cuh:248-676,
 *                                683-694,1169-1282 (block 1024, warp 32, multi-pass)
 *   error bounds ............... source/high_prec/error_bounds.hpp:156-308
 *   exact dot / sum ............ replaces source/high_prec/dot_product_gmp.hpp:108-130
 *                                (GMP mpf loop) with a Kulisch long accumulator
 *   Philox4x32-10 .............. Salmon, Moraes, Dror, Shaw, SC'11 (no reference
 *                                counterpart: the reference seeds cuRAND from
 *                                std::random_device, gpu_vector_operations.h:499-504)
 *   GenDot pair generator ...... Ogita-Rump-Oishi 2005 Alg. 6.1, the reference's variant
 *                                is source/generate_vector_pair.hpp:119-182
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Error-free transformations                                                */
/* ------------------------------------------------------------------------ */

/* Knuth TwoSum: s = fl(a+b), a+b = s+t exactly.  dot_product.h:58-66 */
static inline double two_sum(double a, double b, double *t)
{
    double s = a + b;
    double bs = s - a;
    double as = s - bs;
    *t = (b - bs) + (a - as);
    return s;
}
static inline float two_sumf(float a, float b, float *t)
{
    float s = a + b;
    float bs = s - a;
    float as = s - bs;
    *t = (b - bs) + (a - as);
    return s;
}

/* TwoProd with FMA: p = fl(ab), ab = p+e exactly.  dot_product.h:50-56 */
static inline double two_prod(double a, double b, double *e)
{
    double p = a * b;
    *e = fma(a, b, -p);
    return p;
}
static inline float two_prodf(float a, float b, float *e)
{
    float p = a * b;
    *e = fmaf(a, b, -p);
    return p;
}

ORC_API double orc_two_sum(double a, double b, double *t) { return two_sum(a, b, t); }
ORC_API double orc_two_prod(double a, double b, double *e) { return two_prod(a, b, e); }

/* ------------------------------------------------------------------------ */
/* Sequential reference algorithms                                           */
/* ------------------------------------------------------------------------ */

/* dot_product<T>::dot_naive, dot_product.h:96-104: s += x_j*y_j (two roundings). */
ORC_API double orc_dot_naive(size_t n, const double *x, const double *y)
{
    double s = 0.0;
    for (size_t j = 0; j < n; ++j) s += x[j] * y[j];
    return s;
}

/* dot_product<T>::dot_ogita, dot_product.h:108-121 (ORO Dot2):
 *   p = TwoProd(pi, x_j, y_j); s = TwoSum(t, s, p); c = c + pi + t; return s + c.
 * Note the association (c + pi) + t of the reference's CPU code. */
ORC_API double orc_dot2(size_t n, const double *x, const double *y)
{
    double s = 0.0, c = 0.0;
    for (size_t j = 0; j < n; ++j) {
        double pi, t;
        double p = two_prod(x[j], y[j], &pi);
        s = two_sum(s, p, &t);
        c = c + pi + t;
    }
    return s + c;
}

/* Same algorithm in binary32 (the reference builds it with -DTYPE=float, Makefile:2). */
ORC_API float orc_dot2_f32(size_t n, const float *x, const float *y)
{
    float s = 0.0f, c = 0.0f;
    for (size_t j = 0; j < n; ++j) {
        float pi, t;
        float p = two_prodf(x[j], y[j], &pi);
        s = two_sumf(s, p, &t);
        c = c + pi + t;
    }
    return s + c;
}
ORC_API float orc_dot_naive_f32(size_t n, const float *x, const float *y)
{
    float s = 0.0f;
    for (size_t j = 0; j < n; ++j) s += x[j] * y[j];
    return s;
}

/* sums<T>::sum_naive, sums.hpp:110-121 */
ORC_API double orc_sum_naive(size_t n, const double *x)
{
    double s = 0.0;
    for (size_t j = 0; j < n; ++j) s += x[j];
    return s;
}

/* sums<T>::sum_kahan, sums.hpp:143-154 (FastTwoSum form, sums.hpp:52-57) */
ORC_API double orc_sum_kahan(size_t n, const double *x)
{
    double sum = 0.0, c = 0.0;
    for (size_t j = 0; j < n; ++j) {
        double yv = x[j] + c;
        double s = sum + yv;
        c = yv - (s - sum);
        sum = s;
    }
    return sum;
}

/* sums<T>::sum_rump (ORO Sum2), sums.hpp:163-176: s = TwoSum(e, s, x); c += e. */
ORC_API double orc_sum2(size_t n, const double *x)
{
    double s = 0.0, c = 0.0;
    for (size_t j = 0; j < n; ++j) {
        double e;
        s = two_sum(s, x[j], &e);
        c += e;
    }
    return s + c;
}

/* Compensated sum of |x_j|: threaded_reduction.h:140-147 with use_abs (one part).
 * This is the true sum of absolute values; the reference's GPU asum for real T
 * actually computes Sum2(x) (gpu_reduction_ogita_impl_functions.cuh:214-222). */
ORC_API double orc_asum2(size_t n, const double *x)
{
    double s = 0.0, c = 0.0;
    for (size_t j = 0; j < n; ++j) {
        double e;
        s = two_sum(s, fabs(x[j]), &e);
        c = c + e;
    }
    return s + c;
}

ORC_API float orc_sum2_f32(size_t n, const float *x)
{
    float s = 0.0f, c = 0.0f;
    for (size_t j = 0; j < n; ++j) {
        float e;
        s = two_sumf(s, x[j], &e);
        c += e;
    }
    return s + c;
}

/* ------------------------------------------------------------------------ */
/* Threaded reduction (threaded_reduction.h)                                 */
/* ------------------------------------------------------------------------ */
/* Partition, threaded_reduction.h:418-435: parts contiguous slices of
 * n/parts elements, the last one also takes the remainder.
 * Per-slice Dot2 (threaded_reduction.h:226-238) uses c = c + pi + t.
 * Merge, threaded_reduction.h:240-244, 347:
 *   result = TwoSum(t, result, s); sigma1 = TwoSum(qq, sigma1, c); sigma2 += t + qq
 *   return result + sigma1 + sigma2
 * The reference merges in mutex-acquisition order; this restatement merges in
 * slice order, which is one of the orders the reference can produce and makes
 * the result reproducible. */

typedef struct {
    const double *x, *y;
    size_t begin, end;
    int mode; /* 0 dot2, 1 dot naive, 2 sum2, 3 asum2, 4 sum naive */
    double s, c;
} orc_slice_t;

static void *orc_slice_run(void *arg)
{
    orc_slice_t *w = (orc_slice_t *)arg;
    double s = 0.0, c = 0.0, t;
    size_t j;
    switch (w->mode) {
    case 0:
        for (j = w->begin; j < w->end; ++j) {
            double pi;
            double p = two_prod(w->x[j], w->y[j], &pi);
            s = two_sum(s, p, &t);
            c = c + pi + t;
        }
        break;
    case 1:
        for (j = w->begin; j < w->end; ++j) s += w->x[j] * w->y[j];
        break;
    case 2:
        for (j = w->begin; j < w->end; ++j) {
            s = two_sum(s, w->x[j], &t);
            c = c + t;
        }
        break;
    case 3:
        for (j = w->begin; j < w->end; ++j) {
            s = two_sum(s, fabs(w->x[j]), &t);
            c = c + t;
        }
        break;
    default:
        for (j = w->begin; j < w->end; ++j) s += w->x[j];
        break;
    }
    w->s = s;
    w->c = c;
    return NULL;
}

static double orc_threaded(size_t n, const double *x, const double *y, int parts, int mode)
{
    if (parts < 1) parts = 1;
    orc_slice_t *w = (orc_slice_t *)calloc((size_t)parts, sizeof(orc_slice_t));
    pthread_t *th = (pthread_t *)calloc((size_t)parts, sizeof(pthread_t));
    size_t delta = n / (size_t)parts, rem = n % (size_t)parts, b = 0;
    for (int p = 0; p < parts; ++p) {
        size_t e = b + delta + (p == parts - 1 ? rem : 0);
        w[p].x = x; w[p].y = y; w[p].begin = b; w[p].end = e; w[p].mode = mode;
        b = e;
    }
    for (int p = 1; p < parts; ++p) pthread_create(&th[p], NULL, orc_slice_run, &w[p]);
    orc_slice_run(&w[0]);
    for (int p = 1; p < parts; ++p) pthread_join(th[p], NULL);
    double result = 0.0, sig1 = 0.0, sig2 = 0.0;
    for (int p = 0; p < parts; ++p) {
        if (mode == 1 || mode == 4) {
            result = result + w[p].s;                 /* threaded_reduction.h:119-120 */
        } else {
            double t, qq;
            result = two_sum(result, w[p].s, &t);
            sig1 = two_sum(sig1, w[p].c, &qq);
            sig2 = sig2 + t + qq;
        }
    }
    free(w);
    free(th);
    if (mode == 1 || mode == 4) return result;
    return result + sig1 + sig2;
}

ORC_API double orc_dot2_threaded(size_t n, const double *x, const double *y, int parts)
{ return orc_threaded(n, x, y, parts, 0); }
ORC_API double orc_dot_naive_threaded(size_t n, const double *x, const double *y, int parts)
{ return orc_threaded(n, x, y, parts, 1); }
ORC_API double orc_sum2_threaded(size_t n, const double *x, int parts)
{ return orc_threaded(n, x, x, parts, 2); }
ORC_API double orc_asum2_threaded(size_t n, const double *x, int parts)
{ return orc_threaded(n, x, x, parts, 3); }

/* ------------------------------------------------------------------------ */
/* Reference GPU tree order (gpu_reduction_ogita_impl.cuh)                   */
/* ------------------------------------------------------------------------ */
/* Emulates reduce_dot_ogita_kernel / reduce_sum_ogita_kernel with the launch rule
 * of get_blocks_threads_shmem (:683-694, BLOCK_SIZE 1024) and the host driver
 * reduction_dot / reduction_sum (:1169-1282).  Per thread (:490-518):
 *   s = TwoSum(t, s, p) with p = TwoProd(pi, x_i, y_i); c += (pi + t) + 0
 * for i = b*2T + tid, then i + T, striding 2T*blocks.  Tree (:527-581) for
 * strides 512..64: s = TwoSum(t, s, s'); c = (c + t) + c'.  Then warp-32 part
 * (:583-600): stride 32 as c = c + (t + c'), xor-shuffle 16..1 the same way; the
 * value kept is lane 0's.  Later passes run reduce_sum over the (s, c) pairs:
 *   s = TwoSum(t, s, s_i); c += t + c_i.  Result s + c on the host. */

static inline void rt_merge_tree(double *s, double *c, double s2, double c2)
{
    double t;
    *s = two_sum(*s, s2, &t);
    *c = (*c + t) + c2;
}
static inline void rt_merge_warp(double *s, double *c, double s2, double c2)
{
    double t;
    *s = two_sum(*s, s2, &t);
    *c = *c + (t + c2);
}

static unsigned rt_next_pow2(unsigned v)
{
    --v; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; return ++v;
}

/* Block reduction of T thread values (s[], c[]) following :519-600.
 * The warp-shuffle branch needs tid < 32 lanes; blockSize is the template
 * parameter (= threads).  For blockSize < 64 the shuffle loop runs over all 16..1
 * strides with lanes >= blockSize reading zero-initialised shared memory in the
 * reference's non-shuffle path; the CUDA >= 3.0 shuffle path reads registers of
 * lanes whose (s, c) stayed 0.  Both give the same value as merging with zeros. */
static void rt_block_reduce(unsigned T, double *s, double *c, double *out_s, double *out_c)
{
    for (unsigned st = 512; st >= 64; st >>= 1) {
        if (T >= st * 2) {
            for (unsigned tid = 0; tid < st; ++tid) rt_merge_tree(&s[tid], &c[tid], s[tid + st], c[tid + st]);
        }
    }
    /* warp part: lanes 0..31 */
    double ws[32], wc[32];
    for (unsigned l = 0; l < 32; ++l) { ws[l] = (l < T) ? s[l] : 0.0; wc[l] = (l < T) ? c[l] : 0.0; }
    if (T >= 64) {
        for (unsigned l = 0; l < 32; ++l) rt_merge_warp(&ws[l], &wc[l], s[l + 32], c[l + 32]);
    }
    for (unsigned off = 16; off > 0; off >>= 1) {
        double ns[32], nc[32];
        for (unsigned l = 0; l < 32; ++l) {
            ns[l] = ws[l]; nc[l] = wc[l];
            rt_merge_warp(&ns[l], &nc[l], ws[l ^ off], wc[l ^ off]);
        }
        memcpy(ws, ns, sizeof ws); memcpy(wc, nc, sizeof wc);
    }
    *out_s = ws[0];
    *out_c = wc[0];
}

static void rt_launch(unsigned n, unsigned *threads, unsigned *blocks)
{
    const unsigned B = 1024;
    *threads = (n < B * 2) ? rt_next_pow2((n + 1) / 2) : B;
    *blocks = (n + (*threads * 2 - 1)) / (*threads * 2);
}

/* one pass; mode 0: dot of (x, y); mode 1: sum of pairs (x = s values, y = c values) */
static unsigned rt_pass(unsigned n, const double *x, const double *y, int mode, double *os, double *oc)
{
    unsigned T, G;
    rt_launch(n, &T, &G);
    double *s = (double *)malloc(sizeof(double) * T);
    double *c = (double *)malloc(sizeof(double) * T);
    for (unsigned b = 0; b < G; ++b) {
        for (unsigned tid = 0; tid < T; ++tid) {
            double ss = 0.0, cc = 0.0, t, pi;
            for (unsigned i = b * T * 2 + tid; i < n; i += T * 2 * G) {
                if (mode == 0) {
                    double p = two_prod(x[i], y[i], &pi);
                    ss = two_sum(ss, p, &t);
                    cc += (pi + t) + 0.0;
                } else {
                    ss = two_sum(ss, x[i], &t);
                    cc += t + y[i];
                }
                if (i + T < n) {
                    unsigned k = i + T;
                    if (mode == 0) {
                        double p = two_prod(x[k], y[k], &pi);
                        ss = two_sum(ss, p, &t);
                        cc += (pi + t) + 0.0;
                    } else {
                        ss = two_sum(ss, x[k], &t);
                        cc += t + y[k];
                    }
                }
            }
            s[tid] = ss; c[tid] = cc;
        }
        rt_block_reduce(T, s, c, &os[b], &oc[b]);
    }
    free(s); free(c);
    return G;
}

/* Returns s + c of the reference's multi-pass GPU Dot2 (mode 0) or Sum2 (mode 2,
 * y ignored).  Requires n < 2^31 like the reference (int sizes). */
ORC_API double orc_dot2_refgpu_order(size_t n, const double *x, const double *y, int is_sum)
{
    if (n == 0) return 0.0;
    unsigned T, G;
    rt_launch((unsigned)n, &T, &G);
    double *ps = (double *)malloc(sizeof(double) * G);
    double *pc = (double *)malloc(sizeof(double) * G);
    double *zeros = NULL;
    if (is_sum) {
        zeros = (double *)calloc(n, sizeof(double));
        rt_pass((unsigned)n, x, zeros, 1, ps, pc);
    } else {
        rt_pass((unsigned)n, x, y, 0, ps, pc);
    }
    unsigned sN = G;
    while (sN > 1) {
        unsigned T2, G2;
        rt_launch(sN, &T2, &G2);
        double *ns = (double *)malloc(sizeof(double) * G2);
        double *nc = (double *)malloc(sizeof(double) * G2);
        rt_pass(sN, ps, pc, 1, ns, nc);
        free(ps); free(pc);
        ps = ns; pc = nc;
        sN = G2;
    }
    double r = ps[0] + pc[0];
    free(ps); free(pc); free(zeros);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Error bounds (error_bounds.hpp:156-308), restated with a chosen eps.      */
/* ------------------------------------------------------------------------ */
/* gamma_n = n eps / (1 - n eps), error_bounds.hpp:156-160.  The reference sets
 * eps = numeric_limits<T>::epsilon() (= 2u); pass eps = 2^-53 for the ORO
 * convention.  Indices follow the reference: 0 naive sequential, 1 naive block,
 * 2 naive pairwise, 3 compensated pairwise, 4 ORO (compensated sequential),
 * 5 compensated block (update_real_dots, error_bounds.hpp:288-297). */
static long double bnd_gamma(long double eps, long double n) { return n * eps / (1.0L - n * eps); }

ORC_API void orc_dot_bounds(size_t n, double cond, size_t k_blocks, double eps_d, double *out6)
{
    long double eps = eps_d, C = cond;
    size_t b = (size_t)ceil(1.0 * (double)n / (1.0 * (double)k_blocks));
    long double gk, gb, gl;
    out6[0] = (double)(0.5L * bnd_gamma(eps, (long double)n) * C);
    gk = bnd_gamma(eps, (long double)k_blocks); gb = bnd_gamma(eps, (long double)b);
    out6[1] = (double)(0.5L * (gb + gk + gb * gk) * C);
    gl = bnd_gamma(eps, (long double)((size_t)ceil(log2((double)n)) + 1));
    out6[2] = (double)(0.5L * gl * C);
    gl = bnd_gamma(eps, (long double)((size_t)ceil(log2((double)n)) + 2));
    out6[3] = (double)(eps + 0.5L * gl * gl * C);
    gl = bnd_gamma(eps, (long double)n);
    out6[4] = (double)(eps + 0.5L * gl * gl * C);
    out6[5] = (double)(eps + 0.5L * (2.0L * gk * gk + gb * gb + gk * gb + gk * gb * gb) * C);
}

ORC_API void orc_sum_bounds(size_t n, double cond, size_t k_blocks, double eps_d, double *out6)
{
    long double eps = eps_d, C = cond;
    size_t b = (size_t)ceil(1.0 * (double)n / (1.0 * (double)k_blocks));
    long double gk = bnd_gamma(eps, (long double)k_blocks - 1), gb = bnd_gamma(eps, (long double)b - 1), gl;
    out6[0] = (double)(bnd_gamma(eps, (long double)n - 1) * C);
    out6[1] = (double)((gb + gk + gb * gk) * C);
    gl = bnd_gamma(eps, (long double)(size_t)ceil(log2((double)n)));
    out6[2] = (double)(gl * C);
    gl = bnd_gamma(eps, (long double)((size_t)ceil(log2((double)n)) + 1));
    out6[3] = (double)(eps + gl * gl * C);
    gl = bnd_gamma(eps, (long double)n - 1);
    out6[4] = (double)(eps + gl * gl * C);
    out6[5] = (double)(eps + (2.0L * gk * gk + gb * gb + gk * gb + 2.0L * gk * gb * gb) * C);
}

/* ------------------------------------------------------------------------ */
/* Exact long accumulator (Kulisch) over the full binary64 range             */
/* ------------------------------------------------------------------------ */
/* Bit 0 of the accumulator is 2^-1074.  Limbs hold 32-bit digits in int64_t so
 * that 2^30 additions can be absorbed before carries are propagated.  The result
 * is rounded once, to nearest even.  This replaces the reference's GMP mpf loop
 * (dot_product_gmp.hpp:120-125, 512-1024 bits), which is exact for these inputs
 * too; get_d() there truncates, this rounds to nearest. */

#define KL_LIMBS 72 /* 72*32 = 2304 bits > 2098 + headroom */

typedef struct {
    int64_t d[KL_LIMBS];
    long pending;
    int nonfinite;   /* 1: +inf/-inf/nan seen */
    double nf_sum;
} kacc_t;

static void kacc_init(kacc_t *a) { memset(a, 0, sizeof *a); }

static void kacc_norm(kacc_t *a)
{
    int64_t carry = 0;
    for (int i = 0; i < KL_LIMBS; ++i) {
        int64_t v = a->d[i] + carry;
        carry = v >> 32;                 /* arithmetic shift: floor division */
        a->d[i] = v - (carry << 32);     /* in [0, 2^32) */
    }
    a->d[KL_LIMBS - 1] += carry << 32;   /* keep the sign in the top limb */
    a->pending = 0;
}

static void kacc_add(kacc_t *a, double v)
{
    if (v == 0.0) return;
    if (!isfinite(v)) { a->nonfinite = 1; a->nf_sum += v; return; }
    uint64_t bits;
    memcpy(&bits, &v, 8);
    int neg = (int)(bits >> 63);
    int ex = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1);
    int off;
    if (ex == 0) { off = 0; } else { m |= (1ull << 52); off = ex - 1; }
    /* value = m * 2^(off - 1074); place m at bit 'off' */
    int li = off >> 5, sh = off & 31;
    unsigned __int128 w = (unsigned __int128)m << sh; /* up to 85 bits */
    int64_t p0 = (int64_t)(uint32_t)(w);
    int64_t p1 = (int64_t)(uint32_t)(w >> 32);
    int64_t p2 = (int64_t)(uint32_t)(w >> 64);
    if (neg) { p0 = -p0; p1 = -p1; p2 = -p2; }
    a->d[li] += p0;
    a->d[li + 1] += p1;
    a->d[li + 2] += p2;
    if (++a->pending >= (1l << 30)) kacc_norm(a);
}

/* Round the accumulated value to the nearest double (ties to even). */
static double kacc_round(kacc_t *a)
{
    if (a->nonfinite) return a->nf_sum;
    kacc_norm(a);
    uint32_t mag[KL_LIMBS];
    int neg = a->d[KL_LIMBS - 1] < 0;
    /* two's complement magnitude over KL_LIMBS 32-bit digits */
    uint64_t borrow = 1;
    for (int i = 0; i < KL_LIMBS; ++i) {
        uint32_t dig = (uint32_t)(a->d[i] & 0xffffffffu);
        if (neg) {
            uint64_t v = (uint64_t)(uint32_t)(~dig) + borrow;
            mag[i] = (uint32_t)v;
            borrow = v >> 32;
        } else {
            mag[i] = dig;
        }
    }
    int top = -1;
    for (int i = KL_LIMBS - 1; i >= 0; --i) if (mag[i]) { top = i; break; }
    if (top < 0) return 0.0;
    int msb = top * 32 + (31 - __builtin_clz(mag[top]));   /* bit index of MSB */
    int ulp = msb - 52;
    if (ulp < 0) ulp = 0;                                 /* subnormal range */
    /* q = floor(mag / 2^ulp) (<= 2^53), rem bits for rounding */
    uint64_t q = 0;
    for (int b = msb; b >= ulp; --b) q = (q << 1) | ((mag[b >> 5] >> (b & 31)) & 1u);
    int half = 0, sticky = 0;
    if (ulp > 0) {
        half = (mag[(ulp - 1) >> 5] >> ((ulp - 1) & 31)) & 1u;
        for (int b = ulp - 2; b >= 0 && !sticky; --b) {
            if ((b & 31) == 31 && mag[b >> 5] == 0) { b -= 31; continue; }
            if ((mag[b >> 5] >> (b & 31)) & 1u) sticky = 1;
        }
    }
    if (half && (sticky || (q & 1))) q += 1;
    double r = ldexp((double)q, ulp - 1074);
    return neg ? -r : r;
}

/* Exact sum of x rounded once.  */
ORC_API double orc_exact_sum(size_t n, const double *x)
{
    kacc_t *a = (kacc_t *)malloc(sizeof(kacc_t));
    kacc_init(a);
    for (size_t j = 0; j < n; ++j) kacc_add(a, x[j]);
    double r = kacc_round(a);
    free(a);
    return r;
}

/* Exact dot: every product is split exactly as p + e (TwoProd) and both parts
 * are accumulated.  Valid while no product overflows or underflows below the
 * subnormal range (|x_j y_j| >= 2^-969 or exactly 0). out2 (optional) receives
 * the exact value as an unevaluated double-double (hi, lo), hi = round(exact). */
ORC_API double orc_exact_dot(size_t n, const double *x, const double *y, double *out2)
{
    kacc_t *a = (kacc_t *)malloc(sizeof(kacc_t));
    kacc_init(a);
    for (size_t j = 0; j < n; ++j) {
        double e, p = two_prod(x[j], y[j], &e);
        kacc_add(a, p);
        kacc_add(a, e);
    }
    double hi = kacc_round(a);
    if (out2) {
        kacc_add(a, -hi);
        out2[0] = hi;
        out2[1] = kacc_round(a);
    }
    free(a);
    return hi;
}

ORC_API double orc_exact_dot_f32(size_t n, const float *x, const float *y)
{
    kacc_t *a = (kacc_t *)malloc(sizeof(kacc_t));
    kacc_init(a);
    for (size_t j = 0; j < n; ++j) kacc_add(a, (double)x[j] * (double)y[j]); /* exact: 48-bit product */
    double r = kacc_round(a);
    free(a);
    return r;
}

/* Relative error |res - exact| / |exact| with the difference formed exactly,
 * clipped at 1 like obtain_condition_precision.cpp:46-50.
 * The dot products are recomputed exactly here; for repeated use prefer
 * orc_exact_dot + orc_rel_err_dd. */
ORC_API double orc_rel_err_dd(double res, double hi, double lo)
{
    if (hi == 0.0 && lo == 0.0) return res == 0.0 ? 0.0 : 1.0;
    double t;
    double d = two_sum(res, -hi, &t);
    double err = fabs((d - lo) + t);
    double rel = err / fabs(hi);
    return rel > 1.0 ? 1.0 : rel;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11)                                       */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

ORC_API void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* GenDot-style ill-conditioned pair generator with exactly tracked dot      */
/* ------------------------------------------------------------------------ */
/* Identical arithmetic runs on the GPU (hpb_gendot); the two must agree bit for
 * bit.  Specification:
 *
 *   b = log2(cond) (passed in), hb = b/2, n_head = n - n_tail.
 *   Random words: Philox4x32-10, key = (seed_lo, seed_hi),
 *     counter = (i, stream, 0, 0), stream 0 and 1 give 8 words per element.
 *   k_a = bits 11.. of (w0<<32|w1), k_b of (w2<<32|w3) [stream 0], k_c of
 *     (w4<<32|w5) [stream 1]: 53-bit integers.
 *   unif(k) = k 2^-53 in [0,1);  mant(k) = (2k + 1 - 2^53) 2^-53 in (-1,1), never 0.
 *   head i < n_head:  e = rint(unif(k_a) hb), with e = rint(hb)+1 at i = 0 and
 *     e = 0 at i = n_head-1 (ORO Alg. 6.1); x = mant(k_b) 2^e; y = mant(k_c) 2^e.
 *   tail j = i - n_head: e = rint(hb (1 - j/(n_tail-1))) (0 if n_tail == 1);
 *     x = mant(k_b) 2^e; target r = mant(k_c) 2^e.
 *   Every x and y is a multiple of 2^-53 (grid53 below), so every product is a
 *   multiple of 2^-106 and the running dot is held exactly in a 256-bit two's
 *   complement fixed-point accumulator with LSB 2^-106.
 *   Tail chunks of GD_CHUNK_MIN = 256 elements, the length doubled until there are
 *   at most GD_MAX_CHUNKS = 4096 chunks.  For a tail element that is not the last
 *   of its chunk: y = grid53(fl(fl(r - r_prev) / x)), r_prev = target of the
 *   previous element, or fl(S_head) for the first tail element.  The last element
 *   of each chunk cancels the exact running sum S (head + all earlier elements):
 *   y = grid53(fl(fl(r - fl(S)) / x)).  S at the end is the exact dot.
 *   grid53(v) = v if |v| >= 1/2 else rint(v 2^53) 2^-53.
 */

#define GD_CHUNK_MIN 256
#define GD_MAX_CHUNKS 4096

static size_t gd_chunk_len(size_t n_tail)
{
    size_t c = GD_CHUNK_MIN;
    while ((n_tail + c - 1) / c > GD_MAX_CHUNKS) c *= 2;
    return c;
}

typedef struct { uint64_t w[4]; } acc256_t; /* two's complement, LSB 2^-106 */

static void acc_zero(acc256_t *a) { memset(a, 0, sizeof *a); }

static void acc_add256(acc256_t *a, const acc256_t *b)
{
    unsigned __int128 carry = 0;
    for (int i = 0; i < 4; ++i) {
        unsigned __int128 v = (unsigned __int128)a->w[i] + b->w[i] + carry;
        a->w[i] = (uint64_t)v;
        carry = v >> 64;
    }
}

/* add v (multiple of 2^-106, |v| < 2^149) exactly */
static void acc_add_double(acc256_t *a, double v)
{
    if (v == 0.0) return;
    uint64_t bits;
    memcpy(&bits, &v, 8);
    int neg = (int)(bits >> 63);
    int ex = (int)((bits >> 52) & 0x7ff);
    uint64_t m = (bits & ((1ull << 52) - 1)) | (1ull << 52);
    int sh = ex - 1075 + 106;      /* value = m 2^(ex-1075) */
    if (sh < 0) { m >>= -sh; sh = 0; }
    acc256_t t;
    acc_zero(&t);
    int li = sh >> 6, bs = sh & 63;
    t.w[li] = m << bs;
    if (bs && li + 1 < 4) t.w[li + 1] = m >> (64 - bs);
    if (neg) {
        /* two's complement negate */
        unsigned __int128 carry = 1;
        for (int i = 0; i < 4; ++i) {
            unsigned __int128 s = (unsigned __int128)(~t.w[i]) + carry;
            t.w[i] = (uint64_t)s;
            carry = s >> 64;
        }
    }
    acc_add256(a, &t);
}

/* nearest double (ties to even) of the fixed-point value */
static double acc_round(const acc256_t *a)
{
    uint64_t m[4];
    int neg = (int)(a->w[3] >> 63);
    if (neg) {
        unsigned __int128 carry = 1;
        for (int i = 0; i < 4; ++i) {
            unsigned __int128 s = (unsigned __int128)(~a->w[i]) + carry;
            m[i] = (uint64_t)s;
            carry = s >> 64;
        }
    } else {
        memcpy(m, a->w, sizeof m);
    }
    int top = -1;
    for (int i = 3; i >= 0; --i) if (m[i]) { top = i; break; }
    if (top < 0) return 0.0;
    int msb = top * 64 + 63 - __builtin_clzll(m[top]);
    int ulp = msb - 52;
    if (ulp < 0) ulp = 0;
    uint64_t q = 0;
    for (int b = msb; b >= ulp; --b) q = (q << 1) | ((m[b >> 6] >> (b & 63)) & 1ull);
    int half = 0, sticky = 0;
    if (ulp > 0) {
        half = (int)((m[(ulp - 1) >> 6] >> ((ulp - 1) & 63)) & 1ull);
        for (int b = ulp - 2; b >= 0; --b)
            if ((m[b >> 6] >> (b & 63)) & 1ull) { sticky = 1; break; }
    }
    if (half && (sticky || (q & 1))) q += 1;
    double r = ldexp((double)q, ulp - 106);
    return neg ? -r : r;
}

static inline uint64_t gd_k53(uint32_t hi, uint32_t lo) { return (((uint64_t)hi << 32) | lo) >> 11; }
static inline double gd_unif(uint64_t k) { return (double)k * 0x1p-53; }
static inline double gd_mant(uint64_t k) { return (double)((int64_t)(2 * k + 1) - ((int64_t)1 << 53)) * 0x1p-53; }
static inline double gd_grid53(double v) { return fabs(v) >= 0.5 ? v : rint(v * 0x1p53) * 0x1p-53; }

static void gd_draw(uint64_t seed, uint64_t i, uint64_t *ka, uint64_t *kb, uint64_t *kc)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u};
    uint32_t w[4], v[4];
    orc_philox4x32_10(ctr, key, w);
    ctr[2] = 1u;
    orc_philox4x32_10(ctr, key, v);
    *ka = gd_k53(w[0], w[1]);
    *kb = gd_k53(w[2], w[3]);
    *kc = gd_k53(v[0], v[1]);
}

static double gd_exponent_head(double hb, uint64_t i, uint64_t n_head, uint64_t ka)
{
    if (i == 0) return rint(hb) + 1.0;
    if (i == n_head - 1) return 0.0;
    return rint(gd_unif(ka) * hb);
}
static double gd_exponent_tail(double hb, uint64_t j, uint64_t n_tail)
{
    if (n_tail <= 1) return 0.0;
    return rint(hb * (1.0 - (double)j / (double)(n_tail - 1)));
}

/* Returns 0 on success.  out_exact[4]: the exact dot as 256-bit two's complement
 * limbs (value = limbs * 2^-106); out_d[0] = round(exact dot),
 * out_d[1] = round(sum |x_j y_j|) (exact then rounded), out_d[2] = condition number
 * out_d[1] / |out_d[0]|. */
ORC_API int orc_gendot(size_t n, double log2cond, uint64_t seed, size_t n_tail,
                       double *x, double *y, uint64_t *out_exact, double *out_d)
{
    if (n < 2 || n_tail < 1 || n_tail >= n) return 1;
    double hb = 0.5 * log2cond;
    if (!(hb >= 0.0) || log2cond + 4.0 + log2((double)n) > 147.0) return 2;
    size_t n_head = n - n_tail;
    acc256_t S, A;
    acc_zero(&S); acc_zero(&A);
    for (size_t i = 0; i < n_head; ++i) {
        uint64_t ka, kb, kc;
        gd_draw(seed, i, &ka, &kb, &kc);
        double e = gd_exponent_head(hb, i, n_head, ka);
        x[i] = ldexp(gd_mant(kb), (int)e);
        y[i] = ldexp(gd_mant(kc), (int)e);
        double pe, p = two_prod(x[i], y[i], &pe);
        acc_add_double(&S, p); acc_add_double(&S, pe);
        acc_add_double(&A, fabs(p)); acc_add_double(&A, p < 0 ? -pe : pe);
    }
    double r_prev = acc_round(&S);
    size_t chunk = gd_chunk_len(n_tail);
    size_t nchunks = (n_tail + chunk - 1) / chunk;
    for (size_t ch = 0; ch < nchunks; ++ch) {
        size_t j0 = ch * chunk, j1 = j0 + chunk;
        if (j1 > n_tail) j1 = n_tail;
        for (size_t j = j0; j < j1; ++j) {
            size_t i = n_head + j;
            uint64_t ka, kb, kc;
            gd_draw(seed, i, &ka, &kb, &kc);
            (void)ka;
            double e = gd_exponent_tail(hb, j, n_tail);
            double xi = ldexp(gd_mant(kb), (int)e);
            double r = ldexp(gd_mant(kc), (int)e);
            double yi;
            if (j + 1 < j1) {
                yi = gd_grid53((r - r_prev) / xi);
            } else {
                yi = gd_grid53((r - acc_round(&S)) / xi);
            }
            x[i] = xi; y[i] = yi;
            double pe, p = two_prod(xi, yi, &pe);
            acc_add_double(&S, p); acc_add_double(&S, pe);
            acc_add_double(&A, fabs(p)); acc_add_double(&A, p < 0 ? -pe : pe);
            r_prev = r;
        }
    }
    if (out_exact) memcpy(out_exact, S.w, sizeof S.w);
    if (out_d) {
        out_d[0] = acc_round(&S);
        out_d[1] = acc_round(&A);
        out_d[2] = out_d[0] != 0.0 ? out_d[1] / fabs(out_d[0]) : INFINITY;
    }
    return 0;
}

/* Exact difference (res - exact) for an exact value given as acc256 limbs; res
 * must be a multiple of 2^-106 with |res| < 2^149 (true for any double in
 * [2^-53, 2^149) and for 0), otherwise returns NaN.  Rounded once. */
ORC_API double orc_acc256_minus(const uint64_t *limbs, double res)
{
    acc256_t a;
    memcpy(a.w, limbs, sizeof a.w);
    if (res != 0.0) {
        double q = res * 0x1p106;
        if (!isfinite(q) || fabs(res) >= 0x1p149 || q != rint(q)) return NAN;
        acc_add_double(&a, -res);
    }
    return -acc_round(&a);
}
ORC_API double orc_acc256_round(const uint64_t *limbs)
{
    acc256_t a;
    memcpy(a.w, limbs, sizeof a.w);
    return acc_round(&a);
}

/* ------------------------------------------------------------------------ */
/* The GPU library's summation order, restated on the CPU                    */
/* ------------------------------------------------------------------------ */
/* Same Ogita-Rump-Oishi arithmetic as above (Dot2 step c = c + (pi + t), merge
 * s = TwoSum(s, s'), c = (c + t) + c', the association of
 * gpu_reduction_ogita_impl.cuh:498-501,532-534), evaluated in the blocked order
 * of the library's single-launch reduction.  Tests compare the GPU bit for bit
 * against this, so an indexing or ordering mistake in the kernel shows up as a
 * mismatch rather than hiding inside an error bound.
 *   op: 0 dot2, 1 dot (s = fma(x,y,s)), 2 sum2, 3 sum, 4 asum2, 5 asum,
 *       6 nrm2c (Dot2 of x with x), 7 nrm2 (s = fma(x,x,s))
 *   block, unroll, grid, vecw: the launch shape (hpb_reduce_config).
 * Per thread (b, t): tiles b, b+grid, ... of block*unroll vectors of vecw
 * elements; in a tile, vectors t + u*block for u = 0..unroll-1, elements in
 * vector order.  The partial tile (if any) belongs to block nfull % grid,
 * followed by the n % vecw trailing scalars (thread t takes element
 * nv*vecw + t).  Wave64 butterfly (lane distance 32..1), waves merged in order
 * by thread 0, block partials merged in index order (thread t takes partials
 * t, t+block, ...), then the same block reduction. */

#define HP_DEFINE(T, SUF, TWO_SUM, TWO_PROD, FMA, FABS)                                          \
static inline void hp_acc##SUF(int op, T *s, T *c, T xv, T yv)                                    \
{                                                                                                 \
    T t, pi, p;                                                                                   \
    switch (op) {                                                                                 \
    case 0: p = TWO_PROD(xv, yv, &pi); *s = TWO_SUM(*s, p, &t); *c = *c + (pi + t); break;       \
    case 1: *s = FMA(xv, yv, *s); break;                                                          \
    case 2: *s = TWO_SUM(*s, xv, &t); *c = *c + t; break;                                         \
    case 3: *s = *s + xv; break;                                                                  \
    case 4: *s = TWO_SUM(*s, FABS(xv), &t); *c = *c + t; break;                                   \
    case 5: *s = *s + FABS(xv); break;                                                            \
    case 6: p = TWO_PROD(xv, xv, &pi); *s = TWO_SUM(*s, p, &t); *c = *c + (pi + t); break;       \
    default: *s = FMA(xv, xv, *s); break;                                                         \
    }                                                                                             \
}                                                                                                 \
static inline void hp_merge##SUF(int comp, T *s, T *c, T s2, T c2)                                \
{                                                                                                 \
    if (comp) {                                                                                   \
        T t;                                                                                      \
        *s = TWO_SUM(*s, s2, &t);                                                                 \
        *c = (*c + t) + c2;                                                                       \
    } else {                                                                                      \
        *s = *s + s2;                                                                             \
    }                                                                                             \
}                                                                                                 \
static void hp_block_reduce##SUF(int comp, int block, T *s, T *c, T *os, T *oc)                   \
{                                                                                                 \
    T ns[64], nc[64];                                                                             \
    int nw = block / 64;                                                                          \
    for (int w = 0; w < nw; ++w) {                                                                \
        T *ws = s + 64 * w, *wc = c + 64 * w;                                                     \
        for (int m = 32; m >= 1; m >>= 1) {                                                       \
            for (int l = 0; l < 64; ++l) {                                                        \
                ns[l] = ws[l]; nc[l] = wc[l];                                                     \
                hp_merge##SUF(comp, &ns[l], &nc[l], ws[l ^ m], wc[l ^ m]);                        \
            }                                                                                     \
            memcpy(ws, ns, sizeof ns);                                                            \
            memcpy(wc, nc, sizeof nc);                                                            \
        }                                                                                         \
    }                                                                                             \
    T rs = s[0], rc = c[0];                                                                       \
    for (int w = 1; w < nw; ++w) hp_merge##SUF(comp, &rs, &rc, s[64 * w], c[64 * w]);            \
    *os = rs;                                                                                     \
    *oc = comp ? rc : (T)0;                                                                       \
}                                                                                                 \
ORC_API int orc_hpb_order_reduce##SUF(int op, size_t n, const T *x, const T *y, int block,       \
                                      int unroll, int grid, int vecw, T *out3)                    \
{                                                                                                 \
    if (block % 64 || block <= 0 || unroll <= 0 || grid <= 0 || vecw <= 0) return 1;            \
    int comp = hp_comp(op);                                                                       \
    int bin = (op == 0 || op == 1);                                                               \
    size_t nv = n / (size_t)vecw;                                                                 \
    size_t tile = (size_t)block * (size_t)unroll;                                                 \
    size_t nfull = nv / tile;                                                                     \
    T *s = (T *)calloc((size_t)block, sizeof(T));                                                 \
    T *c = (T *)calloc((size_t)block, sizeof(T));                                                 \
    T *ps = (T *)calloc((size_t)grid, sizeof(T));                                                 \
    T *pc = (T *)calloc((size_t)grid, sizeof(T));                                                 \
    for (int b = 0; b < grid; ++b) {                                                              \
        for (int t = 0; t < block; ++t) {                                                         \
            T ss = 0, cc = 0;                                                                     \
            for (size_t tl = (size_t)b; tl < nfull; tl += (size_t)grid) {                         \
                for (int u = 0; u < unroll; ++u) {                                                \
                    size_t vi = tl * tile + (size_t)t + (size_t)u * (size_t)block;                \
                    for (int w = 0; w < vecw; ++w) {                                              \
                        size_t i = vi * (size_t)vecw + (size_t)w;                                 \
                        hp_acc##SUF(op, &ss, &cc, x[i], bin ? y[i] : (T)0);                       \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
            if ((size_t)b == nfull % (size_t)grid) {                                              \
                for (int u = 0; u < unroll; ++u) {                                                \
                    size_t vi = nfull * tile + (size_t)t + (size_t)u * (size_t)block;             \
                    if (vi < nv)                                                                  \
                        for (int w = 0; w < vecw; ++w) {                                          \
                            size_t i = vi * (size_t)vecw + (size_t)w;                             \
                            hp_acc##SUF(op, &ss, &cc, x[i], bin ? y[i] : (T)0);                   \
                        }                                                                         \
                }                                                                                 \
                if (vecw > 1) {                                                                   \
                    size_t r0 = nv * (size_t)vecw;                                                \
                    if ((size_t)t < n - r0) hp_acc##SUF(op, &ss, &cc, x[r0 + t], bin ? y[r0 + t] : (T)0); \
                }                                                                                 \
            }                                                                                     \
            s[t] = ss;                                                                            \
            c[t] = cc;                                                                            \
        }                                                                                         \
        hp_block_reduce##SUF(comp, block, s, c, &ps[b], &pc[b]);                                  \
    }                                                                                             \
    T fs, fc;                                                                                     \
    if (grid == 1) {                                                                              \
        fs = ps[0];                                                                               \
        fc = pc[0];                                                                               \
    } else {                                                                                      \
        for (int t = 0; t < block; ++t) {                                                         \
            T ss = 0, cc = 0;                                                                     \
            for (int i = t; i < grid; i += block) hp_merge##SUF(comp, &ss, &cc, ps[i], pc[i]);    \
            s[t] = ss;                                                                            \
            c[t] = cc;                                                                            \
        }                                                                                         \
        hp_block_reduce##SUF(comp, block, s, c, &fs, &fc);                                        \
    }                                                                                             \
    out3[0] = fs;                                                                                 \
    out3[1] = fc;                                                                                 \
    out3[2] = fs + fc;                                                                            \
    free(s); free(c); free(ps); free(pc);                                                         \
    return 0;                                                                                     \
}

static inline int hp_comp(int op) { return op == 0 || op == 2 || op == 4 || op == 6; }

HP_DEFINE(double, , two_sum, two_prod, fma, fabs)
HP_DEFINE(float, _f32, two_sumf, two_prodf, fmaf, fabsf)

/* The library's overflow- and underflow-safe norms (hpb_reduce.hip), op 6 nrm2c or
 * 7 nrm2.  M = max |x_i| with NaN entries skipped (fmax).  M = 0, or M in
 * [lo, hi] (2^-400, 2^400 for FP64; 2^-30, 2^30 for FP32): the unscaled result in
 * the library's order, out4 = { s, c, s + c, sqrt(s + c) }.  Otherwise the same
 * order over x_i 2^k, with 2^k the power of two that puts M in [1/2, 1)
 * (ldexp: exact, or rounded once where an entry underflows), scaled back:
 * { s 2^-2k, c 2^-2k, (s + c) 2^-2k, sqrt(s + c) 2^-k }.  M = inf: k = 0 and
 * { s, 0, s, sqrt(s) } (inf, or NaN when x holds a NaN). */
#define NRM_DEFINE(T, SUF, FMAX, FABS, LDEXP, SQRT, LO, HI)                                        \
ORC_API int orc_hpb_order_nrm2##SUF(int op, size_t n, const T *x, int block, int unroll, int grid, \
                                    int vecw, T *out4)                                            \
{                                                                                                 \
    if (op != 6 && op != 7) return 2;                                                             \
    T m = 0;                                                                                      \
    for (size_t i = 0; i < n; ++i) m = FMAX(m, FABS(x[i]));                                       \
    T r3[3];                                                                                      \
    if (m == 0 || ((double)m >= (LO) && (double)m <= (HI))) {                                     \
        int rc = orc_hpb_order_reduce##SUF(op, n, x, x, block, unroll, grid, vecw, r3);           \
        if (rc) return rc;                                                                        \
        out4[0] = r3[0]; out4[1] = r3[1]; out4[2] = r3[2]; out4[3] = SQRT(r3[2]);                 \
        return 0;                                                                                 \
    }                                                                                             \
    int inf_mode = isinf(m), k = 0;                                                               \
    if (!inf_mode) { (void)frexp((double)m, &k); k = -k; }                                        \
    T *xs = (T *)malloc((n ? n : 1) * sizeof(T));                                                 \
    if (!xs) return 3;                                                                            \
    for (size_t i = 0; i < n; ++i) xs[i] = LDEXP(x[i], k);                                        \
    int rc = orc_hpb_order_reduce##SUF(op, n, xs, xs, block, unroll, grid, vecw, r3);             \
    free(xs);                                                                                     \
    if (rc) return rc;                                                                            \
    if (inf_mode) {                                                                               \
        out4[0] = r3[0]; out4[1] = 0; out4[2] = r3[0]; out4[3] = SQRT(r3[0]);                     \
    } else {                                                                                      \
        out4[0] = LDEXP(r3[0], -2 * k);                                                           \
        out4[1] = LDEXP(r3[1], -2 * k);                                                           \
        out4[2] = LDEXP(r3[0] + r3[1], -2 * k);                                                   \
        out4[3] = LDEXP(SQRT(r3[0] + r3[1]), -k);                                                 \
    }                                                                                             \
    return 0;                                                                                     \
}

NRM_DEFINE(double, , fmax, fabs, ldexp, sqrt, 0x1p-400, 0x1p400)
NRM_DEFINE(float, _f32, fmaxf, fabsf, ldexpf, sqrtf, 0x1p-30, 0x1p30)

/* ------------------------------------------------------------------------ */
/* Complex dot x^H y in the library's order (hpb_zdot.hip)                   */
/* ------------------------------------------------------------------------ */
/* Per element: the four real products of conj(a) b by TwoProd, a TwoSum per
 * component, error term (e1 + e2) + t, as in the reference's complex TwoProd
 * (gpu_reduction_ogita_impl_functions.cuh:126-195); accumulation as the real Dot2.
 * Tree: the same tiles, butterflies and block-order merge as orc_hpb_order_reduce
 * with vector width 1 complex element.  comp = 0: fma accumulation.
 * out6 = { s_re, c_re, s_im, c_im, fl(s_re + c_re), fl(s_im + c_im) }. */
#define ZD_DEFINE(T, SUF, TWO_SUM, TWO_PROD, FMA)                                                  \
static inline void zd_acc##SUF(int comp, T *sr, T *cr, T *si, T *ci, T ar, T ai, T br, T bi)       \
{                                                                                                 \
    if (comp) {                                                                                   \
        T e1, e2, e3, e4, t1, t2, t;                                                              \
        T p1 = TWO_PROD(ar, br, &e1), p2 = TWO_PROD(ai, bi, &e2);                                 \
        T p3 = TWO_PROD(ar, bi, &e3), p4 = TWO_PROD(-ai, br, &e4);                                \
        T pr = TWO_SUM(p1, p2, &t1), pir = (e1 + e2) + t1;                                        \
        T pi = TWO_SUM(p3, p4, &t2), pii = (e3 + e4) + t2;                                        \
        *sr = TWO_SUM(*sr, pr, &t); *cr = *cr + (pir + t);                                        \
        *si = TWO_SUM(*si, pi, &t); *ci = *ci + (pii + t);                                        \
    } else {                                                                                      \
        *sr = FMA(ar, br, *sr); *sr = FMA(ai, bi, *sr);                                           \
        *si = FMA(ar, bi, *si); *si = FMA(-ai, br, *si);                                          \
    }                                                                                             \
}                                                                                                 \
ORC_API int orc_zdot_order##SUF(int comp, size_t n, const T *x, const T *y, int block, int unroll, \
                                int grid, T *out6)                                                \
{                                                                                                 \
    if (block % 64 || block <= 0 || unroll <= 0 || grid <= 0) return 1;                          \
    size_t tile = (size_t)block * (size_t)unroll, nfull = n / tile;                               \
    T *s = (T *)calloc((size_t)block, sizeof(T)), *c = (T *)calloc((size_t)block, sizeof(T));     \
    T *s2 = (T *)calloc((size_t)block, sizeof(T)), *c2 = (T *)calloc((size_t)block, sizeof(T));   \
    T *ps = (T *)calloc((size_t)grid * 4, sizeof(T));                                             \
    for (int b = 0; b < grid; ++b) {                                                              \
        for (int t = 0; t < block; ++t) {                                                         \
            T sr = 0, cr = 0, si = 0, ci = 0;                                                     \
            for (size_t tl = (size_t)b; tl < nfull; tl += (size_t)grid)                           \
                for (int u = 0; u < unroll; ++u) {                                                \
                    size_t e = tl * tile + (size_t)t + (size_t)u * (size_t)block;                 \
                    zd_acc##SUF(comp, &sr, &cr, &si, &ci, x[2 * e], x[2 * e + 1], y[2 * e], y[2 * e + 1]); \
                }                                                                                 \
            if ((size_t)b == nfull % (size_t)grid)                                                \
                for (int u = 0; u < unroll; ++u) {                                                \
                    size_t e = nfull * tile + (size_t)t + (size_t)u * (size_t)block;              \
                    if (e < n) zd_acc##SUF(comp, &sr, &cr, &si, &ci, x[2 * e], x[2 * e + 1], y[2 * e], y[2 * e + 1]); \
                }                                                                                 \
            s[t] = sr; c[t] = cr; s2[t] = si; c2[t] = ci;                                         \
        }                                                                                         \
        hp_block_reduce##SUF(comp, block, s, c, &ps[4 * b], &ps[4 * b + 1]);                      \
        hp_block_reduce##SUF(comp, block, s2, c2, &ps[4 * b + 2], &ps[4 * b + 3]);                \
    }                                                                                             \
    for (int t = 0; t < block; ++t) {                                                             \
        T a = 0, ac = 0, bq = 0, bc = 0;                                                          \
        for (int i = t; i < grid; i += block) {                                                   \
            hp_merge##SUF(comp, &a, &ac, ps[4 * i], ps[4 * i + 1]);                               \
            hp_merge##SUF(comp, &bq, &bc, ps[4 * i + 2], ps[4 * i + 3]);                          \
        }                                                                                         \
        s[t] = a; c[t] = ac; s2[t] = bq; c2[t] = bc;                                              \
    }                                                                                             \
    hp_block_reduce##SUF(comp, block, s, c, &out6[0], &out6[1]);                                  \
    hp_block_reduce##SUF(comp, block, s2, c2, &out6[2], &out6[3]);                                \
    out6[4] = comp ? out6[0] + out6[1] : out6[0];                                                 \
    out6[5] = comp ? out6[2] + out6[3] : out6[2];                                                 \
    free(s); free(c); free(s2); free(c2); free(ps);                                               \
    return 0;                                                                                     \
}

ZD_DEFINE(double, , two_sum, two_prod, fma)
ZD_DEFINE(float, _f32, two_sumf, two_prodf, fmaf)

/* ------------------------------------------------------------------------ */
/* Elementwise operations with the rounding the library documents            */
/* ------------------------------------------------------------------------ */
/* gpu_vector_operations_impl.cuh:226-484 computes these as written and leaves
 * fma contraction to nvcc; the library fixes the contraction explicitly and the
 * oracle states the same formulas with C99 fma (correctly rounded).
 *   op  0 fill        s0
 *       1 add_mul_scalar fma(s1, x, s0)
 *       2 scal        s0*x
 *       3 copy        x
 *       4 abs         |x|
 *       5 axpy        fma(s0, x, y)
 *       6 axpby       fma(s0, x, s1*y)
 *       7 axpbypcz    fma(s2, z, fma(s0, x, s1*y))        (z = i2)
 *       8 axpbypcwpdz fma(s3, z, fma(s2, w, fma(s0, x, s1*y)))  (w = i2, z = i3)
 *       9 mulpw       (s0*x)*(s1*y)      10 mulpw_ip x*(s1*y)
 *      11 divpw       (s0*x)/(s1*y)      12 divpw_ip x/(s1*y)
 *      13 axpbyc      TwoProd(s0,x) + TwoProd(s1,y) via TwoSum, one final rounding */
#define MAP_DEFINE(T, SUF, TWO_SUM, TWO_PROD, FMA, FABS)                                          \
ORC_API void orc_map##SUF(int op, size_t n, const T *s, const T *i0, const T *i1, const T *i2,   \
                          const T *i3, T *out)                                                    \
{                                                                                                 \
    for (size_t j = 0; j < n; ++j) {                                                              \
        T x = i0 ? i0[j] : (T)0, y = i1 ? i1[j] : (T)0;                                           \
        T u = i2 ? i2[j] : (T)0, v = i3 ? i3[j] : (T)0, r, e1, e2, t, p1, p2;                     \
        switch (op) {                                                                             \
        case 0: r = s[0]; break;                                                                  \
        case 1: r = FMA(s[1], x, s[0]); break;                                                    \
        case 2: r = s[0] * x; break;                                                              \
        case 3: r = x; break;                                                                     \
        case 4: r = FABS(x); break;                                                               \
        case 5: r = FMA(s[0], x, y); break;                                                       \
        case 6: r = FMA(s[0], x, s[1] * y); break;                                                \
        case 7: r = FMA(s[2], u, FMA(s[0], x, s[1] * y)); break;                                  \
        case 8: r = FMA(s[3], v, FMA(s[2], u, FMA(s[0], x, s[1] * y))); break;                    \
        case 9: r = (s[0] * x) * (s[1] * y); break;                                               \
        case 10: r = x * (s[1] * y); break;                                                       \
        case 11: r = (s[0] * x) / (s[1] * y); break;                                              \
        case 12: r = x / (s[1] * y); break;                                                       \
        default:                                                                                  \
            p1 = TWO_PROD(s[0], x, &e1);                                                          \
            p2 = TWO_PROD(s[1], y, &e2);                                                          \
            r = TWO_SUM(p1, p2, &t);                                                              \
            r = r + ((e1 + e2) + t);                                                              \
            break;                                                                                \
        }                                                                                         \
        out[j] = r;                                                                               \
    }                                                                                             \
}
MAP_DEFINE(double, , two_sum, two_prod, fma, fabs)
MAP_DEFINE(float, _f32, two_sumf, two_prodf, fmaf, fabsf)

/* ------------------------------------------------------------------------ */
/* CSR products                                                              */
/* ------------------------------------------------------------------------ */
/* y = alpha A x + beta y (mode 0) or r = b - A x (mode 1) for zero-based CSR
 * (source/common/csr/matrix.h:91-158), evaluated in the library's row order:
 * lane l of lpr lanes takes entries beg+l, beg+l+lpr, ...; plain lanes
 * accumulate s = fma(a, x, s), compensated lanes run the Dot2 step; lanes are
 * merged by an xor butterfly (distance lpr/2 .. 1, lane 0's value kept); mode 1
 * starts lane 0 at b_i and uses -a.  The result: plain s, compensated fl(s + c);
 * mode 0 then y = v (beta 0, alpha 1), alpha v (beta 0) or fma(alpha, v, beta y). */
ORC_API void orc_csr_spmv(int rows, const int *rp, const int *col, const double *val, const double *x,
                          double alpha, double beta, const double *b, double *y, int lpr, int comp, int mode)
{
    double ls[64], lc[64], ns[64], nc[64];
    for (int r = 0; r < rows; ++r) {
        for (int l = 0; l < lpr; ++l) {
            double s = (mode == 1 && l == 0) ? b[r] : 0.0, c = 0.0;
            for (int k = rp[r] + l; k < rp[r + 1]; k += lpr) {
                double a = mode == 1 ? -val[k] : val[k];
                if (comp) {
                    double pi, t, p = two_prod(a, x[col[k]], &pi);
                    s = two_sum(s, p, &t);
                    c = c + (pi + t);
                } else {
                    s = fma(a, x[col[k]], s);
                }
            }
            ls[l] = s;
            lc[l] = c;
        }
        for (int m = lpr / 2; m >= 1; m >>= 1) {
            for (int l = 0; l < lpr; ++l) {
                ns[l] = ls[l];
                nc[l] = lc[l];
                if (comp) {
                    double t;
                    ns[l] = two_sum(ls[l], ls[l ^ m], &t);
                    nc[l] = (lc[l] + t) + lc[l ^ m];
                } else {
                    ns[l] = ls[l] + ls[l ^ m];
                }
            }
            memcpy(ls, ns, sizeof(double) * (size_t)lpr);
            memcpy(lc, nc, sizeof(double) * (size_t)lpr);
        }
        double v = comp ? ls[0] + lc[0] : ls[0];
        if (mode == 1) y[r] = v;
        else if (beta == 0.0) y[r] = (alpha == 1.0) ? v : alpha * v;
        else y[r] = fma(alpha, v, beta * y[r]);
    }
}
