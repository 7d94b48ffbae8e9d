// hpb_internal.hpp -- library-private handle, status plumbing and launch tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string>

#include "../../include/hpb.h"

struct hpb_tuning_t {
    int block;      // threads per block (multiple of 64)
    int unroll;     // 16-byte vectors per thread per tile
    int grid_max;   // cap on the number of blocks (grid-stride beyond it)
    int flags;      // HPB_TUNE_* bits
};

struct hpb_handle_s {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    int num_cu = 0;
    // Reduction workspace: per-block (s, c) partials of the current launch, the
    // last-block ticket and a device result area.  Ops on one handle run in
    // stream order, so one set suffices.
    void* part = nullptr;
    size_t part_bytes = 0;
    unsigned* ticket = nullptr;     // HPB_TICKETS counters, each returns to 0 after use
    void* dres = nullptr;           // device results, HPB_DRES_BYTES
    void* hres = nullptr;           // pinned host mirror of dres
    int* dflag = nullptr;           // validity flag
    void* nrmfix = nullptr;         // hpb::nrm_fix record of the scaled norms (zero at rest)
    void* hstage = nullptr;         // pinned host staging of the host-side analyses (grown on demand)
    size_t hstage_bytes = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hpb_tuning_t tune_red[2];       // two-input reductions (dot), [0] f64, [1] f32
    hpb_tuning_t tune_red1[2];      // one-input reductions (sum, asum, nrm2)
    hpb_tuning_t tune_map[2];       // elementwise kernels
    hpb_tuning_t tune_multi[2];     // fused multi-dot: unroll = columns per chunk (4 or 8), grid_max = total blocks
};

#define HPB_TICKETS 4
#define HPB_DRES_BYTES (64 * 4 * 8 * 4)
#define HPB_MAX_GRID 4096
#define HPB_MAX_MULTIDOT 64

namespace hpb {

// State passed from a norm launch to its rescaling launch (hpb_reduce.hip): need
// is set when max |x_i| fell outside the range in which the unscaled sum of
// squares can neither overflow nor lose its error terms to underflow; the
// rescaling launch clears it.
struct nrm_fix {
    int need;
    int pad;
    double amax;
};

void set_error(const char* file, int line, const std::string& msg);
// looks up the current device's address of the no-allocation scratch array
// (hpb_multidot_f64_noalloc); called by hpb_create
int noalloc_init();
int hip_status(hipError_t e, const char* file, int line, const char* what);

// Asynchronous launchers used by the device-resident solvers.  Results stay in
// device memory; skip (device int, may be null) makes the launch a no-op while
// *skip != 0, so a captured solver graph can keep replaying after convergence.
template <typename T>
int dot2_async(hpb_handle_t h, size_t n, const T* x, const T* y, T* dev_out4, bool comp, const int* skip);
template <typename T>
int nrm2_async(hpb_handle_t h, size_t n, const T* x, T* dev_out4, bool comp, const int* skip);
template <typename T>
int multidot_async(hpb_handle_t h, int k, const T* const* V, size_t n, const T* w, bool comp, T* dev_out4k,
                   const int* skip);
template <typename T>
int multiaxpy_devcoef(hpb_handle_t h, int k, const T* dev_a, const T* const* V, size_t n, T* y, bool comp,
                      const int* skip);
// Device-resident GMRES (CGS2) / BiCGStab(L <= 8): vectors T, scalars FP64.
// cache (optional) keeps the solver's buffers and captured graph between calls
// with the same operands; release it with dsolve_release.
struct dsolver_base {
    virtual ~dsolver_base() {}
};
template <typename T>
int dsolve(hpb_handle_t h, int rows, int64_t nnz, const int* rp, const int* ci, const T* val, const T* b, T* x,
           const hpb_solver_params& p, hpb_solver_result* res, double* history, int history_cap,
           dsolver_base** cache = nullptr);
void dsolve_release(dsolver_base* cache);
// FP32 BiCGStab(L) iterations with reliable updates of an FP64 solution
int dsolve_reliable(hpb_handle_t h, int rows, int64_t nnz, const int* rp, const int* ci, const float* val32,
                    double mscale, const double* val64, const double* b64, double* x64, float* r32, float* x32,
                    const hpb_solver_params& inner, double rel_tol, hpb_solver_result* res, int* updates);
// power-of-two scale s with max |x_i| s in [0.5, 1) (1 for an empty or zero x)
int absmax_pow2_scale(hpb_handle_t h, size_t n, const double* x, double* scale);
// b == null: y = a A x with a = *dev_alpha (1 when dev_alpha is null); else y = b - A x
template <typename T>
int csr_apply(hpb_handle_t h, int rows, int64_t nnz, const int* rp, const int* col, const T* val, const T* x,
              const T* b, T* y, bool comp, const int* skip, const T* dev_alpha = nullptr);

}  // namespace hpb

#define HPB_HIP(call)                                                            \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess) return hpb::hip_status(_e, __FILE__, __LINE__, #call); \
    } while (0)

#define HPB_REQUIRE(cond, code, msg)                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            hpb::set_error(__FILE__, __LINE__, msg);                             \
            return (code);                                                       \
        }                                                                        \
    } while (0)

// Entry points that launch kernels first drop a stale error left by an earlier,
// unrelated runtime call, so that the check after the launch reports the launch.
#define HPB_CLEAR_STALE() (void)hipGetLastError()

#define HPB_LAUNCH_CHECK()                                                       \
    do {                                                                         \
        hipError_t _e = hipGetLastError();                                       \
        if (_e != hipSuccess) return hpb::hip_status(_e, __FILE__, __LINE__, "kernel launch"); \
    } while (0)
