// hpb_capi.cpp -- handle lifetime, memory, status text, tuning and timers.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "hpb_internal.hpp"

namespace hpb {

static thread_local std::string g_last_error;

void set_error(const char* file, int line, const std::string& msg)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, ":%d: ", line);
    const char* base = std::strrchr(file, '/');
    g_last_error = std::string(base ? base + 1 : file) + buf + msg;
}

int hip_status(hipError_t e, const char* file, int line, const char* what)
{
    set_error(file, line, std::string(what) + " failed: " + hipGetErrorString(e));
    return (int)e;
}

}  // namespace hpb

extern "C" {

const char* hpb_last_error(void) { return hpb::g_last_error.c_str(); }

const char* hpb_version(void) { return "hpb 0.1 gfx950"; }

static void default_tuning(hpb_handle_s* h)
{
    // Shapes measured fastest on MI355X at N = 2^27 (FP64): two-input reductions
    // 512 threads x 2 16-byte loads per input, 3 blocks per 4 CUs, nontemporal
    // loads (7088 GB/s against 6979 with 4 loads, interleaved in one process; the
    // autotuner's choice on another box as well); one-input reductions 4 loads
    // and one block per CU; elementwise
    // 256 threads x 1 vector, 2 blocks per CU, nontemporal loads and stores (axpy
    // 6349 GB/s against 5447 GB/s with 2 vectors and cached accesses).  More
    // resident blocks lower the achieved bandwidth of these streams.
    const int cu = h->num_cu > 0 ? h->num_cu : 256;
    for (int d = 0; d < 2; ++d) {
        h->tune_red[d] = {512, 2, (3 * cu) / 4 > 0 ? (3 * cu) / 4 : 1, HPB_TUNE_NONTEMPORAL};
        h->tune_red1[d] = {512, 4, cu, HPB_TUNE_NONTEMPORAL};
        h->tune_map[d] = {256, 1, 2 * cu, HPB_TUNE_NONTEMPORAL};
        // multi-dot: 8 columns per chunk, 8 blocks per CU over all chunks, nontemporal
        // loads (k = 16, n = 2^24: 6159 GB/s against 5433 with 4 blocks per CU, cached)
        h->tune_multi[d] = {256, 8, 8 * cu, HPB_TUNE_NONTEMPORAL};
    }
}

int hpb_create(hpb_handle_t* out, void* stream)
{
    HPB_REQUIRE(out != nullptr, HPB_ERR_INVALID, "hpb_create: null handle pointer");
    *out = nullptr;
    hpb_handle_s* h = new hpb_handle_s();
    hipError_t e;
    auto fail = [&](hipError_t err, int line, const char* what) {
        hpb_destroy(h);
        return hpb::hip_status(err, __FILE__, line, what);
    };
    if ((e = hipGetDevice(&h->device)) != hipSuccess) return fail(e, __LINE__, "hipGetDevice");
    int ncu = 0;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device)) != hipSuccess)
        return fail(e, __LINE__, "hipDeviceGetAttribute");
    h->num_cu = ncu;
    if (stream) {
        h->stream = (hipStream_t)stream;
    } else {
        if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
            return fail(e, __LINE__, "hipStreamCreate");
        h->own_stream = true;
    }
    h->part_bytes = (size_t)HPB_MAX_GRID * HPB_MAX_MULTIDOT * 2 * sizeof(double);
    if ((e = hipMalloc(&h->part, h->part_bytes)) != hipSuccess) return fail(e, __LINE__, "hipMalloc(part)");
    if ((e = hipMalloc((void**)&h->ticket, HPB_TICKETS * sizeof(unsigned))) != hipSuccess)
        return fail(e, __LINE__, "hipMalloc(ticket)");
    if ((e = hipMemset(h->ticket, 0, HPB_TICKETS * sizeof(unsigned))) != hipSuccess)
        return fail(e, __LINE__, "hipMemset(ticket)");
    if ((e = hipMalloc(&h->dres, HPB_DRES_BYTES)) != hipSuccess) return fail(e, __LINE__, "hipMalloc(dres)");
    if ((e = hipHostMalloc(&h->hres, HPB_DRES_BYTES, hipHostMallocDefault)) != hipSuccess)
        return fail(e, __LINE__, "hipHostMalloc(hres)");
    if ((e = hipMalloc((void**)&h->dflag, 64)) != hipSuccess) return fail(e, __LINE__, "hipMalloc(flag)");
    if ((e = hipMalloc(&h->nrmfix, sizeof(hpb::nrm_fix))) != hipSuccess) return fail(e, __LINE__, "hipMalloc(nrmfix)");
    if ((e = hipMemset(h->nrmfix, 0, sizeof(hpb::nrm_fix))) != hipSuccess) return fail(e, __LINE__, "hipMemset(nrmfix)");
    for (int i = 0; i < 2; ++i)
        if ((e = hipEventCreate(&h->ev[i])) != hipSuccess) return fail(e, __LINE__, "hipEventCreate");
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, __LINE__, "hipDeviceSynchronize");
    default_tuning(h);
    if (int rc = hpb::noalloc_init()) {
        hpb_destroy(h);
        return rc;
    }
    *out = h;
    return HPB_OK;
}

int hpb_destroy(hpb_handle_t h)
{
    if (!h) return HPB_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->part) (void)hipFree(h->part);
    if (h->ticket) (void)hipFree(h->ticket);
    if (h->dres) (void)hipFree(h->dres);
    if (h->hres) (void)hipHostFree(h->hres);
    if (h->dflag) (void)hipFree(h->dflag);
    if (h->nrmfix) (void)hipFree(h->nrmfix);
    if (h->hstage) (void)hipHostFree(h->hstage);
    for (int i = 0; i < 2; ++i)
        if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return HPB_OK;
}

int hpb_get_stream(hpb_handle_t h, void** stream)
{
    HPB_REQUIRE(h && stream, HPB_ERR_INVALID, "hpb_get_stream: null argument");
    *stream = (void*)h->stream;
    return HPB_OK;
}

int hpb_synchronize(hpb_handle_t h)
{
    HPB_REQUIRE(h, HPB_ERR_INVALID, "hpb_synchronize: null handle");
    HPB_HIP(hipStreamSynchronize(h->stream));
    return HPB_OK;
}

int hpb_device_info(hpb_handle_t h, int* num_cu, int* device)
{
    HPB_REQUIRE(h, HPB_ERR_INVALID, "hpb_device_info: null handle");
    if (num_cu) *num_cu = h->num_cu;
    if (device) *device = h->device;
    return HPB_OK;
}

int hpb_malloc(void** p, size_t bytes)
{
    HPB_REQUIRE(p, HPB_ERR_INVALID, "hpb_malloc: null pointer");
    *p = nullptr;
    if (bytes == 0) return HPB_OK;
    HPB_HIP(hipMalloc(p, bytes));
    return HPB_OK;
}

int hpb_free(void* p)
{
    if (p) HPB_HIP(hipFree(p));
    return HPB_OK;
}

int hpb_memcpy_h2d(hpb_handle_t h, void* dst, const void* src, size_t bytes)
{
    HPB_REQUIRE(h && (bytes == 0 || (dst && src)), HPB_ERR_INVALID, "hpb_memcpy_h2d: null argument");
    if (bytes == 0) return HPB_OK;
    HPB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    return HPB_OK;
}

int hpb_memcpy_d2h(hpb_handle_t h, void* dst, const void* src, size_t bytes)
{
    HPB_REQUIRE(h && (bytes == 0 || (dst && src)), HPB_ERR_INVALID, "hpb_memcpy_d2h: null argument");
    if (bytes == 0) return HPB_OK;
    HPB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    return HPB_OK;
}

int hpb_memcpy_d2d(hpb_handle_t h, void* dst, const void* src, size_t bytes)
{
    HPB_REQUIRE(h && (bytes == 0 || (dst && src)), HPB_ERR_INVALID, "hpb_memcpy_d2d: null argument");
    if (bytes == 0) return HPB_OK;
    HPB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream));
    return HPB_OK;
}

int hpb_set_tuning(hpb_handle_t h, int kind, int dtype, int block, int unroll, int grid_max, int flags)
{
    HPB_REQUIRE(h, HPB_ERR_INVALID, "hpb_set_tuning: null handle");
    HPB_REQUIRE(kind >= 0 && kind <= 3, HPB_ERR_INVALID, "hpb_set_tuning: kind must be 0, 1, 2 or 3");
    HPB_REQUIRE(dtype == 0 || dtype == 1, HPB_ERR_INVALID, "hpb_set_tuning: dtype must be 0 or 1");
    HPB_REQUIRE(block == 256 || block == 512, HPB_ERR_UNSUPPORTED, "hpb_set_tuning: block must be 256 or 512");
    HPB_REQUIRE(unroll == 1 || unroll == 2 || unroll == 4 || unroll == 8, HPB_ERR_UNSUPPORTED,
                "hpb_set_tuning: unroll must be 1, 2, 4 or 8");
    HPB_REQUIRE(grid_max >= 1 && grid_max <= HPB_MAX_GRID, HPB_ERR_INVALID, "hpb_set_tuning: grid_max out of range");
    HPB_REQUIRE((flags & ~HPB_TUNE_NONTEMPORAL) == 0, HPB_ERR_UNSUPPORTED, "hpb_set_tuning: unknown flags");
    HPB_REQUIRE(kind != 3 || (block == 256 && (unroll == 4 || unroll == 8)), HPB_ERR_UNSUPPORTED,
                "hpb_set_tuning: the multi-dot takes block 256 and 4 or 8 columns per chunk");
    hpb_tuning_t t = {block, unroll, grid_max, flags};
    if (kind == 0) h->tune_red[dtype] = t;
    else if (kind == 2) h->tune_red1[dtype] = t;
    else if (kind == 3) h->tune_multi[dtype] = t;
    else h->tune_map[dtype] = t;
    return HPB_OK;
}

int hpb_get_tuning(hpb_handle_t h, int kind, int dtype, int* block, int* unroll, int* grid_max, int* flags)
{
    HPB_REQUIRE(h && block && unroll && grid_max, HPB_ERR_INVALID, "hpb_get_tuning: null argument");
    HPB_REQUIRE((kind >= 0 && kind <= 3) && (dtype == 0 || dtype == 1), HPB_ERR_INVALID, "hpb_get_tuning: bad kind/dtype");
    const hpb_tuning_t& t = kind == 0 ? h->tune_red[dtype]
                            : kind == 2 ? h->tune_red1[dtype]
                            : kind == 3 ? h->tune_multi[dtype]
                                        : h->tune_map[dtype];
    *block = t.block;
    *unroll = t.unroll;
    *grid_max = t.grid_max;
    if (flags) *flags = t.flags;
    return HPB_OK;
}

int hpb_timer_start(hpb_handle_t h)
{
    HPB_REQUIRE(h, HPB_ERR_INVALID, "hpb_timer_start: null handle");
    HPB_HIP(hipEventRecord(h->ev[0], h->stream));
    return HPB_OK;
}

int hpb_timer_stop(hpb_handle_t h, float* ms)
{
    HPB_REQUIRE(h && ms, HPB_ERR_INVALID, "hpb_timer_stop: null argument");
    HPB_HIP(hipEventRecord(h->ev[1], h->stream));
    HPB_HIP(hipEventSynchronize(h->ev[1]));
    HPB_HIP(hipEventElapsedTime(ms, h->ev[0], h->ev[1]));
    return HPB_OK;
}

}  // extern "C"
