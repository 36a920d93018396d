// Measurement helpers: a copy kernel that calibrates the achievable HBM
// bandwidth, event-timed stepping for throughput and roofline figures, and
// the fused kernel's tuning knobs.
#include <string>

#include "state.hpp"

namespace plb {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// One 16-byte element per thread, non-temporal both ways (the fastest plain
// copy measured on MI355X): every byte read once and written once.
__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace
}  // namespace plb

using namespace plb;

extern "C" {

int plb_set_tuning(plb_state* s, int key, int value) {
    if (!s) return fail(PLB_ERR_INVALID_ARGUMENT, "null state");
    switch (key) {
        case PLB_TUNE_BLOCK_X:
            if (value != 0 && value != 64 && value != 128 && value != 256)
                return fail(PLB_ERR_INVALID_ARGUMENT, "block width must be 0, 64, 128 or 256");
            s->tune_tx = value;
            break;
        case PLB_TUNE_NT_LOAD: s->tune_nt_load = value != 0; break;
        case PLB_TUNE_NT_STORE: s->tune_nt_store = value != 0; break;
        case PLB_TUNE_SCHEME:
            if (value != PLB_SCHEME_AUTO && value != PLB_SCHEME_PUSH && value != PLB_SCHEME_PULL)
                return fail(PLB_ERR_INVALID_ARGUMENT, "unknown streaming scheme");
            s->tune_scheme = value;
            break;
        case PLB_TUNE_TILE_ORDER:
            if (value < 0 || value > 64 || (value > 0 && (value & (value - 1))))
                return fail(PLB_ERR_INVALID_ARGUMENT, "tile order must be 0 or a power of two up to 64");
            s->tune_order = value;
            break;
        case PLB_TUNE_XCD:
            if (value < -1 || value > 2) return fail(PLB_ERR_INVALID_ARGUMENT, "tile mapping must be -1, 0, 1 or 2");
            s->tune_xcd = value;
            break;
        case PLB_TUNE_INDEX64: s->tune_index64 = value != 0; break;
        case PLB_TUNE_LITERAL_FUSE: s->tune_lit_fuse = value != 0; break;
        case PLB_TUNE_LITERAL_AVX2: s->tune_lit_avx2 = value != 0; break;
        case PLB_TUNE_VECTOR:
            if (value != -1 && value != 0 && value != 1 && value != 2 && value != 4)
                return fail(PLB_ERR_INVALID_ARGUMENT, "row pull width must be -1, 0, 1, 2 or 4 cells per lane");
            s->tune_vector = value;
            break;
        default: return fail(PLB_ERR_INVALID_ARGUMENT, "unknown tuning key");
    }
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
        s->graph_steps = 0;
    }
    return PLB_OK;
}

int plb_step_scheme(const plb_state* s, int* scheme) {
    if (!s || !scheme) return fail(PLB_ERR_INVALID_ARGUMENT, "null argument");
    if (s->mode != PLB_MODE_FUSED) return fail(PLB_ERR_LOGIC, "literal mode has no fused scheme");
    *scheme = fused_uses_pull(s) ? PLB_SCHEME_PULL : PLB_SCHEME_PUSH;
    return PLB_OK;
}

int plb_time_steps(plb_state* s, int nsteps, int steps_per_graph, double* ms) {
    if (!s || !ms) return fail(PLB_ERR_INVALID_ARGUMENT, "null argument");
    PLB_HIP_TRY(hipSetDevice(s->device));
    hipEvent_t a, b;
    PLB_HIP_TRY(hipEventCreate(&a));
    PLB_HIP_TRY(hipEventCreate(&b));
    int rc = PLB_OK;
    hipError_t e = hipEventRecord(a, s->stream);
    if (e == hipSuccess) {
        rc = steps_per_graph > 0 ? plb_step_graph(s, nsteps, steps_per_graph) : plb_step(s, nsteps);
        e = hipEventRecord(b, s->stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(PLB_ERR_HIP, hipGetErrorString(e));
    *ms = t;
    return PLB_OK;
}

int plb_bench_copy(int device, size_t bytes, int iters, double* gbps, double* ms_per_copy) {
    if (!gbps || iters <= 0 || bytes < 16) return fail(PLB_ERR_INVALID_ARGUMENT, "bad copy benchmark arguments");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PLB_ERR_NO_DEVICE, "no HIP device");
    PLB_HIP_TRY(hipSetDevice(device));
    const uint64_t elems = bytes / 16;
    void *src = nullptr, *dst = nullptr;
    PLB_HIP_TRY(hipMalloc(&src, elems * 16));
    hipError_t e = hipMalloc(&dst, elems * 16);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return fail(PLB_ERR_OOM, hipGetErrorString(e));
    }
    hipStream_t st;
    e = hipStreamCreate(&st);
    hipEvent_t a, b;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipMemsetAsync(src, 1, elems * 16, st);
    if (e == hipSuccess) e = hipMemsetAsync(dst, 0, elems * 16, st);
    const unsigned nb = blocks_for(elems, 256);
    float t = 0.f;
    if (e == hipSuccess) {
        for (int w = 0; w < 2; ++w)
            hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, st, (const u32x4*)src, (u32x4*)dst, elems);
        e = hipEventRecord(a, st);
        for (int i = 0; i < iters && e == hipSuccess; ++i) {
            hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, st, (const u32x4*)src, (u32x4*)dst, elems);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    }
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipStreamDestroy(st);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return fail(PLB_ERR_HIP, hipGetErrorString(e));
    const double per = (double)t / iters;
    *gbps = 2.0 * (double)(elems * 16) / (per * 1e-3) / 1e9;
    if (ms_per_copy) *ms_per_copy = per;
    return PLB_OK;
}

}  // extern "C"
