// Measurement helpers: a copy kernel that calibrates the achievable HBM
// bandwidth, the memory-pattern kernels that bound the fused step, event-
// timed stepping for throughput and roofline figures, and the fused kernel's
// tuning knobs.
#include <string>
#include <type_traits>

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

template <class S, int VEC, bool NT>
__device__ inline void pat_load(const S* p, S (&v)[VEC]) {
    typedef S SV __attribute__((ext_vector_type(VEC)));
    if constexpr (VEC == 1) {
        v[0] = NT ? __builtin_nontemporal_load(p) : *p;
    } else {
        const SV w = NT ? __builtin_nontemporal_load(reinterpret_cast<const SV*>(p)) : *reinterpret_cast<const SV*>(p);
#pragma unroll
        for (int j = 0; j < VEC; ++j) v[j] = w[j];
    }
}

template <class S, int VEC>
__device__ inline void pat_store(S* p, const S (&v)[VEC]) {
    typedef S SV __attribute__((ext_vector_type(VEC)));
    if constexpr (VEC == 1) {
        __builtin_nontemporal_store(v[0], p);
    } else {
        SV w;
#pragma unroll
        for (int j = 0; j < VEC; ++j) w[j] = v[j];
        __builtin_nontemporal_store(w, reinterpret_cast<SV*>(p));
    }
}

// The loads and stores of one fused pull step without its arithmetic, in the
// fused layout (population k of cell c at k * ps + c, x fastest), VEC
// consecutive cells per lane, raw bits (S of the storage width):
//   PAT 0  every population read and written at its own cell: a Q-stream
//          copy, non-temporal both ways;
//   PAT 1  the row kernel's pulls: population k read from row (y - c_y,
//          z - c_z) at the same x (the x shift is a register move), the
//          cell's flag byte, non-temporal pulls and stores (its pattern
//          without interior solids);
//   PAT 2  PAT 1 with the row kernel's pattern around interior solids:
//          ordinary pulls, the cell's own reverse slots of every population
//          and a 4-byte neighbour mask per cell, selected by the mask.
// Every index is in range: c0 + VEC <= ncell (ncell is a multiple of X, X of
// VEC), rows wrap into [0, Y) and [0, Z), k * ps + c < Q * ps.
template <int Q, class S, int VEC, int PAT>
__global__ __launch_bounds__(256) void k_pattern(const S* __restrict__ A, S* __restrict__ B,
                                                 const uint8_t* __restrict__ flags,
                                                 const uint32_t* __restrict__ nbmask, int X, int Y, int Z,
                                                 uint64_t ps, uint64_t ncell) {
    const uint64_t c0 = ((uint64_t)blockIdx.x * 256u + threadIdx.x) * VEC;
    if (c0 >= ncell) return;
    const uint64_t row = c0 / (uint64_t)X;
    const int x0 = (int)(c0 - row * (uint64_t)X);
    const int y = (int)(row % (uint64_t)Y), z = (int)(row / (uint64_t)Y);
    S v[Q][VEC];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        uint64_t src = c0;
        if (PAT > 0) {
            int yy = y - Lat<Q>::cy(k), zz = z - Lat<Q>::cz(k);
            yy = yy < 0 ? yy + Y : (yy >= Y ? yy - Y : yy);
            zz = zz < 0 ? zz + Z : (zz >= Z ? zz - Z : zz);
            src = ((uint64_t)zz * Y + yy) * X + x0;
        }
        pat_load<S, VEC, PAT != 2>(A + k * ps + src, v[k]);
    }
    if (PAT == 2) {
        S own[Q][VEC];
#pragma unroll
        for (int k = 1; k < Q; ++k) pat_load<S, VEC, false>(A + rev<Q>(k) * ps + c0, own[k]);
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
            const uint32_t m = nbmask[c0 + j];
#pragma unroll
            for (int k = 1; k < Q; ++k)
                if ((m >> rev<Q>(k)) & 1u) v[k][j] = own[k][j];
        }
    } else if (PAT == 1) {
#pragma unroll
        for (int j = 0; j < VEC; ++j)
            if (flags[c0 + j])
#pragma unroll
                for (int k = 0; k < Q; ++k) v[k][j] = 0;
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) pat_store<S, VEC>(B + k * ps + c0, v[k]);
}

template <int Q, class S, int VEC, int PAT>
void launch_pattern(const void* A, void* B, const uint8_t* flags, const uint32_t* mask, int X, int Y, int Z,
                    uint64_t ps, uint64_t ncell, hipStream_t st) {
    const unsigned nb = (unsigned)((ncell / VEC + 255) / 256);
    hipLaunchKernelGGL((k_pattern<Q, S, VEC, PAT>), dim3(nb), dim3(256), 0, st, (const S*)A, (S*)B, flags, mask, X,
                       Y, Z, ps, ncell);
}

// The launch for (q, storage bytes, cells per lane, pattern), or false.
bool pattern_launch(int q, int sbytes, int vec, int pat, const void* A, void* B, const uint8_t* flags,
                    const uint32_t* mask, int X, int Y, int Z, uint64_t ps, uint64_t ncell, hipStream_t st) {
    bool ok = false;
    auto by_pat = [&](auto qc, auto sc, auto vc) {
        constexpr int Qv = decltype(qc)::value, V = decltype(vc)::value;
        using S = decltype(sc);
        if (pat == 0) launch_pattern<Qv, S, V, 0>(A, B, flags, mask, X, Y, Z, ps, ncell, st);
        if (pat == 1) launch_pattern<Qv, S, V, 1>(A, B, flags, mask, X, Y, Z, ps, ncell, st);
        if (pat == 2) launch_pattern<Qv, S, V, 2>(A, B, flags, mask, X, Y, Z, ps, ncell, st);
        ok = pat >= 0 && pat <= 2;
    };
    auto by_vec = [&](auto qc, auto sc) {
        if (vec == 1) by_pat(qc, sc, std::integral_constant<int, 1>());
        if (vec == 2) by_pat(qc, sc, std::integral_constant<int, 2>());
        if (vec == 4) by_pat(qc, sc, std::integral_constant<int, 4>());
    };
    auto by_size = [&](auto qc) {
        if (sbytes == 2) by_vec(qc, uint16_t());
        if (sbytes == 4) by_vec(qc, uint32_t());
        if (sbytes == 8) by_vec(qc, uint64_t());
    };
    if (q == 19) by_size(std::integral_constant<int, 19>());
    if (q == 27) by_size(std::integral_constant<int, 27>());
    return ok;
}

}  // namespace
}  // namespace plb

using namespace plb;

extern "C" {

int plb_set_tuning(plb_state* s, int key, int value) {
    if (!s) return fail(PLB_ERR_INVALID_ARGUMENT, "null state");
    if (s->fused_deferred) {  // a step the staged calls deferred runs with the old settings
        PLB_HIP_TRY(hipSetDevice(s->device));
        const int rc = fused_flush_deferred(s);
        if (rc) return rc;
    }
    switch (key) {
        case PLB_TUNE_BLOCK_X:
            if (value != 0 && value != 64 && value != 128 && value != 256)
                return fail(PLB_ERR_INVALID_ARGUMENT, "block width must be 0, 64, 128 or 256");
            s->tune_tx = value;
            break;
        case PLB_TUNE_NT_LOAD: s->tune_nt_load = value != 0; break;
        case PLB_TUNE_NT_STORE: s->tune_nt_store = value != 0; break;
        case PLB_TUNE_SCHEME:
            if (value != PLB_SCHEME_AUTO && value != PLB_SCHEME_PUSH && value != PLB_SCHEME_PULL &&
                value != PLB_SCHEME_AA)
                return fail(PLB_ERR_INVALID_ARGUMENT, "unknown streaming scheme");
            // the scheme follows the number of population arrays the state holds
            if (s->aa && value != PLB_SCHEME_AA && value != PLB_SCHEME_AUTO)
                return fail(PLB_ERR_LOGIC, "a one-array state (PLB_MODE_FUSED_AA) steps with PLB_SCHEME_AA only");
            if (!s->aa && value == PLB_SCHEME_AA)
                return fail(PLB_ERR_LOGIC, "PLB_SCHEME_AA needs a state created with PLB_MODE_FUSED_AA");
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
        case PLB_TUNE_TWO_STEP:
            if (value < -1 || value > 1) return fail(PLB_ERR_INVALID_ARGUMENT, "two-step setting must be -1, 0 or 1");
            s->tune_two = value;
            break;
        case PLB_TUNE_TWO_TILE_X:
            if (value != 0 && value != 32 && value != 64)
                return fail(PLB_ERR_INVALID_ARGUMENT, "two-step tile width must be 0, 32 or 64");
            s->tune_two_bx = value;
            break;
        case PLB_TUNE_TWO_ZCHUNK:
            if (value < 0) return fail(PLB_ERR_INVALID_ARGUMENT, "z-chunk length must be 0 (automatic) or positive");
            s->tune_two_zc = value;
            break;
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
    if (s->aa)
        *scheme = PLB_SCHEME_AA;
    else if (!fused_uses_pull(s))
        *scheme = PLB_SCHEME_PUSH;
    else
        *scheme = fused_two_eligible(s) ? PLB_SCHEME_PULL_TWO : PLB_SCHEME_PULL;
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

int plb_bench_pattern(int device, int q, int x, int y, int z, int sbytes, int vec, int pattern, int iters,
                      double* gbps, double* ms) {
    if (!gbps || iters <= 0 || (q != 19 && q != 27) || (sbytes != 2 && sbytes != 4 && sbytes != 8) ||
        (vec != 1 && vec != 2 && vec != 4) || pattern < 0 || pattern > 2 || x <= 0 || y <= 0 || z <= 0 ||
        x % vec != 0)
        return fail(PLB_ERR_INVALID_ARGUMENT, "bad pattern benchmark arguments");
    const uint64_t ncell = (uint64_t)x * y * z;
    if (ncell >= (1ull << 32) / 4) return fail(PLB_ERR_INVALID_ARGUMENT, "pattern benchmark domain too large");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PLB_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(PLB_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    PLB_HIP_TRY(hipSetDevice(device));
    // the fused layout's population stride (plb_state_create's default pad)
    const uint64_t ps = ((ncell + 63) / 64) * 64 + 2048 + 64;
    const size_t pop_bytes = (size_t)q * ps * sbytes;
    void *A = nullptr, *B = nullptr, *flags = nullptr, *mask = nullptr;
    hipError_t e = hipMalloc(&A, pop_bytes);
    if (e == hipSuccess) e = hipMalloc(&B, pop_bytes);
    if (e == hipSuccess) e = hipMalloc(&flags, ncell);
    if (e == hipSuccess) e = hipMalloc(&mask, ncell * 4);
    hipStream_t st = nullptr;
    if (e == hipSuccess) e = hipStreamCreate(&st);
    hipEvent_t a = nullptr, b = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipMemsetAsync(A, 1, pop_bytes, st);
    if (e == hipSuccess) e = hipMemsetAsync(B, 0, pop_bytes, st);
    if (e == hipSuccess) e = hipMemsetAsync(flags, 0, ncell, st);
    if (e == hipSuccess) e = hipMemsetAsync(mask, 0, ncell * 4, st);
    float t = 0.f;
    bool launched = true;
    if (e == hipSuccess) {
        for (int w = 0; w < 2 && launched; ++w)
            launched = pattern_launch(q, sbytes, vec, pattern, A, B, (const uint8_t*)flags, (const uint32_t*)mask, x,
                                      y, z, ps, ncell, st);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(a, st);
        for (int i = 0; i < iters && e == hipSuccess && launched; ++i) {
            pattern_launch(q, sbytes, vec, pattern, A, B, (const uint8_t*)flags, (const uint32_t*)mask, x, y, z, ps,
                           ncell, st);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    }
    if (st) (void)hipStreamSynchronize(st);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (st) (void)hipStreamDestroy(st);
    for (void* p : {A, B, flags, mask})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? PLB_ERR_OOM : PLB_ERR_HIP, hipGetErrorString(e));
    const double per = (double)t / iters;
    *gbps = (double)ncell * (2.0 * q * sbytes + 1.0) / (per * 1e-3) / 1e9;
    if (ms) *ms = per;
    return PLB_OK;
}

}  // extern "C"
