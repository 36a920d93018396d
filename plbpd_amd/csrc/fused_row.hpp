// Row pull step of the fused mode: the same step as k_pull in fused.hip
// (gather from x - c_k with halfway bounce-back, BGK with the velocity-shift
// force, whole-line stores of the cell's own slots; CollisionSRT.cpp:153-176,
// Streaming.cpp:63-74, Boundary.cpp:189-221), with every population load
// aligned.
//
// An x-row of X = 64 * VEC * WPR cells is cut into WPR wavefront segments;
// lane l of segment s owns cells x0 .. x0 + VEC - 1, x0 = (64 s + l) VEC.
// A block takes 4 rows of one segment, 2 rows of two segments, or one row of
// three or more segments (row_block_waves).  Each population row
// (y - c_y, z - c_z) is loaded unshifted, so every cache line is read by
// exactly one wavefront (which is what lets these loads be non-temporal where
// there are no interior solids).  The x shift is done in registers: inside a
// lane by moving values between its VEC cells, across lanes by a rotate, and
// across segments through LDS.  The segments of a row wrap around, which is
// the periodic x seam; a walled x axis has solid wall cells there, whose
// values bounce-back replaces.  Each lattice instantiates these templates in
// its own file (fused_row_q9.hip, fused_row_q19.hip, fused_row_q27.hip).
#pragma once

#include <type_traits>

#include "fused_kernels.hpp"

namespace plb {

namespace {

// Index of direction k among the directions with c_x != 0.
template <int Q>
__host__ __device__ constexpr int xslot(int k) {
    int n = 0;
    for (int j = 0; j < k; ++j) n += Lat<Q>::cx(j) != 0;
    return n;
}
template <int Q>
__host__ __device__ constexpr int nxdirs() { return xslot<Q>(Q); }

template <class S, int VEC, bool NT>
__device__ inline void vload(const S* p, S (&v)[VEC]) {
    typedef S SV __attribute__((ext_vector_type(VEC)));
    if constexpr (VEC == 1) {
        v[0] = NT ? __builtin_nontemporal_load(p) : *p;
    } else {
        const SV* q = reinterpret_cast<const SV*>(p);
        const SV w = NT ? __builtin_nontemporal_load(q) : *q;
#pragma unroll
        for (int j = 0; j < VEC; ++j) v[j] = w[j];
    }
}
template <class S, int VEC, bool NT>
__device__ inline void vstore(S* p, const S (&v)[VEC]) {
    typedef S SV __attribute__((ext_vector_type(VEC)));
    if constexpr (VEC == 1) {
        if (NT)
            __builtin_nontemporal_store(v[0], p);
        else
            *p = v[0];
    } else {
        SV w;
#pragma unroll
        for (int j = 0; j < VEC; ++j) w[j] = v[j];
        if (NT)
            __builtin_nontemporal_store(w, reinterpret_cast<SV*>(p));
        else
            *reinterpret_cast<SV*>(p) = w;
    }
}

// Moving-wall terms subtracted when the slots that point into a wall are
// written (the pull-scheme bounce-back reads them back from the same cell;
// the same values as the push step, see wall_terms in fused.hip).
template <int Q, class P>
__device__ inline void row_wall_terms(typename P::T (&f)[Q], typename P::T rho, const uint8_t* __restrict__ flags,
                                      uint32_t x, uint32_t y, uint32_t z, const FusedArgs<typename P::T>& a) {
    using T = typename P::T;
    const uint32_t xo[3] = {x == 0 ? a.X - 1 : x - 1, x, x + 1 == a.X ? 0 : x + 1};
    const uint32_t yo[3] = {y == 0 ? a.Y - 1 : y - 1, y, y + 1 == a.Y ? 0 : y + 1};
    const uint32_t zo[3] = {z == 0 ? a.Z - 1 : z - 1, z, z + 1 == a.Z ? 0 : z + 1};
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
        // the neighbour x + c_k on a wall face (interior solids take no term)
        const bool wall = (cx > 0 && a.bc.wall_hi[0] && x + 2 == a.X) || (cx < 0 && a.bc.wall_lo[0] && x == 1) ||
                          (cy > 0 && a.bc.wall_hi[1] && y + 2 == a.Y) || (cy < 0 && a.bc.wall_lo[1] && y == 1) ||
                          (cz > 0 && a.bc.wall_hi[2] && z + 2 == a.Z) || (cz < 0 && a.bc.wall_lo[2] && z == 1);
        if (wall) {
            const uint64_t d = ((uint64_t)zo[1 + cz] * a.Y + yo[1 + cy]) * a.X + xo[1 + cx];
            const int fl = flags[d];
            const T cub = (T)cx * a.bc.vel[fl][0] + (T)cy * a.bc.vel[fl][1] + (T)cz * a.bc.vel[fl][2];
            const T term = (T)2 * w_scaled<Q, P>(k) * rho * cub * (T)3;
            f[k] = f[k] - term;
        }
    }
}

template <int Q, int V, int VEC, int WPR, int NB, bool NTS>
__global__ __launch_bounds__(64 * row_block_waves(WPR)) void k_pull_row(const typename Prec<V>::S* __restrict__ A,
                                                     typename Prec<V>::S* __restrict__ B,
                                                     const uint8_t* __restrict__ flags,
                                                     const uint32_t* __restrict__ nbmask,
                                                     FusedArgs<typename Prec<V>::T> a) {
    using P = Prec<V>;
    using T = typename P::T;
    using S = typename P::S;
    constexpr uint32_t kWaves = row_block_waves(WPR);
    constexpr uint32_t kRows = kWaves / WPR;
    constexpr int NX = nxdirs<Q>();
    // segment edges: first value (for c_x < 0) and last value (c_x > 0)
    __shared__ T edge_first[kWaves][NX];
    __shared__ T edge_last[kWaves][NX];

    const uint32_t t = tile_of_block(a);
    uint32_t ty, tz;
    tile_rows(a, t, ty, tz);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t seg = wave % WPR, rib = wave / WPR;
    const uint32_t y = ty * kRows + rib, z = tz;
    const uint32_t x0 = (seg * 64u + lane) * VEC;
    const uint64_t XY = (uint64_t)a.X * a.Y;
    const uint64_t c0 = (uint64_t)z * XY + (uint64_t)y * a.X + x0;

    // Every load is issued before anything waits on memory: the pulled rows,
    // around interior solids the cell's own reverse slots (the bounce-back
    // values; a row that another wavefront pulls from, so these loads stay
    // ordinary and are made whether or not a lane needs them), then the
    // cell flags or neighbour masks.  Every index is in range for an
    // existing row.  Only then does a wavefront with no fluid cell drop out;
    // it still takes part in the edge exchange (one barrier per block).  On
    // MI355X, waiting for the masks first, and loading only the reverse rows
    // some lane needs, made the 2-byte step 35% slower and the 8-byte step 2%.
    // Without interior solids every line is read once and the pulls are
    // non-temporal.
    constexpr bool NTL = NB == 0;
    T f[Q][VEC];
    S own[Q][VEC];
    uint32_t m[VEC];
    bool ss[VEC];
    bool any_fluid = false;
    const bool row_ok = y < a.Y;
    if (row_ok) {
        const uint32_t yo[3] = {y == 0 ? a.Y - 1 : y - 1, y, y + 1 == a.Y ? 0 : y + 1};
        const uint32_t zo[3] = {z == 0 ? a.Z - 1 : z - 1, z, z + 1 == a.Z ? 0 : z + 1};
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const int cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
            const uint64_t srow = (uint64_t)zo[1 - cz] * XY + (uint64_t)yo[1 - cy] * a.X;
            S v[VEC];
            vload<S, VEC, NTL>(A + (uint64_t)k * a.pstride + srow + x0, v);
#pragma unroll
            for (int j = 0; j < VEC; ++j) f[k][j] = P::load_raw(v[j]);
        }
        if (NB) {
#pragma unroll
            for (int k = 1; k < Q; ++k) vload<S, VEC, false>(A + (uint64_t)rev_k<Q>(k) * a.pstride + c0, own[k]);
        }
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
            if (NB) {
                m[j] = nbmask[c0 + j];
                ss[j] = m[j] & 1u;
            } else {
                m[j] = 0;
                ss[j] = flags[c0 + j] != 0;
            }
            any_fluid |= !ss[j];
        }
    }
    const bool active = row_ok && __any(any_fluid);
    // A wavefront whose own cells are all solid still publishes its edge
    // values: a diagonal population it loaded comes from another row, where
    // the source cell can be fluid, and the neighbouring segments pull it.
    if (row_ok && (lane == 0 || lane == 63)) {
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            if (Lat<Q>::cx(k) < 0 && lane == 0) edge_first[wave][xslot<Q>(k)] = f[k][0];
            if (Lat<Q>::cx(k) > 0 && lane == 63) edge_last[wave][xslot<Q>(k)] = f[k][VEC - 1];
        }
    }
    if (WPR > 1) __syncthreads();
    if (!active) return;

    const uint32_t wprev = rib * WPR + (seg + WPR - 1) % WPR;
    const uint32_t wnext = rib * WPR + (seg + 1) % WPR;
    const int lane_prev = (int)((lane + 63u) & 63u), lane_next = (int)((lane + 1u) & 63u);
    const bool wyr[3] = {a.bc.wall_lo[1] && y == 1, false, a.bc.wall_hi[1] && y + 2 == a.Y};
    const bool wzr[3] = {a.bc.wall_lo[2] && z == 1, false, a.bc.wall_hi[2] && z + 2 == a.Z};
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
        T g[VEC];
        if (cx > 0) {
            // source x - 1: the previous cell of this lane, of the previous
            // lane, or of the previous segment's last lane
            T from = __shfl(f[k][VEC - 1], lane_prev);
            if (WPR > 1 && lane == 0) from = edge_last[wprev][xslot<Q>(k)];
            g[0] = from;
#pragma unroll
            for (int j = 1; j < VEC; ++j) g[j] = f[k][j - 1];
        } else if (cx < 0) {
            T from = __shfl(f[k][0], lane_next);
            if (WPR > 1 && lane == 63) from = edge_first[wnext][xslot<Q>(k)];
#pragma unroll
            for (int j = 0; j < VEC - 1; ++j) g[j] = f[k][j + 1];
            g[VEC - 1] = from;
        } else {
#pragma unroll
            for (int j = 0; j < VEC; ++j) g[j] = f[k][j];
        }
        // bounce-back where the source cell x - c_k is solid
        bool bb[VEC];
        bool any_bb = false;
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
            if (NB) {
                bb[j] = ((m[j] >> rev_k<Q>(k)) & 1u) != 0;
            } else {
                const uint32_t x = x0 + j;
                bb[j] = wyr[1 - cy] || wzr[1 - cz] || (cx > 0 && a.bc.wall_lo[0] && x == 1) ||
                        (cx < 0 && a.bc.wall_hi[0] && x + 2 == a.X);
            }
            any_bb |= bb[j] && !ss[j];
        }
        if (NB) {
            if (k > 0) {
#pragma unroll
                for (int j = 0; j < VEC; ++j)
                    if (bb[j]) g[j] = P::load_raw(own[k][j]);
            }
        } else if (__any(any_bb)) {
            S ov[VEC];
            vload<S, VEC, false>(A + (uint64_t)rev_k<Q>(k) * a.pstride + c0, ov);
#pragma unroll
            for (int j = 0; j < VEC; ++j)
                if (bb[j]) g[j] = P::load_raw(ov[j]);
        }
#pragma unroll
        for (int j = 0; j < VEC; ++j) f[k][j] = g[j];
    }
    // collide, cell by cell
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
        T fj[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) fj[k] = f[k][j];
        const T rho = bgk<Q, P>(fj, a, true);
        if (a.bc.moving && !ss[j]) row_wall_terms<Q, P>(fj, rho, flags, x0 + j, y, z, a);
#pragma unroll
        for (int k = 0; k < Q; ++k) f[k][j] = fj[k];
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        S o[VEC];
#pragma unroll
        for (int j = 0; j < VEC; ++j) o[j] = P::store_raw(ss[j] ? (T)0 : f[k][j]);
        vstore<S, VEC, NTS>(B + (uint64_t)k * a.pstride + c0, o);
    }
}

template <int Q, int V, int VEC, int WPR>
void launch_one(const FusedArgs<typename Prec<V>::T>& a, bool nts, bool nb, const typename Prec<V>::S* A,
                typename Prec<V>::S* B, const uint8_t* flags, const uint32_t* nbmask, unsigned nblocks,
                hipStream_t stream) {
    const dim3 grid(nblocks), block(64 * row_block_waves(WPR));
#define PLB_ROW(NBv, NTv)                                                                                   \
    hipLaunchKernelGGL((k_pull_row<Q, V, VEC, WPR, NBv, NTv>), grid, block, 0, stream, A, B, flags, nbmask, a)
    if (nb) {
        if (nts) PLB_ROW(1, true);
        else PLB_ROW(1, false);
    } else {
        if (nts) PLB_ROW(0, true);
        else PLB_ROW(0, false);
    }
#undef PLB_ROW
}

}  // namespace

// The row kernel of shape kRowShapes[I] if it matches (vec, X), else the next
// shape: the dispatch is generated from the table, so the two cannot differ.
template <int Q, int V, size_t I = 0>
bool launch_row_shape(const FusedArgs<typename Prec<V>::T>& a, int vec, bool nts, bool nb,
                      const typename Prec<V>::S* A, typename Prec<V>::S* B, const uint8_t* flags,
                      const uint32_t* nbmask, unsigned nblocks, hipStream_t stream) {
    if constexpr (I < sizeof(kRowShapes) / sizeof(kRowShapes[0])) {
        constexpr RowShape r = kRowShapes[I];
        if (vec == r.vec && a.X == 64u * r.vec * r.wpr) {
            launch_one<Q, V, r.vec, r.wpr>(a, nts, nb, A, B, flags, nbmask, nblocks, stream);
            return true;
        }
        return launch_row_shape<Q, V, I + 1>(a, vec, nts, nb, A, B, flags, nbmask, nblocks, stream);
    } else {
        return false;
    }
}

template <int Q, int V>
bool launch_pull_row(const FusedArgs<typename Prec<V>::T>& a, int vec, bool nts, bool nb,
                     const typename Prec<V>::S* A, typename Prec<V>::S* B, const uint8_t* flags,
                     const uint32_t* nbmask, unsigned nblocks, hipStream_t stream) {
    return launch_row_shape<Q, V>(a, vec, nts, nb, A, B, flags, nbmask, nblocks, stream);
}

#define PLB_ROW_INST(Q, V)                                                                                  \
    template bool launch_pull_row<Q, V>(const FusedArgs<typename Prec<V>::T>&, int, bool, bool,            \
                                        const typename Prec<V>::S*, typename Prec<V>::S*, const uint8_t*,  \
                                        const uint32_t*, unsigned, hipStream_t);
#define PLB_ROW_INST_Q(Q) \
    PLB_ROW_INST(Q, kF64) PLB_ROW_INST(Q, kF32) PLB_ROW_INST(Q, kF64A32) PLB_ROW_INST(Q, kF32S) PLB_ROW_INST(Q, kF16S)

}  // namespace plb
