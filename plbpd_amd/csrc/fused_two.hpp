// Two pull steps per pass over HBM (temporal blocking) for the fused mode:
// the numbers of two k_pull launches (gather from x - c_k with halfway
// bounce-back, BGK with the velocity-shift force, the moving-wall term where
// a wall face is hit; CollisionSRT.cpp:153-176, Streaming.cpp:63-74,
// Boundary.cpp:189-221), bit for bit, from one read of the populations of
// step t and one write of those of step t + 2.
//
// A workgroup owns a tile of 64 x-cells by BY y-rows and sweeps it along z
// over a chunk of planes.  At each plane zi it computes step t + 1 on the
// tile grown by one cell in x and y (66 x (BY + 2) cells, one per thread),
// gathering step t from HBM, and keeps the result in LDS, in a ring of three
// planes.  It then computes step t + 2 on the tile at plane zi - 1, gathering
// from the ring (planes zi - 2, zi - 1, zi), and writes it to the other
// population array.  The grown border is computed by both neighbouring tiles,
// so workgroups never wait for each other: every workgroup reads only A and
// writes only its own cells of B.  The step t + 1 values pass through the
// storage type on their way into LDS, exactly as a single step stores them,
// so the result equals two single steps bit for bit.  The loads of plane
// zi + 1 are issued before the step t + 2 work of plane zi - 1, whose
// arithmetic hides their latency.  Each lattice instantiates these templates
// in its own file (fused_two_q19.hip, fused_two_q27.hip).
#pragma once

#include <algorithm>

#include "fused_kernels.hpp"

namespace plb {

namespace {

// Step t loads without interior solids: population k from x - c_k, or the
// cell's own slot rev(k) where x - c_k is a wall layer (known from the
// coordinates, so no load waits on a flag).
template <int Q, int K, int V, class I>
__device__ inline void pull_gather_walls(typename Prec<V>::S (&fv)[Q], const typename Prec<V>::S* __restrict__ A,
                                         I c, I ps, const Seams<I>& s) {
    if constexpr (K < Q) {
        const I moved = (I)K * ps + (c + pull_offset<Q, K>(s));
        const I own = (I)rev_k<Q>(K) * ps + c;
        fv[K] = *elem(A, pull_wall<Q, K>(s) ? own : moved);
        pull_gather_walls<Q, K + 1, V>(fv, A, c, ps, s);
    }
}

// Step t + 1 values for the step t + 2 cell at ring index li: population k
// from li - c_k in the plane zo - c_k,z (ring slots sm1, s0, sp for planes
// zo - 1, zo, zo + 1), or the cell's own slot rev(k) in plane zo where
// x - c_k is not fluid.
template <int Q, int K, int V, int NB, int NI, int RX, class I>
__device__ inline void ring_gather(typename Prec<V>::T (&f)[Q], const typename Prec<V>::S* ring, I li, uint32_t sp,
                                   uint32_t s0, uint32_t sm1, uint32_t m, const Seams<I>& s) {
    if constexpr (K < Q) {
        constexpr int R = rev_k<Q>(K);
        constexpr int cx = Lat<Q>::cx(K), cy = Lat<Q>::cy(K), cz = Lat<Q>::cz(K);
        const bool solid = NB ? ((m >> R) & 1u) != 0 : pull_wall<Q, K>(s);
        const uint32_t src = cz > 0 ? sm1 : (cz < 0 ? sp : s0);
        const I moved = (src * (uint32_t)Q + (uint32_t)K) * (uint32_t)NI + li - (I)(cx + cy * RX);
        const I own = (s0 * (uint32_t)Q + (uint32_t)R) * (uint32_t)NI + li;
        f[K] = Prec<V>::load_raw(ring[solid ? own : moved]);
        ring_gather<Q, K + 1, V, NB, NI, RX>(f, ring, li, sp, s0, sm1, m, s);
    }
}

// Workgroup barrier that orders LDS only.  The global stores of a step t + 2
// cell are read by no other thread of the launch, so the barrier need not
// wait for them -- nor for the step t loads of the next plane, which are in
// flight across it (__syncthreads would wait for every outstanding global
// access).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int Q, int V, int NB, bool MOV, bool NTS, int BY>
__global__ __launch_bounds__(two_threads(BY)) void k_pull2(const typename Prec<V>::S* __restrict__ A,
                                                           typename Prec<V>::S* __restrict__ B,
                                                           const uint8_t* __restrict__ flags,
                                                           const uint32_t* __restrict__ nbmask,
                                                           FusedArgs<typename Prec<V>::T> a, TwoGeom g) {
    using P = Prec<V>;
    using S = typename P::S;
    using T = typename P::T;
    using I = uint32_t;
    constexpr int RX = 66, NI = two_cells(BY);
    __shared__ S ring[3 * Q * NI];
    __shared__ uint32_t mring[3 * NI];  // the step t + 1 cells' masks or flags

    const uint32_t tid = threadIdx.x;
    uint32_t tx, ty, tz;
    if (g.xcd) {
        const uint32_t j = blockIdx.x & 7u, m = blockIdx.x >> 3, sy = g.nty >> 3;
        tx = m % g.ntx;
        const uint32_t r = m / g.ntx;
        ty = j * sy + r % sy;
        tz = r / sy;
    } else {
        tx = blockIdx.x % g.ntx;
        const uint32_t r = blockIdx.x / g.ntx;
        ty = r % g.nty;
        tz = r / g.nty;
    }
    const uint32_t X = a.X, Y = a.Y, Z = a.Z;
    const I XY = X * Y, ps = (I)a.pstride;
    const uint32_t x0 = tx * 64u, y0 = ty * (uint32_t)BY, z0 = tz * g.zc;
    const uint32_t nplanes = std::min(g.zc, Z - z0) + 2u;

    // step t + 1: this thread's cell of the grown tile (wrapped coordinates)
    const bool in_i = tid < (uint32_t)NI;
    const uint32_t ix = tid % RX, iy = tid / RX;
    const uint32_t gx = (x0 + ix + X - 1u) % X, gy = (y0 + iy + Y - 1u) % Y;
    // step t + 2: this thread's cell of the tile
    const uint32_t ox = tid & 63u, oy = tid >> 6;
    const bool in_o = oy < (uint32_t)BY && y0 + oy < Y;
    const uint32_t px = x0 + ox, py = y0 + oy;
    const I li = (oy + 1u) * RX + ox + 1u;

    // loads of step t for the plane of iteration n (prefetched one iteration ahead)
    S fv[Q], ow[Q];
    uint32_t mi = 0;  // neighbour mask (NB) or flag of the step t + 1 cell
    auto fetch = [&](uint32_t n) {
        const uint32_t zi = (z0 + n + Z - 1u) % Z;
        if (in_i) {
            const I c = zi * XY + gy * X + gx;
            const Seams<I> sm = seams<I>(gx, gy, zi, a);
            if (NB) {
                mi = nbmask[c];
                pull_gather_both<Q, 0, V>(fv, ow, A, c, ps, sm);
            } else {
                mi = flags[c];
                pull_gather_walls<Q, 0, V>(fv, A, c, ps, sm);
            }
        }
    };

    fetch(0);
    for (uint32_t n = 0; n < nplanes; ++n) {
        const uint32_t slot = n % 3u;
        // step t + 1 at plane zi into the ring
        if (in_i) {
            mring[slot * NI + tid] = mi;
            const bool self_solid = NB ? (mi & 1u) != 0 : mi != 0;
            if (!self_solid) {
                T f[Q];
#pragma unroll
                for (int k = 0; k < Q; ++k)
                    f[k] = P::load_raw(NB && k > 0 && ((mi >> rev_k<Q>(k)) & 1u) ? ow[k] : fv[k]);
                const T rho = bgk<Q, P>(f, a, true);
                if (MOV && a.bc.moving) {
                    const uint32_t zi = (z0 + n + Z - 1u) % Z;
                    const I c = zi * XY + gy * X + gx;
                    wall_terms<Q, 1, P>(f, rho, flags, c, seams<I>(gx, gy, zi, a), a.bc);
                }
#pragma unroll
                for (int k = 0; k < Q; ++k) ring[(slot * Q + k) * NI + tid] = P::store_raw(f[k]);
            }
        }
        if (n + 1u < nplanes) fetch(n + 1u);
        lds_barrier();
        // step t + 2 at plane zo = zi - 1 from the ring
        if (n >= 2u && in_o) {
            const uint32_t zo = z0 + n - 2u;
            const I c = zo * XY + py * X + px;
            const uint32_t sp = slot, s0 = (n + 2u) % 3u, sm1 = (n + 1u) % 3u;
            // this cell was a step t + 1 cell of the previous plane
            const uint32_t mcur = mring[s0 * NI + li];
            const bool self_solid = NB ? (mcur & 1u) != 0 : mcur != 0;
            if (!__all(self_solid)) {
                const Seams<I> sm = seams<I>(px, py, zo, a);
                T f[Q];
                ring_gather<Q, 0, V, NB, NI, RX>(f, ring, li, sp, s0, sm1, mcur, sm);
                if (!self_solid) {
                    const T rho = bgk<Q, P>(f, a, true);
                    if (MOV && a.bc.moving) wall_terms<Q, 1, P>(f, rho, flags, c, sm, a.bc);
                }
                own_store<Q, V, NTS>(f, B, c, ps, self_solid);
            }
        }
        lds_barrier();
    }
}

}  // namespace

template <int Q, int V>
bool launch_pull2(const FusedArgs<typename Prec<V>::T>& a, const TwoGeom& g, bool nb, bool mov, bool nts,
                  const typename Prec<V>::S* A, typename Prec<V>::S* B, const uint8_t* flags,
                  const uint32_t* nbmask, hipStream_t stream) {
    constexpr int BY = two_rows(Q, (int)sizeof(typename Prec<V>::S));
    if constexpr (BY == 0) {
        return false;
    } else {
        const dim3 grid(g.ntx * g.nty * g.nzc), block(two_threads(BY));
#define PLB_TWO(NBv, MVv, NTv) \
    hipLaunchKernelGGL((k_pull2<Q, V, NBv, MVv, NTv, BY>), grid, block, 0, stream, A, B, flags, nbmask, a, g)
        if (nb && mov && nts) PLB_TWO(1, true, true);
        else if (nb && mov) PLB_TWO(1, true, false);
        else if (nb && nts) PLB_TWO(1, false, true);
        else if (nb) PLB_TWO(1, false, false);
        else if (mov && nts) PLB_TWO(0, true, true);
        else if (mov) PLB_TWO(0, true, false);
        else if (nts) PLB_TWO(0, false, true);
        else PLB_TWO(0, false, false);
#undef PLB_TWO
        return true;
    }
}

#define PLB_TWO_INST(Q, V)                                                                                  \
    template bool launch_pull2<Q, V>(const FusedArgs<typename Prec<V>::T>&, const TwoGeom&, bool, bool, bool, \
                                     const typename Prec<V>::S*, typename Prec<V>::S*, const uint8_t*,       \
                                     const uint32_t*, hipStream_t);
#define PLB_TWO_INST_Q(Q) \
    PLB_TWO_INST(Q, kF64) PLB_TWO_INST(Q, kF32) PLB_TWO_INST(Q, kF64A32) PLB_TWO_INST(Q, kF32S) PLB_TWO_INST(Q, kF16S)

}  // namespace plb
