// Two pull steps per pass over HBM (temporal blocking) for the fused mode:
// the numbers of two k_pull launches (gather from x - c_k with halfway
// bounce-back, BGK with the velocity-shift force, the moving-wall term where
// a wall face is hit; CollisionSRT.cpp:153-176, Streaming.cpp:63-74,
// Boundary.cpp:189-221), bit for bit, from one read of the populations of
// step t and one write of those of step t + 2.
//
// A workgroup owns a tile of BX x-cells by BY y-rows and sweeps it along z
// over a chunk of planes.  At each plane zi it computes step t + 1 on the
// tile grown by one cell in x and y ((BX + 2) x (BY + 2) cells, one per
// thread), gathering step t from HBM, and keeps the result in an LDS ring
// (ZRing: the directions with c_z = +1 for three planes, the others for
// two).  It then computes step t + 2 on the tile at plane zi - 1, gathering
// from the ring (planes zi - 2, zi - 1, zi), and writes it to the other
// population array.  The grown border is computed by both neighbouring tiles,
// so workgroups never wait for each other: every workgroup reads only A and
// writes only its own cells of B.  The step t + 1 values pass through the
// storage type on their way into LDS, exactly as a single step stores them,
// so the result equals two single steps bit for bit.  The loads of plane
// zi + 1 are issued before the step t + 2 work of plane zi - 1, whose
// arithmetic hides their latency; where two_split() holds, those of the
// c_z = +1 directions a plane earlier still, so that the bounce-back reads of
// a cache line fall in the iteration of its pulls.  Each lattice instantiates
// these templates in its own file (fused_two_q19.hip, fused_two_q27.hip).
#pragma once

#include <algorithm>
#include <type_traits>

#include "fused_kernels.hpp"

namespace plb {

namespace {

// Element (population k, cell offset i) of a population array.  W = false:
// one 32-bit element offset k * pstride + i from the array base (every byte
// offset of the array fits 32 bits); W = true: the population's base pointer
// (a 64-bit scalar) plus the 32-bit cell offset (every byte offset within one
// population fits 32 bits), for arrays beyond 4 GiB.
template <bool W, class S>
__device__ inline S* pop_elem(S* A, uint32_t k, uint32_t ps, uint32_t i) {
    if constexpr (W)
        return elem(A + (uint64_t)k * ps, i);
    else
        return elem(A, k * ps + i);
}

// Direction groups of the step t gathers: kAll, or, where the c_z = +1
// directions are issued a plane ahead of the others, kUp (c_z = +1) and
// kFlat (c_z = 0 and -1).
enum : int { kAll = 0, kUp = 1, kFlat = 2 };
template <int Q, int K, int G>
__device__ constexpr bool in_group() {
    return G == kAll || (Lat<Q>::cz(K) > 0) == (G == kUp);
}

// Step t loads without interior solids, for the directions of group G:
// population k from x - c_k, or the cell's own slot rev(k) where x - c_k is
// a wall layer (known from the coordinates, so no load waits on a flag).
template <int Q, int K, int V, bool W, int G>
__device__ inline void gather_walls(typename Prec<V>::S (&fv)[Q], const typename Prec<V>::S* __restrict__ A,
                                    uint32_t c, uint32_t ps, const Seams<uint32_t>& s) {
    if constexpr (K < Q && !in_group<Q, K, G>()) {
        gather_walls<Q, K + 1, V, W, G>(fv, A, c, ps, s);
    } else if constexpr (K < Q) {
        const bool wall = pull_wall<Q, K>(s);
        if constexpr (W) {
            fv[K] = *pop_elem<W>(A, wall ? (uint32_t)rev_k<Q>(K) : (uint32_t)K, ps,
                                 wall ? c : c + pull_offset<Q, K>(s));
        } else {
            const uint32_t moved = (uint32_t)K * ps + (c + pull_offset<Q, K>(s));
            const uint32_t own = (uint32_t)rev_k<Q>(K) * ps + c;
            fv[K] = *elem(A, wall ? own : moved);
        }
        gather_walls<Q, K + 1, V, W, G>(fv, A, c, ps, s);
    }
}

// Step t loads around interior solids, for the directions of group G:
// population k from x - c_k, or the cell's own slot rev(k) where the
// neighbour mask m says x - c_k is not fluid (one load per population).
template <int Q, int K, int V, bool W, int G>
__device__ inline void gather_mask(typename Prec<V>::S (&fv)[Q], const typename Prec<V>::S* __restrict__ A,
                                   uint32_t c, uint32_t ps, uint32_t m, const Seams<uint32_t>& s) {
    if constexpr (K < Q && !in_group<Q, K, G>()) {
        gather_mask<Q, K + 1, V, W, G>(fv, A, c, ps, m, s);
    } else if constexpr (K < Q) {
        constexpr uint32_t R = (uint32_t)rev_k<Q>(K);
        const bool solid = ((m >> R) & 1u) != 0;
        if constexpr (W) {
            fv[K] = *pop_elem<W>(A, solid ? R : (uint32_t)K, ps, solid ? c : c + pull_offset<Q, K>(s));
        } else {
            // byte offsets (c, ps and the seams of s in bytes): the population
            // parts are uniform (a selected population number times pstride
            // would be a per-lane multiply) and the cell parts shared between
            // directions; the mask bit, sign-extended to all ones, selects
            // bit by bit
            using S = typename Prec<V>::S;
            (void)solid;
            const uint32_t moved = (uint32_t)K * ps + (c + pull_offset<Q, K>(s));
            const uint32_t own = R * ps + c;
            const uint32_t sel = (uint32_t)__builtin_amdgcn_sbfe((int)m, (int)R, 1);
            fv[K] = *reinterpret_cast<const S*>(reinterpret_cast<const char*>(A) + (moved ^ ((moved ^ own) & sel)));
        }
        gather_mask<Q, K + 1, V, W, G>(fv, A, c, ps, m, s);
    }
}

// Step t + 2 of a cell into its own slots (zeros for a cell that is not fluid).
// Each population's base pointer is a scalar and the cell's byte offset is
// shared by all of them.  (The compiler turns these into one 64-bit address
// add per store, where per-population 32-bit offsets from the array base
// kept about 30 loop-carried offsets in vector registers; forcing the base
// into a scalar operand, as the cell kernel does, compiled this kernel to
// more instructions.)
template <int Q, int V, bool NT>
__device__ inline void store_own(const typename Prec<V>::T (&f)[Q], typename Prec<V>::S* __restrict__ B, uint32_t c,
                                 uint32_t ps, bool zero) {
    using P = Prec<V>;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const typename P::S v = P::store_raw(zero ? (typename P::T)0 : f[k]);
        typename P::S* p = pop_elem<true>(B, (uint32_t)k, ps, c);
        if (NT)
            __builtin_nontemporal_store(v, p);
        else
            *p = v;
    }
}

// Ring plane of direction k in its group's slot: element offset of cell 0.
template <int Q, int K, int NI>
__device__ inline uint32_t ring_at(uint32_t slot) {
    using Z = ZRing<Q>;
    constexpr int g = Z::grp(K);
    return ((uint32_t)Z::base(g) + slot * (uint32_t)Z::size(g) + (uint32_t)Z::idx(K)) * (uint32_t)NI;
}

// Step t + 1 of a cell into the ring: direction k to its group's slot for
// this plane (a3 = plane mod 3 for the c_z = +1 group, a2 = plane mod 2 for
// the others).
template <int Q, int K, int V, int NI>
__device__ inline void ring_store(const typename Prec<V>::T (&f)[Q], typename Prec<V>::S* ring, uint32_t cell,
                                  uint32_t a3, uint32_t a2) {
    if constexpr (K < Q) {
        ring[ring_at<Q, K, NI>(ZRing<Q>::grp(K) == 0 ? a3 : a2) + cell] = Prec<V>::store_raw(f[K]);
        ring_store<Q, K + 1, V, NI>(f, ring, cell, a3, a2);
    }
}

// Step t + 1 values for the step t + 2 cell at ring index li (plane zo):
// population k from li - c_k in plane zo - c_k,z, or the cell's own slot
// rev(k) in plane zo where x - c_k is not fluid.  Slots: pa = the c_z = +1
// group's slot of plane zo - 1, oa = its slot of plane zo, b0 = the other
// groups' slot of plane zo, c1 = the c_z = -1 group's slot of plane zo + 1.
template <int Q, int K, int V, int NB, int NI, int RX>
__device__ inline void ring_gather(typename Prec<V>::T (&f)[Q], const typename Prec<V>::S* ring, uint32_t li,
                                   uint32_t pa, uint32_t oa, uint32_t b0, uint32_t c1, uint32_t m,
                                   const Seams<uint32_t>& s) {
    if constexpr (K < Q) {
        using Z = ZRing<Q>;
        constexpr int R = rev_k<Q>(K);
        constexpr int cx = Lat<Q>::cx(K), cy = Lat<Q>::cy(K);
        constexpr int g = Z::grp(K), gr = Z::grp(R);
        const uint32_t moved = ring_at<Q, K, NI>(g == 0 ? pa : (g == 1 ? b0 : c1)) + li - (uint32_t)(cx + cy * RX);
        const uint32_t own = ring_at<Q, R, NI>(gr == 0 ? oa : b0) + li;
        if constexpr (NB != 0) {
            // own - moved is uniform (li cancels) and, like every ring index,
            // far inside the signed 24-bit range: one 24-bit multiply-add by
            // the mask bit instead of a compare and a select
            const int d = __mul24((int)((m >> R) & 1u), (int)(own - moved));
            f[K] = Prec<V>::load_raw(ring[moved + (uint32_t)d]);
        } else {
            f[K] = Prec<V>::load_raw(ring[pull_wall<Q, K>(s) ? own : moved]);
        }
        ring_gather<Q, K + 1, V, NB, NI, RX>(f, ring, li, pa, oa, b0, c1, m, s);
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

template <int Q, int V, int NB, bool MOV, int BX, int BY, bool W, bool SPLIT>
__global__ __launch_bounds__(two_threads(BX, BY), 1) void k_pull2(const typename Prec<V>::S* __restrict__ A,
                                                           typename Prec<V>::S* __restrict__ B,
                                                           const uint8_t* __restrict__ flags,
                                                           const uint32_t* __restrict__ nbmask,
                                                           FusedArgs<typename Prec<V>::T> a, TwoGeom g) {
    using P = Prec<V>;
    using S = typename P::S;
    using T = typename P::T;
    using I = uint32_t;
    constexpr int RX = BX + 2, NI = two_cells(BX, BY);
    __shared__ S ring[ZRing<Q>::planes() * NI];
    __shared__ uint32_t mring[2 * NI];  // the step t + 1 cells' masks or flags, two planes

    const uint32_t tid = threadIdx.x;
    uint32_t tx, ty, tz;
    if (g.xcd) {
        // workgroup i runs on XCD i mod 8, which takes pairs [j kx, (j + 1) kx)
        const uint32_t j = blockIdx.x & 7u, m = blockIdx.x >> 3;
        tx = m % g.ntx;
        const uint32_t pr = j * g.kx + m / g.ntx;
        if (pr >= g.nty * g.nzc) return;  // past the last pair (the whole workgroup)
        ty = pr % g.nty;
        tz = pr / g.nty;
    } else {
        tx = blockIdx.x % g.ntx;
        const uint32_t r = blockIdx.x / g.ntx;
        ty = r % g.nty;
        tz = r / g.nty;
    }
    const uint32_t X = a.X, Y = a.Y, Z = a.Z;
    const I XY = X * Y, ps = (I)a.pstride;
    const uint32_t x0 = tx * (uint32_t)BX, y0 = ty * (uint32_t)BY, z0 = tz * g.zc;
    const uint32_t nplanes = std::min(g.zc, Z - z0) + 2u;

    // step t + 1: this thread's cell of the grown tile (wrapped coordinates)
    const bool in_i = tid < (uint32_t)NI;
    const uint32_t ix = tid % RX, iy = tid / RX;
    const uint32_t gx = (x0 + ix + X - 1u) % X, gy = (y0 + iy + Y - 1u) % Y;
    // step t + 2: this thread's cell of the tile
    const uint32_t ox = tid % (uint32_t)BX, oy = tid / (uint32_t)BX;
    const bool in_o = oy < (uint32_t)BY && y0 + oy < Y;
    const uint32_t px = x0 + ox, py = y0 + oy;
    const I li = (oy + 1u) * RX + ox + 1u;

    // Loads of step t, issued an iteration ahead of the plane that uses them.
    // SPLIT issues the directions with c_z = +1 two iterations ahead instead,
    // so that the bounce-back read of a cell's own slot rev(k) falls in the
    // same iteration as the pulls from the same cache line: a line of a
    // c_z = +1 population in plane p is pulled by the cells of plane p + 1
    // and read back by those of plane p for their c_z = -1 directions, and a
    // line of a c_z = -1 population is pulled by the cells of plane p - 1 and
    // read back, for their c_z = +1 directions, by those of plane p.  (The
    // c_z = +1 values then stay in registers for an extra iteration.)
    // Around interior solids the gather addresses depend on the neighbour
    // mask, loaded at the start of the iteration before the gather.
    //
    // Neighbour mask (NB) or flag of this thread's step t + 1 cell in plane n.
    // Every thread loads, also past the grown tile (its wrapped coordinates
    // stay in the lattice; the values are not used): loads under a branch
    // made the compiler wait for loads in flight where the branches merge.
    auto load_mask = [&](uint32_t n) __attribute__((always_inline)) -> uint32_t {
        const I c = ((z0 + n + Z - 1u) % Z) * XY + gy * X + gx;
        return NB ? nbmask[c] : (uint32_t)flags[c];
    };
    // the loads of plane n's directions of group G into fv
    auto fetch = [&](auto group, uint32_t n, S (&fv)[Q], uint32_t m) __attribute__((always_inline)) {
        constexpr int G = decltype(group)::value;
        const uint32_t zi = (z0 + n + Z - 1u) % Z;
        const I c = zi * XY + gy * X + gx;
        const Seams<I> sm = seams<I>(gx, gy, zi, a);
        if (NB && !W) {
            // in bytes: one vector add per cell part, none per direction
            constexpr I sz = (I)sizeof(S);
            Seams<I> sb = sm;
            sb.xm *= sz, sb.xp *= sz, sb.ym *= sz, sb.yp *= sz, sb.zm *= sz, sb.zp *= sz;
            gather_mask<Q, 0, V, W, G>(fv, A, c * sz, ps * sz, m, sb);
        } else if (NB) {
            gather_mask<Q, 0, V, W, G>(fv, A, c, ps, m, sm);
        } else {
            gather_walls<Q, 0, V, W, G>(fv, A, c, ps, sm);
        }
    };
    using All = std::integral_constant<int, kAll>;
    using Up = std::integral_constant<int, kUp>;
    using Flat = std::integral_constant<int, kFlat>;

    // Step t + 2 of the tile's cell at plane n - 2 of the chunk (iteration n):
    // from the ring planes of iterations n - 2, n - 1 and n, with the mask of
    // its step t + 1 cell (stored in mring by iteration n - 1).
    auto step2 = [&](uint32_t n) __attribute__((always_inline)) {
        const uint32_t zo = z0 + n - 2u;
        const I c = zo * XY + py * X + px;
        const uint32_t pa = (n + 1u) % 3u, oa = (n + 2u) % 3u, b0 = (n + 1u) & 1u, c1 = n & 1u;
        const uint32_t mo = mring[b0 * NI + li];
        const bool self_solid = NB ? (mo & 1u) != 0 : mo != 0;
        if (!__all(self_solid)) {
            const Seams<I> sm = seams<I>(px, py, zo, a);
            T f[Q];
            ring_gather<Q, 0, V, NB, NI, RX>(f, ring, li, pa, oa, b0, c1, mo, sm);
            if (!self_solid) {
                const T rho = bgk<Q, P>(f, a, true);
                if (MOV && a.bc.moving) moving_wall_terms<Q, P>(f, rho, sm, a);
            }
            if (a.nt_store)
                store_own<Q, V, true>(f, B, c, ps, self_solid);
            else
                store_own<Q, V, false>(f, B, c, ps, self_solid);
        }
    };
    // Step t + 1 of this thread's cell at plane n from the loads in fv, into
    // the ring slots a3 (c_z = +1 group) and a2 (the others).
    auto step1 = [&](uint32_t n, const S (&fv)[Q], uint32_t a3, uint32_t a2) __attribute__((always_inline)) {
        T f[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) f[k] = P::load_raw(fv[k]);
        const T rho = bgk<Q, P>(f, a, true);
        if (MOV && a.bc.moving) moving_wall_terms<Q, P>(f, rho, seams<I>(gx, gy, (z0 + n + Z - 1u) % Z, a), a);
        ring_store<Q, 0, V, NI>(f, ring, tid, a3, a2);
    };

    if constexpr (!SPLIT) {
        // One plane n: the step t + 1 cell of plane n from the loads in fv
        // and the mask mcur; the mask of plane n + 1 into mnext and its loads
        // into fv; the step t + 2 cell of plane n - 1.  The loop runs it twice
        // per trip with the two mask registers swapped, so that no register
        // that a load in flight writes is copied (a copy would wait for the
        // load).  (Written with one captured register set: the same body with
        // the set passed by reference compiled to longer waits.)
        S fv[Q];
        auto plane = [&](uint32_t n, uint32_t& mcur, uint32_t& mnext) __attribute__((always_inline)) {
            // (plane numbers wrap modulo Z, so the loads past the chunk's
            // last plane stay in bounds; their values are not used)
            mnext = load_mask(n + 1u);
            // ring slots of plane n (relative plane numbers: only n mod 3 and mod 2 matter)
            const uint32_t a3 = n % 3u, a2 = n & 1u;
            if (in_i) {
                // the step t + 2 cells read this mask in the next iteration
                mring[a2 * NI + tid] = mcur;
                const bool self_solid = NB ? (mcur & 1u) != 0 : mcur != 0;
                if (!self_solid) step1(n, fv, a3, a2);
            }
            fetch(All{}, n + 1u, fv, mnext);
            lds_barrier();
            if (n >= 2u && in_o) step2(n);
            lds_barrier();
        };
        uint32_t m0 = load_mask(0), m1 = 0;
        fetch(All{}, 0, fv, m0);
        for (uint32_t n = 0; n < nplanes; n += 2u) {
            plane(n, m0, m1);
            if (n + 1u < nplanes) plane(n + 1u, m1, m0);
        }
    } else {
        // One plane n: the step t + 1 cell of plane n from fc and the mask mc
        // of plane n; the c_z = 0 and -1 directions of plane n + 1 into fn
        // (its c_z = +1 directions are in flight) with the mask mn of plane
        // n + 1, loaded an iteration ago; the mask of plane n + 2 into mc and
        // its c_z = +1 directions into fc; the step t + 2 cell of plane
        // n - 1.  The loop runs it twice per trip with the register sets
        // swapped (odd: fc = fv1, fn = fv0), so that no register that a load
        // in flight writes is copied.
        S fv0[Q], fv1[Q];
        auto plane = [&](auto odd, uint32_t n, uint32_t& mc, uint32_t& mn) __attribute__((always_inline)) {
            S (&fc)[Q] = decltype(odd)::value ? fv1 : fv0;
            S (&fn)[Q] = decltype(odd)::value ? fv0 : fv1;
            const uint32_t a3 = n % 3u, a2 = n & 1u;
            if (in_i) mring[a2 * NI + tid] = mc;
            const bool self_solid = NB ? (mc & 1u) != 0 : mc != 0;
            mc = load_mask(n + 2u);
            if (in_i && !self_solid) step1(n, fc, a3, a2);
            fetch(Flat{}, n + 1u, fn, mn);
            fetch(Up{}, n + 2u, fc, mc);
            lds_barrier();
            if (n >= 2u && in_o) step2(n);
            lds_barrier();
        };
        uint32_t m0 = load_mask(0), m1 = load_mask(1);
        fetch(All{}, 0, fv0, m0);
        fetch(Up{}, 1, fv1, m1);
        for (uint32_t n = 0; n < nplanes; n += 2u) {
            plane(std::false_type{}, n, m0, m1);
            if (n + 1u < nplanes) plane(std::true_type{}, n + 1u, m1, m0);
        }
    }
}

}  // namespace

template <int Q, int V, int BX, bool W>
bool launch_pull2_bx(const FusedArgs<typename Prec<V>::T>& a, const TwoGeom& g, bool nb, bool mov,
                     const typename Prec<V>::S* A, typename Prec<V>::S* B, const uint8_t* flags,
                     const uint32_t* nbmask, hipStream_t stream) {
    constexpr int sb = (int)sizeof(typename Prec<V>::S), BY = two_rows(Q, sb, BX);
    if constexpr (BY == 0) {
        return false;
    } else {
        const dim3 grid(two_blocks(g)), block(two_threads(BX, BY));
#define PLB_TWO(NBv, MVv)                                                                               \
    hipLaunchKernelGGL((k_pull2<Q, V, NBv, MVv, BX, BY, W, two_split(Q, sb, NBv != 0)>), grid, block, 0, \
                       stream, A, B, flags, nbmask, a, g)
        if (nb && mov) PLB_TWO(1, true);
        else if (nb) PLB_TWO(1, false);
        else if (mov) PLB_TWO(0, true);
        else PLB_TWO(0, false);
#undef PLB_TWO
        return true;
    }
}

template <int Q, int V>
bool launch_pull2(const FusedArgs<typename Prec<V>::T>& a, const TwoGeom& g, bool nb, bool mov,
                  const typename Prec<V>::S* A, typename Prec<V>::S* B, const uint8_t* flags,
                  const uint32_t* nbmask, hipStream_t stream) {
    if (g.bx == 64 && !g.wide) return launch_pull2_bx<Q, V, 64, false>(a, g, nb, mov, A, B, flags, nbmask, stream);
    if (g.bx == 64) return launch_pull2_bx<Q, V, 64, true>(a, g, nb, mov, A, B, flags, nbmask, stream);
    if (g.bx == 32 && !g.wide) return launch_pull2_bx<Q, V, 32, false>(a, g, nb, mov, A, B, flags, nbmask, stream);
    if (g.bx == 32) return launch_pull2_bx<Q, V, 32, true>(a, g, nb, mov, A, B, flags, nbmask, stream);
    return false;
}

#define PLB_TWO_INST(Q, V)                                                                                  \
    template bool launch_pull2<Q, V>(const FusedArgs<typename Prec<V>::T>&, const TwoGeom&, bool, bool,       \
                                     const typename Prec<V>::S*, typename Prec<V>::S*, const uint8_t*,       \
                                     const uint32_t*, hipStream_t);
#define PLB_TWO_INST_Q(Q) \
    PLB_TWO_INST(Q, kF64) PLB_TWO_INST(Q, kF32) PLB_TWO_INST(Q, kF64A32) PLB_TWO_INST(Q, kF32S) PLB_TWO_INST(Q, kF16S)

}  // namespace plb
