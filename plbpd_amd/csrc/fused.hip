// Fused mode: one kernel per time step (or per two steps: fused_two.hpp)
// does the work of CollisionSRT,
// Streaming, Boundary::apply_noslip / apply_periodicity and the moments that
// the next collision needs (CollisionSRT.cpp:153-176, Streaming.cpp:63-74,
// Boundary.cpp:189-281, CalcMoments.hpp:22-37), without the stale seam reads
// of the staged step.
//
// Layout in HBM (structure of arrays, no halo):
//   f[k * pstride + c],  c = (z*Y + y)*X + x,  x fastest.
// Two schemes give the same numbers.  The pull step (the default; k_pull here,
// the row kernel of fused_row.hpp, and two steps per launch in fused_two.hpp)
// gathers each population from x - c_k and writes the cell's own slots, so
// every store is a whole aligned line.
// The push step (k_push) reads the cell's own slots and writes each
// population at x + c_k.  One thread per cell in both cell kernels: a block
// covers TX x-consecutive cells of TY rows, so a wavefront touches 64
// consecutive cells of one population.  Periodic axes wrap by index
// arithmetic; a neighbour that is a wall or solid sends the population back
// to the cell's own reverse slot (halfway bounce-back, with the moving-wall
// term of Boundary.cpp:213-217).
#include <algorithm>
#include <type_traits>
#include <vector>

#include "fused_kernels.hpp"

namespace plb {

namespace {

template <class T>
struct Nbr {
    uint32_t xo[3], yo[3], zo[3];
    bool wx[3], wy[3], wz[3];
};

// Neighbour coordinates with periodic wrap (a walled axis never wraps for a
// fluid cell: the wall layer sits between it and the domain edge), and which
// neighbour layers are wall layers.
template <class T>
__device__ inline Nbr<T> neighbours(const Cell& p, const FusedArgs<T>& a) {
    Nbr<T> n;
    n.xo[0] = p.x == 0 ? a.X - 1 : p.x - 1;
    n.xo[1] = p.x;
    n.xo[2] = p.x + 1 == a.X ? 0 : p.x + 1;
    n.yo[0] = p.y == 0 ? a.Y - 1 : p.y - 1;
    n.yo[1] = p.y;
    n.yo[2] = p.y + 1 == a.Y ? 0 : p.y + 1;
    n.zo[0] = p.z == 0 ? a.Z - 1 : p.z - 1;
    n.zo[1] = p.z;
    n.zo[2] = p.z + 1 == a.Z ? 0 : p.z + 1;
    n.wx[0] = a.bc.wall_lo[0] && n.xo[0] == 0;
    n.wx[1] = false;
    n.wx[2] = a.bc.wall_hi[0] && n.xo[2] == a.X - 1;
    n.wy[0] = a.bc.wall_lo[1] && n.yo[0] == 0;
    n.wy[1] = false;
    n.wy[2] = a.bc.wall_hi[1] && n.yo[2] == a.Y - 1;
    n.wz[0] = a.bc.wall_lo[2] && n.zo[0] == 0;
    n.wz[1] = false;
    n.wz[2] = a.bc.wall_hi[2] && n.zo[2] == a.Z - 1;
    return n;
}

// Push step: each fluid cell loads its own populations (aligned), collides,
// and stores each one at the neighbour it moves to, or reflected into its own
// reverse slot when that neighbour is solid (halfway bounce-back with the
// moving-wall term of Boundary.cpp:213-217).
template <int Q, int V, int NB, bool COLLIDE>
__global__ __launch_bounds__(kBlock) void k_push(const typename Prec<V>::S* __restrict__ A,
                                                 typename Prec<V>::S* __restrict__ B,
                                                 const uint8_t* __restrict__ flags,
                                                 const uint32_t* __restrict__ nbmask,
                                                 FusedArgs<typename Prec<V>::T> a) {
    using P = Prec<V>;
    using T = typename P::T;
    const Cell p = cell_of(a);
    if (!p.valid) return;
    const uint64_t c = p.c;
    uint32_t m = 0;
    bool self_solid;
    if (NB) {
        m = nbmask[c];
        self_solid = m & 1u;
    } else {
        self_solid = flags[c] != 0;
    }
    // A wavefront made only of solid cells (a wall row) has nothing to do;
    // otherwise solid lanes write zeros into their own slots.
    if (__all(self_solid)) return;

    T f[Q];
    if (self_solid) {
#pragma unroll
        for (int k = 0; k < Q; ++k) f[k] = (T)0;
    } else if (a.nt_load) {
#pragma unroll
        for (int k = 0; k < Q; ++k) f[k] = P::load_raw(__builtin_nontemporal_load(A + (uint64_t)k * a.pstride + c));
    } else {
#pragma unroll
        for (int k = 0; k < Q; ++k) f[k] = P::load_raw(A[(uint64_t)k * a.pstride + c]);
    }
    T rho = (T)0;
    if (COLLIDE || a.bc.moving) rho = bgk<Q, P>(f, a, COLLIDE);

    const Nbr<T> nb = neighbours(p, a);
    const uint64_t XY = (uint64_t)a.X * a.Y;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
        const uint64_t d = (uint64_t)nb.zo[cz + 1] * XY + (uint64_t)nb.yo[cy + 1] * a.X + nb.xo[cx + 1];
        const bool solid = NB ? ((m >> k) & 1u) != 0 : (nb.wx[cx + 1] || nb.wy[cy + 1] || nb.wz[cz + 1]);
        T v = f[k];
        // the moving-wall term where a wall face (not an interior solid) moves
        const bool face_wall = nb.wx[cx + 1] || nb.wy[cy + 1] || nb.wz[cz + 1];
        if (a.bc.moving && face_wall && !self_solid) {
            const int fl = flags[d];
            const T cub = (T)cx * a.bc.vel[fl][0] + (T)cy * a.bc.vel[fl][1] + (T)cz * a.bc.vel[fl][2];
            const T term = (T)2 * w_scaled<Q, P>(k) * rho * cub * (T)3;
            v = v - term;
        }
        // one store per direction, whatever the lane's case
        uint64_t dst = solid ? (uint64_t)rev_k<Q>(k) * a.pstride + c : (uint64_t)k * a.pstride + d;
        if (self_solid) {
            dst = (uint64_t)k * a.pstride + c;
            v = (T)0;
        }
        if (a.nt_store)
            __builtin_nontemporal_store(P::store_raw(v), B + dst);
        else
            B[dst] = P::store_raw(v);
    }
}

// Pull step on post-collision storage: each fluid cell
// gathers population k from x - c_k, or from its own reverse slot when that
// cell is solid (the same halfway bounce-back as the push step), then
// collides and stores every population in its own slot.  A wavefront writes
// only its own row, so every cache line of B is written whole, by one
// instruction, even around interior solids.  COLLIDE = false gathers only
// (post-collision -> pre-collision populations).
//
// Addressing is kept to a few integer adds per population: the direction
// offsets are per-cell constants (seams()), indices are 32-bit (I = uint32_t)
// whenever Q * pstride fits, and loads are unconditional (every index is in
// range; solid lanes discard what they load and store zeros).
template <int Q, int V, int NB, bool COLLIDE, class I, bool NTS, bool MOV>
__global__ __launch_bounds__(kBlock) void k_pull(const typename Prec<V>::S* __restrict__ A,
                                                 typename Prec<V>::S* __restrict__ B,
                                                 const uint8_t* __restrict__ flags,
                                                 const uint32_t* __restrict__ nbmask,
                                                 FusedArgs<typename Prec<V>::T> a) {
    using P = Prec<V>;
    using T = typename P::T;
    const Cell p = cell_of(a);
    if (!p.valid) return;
    const I c = (I)p.c;
    const Seams<I> sm = seams<I>(p.x, p.y, p.z, a);
    const I ps = (I)a.pstride;
    T f[Q];
    uint32_t m = 0;
    bool self_solid;
    if (NB && sizeof(typename P::S) == 2) {
        typename P::S fv[Q], own[Q];
        pull_gather_both<Q, 0, V>(fv, own, A, c, ps, sm);
        m = nbmask[c];
        self_solid = m & 1u;
        if (__all(self_solid)) return;
#pragma unroll
        for (int k = 0; k < Q; ++k)
            f[k] = P::load_raw(k > 0 && ((m >> rev_k<Q>(k)) & 1u) ? own[k] : fv[k]);
    } else if (NB) {
        // the gather addresses depend on the mask
        m = nbmask[c];
        self_solid = m & 1u;
        if (__all(self_solid)) return;
        pull_gather<Q, 0, V, NB, false>(f, A, c, ps, m, sm);
    } else {
        // the addresses do not depend on the flag: every load is issued
        // before the first wait (a wavefront of wall cells loads in vain);
        // a wavefront none of whose cells borders a wall face pulls every
        // population without the bounce-back select
        if (__any(sm.wxm || sm.wxp || sm.wym || sm.wyp || sm.wzm || sm.wzp))
            pull_gather<Q, 0, V, NB, false>(f, A, c, ps, m, sm);
        else
            pull_gather<Q, 0, V, NB, false, false>(f, A, c, ps, m, sm);
        self_solid = flags[c] != 0;
        if (__all(self_solid)) return;
    }
    if (COLLIDE) {
        const T rho = bgk<Q, P>(f, a, true);
        // MOV: an instantiation for boundaries with a moving wall, so that
        // the others carry no wall-term code (compiled in everywhere, the
        // dormant branch cost the 2-byte step about 5%)
        if (MOV && a.bc.moving && !self_solid) moving_wall_terms<Q, P>(f, rho, sm, a);
    }
    own_store<Q, V, NTS>(f, B, c, ps, self_solid);
}

// One-array (AA-pattern) steps.  A state created with PLB_MODE_FUSED_AA keeps
// one population array and alternates two in-place steps:
//   even: from form P (population k of cell x in its own slot k, before
//         collision) each cell reads its own slots, collides, and writes f*_k
//         into its own slot rev(k): form S;
//   odd:  from form S each fluid cell gathers f_k = S[rev k][x - c_k] (the
//         neighbour's f*_k), or S[k][x] (its own f*_{rev k}, halfway
//         bounce-back) where x - c_k is not fluid; collides; and writes f*_j
//         to the slot it reads for population rev(j): P[j][x + c_j], or
//         P[rev j][x] where x + c_j is not fluid -- form P again.
// Each odd-step cell writes exactly the slots it read and no other cell
// touches them, so the update needs no second array.  The moving-wall term
// of Boundary.cpp:213-217 is subtracted when the value that enters a wall is
// written, with the pull step's expression (wall_terms), so every step gives
// the pull and push steps' numbers bit for bit.  COLLIDE = false streams
// without colliding (the stream-only step of the staged calls).
template <int Q, int V, bool COLLIDE, class I, bool MOV>
__global__ __launch_bounds__(kBlock) void k_aa_even(typename Prec<V>::S* __restrict__ A,
                                                    const uint8_t* __restrict__ flags,
                                                    FusedArgs<typename Prec<V>::T> a) {
    using P = Prec<V>;
    using T = typename P::T;
    const Cell p = cell_of(a);
    if (!p.valid) return;
    const I c = (I)p.c;
    const I ps = (I)a.pstride;
    T f[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) f[k] = P::load_raw(*elem(A, (I)k * ps + c));
    const bool self_solid = flags[c] != 0;
    if (__all(self_solid)) return;
    const T rho = bgk<Q, P>(f, a, COLLIDE);
    if (MOV && a.bc.moving && !self_solid) {
        const Seams<I> sm = seams<I>(p.x, p.y, p.z, a);
        moving_wall_terms<Q, P>(f, rho, sm, a);
    }
    // solid lanes store zeros, so that every line is written whole
#pragma unroll
    for (int k = 0; k < Q; ++k)
        *elem(A, (I)rev_k<Q>(k) * ps + c) = P::store_raw(self_solid ? (T)0 : f[k]);
}

// The slot of population k that an odd step (or the out-of-place gather)
// reads for cell c: S[rev k][x - c_k], or S[k][x] where x - c_k is not fluid.
template <int Q, int K, int NB, class I>
__device__ inline I aa_slot(I c, I ps, uint32_t m, const Seams<I>& s) {
    constexpr int R = rev_k<Q>(K);
    const bool solid = NB ? ((m >> R) & 1u) != 0 : pull_wall<Q, K>(s);
    const I moved = (I)R * ps + (c + pull_offset<Q, K>(s));
    const I own = (I)K * ps + c;
    return moved ^ ((moved ^ own) & ((I)0 - (I)solid));
}

template <int Q, int K, int NB, class I>
__device__ inline void aa_slots(I (&slot)[Q], I c, I ps, uint32_t m, const Seams<I>& s) {
    if constexpr (K < Q) {
        slot[K] = aa_slot<Q, K, NB, I>(c, ps, m, s);
        aa_slots<Q, K + 1, NB, I>(slot, c, ps, m, s);
    }
}

template <int Q, int V, int NB, bool COLLIDE, class I, bool MOV>
__global__ __launch_bounds__(kBlock) void k_aa_odd(typename Prec<V>::S* __restrict__ A,
                                                   const uint8_t* __restrict__ flags,
                                                   const uint32_t* __restrict__ nbmask,
                                                   FusedArgs<typename Prec<V>::T> a) {
    using P = Prec<V>;
    using T = typename P::T;
    const Cell p = cell_of(a);
    if (!p.valid) return;
    const I c = (I)p.c;
    uint32_t m = 0;
    bool self_solid;
    if (NB) {
        m = nbmask[c];
        self_solid = m & 1u;
    } else {
        self_solid = flags[c] != 0;
    }
    // a solid cell owns no slot of the update: it neither reads nor writes
    if (self_solid) return;
    const Seams<I> sm = seams<I>(p.x, p.y, p.z, a);
    const I ps = (I)a.pstride;
    I slot[Q];
    aa_slots<Q, 0, NB, I>(slot, c, ps, m, sm);
    T f[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) f[k] = P::load_raw(*elem(A, slot[k]));
    const T rho = bgk<Q, P>(f, a, COLLIDE);
    if (MOV && a.bc.moving) moving_wall_terms<Q, P>(f, rho, sm, a);
    // f*_j goes where population rev(j) was read
#pragma unroll
    for (int j = 0; j < Q; ++j) *elem(A, slot[rev_k<Q>(j)]) = P::store_raw(f[j]);
}

// Form S -> form P out of place (reading a state between an even and an odd
// step): B[k][x] = the value the odd step would gather; solid cells zero.
template <int Q, int V, int NB, class I>
__global__ __launch_bounds__(kBlock) void k_aa_gather(const typename Prec<V>::S* __restrict__ A,
                                                      typename Prec<V>::S* __restrict__ B,
                                                      const uint8_t* __restrict__ flags,
                                                      const uint32_t* __restrict__ nbmask,
                                                      FusedArgs<typename Prec<V>::T> a) {
    using P = Prec<V>;
    const Cell p = cell_of(a);
    if (!p.valid) return;
    const I c = (I)p.c;
    uint32_t m = 0;
    bool self_solid;
    if (NB) {
        m = nbmask[c];
        self_solid = m & 1u;
    } else {
        self_solid = flags[c] != 0;
    }
    const Seams<I> sm = seams<I>(p.x, p.y, p.z, a);
    const I ps = (I)a.pstride;
    I slot[Q];
    aa_slots<Q, 0, NB, I>(slot, c, ps, m, sm);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const typename P::S v = *elem(A, slot[k]);
        *elem(B, (I)k * ps + c) = self_solid ? P::store_raw((typename P::T)0) : v;
    }
}

// Collision only, in place (pre-collision -> post-collision storage).
template <int Q, int V>
__global__ __launch_bounds__(kBlock) void k_collide_only(typename Prec<V>::S* __restrict__ A,
                                                         const uint8_t* __restrict__ flags,
                                                         FusedArgs<typename Prec<V>::T> a) {
    using P = Prec<V>;
    using T = typename P::T;
    const Cell p = cell_of(a);
    if (!p.valid || flags[p.c]) return;
    T f[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) f[k] = P::load_raw(A[(uint64_t)k * a.pstride + p.c]);
    const T rho = bgk<Q, P>(f, a, true);
    if (a.bc.moving) {
        // the first pull-scheme step: the same write-time moving-wall terms as k_pull
        const Seams<uint64_t> sm = seams<uint64_t>(p.x, p.y, p.z, a);
        moving_wall_terms<Q, P>(f, rho, sm, a);
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) A[(uint64_t)k * a.pstride + p.c] = P::store_raw(f[k]);
}

// Initial state: populations from feq(rho0, u(cell)) evaluated in double
// (EqlbDist.hpp:26-38), stored in the variant's form.
template <int Q, int V>
__global__ void k_finit(typename Prec<V>::S* __restrict__ f, uint32_t X, uint32_t Y, uint32_t Z,
                        uint64_t pstride, double rho0, double u0x, double u0y, double u0z, int kind,
                        uint64_t seed, double amp, const double* tg) {
    using P = Prec<V>;
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)X * Y * Z;
    if (c >= n) return;
    const int x = (int)(c % X), y = (int)((c / X) % Y), z = (int)(c / ((uint64_t)X * Y));
    const double u0[3] = {u0x, u0y, u0z};
    double uc[3];
    cell_velocity(kind, seed, amp, X, Y, Z, u0, tg, x, y, z, uc);
    const double usq = uc[0] * uc[0] + uc[1] * uc[1] + uc[2] * uc[2];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        double v = feq_ref<Q, double>(k, rho0, uc, usq);
        if (P::shifted) v = v - Lat<Q>::w(k);
        typename P::S s;
        if (V == kF16S)
            s = (typename P::S)((float)v * 0x1.0p15f);
        else
            s = (typename P::S)v;
        f[(uint64_t)k * pstride + c] = s;
    }
}

// rho, u (double, SoA) from the populations; solids report rho_s and the wall
// velocity, as Boundary::reset leaves wall cells (Boundary.cpp:58-89).
template <int Q, int V>
__global__ void k_fmoments(const typename Prec<V>::S* __restrict__ f, const uint8_t* __restrict__ flags,
                           uint64_t n, uint64_t pstride, double rho_s, FusedBC<double> bc,
                           double* __restrict__ rho, double* __restrict__ u) {
    using P = Prec<V>;
    using T = typename P::T;
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int fl = flags[c];
    if (fl) {
        rho[c] = rho_s;
        for (int i = 0; i < 3; ++i) u[i * n + c] = bc.vel[fl][i];
        return;
    }
    T fk[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) fk[k] = P::load(f[(uint64_t)k * pstride + c]);
    T r, j0, j1, j2;
    moment_sums<Q, T>(fk, r, j0, j1, j2);
    if (P::shifted) r = (T)1 + r;
    const T ri = (T)1 / r;
    rho[c] = (double)r;
    u[c] = (double)(j0 * ri);
    u[n + c] = (double)(j1 * ri);
    u[2 * n + c] = (double)(j2 * ri);
}

// Populations: device SoA <-> host AoS, plain f in H (shift added back).
template <int Q, int V, class H>
__global__ void k_fpop(typename Prec<V>::S* __restrict__ f, H* __restrict__ host, uint32_t X,
                       uint32_t Y, uint32_t Z, uint64_t pstride, int reference_layout, int upload) {
    using P = Prec<V>;
    const uint64_t n = (uint64_t)X * Y * Z;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * Q;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t % Q);
        const uint64_t c = t / Q;
        uint64_t h = t;
        if (reference_layout) {
            const uint64_t x = c % X, y = (c / X) % Y, z = c / ((uint64_t)X * Y);
            h = ((((z + 1) * (Y + 2)) + y + 1) * (X + 2) + x + 1) * Q + k;
        }
        if (upload) {
            double v = (double)host[h];
            if (P::shifted) v = v - Lat<Q>::w(k);
            typename P::S s;
            if (V == kF16S)
                s = (typename P::S)((float)v * 0x1.0p15f);
            else
                s = (typename P::S)v;
            f[(uint64_t)k * pstride + c] = s;
        } else {
            double v = (double)P::load(f[(uint64_t)k * pstride + c]);
            if (P::shifted) v = v + Lat<Q>::w(k);
            host[h] = (H)v;
        }
    }
}

// moments (double SoA) -> host AoS in H
template <class H>
__global__ void k_fmom_out(const double* __restrict__ src, H* __restrict__ host, uint32_t X,
                           uint32_t Y, uint32_t Z, int vlen, int reference_layout) {
    const uint64_t n = (uint64_t)X * Y * Z;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * vlen;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int v = (int)(t % vlen);
        const uint64_t c = t / vlen;
        uint64_t h = t;
        if (reference_layout) {
            const uint64_t x = c % X, y = (c / X) % Y, z = c / ((uint64_t)X * Y);
            h = ((((z + 1) * (Y + 2)) + y + 1) * (X + 2) + x + 1) * vlen + v;
        }
        host[h] = (H)src[v * n + c];
    }
}

template <class S>
__global__ void k_fraw(const S* __restrict__ f, S* __restrict__ out, uint64_t n, uint64_t pstride, int q) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * q;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t % q);
        const uint64_t c = t / q;
        out[t] = f[(uint64_t)k * pstride + c];
    }
}

// Neighbour-solid mask for domains with interior solids: bit 0 = the cell
// itself is not fluid, bit k = neighbour x + c_k is not fluid.
template <int Q>
__global__ void k_nbmask(const uint8_t* __restrict__ flags, uint32_t* __restrict__ mask, uint32_t X,
                         uint32_t Y, uint32_t Z) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)X * Y * Z;
    if (c >= n) return;
    const uint32_t x = c % X, y = (c / X) % Y, z = c / ((uint64_t)X * Y);
    uint32_t m = flags[c] ? 1u : 0u;
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const uint32_t xd = (x + X + Lat<Q>::cx(k)) % X, yd = (y + Y + Lat<Q>::cy(k)) % Y,
                       zd = (z + Z + Lat<Q>::cz(k)) % Z;
        if (flags[((uint64_t)zd * Y + yd) * X + xd]) m |= 1u << k;
    }
    mask[c] = m;
}

template <int Q, int V>
__global__ void k_fglobal_partial(const typename Prec<V>::S* __restrict__ f,
                                  const uint8_t* __restrict__ flags, uint64_t n, uint64_t pstride,
                                  double* __restrict__ partial) {
    using P = Prec<V>;
    __shared__ double sh[4][kBlock];
    double acc[4] = {0, 0, 0, 0};
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (uint64_t)gridDim.x * blockDim.x) {
        if (flags[c]) continue;
        double r = 0, j0 = 0, j1 = 0, j2 = 0;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const double fk = (double)P::load(f[(uint64_t)k * pstride + c]);
            r += fk;
            j0 += fk * Lat<Q>::cx(k);
            j1 += fk * Lat<Q>::cy(k);
            j2 += fk * Lat<Q>::cz(k);
        }
        if (P::shifted) r = 1.0 + r;
        acc[0] += r;
        acc[1] += j0;
        acc[2] += j1;
        acc[3] += j2;
    }
    for (int d = 0; d < 4; ++d) sh[d][threadIdx.x] = acc[d];
    __syncthreads();
    for (unsigned s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 4; ++d) sh[d][threadIdx.x] += sh[d][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 4; ++d) partial[4 * blockIdx.x + d] = sh[d][0];
}

__global__ void k_fglobal_final(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a[4] = {0, 0, 0, 0};
    for (int b = 0; b < nblocks; ++b)
        for (int d = 0; d < 4; ++d) a[d] += partial[4 * b + d];
    for (int d = 0; d < 4; ++d) out[d] = a[d];
}

template <class Fn>
int with_qv(int q, int v, Fn&& fn) {
    auto vsel = [&](auto qc) {
        switch (v) {
            case kF64: return fn(qc, std::integral_constant<int, kF64>());
            case kF32: return fn(qc, std::integral_constant<int, kF32>());
            case kF64A32: return fn(qc, std::integral_constant<int, kF64A32>());
            case kF32S: return fn(qc, std::integral_constant<int, kF32S>());
            case kF16S: return fn(qc, std::integral_constant<int, kF16S>());
            default: return fail(PLB_ERR_INVALID_ARGUMENT, "unknown storage variant");
        }
    };
    switch (q) {
        case 9: return vsel(std::integral_constant<int, 9>());
        case 19: return vsel(std::integral_constant<int, 19>());
        case 27: return vsel(std::integral_constant<int, 27>());
        default: return fail(PLB_ERR_LOGIC, "Model with q=" + std::to_string(q) + " not yet implemented");
    }
}

int ensure_scratch(plb_state* s, size_t bytes) {
    if (s->scratch_bytes >= bytes) return PLB_OK;
    if (s->scratch) PLB_HIP_TRY(hipFree(s->scratch));
    s->scratch = nullptr;
    s->scratch_bytes = 0;
    PLB_HIP_TRY(hipMalloc(&s->scratch, bytes));
    s->scratch_bytes = bytes;
    return PLB_OK;
}

bool lattice_moves(int q, int axis) {
    if (axis == 1 && q == 9) return false;  // D2Q9 lives in the x-z plane
    return true;
}

}  // namespace

int check_fused_bc(int q, int X, int Y, int Z, const plb_boundary* bc) {
    const int dims[3] = {X, Y, Z};
    for (int a = 0; a < 3; ++a) {
        if (!lattice_moves(q, a)) continue;
        const int hi = bc->type[2 * a], lo = bc->type[2 * a + 1];
        const bool per = hi == kPeriodic && lo == kPeriodic;
        const bool wall = hi == kNoslip && lo == kNoslip;
        if (!per && !wall)
            return fail(PLB_ERR_LOGIC,
                        "fused mode needs each axis periodic on both faces or no-slip on both faces");
        if (wall && dims[a] < 3) return fail(PLB_ERR_INVALID_ARGUMENT, "walled axis shorter than 3 cells");
    }
    for (int f = 0; f < 6; ++f)
        if (bc->type[f] == kNoslip && !bc->has_vel[f])
            return fail(PLB_ERR_OUT_OF_RANGE, "no-slip face without a prescribed velocity");
    return PLB_OK;
}

template <class T>
void make_fused_bc(const plb_boundary& bc, int q, FusedBC<T>* out) {
    *out = FusedBC<T>();
    for (int a = 0; a < 3; ++a) {
        const bool moves = lattice_moves(q, a);
        out->periodic[a] = moves && bc.type[2 * a] == kPeriodic && bc.type[2 * a + 1] == kPeriodic;
        out->wall_hi[a] = moves && bc.type[2 * a] == kNoslip;
        out->wall_lo[a] = moves && bc.type[2 * a + 1] == kNoslip;
    }
    out->moving = 0;
    for (int i = 0; i < 8; ++i)
        for (int d = 0; d < 3; ++d) out->vel[i][d] = (T)0;
    for (int f = 0; f < 6; ++f) {
        if (bc.type[f] != kNoslip) continue;
        for (int d = 0; d < 3; ++d) {
            out->vel[1 + f][d] = (T)bc.vel[f][d];
            if (bc.vel[f][d] != 0.0) out->moving = 1;
        }
    }
}
template void make_fused_bc<double>(const plb_boundary&, int, FusedBC<double>*);
template void make_fused_bc<float>(const plb_boundary&, int, FusedBC<float>*);

int fused_build_flags(plb_state* s, const plb_boundary* bc, const uint8_t* mask) {
    int rc = check_fused_bc(s->q, s->X, s->Y, s->Z, bc);
    if (rc) return rc;
    // Populations between pull steps still owe the streaming of the last
    // step; it streams through the walls and solids of that step, so it is
    // finished with the old flags before they change (as the push step, which
    // streams within the step, would have).
    rc = fused_to_pre(s);
    if (rc) return rc;
    const uint64_t n = s->ncell;
    std::vector<uint8_t> fl(n);
    bool any_solid = false;
    for (uint32_t z = 0; z < (uint32_t)s->Z; ++z)
        for (uint32_t y = 0; y < (uint32_t)s->Y; ++y)
            for (uint32_t x = 0; x < (uint32_t)s->X; ++x) {
                const uint64_t c = ((uint64_t)z * s->Y + y) * s->X + x;
                int f8 = (mask && mask[c]) ? kFlagSolid : kFlagFluid;
                if (f8) any_solid = true;
                const bool on[6] = {x == (uint32_t)s->X - 1, x == 0, y == (uint32_t)s->Y - 1,
                                    y == 0, z == (uint32_t)s->Z - 1, z == 0};
                // the last face wins: wall_terms (fused_kernels.hpp) derives a
                // wall cell's flag from its coordinates by this same rule
                for (int f = 0; f < 6; ++f)
                    if (bc->type[f] == kNoslip && on[f] && lattice_moves(s->q, f / 2)) f8 = 1 + f;
                fl[c] = (uint8_t)f8;
            }
    PLB_HIP_TRY(hipMemcpyAsync(s->flags, fl.data(), n, hipMemcpyHostToDevice, s->stream));
    if (any_solid) {
        if (!s->nbmask) PLB_HIP_TRY(hipMalloc(&s->nbmask, n * sizeof(uint32_t)));
        rc = with_qv(s->q, kF64, [&](auto qc, auto) {
            constexpr int Q = decltype(qc)::value;
            hipLaunchKernelGGL((k_nbmask<Q>), dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s->stream,
                               (const uint8_t*)s->flags, s->nbmask, (uint32_t)s->X, (uint32_t)s->Y,
                               (uint32_t)s->Z);
            PLB_HIP_TRY(hipGetLastError());
            return PLB_OK;
        });
        if (rc) return rc;
    }
    PLB_HIP_TRY(hipStreamSynchronize(s->stream));
    s->has_solids = any_solid;
    s->bc = *bc;
    s->bc_set = true;
    return PLB_OK;
}

int fused_initialize(plb_state* s, double rho0, const double u0[3], const plb_init_spec* init,
                     const double* tg_dev) {
    const int kind = init ? init->kind : PLB_INIT_UNIFORM;
    const uint64_t seed = init ? init->seed : 0;
    const double amp = init ? init->amp : 0.0;
    return with_qv(s->q, s->variant, [&](auto qc, auto vc) {
        constexpr int Q = decltype(qc)::value;
        constexpr int V = decltype(vc)::value;
        using S = typename Prec<V>::S;
        s->post_collision = false;
        s->aa_odd = false;
        hipLaunchKernelGGL((k_finit<Q, V>), dim3(blocks_for(s->ncell, kBlock)), dim3(kBlock), 0,
                           s->stream, (S*)s->buf[s->cur], (uint32_t)s->X, (uint32_t)s->Y,
                           (uint32_t)s->Z, (uint64_t)s->pstride, rho0, u0[0], u0[1], u0[2], kind, seed,
                           amp, tg_dev);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
}

namespace {

enum KernelKind { kPush, kPull, kGather, kStreamOnly, kCollideOnly, kAaEven, kAaOdd, kAaGather, kPull2 };

// Chunks of a two-step launch along z, and the workgroup -> tile mapping.  A
// workgroup holds a compute unit's LDS, so the launch runs in rounds of the
// workgroups resident at once (one per compute unit), and each chunk
// of zc planes costs zc + 2 planes of work (the grown planes at its ends):
// the chunk count minimises rounds x (zc + 2).  With the XCD mapping
// (PLB_TUNE_XCD not 0, compute units in multiples of 8) workgroup i runs on
// XCD i mod 8, and XCD j takes the j-th of eight contiguous runs of the
// (z-chunk, y-tile) pairs, ordered y-tile first, with every x-tile of each:
// neighbouring tiles, whose grown borders overlap, then share that XCD's
// L2.  The runs hold kx = ceil(pairs / 8) pairs (the last may be shorter),
// so the rounds are counted per XCD.
TwoGeom two_geom(const plb_state* s) {
    const int bx = fused_two_tile_x(s);
    const int by = two_rows(s->q, (int)s->sbytes, bx);
    TwoGeom g;
    g.bx = (uint32_t)bx;
    g.wide = (s->tune_index64 || (uint64_t)s->q * s->pstride * s->sbytes > 0xffffffffull) ? 1u : 0u;
    g.ntx = (uint32_t)s->X / (uint32_t)bx;
    g.nty = ((uint32_t)s->Y + (uint32_t)by - 1u) / (uint32_t)by;
    const uint64_t ntx = g.ntx, nty = g.nty, Z = (uint64_t)s->Z;
    const uint64_t ncu = (uint64_t)s->num_cu;  // workgroups resident at once (one per compute unit)
    const bool xcd = s->tune_xcd != 0 && s->num_cu % 8 == 0;
    auto rounds = [&](uint64_t chunks) -> uint64_t {
        if (!xcd) return (ntx * nty * chunks + ncu - 1) / ncu;
        const uint64_t kx = (nty * chunks + 7) / 8, per = ncu / 8;
        return (ntx * kx + per - 1) / per;
    };
    uint64_t zc = Z;
    if (s->tune_two_zc > 0) {
        zc = std::min<uint64_t>((uint64_t)s->tune_two_zc, Z);
    } else {
        uint64_t best = ~0ull;
        for (uint64_t n = 1; n <= 32 && n <= Z; ++n) {
            const uint64_t c = (Z + n - 1) / n, chunks = (Z + c - 1) / c;
            const uint64_t cost = rounds(chunks) * (c + 2);
            if (cost < best) best = cost, zc = c;
        }
    }
    g.zc = (uint32_t)zc;
    g.nzc = (uint32_t)((Z + zc - 1) / zc);
    // with fewer than eight pairs some XCDs would get none
    const uint64_t pairs = nty * g.nzc;
    g.xcd = xcd && pairs >= 8 ? 1u : 0u;
    g.kx = (uint32_t)((pairs + 7) / 8);
    return g;
}

// Launch geometry and arguments shared by every fused-layout kernel.
int launch_kind(plb_state* s, KernelKind kind) {
    if (!s->bc_set) return fail(PLB_ERR_LOGIC, "fused step before the boundary was set");
    // the row pull kernel (fused_row.hpp) takes rows of 64 * vec * wpr cells
    // (kRowShapes).  Automatic choice, from measurement on MI355X (256^3
    // channels): for 4- and 8-byte storage 1 cell per lane without interior
    // solids; around solids 4 (D3Q19 and D2Q9, 4-byte storage) or 2 (8-byte
    // storage, and D3Q27, whose 4-cell kernels need all 256 VGPRs: one
    // wavefront per SIMD, 20% slower for f32s), else the next width with a
    // kernel for X.  2-byte storage takes the cell-per-thread kernel, which
    // measured as fast as the best row width for D3Q19 and faster for D3Q27.
    auto row_ok = [&](int v) { return v > 0 && s->pstride % (size_t)v == 0 && pull_row_rows(v, s->X) > 0; };
    int vec = 0;
    if (kind == kPull && s->tune_vector >= 0) {
        vec = row_ok(s->tune_vector) ? s->tune_vector : 0;
    } else if (kind == kPull && s->sbytes != 2) {
        static const int k421[] = {4, 2, 1}, k21[] = {2, 1}, kOpen[] = {1, 2};
        const bool wide = s->q == 27 || s->sbytes == 8;
        const int* pref = kOpen;
        int npref = 2;
        if (s->has_solids) pref = wide ? k21 : k421, npref = wide ? 2 : 3;
        for (int i = 0; i < npref && !vec; ++i)
            if (row_ok(pref[i])) vec = pref[i];
    }
    const int row_rows = vec ? pull_row_rows(vec, s->X) : 0;
    unsigned tx = 64;
    if (vec)
        tx = 256;
    else if (s->tune_tx > 0)
        tx = (unsigned)s->tune_tx;
    else if (s->X % 256 == 0)
        tx = 256;
    else if (s->X % 128 == 0)
        tx = 128;
    const unsigned ty = vec ? (unsigned)row_rows : kBlock / tx;
    const uint32_t ntx = vec ? 1u : (uint32_t)((s->X + tx - 1) / tx);
    const uint32_t nty = (uint32_t)((s->Y + ty - 1) / ty);
    const uint64_t ntiles = (uint64_t)ntx * nty * s->Z;
    if (ntiles >= (1ull << 31)) return fail(PLB_ERR_INVALID_ARGUMENT, "domain too large for one launch");
    const int rc = with_qv(s->q, s->variant, [&](auto qc, auto vc) {
        constexpr int Q = decltype(qc)::value;
        constexpr int V = decltype(vc)::value;
        using P = Prec<V>;
        using S = typename P::S;
        using T = typename P::T;
        FusedArgs<T> a;
        a.X = s->X;
        a.Y = s->Y;
        a.Z = s->Z;
        a.pstride = s->pstride;
        a.tx_shift = (uint32_t)__builtin_ctz(tx);
        a.tx_mask = tx - 1;
        a.ntx = ntx;
        a.nty = nty;
        a.ntiles = (uint32_t)ntiles;
        a.order = 0;
        if (s->tune_order > 0 && nty % (uint32_t)s->tune_order == 0 && s->Z % s->tune_order == 0)
            a.order = (uint32_t)s->tune_order;
        a.order_shift = a.order ? (uint32_t)__builtin_ctz(a.order) : 0u;
        a.dntx = make_fastdiv(ntx);
        a.dnty = make_fastdiv(nty);
        a.dnby = make_fastdiv(a.order ? nty / a.order : 1u);
        // Automatic tile mapping.  Around interior solids the bounce-back loads
        // read lines again that the y +- 1 and z +- 1 rows read; XCD y-stripes
        // (2) keep those rows on one L2.  For the cell kernel that cuts HBM
        // reads from 1.29x to 1.04x the algorithmic volume (f16s) at the same
        // speed, so it is the cell kernel's default around solids.  The row
        // kernel loads a row's bounce-back lines with its pulled rows, and is
        // 1% (f32s) to 3% (f32) faster with workgroup i on tile i (0), the
        // default everywhere else.
        int xm = s->tune_xcd >= 0 ? s->tune_xcd : (s->has_solids && !vec ? 2 : 0);
        if (xm == 2 && nty % 8 != 0) xm = 0;
        a.xcd = (uint32_t)xm;
        a.dper = make_fastdiv(xm == 2 ? ntx * (nty / 8) : 1u);
        a.nt_load = s->tune_nt_load;
        a.nt_store = s->tune_nt_store;
        a.tau = (T)s->tau;
        if (std::is_same<T, float>::value)  // CollisionSRT.cpp:19-20
            a.om = (T)(1. / (double)a.tau);
        else
            a.om = (T)(1.0 / s->tau);
        a.omm = (T)1 - a.om;
        a.F0 = (T)s->force[0];
        a.F1 = (T)s->force[1];
        a.F2 = (T)s->force[2];
        make_fused_bc<T>(s->bc, s->q, &a.bc);
        fill_wall_cub<Q, T>(a.bc, a.wcub);
        const S* A = (const S*)s->buf[s->cur];
        S* B = (S*)s->buf[s->cur ^ 1];
        const dim3 grid((unsigned)ntiles), block(kBlock);
        const bool nb = s->has_solids;
        // 32-bit indices (and byte offsets) when every byte offset fits and
        // 64-bit ones are not asked for
        const bool idx32 = !s->tune_index64 && (uint64_t)Q * s->pstride * sizeof(S) <= 0xffffffffull;
        switch (kind) {
            case kPush:
                if (nb) hipLaunchKernelGGL((k_push<Q, V, 1, true>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                else hipLaunchKernelGGL((k_push<Q, V, 0, true>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                break;
            case kStreamOnly:
                if (nb) hipLaunchKernelGGL((k_push<Q, V, 1, false>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                else hipLaunchKernelGGL((k_push<Q, V, 0, false>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                break;
            case kPull2:
                if constexpr (Q == 9) {
                    return fail(PLB_ERR_LOGIC, "no two-step kernel for D2Q9");
                } else {
                    if (!launch_pull2<Q, V>(a, two_geom(s), nb, a.bc.moving != 0, A, B, s->flags, s->nbmask,
                                            s->stream))
                        return fail(PLB_ERR_LOGIC, "no two-step kernel for this storage");
                }
                break;
            case kPull:
                if (vec) {
                    if (!launch_pull_row<Q, V>(a, vec, a.nt_store != 0, nb, A, B, s->flags, s->nbmask, grid.x,
                                               s->stream))
                        return fail(PLB_ERR_LOGIC, "no row pull kernel for this row length");
                } else if (idx32) {
                    // non-temporal stores and moving walls by instantiation on the hot path
                    auto go = [&](auto nbc, auto ntc, auto mvc) {
                        constexpr int NBv = decltype(nbc)::value;
                        constexpr bool NTv = decltype(ntc)::value, MVv = decltype(mvc)::value;
                        hipLaunchKernelGGL((k_pull<Q, V, NBv, true, uint32_t, NTv, MVv>), grid, block, 0, s->stream,
                                           A, B, s->flags, s->nbmask, a);
                    };
                    using N0 = std::integral_constant<int, 0>;
                    using N1 = std::integral_constant<int, 1>;
                    using F = std::false_type;
                    using Tr = std::true_type;
                    const bool nts = a.nt_store != 0, mv = a.bc.moving != 0;
                    if (nb) {
                        if (nts) mv ? go(N1{}, Tr{}, Tr{}) : go(N1{}, Tr{}, F{});
                        else mv ? go(N1{}, F{}, Tr{}) : go(N1{}, F{}, F{});
                    } else {
                        if (nts) mv ? go(N0{}, Tr{}, Tr{}) : go(N0{}, Tr{}, F{});
                        else mv ? go(N0{}, F{}, Tr{}) : go(N0{}, F{}, F{});
                    }
                } else {
                    if (nb) hipLaunchKernelGGL((k_pull<Q, V, 1, true, uint64_t, false, true>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                    else hipLaunchKernelGGL((k_pull<Q, V, 0, true, uint64_t, false, true>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                }
                break;
            case kGather:
                if (idx32) {
                    if (nb) hipLaunchKernelGGL((k_pull<Q, V, 1, false, uint32_t, false, false>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                    else hipLaunchKernelGGL((k_pull<Q, V, 0, false, uint32_t, false, false>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                } else {
                    if (nb) hipLaunchKernelGGL((k_pull<Q, V, 1, false, uint64_t, false, false>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                    else hipLaunchKernelGGL((k_pull<Q, V, 0, false, uint64_t, false, false>), grid, block, 0, s->stream, A, B, s->flags, s->nbmask, a);
                }
                break;
            case kCollideOnly:
                hipLaunchKernelGGL((k_collide_only<Q, V>), grid, block, 0, s->stream, (S*)s->buf[s->cur],
                                   (const uint8_t*)s->flags, a);
                break;
            case kAaEven:
            case kAaOdd:
            case kAaGather: {
                S* A1 = (S*)s->buf[s->cur];
                auto go = [&](auto ic, auto nbc, auto colc, auto mvc) {
                    using I = decltype(ic);
                    constexpr int NBv = decltype(nbc)::value;
                    constexpr bool Cv = decltype(colc)::value, MVv = decltype(mvc)::value;
                    if (kind == kAaEven)
                        hipLaunchKernelGGL((k_aa_even<Q, V, Cv, I, MVv>), grid, block, 0, s->stream, A1,
                                           (const uint8_t*)s->flags, a);
                    else if (kind == kAaOdd)
                        hipLaunchKernelGGL((k_aa_odd<Q, V, NBv, Cv, I, MVv>), grid, block, 0, s->stream, A1,
                                           (const uint8_t*)s->flags, s->nbmask, a);
                    else
                        hipLaunchKernelGGL((k_aa_gather<Q, V, NBv, I>), grid, block, 0, s->stream, A1, B,
                                           (const uint8_t*)s->flags, s->nbmask, a);
                };
                using N0 = std::integral_constant<int, 0>;
                using N1 = std::integral_constant<int, 1>;
                using F = std::false_type;
                using Tr = std::true_type;
                const bool mv = a.bc.moving != 0, col = s->aa_collide;
                auto by_nb = [&](auto ic) {
                    auto by_col = [&](auto nbc) {
                        if (kind == kAaGather) return go(ic, nbc, F{}, F{});
                        if (col) mv ? go(ic, nbc, Tr{}, Tr{}) : go(ic, nbc, Tr{}, F{});
                        else mv ? go(ic, nbc, F{}, Tr{}) : go(ic, nbc, F{}, F{});
                    };
                    // the even step reads no neighbour, so no mask instantiation
                    if (nb && kind != kAaEven) by_col(N1{});
                    else by_col(N0{});
                };
                if (idx32) by_nb(uint32_t());
                else by_nb(uint64_t());
                break;
            }
        }
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    // every kernel but the in-place ones writes the other buffer
    if (rc == PLB_OK && kind != kCollideOnly && kind != kAaEven && kind != kAaOdd) s->cur ^= 1;
    return rc;
}


}  // namespace

int fused_two_tile_x(const plb_state* s) {
    auto fits = [&](int bx) { return s->X % bx == 0 && two_rows(s->q, (int)s->sbytes, bx) > 0; };
    if (s->tune_two_bx > 0) return fits(s->tune_two_bx) ? s->tune_two_bx : 0;
    const int pref = two_tile_pref(s->q, (int)s->sbytes);
    for (int bx : {pref, 96 - pref})
        if (fits(bx)) return bx;
    return 0;
}

bool fused_two_eligible(const plb_state* s) {
    if (s->tune_two == 0 || s->aa || !fused_uses_pull(s) || s->mode != PLB_MODE_FUSED) return false;
    // automatic: 4- and 8-byte storage, where two steps per pass measured
    // 1.40x (D3Q19 256^3 FP32-shifted, 10% solids), 1.57x (walls only), 1.26x
    // and 1.41x (FP64), 1.35x (D3Q27 FP32-shifted, walls) and 1.21x (D3Q27
    // FP64, walls) faster on MI355X;
    // 2-byte storage ran 1.23x slower with solids (its one-step kernel is
    // closer to the arithmetic limit)
    if (s->tune_two < 0 && s->sbytes == 2) return false;
    // automatic: lattices of at least 2^18 cells (2^19 for D3Q27).  Below
    // that the launch has too few tiles and chunks of a few planes, and one
    // step per launch measured faster (32^3: 0.66-0.91x; 64 x 64 x 32:
    // 0.67-0.96x; D3Q27 64^3: 0.89-0.94x; D3Q19 64^3 0.94-1.27x)
    const uint64_t cells = (uint64_t)s->X * (uint64_t)s->Y * (uint64_t)s->Z;
    if (s->tune_two < 0 && cells < (s->q == 27 ? (1ull << 19) : (1ull << 18))) return false;
    if ((s->q != 19 && s->q != 27) || fused_two_tile_x(s) == 0) return false;
    // offsets within one population array are 32-bit
    if ((uint64_t)s->pstride * s->sbytes > 0xffffffffull) return false;
    return true;
}

int fused_launch_two(plb_state* s) {
    if (!s->post_collision || !fused_two_eligible(s))
        return fail(PLB_ERR_LOGIC, "two-step launch needs an eligible state in post-collision form");
    const int rc = launch_kind(s, kPull2);
    s->moments_valid = false;
    return rc;
}

bool fused_uses_pull(const plb_state* s) {
    // the pull step is the faster one for every storage type and boundary on
    // MI355X (whole-line stores; moving walls by write-time terms)
    return !s->aa && s->tune_scheme != PLB_SCHEME_PUSH;
}

namespace {
// A one-array state between an even and an odd step (form S) is read after
// its populations are gathered into a second array, allocated for the
// purpose; the old array is then freed, so the state holds one array again.
int aa_to_pre(plb_state* s) {
    if (!s->aa_odd) return PLB_OK;
    const size_t pop_bytes = s->pstride * s->q * s->sbytes;
    void* tmp = nullptr;
    PLB_HIP_TRY(hipMalloc(&tmp, pop_bytes));
    s->buf[s->cur ^ 1] = tmp;
    int rc = launch_kind(s, kAaGather);  // writes buf[cur ^ 1] and swaps
    if (rc == PLB_OK) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess) e = hipFree(s->buf[s->cur ^ 1]);
        if (e != hipSuccess) rc = fail(PLB_ERR_HIP, hipGetErrorString(e));
        s->buf[s->cur ^ 1] = nullptr;
        s->aa_odd = false;
        // a captured step graph holds the old array's address
        if (s->graph) (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
        s->graph_steps = 0;
    } else {
        (void)hipFree(tmp);
        s->buf[s->cur ^ 1] = nullptr;
    }
    return rc;
}
}  // namespace

// The populations in buf[cur] are stored either before collision (what the
// reference's n holds after a whole step) or, between pull steps, after the
// collision of the step in progress, whose streaming is still pending.  A
// pull-mode step from the pre-collision form is the collision alone; every
// later step gathers (finishing the pending streaming), then collides.
// Reading the state performs the pending gather first.
int fused_to_pre(plb_state* s) {
    if (s->aa) return aa_to_pre(s);
    if (!s->post_collision) return PLB_OK;
    const int rc = launch_kind(s, kGather);
    if (rc == PLB_OK) s->post_collision = false;
    return rc;
}

int fused_launch(plb_state* s, bool collide) {
    int rc;
    if (s->aa) {
        s->aa_collide = collide;
        rc = launch_kind(s, s->aa_odd ? kAaOdd : kAaEven);
        if (rc == PLB_OK) s->aa_odd = !s->aa_odd;
        s->moments_valid = false;
        return rc;
    }
    if (!collide) {
        rc = fused_to_pre(s);
        if (rc == PLB_OK) rc = launch_kind(s, kStreamOnly);
    } else if (fused_uses_pull(s)) {
        if (!s->post_collision) {
            rc = launch_kind(s, kCollideOnly);
            if (rc == PLB_OK) s->post_collision = true;
        } else {
            rc = launch_kind(s, kPull);
        }
    } else {
        rc = fused_to_pre(s);
        if (rc == PLB_OK) rc = launch_kind(s, kPush);
    }
    s->moments_valid = false;
    return rc;
}

int fused_moments(plb_state* s) {
    if (!s->bc_set) return fail(PLB_ERR_LOGIC, "moments before the boundary was set");
    int prc = fused_to_pre(s);
    if (prc) return prc;
    FusedBC<double> bc;
    make_fused_bc<double>(s->bc, s->q, &bc);
    const int rc = with_qv(s->q, s->variant, [&](auto qc, auto vc) {
        constexpr int Q = decltype(qc)::value;
        constexpr int V = decltype(vc)::value;
        using S = typename Prec<V>::S;
        hipLaunchKernelGGL((k_fmoments<Q, V>), dim3(blocks_for(s->ncell, kBlock)), dim3(kBlock), 0,
                           s->stream, (const S*)s->buf[s->cur], (const uint8_t*)s->flags,
                           (uint64_t)s->ncell, (uint64_t)s->pstride, s->bc.rho_s, bc, (double*)s->rho,
                           (double*)s->u);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc == PLB_OK) s->moments_valid = true;
    return rc;
}

int fused_transfer(plb_state* s, int field, int layout, int dtype, void* host, bool upload) {
    const int ref = layout == PLB_LAYOUT_REFERENCE;
    const uint64_t n = s->ncell;
    if (field == PLB_FIELD_NTMP) return fail(PLB_ERR_LOGIC, "ntmp is not kept in fused mode");
    if (field != PLB_FIELD_N && field != PLB_FIELD_RHO && field != PLB_FIELD_U)
        return fail(PLB_ERR_INVALID_ARGUMENT, "unknown field");
    if (upload && field != PLB_FIELD_N)
        return fail(PLB_ERR_LOGIC, "rho and u are derived from the populations in fused mode");
    const int vlen = field == PLB_FIELD_N ? s->q : (field == PLB_FIELD_RHO ? 1 : 3);
    const uint64_t hcells = ref ? ((uint64_t)s->X + 2) * ((uint64_t)s->Y + 2) * ((uint64_t)s->Z + 2) : n;
    const size_t esz = dtype == PLB_DTYPE_F64 ? 8 : 4;
    const size_t hbytes = hcells * vlen * esz;
    if (field != PLB_FIELD_N && !s->moments_valid) {
        const int rc = fused_moments(s);
        if (rc) return rc;
    }
    if (field == PLB_FIELD_N && !upload) {
        const int rc = fused_to_pre(s);
        if (rc) return rc;
    }
    void* stage = nullptr;
    PLB_HIP_TRY(hipMalloc(&stage, hbytes));
    int rc = PLB_OK;
    // a host array in reference layout keeps its halo: stage it both ways
    if (upload || ref) {
        hipError_t e = hipMemcpyAsync(stage, host, hbytes, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) rc = fail(PLB_ERR_HIP, hipGetErrorString(e));
    }
    if (rc == PLB_OK) {
        if (field == PLB_FIELD_N) {
            rc = with_qv(s->q, s->variant, [&](auto qc, auto vc) {
                constexpr int Q = decltype(qc)::value;
                constexpr int V = decltype(vc)::value;
                using S = typename Prec<V>::S;
                const unsigned nb = stride_blocks(n * Q, kBlock);
                if (dtype == PLB_DTYPE_F64)
                    hipLaunchKernelGGL((k_fpop<Q, V, double>), dim3(nb), dim3(kBlock), 0, s->stream,
                                       (S*)s->buf[s->cur], (double*)stage, (uint32_t)s->X, (uint32_t)s->Y,
                                       (uint32_t)s->Z, (uint64_t)s->pstride, ref, upload ? 1 : 0);
                else
                    hipLaunchKernelGGL((k_fpop<Q, V, float>), dim3(nb), dim3(kBlock), 0, s->stream,
                                       (S*)s->buf[s->cur], (float*)stage, (uint32_t)s->X, (uint32_t)s->Y,
                                       (uint32_t)s->Z, (uint64_t)s->pstride, ref, upload ? 1 : 0);
                PLB_HIP_TRY(hipGetLastError());
                return PLB_OK;
            });
        } else {
            const double* src = field == PLB_FIELD_RHO ? (const double*)s->rho : (const double*)s->u;
            const unsigned nb = stride_blocks(n * vlen, kBlock);
            if (dtype == PLB_DTYPE_F64)
                hipLaunchKernelGGL((k_fmom_out<double>), dim3(nb), dim3(kBlock), 0, s->stream, src,
                                   (double*)stage, (uint32_t)s->X, (uint32_t)s->Y, (uint32_t)s->Z, vlen, ref);
            else
                hipLaunchKernelGGL((k_fmom_out<float>), dim3(nb), dim3(kBlock), 0, s->stream, src,
                                   (float*)stage, (uint32_t)s->X, (uint32_t)s->Y, (uint32_t)s->Z, vlen, ref);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = fail(PLB_ERR_HIP, hipGetErrorString(e));
        }
    }
    if (rc == PLB_OK && !upload) {
        hipError_t e = hipMemcpyAsync(host, stage, hbytes, hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) rc = fail(PLB_ERR_HIP, hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(s->stream);
    (void)hipFree(stage);
    if (rc == PLB_OK && e != hipSuccess) rc = fail(PLB_ERR_HIP, hipGetErrorString(e));
    if (rc == PLB_OK && upload) {
        s->moments_valid = false;
        s->post_collision = false;  // uploaded populations are pre-collision
        s->aa_odd = false;
    }
    return rc;
}

int fused_download_raw(plb_state* s, void* host, size_t bytes) {
    if (s->bc_set) {
        const int prc = fused_to_pre(s);
        if (prc) return prc;
    }
    const size_t need = s->ncell * s->q * s->sbytes;
    if (bytes < need) return fail(PLB_ERR_INVALID_ARGUMENT, "raw buffer too small");
    void* stage = nullptr;
    PLB_HIP_TRY(hipMalloc(&stage, need));
    const unsigned nb = stride_blocks(s->ncell * s->q, kBlock);
    if (s->sbytes == 8)
        hipLaunchKernelGGL((k_fraw<uint64_t>), dim3(nb), dim3(kBlock), 0, s->stream,
                           (const uint64_t*)s->buf[s->cur], (uint64_t*)stage, (uint64_t)s->ncell,
                           (uint64_t)s->pstride, s->q);
    else if (s->sbytes == 4)
        hipLaunchKernelGGL((k_fraw<uint32_t>), dim3(nb), dim3(kBlock), 0, s->stream,
                           (const uint32_t*)s->buf[s->cur], (uint32_t*)stage, (uint64_t)s->ncell,
                           (uint64_t)s->pstride, s->q);
    else
        hipLaunchKernelGGL((k_fraw<uint16_t>), dim3(nb), dim3(kBlock), 0, s->stream,
                           (const uint16_t*)s->buf[s->cur], (uint16_t*)stage, (uint64_t)s->ncell,
                           (uint64_t)s->pstride, s->q);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, stage, need, hipMemcpyDeviceToHost, s->stream);
    hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(PLB_ERR_HIP, hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(PLB_ERR_HIP, hipGetErrorString(e2));
    return PLB_OK;
}

int fused_global(plb_state* s, double out[4]) {
    if (!s->bc_set) return fail(PLB_ERR_LOGIC, "global moments before the boundary was set");
    int prc = fused_to_pre(s);
    if (prc) return prc;
    const int nb = 1024;
    int rc = ensure_scratch(s, (size_t)(4 * nb + 4) * sizeof(double));
    if (rc) return rc;
    double* partial = (double*)s->scratch;
    rc = with_qv(s->q, s->variant, [&](auto qc, auto vc) {
        constexpr int Q = decltype(qc)::value;
        constexpr int V = decltype(vc)::value;
        using S = typename Prec<V>::S;
        hipLaunchKernelGGL((k_fglobal_partial<Q, V>), dim3(nb), dim3(kBlock), 0, s->stream,
                           (const S*)s->buf[s->cur], (const uint8_t*)s->flags, (uint64_t)s->ncell,
                           (uint64_t)s->pstride, partial);
        PLB_HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_fglobal_final, dim3(1), dim3(64), 0, s->stream, partial, nb, partial + 4 * nb);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc) return rc;
    PLB_HIP_TRY(hipMemcpyAsync(out, partial + 4 * nb, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PLB_HIP_TRY(hipStreamSynchronize(s->stream));
    return PLB_OK;
}

}  // namespace plb
