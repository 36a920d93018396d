// Device-side pieces shared by the fused-layout kernels (fused.hip,
// fused_row.hpp and fused_two.hpp): launch arguments, the block -> cell
// mapping, neighbour offsets, gathers and stores, and the BGK update.  Layout in HBM (structure of arrays, no halo):
//   f[k * pstride + c],  c = (z*Y + y)*X + x,  x fastest.
#pragma once

#include "bgk.hpp"
#include "fastdiv.hpp"
#include "state.hpp"

namespace plb {

template <class T>
struct FusedArgs {
    uint32_t X, Y, Z;
    uint64_t pstride;
    uint32_t tx_shift, tx_mask;  // TX = 1 << tx_shift
    uint32_t ntx, nty, ntiles;
    uint32_t order;             // 0: linear tile rows; B: B x B blocks of (y, z) rows
    uint32_t order_shift;       // log2(order)
    FastDiv dntx, dnty, dper, dnby;  // by ntx, nty, ntx * nty / 8, nty / order
    uint32_t xcd;               // tile mapping: 0 block i -> tile i; 1 XCD runs; 2 XCD y-stripes
    int32_t nt_load, nt_store;  // non-temporal loads / stores
    T tau, om, omm, F0, F1, F2;
    FusedBC<T> bc;
    // c_k . u_w for each direction k and axis: u_w the velocity of the wall
    // face on the side c_k points to along that axis (zero where c_k has no
    // component along it), in T, as (cx u0 + cy u1) + cz u2 (wall_terms)
    T wcub[27][3];
};

namespace {

constexpr unsigned kBlock = 256;

template <int Q>
__host__ __device__ constexpr int rev_k(int k) { return rev<Q>(k); }

// Workgroup i runs on XCD i % 8 (each XCD has its own L2).  Tile order:
//   0: workgroup i takes tile i; the chip sweeps one contiguous window.
//   1: each XCD takes a contiguous run of tiles (eight separate windows).
//   2: each XCD takes a fixed stripe of nty / 8 tile rows in every z-plane,
//      z-plane after z-plane: one window for the chip, while the rows a row
//      exchanges populations with (y +-1, z +-1) are on the same XCD.
//      Needs nty % 8 == 0 (the launcher falls back to 0 otherwise).
template <class T>
__device__ inline uint32_t tile_of_block(const FusedArgs<T>& a) {
    const uint32_t i = blockIdx.x;
    if (a.xcd == 1) {
        const uint32_t n8 = a.ntiles >> 3, r = a.ntiles & 7u, j = i & 7u;
        return j * n8 + (j < r ? j : r) + (i >> 3);
    }
    if (a.xcd == 2) {
        const uint32_t j = i & 7u, m = i >> 3;
        const uint32_t sy = a.nty >> 3;
        const uint32_t z = fdiv(m, a.dper), r = m - z * a.dper.d;
        const uint32_t ry = fdiv(r, a.dntx), tx = r - ry * a.ntx;
        return (z * a.nty + j * sy + ry) * a.ntx + tx;
    }
    return i;
}

// Tile row number -> (ty, tz): linear, or in square blocks of `order` x
// `order` rows of the y-z plane.
template <class T>
__device__ inline void tile_rows(const FusedArgs<T>& a, uint32_t rest, uint32_t& ty, uint32_t& tz) {
    if (a.order) {
        const uint32_t b = rest >> (2 * a.order_shift), within = rest & (a.order * a.order - 1);
        const uint32_t bz = fdiv(b, a.dnby), by = b - bz * a.dnby.d;
        ty = by * a.order + (within & (a.order - 1));
        tz = bz * a.order + (within >> a.order_shift);
    } else {
        tz = fdiv(rest, a.dnty);
        ty = rest - tz * a.nty;
    }
}

struct Cell {
    uint32_t x, y, z;
    uint64_t c;
    bool valid;
};

// block -> tile -> (x, y, z) of this thread's cell.  Tile rows (ty, z) are
// taken either in linear order or in square blocks of `order` x `order` rows
// of the y-z plane, so that the rows a row exchanges populations with (y +-1,
// z +-1) are worked on at about the same time, while their lines are in L2.
template <class T>
__device__ inline Cell cell_of(const FusedArgs<T>& a) {
    const uint32_t t = tile_of_block(a);
    const uint32_t rest = fdiv(t, a.dntx);
    const uint32_t tx = t - rest * a.ntx;
    uint32_t ty, tz;
    tile_rows(a, rest, ty, tz);
    Cell p;
    p.z = tz;
    p.x = (tx << a.tx_shift) + (threadIdx.x & a.tx_mask);
    p.y = ty * (kBlock >> a.tx_shift) + (threadIdx.x >> a.tx_shift);
    p.valid = p.x < a.X && p.y < a.Y;
    p.c = (uint64_t)p.z * a.X * a.Y + (uint64_t)p.y * a.X + p.x;
    return p;
}

// Element offsets from a cell to its x-1, x+1, y-1, y+1, z-1, z+1
// neighbours with the periodic wrap folded in, in the unsigned index type I
// (modular arithmetic: a "negative" offset wraps and the sum lands in range),
// and whether each of those neighbour layers is a wall layer (a walled axis
// never wraps for a fluid cell: the wall layer sits between it and the edge).
template <class I>
struct Seams {
    I xm, xp, ym, yp, zm, zp;
    bool wxm, wxp, wym, wyp, wzm, wzp;
};

template <class I, class T>
__device__ inline Seams<I> seams(uint32_t x, uint32_t y, uint32_t z, const FusedArgs<T>& a) {
    const I X = a.X, XY = (I)a.X * a.Y;
    Seams<I> s;
    s.xm = x == 0 ? (I)(a.X - 1) : (I)0 - (I)1;
    s.xp = x + 1 == a.X ? (I)0 - (I)(a.X - 1) : (I)1;
    s.ym = y == 0 ? (I)(a.Y - 1) * X : (I)0 - X;
    s.yp = y + 1 == a.Y ? (I)0 - (I)(a.Y - 1) * X : X;
    s.zm = z == 0 ? (I)(a.Z - 1) * XY : (I)0 - XY;
    s.zp = z + 1 == a.Z ? (I)0 - (I)(a.Z - 1) * XY : XY;
    s.wxm = a.bc.wall_lo[0] && x == 1;
    s.wxp = a.bc.wall_hi[0] && x + 2 == a.X;
    s.wym = a.bc.wall_lo[1] && y == 1;
    s.wyp = a.bc.wall_hi[1] && y + 2 == a.Y;
    s.wzm = a.bc.wall_lo[2] && z == 1;
    s.wzp = a.bc.wall_hi[2] && z + 2 == a.Z;
    return s;
}

// Offset from a cell to x - c_k (the pull source), and whether that source is
// a wall layer.  The direction components are compile-time constants, so each
// is at most three adds and three ors.
template <int Q, int K, class I>
__device__ inline I pull_offset(const Seams<I>& s) {
    constexpr int cx = Lat<Q>::cx(K), cy = Lat<Q>::cy(K), cz = Lat<Q>::cz(K);
    I d = 0;
    if (cx > 0) d += s.xm;
    if (cx < 0) d += s.xp;
    if (cy > 0) d += s.ym;
    if (cy < 0) d += s.yp;
    if (cz > 0) d += s.zm;
    if (cz < 0) d += s.zp;
    return d;
}
template <int Q, int K, class I>
__device__ inline bool pull_wall(const Seams<I>& s) {
    constexpr int cx = Lat<Q>::cx(K), cy = Lat<Q>::cy(K), cz = Lat<Q>::cz(K);
    bool w = false;
    if (cx > 0) w = w || s.wxm;
    if (cx < 0) w = w || s.wxp;
    if (cy > 0) w = w || s.wym;
    if (cy < 0) w = w || s.wyp;
    if (cz > 0) w = w || s.wzm;
    if (cz < 0) w = w || s.wzp;
    return w;
}

// Element i of a population buffer.  With 32-bit indices the launcher
// guarantees that every byte offset fits 32 bits, and the access is written
// as buffer base + 32-bit byte offset, which the compiler turns into a load
// or store with a scalar base address (no 64-bit address add per access).
template <class S, class I>
__device__ inline const S* elem(const S* base, I i) {
    if constexpr (sizeof(I) == 4)
        return reinterpret_cast<const S*>(reinterpret_cast<const char*>(base) + i * (I)sizeof(S));
    else
        return base + i;
}
template <class S, class I>
__device__ inline S* elem(S* base, I i) {
    if constexpr (sizeof(I) == 4)
        return reinterpret_cast<S*>(reinterpret_cast<char*>(base) + i * (I)sizeof(S));
    else
        return base + i;
}

// The in-plane (x, y) and the z part of pull_offset.
template <int Q, int K, class I>
__device__ inline I pull_offset_xy(const Seams<I>& s) {
    constexpr int cx = Lat<Q>::cx(K), cy = Lat<Q>::cy(K);
    I d = 0;
    if (cx > 0) d += s.xm;
    if (cx < 0) d += s.xp;
    if (cy > 0) d += s.ym;
    if (cy < 0) d += s.yp;
    return d;
}
template <int Q, int K, class I>
__device__ inline I pull_offset_z(const Seams<I>& s) {
    constexpr int cz = Lat<Q>::cz(K);
    return cz > 0 ? s.zm : (cz < 0 ? s.zp : (I)0);
}

// With 32-bit indices (every byte offset of the array fits 32 bits, so every
// element offset fits 31) the gathers and stores below split an address into
// a part that is uniform across a workgroup (the population times pstride,
// and the z shift: a workgroup's cells share their plane) and a per-cell
// part, so that the per-cell vector work is shared between directions
// instead of repeated for each.
// BB = false: no source is solid (a wavefront away from the wall faces).
template <int Q, int K, int V, int NB, bool NT, bool BB = true, class I>
__device__ inline void pull_gather(typename Prec<V>::T (&f)[Q], const typename Prec<V>::S* __restrict__ A,
                                   I c, I ps, uint32_t m, const Seams<I>& s) {
    if constexpr (K < Q) {
        using P = Prec<V>;
        using S = typename P::S;
        constexpr int R = rev_k<Q>(K);
        // source x - c_k; it is solid when neighbour rev(k) is
        const bool solid = NB ? ((m >> R) & 1u) != 0 : pull_wall<Q, K>(s);
        const S* p;
        if constexpr (!BB && sizeof(I) == 4) {
            (void)solid;
            const int64_t zo = (int32_t)pull_offset_z<Q, K>(s);
            p = elem(A + ((int64_t)K * ps + zo), c + pull_offset_xy<Q, K>(s));
        } else if constexpr (!BB) {
            (void)solid;
            p = A + ((I)K * ps + (c + pull_offset<Q, K>(s)));
        } else if constexpr (sizeof(I) == 4) {
            // byte offsets from A, selected bit by bit
            constexpr uint32_t sz = sizeof(S);
            const uint32_t moved = ((uint32_t)K * ps + pull_offset_z<Q, K>(s)) * sz + (c + pull_offset_xy<Q, K>(s)) * sz;
            const uint32_t own = (uint32_t)R * ps * sz + c * sz;
            const uint32_t src = moved ^ ((moved ^ own) & (0u - (uint32_t)solid));
            p = reinterpret_cast<const S*>(reinterpret_cast<const char*>(A) + src);
        } else {
            const I moved = (I)K * ps + (c + pull_offset<Q, K>(s));
            const I own = (I)R * ps + c;
            // branch-free select (a conditional here becomes divergent branches)
            p = A + (moved ^ ((moved ^ own) & ((I)0 - (I)solid)));
        }
        f[K] = NT ? P::load_raw(__builtin_nontemporal_load(p)) : P::load_raw(*p);
        pull_gather<Q, K + 1, V, NB, NT, BB>(f, A, c, ps, m, s);
    }
}

// Around interior solids with 2-byte storage: both candidates of every
// population, the pulled value and the cell's own reverse slot, are loaded
// before the neighbour mask is known (no load address waits on it); the mask
// then selects.  On MI355X this made the 2-byte step 11% (D3Q19) to 13%
// (D3Q27) faster; 4-byte storage was unchanged and 8-byte 2-3% slower, so
// they keep the single load at a mask-selected address.
template <int Q, int K, int V, class I>
__device__ inline void pull_gather_both(typename Prec<V>::S (&fv)[Q], typename Prec<V>::S (&own)[Q],
                                        const typename Prec<V>::S* __restrict__ A, I c, I ps,
                                        const Seams<I>& s) {
    if constexpr (K < Q) {
        if constexpr (sizeof(I) == 4) {
            // population base and z shift as one uniform pointer; the in-plane
            // offset (shared by the directions with the same c_x, c_y) as the
            // vector operand
            const int64_t zo = (int32_t)pull_offset_z<Q, K>(s);
            fv[K] = *elem(A + ((int64_t)K * ps + zo), c + pull_offset_xy<Q, K>(s));
            if constexpr (K > 0) own[K] = *elem(A + (uint64_t)rev_k<Q>(K) * ps, c);
        } else {
            fv[K] = *elem(A, (I)K * ps + (c + pull_offset<Q, K>(s)));
            if constexpr (K > 0) own[K] = *elem(A, (I)rev_k<Q>(K) * ps + c);
        }
        pull_gather_both<Q, K + 1, V>(fv, own, A, c, ps, s);
    }
}

template <int Q, int V, bool NT, class I>
__device__ inline void own_store(const typename Prec<V>::T (&f)[Q], typename Prec<V>::S* __restrict__ B, I c,
                                 I ps, bool zero) {
    using P = Prec<V>;
    using T = typename P::T;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const typename P::S v = P::store_raw(zero ? (T)0 : f[k]);
        // the population's base pointer is uniform, so every store takes the
        // same vector offset; the empty asm keeps the base a scalar operand
        // (else the compiler chains the full 64-bit addresses in vector
        // registers, one 64-bit add per store)
        using GS = __attribute__((address_space(1))) typename P::S;
        uint64_t base = (uint64_t)(B + (uint64_t)k * ps);
        asm("" : "+s"(base));
        GS* p = reinterpret_cast<GS*>(base + (uint64_t)(c * (I)sizeof(typename P::S)));
        if (NT)
            __builtin_nontemporal_store(v, p);
        else
            *p = v;
    }
}

// Moving walls in the pull scheme.  A population f*_k that a fluid cell sends
// towards a wall comes back to the same cell as direction rev(k) at the next
// step (halfway bounce-back), read from the cell's own slot k, and nothing
// else reads that slot.  So the moving-wall term of Boundary.cpp:213-217 is
// subtracted when the slot is written, with the rho and the expression of the
// push step (which subtracts it on the way): the two schemes store the same
// values.  Only wall faces move; interior solids take no term.  The velocity
// is that of the wall cell's flag, which fused_build_flags sets from the
// coordinates alone (1 + the last of the faces x = X - 1, x = 0, y = Y - 1,
// y = 0, z = Z - 1, z = 0 that holds the cell): so c_k . u_w is the host's
// a.wcub entry for the last axis along which x + c_k lies on a wall face,
// taken from the fluid cell's seams with no load (a load of the flag under
// the branch would make the kernel wait for every load in flight before it,
// the two-step kernel's next planes among them).
template <int Q, int K, class P, class I>
__device__ inline void wall_terms(typename P::T (&f)[Q], typename P::T rho, const Seams<I>& s,
                                  const FusedArgs<typename P::T>& a) {
    using T = typename P::T;
    if constexpr (K < Q) {
        constexpr int R = rev_k<Q>(K);
        constexpr int cy = Lat<Q>::cy(K), cz = Lat<Q>::cz(K);
        // neighbour x + c_K (the pull source of direction rev(K)) on a wall face
        if (pull_wall<Q, R>(s)) {
            const bool wz = cz > 0 ? s.wzp : (cz < 0 ? s.wzm : false);
            const bool wy = cy > 0 ? s.wyp : (cy < 0 ? s.wym : false);
            const T cub = wz ? a.wcub[K][2] : (wy ? a.wcub[K][1] : a.wcub[K][0]);
            const T term = (T)2 * w_scaled<Q, P>(K) * rho * cub * (T)3;
            f[K] = f[K] - term;
        }
        wall_terms<Q, K + 1, P>(f, rho, s, a);
    }
}

// wall_terms for a fluid cell, skipped by the wavefronts none of whose cells
// is next to a wall face (the vote keeps a branch the compiler would
// otherwise turn into arithmetic for every cell).
template <int Q, class P, class I>
__device__ inline void moving_wall_terms(typename P::T (&f)[Q], typename P::T rho, const Seams<I>& s,
                                         const FusedArgs<typename P::T>& a) {
    if (__any(s.wxm || s.wxp || s.wym || s.wyp || s.wzm || s.wzp)) wall_terms<Q, 1, P>(f, rho, s, a);
}

// The a.wcub table of wall_terms from the wall velocities of bc.
template <int Q, class T>
void fill_wall_cub(const FusedBC<T>& bc, T (&wcub)[27][3]) {
    for (int k = 0; k < 27; ++k)
        for (int ax = 0; ax < 3; ++ax) wcub[k][ax] = (T)0;
    for (int k = 0; k < Q; ++k) {
        const int c[3] = {Lat<Q>::cx(k), Lat<Q>::cy(k), Lat<Q>::cz(k)};
        for (int ax = 0; ax < 3; ++ax) {
            if (c[ax] == 0) continue;
            // face 2 ax is the high side of the axis, 2 ax + 1 the low side
            const T* u = bc.vel[1 + 2 * ax + (c[ax] > 0 ? 0 : 1)];
            wcub[k][ax] = (T)c[0] * u[0] + (T)c[1] * u[1] + (T)c[2] * u[2];
        }
    }
}

// Moments and, when `collide`, the BGK update with the velocity-shift force
// in place (bgk.hpp).  Returns rho.
template <int Q, class P>
__device__ inline typename P::T bgk(typename P::T (&f)[Q], const FusedArgs<typename P::T>& a, bool collide) {
    return bgk_update<Q, P>(f, a.tau, a.om, a.omm, a.F0, a.F1, a.F2, collide);
}

}  // namespace

// Row pull step (fused_row.hpp): rows of X == 64 * vec * wpr cells, vec cells
// per lane, wpr wavefronts per row.  The shapes it is built for (the launcher
// in fused_row.hpp is generated from this table):
struct RowShape {
    int vec, wpr;
};
constexpr RowShape kRowShapes[] = {{1, 1}, {1, 2}, {1, 4}, {1, 6}, {1, 8}, {2, 1},
                                   {2, 2}, {2, 3}, {2, 4}, {4, 1}, {4, 2}};
// Wavefronts per block: 4 rows of one wavefront, 2 rows of two, else one row.
__host__ __device__ constexpr uint32_t row_block_waves(int wpr) { return wpr >= 3 ? (uint32_t)wpr : 4u; }
// Rows per block of the row pull step for (vec, X), or 0 where it has no kernel.
inline int pull_row_rows(int vec, int X) {
    for (const RowShape& r : kRowShapes)
        if (r.vec == vec && X == 64 * r.vec * r.wpr) return (int)row_block_waves(r.wpr) / r.wpr;
    return 0;
}
// The row pull step; nts = non-temporal stores.  Returns false if no kernel
// exists for (vec, X).
template <int Q, int V>
bool launch_pull_row(const FusedArgs<typename Prec<V>::T>& a, int vec, bool nts, bool nb,
                     const typename Prec<V>::S* A, typename Prec<V>::S* B, const uint8_t* flags,
                     const uint32_t* nbmask, unsigned nblocks, hipStream_t stream);

// Tiles of a two-step launch: ntx x nty tiles of bx x BY cells in every
// chunk of zc planes, nzc chunks; xcd = 1 gives XCD j the j-th of eight runs
// of kx (z-chunk, y-tile) pairs, so that tiles sharing their grown borders
// share an L2.
struct TwoGeom {
    uint32_t ntx, nty, nzc, zc;
    uint32_t xcd;   // 1: XCD j takes the j-th run of kx (z-chunk, y-tile) pairs; 0: workgroup i takes tile i
    uint32_t kx;
    uint32_t bx;    // tile width, 64 or 32 cells (one workgroup per compute unit either way)
    uint32_t wide;  // population arrays beyond 4 GiB: per-population base pointers
};
// Workgroups of a two-step launch (with the XCD mapping, eight runs of kx
// pairs; the workgroups past the last pair return at once).
__host__ __device__ constexpr uint32_t two_blocks(const TwoGeom& g) {
    return g.xcd ? 8u * g.ntx * g.kx : g.ntx * g.nty * g.nzc;
}

// The step t + 1 ring of the two-step kernel.  Directions are grouped by
// c_z (+1, 0, -1).  The step t + 2 cell at plane z pulls the c_z = +1 group
// from plane z - 1, the c_z = 0 group from plane z and the c_z = -1 group
// from plane z + 1, and its own reverse slots (bounce-back) from plane z: a
// plane's c_z = +1 group is read up to two planes after it is computed, the
// other groups up to one, so the groups keep 3, 2 and 2 planes.  (One plane
// more per group would let the step t + 1 cells of the next plane be written
// while the step t + 2 cells of this one are read, with one workgroup barrier
// per plane instead of two; for D3Q19 FP32-shifted its 64 x 7 tiles ran 354 /
// 309 us per step at 256^3 with solids / walls only, against 341 / 288 for
// two barriers and 64 x 8 tiles.)
template <int Q>
struct ZRing {
    __host__ __device__ static constexpr int grp(int k) {
        return Lat<Q>::cz(k) > 0 ? 0 : (Lat<Q>::cz(k) == 0 ? 1 : 2);
    }
    __host__ __device__ static constexpr int size(int g) {
        int n = 0;
        for (int k = 0; k < Q; ++k) n += grp(k) == g;
        return n;
    }
    __host__ __device__ static constexpr int idx(int k) {
        int n = 0;
        for (int j = 0; j < k; ++j) n += grp(j) == grp(k);
        return n;
    }
    __host__ __device__ static constexpr int slots(int g) { return g == 0 ? 3 : 2; }
    __host__ __device__ static constexpr int base(int g) {
        int b = 0;
        for (int h = 0; h < g; ++h) b += size(h) * slots(h);
        return b;
    }
    // population planes the ring holds per cell (43 for D3Q19, 63 for D3Q27)
    __host__ __device__ static constexpr int planes() { return base(3); }
};
__host__ __device__ constexpr int zring_planes(int q) {
    return q == 19 ? ZRing<19>::planes() : (q == 27 ? ZRing<27>::planes() : (q == 9 ? ZRing<9>::planes() : 0));
}

// Cells of a grown tile, and threads of its workgroup (one cell each).
__host__ __device__ constexpr int two_cells(int bx, int by) { return (bx + 2) * (by + 2); }
__host__ __device__ constexpr int two_threads(int bx, int by) { return (two_cells(bx, by) + 63) / 64 * 64; }
// Whether the two-step kernel issues the step t gathers of the c_z = +1
// directions an iteration earlier than the others (fused_two.hpp): D3Q19
// with 4-byte storage, and with 8-byte storage around interior solids (nb).
// The values held for the extra iteration fit there without spills (136
// VGPRs with 4-byte storage, 256 with 8-byte); D3Q27 (9 such directions) and
// 2-byte storage (13-row tiles at 128 VGPRs) spilled, and 8-byte storage
// without interior solids measured 1% slower.
__host__ __device__ constexpr bool two_split(int q, int sbytes, bool nb) {
    return q == 19 && (sbytes == 4 || (sbytes == 8 && nb));
}

// Rows per tile for a lattice, a storage width and a tile width: the largest
// BY whose ring ((bx + 2) x (BY + 2) cells x (zring_planes(q) population
// planes + two 4-byte mask planes)) fits the LDS of a compute unit (160 KiB
// on gfx950; a workgroup takes all of it), within 1024 threads; 0 where no
// BY >= 3 fits.  The split kernel with 4-byte storage takes at most 704
// cells: 64 x 9 tiles (768 threads) ran slower than 64 x 8, and so did
// 64 x 11 (858 threads in 128 VGPRs: D3Q19 256^3 FP32-shifted 356 / 307
// against 341 / 288 us per step with solids / walls only).
__host__ __device__ constexpr int two_rows(int q, int sbytes, int bx) {
    int best = 0;
    const int maxc = two_split(q, sbytes, true) && sbytes == 4 ? 704 : 1024;
    for (int by = 1; by <= 40; ++by)
        if (two_cells(bx, by) * (zring_planes(q) * sbytes + 8) <= 160 * 1024 && two_cells(bx, by) <= maxc) best = by;
    return best >= 3 ? best : 0;
}
// The tile width a lattice and storage width prefer: 32 for 8-byte storage,
// whose taller tile recomputes fewer grown-border cells (D3Q19 32 x 11
// against 64 x 5: 6-13% faster from 128^3 to 384^3; D3Q27 fits 32-cell tiles
// only), 64 otherwise (D3Q19 4-byte 32 x 18 against 64 x 8: 19-22% slower;
// D3Q27 4-byte 32 x 16 against 64 x 7: 7-10% faster at 256^3 but 2.5%
// slower at 384^3).
__host__ __device__ constexpr int two_tile_pref(int q, int sbytes) { return (void)q, sbytes == 8 ? 32 : 64; }

// Two pull steps in one pass (fused_two.hpp), tiles of g.bx x two_rows(Q, S,
// g.bx) cells.  Returns false where no kernel exists for (Q, V, g.bx, g.wide).
template <int Q, int V>
bool launch_pull2(const FusedArgs<typename Prec<V>::T>& a, const TwoGeom& g, bool nb, bool mov,
                  const typename Prec<V>::S* A, typename Prec<V>::S* B, const uint8_t* flags,
                  const uint32_t* nbmask, hipStream_t stream);

}  // namespace plb
