// Device-side pieces shared by the fused-layout kernels (fused.hip and
// fused_row.hpp): launch arguments, the block -> cell mapping and the BGK
// update.  Layout in HBM (structure of arrays, no halo):
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

}  // namespace plb
