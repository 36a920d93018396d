// Batched domains: many independent channel domains (uneven sizes, their own
// tau and force) advanced by one launch per step.  No reference counterpart:
// plbpd runs one domain per PLBPD object (PLBPD.hpp:14-17).
//
// Layout: one pool per population buffer; domain d owns
//   f[off_d + k * ps_d + c],  c = (z*Y_d + y)*X_d + x,  ps_d = N_d rounded up to 64.
// Work: a tile is `tile` consecutive cells of one domain (the last tile of a
// domain may be short).  A launch covers every tile, either one block per
// tile or a persistent grid whose blocks walk the tiles.  By default
// workgroup i takes tile i (a persistent block strides by the grid), so the
// chip sweeps one window of memory; the XCD mapping instead hands each XCD a
// contiguous run of tiles, so a domain's neighbouring rows share an L2
// (batch_tiles).  The per-cell work is the step of fused.hip, pull (the
// default) or push: the same collision, periodic wraps on periodic axes and
// halfway bounce-back on the wall layers of no-slip faces (stationary or
// moving).
#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "bgk.hpp"
#include "fastdiv.hpp"
#include "state.hpp"

struct plb_batch;

namespace plb {
namespace {

constexpr int kMaxBlock = 256;

template <class T>
struct BDom {
    uint32_t X, Y, Z, N;
    FastDiv dX, dY;  // division by X and Y (cell -> coordinates)
    uint64_t off;    // element offset of population 0 in the pool
    uint64_t ps;     // elements between populations
    T tau, om, omm, F0, F1, F2;
};

// Coordinates of cell c of a domain.
template <class T>
__device__ inline void cell_xyz(const BDom<T>& D, uint32_t c, uint32_t& x, uint32_t& y, uint32_t& z) {
    const uint32_t r = fdiv(c, D.dX);
    x = c - r * D.X;
    z = fdiv(r, D.dY);
    y = r - z * D.Y;
}

template <class T>
struct BArgs {
    const BDom<T>* doms;
    const uint2* tiles;  // (domain, first cell)
    uint32_t ntiles;
    uint32_t tile;       // cells per tile
    uint32_t persistent; // 1: gridDim.x blocks walk the tiles
    uint32_t xcd;        // 1: each XCD takes a contiguous run of tiles; 0: one window for the chip
    FusedBC<T> bc;       // shared boundary kinds
};

__device__ inline uint32_t xcd_worker(uint32_t i, uint32_t n) {
    const uint32_t n8 = n >> 3, r = n & 7u, j = i & 7u;
    return j * n8 + (j < r ? j : r) + (i >> 3);
}

template <int Q, int V>
__device__ inline void batch_cell(const typename Prec<V>::S* __restrict__ A, typename Prec<V>::S* __restrict__ B,
                                  const BDom<typename Prec<V>::T>& D, const FusedBC<typename Prec<V>::T>& bc,
                                  uint32_t c) {
    using P = Prec<V>;
    using T = typename P::T;
    uint32_t x, y, z;
    cell_xyz(D, c, x, y, z);
    // wall layers of no-slip faces are solid
    const bool solid_self = (bc.wall_lo[0] && x == 0) || (bc.wall_hi[0] && x == D.X - 1) ||
                            (bc.wall_lo[1] && y == 0) || (bc.wall_hi[1] && y == D.Y - 1) ||
                            (bc.wall_lo[2] && z == 0) || (bc.wall_hi[2] && z == D.Z - 1);
    if (solid_self) return;
    const typename P::S* a = A + D.off;
    typename P::S* b = B + D.off;
    T f[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) f[k] = P::load_raw(a[(uint64_t)k * D.ps + c]);
    const T rr = bgk_update<Q, P>(f, D.tau, D.om, D.omm, D.F0, D.F1, D.F2, true);
    const uint32_t xo[3] = {x == 0 ? D.X - 1 : x - 1, x, x + 1 == D.X ? 0 : x + 1};
    const uint32_t yo[3] = {y == 0 ? D.Y - 1 : y - 1, y, y + 1 == D.Y ? 0 : y + 1};
    const uint32_t zo[3] = {z == 0 ? D.Z - 1 : z - 1, z, z + 1 == D.Z ? 0 : z + 1};
    const bool wx[3] = {bc.wall_lo[0] && xo[0] == 0, false, bc.wall_hi[0] && xo[2] == D.X - 1};
    const bool wy[3] = {bc.wall_lo[1] && yo[0] == 0, false, bc.wall_hi[1] && yo[2] == D.Y - 1};
    const bool wz[3] = {bc.wall_lo[2] && zo[0] == 0, false, bc.wall_hi[2] && zo[2] == D.Z - 1};
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
        const bool solid = wx[cx + 1] || wy[cy + 1] || wz[cz + 1];
        T v = f[k];
        if (bc.moving && solid) {
            // the wall's face: the last no-slip face in reference order wins
            int face = -1;
            const bool on[6] = {xo[cx + 1] == D.X - 1 && bc.wall_hi[0], xo[cx + 1] == 0 && bc.wall_lo[0],
                                yo[cy + 1] == D.Y - 1 && bc.wall_hi[1], yo[cy + 1] == 0 && bc.wall_lo[1],
                                zo[cz + 1] == D.Z - 1 && bc.wall_hi[2], zo[cz + 1] == 0 && bc.wall_lo[2]};
            for (int fc = 0; fc < 6; ++fc)
                if (on[fc]) face = fc;
            const T* ub = bc.vel[1 + face];
            const T cub = (T)cx * ub[0] + (T)cy * ub[1] + (T)cz * ub[2];
            v = v - (T)2 * w_scaled<Q, P>(k) * rr * cub * (T)3;
        }
        const uint32_t d = (zo[cz + 1] * D.Y + yo[cy + 1]) * D.X + xo[cx + 1];
        const uint64_t dst = solid ? (uint64_t)rev<Q>(k) * D.ps + c : (uint64_t)k * D.ps + d;
        b[dst] = P::store_raw(v);
    }
}

// The face a wall cell belongs to: the last no-slip face in reference order
// wins (Boundary.cpp:17-18), as the flags of fused.hip.
template <class T>
__device__ inline int wall_face(const BDom<T>& D, const FusedBC<T>& bc, uint32_t x, uint32_t y, uint32_t z) {
    int face = -1;
    const bool on[6] = {x == D.X - 1 && bc.wall_hi[0], x == 0 && bc.wall_lo[0], y == D.Y - 1 && bc.wall_hi[1],
                        y == 0 && bc.wall_lo[1], z == D.Z - 1 && bc.wall_hi[2], z == 0 && bc.wall_lo[2]};
    for (int fc = 0; fc < 6; ++fc)
        if (on[fc]) face = fc;
    return face;
}

// Pull step of one cell (see k_pull in fused.hip): gather population k from
// x - c_k, or from the cell's own slot rev(k) where that neighbour is a wall
// layer (halfway bounce-back); then, when COLLIDE, the BGK update and, for a
// moving wall, the wall term subtracted from the slot that the cell reads
// back at the next step; store into the cell's own slots.  Wall cells store
// zeros, so that every line of B is written whole.
template <int Q, int V, bool COLLIDE>
__device__ inline void batch_pull_cell(const typename Prec<V>::S* __restrict__ A,
                                       typename Prec<V>::S* __restrict__ B, const BDom<typename Prec<V>::T>& D,
                                       const FusedBC<typename Prec<V>::T>& bc, uint32_t c) {
    using P = Prec<V>;
    using T = typename P::T;
    uint32_t x, y, z;
    cell_xyz(D, c, x, y, z);
    const typename P::S* a = A + D.off;
    typename P::S* b = B + D.off;
    const bool solid_self = (bc.wall_lo[0] && x == 0) || (bc.wall_hi[0] && x == D.X - 1) ||
                            (bc.wall_lo[1] && y == 0) || (bc.wall_hi[1] && y == D.Y - 1) ||
                            (bc.wall_lo[2] && z == 0) || (bc.wall_hi[2] && z == D.Z - 1);
    if (solid_self) {
#pragma unroll
        for (int k = 0; k < Q; ++k) b[(uint64_t)k * D.ps + c] = P::store_raw((T)0);
        return;
    }
    const uint32_t xo[3] = {x == 0 ? D.X - 1 : x - 1, x, x + 1 == D.X ? 0 : x + 1};
    const uint32_t yo[3] = {y == 0 ? D.Y - 1 : y - 1, y, y + 1 == D.Y ? 0 : y + 1};
    const uint32_t zo[3] = {z == 0 ? D.Z - 1 : z - 1, z, z + 1 == D.Z ? 0 : z + 1};
    const bool wx[3] = {bc.wall_lo[0] && xo[0] == 0, false, bc.wall_hi[0] && xo[2] == D.X - 1};
    const bool wy[3] = {bc.wall_lo[1] && yo[0] == 0, false, bc.wall_hi[1] && yo[2] == D.Y - 1};
    const bool wz[3] = {bc.wall_lo[2] && zo[0] == 0, false, bc.wall_hi[2] && zo[2] == D.Z - 1};
    T f[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
        const bool bb = wx[1 - cx] || wy[1 - cy] || wz[1 - cz];
        const uint32_t s = (zo[1 - cz] * D.Y + yo[1 - cy]) * D.X + xo[1 - cx];
        const uint64_t src = bb ? (uint64_t)rev<Q>(k) * D.ps + c : (uint64_t)k * D.ps + s;
        f[k] = P::load_raw(a[src]);
    }
    if (COLLIDE) {
        const T rr = bgk_update<Q, P>(f, D.tau, D.om, D.omm, D.F0, D.F1, D.F2, true);
        if (bc.moving) {
#pragma unroll
            for (int k = 1; k < Q; ++k) {
                const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
                if (wx[cx + 1] || wy[cy + 1] || wz[cz + 1]) {
                    const int face = wall_face(D, bc, xo[cx + 1], yo[cy + 1], zo[cz + 1]);
                    const T* ub = bc.vel[1 + face];
                    const T cub = (T)cx * ub[0] + (T)cy * ub[1] + (T)cz * ub[2];
                    f[k] = f[k] - (T)2 * w_scaled<Q, P>(k) * rr * cub * (T)3;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) b[(uint64_t)k * D.ps + c] = P::store_raw(f[k]);
}

// The first pull-scheme step from pre-collision storage: the collision alone,
// in place, with the same write-time wall terms.
template <int Q, int V>
__device__ inline void batch_collide_cell(typename Prec<V>::S* A, const BDom<typename Prec<V>::T>& D,
                                          const FusedBC<typename Prec<V>::T>& bc, uint32_t c) {
    using P = Prec<V>;
    using T = typename P::T;
    uint32_t x, y, z;
    cell_xyz(D, c, x, y, z);
    const bool solid_self = (bc.wall_lo[0] && x == 0) || (bc.wall_hi[0] && x == D.X - 1) ||
                            (bc.wall_lo[1] && y == 0) || (bc.wall_hi[1] && y == D.Y - 1) ||
                            (bc.wall_lo[2] && z == 0) || (bc.wall_hi[2] && z == D.Z - 1);
    if (solid_self) return;
    typename P::S* a = A + D.off;
    T f[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) f[k] = P::load_raw(a[(uint64_t)k * D.ps + c]);
    const T rr = bgk_update<Q, P>(f, D.tau, D.om, D.omm, D.F0, D.F1, D.F2, true);
    if (bc.moving) {
        const uint32_t xo[3] = {x == 0 ? D.X - 1 : x - 1, x, x + 1 == D.X ? 0 : x + 1};
        const uint32_t yo[3] = {y == 0 ? D.Y - 1 : y - 1, y, y + 1 == D.Y ? 0 : y + 1};
        const uint32_t zo[3] = {z == 0 ? D.Z - 1 : z - 1, z, z + 1 == D.Z ? 0 : z + 1};
        const bool wx[3] = {bc.wall_lo[0] && xo[0] == 0, false, bc.wall_hi[0] && xo[2] == D.X - 1};
        const bool wy[3] = {bc.wall_lo[1] && yo[0] == 0, false, bc.wall_hi[1] && yo[2] == D.Y - 1};
        const bool wz[3] = {bc.wall_lo[2] && zo[0] == 0, false, bc.wall_hi[2] && zo[2] == D.Z - 1};
#pragma unroll
        for (int k = 1; k < Q; ++k) {
            const int cx = Lat<Q>::cx(k), cy = Lat<Q>::cy(k), cz = Lat<Q>::cz(k);
            if (wx[cx + 1] || wy[cy + 1] || wz[cz + 1]) {
                const int face = wall_face(D, bc, xo[cx + 1], yo[cy + 1], zo[cz + 1]);
                const T* ub = bc.vel[1 + face];
                const T cub = (T)cx * ub[0] + (T)cy * ub[1] + (T)cz * ub[2];
                f[k] = f[k] - (T)2 * w_scaled<Q, P>(k) * rr * cub * (T)3;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) a[(uint64_t)k * D.ps + c] = P::store_raw(f[k]);
}

// Kinds of batched launch.
enum BatchKind { kBPush = 0, kBPull = 1, kBGather = 2, kBCollide = 3 };

// The tiles of one launch: with `xcd`, XCD i % 8 takes a contiguous run of
// tiles (its own window of memory, neighbouring rows in its L2); without it
// workgroup i takes tile i (persistent blocks stride by the grid), so the chip
// sweeps one contiguous window.  fn(domain, cell) for every cell.
template <class T, class Fn>
__device__ inline void batch_tiles(const BArgs<T>& a, uint32_t t_begin, uint32_t t_end, Fn&& fn) {
    uint32_t lo, hi, stride = 1;
    const uint32_t span = t_end - t_begin;
    if (a.persistent) {
        if (a.xcd) {
            const uint32_t w = xcd_worker(blockIdx.x, gridDim.x);
            lo = t_begin + (uint32_t)((uint64_t)span * w / gridDim.x);
            hi = t_begin + (uint32_t)((uint64_t)span * (w + 1) / gridDim.x);
        } else {
            lo = t_begin + blockIdx.x;
            hi = t_end;
            stride = gridDim.x;
        }
    } else {
        lo = t_begin + (a.xcd ? xcd_worker(blockIdx.x, span) : blockIdx.x);
        hi = lo + 1;
    }
    for (uint32_t t = lo; t < hi; t += stride) {
        const uint2 te = a.tiles[t];
        const BDom<T> D = a.doms[te.x];
        for (uint32_t i = threadIdx.x; i < a.tile; i += blockDim.x) {
            const uint32_t c = te.y + i;
            if (c >= D.N) break;
            fn(D, c);
        }
    }
}

// One step (push or pull) or the gather, A -> B.
template <int Q, int V, int KIND>
__global__ __launch_bounds__(kMaxBlock) void k_batch(const typename Prec<V>::S* __restrict__ A,
                                                     typename Prec<V>::S* __restrict__ B,
                                                     BArgs<typename Prec<V>::T> a, uint32_t t_begin,
                                                     uint32_t t_end) {
    batch_tiles(a, t_begin, t_end, [&](const BDom<typename Prec<V>::T>& D, uint32_t c) {
        if (KIND == kBPush)
            batch_cell<Q, V>(A, B, D, a.bc, c);
        else if (KIND == kBPull)
            batch_pull_cell<Q, V, true>(A, B, D, a.bc, c);
        else
            batch_pull_cell<Q, V, false>(A, B, D, a.bc, c);
    });
}

// The collision alone, in place.
template <int Q, int V>
__global__ __launch_bounds__(kMaxBlock) void k_batch_collide(typename Prec<V>::S* A, BArgs<typename Prec<V>::T> a,
                                                             uint32_t t_begin, uint32_t t_end) {
    batch_tiles(a, t_begin, t_end,
                [&](const BDom<typename Prec<V>::T>& D, uint32_t c) { batch_collide_cell<Q, V>(A, D, a.bc, c); });
}

template <int Q, int V>
__global__ void k_batch_init(typename Prec<V>::S* __restrict__ f, const BDom<typename Prec<V>::T>* doms,
                             const uint2* __restrict__ tiles, uint32_t ntiles, uint32_t tile, double rho0,
                             double u0x, double u0y, double u0z, int kind, uint64_t seed, double amp) {
    using P = Prec<V>;
    const uint32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const uint2 te = tiles[t];
    const BDom<typename P::T> D = doms[te.x];
    for (uint32_t i = threadIdx.x; i < tile; i += blockDim.x) {
        const uint32_t c = te.y + i;
        if (c >= D.N) break;
        const int x = (int)(c % D.X), y = (int)((c / D.X) % D.Y), z = (int)(c / (D.X * D.Y));
        const double u0[3] = {u0x, u0y, u0z};
        double uc[3];
        cell_velocity(kind, seed + te.x, amp, D.X, D.Y, D.Z, u0, nullptr, x, y, z, uc);
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
            f[D.off + (uint64_t)k * D.ps + c] = s;
        }
    }
}

// rho (N) and u (3N, AoS) of one domain, double; walls report rho_s, u_b.
template <int Q, int V>
__global__ void k_batch_moments(const typename Prec<V>::S* __restrict__ f, BDom<typename Prec<V>::T> D,
                                FusedBC<double> bc, double rho_s, double* __restrict__ rho,
                                double* __restrict__ u) {
    using P = Prec<V>;
    using T = typename P::T;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.N) return;
    const uint32_t x = c % D.X, y = (c / D.X) % D.Y, z = c / (D.X * D.Y);
    int face = -1;
    const bool on[6] = {bc.wall_hi[0] && x == D.X - 1, bc.wall_lo[0] && x == 0, bc.wall_hi[1] && y == D.Y - 1,
                        bc.wall_lo[1] && y == 0, bc.wall_hi[2] && z == D.Z - 1, bc.wall_lo[2] && z == 0};
    for (int fc = 0; fc < 6; ++fc)
        if (on[fc]) face = fc;
    if (face >= 0) {
        rho[c] = rho_s;
        for (int i = 0; i < 3; ++i) u[3 * (uint64_t)c + i] = bc.vel[1 + face][i];
        return;
    }
    T fl[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) fl[k] = P::load(f[D.off + (uint64_t)k * D.ps + c]);
    T r, j0, j1, j2;
    moment_sums<Q, T>(fl, r, j0, j1, j2);
    if (P::shifted) r = (T)1 + r;
    const T ri = (T)1 / r;
    rho[c] = (double)r;
    u[3 * (uint64_t)c] = (double)(j0 * ri);
    u[3 * (uint64_t)c + 1] = (double)(j1 * ri);
    u[3 * (uint64_t)c + 2] = (double)(j2 * ri);
}

template <class Fn>
int with_qv_b(int q, int v, Fn&& fn) {
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

}  // namespace
}  // namespace plb

using namespace plb;

struct plb_batch {
    int q, variant, device, ndom;
    int map;     // tile order of a launch: 0 one window, 1 XCD-contiguous runs
    int scheme;  // PLB_SCHEME_*: pull (default) or push
    bool post;   // pull scheme: buffers hold post-collision populations
    std::vector<plb_domain_desc> desc;
    std::vector<uint64_t> off, ps;
    std::vector<uint32_t> tile_first;  // first tile of each domain (+ total at the end)
    plb_boundary bc;
    size_t sbytes;
    uint64_t pool_elems;
    void* buf[2];
    int cur;
    void* doms;       // BDom<T> array on the device
    uint2* tiles;     // tile table on the device
    uint32_t ntiles;
    int tile;         // cells per tile
    int grid;         // persistent grid blocks (0 = one block per tile)
    hipStream_t stream;
    hipGraphExec_t graph;
    int graph_steps, graph_parity;
    uint64_t total_cells, fluid_cells;
};

namespace plb {
namespace {

int batch_layout(plb_batch* b, int tile) {
    std::vector<uint2> t;
    b->tile_first.assign(b->ndom + 1, 0);
    for (int d = 0; d < b->ndom; ++d) {
        b->tile_first[d] = (uint32_t)t.size();
        const uint64_t n = (uint64_t)b->desc[d].x * b->desc[d].y * b->desc[d].z;
        for (uint64_t c = 0; c < n; c += tile) t.push_back(make_uint2((uint32_t)d, (uint32_t)c));
    }
    b->tile_first[b->ndom] = (uint32_t)t.size();
    if (b->tiles) PLB_HIP_TRY(hipFree(b->tiles));
    b->tiles = nullptr;
    PLB_HIP_TRY(hipMalloc(&b->tiles, t.size() * sizeof(uint2)));
    PLB_HIP_TRY(hipMemcpy(b->tiles, t.data(), t.size() * sizeof(uint2), hipMemcpyHostToDevice));
    b->ntiles = (uint32_t)t.size();
    b->tile = tile;
    return PLB_OK;
}

template <class T>
int upload_doms(plb_batch* b) {
    std::vector<BDom<T>> h(b->ndom);
    for (int d = 0; d < b->ndom; ++d) {
        const plb_domain_desc& s = b->desc[d];
        BDom<T>& o = h[d];
        o.X = s.x;
        o.Y = s.y;
        o.Z = s.z;
        o.N = (uint32_t)((uint64_t)s.x * s.y * s.z);
        o.dX = make_fastdiv((uint32_t)s.x);
        o.dY = make_fastdiv((uint32_t)s.y);
        o.off = b->off[d];
        o.ps = b->ps[d];
        o.tau = (T)s.tau;
        if (std::is_same<T, float>::value)
            o.om = (T)(1. / (double)o.tau);
        else
            o.om = (T)(1.0 / s.tau);
        o.omm = (T)1 - o.om;
        o.F0 = (T)s.force[0];
        o.F1 = (T)s.force[1];
        o.F2 = (T)s.force[2];
    }
    if (b->doms) PLB_HIP_TRY(hipFree(b->doms));
    b->doms = nullptr;
    PLB_HIP_TRY(hipMalloc(&b->doms, h.size() * sizeof(BDom<T>)));
    PLB_HIP_TRY(hipMemcpy(b->doms, h.data(), h.size() * sizeof(BDom<T>), hipMemcpyHostToDevice));
    return PLB_OK;
}

void drop_batch_graph(plb_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_steps = 0;
}

// One launch over tiles [t0, t1) (all domains, or one domain for the baseline).
int batch_launch(plb_batch* b, uint32_t t0, uint32_t t1, bool persistent_ok, int kind) {
    const int rc = with_qv_b(b->q, b->variant, [&](auto qc, auto vc) {
        constexpr int Q = decltype(qc)::value;
        constexpr int V = decltype(vc)::value;
        using P = Prec<V>;
        using T = typename P::T;
        using S = typename P::S;
        BArgs<T> a;
        a.doms = (const BDom<T>*)b->doms;
        a.tiles = b->tiles;
        a.ntiles = b->ntiles;
        a.tile = (uint32_t)b->tile;
        const uint32_t span = t1 - t0;
        const bool pers = persistent_ok && b->grid > 0 && (uint32_t)b->grid < span;
        a.persistent = pers ? 1u : 0u;
        a.xcd = b->map ? 1u : 0u;
        make_fused_bc<T>(b->bc, b->q, &a.bc);
        const unsigned threads = (unsigned)std::min(b->tile, kMaxBlock);
        const unsigned grid = pers ? (unsigned)b->grid : span;
        const S* A = (const S*)b->buf[b->cur];
        S* B = (S*)b->buf[b->cur ^ 1];
        switch (kind) {
            case kBPush:
                hipLaunchKernelGGL((k_batch<Q, V, kBPush>), dim3(grid), dim3(threads), 0, b->stream, A, B, a, t0, t1);
                break;
            case kBPull:
                hipLaunchKernelGGL((k_batch<Q, V, kBPull>), dim3(grid), dim3(threads), 0, b->stream, A, B, a, t0, t1);
                break;
            case kBGather:
                hipLaunchKernelGGL((k_batch<Q, V, kBGather>), dim3(grid), dim3(threads), 0, b->stream, A, B, a, t0,
                                   t1);
                break;
            default:  // in place on the current buffer
                hipLaunchKernelGGL((k_batch_collide<Q, V>), dim3(grid), dim3(threads), 0, b->stream,
                                   (S*)b->buf[b->cur], a, t0, t1);
                break;
        }
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    return rc;
}

bool batch_pull(const plb_batch* b) { return b->scheme != PLB_SCHEME_PUSH; }

// Pull-scheme storage between steps is post-collision with the streaming of
// the step pending (see fused_to_pre in fused.hip); the gather finishes it.
int batch_to_pre(plb_batch* b) {
    if (!b->post) return PLB_OK;
    const int rc = batch_launch(b, 0, b->ntiles, true, kBGather);
    if (rc) return rc;
    b->cur ^= 1;
    b->post = false;
    return PLB_OK;
}

// One step over tiles [t0, t1) of every domain range given, in the batch's
// scheme.  A pull-scheme step from pre-collision storage is the collision
// alone (in place); later pull steps gather, then collide.
template <class Launch>
int batch_one_step(plb_batch* b, Launch&& launch) {
    int rc;
    if (batch_pull(b)) {
        if (!b->post) {
            rc = launch(kBCollide);
            if (rc) return rc;
            b->post = true;
            return PLB_OK;
        }
        rc = launch(kBPull);
    } else {
        rc = batch_to_pre(b);
        if (rc) return rc;
        rc = launch(kBPush);
    }
    if (rc) return rc;
    b->cur ^= 1;
    return PLB_OK;
}

int batch_step_plain(plb_batch* b, int nsteps) {
    for (int i = 0; i < nsteps; ++i) {
        const int rc = batch_one_step(b, [&](int kind) { return batch_launch(b, 0, b->ntiles, true, kind); });
        if (rc) return rc;
    }
    return PLB_OK;
}

}  // namespace
}  // namespace plb

extern "C" {

int plb_batch_create(int q, int variant, int ndom, const plb_domain_desc* doms, const plb_boundary* bc,
                     int device, plb_batch** out) {
    if (!out || !doms || !bc || ndom <= 0) return fail(PLB_ERR_INVALID_ARGUMENT, "bad batch arguments");
    *out = nullptr;
    if (q != 9 && q != 19 && q != 27) return fail(PLB_ERR_LOGIC, "Model with q=" + std::to_string(q) + " not yet implemented");
    if (storage_bytes(variant) == 0) return fail(PLB_ERR_INVALID_ARGUMENT, "unknown storage variant");
    uint64_t pool = 0, total = 0, fluid = 0;
    std::vector<uint64_t> off(ndom), ps(ndom);
    for (int d = 0; d < ndom; ++d) {
        const plb_domain_desc& s = doms[d];
        if (s.x <= 0 || s.y <= 0 || s.z <= 0) return fail(PLB_ERR_INVALID_ARGUMENT, "argument is zero");
        // each factor is below 2^31, so neither product can wrap
        const uint64_t xy = (uint64_t)s.x * (uint64_t)s.y;
        if (xy >= (1ull << 31) || xy * (uint64_t)s.z >= (1ull << 31))
            return fail(PLB_ERR_INVALID_ARGUMENT, "domain too large for a batch");
        const uint64_t n = xy * (uint64_t)s.z;
        if (!(s.tau > 0.5)) return fail(PLB_ERR_INVALID_ARGUMENT, "tau must exceed 1/2");
        int rc = check_fused_bc(q, s.x, s.y, s.z, bc);
        if (rc) return rc;
        off[d] = pool;
        ps[d] = (n + 63) / 64 * 64;
        pool += ps[d] * q;  // below 2^31 * 27 per domain
        if (pool >= (1ull << 56)) return fail(PLB_ERR_INVALID_ARGUMENT, "batch too large");
        total += n;
        // wall layers of no-slip faces are not fluid (D2Q9 never moves in y)
        const uint64_t fx = s.x - (bc->type[0] == kNoslip ? 2 : 0);
        const uint64_t fy = s.y - (q != 9 && bc->type[2] == kNoslip ? 2 : 0);
        const uint64_t fz = s.z - (bc->type[4] == kNoslip ? 2 : 0);
        fluid += fx * fy * fz;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PLB_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(PLB_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    PLB_HIP_TRY(hipSetDevice(device));
    plb_batch* b = new plb_batch();
    b->q = q;
    b->variant = variant;
    b->device = device;
    b->ndom = ndom;
    b->desc.assign(doms, doms + ndom);
    b->off = off;
    b->ps = ps;
    b->bc = *bc;
    b->sbytes = storage_bytes(variant);
    b->pool_elems = pool;
    b->buf[0] = b->buf[1] = nullptr;
    b->cur = 0;
    b->doms = nullptr;
    b->tiles = nullptr;
    b->grid = 0;
    b->map = 0;
    b->scheme = PLB_SCHEME_AUTO;
    b->post = false;
    b->graph = nullptr;
    b->graph_steps = 0;
    b->graph_parity = 0;
    b->total_cells = total;
    b->fluid_cells = fluid;
    auto bail = [&](int code, const std::string& m) {
        plb_batch_destroy(b);
        return fail(code, m);
    };
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        delete b;
        return fail(PLB_ERR_HIP, "stream creation failed");
    }
    for (int i = 0; i < 2; ++i) {
        if (hipMalloc(&b->buf[i], pool * b->sbytes) != hipSuccess) return bail(PLB_ERR_OOM, "batch pool");
        if (hipMemset(b->buf[i], 0, pool * b->sbytes) != hipSuccess) return bail(PLB_ERR_HIP, "batch pool clear");
    }
    int rc = batch_layout(b, 256);
    if (rc) return bail(rc, plb_last_error());
    rc = variant == kF64 ? upload_doms<double>(b) : upload_doms<float>(b);
    if (rc) return bail(rc, plb_last_error());
    *out = b;
    return PLB_OK;
}

int plb_batch_destroy(plb_batch* b) {
    if (!b) return PLB_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    void* ptrs[] = {b->buf[0], b->buf[1], b->doms, b->tiles};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PLB_OK;
}

int plb_batch_initialize(plb_batch* b, double rho0, const double u0[3], const plb_init_spec* init) {
    if (!b || !u0) return fail(PLB_ERR_INVALID_ARGUMENT, "null argument");
    PLB_HIP_TRY(hipSetDevice(b->device));
    const int kind = init ? init->kind : PLB_INIT_UNIFORM;
    if (kind == PLB_INIT_TAYLOR_GREEN) return fail(PLB_ERR_INVALID_ARGUMENT, "Taylor-Green start is per-domain only");
    drop_batch_graph(b);
    const int rc = with_qv_b(b->q, b->variant, [&](auto qc, auto vc) {
        constexpr int Q = decltype(qc)::value;
        constexpr int V = decltype(vc)::value;
        using P = Prec<V>;
        hipLaunchKernelGGL((k_batch_init<Q, V>), dim3(b->ntiles), dim3(std::min(b->tile, kMaxBlock)), 0, b->stream,
                           (typename P::S*)b->buf[b->cur], (const BDom<typename P::T>*)b->doms, b->tiles,
                           b->ntiles, (uint32_t)b->tile, rho0, u0[0], u0[1], u0[2], kind,
                           init ? init->seed : 0, init ? init->amp : 0.0);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    if (rc) return rc;
    b->post = false;
    PLB_HIP_TRY(hipStreamSynchronize(b->stream));
    return PLB_OK;
}

int plb_batch_set_scheme(plb_batch* b, int scheme) {
    if (!b) return fail(PLB_ERR_INVALID_ARGUMENT, "null batch");
    if (scheme != PLB_SCHEME_AUTO && scheme != PLB_SCHEME_PUSH && scheme != PLB_SCHEME_PULL)
        return fail(PLB_ERR_INVALID_ARGUMENT, "unknown streaming scheme");
    PLB_HIP_TRY(hipSetDevice(b->device));
    drop_batch_graph(b);
    b->scheme = scheme;
    return PLB_OK;
}

int plb_batch_set_tiling(plb_batch* b, int tile_cells, int grid_blocks) {
    if (!b) return fail(PLB_ERR_INVALID_ARGUMENT, "null batch");
    if (tile_cells < 64 || tile_cells > 4096 || tile_cells % 64)
        return fail(PLB_ERR_INVALID_ARGUMENT, "tile_cells must be a multiple of 64 in [64, 4096]");
    if (grid_blocks < 0) return fail(PLB_ERR_INVALID_ARGUMENT, "negative grid");
    PLB_HIP_TRY(hipSetDevice(b->device));
    PLB_HIP_TRY(hipStreamSynchronize(b->stream));
    drop_batch_graph(b);
    b->grid = grid_blocks;
    if (tile_cells != b->tile) return batch_layout(b, tile_cells);
    return PLB_OK;
}

int plb_batch_get_tiling(const plb_batch* b, int* tile_cells, int* grid_blocks) {
    if (!b) return fail(PLB_ERR_INVALID_ARGUMENT, "null batch");
    if (tile_cells) *tile_cells = b->tile;
    if (grid_blocks) *grid_blocks = b->grid;
    return PLB_OK;
}

int plb_batch_set_mapping(plb_batch* b, int xcd_runs) {
    if (!b) return fail(PLB_ERR_INVALID_ARGUMENT, "null batch");
    if (xcd_runs != 0 && xcd_runs != 1) return fail(PLB_ERR_INVALID_ARGUMENT, "mapping must be 0 or 1");
    PLB_HIP_TRY(hipSetDevice(b->device));
    PLB_HIP_TRY(hipStreamSynchronize(b->stream));
    drop_batch_graph(b);
    b->map = xcd_runs;
    return PLB_OK;
}

int plb_batch_get_mapping(const plb_batch* b, int* xcd_runs) {
    if (!b || !xcd_runs) return fail(PLB_ERR_INVALID_ARGUMENT, "null argument");
    *xcd_runs = b->map;
    return PLB_OK;
}

int plb_batch_step(plb_batch* b, int nsteps, int steps_per_graph) {
    if (!b || nsteps < 0) return fail(PLB_ERR_INVALID_ARGUMENT, "bad arguments");
    PLB_HIP_TRY(hipSetDevice(b->device));
    if (steps_per_graph <= 0) return batch_step_plain(b, nsteps);
    if (steps_per_graph & 1) return fail(PLB_ERR_INVALID_ARGUMENT, "steps_per_graph must be even");
    // the captured steps keep one storage form: post-collision for the pull
    // step (from pre-collision storage its first step, the collision alone,
    // runs here), pre-collision for the push step
    if (batch_pull(b)) {
        if (!b->post && nsteps > 0) {
            const int rc = batch_step_plain(b, 1);
            if (rc) return rc;
            --nsteps;
        }
    } else {
        const int rc = batch_to_pre(b);
        if (rc) return rc;
    }
    const int reps = nsteps / steps_per_graph;
    if (reps > 0 && (!b->graph || b->graph_steps != steps_per_graph || b->graph_parity != b->cur)) {
        drop_batch_graph(b);
        PLB_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int lrc = batch_step_plain(b, steps_per_graph);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (lrc) {
            if (g) (void)hipGraphDestroy(g);
            return lrc;
        }
        if (e != hipSuccess) return fail(PLB_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        const hipError_t e2 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            b->graph = nullptr;
            return fail(PLB_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e2));
        }
        b->graph_steps = steps_per_graph;
        b->graph_parity = b->cur;
    }
    for (int r = 0; r < reps; ++r) PLB_HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    return batch_step_plain(b, nsteps - reps * steps_per_graph);
}

int plb_batch_step_per_domain(plb_batch* b, int nsteps) {
    if (!b || nsteps < 0) return fail(PLB_ERR_INVALID_ARGUMENT, "bad arguments");
    PLB_HIP_TRY(hipSetDevice(b->device));
    for (int i = 0; i < nsteps; ++i) {
        const int rc = batch_one_step(b, [&](int kind) {
            for (int d = 0; d < b->ndom; ++d) {
                const int lrc = batch_launch(b, b->tile_first[d], b->tile_first[d + 1], false, kind);
                if (lrc) return lrc;
            }
            return PLB_OK;
        });
        if (rc) return rc;
    }
    return PLB_OK;
}

int plb_batch_autotune(plb_batch* b, int steps, double* results, int max_results, int* nresults) {
    if (!b || steps <= 0) return fail(PLB_ERR_INVALID_ARGUMENT, "bad arguments");
    PLB_HIP_TRY(hipSetDevice(b->device));
    int cu = 256;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, b->device) == hipSuccess) cu = p.multiProcessorCount;
    const int tiles[] = {64, 128, 256, 512, 1024, 2048};
    const int grids[] = {0, 2 * cu, 4 * cu, 8 * cu, 16 * cu};
    const int t_keep = b->tile, g_keep = b->grid, m_keep = b->map;
    double best = 1e30;
    int best_t = t_keep, best_g = g_keep, best_m = m_keep, n = 0;
    hipEvent_t e0, e1;
    PLB_HIP_TRY(hipEventCreate(&e0));
    PLB_HIP_TRY(hipEventCreate(&e1));
    int rc = PLB_OK;
    // every candidate runs an even number of steps (the total the caller sees
    // is even; the buffer parity is tracked in b->cur either way)
    const int even = steps + (steps & 1);
    for (int t : tiles) {
        for (int g : grids) {
          for (int m : {0, 1}) {
            rc = plb_batch_set_tiling(b, t, g);
            if (rc) break;
            b->map = m;
            rc = batch_step_plain(b, 2);
            if (rc) break;
            hipError_t e = hipEventRecord(e0, b->stream);
            if (e == hipSuccess) rc = batch_step_plain(b, even);
            if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) rc = fail(PLB_ERR_HIP, hipGetErrorString(e));
            if (rc) break;
            const double per = (double)ms / even;
            if (results && n < max_results) {
                results[4 * n] = t;
                results[4 * n + 1] = g;
                results[4 * n + 2] = m;
                results[4 * n + 3] = per;
            }
            ++n;
            if (per < best) {
                best = per;
                best_t = t;
                best_g = g;
                best_m = m;
            }
          }
          if (rc) break;
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (nresults) *nresults = std::min(n, max_results);
    b->map = best_m;
    return plb_batch_set_tiling(b, best_t, best_g);
}

int plb_batch_download_moments(plb_batch* b, int d, double* rho, double* u) {
    if (!b || !rho || !u || d < 0 || d >= b->ndom) return fail(PLB_ERR_INVALID_ARGUMENT, "bad arguments");
    PLB_HIP_TRY(hipSetDevice(b->device));
    const int prc = batch_to_pre(b);
    if (prc) return prc;
    const uint64_t n = (uint64_t)b->desc[d].x * b->desc[d].y * b->desc[d].z;
    double* dev = nullptr;
    PLB_HIP_TRY(hipMalloc(&dev, n * 4 * sizeof(double)));
    FusedBC<double> bc;
    make_fused_bc<double>(b->bc, b->q, &bc);
    int rc = with_qv_b(b->q, b->variant, [&](auto qc, auto vc) {
        constexpr int Q = decltype(qc)::value;
        constexpr int V = decltype(vc)::value;
        using P = Prec<V>;
        using T = typename P::T;
        BDom<T> D;
        D.X = b->desc[d].x;
        D.Y = b->desc[d].y;
        D.Z = b->desc[d].z;
        D.N = (uint32_t)n;
        D.off = b->off[d];
        D.ps = b->ps[d];
        D.tau = D.om = D.omm = D.F0 = D.F1 = D.F2 = (T)0;
        hipLaunchKernelGGL((k_batch_moments<Q, V>), dim3(blocks_for(n, 256)), dim3(256), 0, b->stream,
                           (const typename P::S*)b->buf[b->cur], D, bc, b->bc.rho_s, dev, dev + n);
        PLB_HIP_TRY(hipGetLastError());
        return PLB_OK;
    });
    hipError_t e = hipSuccess;
    if (rc == PLB_OK) e = hipMemcpyAsync(rho, dev, n * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    if (rc == PLB_OK && e == hipSuccess)
        e = hipMemcpyAsync(u, dev + n, 3 * n * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    const hipError_t e2 = hipStreamSynchronize(b->stream);
    (void)hipFree(dev);
    if (rc) return rc;
    if (e != hipSuccess) return fail(PLB_ERR_HIP, hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(PLB_ERR_HIP, hipGetErrorString(e2));
    return PLB_OK;
}

int plb_batch_cells(const plb_batch* b, uint64_t* total_cells, uint64_t* fluid_cells) {
    if (!b) return fail(PLB_ERR_INVALID_ARGUMENT, "null batch");
    if (total_cells) *total_cells = b->total_cells;
    if (fluid_cells) *fluid_cells = b->fluid_cells;
    return PLB_OK;
}

int plb_batch_synchronize(plb_batch* b) {
    if (!b) return fail(PLB_ERR_INVALID_ARGUMENT, "null batch");
    PLB_HIP_TRY(hipSetDevice(b->device));
    PLB_HIP_TRY(hipStreamSynchronize(b->stream));
    return PLB_OK;
}

}  // extern "C"
