// Internal state of a plbpd device domain (the device side of SimData) and the
// helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/plbpd_hip.h"
#include "lattice.hpp"

namespace plb {

// ------------------------------------------------------------------ errors --
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PLB_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::plb::fail(e_ == hipErrorOutOfMemory ? PLB_ERR_OOM : PLB_ERR_HIP,   \
                               std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

// ------------------------------------------------------------- boundaries --
enum { kUnset = 0, kPeriodic = 1, kNoslip = 2 };

// Per-cell flag in the fused layout: 0 fluid, 1 + face for the wall layer of a
// no-slip face (the last face in reference order wins), 7 interior solid.
enum { kFlagFluid = 0, kFlagSolid = 7 };

// Fused-kernel view of a boundary: per axis periodic or walled.
template <class T>
struct FusedBC {
    int32_t periodic[3];
    int32_t wall_lo[3], wall_hi[3];
    int32_t moving;     // any wall with nonzero velocity
    T vel[8][3];        // wall velocity by flag value
};

// ------------------------------------------------------------ timers -------
struct TimerPair {
    int stage;
    hipEvent_t start, stop;
};

// ------------------------------------------------ deferred literal calls ---
enum LitOpKind { kLitMoments, kLitResetVelocity, kLitResetDensity, kLitCollide };
struct LitOp {
    LitOpKind kind;
    double tau;        // kLitCollide
    double force[3];   // kLitCollide
    plb_boundary bc;   // the resets
};

}  // namespace plb

struct plb_state {
    int q, X, Y, Z;
    int variant, mode, device;
    size_t ncell;    // literal: halo'd cells; fused: compact cells
    size_t pstride;  // elements between population arrays
    size_t sbytes;   // bytes per stored population value
    void* buf[2];    // population buffers
    int cur;         // buf[cur] holds n (SimData::n), buf[cur ^ 1] ntmp
    void* rho;       // literal: R[ncell]; fused: double[ncell]
    void* u;         // literal: R[3][ncell]; fused: double[3][ncell]
    uint8_t* flags;  // fused: per-cell flag
    uint32_t* nbmask;  // fused with interior solids: bit k = neighbour k solid
    bool has_solids;
    std::vector<uint8_t> mask_host;  // interior-solid mask kept to rebuild flags
    bool bc_set;
    plb_boundary bc;
    double tau;
    double force[3];
    bool configured;
    bool collide_pending;      // fused: the next fused launch collides first
    // fused: one collide+stream step that plb_stream recorded without
    // launching it; the next plb_stream runs it with its own step as one
    // two-step launch, and any call that needs it done runs it first
    bool fused_deferred;
    // literal: calls deferred so that a following push can fold them into
    // its kernel (at most moments, reset velocity, reset density, collide, in
    // that order); any other call runs them first, in order
    std::vector<plb::LitOp> lit_queue;
    // literal: the last collide+push kernel stored its in-place values on the
    // outer layer only; ntmp off that layer is n(c + c_k, k) until restored
    bool lit_ntmp_folded;
    bool moments_valid;
    void* scratch;  // reductions, face staging
    size_t scratch_bytes;
    hipStream_t stream;
    // graph of a fixed number of fused steps
    hipGraphExec_t graph;
    int graph_steps;
    int graph_parity;
    // tuning knobs of the fused kernel (plb_set_tuning)
    int tune_tx;  // block width in x (64, 128, 256); 0 = pick from X
    int tune_nt_load, tune_nt_store;
    int tune_scheme;      // PLB_SCHEME_*
    int tune_order;       // tile-row order: 0 linear, B: B x B (y, z) blocks
    int tune_xcd;         // tile mapping: 0 block i -> tile i, 1 XCD runs, 2 XCD y-stripes, -1 auto
    int tune_index64;     // 1: 64-bit index kernels even where 32-bit offsets fit
    int tune_vector;      // row pull kernel cells per lane (1, 2, 4) where X allows; 0 off; -1 auto
    int tune_lit_fuse;    // literal: 1 collision and push in one kernel when they follow each other
    int tune_two;         // fused pull: two steps per launch where a kernel exists; 0 off, 1 on, -1 auto (4- and 8-byte storage)
    int tune_two_zc;      // planes per z-chunk of the two-step launch; 0 = chosen from the grid
    int tune_two_bx;      // two-step tile width: 32 or 64 cells; 0 = automatic
    int num_cu;           // compute units of the state's device
    bool post_collision;  // fused: buf[cur] holds post-collision populations (pull step)
    // one-array state (PLB_MODE_FUSED_AA): buf[cur] only; aa_odd = after an
    // even step, f*_k swapped into slot rev(k); aa_collide = the step in
    // progress collides
    bool aa, aa_odd, aa_collide;
    // timers
    bool timers_on;
    std::vector<plb::TimerPair> pending;
    std::vector<hipEvent_t> free_events;
    double stage_ms[PLB_NSTAGES];
};

namespace plb {

// timer helpers (capi.hip)
int timer_begin(plb_state* s, int stage, TimerPair* tp);
int timer_end(plb_state* s, TimerPair* tp);

// fused mode: two pull steps in one launch (fused.hip).  Eligible when the
// state steps with the pull scheme and a kernel exists for its lattice,
// storage and row length; the launch needs post-collision storage.
bool fused_two_eligible(const plb_state* s);
// Runs the step plb_stream deferred, if any (capi.hip).
int fused_flush_deferred(plb_state* s);
// The two-step tile width for the state (32 or 64 cells), or 0 where none applies.
int fused_two_tile_x(const plb_state* s);
int fused_launch_two(plb_state* s);

// literal mode (literal.hip)
int lit_initialize(plb_state* s, double rho0, const double u0[3], const plb_init_spec* init,
                   const double* tg_dev);
int lit_collide(plb_state* s, double tau, const double F[3]);
int lit_push(plb_state* s);
int lit_collide_push(plb_state* s, double tau, const double F[3]);
// bc_vel: the faces whose velocity is reset; bc_rho: those whose density is
int lit_moments_collide_push(plb_state* s, double tau, const double F[3], const plb_boundary* bc_vel,
                             const plb_boundary* bc_rho);
int lit_noslip(plb_state* s, const plb_boundary* bc);
// The same pass staged in the library's static device buffer: no allocation.
int lit_noslip_noalloc(plb_state* s, const plb_boundary* bc);
int lit_periodic(plb_state* s, const plb_boundary* bc);
int lit_moments(plb_state* s);
int lit_reset_velocity(plb_state* s, const plb_boundary* bc);
int lit_reset_density(plb_state* s, const plb_boundary* bc);
int lit_transfer(plb_state* s, int field, int layout, int dtype, void* host, bool upload);
int lit_global(plb_state* s, double out[4]);

// fused mode (fused.hip)
int fused_initialize(plb_state* s, double rho0, const double u0[3], const plb_init_spec* init,
                     const double* tg_dev);
int fused_build_flags(plb_state* s, const plb_boundary* bc, const uint8_t* mask);
int fused_launch(plb_state* s, bool collide);
bool fused_uses_pull(const plb_state* s);
int fused_to_pre(plb_state* s);
int fused_moments(plb_state* s);
int fused_transfer(plb_state* s, int field, int layout, int dtype, void* host, bool upload);
int fused_download_raw(plb_state* s, void* host, size_t bytes);
int fused_global(plb_state* s, double out[4]);
template <class T>
void make_fused_bc(const plb_boundary& bc, int q, FusedBC<T>* out);

// Validates a boundary for fused semantics (each moving axis periodic on both
// sides or no-slip on both sides; no-slip faces need a velocity).
int check_fused_bc(int q, int X, int Y, int Z, const plb_boundary* bc);

// shared initial-state helpers (capi.hip): Taylor-Green per-axis tables
// sx[X] cx[X] sy[Y] cy[Y] sz[Z] cz[Z], computed on the host with libm.
std::vector<double> tg_tables(int X, int Y, int Z);

// Counter-based uniform in [0,1): splitmix64 finaliser (Steele, Lea, Flood 2014).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__host__ __device__ inline double unit_uniform(uint64_t seed, uint64_t cell, int comp) {
    const uint64_t h = mix64(mix64(seed) + 3ull * cell + (uint64_t)comp);
    return (double)(h >> 11) * 0x1.0p-53;
}

// Initial velocity of interior cell (xi, yi, zi) (0-based), PLB_INIT_* kinds.
__host__ __device__ inline void cell_velocity(int kind, uint64_t seed, double amp, int X, int Y,
                                              int Z, const double* u0, const double* tg, int xi,
                                              int yi, int zi, double u[3]) {
    if (kind == PLB_INIT_PERTURBED) {
        const uint64_t cell = ((uint64_t)zi * Y + yi) * X + xi;
        for (int i = 0; i < 3; ++i) u[i] = u0[i] + amp * (2.0 * unit_uniform(seed, cell, i) - 1.0);
    } else if (kind == PLB_INIT_TAYLOR_GREEN) {
        const double* sx = tg;
        const double* cx = sx + X;
        const double* cy = cx + X + Y;
        const double* sz = cy + Y;
        const double* cz = sz + Z;
        u[0] = ((u0[0] * sx[xi]) * cz[zi]) * cy[yi];
        u[1] = 0.0;
        u[2] = ((-u0[0] * cx[xi]) * sz[zi]) * cy[yi];
    } else {
        u[0] = u0[0];
        u[1] = u0[1];
        u[2] = u0[2];
    }
}

// EqlbDist::operator() (EqlbDist.hpp:26-38) in R: usq and cu in R, the bracket
// in double, w*rho in R, result rounded to R.
template <int Q, class R>
__host__ __device__ inline R feq_ref(int k, R rho, const R u[3], R usq) {
    const R cu = (R)Lat<Q>::cx(k) * u[0] + (R)Lat<Q>::cy(k) * u[1] + (R)Lat<Q>::cz(k) * u[2];
    const R wr = (R)Lat<Q>::w(k) * rho;
    return (R)((double)wr * (1.0 + 3.0 * (double)cu + 4.5 * (double)cu * (double)cu -
                             1.5 * (double)usq));
}

inline unsigned blocks_for(size_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }

// Grid of a grid-stride kernel over n items: one item per thread up to 2^20
// blocks, more items per thread beyond (a launch holds fewer than 2^32
// work-items, and a layout transfer of a large state has more values).
inline unsigned stride_blocks(uint64_t n, unsigned threads) {
    const uint64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < (1u << 20) ? (b ? b : 1) : (1u << 20));
}

}  // namespace plb
