/* plbpd_hip.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann BGK step.
 *
 * This is the drop-in boundary behind plbpd's operator classes.  Each entry
 * point below names the reference interface it replaces (paths relative to
 * the plbpd source tree).  Plain C: opaque handles, plain pointers and sizes,
 * int status codes (0 = success).  No C++ exception crosses this boundary;
 * plb_last_error() returns the message of the last failure on the calling
 * thread, and the status code says which exception the C++ wrappers rethrow:
 *
 *   PLB_ERR_INVALID_ARGUMENT  std::invalid_argument (Field.hpp:28-31, Flow.cpp:15)
 *   PLB_ERR_LOGIC             std::logic_error (LBModel.cpp:17, Streaming.cpp:18-24,
 *                             96-102, Boundary.cpp:146, SimData.cpp:162-164)
 *   PLB_ERR_OUT_OF_RANGE      std::out_of_range (Boundary.cpp:197, _velocity.at)
 *   PLB_ERR_HIP / _NO_DEVICE / _OOM   device failures (no reference counterpart)
 *
 * Work is queued on one HIP stream per state; results are observable only
 * after plb_download / plb_global_moments / plb_synchronize.
 */
#ifndef PLBPD_HIP_H_
#define PLBPD_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLB_OK 0
#define PLB_ERR_INVALID_ARGUMENT 1
#define PLB_ERR_LOGIC 2
#define PLB_ERR_OUT_OF_RANGE 3
#define PLB_ERR_HIP 4
#define PLB_ERR_NO_DEVICE 5
#define PLB_ERR_OOM 6

/* Storage variants.  The first type is what populations are stored as in
 * HBM, the second the collision arithmetic.  Shifted variants store
 * g = f - w_k; F16S stores half(g * 2^15). */
#define PLB_F64 0     /* f64 / f64 : the FP64 contract                     */
#define PLB_F32 1     /* f32 / f32 (literal mode: the reference's rounding) */
#define PLB_F64_A32 2 /* f64 / f32                                          */
#define PLB_F32S 3    /* f32 shifted / f32                                  */
#define PLB_F16S 4    /* f16 shifted, scaled / f32                          */

/* Step semantics.  LITERAL reproduces the reference's staged step
 * (PLBPD.cpp:23-30) on halo'd arrays, stale seam reads included.  FUSED runs
 * one collide+push kernel per step with periodic wraps and halfway
 * bounce-back folded in through a per-cell flag (no stale reads). */
#define PLB_MODE_LITERAL 0
#define PLB_MODE_FUSED 1
/* FUSED semantics with one population array instead of two (the AA pattern:
 * steps alternate between an in-place local step and an in-place
 * gather/scatter step): half the population memory, the same numbers.
 * Reading the state after an odd number of steps gathers it into a
 * temporary second array first.  Its scheme is PLB_SCHEME_AA. */
#define PLB_MODE_FUSED_AA 2

/* Face order of Boundary.cpp:17-18. */
#define PLB_EAST 0
#define PLB_WEST 1
#define PLB_NORTH 2
#define PLB_SOUTH 3
#define PLB_UP 4
#define PLB_DOWN 5
#define PLB_FACE_UNSET 0
#define PLB_FACE_PERIODIC 1
#define PLB_FACE_NOSLIP 2

#define PLB_FIELD_RHO 0
#define PLB_FIELD_U 1
#define PLB_FIELD_N 2
#define PLB_FIELD_NTMP 3

/* PLB_LAYOUT_REFERENCE: the reference's AoS arrays with one halo layer,
 * index ((z*(Y+2) + y)*(X+2) + x)*vlen + v (Field.hpp:91-93).
 * PLB_LAYOUT_COMPACT: the same AoS order without the halo. */
#define PLB_LAYOUT_REFERENCE 0
#define PLB_LAYOUT_COMPACT 1

#define PLB_DTYPE_F64 0
#define PLB_DTYPE_F32 1

#define PLB_INIT_UNIFORM 0      /* Domain::initialize (Domain.cpp:46-64)          */
#define PLB_INIT_PERTURBED 1    /* u0 + amp*U(-1,1), splitmix64(seed, cell, comp) */
#define PLB_INIT_TAYLOR_GREEN 2 /* x-z Taylor-Green vortex of amplitude u0[0]      */

#define PLB_NSTAGES 8
#define PLB_STAGE_COLLIDE 0
#define PLB_STAGE_STREAM 1
#define PLB_STAGE_NOSLIP 2
#define PLB_STAGE_PERIODIC 3
#define PLB_STAGE_MOMENTS 4
#define PLB_STAGE_RESET_VELOCITY 5
#define PLB_STAGE_RESET_DENSITY 6
#define PLB_STAGE_FUSED 7

typedef struct plb_state plb_state;
typedef struct plb_batch plb_batch;

typedef struct {
    int32_t q;        /* 9 (D2Q9, x-z plane), 19 or 27 */
    int32_t x, y, z;  /* interior extents, Domain::get_dimensions order (Domain.cpp:26-28) */
    int32_t variant;  /* PLB_F64 ... PLB_F16S */
    int32_t mode;     /* PLB_MODE_LITERAL or PLB_MODE_FUSED */
    int32_t device;   /* HIP device ordinal */
    int32_t pop_pad;  /* fused layout: extra elements between population arrays; <0 = default */
} plb_state_desc;

/* Boundary(lbmodel, domain, solid_density, BoundaryType, BoundaryVelocity)
 * (Boundary.hpp:55-60): the unordered maps become per-face arrays. */
typedef struct {
    int32_t type[6];    /* PLB_FACE_* per face */
    int32_t has_vel[6]; /* face present in BoundaryVelocity */
    double vel[6][3];
    double rho_s;       /* solid_density */
} plb_boundary;

typedef struct {
    int32_t kind; /* PLB_INIT_* */
    uint64_t seed;
    double amp;
} plb_init_spec;

/* ---- library ------------------------------------------------------------ */
/* Message of the last failure on the calling thread (the what() of the
 * exception the reference would throw). */
const char* plb_last_error(void);
/* Library version (100 = 1.0.0).  No reference counterpart. */
int plb_version(void);
/* Device name, CU count, HBM bytes of the device the library would use.  No
 * reference counterpart. */
int plb_device_info(int device, char* name, size_t name_len, int* num_cu, size_t* hbm_bytes);

/* LBModel::get_directional_velocities / get_directional_weights / get_reverse
 * (LBModel.hpp:29-34, LBModel.cpp:23-165); unknown q -> PLB_ERR_LOGIC
 * (LBModel.cpp:17).  c: q*3 int32, w: q doubles, rev: q uint32. */
int plb_lattice_tables(int q, int32_t* c, double* w, uint32_t* rev);

/* ---- state: SimData (SimData.hpp:9-31) on the device ---------------------- */
/* SimData(domain_dimensions, num_directions) (SimData.cpp:22-30): allocates
 * the device population buffers (A/B, the n/ntmp pair of SimData.hpp:20-21),
 * rho and u.  A zero extent -> PLB_ERR_INVALID_ARGUMENT (Field.hpp:28-31).
 * Sizes no device holds are refused by size alone (PLB_ERR_INVALID_ARGUMENT,
 * "domain too large"), before any device call: fused states take fewer than
 * 2^32 cells, literal states fewer than 2^40 halo'd cells, with halo'd x and
 * y edges below 2^31. */
int plb_state_create(const plb_state_desc* desc, plb_state** out);
int plb_state_destroy(plb_state* s);
/* Bytes of device memory the state holds.  No reference counterpart. */
int plb_state_bytes(const plb_state* s, size_t* bytes);

/* Host <-> device copies of SimData's fields (SimData.hpp:18-21).  In fused
 * mode the halo of PLB_LAYOUT_REFERENCE is left untouched on download and
 * ignored on upload; NTMP is not available in fused mode. */
int plb_upload(plb_state* s, int field, int layout, int dtype, const void* host);
int plb_download(plb_state* s, int field, int layout, int dtype, void* host);
/* Current populations in stored form (compact AoS, storage element type),
 * for bit-level comparison of reduced-precision storage.  No reference
 * counterpart (the reference stores float only). */
int plb_download_raw(plb_state* s, void* host, size_t bytes);

/* ---- set-up ------------------------------------------------------------- */
/* PLBPD::initialize = Domain::initialize + Boundary::reset (PLBPD.cpp:16-19,
 * Domain.cpp:46-64, Boundary.cpp:58-89).  bc may be NULL (no reset). */
int plb_domain_initialize(plb_state* s, double rho0, const double u0[3],
                          const plb_init_spec* init, const plb_boundary* bc);
/* Boundary set-up for the fused kernel: flags wall layers of no-slip faces and
 * validates periodic pairs.  mask: optional x*y*z bytes (nonzero = interior
 * solid, compact order) or NULL.  No reference counterpart for the mask. */
int plb_set_boundary(plb_state* s, const plb_boundary* bc, const uint8_t* mask);

/* ---- one entry per reference call of PLBPD::march_in_time (PLBPD.cpp:24-29) */
/* CollisionSRT::operator() (CollisionSRT.cpp:25-33, 153-176).  tau and the
 * external force F are passed directly (tau = 3 nu + 1/2 in the reference,
 * CollisionSRT.cpp:19).  Fused mode: defers to the next plb_stream.  Literal
 * mode (PLB_TUNE_LITERAL_FUSE on, the default): the collision runs with the
 * next call on the state -- as one kernel with the push when that call is
 * plb_stream(s, "push"), else before that call does anything -- so every
 * value a caller can observe is the one the immediate collision gives; a
 * launch failure of the collision is then reported by that next call. */
int plb_collide_srt(plb_state* s, double tau, const double force[3]);
/* Streaming::operator()(simdata, stream_type) (Streaming.cpp:31-42, 63-94):
 * "push" streams n into ntmp and swaps them; "pull" and anything else ->
 * PLB_ERR_LOGIC as in the reference (Streaming.cpp:17-23, 96-102), after the
 * literal calls deferred before it have run.  A literal "push" that follows
 * a deferred collision runs it in the same kernel (with the deferred moments
 * and resets of the step before, see plb_calc_moments).  Fused
 * mode: runs the fused kernel for "push" or "pull" (the pull streaming the
 * reference leaves unimplemented; both give the same numbers); "pull" while
 * PLB_TUNE_SCHEME forces push -> PLB_ERR_LOGIC. */
int plb_stream(plb_state* s, const char* stream_type);
/* Boundary::apply_noslip (Boundary.cpp:101-111, 189-221). */
int plb_apply_noslip(plb_state* s, const plb_boundary* bc);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  plb_apply_noslip on
 * a literal state (the same values) with its staging in a static device
 * buffer of 16 MiB that every GPU the library runs on holds until the
 * process ends, instead of the state's scratch buffer, which plb_apply_noslip
 * allocates on first use.  Calls no hipMalloc or hipFree and creates no
 * event: the calls deferred before it and the pass itself are not timed.
 * Each no-slip face needs (face cells) * (q - 1) * sizeof(value) bytes; a
 * larger face -> PLB_ERR_INVALID_ARGUMENT before anything runs.  Fused
 * states -> PLB_ERR_LOGIC.  Every call on one GPU shares the buffer, so calls
 * that overlap on one GPU (states on different streams, other threads) race;
 * nothing in it needs to start at zero. */
int plb_apply_noslip_noalloc(plb_state* s, const plb_boundary* bc);
/* Boundary::apply_periodicity (Boundary.cpp:91-99, 223-281). */
int plb_apply_periodicity(plb_state* s, const plb_boundary* bc);
/* CalcMoments::operator() (CalcMoments.hpp:46-73).  Literal mode with
 * PLB_TUNE_LITERAL_FUSE on: the moments, and the resets that directly follow
 * them, run with the next call on the state -- folded into the collide+push
 * kernel when that sequence is collide then plb_stream(s, "push"), else
 * before the call does anything -- so every observable value is unchanged. */
int plb_calc_moments(plb_state* s);
/* Boundary::reset / reset_velocity / reset_density (Boundary.cpp:58-89).
 * Deferred in literal mode directly after plb_calc_moments, as described
 * there. */
int plb_reset(plb_state* s, const plb_boundary* bc);
int plb_reset_velocity(plb_state* s, const plb_boundary* bc);
int plb_reset_density(plb_state* s, const plb_boundary* bc);

/* ---- fused / whole-loop entry points -------------------------------------- */
/* Parameters used by plb_step: tau, force and boundary. */
int plb_configure(plb_state* s, double tau, const double force[3], const plb_boundary* bc);
/* nsteps of PLBPD::march_in_time's loop body (PLBPD.cpp:23-30).  Literal
 * mode: the six staged calls; fused mode: one kernel per step. */
int plb_step(plb_state* s, int nsteps);
/* As plb_step, with the loop captured in a HIP graph of steps_per_graph
 * (even) steps, replayed; remainder steps run as plain launches. */
int plb_step_graph(plb_state* s, int nsteps, int steps_per_graph);
/* FP64 global reductions over fluid cells: out = {sum rho, sum rho ux,
 * sum rho uy, sum rho uz}.  No reference counterpart. */
int plb_global_moments(plb_state* s, double out[4]);
int plb_synchronize(plb_state* s);

/* ---- timers: the static per-class timers of the reference (Collision.cpp:3,
 * Streaming.cpp:12, Boundary.cpp:12-16, CalcMoments.cpp:5), measured with HIP
 * events on the state's stream.  ms_out: PLB_NSTAGES accumulated values.  A
 * literal collision run together with its push is counted under
 * PLB_STAGE_COLLIDE; PLB_STAGE_STREAM counts the pushes run alone.  Moments
 * and resets folded into the next step's kernel (inside a literal plb_step,
 * or deferred as plb_calc_moments describes) are counted under
 * PLB_STAGE_COLLIDE too. */
int plb_enable_timers(plb_state* s, int on);
int plb_stage_times(plb_state* s, double* ms_out);

/* ---- measurement (no reference counterpart; the reference times its loop
 * with tbb::tick_count, PLBPD.cpp:22, 31) ---------------------------------- */
/* Runs nsteps as plb_step (or plb_step_graph when steps_per_graph > 0)
 * between two HIP events on the state's stream; ms = elapsed time. */
int plb_time_steps(plb_state* s, int nsteps, int steps_per_graph, double* ms);
/* Copy-kernel calibration: copies `bytes` device-to-device `iters` times;
 * gbps counts bytes read plus bytes written. */
int plb_bench_copy(int device, size_t bytes, int iters, double* gbps, double* ms_per_copy);
/* Memory-pattern ceiling of the fused step: the loads and stores of one
 * step over a q-population x*y*z lattice (q 19 or 27) in the fused layout,
 * with sbytes-wide raw values (2, 4, 8) and vec consecutive cells per lane
 * (1, 2, 4; x a multiple of vec), without the collision.  pattern 0: a
 * q-stream copy (every population read and written at its own cell);
 * 1: the row pull kernel's loads (population k from row y - c_y, z - c_z)
 * and flag bytes; 2: pattern 1 with the kernel's loads around interior
 * solids (the cell's own reverse slots and a 4-byte neighbour mask).  gbps
 * counts the step's algorithmic bytes, x*y*z*(2*q*sbytes + 1), over the
 * mean launch time. */
int plb_bench_pattern(int device, int q, int x, int y, int z, int sbytes, int vec, int pattern, int iters,
                      double* gbps, double* ms);
/* Fused-kernel tuning knobs (no reference counterpart).  The defaults are
 * the settings measured fastest on MI355X for D3Q19 256^3; no knob changes
 * results; a change drops a captured step graph. */
#define PLB_TUNE_BLOCK_X 0  /* cells per block row: 64, 128, 256 (0 = from x)   */
#define PLB_TUNE_NT_LOAD 1  /* non-temporal population loads of the push kernel */
                            /* (0/1; default 0).  The cell pull kernel's loads   */
                            /* are ordinary (its shifted lines are shared with   */
                            /* the next wavefront); the row pull kernel's are    */
                            /* non-temporal without interior solids.            */
#define PLB_TUNE_NT_STORE 2 /* non-temporal population stores (0/1; default 1) */
#define PLB_TUNE_SCHEME 3   /* streaming scheme of the fused step:              */
#define PLB_SCHEME_AUTO 0   /*   pull (default; moving walls included)         */
#define PLB_SCHEME_PUSH 1   /*   collide, then scatter to neighbours            */
#define PLB_SCHEME_PULL 2   /*   gather from neighbours, then collide           */
#define PLB_SCHEME_AA 3     /*   one array (PLB_MODE_FUSED_AA states only)      */
#define PLB_SCHEME_PULL_TWO 4 /* reported only: pull, two steps per launch       */
                              /* (chosen by PLB_TUNE_TWO_STEP)                   */
#define PLB_TUNE_TILE_ORDER 4 /* rows in B x B blocks of the y-z plane (B a power */
                              /* of two dividing the row counts; default 0 =      */
                              /* linear)                                          */
#define PLB_TUNE_VECTOR 5     /* row pull kernel with 1, 2 or 4 aligned cells per */
                              /* lane (x = 64, 128, 256, 384, 512 for 1; 128,     */
                              /* 256, 384, 512 for 2; 256, 512 for 4); 0 =        */
                              /* cell-per-thread pull kernel, also                */
                              /* where the row kernel has no shape for x; -1      */
                              /* (default) = for 4- and 8-byte storage 1 without  */
                              /* interior solids, around them 4 (4-byte) or 2     */
                              /* (8-byte, and every D3Q27 state), else the next   */
                              /* width with a shape for x; the cell kernel for    */
                              /* 2-byte storage                                   */
#define PLB_TUNE_XCD 6        /* tile mapping: 0 workgroup i takes tile i; 1 each */
                              /* XCD sweeps its own contiguous run of tiles; 2    */
                              /* each XCD takes a y-stripe of every z-plane (rows */
                              /* % 8 == 0, else 0); -1 (default) = 2 for the cell */
                              /* kernels around interior solids, else 0.  The     */
                              /* two-step kernel: 0 workgroup i takes tile i;     */
                              /* otherwise (default) each XCD takes one of eight  */
                              /* contiguous runs of its (z-chunk, y-tile) pairs   */
#define PLB_TUNE_INDEX64 7    /* 1: the cell kernels' 64-bit index instantiations, and */
                              /* the two-step kernel's per-population base pointers,   */
                              /* even where 32-bit byte offsets fit (they run by       */
                              /* themselves for buffers over 4 GiB); default 0         */
#define PLB_TUNE_LITERAL_FUSE 8 /* literal mode: 1 (default) runs a collision and the */
                                /* push that follows it as one kernel (same values in */
                                /* n and ntmp, one pass fewer over n), and computes   */
                                /* each step's moments and resets in the next step's */
                                /* kernel (same rho and u), inside plb_step and for   */
                                /* the staged calls alike; 0 runs one kernel per      */
                                /* reference call                                      */
#define PLB_TUNE_TWO_STEP 9     /* fused pull step: 1 advances two steps per launch   */
                                /* (one read and one write of the populations per two */
                                /* steps, the same values bit for bit) wherever a     */
                                /* kernel exists: D3Q19 and D3Q27; x a multiple of 32; */
                                /* each population's array below 4 GiB; 0 one step per */
                                /* launch; -1 (default) automatic: on for             */
                                /* 4- and 8-byte storage on lattices of at least 2^18 */
                                /* cells (2^19 for D3Q27).  The staged calls pair     */
                                /* whole steps (plb_stream records one step and the   */
                                /* next plb_stream launches both; any call that reads */
                                /* or changes the state runs a recorded step first)   */
#define PLB_TUNE_TWO_ZCHUNK 10  /* planes per z-chunk of a two-step launch (0: chosen */
                                /* from the grid and the compute-unit count)          */
#define PLB_TUNE_TWO_TILE_X 11  /* two-step tile width: 64 or 32 cells (one          */
                                /* workgroup per compute unit either way, 32-cell     */
                                /* tiles taller); 0 automatic: 32 for 8-byte storage, */
                                /* else 64                                            */
int plb_set_tuning(plb_state* s, int key, int value);
/* The streaming scheme the next fused steps will use: PLB_SCHEME_PUSH,
 * PLB_SCHEME_PULL, PLB_SCHEME_PULL_TWO (pull, two steps per launch, for
 * pairs of steps) or PLB_SCHEME_AA. */
int plb_step_scheme(const plb_state* s, int* scheme);

/* ---- batched domains (no reference counterpart: one domain per PLBPD,
 * PLBPD.hpp:14-17) --------------------------------------------------------- */
typedef struct {
    int32_t x, y, z;
    double tau;
    double force[3];
} plb_domain_desc;

/* ndom independent domains sharing q, variant and boundary kinds, packed in
 * one pool of device memory.  bc applies to every domain (fused semantics).
 * Each domain has fewer than 2^31 cells (else PLB_ERR_INVALID_ARGUMENT,
 * "domain too large for a batch"). */
int plb_batch_create(int q, int variant, int ndom, const plb_domain_desc* doms,
                     const plb_boundary* bc, int device, plb_batch** out);
int plb_batch_destroy(plb_batch* b);
int plb_batch_initialize(plb_batch* b, double rho0, const double u0[3], const plb_init_spec* init);
/* Tile shape: cells per tile (multiple of 64 in [64, 4096]) and the
 * persistent grid size in blocks (0 = one block per tile). */
int plb_batch_set_tiling(plb_batch* b, int tile_cells, int grid_blocks);
int plb_batch_get_tiling(const plb_batch* b, int* tile_cells, int* grid_blocks);
/* Streaming scheme of the batched step: PLB_SCHEME_AUTO / PLB_SCHEME_PULL
 * (default: gather, whole-line stores, moving walls by write-time terms) or
 * PLB_SCHEME_PUSH. */
int plb_batch_set_scheme(plb_batch* b, int scheme);
/* Tile order of a launch: 0 (default) = workgroup i takes tile i (persistent
 * blocks stride by the grid), one window of memory for the chip; 1 = each XCD
 * takes a contiguous run of tiles. */
int plb_batch_set_mapping(plb_batch* b, int xcd_runs);
int plb_batch_get_mapping(const plb_batch* b, int* xcd_runs);
/* Times each candidate tiling and mapping over `steps` steps and keeps the
 * fastest.  results: up to max_results rows of {tile_cells, grid_blocks,
 * mapping, ms_per_step}; *nresults = the rows written.  The candidates run
 * real steps on the batch: every candidate (60: 6 tile sizes x 5 grids x 2
 * mappings) advances all domains by 2 + (steps rounded up to even) steps, so
 * the call leaves the simulation that many steps times 60 further on (an
 * even count); initialise the batch again to start from its initial state. */
int plb_batch_autotune(plb_batch* b, int steps, double* results, int max_results, int* nresults);
/* nsteps of all domains: one launch per step (graph-captured when
 * steps_per_graph > 0). */
int plb_batch_step(plb_batch* b, int nsteps, int steps_per_graph);
/* Baseline: the same steps with one launch per domain per step. */
int plb_batch_step_per_domain(plb_batch* b, int nsteps);
/* rho (x*y*z) and u (x*y*z*3) of domain d, compact AoS, double. */
int plb_batch_download_moments(plb_batch* b, int d, double* rho, double* u);
int plb_batch_cells(const plb_batch* b, uint64_t* total_cells, uint64_t* fluid_cells);
int plb_batch_synchronize(plb_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* PLBPD_HIP_H_ */
