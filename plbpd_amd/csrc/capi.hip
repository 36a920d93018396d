// C ABI of the plbpd HIP library: state lifecycle, one entry point per
// reference operator call, fused stepping, HIP-graph replay, timers.
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "state.hpp"

namespace plb {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

std::vector<double> tg_tables(int X, int Y, int Z) {
    std::vector<double> t(2 * (size_t)(X + Y + Z));
    double* p = t.data();
    const int L[3] = {X, Y, Z};
    for (int a = 0; a < 3; ++a) {
        const double k = 2.0 * M_PI / (double)L[a];
        for (int i = 0; i < L[a]; ++i) p[i] = std::sin(k * ((double)i + 0.5));
        for (int i = 0; i < L[a]; ++i) p[L[a] + i] = std::cos(k * ((double)i + 0.5));
        p += 2 * L[a];
    }
    return t;
}

int timer_begin(plb_state* s, int stage, TimerPair* tp) {
    tp->stage = -1;
    if (!s->timers_on) return PLB_OK;
    auto take = [&](hipEvent_t* e) -> int {
        if (!s->free_events.empty()) {
            *e = s->free_events.back();
            s->free_events.pop_back();
            return PLB_OK;
        }
        PLB_HIP_TRY(hipEventCreate(e));
        return PLB_OK;
    };
    int rc = take(&tp->start);
    if (rc) return rc;
    rc = take(&tp->stop);
    if (rc) return rc;
    tp->stage = stage;
    PLB_HIP_TRY(hipEventRecord(tp->start, s->stream));
    return PLB_OK;
}

namespace {
int resolve_timers(plb_state* s, size_t keep) {
    size_t n = s->pending.size() > keep ? s->pending.size() - keep : 0;
    for (size_t i = 0; i < n; ++i) {
        TimerPair& tp = s->pending[i];
        PLB_HIP_TRY(hipEventSynchronize(tp.stop));
        float ms = 0.f;
        PLB_HIP_TRY(hipEventElapsedTime(&ms, tp.start, tp.stop));
        s->stage_ms[tp.stage] += ms;
        s->free_events.push_back(tp.start);
        s->free_events.push_back(tp.stop);
    }
    s->pending.erase(s->pending.begin(), s->pending.begin() + n);
    return PLB_OK;
}
}  // namespace

int timer_end(plb_state* s, TimerPair* tp) {
    if (tp->stage < 0) return PLB_OK;
    PLB_HIP_TRY(hipEventRecord(tp->stop, s->stream));
    s->pending.push_back(*tp);
    if (s->pending.size() > 1024) return resolve_timers(s, 256);
    return PLB_OK;
}

namespace {

// Entry into a state's call that leaves the deferred literal calls queued:
// plb_collide_srt and the resets may extend the queue, and plb_stream folds
// it into its push.  Every other entry point uses enter(), which runs it.
int enter_keep(plb_state* s) {
    if (!s) return fail(PLB_ERR_INVALID_ARGUMENT, "null state");
    PLB_HIP_TRY(hipSetDevice(s->device));
    return PLB_OK;
}

void drop_graph(plb_state* s) {
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    s->graph = nullptr;
    s->graph_steps = 0;
}

template <class Fn>
int timed(plb_state* s, int stage, Fn&& fn) {
    TimerPair tp;
    int rc = timer_begin(s, stage, &tp);
    if (rc) return rc;
    rc = fn();
    const int rc2 = timer_end(s, &tp);
    return rc ? rc : rc2;
}

// Literal calls deferred in s->lit_queue run, in order, before any other call
// observes or changes the state; plb_stream("push") instead folds them into
// its kernel.
int lit_flush(plb_state* s) {
    std::vector<LitOp> q;
    q.swap(s->lit_queue);
    for (const LitOp& op : q) {
        int rc;
        switch (op.kind) {
            case kLitMoments: rc = timed(s, PLB_STAGE_MOMENTS, [&] { return lit_moments(s); }); break;
            case kLitResetVelocity:
                rc = timed(s, PLB_STAGE_RESET_VELOCITY, [&] { return lit_reset_velocity(s, &op.bc); });
                break;
            case kLitResetDensity:
                rc = timed(s, PLB_STAGE_RESET_DENSITY, [&] { return lit_reset_density(s, &op.bc); });
                break;
            default:
                rc = timed(s, PLB_STAGE_COLLIDE, [&] { return lit_collide(s, op.tau, op.force); });
                break;
        }
        if (rc) return rc;
    }
    return PLB_OK;
}

bool lit_queue_is(const plb_state* s, std::initializer_list<LitOpKind> kinds) {
    if (s->lit_queue.size() != kinds.size()) return false;
    size_t i = 0;
    for (LitOpKind k : kinds)
        if (s->lit_queue[i++].kind != k) return false;
    return true;
}

LitOp lit_op(LitOpKind kind, const plb_boundary* bc = nullptr, double tau = 0.0, const double* force = nullptr) {
    LitOp op;
    std::memset(&op, 0, sizeof(op));
    op.kind = kind;
    if (bc) op.bc = *bc;
    op.tau = tau;
    if (force)
        for (int i = 0; i < 3; ++i) op.force[i] = force[i];
    return op;
}

int enter(plb_state* s) {
    int rc = enter_keep(s);
    if (rc) return rc;
    rc = lit_flush(s);
    if (rc) return rc;
    return fused_flush_deferred(s);
}

bool same_bc(const plb_boundary& a, const plb_boundary& b) {
    return std::memcmp(&a, &b, sizeof(plb_boundary)) == 0;
}

int ensure_fused_bc(plb_state* s, const plb_boundary* bc, std::vector<uint8_t>* mask) {
    if (s->bc_set && same_bc(s->bc, *bc)) return PLB_OK;
    // a deferred step runs with the boundary it was taken with
    const int rc = fused_flush_deferred(s);
    if (rc) return rc;
    drop_graph(s);
    return fused_build_flags(s, bc, mask && !mask->empty() ? mask->data() : nullptr);
}

}  // namespace

int fused_flush_deferred(plb_state* s) {
    if (!s->fused_deferred) return PLB_OK;
    s->fused_deferred = false;
    return timed(s, PLB_STAGE_FUSED, [&] { return fused_launch(s, true); });
}

}  // namespace plb

using namespace plb;

extern "C" {

const char* plb_last_error(void) { return g_error.c_str(); }

int plb_version(void) { return 100; }

int plb_device_info(int device, char* name, size_t name_len, int* num_cu, size_t* hbm_bytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PLB_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(PLB_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    hipDeviceProp_t p;
    PLB_HIP_TRY(hipGetDeviceProperties(&p, device));
    if (name && name_len) {
        std::strncpy(name, p.name[0] ? p.name : p.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (num_cu) *num_cu = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    return PLB_OK;
}

int plb_lattice_tables(int q, int32_t* c, double* w, uint32_t* rev) {
    auto fill = [&](auto lat) {
        using L = decltype(lat);
        for (int k = 0; k < L::q; ++k) {
            if (c) {
                c[3 * k] = L::cx(k);
                c[3 * k + 1] = L::cy(k);
                c[3 * k + 2] = L::cz(k);
            }
            if (w) w[k] = L::w(k);
            if (rev) rev[k] = (uint32_t)plb::rev<L::q>(k);
        }
        return PLB_OK;
    };
    switch (q) {
        case 9: return fill(Lat<9>());
        case 19: return fill(Lat<19>());
        case 27: return fill(Lat<27>());
        default: return fail(PLB_ERR_LOGIC, "Model with q=" + std::to_string(q) + " not yet implemented");
    }
}

int plb_state_create(const plb_state_desc* d, plb_state** out) {
    if (!d || !out) return fail(PLB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (d->q != 9 && d->q != 19 && d->q != 27)
        return fail(PLB_ERR_LOGIC, "Model with q=" + std::to_string(d->q) + " not yet implemented");
    if (d->x <= 0 || d->y <= 0 || d->z <= 0) return fail(PLB_ERR_INVALID_ARGUMENT, "argument is zero");
    if (d->mode != PLB_MODE_LITERAL && d->mode != PLB_MODE_FUSED && d->mode != PLB_MODE_FUSED_AA)
        return fail(PLB_ERR_INVALID_ARGUMENT, "unknown mode");
    // Cell counts that no device holds are refused before any size product
    // can wrap: the fused kernels number cells, rows and tiles in 32 bits, and
    // the literal kernels keep halo'd row lengths in int.
    {
        const uint64_t h = d->mode == PLB_MODE_LITERAL ? 2 : 0;
        const uint64_t ex = (uint64_t)d->x + h, ey = (uint64_t)d->y + h, ez = (uint64_t)d->z + h;
        const uint64_t lim = d->mode == PLB_MODE_LITERAL ? (1ull << 40) : (1ull << 32);
        // ex, ey < 2^31 keep ex * ey from wrapping; the division keeps the
        // product with ez from wrapping (ex * ey * ez <= lim - 1 exactly then)
        if (ex >= (1ull << 31) || ey >= (1ull << 31) || ex * ey > (lim - 1) / ez)
            return fail(PLB_ERR_INVALID_ARGUMENT, "domain too large");
    }
    if (storage_bytes(d->variant) == 0) return fail(PLB_ERR_INVALID_ARGUMENT, "unknown storage variant");
    if (d->mode == PLB_MODE_LITERAL && d->variant != kF64 && d->variant != kF32)
        return fail(PLB_ERR_INVALID_ARGUMENT, "literal mode supports PLB_F64 and PLB_F32 only");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PLB_ERR_NO_DEVICE, "no HIP device");
    if (d->device < 0 || d->device >= ndev) return fail(PLB_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    PLB_HIP_TRY(hipSetDevice(d->device));

    plb_state* s = new plb_state();
    s->q = d->q;
    s->X = d->x;
    s->Y = d->y;
    s->Z = d->z;
    s->variant = d->variant;
    // a one-array state has fused semantics; only its step differs
    s->mode = d->mode == PLB_MODE_FUSED_AA ? PLB_MODE_FUSED : d->mode;
    s->aa = d->mode == PLB_MODE_FUSED_AA;
    s->aa_odd = false;
    s->aa_collide = true;
    s->device = d->device;
    s->sbytes = storage_bytes(d->variant);
    s->cur = 0;
    s->buf[0] = s->buf[1] = nullptr;
    s->rho = s->u = nullptr;
    s->flags = nullptr;
    s->nbmask = nullptr;
    s->has_solids = false;
    s->bc_set = false;
    std::memset(&s->bc, 0, sizeof(s->bc));
    s->tau = 1.0;
    s->force[0] = s->force[1] = s->force[2] = 0.0;
    s->configured = false;
    s->collide_pending = false;
    s->fused_deferred = false;
    s->lit_ntmp_folded = false;
    s->moments_valid = false;
    s->scratch = nullptr;
    s->scratch_bytes = 0;
    s->graph = nullptr;
    s->graph_steps = 0;
    s->graph_parity = 0;
    s->timers_on = false;
    s->tune_tx = 0;
    // defaults measured fastest on MI355X for D3Q19 and D3Q27 at 256^3 and
    // 384^3 (FP64, FP32, FP16 storage, with and without interior solids):
    // non-temporal stores, linear row order, the row pull kernel for 4- and
    // 8-byte storage where the row length allows it (width by storage size,
    // lattice and solids) and the cell kernel for 2-byte storage; workgroup
    // i on tile i, except XCD y-stripes for the cell kernel around solids
    s->tune_nt_load = 0;
    s->tune_nt_store = 1;
    s->tune_scheme = PLB_SCHEME_AUTO;
    s->tune_order = 0;
    s->tune_vector = -1;
    s->tune_xcd = -1;
    s->tune_index64 = 0;
    s->tune_lit_fuse = 1;
    s->tune_two = -1;
    s->tune_two_zc = 0;
    s->tune_two_bx = 0;
    s->num_cu = 256;
    {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, d->device) == hipSuccess && cu > 0)
            s->num_cu = cu;
    }
    s->post_collision = false;
    for (double& v : s->stage_ms) v = 0.0;

    const uint64_t X = d->x, Y = d->y, Z = d->z;
    size_t mom_bytes;
    if (s->aa) s->tune_scheme = PLB_SCHEME_AA;
    if (d->mode == PLB_MODE_LITERAL) {
        s->ncell = (X + 2) * (Y + 2) * (Z + 2);
        s->pstride = s->ncell;
        mom_bytes = s->ncell * s->sbytes;
    } else {
        s->ncell = X * Y * Z;
        const size_t pad = d->pop_pad >= 0 ? (size_t)d->pop_pad : (size_t)(2048 + 64);
        s->pstride = ((s->ncell + 63) / 64) * 64 + pad;
        mom_bytes = s->ncell * sizeof(double);
    }
    auto cleanup = [&](int code, const std::string& msg) {
        plb_state_destroy(s);
        return fail(code, msg);
    };
    hipError_t e = hipStreamCreate(&s->stream);
    if (e != hipSuccess) {
        delete s;
        return fail(PLB_ERR_HIP, hipGetErrorString(e));
    }
    const size_t pop_bytes = s->pstride * s->q * s->sbytes;
    for (int b = 0; b < (s->aa ? 1 : 2); ++b) {
        e = hipMalloc(&s->buf[b], pop_bytes);
        if (e != hipSuccess) return cleanup(PLB_ERR_OOM, "population buffer: " + std::string(hipGetErrorString(e)));
        e = hipMemset(s->buf[b], 0, pop_bytes);
        if (e != hipSuccess) return cleanup(PLB_ERR_HIP, hipGetErrorString(e));
    }
    e = hipMalloc(&s->rho, mom_bytes);
    if (e == hipSuccess) e = hipMalloc(&s->u, 3 * mom_bytes);
    if (e == hipSuccess) e = hipMemset(s->rho, 0, mom_bytes);
    if (e == hipSuccess) e = hipMemset(s->u, 0, 3 * mom_bytes);
    if (e == hipSuccess && s->mode == PLB_MODE_FUSED) {
        e = hipMalloc(&s->flags, s->ncell);
        if (e == hipSuccess) e = hipMemset(s->flags, 0, s->ncell);
    }
    if (e != hipSuccess) return cleanup(PLB_ERR_OOM, hipGetErrorString(e));
    *out = s;
    return PLB_OK;
}

int plb_state_destroy(plb_state* s) {
    if (!s) return PLB_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    for (auto& tp : s->pending) {
        (void)hipEventDestroy(tp.start);
        (void)hipEventDestroy(tp.stop);
    }
    for (auto e : s->free_events) (void)hipEventDestroy(e);
    void* ptrs[] = {s->buf[0], s->buf[1], s->rho, s->u, s->flags, s->nbmask, s->scratch};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return PLB_OK;
}

int plb_state_bytes(const plb_state* s, size_t* bytes) {
    if (!s || !bytes) return fail(PLB_ERR_INVALID_ARGUMENT, "null argument");
    size_t b = (s->aa ? 1 : 2) * s->pstride * s->q * s->sbytes;
    if (s->mode == PLB_MODE_LITERAL)
        b += 4 * s->ncell * s->sbytes;
    else
        b += 4 * s->ncell * sizeof(double) + s->ncell + (s->nbmask ? 4 * s->ncell : 0);
    *bytes = b + s->scratch_bytes;
    return PLB_OK;
}

int plb_upload(plb_state* s, int field, int layout, int dtype, const void* host) {
    int rc = enter(s);
    if (rc) return rc;
    if (!host) return fail(PLB_ERR_INVALID_ARGUMENT, "null host pointer");
    if (dtype != PLB_DTYPE_F64 && dtype != PLB_DTYPE_F32) return fail(PLB_ERR_INVALID_ARGUMENT, "unknown dtype");
    if (layout != PLB_LAYOUT_REFERENCE && layout != PLB_LAYOUT_COMPACT)
        return fail(PLB_ERR_INVALID_ARGUMENT, "unknown layout");
    drop_graph(s);
    if (s->mode == PLB_MODE_LITERAL) return lit_transfer(s, field, layout, dtype, const_cast<void*>(host), true);
    return fused_transfer(s, field, layout, dtype, const_cast<void*>(host), true);
}

int plb_download(plb_state* s, int field, int layout, int dtype, void* host) {
    int rc = enter(s);
    if (rc) return rc;
    if (!host) return fail(PLB_ERR_INVALID_ARGUMENT, "null host pointer");
    if (dtype != PLB_DTYPE_F64 && dtype != PLB_DTYPE_F32) return fail(PLB_ERR_INVALID_ARGUMENT, "unknown dtype");
    if (layout != PLB_LAYOUT_REFERENCE && layout != PLB_LAYOUT_COMPACT)
        return fail(PLB_ERR_INVALID_ARGUMENT, "unknown layout");
    if (s->mode == PLB_MODE_LITERAL) return lit_transfer(s, field, layout, dtype, host, false);
    return fused_transfer(s, field, layout, dtype, host, false);
}

int plb_download_raw(plb_state* s, void* host, size_t bytes) {
    int rc = enter(s);
    if (rc) return rc;
    if (s->mode != PLB_MODE_FUSED) return fail(PLB_ERR_LOGIC, "raw download is for the fused layout");
    return fused_download_raw(s, host, bytes);
}

int plb_domain_initialize(plb_state* s, double rho0, const double u0[3], const plb_init_spec* init,
                          const plb_boundary* bc) {
    int rc = enter(s);
    if (rc) return rc;
    if (!u0) return fail(PLB_ERR_INVALID_ARGUMENT, "null u0");
    drop_graph(s);
    double* tg = nullptr;
    if (init && init->kind == PLB_INIT_TAYLOR_GREEN) {
        const std::vector<double> t = tg_tables(s->X, s->Y, s->Z);
        PLB_HIP_TRY(hipMalloc(&tg, t.size() * sizeof(double)));
        PLB_HIP_TRY(hipMemcpy(tg, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    } else if (init && init->kind != PLB_INIT_UNIFORM && init->kind != PLB_INIT_PERTURBED) {
        return fail(PLB_ERR_INVALID_ARGUMENT, "unknown initial-state kind");
    }
    rc = s->mode == PLB_MODE_LITERAL ? lit_initialize(s, rho0, u0, init, tg) : fused_initialize(s, rho0, u0, init, tg);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (tg) (void)hipFree(tg);
    if (rc) return rc;
    if (e != hipSuccess) return fail(PLB_ERR_HIP, hipGetErrorString(e));
    s->moments_valid = false;
    s->collide_pending = false;
    if (bc) return plb_reset(s, bc);
    return PLB_OK;
}

int plb_set_boundary(plb_state* s, const plb_boundary* bc, const uint8_t* mask) {
    int rc = enter(s);
    if (rc) return rc;
    if (!bc) return fail(PLB_ERR_INVALID_ARGUMENT, "null boundary");
    if (s->mode != PLB_MODE_FUSED) {
        if (mask) return fail(PLB_ERR_LOGIC, "interior solids need fused mode");
        s->bc = *bc;
        s->bc_set = true;
        return PLB_OK;
    }
    // a pull-scheme state finishes its pending streaming with the old flags
    rc = fused_to_pre(s);
    if (rc) return rc;
    std::vector<uint8_t>& m = s->mask_host;
    if (mask)
        m.assign(mask, mask + s->ncell);
    else
        m.clear();
    drop_graph(s);
    s->bc_set = false;
    return fused_build_flags(s, bc, mask);
}

int plb_collide_srt(plb_state* s, double tau, const double force[3]) {
    int rc = enter_keep(s);
    if (rc) return rc;
    if (s->mode == PLB_MODE_LITERAL) {
        // a collision waits for the push after nothing or after the moments
        // and resets of the step before
        if (!s->tune_lit_fuse ||
            !(s->lit_queue.empty() || lit_queue_is(s, {kLitMoments, kLitResetVelocity, kLitResetDensity})))
            if ((rc = lit_flush(s))) return rc;
        if (!force) return fail(PLB_ERR_INVALID_ARGUMENT, "null force");
        if (!s->tune_lit_fuse)
            return timed(s, PLB_STAGE_COLLIDE, [&] { return lit_collide(s, tau, force); });
        s->lit_queue.push_back(lit_op(kLitCollide, nullptr, tau, force));
        return PLB_OK;
    }
    if (!force) return fail(PLB_ERR_INVALID_ARGUMENT, "null force");
    if (s->tau != tau || std::memcmp(s->force, force, sizeof(s->force)) != 0) {
        // a deferred step runs with the relaxation time and force it was taken with
        if ((rc = fused_flush_deferred(s))) return rc;
        drop_graph(s);
    }
    s->tau = tau;
    for (int i = 0; i < 3; ++i) s->force[i] = force[i];
    s->collide_pending = true;
    return PLB_OK;
}

int plb_stream(plb_state* s, const char* stream_type) {
    int rc = enter_keep(s);
    if (rc) return rc;
    const std::string t = stream_type ? stream_type : "";
    // a queue left from before the fusion knob was turned off runs call by call
    const bool fold = s->tune_lit_fuse && t == "push";
    if (fold && lit_queue_is(s, {kLitMoments, kLitResetVelocity, kLitResetDensity, kLitCollide})) {
        const std::vector<LitOp> q = std::move(s->lit_queue);
        s->lit_queue.clear();
        return timed(s, PLB_STAGE_COLLIDE,
                     [&] {
                         return lit_moments_collide_push(s, q[3].tau, q[3].force, &q[1].bc, &q[2].bc);
                     });
    }
    if (fold && lit_queue_is(s, {kLitCollide})) {
        const LitOp c = s->lit_queue[0];
        s->lit_queue.clear();
        return timed(s, PLB_STAGE_COLLIDE, [&] { return lit_collide_push(s, c.tau, c.force); });
    }
    // every other request finds the deferred calls done, as the reference's does
    rc = lit_flush(s);
    if (rc) return rc;
    if (t != "push" && t != "pull") return fail(PLB_ERR_LOGIC, "Invalid stream type: " + t);
    if (s->mode == PLB_MODE_LITERAL) {
        if (t == "pull") return fail(PLB_ERR_LOGIC, "Streaming::_pull_tbb has not yet been implemented");
        return timed(s, PLB_STAGE_STREAM, [&] { return lit_push(s); });
    }
    // Fused mode: both schemes give the same numbers; "pull" asks for the
    // pull kernels, which the push scheme knob would contradict.
    if (t == "pull" && !s->aa && !fused_uses_pull(s))
        return fail(PLB_ERR_LOGIC, "stream type pull with the push scheme selected (PLB_TUNE_SCHEME)");
    const bool collide = s->collide_pending;
    s->collide_pending = false;
    // Whole steps of the staged calls pair up into two-step launches: the
    // first is recorded, the second launches both.  Every call that observes
    // or changes the state runs a recorded step first (enter(), and
    // ensure_fused_bc and plb_collide_srt when they change a parameter).
    if (collide && s->post_collision && fused_two_eligible(s)) {
        s->moments_valid = false;
        if (!s->fused_deferred) {
            s->fused_deferred = true;
            return PLB_OK;
        }
        s->fused_deferred = false;
        return timed(s, PLB_STAGE_FUSED, [&] { return fused_launch_two(s); });
    }
    if ((rc = fused_flush_deferred(s))) return rc;
    return timed(s, PLB_STAGE_FUSED, [&] { return fused_launch(s, collide); });
}

int plb_apply_noslip(plb_state* s, const plb_boundary* bc) {
    // fused: a check of the folded boundary, which keeps a deferred step
    int rc = s && s->mode == PLB_MODE_LITERAL ? enter(s) : enter_keep(s);
    if (rc) return rc;
    if (!bc) return fail(PLB_ERR_INVALID_ARGUMENT, "null boundary");
    if (s->mode == PLB_MODE_LITERAL) return timed(s, PLB_STAGE_NOSLIP, [&] { return lit_noslip(s, bc); });
    // folded into the fused kernel; the boundary must be the one it uses
    for (int f = 0; f < 6; ++f)
        if (bc->type[f] == kNoslip && !bc->has_vel[f])
            return fail(PLB_ERR_OUT_OF_RANGE, "no-slip face without a prescribed velocity");
    return ensure_fused_bc(s, bc, &s->mask_host);
}

int plb_apply_noslip_noalloc(plb_state* s, const plb_boundary* bc) {
    int rc = enter_keep(s);
    if (rc) return rc;
    if (!bc) return fail(PLB_ERR_INVALID_ARGUMENT, "null boundary");
    if (s->mode != PLB_MODE_LITERAL)
        return fail(PLB_ERR_LOGIC, "plb_apply_noslip_noalloc is for literal states (the fused step folds no-slip in)");
    // The deferred calls and this pass run untimed: a timer may need a new
    // event, and creating one can allocate.
    const bool timers = s->timers_on;
    s->timers_on = false;
    rc = lit_flush(s);
    if (rc == PLB_OK) rc = lit_noslip_noalloc(s, bc);
    s->timers_on = timers;
    return rc;
}

int plb_apply_periodicity(plb_state* s, const plb_boundary* bc) {
    int rc = s && s->mode == PLB_MODE_LITERAL ? enter(s) : enter_keep(s);
    if (rc) return rc;
    if (!bc) return fail(PLB_ERR_INVALID_ARGUMENT, "null boundary");
    if (s->mode == PLB_MODE_LITERAL) return timed(s, PLB_STAGE_PERIODIC, [&] { return lit_periodic(s, bc); });
    return ensure_fused_bc(s, bc, &s->mask_host);
}

int plb_calc_moments(plb_state* s) {
    int rc = s && s->mode == PLB_MODE_LITERAL ? enter(s) : enter_keep(s);
    if (rc) return rc;
    if (s->mode == PLB_MODE_LITERAL) {
        if (!s->tune_lit_fuse) return timed(s, PLB_STAGE_MOMENTS, [&] { return lit_moments(s); });
        s->lit_queue.push_back(lit_op(kLitMoments));  // enter() left the queue empty
        return PLB_OK;
    }
    s->moments_valid = false;  // computed on demand from the populations
    return PLB_OK;
}

int plb_reset_velocity(plb_state* s, const plb_boundary* bc) {
    int rc = enter_keep(s);
    if (rc) return rc;
    if (s->tune_lit_fuse && bc && lit_queue_is(s, {kLitMoments})) {
        s->lit_queue.push_back(lit_op(kLitResetVelocity, bc));
        return PLB_OK;
    }
    if ((rc = lit_flush(s))) return rc;
    if (!bc) return fail(PLB_ERR_INVALID_ARGUMENT, "null boundary");
    if (s->mode == PLB_MODE_LITERAL)
        return timed(s, PLB_STAGE_RESET_VELOCITY, [&] { return lit_reset_velocity(s, bc); });
    return ensure_fused_bc(s, bc, &s->mask_host);
}

int plb_reset_density(plb_state* s, const plb_boundary* bc) {
    int rc = enter_keep(s);
    if (rc) return rc;
    if (s->tune_lit_fuse && bc && lit_queue_is(s, {kLitMoments, kLitResetVelocity})) {
        s->lit_queue.push_back(lit_op(kLitResetDensity, bc));
        return PLB_OK;
    }
    if ((rc = lit_flush(s))) return rc;
    if (!bc) return fail(PLB_ERR_INVALID_ARGUMENT, "null boundary");
    if (s->mode == PLB_MODE_LITERAL)
        return timed(s, PLB_STAGE_RESET_DENSITY, [&] { return lit_reset_density(s, bc); });
    return ensure_fused_bc(s, bc, &s->mask_host);
}

int plb_reset(plb_state* s, const plb_boundary* bc) {
    int rc = plb_reset_velocity(s, bc);
    if (rc) return rc;
    return plb_reset_density(s, bc);
}

int plb_configure(plb_state* s, double tau, const double force[3], const plb_boundary* bc) {
    int rc = enter(s);
    if (rc) return rc;
    if (!force || !bc) return fail(PLB_ERR_INVALID_ARGUMENT, "null argument");
    if (!(tau > 0.5)) return fail(PLB_ERR_INVALID_ARGUMENT, "tau must exceed 1/2");
    if (s->tau != tau || std::memcmp(s->force, force, sizeof(s->force)) != 0) drop_graph(s);
    s->tau = tau;
    for (int i = 0; i < 3; ++i) s->force[i] = force[i];
    if (s->mode == PLB_MODE_FUSED) {
        rc = ensure_fused_bc(s, bc, &s->mask_host);
        if (rc) return rc;
    } else {
        for (int f = 0; f < 6; ++f)
            if (bc->type[f] == kNoslip && !bc->has_vel[f])
                return fail(PLB_ERR_OUT_OF_RANGE, "no-slip face without a prescribed velocity");
        s->bc = *bc;
        s->bc_set = true;
    }
    s->configured = true;
    return PLB_OK;
}

int plb_step(plb_state* s, int nsteps) {
    int rc = enter(s);
    if (rc) return rc;
    if (!s->configured) return fail(PLB_ERR_LOGIC, "plb_step before plb_configure");
    if (nsteps < 0) return fail(PLB_ERR_INVALID_ARGUMENT, "negative step count");
    const double* F = s->force;
    for (int i = 0; i < nsteps; ++i) {
        if (s->mode == PLB_MODE_LITERAL) {
            // PLBPD::march_in_time loop body (PLBPD.cpp:24-29).  With the
            // kernels folded, the moments and resets of every step but the
            // last are computed by the next step's collision kernel, which
            // stores rho and u as those stages would before it collides.
            const plb_boundary* bc = &s->bc;
            const bool fold = s->tune_lit_fuse != 0;
            if (fold && i > 0) {
                if ((rc = timed(s, PLB_STAGE_COLLIDE,
                                [&] { return lit_moments_collide_push(s, s->tau, F, bc, bc); })))
                    return rc;
            } else if (fold) {
                if ((rc = timed(s, PLB_STAGE_COLLIDE, [&] { return lit_collide_push(s, s->tau, F); })))
                    return rc;
            } else {
                if ((rc = timed(s, PLB_STAGE_COLLIDE, [&] { return lit_collide(s, s->tau, F); }))) return rc;
                if ((rc = timed(s, PLB_STAGE_STREAM, [&] { return lit_push(s); }))) return rc;
            }
            if ((rc = timed(s, PLB_STAGE_NOSLIP, [&] { return lit_noslip(s, bc); }))) return rc;
            if ((rc = timed(s, PLB_STAGE_PERIODIC, [&] { return lit_periodic(s, bc); }))) return rc;
            if (fold && i + 1 < nsteps) continue;
            if ((rc = timed(s, PLB_STAGE_MOMENTS, [&] { return lit_moments(s); }))) return rc;
            if ((rc = timed(s, PLB_STAGE_RESET_VELOCITY, [&] { return lit_reset_velocity(s, bc); }))) return rc;
            if ((rc = timed(s, PLB_STAGE_RESET_DENSITY, [&] { return lit_reset_density(s, bc); }))) return rc;
        } else if (i + 1 < nsteps && s->post_collision && fused_two_eligible(s)) {
            // two steps in one pass over the populations
            if ((rc = timed(s, PLB_STAGE_FUSED, [&] { return fused_launch_two(s); }))) return rc;
            ++i;
        } else {
            if ((rc = timed(s, PLB_STAGE_FUSED, [&] { return fused_launch(s, true); }))) return rc;
        }
    }
    s->collide_pending = false;
    return PLB_OK;
}

int plb_step_graph(plb_state* s, int nsteps, int steps_per_graph) {
    int rc = enter(s);
    if (rc) return rc;
    if (!s->configured) return fail(PLB_ERR_LOGIC, "plb_step_graph before plb_configure");
    if (s->mode != PLB_MODE_FUSED) return plb_step(s, nsteps);
    if (steps_per_graph <= 0 || (steps_per_graph & 1))
        return fail(PLB_ERR_INVALID_ARGUMENT, "steps_per_graph must be positive and even");
    if (nsteps < 0) return fail(PLB_ERR_INVALID_ARGUMENT, "negative step count");
    // The captured steps keep one storage form: post-collision for the pull
    // step (from the pre-collision form its first step is the collision
    // alone, run here), pre-collision for the push step, and for a one-array
    // state the form after an even step count (an odd step is run here first
    // if needed; a replay of an even number of steps returns to the form).
    if (s->aa) {
        if (s->aa_odd && nsteps > 0) {
            rc = timed(s, PLB_STAGE_FUSED, [&] { return fused_launch(s, true); });
            if (rc) return rc;
            --nsteps;
        }
        rc = fused_to_pre(s);
        if (rc) return rc;
    } else if (fused_uses_pull(s)) {
        if (!s->post_collision && nsteps > 0) {
            rc = timed(s, PLB_STAGE_FUSED, [&] { return fused_launch(s, true); });
            if (rc) return rc;
            --nsteps;
        }
    } else {
        rc = fused_to_pre(s);
        if (rc) return rc;
    }
    const int reps = nsteps / steps_per_graph;
    if (reps > 0 && (!s->graph || s->graph_steps != steps_per_graph || s->graph_parity != s->cur)) {
        drop_graph(s);
        const bool timers = s->timers_on;
        s->timers_on = false;
        PLB_HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        // two-step launches where eligible, in a number that keeps the count
        // of buffer swaps even (the graph must leave buf[cur] where it began)
        int pairs = !s->aa && s->post_collision && fused_two_eligible(s) ? steps_per_graph / 2 : 0;
        int singles = steps_per_graph - 2 * pairs;
        if (((pairs + singles) & 1) && pairs > 0) {
            --pairs;
            singles += 2;
        }
        int lrc = PLB_OK;
        for (int i = 0; i < pairs && lrc == PLB_OK; ++i) lrc = fused_launch_two(s);
        for (int i = 0; i < singles && lrc == PLB_OK; ++i) lrc = fused_launch(s, true);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        s->timers_on = timers;
        if (lrc) {
            if (g) (void)hipGraphDestroy(g);
            return lrc;
        }
        if (e != hipSuccess) return fail(PLB_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            s->graph = nullptr;
            return fail(PLB_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        s->graph_steps = steps_per_graph;
        s->graph_parity = s->cur;  // an even number of swaps leaves the parity unchanged
    }
    for (int r = 0; r < reps; ++r) {
        rc = timed(s, PLB_STAGE_FUSED, [&]() -> int {
            PLB_HIP_TRY(hipGraphLaunch(s->graph, s->stream));
            return PLB_OK;
        });
        if (rc) return rc;
    }
    s->moments_valid = false;
    const int rem = nsteps - reps * steps_per_graph;
    return rem ? plb_step(s, rem) : PLB_OK;
}

int plb_global_moments(plb_state* s, double out[4]) {
    int rc = enter(s);
    if (rc) return rc;
    if (!out) return fail(PLB_ERR_INVALID_ARGUMENT, "null output");
    if (s->mode == PLB_MODE_LITERAL) return lit_global(s, out);
    return fused_global(s, out);
}

int plb_synchronize(plb_state* s) {
    int rc = enter(s);
    if (rc) return rc;
    PLB_HIP_TRY(hipStreamSynchronize(s->stream));
    return PLB_OK;
}

int plb_enable_timers(plb_state* s, int on) {
    int rc = enter(s);
    if (rc) return rc;
    s->timers_on = on != 0;
    return PLB_OK;
}

int plb_stage_times(plb_state* s, double* ms_out) {
    int rc = enter(s);
    if (rc) return rc;
    rc = resolve_timers(s, 0);
    if (rc) return rc;
    if (ms_out)
        for (int i = 0; i < PLB_NSTAGES; ++i) ms_out[i] = s->stage_ms[i];
    return PLB_OK;
}

}  // extern "C"
