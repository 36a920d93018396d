#include "SimData.hpp"

#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "plbpd_state.hpp"

namespace {
constexpr int kDtype = std::is_same<plbpd::real, double>::value ? PLB_DTYPE_F64 : PLB_DTYPE_F32;
}

SimData::SimData(const std::tuple<size_t, size_t, size_t> domain_dimensions, const size_t num_directions)
    : SimData(domain_dimensions, num_directions, plbpd::default_device_config()) {}

SimData::SimData(const std::tuple<size_t, size_t, size_t> domain_dimensions, const size_t num_directions,
                 const plbpd::DeviceConfig& cfg)
    : rho(nullptr), u(nullptr), n(nullptr), ntmp(nullptr), dims_(domain_dimensions),
      kdim_(num_directions), cfg_(cfg), state_(nullptr) {
    bootstrap();
}

SimData::SimData(SimData&& other) noexcept
    : rho(other.rho), u(other.u), n(other.n), ntmp(other.ntmp), dims_(other.dims_),
      kdim_(other.kdim_), cfg_(other.cfg_), state_(other.state_) {
    other.rho = nullptr;
    other.u = nullptr;
    other.n = nullptr;
    other.ntmp = nullptr;
    other.state_ = nullptr;
}

SimData::~SimData() {
    delete rho;
    delete u;
    delete n;
    delete ntmp;
    if (state_) {
        plbpd::retire_state(state_);
        plb_state_destroy(state_);
    }
}

void SimData::bootstrap() {
    size_t X, Y, Z;
    std::tie(X, Y, Z) = dims_;
    // The device state takes int32 edges; a wider one would be truncated.
    if (X > (size_t)INT32_MAX || Y > (size_t)INT32_MAX || Z > (size_t)INT32_MAX)
        throw std::invalid_argument("domain too large");
    // A constructor that throws runs no destructor: free what was built.
    try {
        rho = new Field::ScalarField<plbpd::real, 1>(Z, Y, X, 0);
        u = new Field::VectorField<plbpd::real, 1>(Z, Y, X, 3, 0);
        n = new Field::VectorField<plbpd::real, 1>(Z, Y, X, kdim_, 0);
        ntmp = new Field::VectorField<plbpd::real, 1>(Z, Y, X, kdim_, 0);
        plb_state_desc d;
        d.q = (int32_t)kdim_;
        d.x = (int32_t)X;
        d.y = (int32_t)Y;
        d.z = (int32_t)Z;
        d.variant = cfg_.variant;
        d.mode = cfg_.mode;
        d.device = cfg_.device;
        d.pop_pad = cfg_.pop_pad;
        plbpd::check(plb_state_create(&d, &state_));
        if (cfg_.device_timers) plbpd::check(plb_enable_timers(state_, 1));
        plbpd::register_state(state_);
    } catch (...) {
        delete rho;
        delete u;
        delete n;
        delete ntmp;
        rho = nullptr;
        u = nullptr;
        n = nullptr;
        ntmp = nullptr;
        if (state_) plb_state_destroy(state_);
        state_ = nullptr;
        throw;
    }
}

void SimData::sync_to_host() {
    plbpd::check(plb_download(state_, PLB_FIELD_RHO, PLB_LAYOUT_REFERENCE, kDtype, rho->data()));
    plbpd::check(plb_download(state_, PLB_FIELD_U, PLB_LAYOUT_REFERENCE, kDtype, u->data()));
    plbpd::check(plb_download(state_, PLB_FIELD_N, PLB_LAYOUT_REFERENCE, kDtype, n->data()));
    if (cfg_.mode == PLB_MODE_LITERAL)
        plbpd::check(plb_download(state_, PLB_FIELD_NTMP, PLB_LAYOUT_REFERENCE, kDtype, ntmp->data()));
}

void SimData::sync_to_device() {
    plbpd::check(plb_upload(state_, PLB_FIELD_N, PLB_LAYOUT_REFERENCE, kDtype, n->data()));
    if (cfg_.mode == PLB_MODE_LITERAL) {
        plbpd::check(plb_upload(state_, PLB_FIELD_RHO, PLB_LAYOUT_REFERENCE, kDtype, rho->data()));
        plbpd::check(plb_upload(state_, PLB_FIELD_U, PLB_LAYOUT_REFERENCE, kDtype, u->data()));
        plbpd::check(plb_upload(state_, PLB_FIELD_NTMP, PLB_LAYOUT_REFERENCE, kDtype, ntmp->data()));
    }
}

// The file's datasets come from FP64 downloads of the device state, so an
// FP64 state is written exactly whatever the host mirror's type; the mirror
// is refreshed as well (the reference writes the fields it holds).
void SimData::write_state(std::string dump_file) {
    sync_to_host();
    size_t X, Y, Z;
    std::tie(X, Y, Z) = dims_;
    const size_t cells = (X + 2) * (Y + 2) * (Z + 2);
    const bool literal = cfg_.mode == PLB_MODE_LITERAL;
    std::vector<double> r(cells), v(3 * cells), f(kdim_ * cells), t(literal ? kdim_ * cells : 0);
    plbpd::check(plb_download(state_, PLB_FIELD_RHO, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, r.data()));
    plbpd::check(plb_download(state_, PLB_FIELD_U, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, v.data()));
    plbpd::check(plb_download(state_, PLB_FIELD_N, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, f.data()));
    if (literal) plbpd::check(plb_download(state_, PLB_FIELD_NTMP, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, t.data()));
    plbpd::write_state_file(dump_file, X, Y, Z, kdim_, r.data(), v.data(), f.data(), literal ? t.data() : nullptr);
}

// Uploads the file's FP64 arrays to the device state (ntmp only when the file
// holds it), then refreshes the host mirror.
void SimData::read_state(std::string dump_file) {
    size_t X, Y, Z;
    std::tie(X, Y, Z) = dims_;
    const size_t cells = (X + 2) * (Y + 2) * (Z + 2);
    std::vector<double> r(cells), v(3 * cells), f(kdim_ * cells), t(kdim_ * cells);
    const bool has_ntmp = plbpd::read_state_file(dump_file, X, Y, Z, kdim_, r.data(), v.data(), f.data(), t.data());
    plbpd::check(plb_upload(state_, PLB_FIELD_N, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, f.data()));
    if (cfg_.mode == PLB_MODE_LITERAL) {
        plbpd::check(plb_upload(state_, PLB_FIELD_RHO, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, r.data()));
        plbpd::check(plb_upload(state_, PLB_FIELD_U, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, v.data()));
        if (has_ntmp) plbpd::check(plb_upload(state_, PLB_FIELD_NTMP, PLB_LAYOUT_REFERENCE, PLB_DTYPE_F64, t.data()));
    }
    sync_to_host();
}
