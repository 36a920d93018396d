#include "Boundary.hpp"

#include <chrono>
#include <cstring>

namespace chrono = std::chrono;

float Boundary::_time_noslip = 0.0f;
float Boundary::_time_periodicity = 0.0f;
float Boundary::_time_reset = 0.0f;
float Boundary::_time_velocity_reset = 0.0f;
float Boundary::_time_density_reset = 0.0f;
// face order of the reference (Boundary.cpp:17-18) = PLB_EAST ... PLB_DOWN
const std::vector<std::string> Boundary::_directions = {"east", "west", "north", "south", "up", "down"};

namespace {
template <class Fn>
float timed(Fn&& fn) {
    const auto start = chrono::system_clock::now();
    fn();
    const chrono::duration<float> elapsed = chrono::system_clock::now() - start;
    return elapsed.count();
}
}  // namespace

Boundary::Boundary(const LBModel* lbmodel, const Domain* domain, const float solid_density, BoundaryType type,
                   BoundaryVelocity velocity)
    : _type(type), _velocity(velocity) {
    (void)lbmodel;
    (void)domain;
    std::memset(&bc_, 0, sizeof(bc_));
    for (int f = 0; f < 6; ++f) {
        const auto t = _type.find(_directions[f]);
        if (t != _type.end()) {
            if (t->second == "periodic") bc_.type[f] = PLB_FACE_PERIODIC;
            if (t->second == "noslip") bc_.type[f] = PLB_FACE_NOSLIP;
        }
        const auto v = _velocity.find(_directions[f]);
        if (v != _velocity.end()) {
            bc_.has_vel[f] = 1;
            for (int i = 0; i < 3; ++i) bc_.vel[f][i] = v->second[i];
        }
    }
    bc_.rho_s = solid_density;
}

Boundary::~Boundary() {}

const BoundaryType Boundary::get_boundary_type() const { return _type; }
const BoundaryVelocity Boundary::get_boundary_velocity() const { return _velocity; }

void Boundary::reset(SimData& simdata) const {
    _time_reset += timed([&] {
        reset_velocity(simdata);
        reset_density(simdata);
    });
}

void Boundary::reset_velocity(SimData& simdata) const {
    _time_velocity_reset += timed([&] { plbpd::check(plb_reset_velocity(simdata.device(), &bc_)); });
}

void Boundary::reset_density(SimData& simdata) const {
    _time_density_reset += timed([&] { plbpd::check(plb_reset_density(simdata.device(), &bc_)); });
}

void Boundary::apply_periodicity(SimData& simdata) const {
    _time_periodicity += timed([&] { plbpd::check(plb_apply_periodicity(simdata.device(), &bc_)); });
}

void Boundary::apply_noslip(SimData& simdata) const {
    _time_noslip += timed([&] { plbpd::check(plb_apply_noslip(simdata.device(), &bc_)); });
}

void Boundary::set_solid_mask(SimData& simdata, const uint8_t* mask) const {
    plbpd::check(plb_set_boundary(simdata.device(), &bc_, mask));
}

float Boundary::get_time_noslip() const { return _time_noslip; }
float Boundary::get_time_periodicity() const { return _time_periodicity; }
float Boundary::get_time_reset() const { return _time_reset; }
float Boundary::get_time_velocity_reset() const { return _time_velocity_reset; }
float Boundary::get_time_density_reset() const { return _time_density_reset; }
float Boundary::get_total_time() const { return _time_noslip + _time_periodicity + _time_reset; }
