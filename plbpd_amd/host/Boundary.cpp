#include "Boundary.hpp"

#include <cstring>

// face order of the reference (Boundary.cpp:17-18) = PLB_EAST ... PLB_DOWN
const std::vector<std::string> Boundary::_directions = {"east", "west", "north", "south", "up", "down"};

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
    reset_velocity(simdata);
    reset_density(simdata);
}

void Boundary::reset_velocity(SimData& simdata) const {
    plbpd::check(plb_reset_velocity(simdata.device(), &bc_));
}

void Boundary::reset_density(SimData& simdata) const {
    plbpd::check(plb_reset_density(simdata.device(), &bc_));
}

void Boundary::apply_periodicity(SimData& simdata) const {
    plbpd::check(plb_apply_periodicity(simdata.device(), &bc_));
}

void Boundary::apply_noslip(SimData& simdata) const { plbpd::check(plb_apply_noslip(simdata.device(), &bc_)); }

void Boundary::set_solid_mask(SimData& simdata, const uint8_t* mask) const {
    plbpd::check(plb_set_boundary(simdata.device(), &bc_, mask));
}

// The timers report GPU seconds of the stages on every device state
// (resets folded into a literal collide + push kernel count under the
// collision).  reset covers both resets, as Boundary.cpp:58-66 times them.
float Boundary::get_time_noslip() const { return (float)plbpd::stage_seconds()[PLB_STAGE_NOSLIP]; }
float Boundary::get_time_periodicity() const { return (float)plbpd::stage_seconds()[PLB_STAGE_PERIODIC]; }
float Boundary::get_time_velocity_reset() const {
    return (float)plbpd::stage_seconds()[PLB_STAGE_RESET_VELOCITY];
}
float Boundary::get_time_density_reset() const { return (float)plbpd::stage_seconds()[PLB_STAGE_RESET_DENSITY]; }
float Boundary::get_time_reset() const {
    const auto t = plbpd::stage_seconds();
    return (float)(t[PLB_STAGE_RESET_VELOCITY] + t[PLB_STAGE_RESET_DENSITY]);
}
float Boundary::get_total_time() const {
    const auto t = plbpd::stage_seconds();
    return (float)(t[PLB_STAGE_NOSLIP] + t[PLB_STAGE_PERIODIC] + t[PLB_STAGE_RESET_VELOCITY] +
                   t[PLB_STAGE_RESET_DENSITY]);
}
