#include "Domain.hpp"

Domain::Domain(const LBModel* lbmodel, size_t xdim, size_t ydim, size_t zdim, float fluid_viscosity,
               float fluid_density, std::array<float, 3> init_flow_velocity, std::array<float, 3> external_force)
    : lbmodel_(lbmodel), x_(xdim), y_(ydim), z_(zdim), nu_(fluid_viscosity), rho_(fluid_density),
      u0_(init_flow_velocity), F_(external_force) {
    tau_ = (double)(3.0f * fluid_viscosity + 0.5f);  // CollisionSRT.cpp:19, in float
    rho64_ = fluid_density;
    for (int i = 0; i < 3; ++i) {
        u064_[i] = init_flow_velocity[i];
        F64_[i] = external_force[i];
    }
    init_ = {PLB_INIT_UNIFORM, 0, 0.0};
}

Domain::Domain(const LBModel* lbmodel, size_t xdim, size_t ydim, size_t zdim, Relaxation tau, double fluid_density,
               std::array<double, 3> init_flow_velocity, std::array<double, 3> external_force)
    : lbmodel_(lbmodel), x_(xdim), y_(ydim), z_(zdim), tau_(tau.tau), rho64_(fluid_density),
      u064_(init_flow_velocity), F64_(external_force) {
    nu_ = (float)((tau.tau - 0.5) / 3.0);
    rho_ = (float)fluid_density;
    for (int i = 0; i < 3; ++i) {
        u0_[i] = (float)init_flow_velocity[i];
        F_[i] = (float)external_force[i];
    }
    init_ = {PLB_INIT_UNIFORM, 0, 0.0};
}

Domain::~Domain() {}

std::tuple<size_t, size_t, size_t> Domain::get_dimensions() const { return std::make_tuple(x_, y_, z_); }
float Domain::get_fluid_viscosity() const { return nu_; }
float Domain::get_fluid_density() const { return rho_; }
const std::array<float, 3>& Domain::get_init_flow_velocity() const { return u0_; }
const std::array<float, 3>& Domain::get_external_force() const { return F_; }

void Domain::initialize(SimData& simdata) const {
    plbpd::check(plb_domain_initialize(simdata.device(), rho64_, u064_.data(), &init_, nullptr));
}
