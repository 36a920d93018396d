// Drop-in Domain (interface of the reference's Domain.hpp:12-39): extents,
// fluid parameters, the t = 0 state.  initialize() fills the device state
// (Domain.cpp:46-64: rho0, u0 and feq(rho0, u0) on every cell, halo
// included).  Extension: a constructor taking the relaxation time and the
// parameters in double, for the FP64 contract (the reference derives
// tau = 3 nu + 1/2 in float, CollisionSRT.cpp:19).
#ifndef PLBPD_AMD_HOST_DOMAIN_HPP_
#define PLBPD_AMD_HOST_DOMAIN_HPP_

#include <array>
#include <string>
#include <tuple>

#include "LBModel.hpp"
#include "SimData.hpp"

class Domain final {
 public:
    struct Relaxation {
        double tau;
    };

    Domain() = delete;
    Domain(const LBModel* lbmodel, size_t xdim, size_t ydim, size_t zdim, float fluid_viscosity,
           float fluid_density, std::array<float, 3> init_flow_velocity, std::array<float, 3> external_force);
    Domain(const LBModel* lbmodel, size_t xdim, size_t ydim, size_t zdim, Relaxation tau, double fluid_density,
           std::array<double, 3> init_flow_velocity, std::array<double, 3> external_force);
    Domain(Domain&) = delete;
    Domain& operator=(Domain&) = delete;
    ~Domain();

    std::tuple<size_t, size_t, size_t> get_dimensions() const;
    float get_fluid_viscosity() const;
    float get_fluid_density() const;
    // Declared by the reference (Domain.hpp:36) and defined nowhere: a
    // Domain holds no solid density (Boundary does), so a call fails to link
    // there and here alike.
    float get_solid_density() const;
    const std::array<float, 3>& get_init_flow_velocity() const;
    const std::array<float, 3>& get_external_force() const;
    void initialize(SimData& simdata) const;

    // Exact parameters the device step uses.
    double get_relaxation_time() const { return tau_; }
    double get_fluid_density_f64() const { return rho64_; }
    const std::array<double, 3>& get_init_flow_velocity_f64() const { return u064_; }
    const std::array<double, 3>& get_external_force_f64() const { return F64_; }
    // Optional per-cell start (PLB_INIT_PERTURBED / PLB_INIT_TAYLOR_GREEN).
    void set_initial_state(const plb_init_spec& spec) { init_ = spec; }

 private:
    const LBModel* lbmodel_;
    size_t x_, y_, z_;
    float nu_, rho_;
    std::array<float, 3> u0_, F_;
    double tau_, rho64_;
    std::array<double, 3> u064_, F64_;
    plb_init_spec init_;
};

#endif  // PLBPD_AMD_HOST_DOMAIN_HPP_
