// Drop-in Boundary (interface of the reference's Boundary.hpp:14-79).  The
// face maps are turned into a plb_boundary once; every call runs on the
// device.  A no-slip face without a velocity entry throws std::out_of_range
// from apply_noslip, as the reference's _velocity.at does (Boundary.cpp:197).
// The timer getters report GPU time (HIP events on every device state).
#ifndef PLBPD_AMD_HOST_BOUNDARY_HPP_
#define PLBPD_AMD_HOST_BOUNDARY_HPP_

#include <array>
#include <memory>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "Domain.hpp"
#include "LBModel.hpp"
#include "SimData.hpp"

using fVectorField = Field::VectorField<plbpd::real, 1>;
using fScalarField = Field::ScalarField<plbpd::real, 1>;
using BoundaryType = std::unordered_map<std::string, std::string>;
using BoundaryVelocity = std::unordered_map<std::string, std::array<float, 3> >;

class Boundary final {
 public:
    Boundary() = delete;
    Boundary(const LBModel* lbmodel, const Domain* domain, const float solid_density, BoundaryType type,
             BoundaryVelocity velocity);
    Boundary(Boundary&) = delete;
    Boundary& operator=(Boundary&) = delete;
    ~Boundary();

    void reset(SimData& simdata) const;
    const BoundaryType get_boundary_type() const;
    const BoundaryVelocity get_boundary_velocity() const;
    void reset_velocity(SimData& simdata) const;
    void reset_density(SimData& simdata) const;
    void apply_periodicity(SimData& simdata) const;
    void apply_noslip(SimData& simdata) const;
    float get_time_noslip() const;
    float get_time_periodicity() const;
    float get_time_reset() const;
    float get_time_velocity_reset() const;
    float get_time_density_reset() const;
    float get_total_time() const;

    // The C-ABI description of this boundary; interior solids for the fused
    // step are registered with set_solid_mask (no reference counterpart).
    const plb_boundary& descriptor() const { return bc_; }
    void set_solid_mask(SimData& simdata, const uint8_t* mask) const;

 private:
    BoundaryType _type;
    BoundaryVelocity _velocity;
    plb_boundary bc_;
    static const std::vector<std::string> _directions;
};

#endif  // PLBPD_AMD_HOST_BOUNDARY_HPP_
