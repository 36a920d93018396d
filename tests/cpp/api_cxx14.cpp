// Compile-time check of the drop-in interface: every public member that the
// reference's headers declare exists in plbpd_amd/host with the same type,
// under C++14 (the reference's standard, src/CMakeLists.txt:6).  Built with
// -DPLBPD_REAL=float it also checks the reference's float fields.  Nothing is
// run; the program only has to compile.
#include <array>
#include <cstdint>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "Boundary.hpp"
#include "CalcMoments.hpp"
#include "Collision.hpp"
#include "CollisionSRT.hpp"
#include "Domain.hpp"
#include "EqlbDist.hpp"
#include "Field.hpp"
#include "Flow.hpp"
#include "LBDynamics.hpp"
#include "LBModel.hpp"
#include "PLBPD.hpp"
#include "SimData.hpp"
#include "Streaming.hpp"

#define SAME_TYPE(member, ...) static_assert(std::is_same<decltype(&member), __VA_ARGS__>::value, #member)

using Dims = std::tuple<size_t, size_t, size_t>;
using F3 = std::array<float, 3>;

// LBModel.hpp:21-31
static_assert(std::is_constructible<LBModel, std::string>::value, "LBModel(std::string)");
static_assert(!std::is_convertible<std::string, LBModel>::value, "LBModel(std::string) is explicit");
static_assert(!std::is_copy_constructible<LBModel>::value, "LBModel is not copyable");
static_assert(std::has_virtual_destructor<LBModel>::value, "LBModel has a virtual destructor");
SAME_TYPE(LBModel::get_model_name, std::string (LBModel::*)() const);
SAME_TYPE(LBModel::get_num_directions, uint32_t (LBModel::*)() const);
SAME_TYPE(LBModel::get_speed_of_sound_squared, float (LBModel::*)() const);
SAME_TYPE(LBModel::get_directional_velocities,
          const std::vector<std::array<int32_t, 3> >& (LBModel::*)() const);
SAME_TYPE(LBModel::get_directional_weights, const std::vector<float>& (LBModel::*)() const);
SAME_TYPE(LBModel::get_reverse, const std::vector<uint32_t>& (LBModel::*)() const);

// Domain.hpp:20-36
static_assert(std::is_constructible<Domain, const LBModel*, size_t, size_t, size_t, float, float, F3, F3>::value,
              "Domain(lbmodel, x, y, z, viscosity, density, velocity, force)");
static_assert(!std::is_copy_constructible<Domain>::value, "Domain is not copyable");
SAME_TYPE(Domain::get_dimensions, Dims (Domain::*)() const);
SAME_TYPE(Domain::get_fluid_viscosity, float (Domain::*)() const);
SAME_TYPE(Domain::get_fluid_density, float (Domain::*)() const);
SAME_TYPE(Domain::get_solid_density, float (Domain::*)() const);
SAME_TYPE(Domain::get_init_flow_velocity, const F3& (Domain::*)() const);
SAME_TYPE(Domain::get_external_force, const F3& (Domain::*)() const);
SAME_TYPE(Domain::initialize, void (Domain::*)(SimData&) const);

// Boundary.hpp:14-17, 52-75
static_assert(std::is_same<BoundaryType, std::unordered_map<std::string, std::string> >::value, "BoundaryType");
static_assert(std::is_same<BoundaryVelocity, std::unordered_map<std::string, F3> >::value, "BoundaryVelocity");
static_assert(std::is_constructible<Boundary, const LBModel*, const Domain*, const float, BoundaryType,
                                    BoundaryVelocity>::value,
              "Boundary(lbmodel, domain, solid_density, type, velocity)");
static_assert(!std::is_copy_constructible<Boundary>::value, "Boundary is not copyable");
SAME_TYPE(Boundary::reset, void (Boundary::*)(SimData&) const);
SAME_TYPE(Boundary::get_boundary_type, const BoundaryType (Boundary::*)() const);
SAME_TYPE(Boundary::get_boundary_velocity, const BoundaryVelocity (Boundary::*)() const);
SAME_TYPE(Boundary::reset_velocity, void (Boundary::*)(SimData&) const);
SAME_TYPE(Boundary::reset_density, void (Boundary::*)(SimData&) const);
SAME_TYPE(Boundary::apply_periodicity, void (Boundary::*)(SimData&) const);
SAME_TYPE(Boundary::apply_noslip, void (Boundary::*)(SimData&) const);
SAME_TYPE(Boundary::get_time_noslip, float (Boundary::*)() const);
SAME_TYPE(Boundary::get_time_periodicity, float (Boundary::*)() const);
SAME_TYPE(Boundary::get_time_reset, float (Boundary::*)() const);
SAME_TYPE(Boundary::get_time_velocity_reset, float (Boundary::*)() const);
SAME_TYPE(Boundary::get_time_density_reset, float (Boundary::*)() const);
SAME_TYPE(Boundary::get_total_time, float (Boundary::*)() const);

// Streaming.hpp:25-33
static_assert(std::is_constructible<Streaming, const LBModel*>::value, "Streaming(lbmodel)");
static_assert(!std::is_convertible<const LBModel*, Streaming>::value, "Streaming(lbmodel) is explicit");
SAME_TYPE(Streaming::operator(), void (Streaming::*)(SimData&, std::string, bool) const);
SAME_TYPE(Streaming::get_total_time, float (Streaming::*)() const);

// Collision.hpp:15-20, CollisionSRT.hpp:38-43
static_assert(std::is_abstract<Collision>::value, "Collision is abstract");
static_assert(std::has_virtual_destructor<Collision>::value, "Collision has a virtual destructor");
static_assert(std::is_base_of<Collision, CollisionSRT>::value, "CollisionSRT derives from Collision");
SAME_TYPE(Collision::get_total_time, float (Collision::*)() const);
static_assert(std::is_constructible<CollisionSRT, const LBModel*, const Domain*>::value,
              "CollisionSRT(lbmodel, domain)");
SAME_TYPE(CollisionSRT::operator(), void (CollisionSRT::*)(SimData&, bool) const);

// CalcMoments.hpp:39-46
static_assert(std::is_constructible<CalcMoments, const LBModel*>::value, "CalcMoments(lbmodel)");
static_assert(!std::is_convertible<const LBModel*, CalcMoments>::value, "CalcMoments(lbmodel) is explicit");
SAME_TYPE(CalcMoments::get_total_time, float (CalcMoments::*)() const);
SAME_TYPE(CalcMoments::operator(), void (CalcMoments::*)(SimData&) const);

// LBDynamics.hpp
static_assert(std::is_constructible<LBDynamics, const LBModel*, const Domain*>::value,
              "LBDynamics(lbmodel, domain)");
SAME_TYPE(LBDynamics::collide, void (LBDynamics::*)(SimData&, bool) const);
SAME_TYPE(LBDynamics::stream, void (LBDynamics::*)(SimData&, std::string, bool) const);
SAME_TYPE(LBDynamics::get_time_collide, float (LBDynamics::*)() const);
SAME_TYPE(LBDynamics::get_time_stream, float (LBDynamics::*)() const);
SAME_TYPE(LBDynamics::get_total_time, float (LBDynamics::*)() const);

// PLBPD.hpp
static_assert(std::is_constructible<PLBPD, const LBModel*, const Domain*, const Boundary*>::value,
              "PLBPD(lbmodel, domain, boundary)");
SAME_TYPE(PLBPD::initialize, void (PLBPD::*)(SimData&) const);
SAME_TYPE(PLBPD::march_in_time, void (PLBPD::*)(size_t, SimData&) const);
SAME_TYPE(PLBPD::print_times, void (PLBPD::*)() const);

// Flow.hpp
static_assert(std::is_constructible<Flow, std::string>::value, "Flow(std::string)");
static_assert(!std::is_convertible<std::string, Flow>::value, "Flow(std::string) is explicit");
SAME_TYPE(Flow::get_domain, const Domain* (Flow::*)() const);
SAME_TYPE(Flow::get_boundary, const Boundary* (Flow::*)() const);
SAME_TYPE(Flow::get_lbmodel, const LBModel* (Flow::*)() const);
SAME_TYPE(Flow::get_num_timesteps, size_t (Flow::*)() const);

// SimData.hpp:9-31: the public fields and the constructor.
static_assert(std::is_constructible<SimData, Dims, size_t>::value, "SimData(dimensions, num_directions)");
static_assert(std::is_same<decltype(SimData::rho), Field::ScalarField<plbpd::real, 1>*>::value, "SimData::rho");
static_assert(std::is_same<decltype(SimData::u), Field::VectorField<plbpd::real, 1>*>::value, "SimData::u");
static_assert(std::is_same<decltype(SimData::n), Field::VectorField<plbpd::real, 1>*>::value, "SimData::n");
static_assert(std::is_same<decltype(SimData::ntmp), Field::VectorField<plbpd::real, 1>*>::value, "SimData::ntmp");
SAME_TYPE(SimData::write_state, void (SimData::*)(std::string));

// Field.hpp:19-149
using VF = Field::VectorField<plbpd::real, 1>;
using SF = Field::ScalarField<plbpd::real, 1>;
using R = plbpd::real;
static_assert(std::is_base_of<VF, SF>::value, "ScalarField derives from VectorField");
static_assert(std::is_constructible<VF, uint32_t, uint32_t, uint32_t, uint32_t, R>::value,
              "VectorField(z, y, x, v, value)");
static_assert(std::is_constructible<SF, uint32_t, uint32_t, uint32_t, R>::value, "ScalarField(z, y, x, value)");
SAME_TYPE(VF::get_dimensions, std::tuple<uint32_t, uint32_t, uint32_t> (VF::*)() const);
SAME_TYPE(VF::get_vector_length, uint32_t (VF::*)() const);
SAME_TYPE(VF::get_extents, const Field::FieldExtents& (VF::*)() const);
static_assert(std::is_same<decltype(std::declval<VF&>().sub2ind(0, 0, 0, 0)), uint32_t>::value, "VF::sub2ind");
static_assert(std::is_same<decltype(std::declval<VF&>().at(0, 0, 0, 0)), R&>::value, "VF::at");
static_assert(std::is_same<decltype(std::declval<const VF&>().at(0, 0, 0, 0)), const R&>::value, "VF::at const");
static_assert(std::is_same<decltype(std::declval<VF&>().get(0, 0, 0, 0)), R*>::value, "VF::get");
static_assert(std::is_same<decltype(std::declval<const VF&>().get(0, 0, 0, 0)), const R*>::value, "VF::get const");
static_assert(std::is_same<decltype(std::declval<SF&>().sub2ind(0, 0, 0)), uint32_t>::value, "SF::sub2ind");
static_assert(std::is_same<decltype(std::declval<SF&>().at(0, 0, 0)), R&>::value, "SF::at");
static_assert(std::is_same<decltype(std::declval<const SF&>().at(0, 0, 0)), const R&>::value, "SF::at const");
static_assert(std::is_same<decltype(std::declval<SF&>().get(0, 0, 0)), R*>::value, "SF::get");
static_assert(std::is_same<decltype(std::declval<const SF&>().get(0, 0, 0)), const R*>::value, "SF::get const");
static_assert(std::is_same<decltype(std::declval<VF&>().get_neighbor(0, 0, 0, std::array<int32_t, 3>{})),
                           std::tuple<uint32_t, uint32_t, uint32_t> >::value,
              "VF::get_neighbor");

// EqlbDist.hpp
static_assert(std::is_constructible<EqlbDist, const LBModel*>::value, "EqlbDist(lbmodel)");
static_assert(!std::is_convertible<const LBModel*, EqlbDist>::value, "EqlbDist(lbmodel) is explicit");
static_assert(std::is_same<decltype(std::declval<EqlbDist&>()(0.0f, std::declval<const F3&>(),
                                                              std::declval<std::vector<float>&>())),
                           void>::value,
              "EqlbDist::operator()(rho, u, n)");

int main() { return 0; }
