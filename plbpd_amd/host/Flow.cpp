#include "Flow.hpp"

#include <fstream>
#include <stdexcept>

Flow::Flow(std::string flow_type) {
    const std::array<float, 3> zero = {0.0f, 0.0f, 0.0f};
    if (flow_type == "Couette2D")  // Flow.cpp:20-39
        channel("D2Q9", 100, 1, 25, 0.5f, zero, false, {0.5f, 0.0f, 0.0f}, 20000);
    else if (flow_type == "Couette3D")  // Flow.cpp:41-62
        channel("D3Q19", 200, 50, 50, 0.5f, zero, true, {0.5f, 0.0f, 0.0f}, 200);
    else if (flow_type == "Poiseuille2D")  // Flow.cpp:64-83
        channel("D2Q9", 256, 1, 48, 0.1f, {1.0e-5f, 0.0f, 0.0f}, false, zero, 5000);
    else if (flow_type == "Poiseuille3D")  // Flow.cpp:85-106
        channel("D3Q19", 256, 256, 48, 0.1f, {1.0e-5f, 0.0f, 0.0f}, true, zero, 10);
    else if (std::ifstream(flow_type).good())
        from_scenario(plbpd::read_scenario(flow_type));
    else
        throw std::invalid_argument("Flow type [" + flow_type + "] not recognized");
}

void Flow::from_scenario(const plbpd::Scenario& s) {
    _scenario.reset(new plbpd::Scenario(s));
    _lbmodel.reset(new LBModel(s.lattice));
    if (s.has_tau) {
        _domain.reset(new Domain(_lbmodel.get(), s.x, s.y, s.z, Domain::Relaxation{s.tau}, s.density, s.velocity,
                                 s.force));
    } else {
        const auto f3 = [](const std::array<double, 3>& v) {
            return std::array<float, 3>{(float)v[0], (float)v[1], (float)v[2]};
        };
        _domain.reset(new Domain(_lbmodel.get(), s.x, s.y, s.z, s.viscosity, (float)s.density, f3(s.velocity),
                                 f3(s.force)));
    }
    _domain->set_initial_state(s.init);
    BoundaryType type(s.boundary_type.begin(), s.boundary_type.end());
    BoundaryVelocity velocity;
    for (const auto& f : s.boundary_type)
        if (f.second == "noslip") velocity[f.first] = {0.0f, 0.0f, 0.0f};
    for (const auto& v : s.wall_velocity)
        velocity[v.first] = {(float)v.second[0], (float)v.second[1], (float)v.second[2]};
    _boundary.reset(new Boundary(_lbmodel.get(), _domain.get(), s.solid_density, type, velocity));
    _num_timesteps = (uint32_t)s.timesteps;
}

void Flow::channel(const char* model, size_t x, size_t y, size_t z, float nu, std::array<float, 3> force,
                   bool periodic_y, std::array<float, 3> lid, uint32_t steps) {
    _lbmodel.reset(new LBModel(model));
    _domain.reset(new Domain(_lbmodel.get(), x, y, z, nu, 1.0f, {0.0f, 0.0f, 0.0f}, force));
    BoundaryType type = {{"east", "periodic"}, {"west", "periodic"}, {"up", "noslip"}, {"down", "noslip"}};
    if (periodic_y) {
        type["north"] = "periodic";
        type["south"] = "periodic";
    }
    BoundaryVelocity velocity = {{"up", lid}, {"down", {0.0f, 0.0f, 0.0f}}};
    _boundary.reset(new Boundary(_lbmodel.get(), _domain.get(), 1.0f, type, velocity));
    _num_timesteps = steps;
}

Flow::~Flow() {}
const Domain* Flow::get_domain() const { return _domain.get(); }
const Boundary* Flow::get_boundary() const { return _boundary.get(); }
const LBModel* Flow::get_lbmodel() const { return _lbmodel.get(); }
size_t Flow::get_num_timesteps() const { return _num_timesteps; }
