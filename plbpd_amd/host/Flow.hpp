// Scenario presets (interface of the reference's Flow.hpp:11-31): Couette and
// Poiseuille channels in 2D and 3D with the reference's parameters
// (Flow.cpp:20-106).  A flow_type that is not a preset name is read as a
// scenario file (plbpd_scenario.hpp), the config file the reference stubs out
// (Domain.cpp:7-11).  Anything else throws std::invalid_argument
// (Flow.cpp:15).
#ifndef PLBPD_AMD_HOST_FLOW_HPP_
#define PLBPD_AMD_HOST_FLOW_HPP_

#include <memory>
#include <string>

#include "Boundary.hpp"
#include "Domain.hpp"
#include "LBModel.hpp"
#include "plbpd_scenario.hpp"

class Flow final {
 public:
    explicit Flow(std::string flow_type);
    Flow(Flow&) = delete;
    Flow& operator=(Flow&) = delete;
    ~Flow();
    const Domain* get_domain() const;
    const Boundary* get_boundary() const;
    const LBModel* get_lbmodel() const;
    size_t get_num_timesteps() const;
    // The scenario file this flow was read from, or nullptr for a preset.
    const plbpd::Scenario* get_scenario() const { return _scenario.get(); }

 private:
    void channel(const char* model, size_t x, size_t y, size_t z, float nu, std::array<float, 3> force,
                 bool periodic_y, std::array<float, 3> lid, uint32_t steps);
    void from_scenario(const plbpd::Scenario& s);
    std::unique_ptr<LBModel> _lbmodel;
    std::unique_ptr<Domain> _domain;
    std::unique_ptr<Boundary> _boundary;
    std::unique_ptr<plbpd::Scenario> _scenario;
    uint32_t _num_timesteps;
};

#endif  // PLBPD_AMD_HOST_FLOW_HPP_
