// Drop-in SimData (interface of the reference's SimData.hpp:9-31).  The
// simulation state lives on the GPU (a plb_state); the public rho / u / n /
// ntmp fields are a host mirror in the reference's halo'd AoS layout, valid
// after sync_to_host() (write_state calls it).  Host edits reach the device
// through sync_to_device().
#ifndef PLBPD_AMD_HOST_SIMDATA_HPP_
#define PLBPD_AMD_HOST_SIMDATA_HPP_

#include <string>
#include <tuple>

#include "Field.hpp"
#include "plbpd_device.hpp"

class SimData final {
 public:
    Field::ScalarField<plbpd::real, 1>* rho;  // 1 -> number of buffer layers
    Field::VectorField<plbpd::real, 1>* u;
    Field::VectorField<plbpd::real, 1>* n;
    Field::VectorField<plbpd::real, 1>* ntmp;

    SimData() = delete;
    // dimensions are (x, y, z) as Domain::get_dimensions returns them
    SimData(const std::tuple<size_t, size_t, size_t> domain_dimensions, const size_t num_directions);
    SimData(const std::tuple<size_t, size_t, size_t> domain_dimensions, const size_t num_directions,
            const plbpd::DeviceConfig& cfg);
    SimData(const SimData&) = delete;
    // Takes over the device state and the host mirror; the source keeps
    // nothing and may only be destroyed.  The reference's main.cpp:22-23
    // writes `auto simdata = SimData(...)`, which C++14 (the reference's
    // standard, src/CMakeLists.txt:6) accepts only with a copy or move
    // constructor to elide.
    SimData(SimData&& other) noexcept;
    SimData& operator=(SimData&) = delete;
    ~SimData();

    // Dumps /Density (Z+2, Y+2, X+2) and /FlowVelocity (Z+2, Y+2, X+2, 3)
    // as float32, the datasets of the reference's write_state
    // (SimData.cpp:83-160), plus /ParticleDistribution (Z+2, Y+2, X+2, Q),
    // in HDF5 when the library is built with it, else in the raw format of
    // plbpd_state.hpp.
    void write_state(std::string dump_file);
    // Reads a file written by write_state back into the device state (the
    // reference's _read_state, SimData.cpp:162-164, is not implemented).
    void read_state(std::string dump_file);

    void sync_to_host();
    void sync_to_device();
    plb_state* device() const { return state_; }
    const plbpd::DeviceConfig& device_config() const { return cfg_; }
    size_t num_directions() const { return kdim_; }
    std::tuple<size_t, size_t, size_t> dimensions() const { return dims_; }

 private:
    void bootstrap();
    std::tuple<size_t, size_t, size_t> dims_;
    size_t kdim_;
    plbpd::DeviceConfig cfg_;
    plb_state* state_;
};

#endif  // PLBPD_AMD_HOST_SIMDATA_HPP_
