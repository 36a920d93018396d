// Scenario files: a run described in text instead of code.  The reference
// hard-codes its runs (main.cpp:15 picks a Flow preset, Flow.cpp:20-106) and
// stubs out reading a domain from a config file (Domain.cpp:7-11); Flow reads
// these files (Flow::Flow with a path instead of a preset name).
//
// Format: one `key = value` per line, `#` starts a comment, keys in any order.
//   lattice        = D2Q9 | D3Q19 | D3Q27                       (required)
//   dimensions     = X Y Z                                       (required)
//   viscosity      = nu        (tau = 3 nu + 1/2 in float, as the reference)
//   tau            = tau       (given directly, in double; instead of viscosity)
//   density        = rho0                                        (default 1)
//   velocity       = ux uy uz  initial velocity                  (default 0)
//   force          = Fx Fy Fz  body force                        (default 0)
//   boundary.<face> = periodic | noslip   face: east west north south up down
//   velocity.<face> = ux uy uz  wall velocity of a no-slip face  (default 0)
//   solid_density  = rho_s                                       (default 1)
//   timesteps      = N                                           (default 0)
//   variant        = f64 | f32 | f64a32 | f32s | f16s            (device storage)
//   mode           = literal | fused                             (step semantics)
//   init           = uniform | perturbed <seed> <amp> | taylor-green
//   solids.file    = path      raw X*Y*Z bytes, x fastest, nonzero = solid
//   solids.fraction = p        Bernoulli(p) solid voxels ...
//   solids.seed    = s         ... from a counter-based generator (splitmix64)
// A malformed line, an unknown key or a missing required key throws
// std::invalid_argument naming the file and line.
#ifndef PLBPD_AMD_HOST_PLBPD_SCENARIO_HPP_
#define PLBPD_AMD_HOST_PLBPD_SCENARIO_HPP_

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "plbpd_device.hpp"

namespace plbpd {

struct Scenario {
    std::string lattice;
    size_t x = 0, y = 0, z = 0;
    bool has_tau = false;
    double tau = 0.0;
    float viscosity = 0.0f;
    double density = 1.0;
    std::array<double, 3> velocity{};
    std::array<double, 3> force{};
    std::map<std::string, std::string> boundary_type;             // face -> periodic / noslip
    std::map<std::string, std::array<double, 3>> wall_velocity;   // face -> u_b
    float solid_density = 1.0f;
    size_t timesteps = 0;
    bool has_variant = false, has_mode = false;
    int variant = PLB_F64;
    int mode = PLB_MODE_LITERAL;
    plb_init_spec init{PLB_INIT_UNIFORM, 0, 0.0};
    std::string solids_file;
    double solid_fraction = 0.0;
    uint64_t solid_seed = 0;
};

// Parses a scenario file (format above).
Scenario read_scenario(const std::string& path);

// The interior-solid mask the scenario asks for (X*Y*Z bytes, x fastest), or
// an empty vector.  Reading solids.file, or drawing solids.fraction with the
// counter-based generator, so the same file gives the same mask anywhere.
std::vector<uint8_t> scenario_solids(const Scenario& s);

}  // namespace plbpd

#endif  // PLBPD_AMD_HOST_PLBPD_SCENARIO_HPP_
