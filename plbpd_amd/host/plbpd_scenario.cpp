#include "plbpd_scenario.hpp"

#include <fstream>
#include <set>
#include <sstream>
#include <stdexcept>

namespace plbpd {

namespace {

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

struct Line {
    const std::string& file;
    int number;
    [[noreturn]] void fail(const std::string& what) const {
        throw std::invalid_argument(file + ":" + std::to_string(number) + ": " + what);
    }
};

// Exactly n numbers of type V in `value`, nothing else.
template <class V, size_t N>
std::array<V, N> numbers(const std::string& value, const Line& ln) {
    std::istringstream in(value);
    std::array<V, N> out{};
    for (size_t i = 0; i < N; ++i)
        if (!(in >> out[i])) ln.fail("expected " + std::to_string(N) + " number(s): '" + value + "'");
    std::string rest;
    if (in >> rest) ln.fail("unexpected '" + rest + "'");
    return out;
}

const std::set<std::string> kFaces = {"east", "west", "north", "south", "up", "down"};

// splitmix64 finaliser (Steele, Lea and Flood, OOPSLA 2014).
uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

Scenario read_scenario(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument("cannot open scenario file " + path);
    Scenario s;
    bool have_dims = false, have_visc = false;
    std::string raw;
    for (int number = 1; std::getline(in, raw); ++number) {
        const Line ln{path, number};
        const std::string line = trim(raw.substr(0, raw.find('#')));
        if (line.empty()) continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) ln.fail("expected 'key = value'");
        const std::string key = trim(line.substr(0, eq)), value = trim(line.substr(eq + 1));
        if (key == "lattice") {
            s.lattice = value;
        } else if (key == "dimensions") {
            const auto d = numbers<long long, 3>(value, ln);
            if (d[0] <= 0 || d[1] <= 0 || d[2] <= 0) ln.fail("dimensions must be positive");
            s.x = (size_t)d[0];
            s.y = (size_t)d[1];
            s.z = (size_t)d[2];
            have_dims = true;
        } else if (key == "viscosity") {
            s.viscosity = numbers<float, 1>(value, ln)[0];
            have_visc = true;
        } else if (key == "tau") {
            s.tau = numbers<double, 1>(value, ln)[0];
            if (!(s.tau > 0.5)) ln.fail("tau must exceed 1/2");
            s.has_tau = true;
        } else if (key == "density") {
            s.density = numbers<double, 1>(value, ln)[0];
        } else if (key == "velocity") {
            s.velocity = numbers<double, 3>(value, ln);
        } else if (key == "force") {
            s.force = numbers<double, 3>(value, ln);
        } else if (key.rfind("boundary.", 0) == 0) {
            const std::string face = key.substr(9);
            if (!kFaces.count(face)) ln.fail("unknown face '" + face + "'");
            if (value != "periodic" && value != "noslip") ln.fail("boundary type must be periodic or noslip");
            s.boundary_type[face] = value;
        } else if (key.rfind("velocity.", 0) == 0) {
            const std::string face = key.substr(9);
            if (!kFaces.count(face)) ln.fail("unknown face '" + face + "'");
            s.wall_velocity[face] = numbers<double, 3>(value, ln);
        } else if (key == "solid_density") {
            s.solid_density = numbers<float, 1>(value, ln)[0];
        } else if (key == "timesteps") {
            const auto n = numbers<long long, 1>(value, ln)[0];
            if (n < 0) ln.fail("timesteps must not be negative");
            s.timesteps = (size_t)n;
        } else if (key == "variant") {
            const std::map<std::string, int> v = {{"f64", PLB_F64}, {"f32", PLB_F32}, {"f64a32", PLB_F64_A32},
                                                  {"f32s", PLB_F32S}, {"f16s", PLB_F16S}};
            const auto it = v.find(value);
            if (it == v.end()) ln.fail("unknown storage variant '" + value + "'");
            s.variant = it->second;
            s.has_variant = true;
        } else if (key == "mode") {
            if (value == "literal") s.mode = PLB_MODE_LITERAL;
            else if (value == "fused") s.mode = PLB_MODE_FUSED;
            else ln.fail("mode must be literal or fused");
            s.has_mode = true;
        } else if (key == "init") {
            std::istringstream iv(value);
            std::string kind;
            iv >> kind;
            if (kind == "uniform") {
                s.init = {PLB_INIT_UNIFORM, 0, 0.0};
            } else if (kind == "taylor-green") {
                s.init = {PLB_INIT_TAYLOR_GREEN, 0, 0.0};
            } else if (kind == "perturbed") {
                unsigned long long seed = 0;
                double amp = 0.0;
                if (!(iv >> seed >> amp)) ln.fail("init = perturbed <seed> <amplitude>");
                s.init = {PLB_INIT_PERTURBED, seed, amp};
            } else {
                ln.fail("unknown init '" + kind + "'");
            }
            std::string rest;
            if (iv >> rest) ln.fail("unexpected '" + rest + "'");
        } else if (key == "solids.file") {
            s.solids_file = value;
        } else if (key == "solids.fraction") {
            s.solid_fraction = numbers<double, 1>(value, ln)[0];
            if (s.solid_fraction < 0.0 || s.solid_fraction > 1.0) ln.fail("solids.fraction must be in [0, 1]");
        } else if (key == "solids.seed") {
            s.solid_seed = numbers<unsigned long long, 1>(value, ln)[0];
        } else {
            ln.fail("unknown key '" + key + "'");
        }
    }
    if (s.lattice.empty()) throw std::invalid_argument(path + ": missing 'lattice'");
    if (!have_dims) throw std::invalid_argument(path + ": missing 'dimensions'");
    if (!have_visc && !s.has_tau) throw std::invalid_argument(path + ": missing 'viscosity' or 'tau'");
    return s;
}

std::vector<uint8_t> scenario_solids(const Scenario& s) {
    const size_t n = s.x * s.y * s.z;
    std::vector<uint8_t> mask;
    if (!s.solids_file.empty()) {
        std::ifstream in(s.solids_file, std::ios::binary);
        if (!in) throw std::invalid_argument("cannot open solids file " + s.solids_file);
        mask.resize(n);
        in.read(reinterpret_cast<char*>(mask.data()), (std::streamsize)n);
        if ((size_t)in.gcount() != n)
            throw std::invalid_argument("solids file " + s.solids_file + " holds fewer than X*Y*Z bytes");
    } else if (s.solid_fraction > 0.0) {
        mask.resize(n);
        for (size_t c = 0; c < n; ++c) {
            const uint64_t v = mix64(s.solid_seed * 0x9e3779b97f4a7c15ull + c);
            mask[c] = (double)(v >> 11) * 0x1.0p-53 < s.solid_fraction ? 1 : 0;
        }
    }
    return mask;
}

}  // namespace plbpd
