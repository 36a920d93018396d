#include "plbpd_state.hpp"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#ifdef PLBPD_HAVE_HDF5
#include "hdf5.h"
#endif

namespace plbpd {
namespace {

template <class Out, class In>
std::vector<Out> convert(const In* p, size_t n) {
    std::vector<Out> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = (Out)p[i];
    return v;
}

#ifdef PLBPD_HAVE_HDF5
void h5_write(hid_t file, const char* name, hid_t type, int rank, const hsize_t* dims, const void* data) {
    const hid_t space = H5Screate_simple(rank, dims, nullptr);
    const hid_t set = H5Dcreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t st = set < 0 ? -1 : H5Dwrite(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (set >= 0) H5Dclose(set);
    H5Sclose(space);
    if (set < 0 || st < 0) throw std::runtime_error(std::string("HDF5 write failed: ") + name);
}

void h5_read(hid_t file, const char* name, hid_t type, size_t count, void* data) {
    const hid_t set = H5Dopen2(file, name, H5P_DEFAULT);
    if (set < 0) throw std::runtime_error(std::string("HDF5 dataset missing: ") + name);
    const hid_t space = H5Dget_space(set);
    const hssize_t n = H5Sget_simple_extent_npoints(space);
    herr_t st = -1;
    if (n == (hssize_t)count) st = H5Dread(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Sclose(space);
    H5Dclose(set);
    if (st < 0) throw std::runtime_error(std::string("HDF5 read failed or size mismatch: ") + name);
}
#endif

}  // namespace

bool state_files_are_hdf5() {
#ifdef PLBPD_HAVE_HDF5
    return true;
#else
    return false;
#endif
}

void write_state_file(const std::string& path, size_t X, size_t Y, size_t Z, size_t Q, const double* rho,
                      const double* u, const double* n, const double* ntmp) {
    const size_t cells = (X + 2) * (Y + 2) * (Z + 2);
    const std::vector<float> r32 = convert<float>(rho, cells);
    const std::vector<float> u32 = convert<float>(u, 3 * cells);
    const std::vector<double> r64(rho, rho + cells);
    const std::vector<double> u64(u, u + 3 * cells);
    const std::vector<double> n64(n, n + Q * cells);
    const std::vector<double> t64 = ntmp ? std::vector<double>(ntmp, ntmp + Q * cells) : std::vector<double>();
#ifdef PLBPD_HAVE_HDF5
    const hid_t file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) throw std::runtime_error("cannot create " + path);
    const hsize_t d3[3] = {Z + 2, Y + 2, X + 2};
    const hsize_t d4[4] = {Z + 2, Y + 2, X + 2, 3};
    const hsize_t dq[4] = {Z + 2, Y + 2, X + 2, Q};
    try {
        h5_write(file, "/Density", H5T_NATIVE_FLOAT, 3, d3, r32.data());
        h5_write(file, "/FlowVelocity", H5T_NATIVE_FLOAT, 4, d4, u32.data());
        h5_write(file, "/ParticleDistribution", H5T_NATIVE_DOUBLE, 4, dq, n64.data());
        h5_write(file, "/DensityF64", H5T_NATIVE_DOUBLE, 3, d3, r64.data());
        h5_write(file, "/FlowVelocityF64", H5T_NATIVE_DOUBLE, 4, d4, u64.data());
        if (ntmp) h5_write(file, "/ParticleDistributionTmp", H5T_NATIVE_DOUBLE, 4, dq, t64.data());
    } catch (...) {
        H5Fclose(file);
        throw;
    }
    H5Fclose(file);
#else
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot create " + path);
    const uint64_t hdr[5] = {X, Y, Z, Q, ntmp ? 1u : 0u};
    bool ok = std::fwrite("PLBPDST2", 1, 8, f) == 8 && std::fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
              std::fwrite(r32.data(), 4, cells, f) == cells && std::fwrite(u32.data(), 4, 3 * cells, f) == 3 * cells &&
              std::fwrite(n64.data(), 8, Q * cells, f) == Q * cells &&
              std::fwrite(r64.data(), 8, cells, f) == cells && std::fwrite(u64.data(), 8, 3 * cells, f) == 3 * cells &&
              (!ntmp || std::fwrite(t64.data(), 8, Q * cells, f) == Q * cells);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw std::runtime_error("write failed: " + path);
#endif
}

bool read_state_file(const std::string& path, size_t X, size_t Y, size_t Z, size_t Q, double* rho, double* u,
                     double* n, double* ntmp) {
    const size_t cells = (X + 2) * (Y + 2) * (Z + 2);
    std::vector<double> r64(cells), u64(3 * cells), n64(Q * cells), t64;
    bool has_tmp = false;
#ifdef PLBPD_HAVE_HDF5
    const hid_t file = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) throw std::runtime_error("cannot open " + path);
    try {
        h5_read(file, "/ParticleDistribution", H5T_NATIVE_DOUBLE, Q * cells, n64.data());
        if (H5Lexists(file, "/DensityF64", H5P_DEFAULT) > 0) {
            h5_read(file, "/DensityF64", H5T_NATIVE_DOUBLE, cells, r64.data());
            h5_read(file, "/FlowVelocityF64", H5T_NATIVE_DOUBLE, 3 * cells, u64.data());
        } else {  // a file with the reference's float32 datasets only
            h5_read(file, "/Density", H5T_NATIVE_DOUBLE, cells, r64.data());
            h5_read(file, "/FlowVelocity", H5T_NATIVE_DOUBLE, 3 * cells, u64.data());
        }
        if (H5Lexists(file, "/ParticleDistributionTmp", H5P_DEFAULT) > 0) {
            t64.resize(Q * cells);
            h5_read(file, "/ParticleDistributionTmp", H5T_NATIVE_DOUBLE, Q * cells, t64.data());
            has_tmp = true;
        }
    } catch (...) {
        H5Fclose(file);
        throw;
    }
    H5Fclose(file);
#else
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    char tag[8];
    uint64_t hdr[5];
    std::vector<float> r32(cells), u32(3 * cells);
    bool ok = std::fread(tag, 1, 8, f) == 8 && std::memcmp(tag, "PLBPDST2", 8) == 0 &&
              std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == X && hdr[1] == Y && hdr[2] == Z && hdr[3] == Q &&
              std::fread(r32.data(), 4, cells, f) == cells && std::fread(u32.data(), 4, 3 * cells, f) == 3 * cells &&
              std::fread(n64.data(), 8, Q * cells, f) == Q * cells &&
              std::fread(r64.data(), 8, cells, f) == cells && std::fread(u64.data(), 8, 3 * cells, f) == 3 * cells;
    if (ok && hdr[4]) {
        t64.resize(Q * cells);
        ok = std::fread(t64.data(), 8, Q * cells, f) == Q * cells;
        has_tmp = true;
    }
    std::fclose(f);
    if (!ok) throw std::runtime_error("not a state file of this shape: " + path);
#endif
    std::copy(r64.begin(), r64.end(), rho);
    std::copy(u64.begin(), u64.end(), u);
    std::copy(n64.begin(), n64.end(), n);
    if (has_tmp && ntmp) std::copy(t64.begin(), t64.end(), ntmp);
    return has_tmp;
}

}  // namespace plbpd
