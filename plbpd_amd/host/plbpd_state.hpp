// State files: the datasets of the reference's SimData::write_state
// (SimData.cpp:83-160): /Density (Z+2, Y+2, X+2) and /FlowVelocity
// (Z+2, Y+2, X+2, 3) as float32, halo included; plus /ParticleDistribution
// (Z+2, Y+2, X+2, Q) as float64, which the reference leaves commented out;
// and, for an exact resume, /DensityF64, /FlowVelocityF64 and (when given)
// /ParticleDistributionTmp, the ntmp array that the staged step reads at
// periodic seams.  HDF5 when built with it (PLBPD_HAVE_HDF5), else a raw
// file: the 8-byte tag "PLBPDST2", five uint64 (X, Y, Z, Q, has_ntmp), then
// the arrays in the order above.
#ifndef PLBPD_AMD_HOST_PLBPD_STATE_HPP_
#define PLBPD_AMD_HOST_PLBPD_STATE_HPP_

#include <cstddef>
#include <string>

#include "plbpd_device.hpp"

namespace plbpd {

// The arrays are FP64 whatever the host mirror's type, so that a state file
// holds the device state exactly (FP64 states included) for a resume.
void write_state_file(const std::string& path, size_t X, size_t Y, size_t Z, size_t Q, const double* rho,
                      const double* u, const double* n, const double* ntmp);
// Returns true when the file held ntmp (then written to `ntmp` if non-null).
bool read_state_file(const std::string& path, size_t X, size_t Y, size_t Z, size_t Q, double* rho, double* u,
                     double* n, double* ntmp);
bool state_files_are_hdf5();

}  // namespace plbpd

#endif  // PLBPD_AMD_HOST_PLBPD_STATE_HPP_
