// Device configuration and error mapping shared by the drop-in classes.
//
// The drop-in keeps the reference's class names and signatures; what the
// reference leaves implicit (where the arrays live, in which precision, and
// which step semantics) is chosen here.  SimData's two-argument constructor
// (SimData.hpp:22-24 of the reference) uses the process-wide default set by
// plbpd::set_default_device_config().
#ifndef PLBPD_AMD_HOST_PLBPD_DEVICE_HPP_
#define PLBPD_AMD_HOST_PLBPD_DEVICE_HPP_

#include <array>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "plbpd_hip.h"

// Element type of the host mirror fields.  float by default, as the
// reference's SimData holds them (SimData.hpp:18-21); build with
// -DPLBPD_REAL=double for the FP64 contract (FP64 device state read back
// without rounding).
#ifndef PLBPD_REAL
#define PLBPD_REAL float
#endif

namespace plbpd {

using real = PLBPD_REAL;

struct DeviceConfig {
    // PLB_F64, PLB_F32, PLB_F64_A32, PLB_F32S, PLB_F16S.  By default the
    // device computes in the host fields' type: FP32 with the reference's
    // own rounding for float fields, the FP64 contract for double fields.
    int variant = std::is_same<real, double>::value ? PLB_F64 : PLB_F32;
    int mode = PLB_MODE_LITERAL;    // PLB_MODE_LITERAL or PLB_MODE_FUSED
    int device = 0;
    int pop_pad = -1;
    // Time every stage on the GPU with HIP events (plb_enable_timers).  These
    // times back the reference's per-class timer getters (Collision,
    // Streaming, Boundary, CalcMoments), which PLBPD::print_times reports.
    bool device_timers = true;
};

void set_default_device_config(const DeviceConfig& cfg);
DeviceConfig default_device_config();

// GPU seconds per stage (PLB_STAGE_*) summed over every device state the
// process has created: the states alive now and those already destroyed,
// as the reference's static per-class timers add up over all instances
// (Collision.cpp:3, Streaming.cpp:12, Boundary.cpp:12-16, CalcMoments.cpp:5).
// Reading waits for the timed work to finish.
void register_state(plb_state* s);
void retire_state(plb_state* s);  // adds the state's times to the retired sums
std::array<double, PLB_NSTAGES> stage_seconds();

// Throws the exception type the reference throws for the status code.
void check(int rc);

}  // namespace plbpd

#endif  // PLBPD_AMD_HOST_PLBPD_DEVICE_HPP_
