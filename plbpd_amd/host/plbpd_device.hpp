// Device configuration and error mapping shared by the drop-in classes.
//
// The drop-in keeps the reference's class names and signatures; what the
// reference leaves implicit (where the arrays live, in which precision, and
// which step semantics) is chosen here.  SimData's two-argument constructor
// (SimData.hpp:22-24 of the reference) uses the process-wide default set by
// plbpd::set_default_device_config().
#ifndef PLBPD_AMD_HOST_PLBPD_DEVICE_HPP_
#define PLBPD_AMD_HOST_PLBPD_DEVICE_HPP_

#include <stdexcept>
#include <string>

#include "plbpd_hip.h"

#ifndef PLBPD_REAL
#define PLBPD_REAL double
#endif

namespace plbpd {

// Element type of the host mirror fields (the reference uses float; the
// FP64 contract uses double).
using real = PLBPD_REAL;

struct DeviceConfig {
    int variant = PLB_F64;          // PLB_F64, PLB_F32, PLB_F64_A32, PLB_F32S, PLB_F16S
    int mode = PLB_MODE_LITERAL;    // PLB_MODE_LITERAL or PLB_MODE_FUSED
    int device = 0;
    int pop_pad = -1;
    // Time every stage on the GPU with HIP events (plb_enable_timers): the
    // device-side counterpart of the reference's per-class timers, which
    // PLBPD::print_device_times reports.  Off by default (an event pair per
    // call).
    bool device_timers = false;
};

void set_default_device_config(const DeviceConfig& cfg);
DeviceConfig default_device_config();

// Throws the exception type the reference throws for the status code.
void check(int rc);

}  // namespace plbpd

#endif  // PLBPD_AMD_HOST_PLBPD_DEVICE_HPP_
