// Drop-in Streaming (interface of the reference's Streaming.hpp:10-34):
// "push" streams and swaps n / ntmp on the device; "pull" and any other
// string throw std::logic_error as in the reference (Streaming.cpp:18-24,
// 96-102).  With the fused device mode, "pull" runs the fused pull step (the
// same numbers as "push").
#ifndef PLBPD_AMD_HOST_STREAMING_HPP_
#define PLBPD_AMD_HOST_STREAMING_HPP_

#include <string>
#include <vector>

#include "LBModel.hpp"
#include "SimData.hpp"

class Streaming final {
 public:
    explicit Streaming(const LBModel* lbmodel);
    Streaming(Streaming&) = delete;
    Streaming& operator=(Streaming&) = delete;
    ~Streaming();
    void operator()(SimData& simdata, std::string stream_type, bool reference = false) const;
    float get_total_time() const;
};

#endif  // PLBPD_AMD_HOST_STREAMING_HPP_
