#include "Streaming.hpp"

Streaming::Streaming(const LBModel* lbmodel) { (void)lbmodel; }
Streaming::~Streaming() {}

// GPU seconds of the pushes run alone and of the fused steps (a fused state
// streams and collides in one kernel, launched by this call).
float Streaming::get_total_time() const {
    const auto t = plbpd::stage_seconds();
    return (float)(t[PLB_STAGE_STREAM] + t[PLB_STAGE_FUSED]);
}

void Streaming::operator()(SimData& simdata, std::string stream_type, bool reference) const {
    (void)reference;
    plbpd::check(plb_stream(simdata.device(), stream_type.c_str()));
}
