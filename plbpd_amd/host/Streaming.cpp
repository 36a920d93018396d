#include "Streaming.hpp"

#include <chrono>

namespace chrono = std::chrono;

float Streaming::_time = 0.0f;

Streaming::Streaming(const LBModel* lbmodel) { (void)lbmodel; }
Streaming::~Streaming() {}
float Streaming::get_total_time() const { return _time; }

void Streaming::operator()(SimData& simdata, std::string stream_type, bool reference) const {
    (void)reference;
    const auto start = chrono::system_clock::now();
    plbpd::check(plb_stream(simdata.device(), stream_type.c_str()));
    const chrono::duration<float> elapsed = chrono::system_clock::now() - start;
    Streaming::_time += elapsed.count();
}
