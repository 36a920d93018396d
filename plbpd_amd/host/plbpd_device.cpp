#include "plbpd_device.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace plbpd {
namespace {
std::mutex g_mu;
DeviceConfig g_cfg;

std::mutex g_timer_mu;
std::vector<plb_state*> g_live;
std::array<double, PLB_NSTAGES> g_retired{};

void add_times(plb_state* s, std::array<double, PLB_NSTAGES>& acc) {
    double ms[PLB_NSTAGES];
    check(plb_stage_times(s, ms));
    for (int i = 0; i < PLB_NSTAGES; ++i) acc[i] += ms[i] * 1e-3;
}
}  // namespace

void register_state(plb_state* s) {
    std::lock_guard<std::mutex> lk(g_timer_mu);
    g_live.push_back(s);
}

void retire_state(plb_state* s) {
    std::lock_guard<std::mutex> lk(g_timer_mu);
    const auto it = std::find(g_live.begin(), g_live.end(), s);
    if (it == g_live.end()) return;
    g_live.erase(it);
    try {
        add_times(s, g_retired);
    } catch (...) {  // a failing device keeps the times it had
    }
}

std::array<double, PLB_NSTAGES> stage_seconds() {
    std::lock_guard<std::mutex> lk(g_timer_mu);
    std::array<double, PLB_NSTAGES> t = g_retired;
    for (plb_state* s : g_live) add_times(s, t);
    return t;
}

void set_default_device_config(const DeviceConfig& cfg) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cfg = cfg;
}

DeviceConfig default_device_config() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_cfg;
}

void check(int rc) {
    if (rc == PLB_OK) return;
    const std::string msg = plb_last_error();
    switch (rc) {
        case PLB_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case PLB_ERR_LOGIC: throw std::logic_error(msg);
        case PLB_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
        default: throw std::runtime_error("plbpd device error: " + msg);
    }
}

}  // namespace plbpd
