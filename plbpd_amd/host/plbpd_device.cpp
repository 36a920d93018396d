#include "plbpd_device.hpp"

#include <mutex>

namespace plbpd {
namespace {
std::mutex g_mu;
DeviceConfig g_cfg;
}  // namespace

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
