#include "LBModel.hpp"

#include <stdexcept>

#include "plbpd_device.hpp"

LBModel::LBModel(std::string model_name) : name_(model_name) {
    int q;
    if (name_ == "D2Q9")
        q = 9;
    else if (name_ == "D3Q19")
        q = 19;
    else if (name_ == "D3Q27")
        q = 27;
    else
        throw std::logic_error("Model " + name_ + " not yet implemented");
    q_ = (uint32_t)q;
    cs2_ = 1. / 3.;
    std::vector<int32_t> c(3 * q);
    w64_.assign(q, 0.0);
    rev_.assign(q, 0);
    plbpd::check(plb_lattice_tables(q, c.data(), w64_.data(), rev_.data()));
    c_.resize(q);
    w_.resize(q);
    for (int k = 0; k < q; ++k) {
        c_[k] = {c[3 * k], c[3 * k + 1], c[3 * k + 2]};
        w_[k] = (float)w64_[k];
    }
}

LBModel::~LBModel() {}
std::string LBModel::get_model_name() const { return name_; }
uint32_t LBModel::get_num_directions() const { return q_; }
float LBModel::get_speed_of_sound_squared() const { return cs2_; }
const std::vector<std::array<int32_t, 3> >& LBModel::get_directional_velocities() const { return c_; }
const std::vector<float>& LBModel::get_directional_weights() const { return w_; }
const std::vector<uint32_t>& LBModel::get_reverse() const { return rev_; }
const std::vector<double>& LBModel::get_directional_weights_f64() const { return w64_; }
