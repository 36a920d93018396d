// Drop-in LBModel (interface of the reference's LBModel.hpp:8-35): lattice
// name, Q, c_s^2 and the velocity / weight / reverse tables, read from the
// HIP library so host and device use one set of constants.  Unknown names
// throw std::logic_error (LBModel.cpp:17).
#ifndef PLBPD_AMD_HOST_LBMODEL_HPP_
#define PLBPD_AMD_HOST_LBMODEL_HPP_

#include <array>
#include <cstdint>
#include <string>
#include <vector>

class LBModel final {
 public:
    LBModel() = delete;
    explicit LBModel(std::string model_name);
    LBModel(LBModel&) = delete;
    LBModel& operator=(LBModel&) = delete;
    virtual ~LBModel();
    std::string get_model_name() const;
    uint32_t get_num_directions() const;
    float get_speed_of_sound_squared() const;
    const std::vector<std::array<int32_t, 3> >& get_directional_velocities() const;
    const std::vector<float>& get_directional_weights() const;
    const std::vector<uint32_t>& get_reverse() const;
    // The weights as the exact double values (the FP64 path uses these).
    const std::vector<double>& get_directional_weights_f64() const;

 private:
    std::string name_;
    uint32_t q_;
    float cs2_;
    std::vector<std::array<int32_t, 3> > c_;
    std::vector<float> w_;
    std::vector<double> w64_;
    std::vector<uint32_t> rev_;
};

#endif  // PLBPD_AMD_HOST_LBMODEL_HPP_
