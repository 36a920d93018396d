// Host-side field of the drop-in: the interface of the reference's
// Field::VectorField / ScalarField (Field.hpp:13-149), used as the host mirror
// of the device arrays.  Layout and index are the reference's:
// (z, y, x, v) with `num_buffer_layers` halo cells on each side,
// index = z*(vlen*X*Y) + y*(vlen*X) + x*vlen + v (Field.hpp:91-93), extents
// [nbuf, len + nbuf) per axis (Field.hpp:49-56), a zero extent throws
// std::invalid_argument (Field.hpp:28-31).
#ifndef PLBPD_AMD_HOST_FIELD_HPP_
#define PLBPD_AMD_HOST_FIELD_HPP_

#include <array>
#include <cstdint>
#include <stdexcept>
#include <tuple>
#include <vector>

namespace Field {

struct FieldExtents {
    uint32_t zbegin, zend;
    uint32_t ybegin, yend;
    uint32_t xbegin, xend;
};

template <typename T, uint32_t num_buffer_layers>
class VectorField {
 public:
    VectorField(const uint32_t zlen, const uint32_t ylen, const uint32_t xlen, const uint32_t vlen,
                const T& value) {
        if (zlen == 0 || ylen == 0 || xlen == 0 || vlen == 0) throw std::invalid_argument("argument is zero");
        nz_ = zlen + 2 * num_buffer_layers;
        ny_ = ylen + 2 * num_buffer_layers;
        nx_ = xlen + 2 * num_buffer_layers;
        nv_ = vlen;
        data_.assign((size_t)nz_ * ny_ * nx_ * nv_, value);
        ext_ = {num_buffer_layers, nz_ - num_buffer_layers, num_buffer_layers,
                ny_ - num_buffer_layers, num_buffer_layers, nx_ - num_buffer_layers};
    }
    VectorField(VectorField&) = delete;
    VectorField& operator=(VectorField&) = delete;
    ~VectorField() {}

    std::tuple<uint32_t, uint32_t, uint32_t> get_dimensions() const { return std::make_tuple(nz_, ny_, nx_); }
    uint32_t get_vector_length() const { return nv_; }
    const FieldExtents& get_extents() const { return ext_; }

    uint32_t sub2ind(uint32_t z, uint32_t y, uint32_t x, uint32_t v) const {
        return ((z * ny_ + y) * nx_ + x) * nv_ + v;
    }
    T& at(uint32_t z, uint32_t y, uint32_t x, uint32_t v) { return data_[sub2ind(z, y, x, v)]; }
    const T& at(uint32_t z, uint32_t y, uint32_t x, uint32_t v) const { return data_[sub2ind(z, y, x, v)]; }
    T* get(uint32_t z, uint32_t y, uint32_t x, uint32_t v) { return &data_[sub2ind(z, y, x, v)]; }
    const T* get(uint32_t z, uint32_t y, uint32_t x, uint32_t v) const { return &data_[sub2ind(z, y, x, v)]; }

    // neighbour (z + c_z, y + c_y, x + c_x) of a cell (Field.hpp:111-115)
    std::tuple<uint32_t, uint32_t, uint32_t> get_neighbor(uint32_t z, uint32_t y, uint32_t x,
                                                          const std::array<int32_t, 3>& ck) const {
        return std::make_tuple(z + ck[2], y + ck[1], x + ck[0]);
    }

    // Contiguous storage, for transfers to and from the device.
    T* data() { return data_.data(); }
    const T* data() const { return data_.data(); }
    size_t size() const { return data_.size(); }

 private:
    uint32_t nz_, ny_, nx_, nv_;
    std::vector<T> data_;
    FieldExtents ext_;
};

template <typename T, uint32_t num_buffer_layers>
class ScalarField : public VectorField<T, num_buffer_layers> {
    using Base = VectorField<T, num_buffer_layers>;

 public:
    ScalarField(uint32_t zlen, uint32_t ylen, uint32_t xlen, const T& value) : Base(zlen, ylen, xlen, 1, value) {}
    ScalarField(ScalarField&) = delete;
    ScalarField& operator=(ScalarField&) = delete;
    ~ScalarField() {}

    uint32_t sub2ind(uint32_t z, uint32_t y, uint32_t x) const { return Base::sub2ind(z, y, x, 0); }
    T& at(uint32_t z, uint32_t y, uint32_t x) { return Base::at(z, y, x, 0); }
    const T& at(uint32_t z, uint32_t y, uint32_t x) const { return Base::at(z, y, x, 0); }
    T* get(uint32_t z, uint32_t y, uint32_t x) { return Base::get(z, y, x, 0); }
    const T* get(uint32_t z, uint32_t y, uint32_t x) const { return Base::get(z, y, x, 0); }
};

}  // namespace Field

#endif  // PLBPD_AMD_HOST_FIELD_HPP_
