// Drop-in Collision base (interface of the reference's Collision.hpp:10-21).
// The static timer counts GPU time: get_total_time adds the HIP-event time of
// every collision on a device state to what subclasses add to _time.
#ifndef PLBPD_AMD_HOST_COLLISION_HPP_
#define PLBPD_AMD_HOST_COLLISION_HPP_

#include <chrono>
#include <string>

#include "LBModel.hpp"
#include "SimData.hpp"

class Collision {
 protected:
    static float _time;

 public:
    Collision();
    Collision(Collision&) = delete;
    Collision& operator=(Collision&) = delete;
    virtual ~Collision();
    virtual void operator()(SimData& simdata, bool reference) const = 0;
    float get_total_time() const;
};

#endif  // PLBPD_AMD_HOST_COLLISION_HPP_
