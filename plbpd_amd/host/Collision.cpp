#include "Collision.hpp"

float Collision::_time = 0.0f;
Collision::Collision() {}
Collision::~Collision() {}
// Host seconds that subclasses add to _time (as the reference's do,
// CollisionSRT.cpp:31-32) plus the GPU time of the collisions run on device
// states; a literal collision run in one kernel with its push counts here.
float Collision::get_total_time() const {
    return _time + (float)plbpd::stage_seconds()[PLB_STAGE_COLLIDE];
}
