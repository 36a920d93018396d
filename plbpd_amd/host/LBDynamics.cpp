#include "LBDynamics.hpp"

LBDynamics::LBDynamics(const LBModel* lbmodel, const Domain* domain)
    : _collide(new CollisionSRT(lbmodel, domain)), _stream(new Streaming(lbmodel)) {}
LBDynamics::~LBDynamics() {}
void LBDynamics::collide(SimData& simdata, bool reference) const { (*_collide)(simdata, reference); }
void LBDynamics::stream(SimData& simdata, std::string stream_type, bool reference) const {
    (*_stream)(simdata, stream_type, reference);
}
float LBDynamics::get_time_collide() const { return _collide->get_total_time(); }
float LBDynamics::get_time_stream() const { return _stream->get_total_time(); }
float LBDynamics::get_total_time() const { return get_time_collide() + get_time_stream(); }
