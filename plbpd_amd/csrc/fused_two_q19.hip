// Two pull steps per pass (fused_two.hpp) for the D3Q19 lattice, every storage variant.
#include "fused_two.hpp"

namespace plb {
PLB_TWO_INST_Q(19)
}  // namespace plb
