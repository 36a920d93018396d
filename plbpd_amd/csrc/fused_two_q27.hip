// Two pull steps per pass (fused_two.hpp) for the D3Q27 lattice, every storage variant
// (8-byte storage in 32-cell tiles only: its ring leaves a 64-cell tile too few rows).
#include "fused_two.hpp"

namespace plb {
PLB_TWO_INST_Q(27)
}  // namespace plb
