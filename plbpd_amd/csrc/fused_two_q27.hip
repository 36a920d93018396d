// Two pull steps per pass (fused_two.hpp) for the D3Q27 lattice, every storage variant
// (8-byte storage has no kernel: its ring leaves too few rows per tile).
#include "fused_two.hpp"

namespace plb {
PLB_TWO_INST_Q(27)
}  // namespace plb
