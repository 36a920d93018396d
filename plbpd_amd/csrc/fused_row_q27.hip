// Row pull step (fused_row.hpp) for the D3Q27 lattice, every storage variant.
#include "fused_row.hpp"

namespace plb {
PLB_ROW_INST_Q(27)
}  // namespace plb
