// Row pull step (fused_row.hpp) for the D2Q9 lattice, every storage variant.
#include "fused_row.hpp"

namespace plb {
PLB_ROW_INST_Q(9)
}  // namespace plb
