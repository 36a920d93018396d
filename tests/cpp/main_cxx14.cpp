// The reference's main() call sequence (main.cpp:13-31), built as C++14 --
// the reference's language standard (src/CMakeLists.txt:6) -- against the
// drop-in headers.  It shows that a maintainer's unchanged main() compiles
// and runs with plbpd_amd/host ahead of src/ on the include path:
// copy-initialising SimData from a temporary (main.cpp:22-23) needs an
// accessible copy or move constructor before C++17.  The program writes
// init_state.h5 and final_state.h5 into the working directory.
#include <memory>

#include "Boundary.hpp"
#include "Domain.hpp"
#include "Flow.hpp"
#include "LBModel.hpp"
#include "PLBPD.hpp"
#include "SimData.hpp"

#if __cplusplus != 201402L
#error "this program checks the drop-in headers under C++14"
#endif

int main() {
    const auto flow = std::make_unique<Flow>("Couette3D");
    const auto lbmodel = flow->get_lbmodel();
    const auto domain = flow->get_domain();
    const auto boundary = flow->get_boundary();
    const auto problem = std::make_unique<PLBPD>(lbmodel, domain, boundary);
    auto simdata = SimData(domain->get_dimensions(), lbmodel->get_num_directions());
    problem->initialize(simdata);
    simdata.write_state("init_state.h5");
    problem->march_in_time(flow->get_num_timesteps(), simdata);
    simdata.write_state("final_state.h5");
    problem->print_times();
}
