// oracle_les.cpp -- les_chan / les_chan_th (for/les.f90) on the CPU.
// TEST INFRASTRUCTURE ONLY (see oracle.cpp).
#include "oracle_internal.h"

namespace orc {

int Solver::les_chan(double) { return DGPU_EUNSUP; }
int Solver::les_chan_th(double) { return DGPU_EUNSUP; }

}  // namespace orc
