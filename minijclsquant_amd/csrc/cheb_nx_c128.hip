// complex128 neighbour-exchange step kernels (variants 48..51): one vector
// per lane, 4 or 8 lanes per row (groups of 4 and 8 vectors).
#include "cheb_nx.h"

namespace jkpm {
namespace step {

StepEntry pick_step_nx_c128(int rb, int nl, int k, int base, bool dbl) {
    if (rb == 4) return nx::entry<double, 1, 4>(nl, k, base, dbl);
    if (rb == 8) return nx::entry<double, 1, 8>(nl, k, base, dbl);
    return {nullptr, 0, 0, 1};
}

}  // namespace step
}  // namespace jkpm
