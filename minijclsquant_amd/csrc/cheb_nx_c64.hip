// complex64 neighbour-exchange step kernels (variants 48..51): two vectors
// per lane, 4 or 8 lanes per row (groups of 8 and 16 vectors).
#include "cheb_nx.h"

namespace jkpm {
namespace step {

StepEntry pick_step_nx_c64(int rb, bool wide, int nl, int k, int base, bool dbl) {
    if (!wide) return {nullptr, 0, 0, 1};
    if (rb == 8) return nx::entry<float, 2, 4>(nl, k, base, dbl);
    if (rb == 16) return nx::entry<float, 2, 8>(nl, k, base, dbl);
    return {nullptr, 0, 0, 1};
}

}  // namespace step
}  // namespace jkpm
