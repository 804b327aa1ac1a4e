// Neighbour-exchange (NX) twins of the step kernel: launch variants 48..51,
// instantiated for one value type per translation unit (cheb_nx_c64.hip,
// cheb_nx_c128.hip) so the two compile in parallel.
#pragma once

#include "cheb_step.h"

namespace jkpm {
namespace step {
namespace nx {

// Single-form kernels keep the first-step accumulators (F0), as every
// single-form variant does; the doubling-form ones have none (the first step
// of a doubling recursion runs the base kernel).
template <typename Real, int VPL, int LPR, int KT, bool NT, bool NS>
const void *ptr(bool dbl) {
    if (dbl)
        return reinterpret_cast<const void *>(
            &cheb_step_kernel<Real, VPL, LPR, 2, KT, 1, NT, true, NS, false, false, true>);
    return reinterpret_cast<const void *>(
        &cheb_step_kernel<Real, VPL, LPR, 1, KT, 1, NT, false, NS, false, true, true>);
}

// base: the gathering variant the kernel twins (0, 2, 12 or 14)
template <typename Real, int VPL, int LPR, int KT>
const void *of_base(int base, bool dbl) {
    switch (base) {
        case 0: return ptr<Real, VPL, LPR, KT, false, false>(dbl);
        case 2: return ptr<Real, VPL, LPR, KT, true, false>(dbl);
        case 12: return ptr<Real, VPL, LPR, KT, false, true>(dbl);
        case 14: return ptr<Real, VPL, LPR, KT, true, true>(dbl);
        default: return nullptr;
    }
}

template <typename Real, int VPL, int LPR>
StepEntry entry(int nl, int k, int base, bool dbl) {
    if (nl != 1 && !dbl) return {nullptr, 0, 0, 1};
    const void *f = k == 4 ? of_base<Real, VPL, LPR, 4>(base, dbl)
                  : k == 5 ? of_base<Real, VPL, LPR, 5>(base, dbl)
                           : nullptr;
    return {f, VPL, LPR, 1};
}

}  // namespace nx
}  // namespace step
}  // namespace jkpm
