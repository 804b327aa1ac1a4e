// complex64 instantiations of the Chebyshev step kernel and their dispatch.
// Built three times, so the parts compile in parallel: as itself (NTM =
// false: every kernel, and pick_step_c64); included from cheb_c64_ntm.hip
// with JKPM_STEP_NTM_UNIT defined, for the tuned shapes with non-temporal
// matrix loads (pick_step_c64_ntm); and included from cheb_c64_rv.hip with
// JKPM_STEP_RV_UNIT defined, for the real-value twins of the tuned shapes
// (pick_step_c64_rv).
#include "cheb_step.h"

namespace jkpm {
namespace step {
namespace {

using Real = float;

#ifdef JKPM_STEP_NTM_UNIT
constexpr bool kNTM = true;
#else
constexpr bool kNTM = false;
#endif
#ifdef JKPM_STEP_RV_UNIT
constexpr bool kRV = true;
#else
constexpr bool kRV = false;
#endif

// Single-form kernels keep the first-step accumulators (F0) in every
// variant: without them the complex64 plain kernel measured slower (242.5
// against 208.3 us per launch on the headline workload at the same launch
// shape, with 68 against 80 VGPRs), and the complex128 ones gained nothing.
// The doubling-form kernels drop them (86 against 104 VGPRs, 5 waves per
// SIMD instead of 4, 179 against 186 us per launch).
template <int VPL, int LPR, int NL, int KT, int UNR, bool NT, bool NS = false, bool NTM = false,
          bool F0 = true>
const void *ptr() {
    return reinterpret_cast<const void *>(
        &cheb_step_kernel<Real, VPL, LPR, NL, KT, UNR, NT, false, NS, NTM, F0, kRV>);
}

template <int VPL, int LPR, int KT, int UNR, bool NT, bool NS = false, bool NTM = false,
          bool F0 = false>
const void *dbl_ptr() {
    return reinterpret_cast<const void *>(
        &cheb_step_kernel<Real, VPL, LPR, 2, KT, UNR, NT, true, NS, NTM, F0, kRV>);
}

#ifdef JKPM_STEP_RV_UNIT
// Real-value twins: the tuned single-form and doubling-form kernels of
// widths 4 and 5 and the doubling form's first-step kernel; NULL for the
// rest (they keep reading the complex values).
template <int VPL, int LPR, int KT>
const void *rv_tuned(int v, bool dbl) {
    const bool ns = v >= 12;
    switch (v & 3) {
        case 1:
            return dbl ? (ns ? dbl_ptr<VPL, LPR, KT, 2, false, true, kNTM>() : dbl_ptr<VPL, LPR, KT, 2, false, false, kNTM>())
                       : (ns ? ptr<VPL, LPR, 1, KT, 2, false, true, kNTM>() : ptr<VPL, LPR, 1, KT, 2, false, false, kNTM>());
        case 2:
            return dbl ? (ns ? dbl_ptr<VPL, LPR, KT, 1, true, true, kNTM>() : dbl_ptr<VPL, LPR, KT, 1, true, false, kNTM>())
                       : (ns ? ptr<VPL, LPR, 1, KT, 1, true, true, kNTM>() : ptr<VPL, LPR, 1, KT, 1, true, false, kNTM>());
        case 3:
            return dbl ? (ns ? dbl_ptr<VPL, LPR, KT, 2, true, true, kNTM>() : dbl_ptr<VPL, LPR, KT, 2, true, false, kNTM>())
                       : (ns ? ptr<VPL, LPR, 1, KT, 2, true, true, kNTM>() : ptr<VPL, LPR, 1, KT, 2, true, false, kNTM>());
        default:
            return dbl ? (ns ? dbl_ptr<VPL, LPR, KT, 1, false, true, kNTM>() : dbl_ptr<VPL, LPR, KT, 1, false, false, kNTM>())
                       : (ns ? ptr<VPL, LPR, 1, KT, 1, false, true, kNTM>() : ptr<VPL, LPR, 1, KT, 1, false, false, kNTM>());
    }
}

template <int VPL, int LPR>
StepEntry rv_entry(int nl, int k, int v, bool dbl, bool first) {
    if ((k != 4 && k != 5) || (!dbl && nl != 1)) return {nullptr, 0, 0, 1};
    if (first && dbl) {
        if constexpr (kNTM) return {nullptr, 0, 0, 1};
        const void *f = k == 4 ? dbl_ptr<VPL, LPR, 4, 1, false, false, false, true>()
                               : dbl_ptr<VPL, LPR, 5, 1, false, false, false, true>();
        return {f, VPL, LPR, 1, true};
    }
    const void *f = k == 4 ? rv_tuned<VPL, LPR, 4>(v, dbl) : rv_tuned<VPL, LPR, 5>(v, dbl);
    return {f, VPL, LPR, (v & 1) ? 2 : 1, true};
}
#else
// First-step kernels: the base kernel of the shape with the step-0
// accumulators (F0), whatever the tuned variant (one launch per recursion).
template <int VPL, int LPR>
StepEntry first_entry(int nl, int k, bool dbl) {
    const void *f;
    if (dbl)
        f = k == 4 ? dbl_ptr<VPL, LPR, 4, 1, false, false, false, true>()
          : k == 5 ? dbl_ptr<VPL, LPR, 5, 1, false, false, false, true>()
                   : dbl_ptr<VPL, LPR, 0, 1, false, false, false, true>();
    else if (nl == 2)
        f = k == 4 ? ptr<VPL, LPR, 2, 4, 1, false, false, false, true>()
          : k == 5 ? ptr<VPL, LPR, 2, 5, 1, false, false, false, true>()
                   : ptr<VPL, LPR, 2, 0, 1, false, false, false, true>();
    else
        f = k == 4 ? ptr<VPL, LPR, 1, 4, 1, false, false, false, true>()
          : k == 5 ? ptr<VPL, LPR, 1, 5, 1, false, false, false, true>()
                   : ptr<VPL, LPR, 1, 0, 1, false, false, false, true>();
    return {f, VPL, LPR, 1};
}

template <int VPL, int LPR, int KT, bool NS>
const void *dbl_tuned_ns(int v) {
    switch (v & 3) {
        case 1: return dbl_ptr<VPL, LPR, KT, 2, false, NS, kNTM>();
        case 2: return dbl_ptr<VPL, LPR, KT, 1, true, NS, kNTM>();
        case 3: return dbl_ptr<VPL, LPR, KT, 2, true, NS, kNTM>();
        default: return dbl_ptr<VPL, LPR, KT, 1, false, NS, kNTM>();
    }
}

template <int VPL, int LPR, int KT>
const void *dbl_tuned(int v) {
    return v >= 12 ? dbl_tuned_ns<VPL, LPR, KT, true>(v) : dbl_tuned_ns<VPL, LPR, KT, false>(v);
}

// Doubling-form kernels: (unroll, non-temporal) variants for widths 4, 5.
template <int VPL, int LPR>
StepEntry dbl_entry(int k, int v) {
    const int u = (v & 1) ? 2 : 1;
    if (k == 4) return {dbl_tuned<VPL, LPR, 4>(v), VPL, LPR, u};
    if (k == 5) return {dbl_tuned<VPL, LPR, 5>(v), VPL, LPR, u};
    if constexpr (kNTM) return {nullptr, 0, 0, 1};
    else return {dbl_ptr<VPL, LPR, 0, 1, false>(), VPL, LPR, 1};
}

template <int VPL, int LPR, int KT, bool NS>
const void *tuned_ns(int v) {
    switch (v & 3) {
        case 1: return ptr<VPL, LPR, 1, KT, 2, false, NS, kNTM>();
        case 2: return ptr<VPL, LPR, 1, KT, 1, true, NS, kNTM>();
        case 3: return ptr<VPL, LPR, 1, KT, 2, true, NS, kNTM>();
        default: return ptr<VPL, LPR, 1, KT, 1, false, NS, kNTM>();
    }
}

template <int VPL, int LPR, int KT>
const void *tuned(int v) {
    return v >= 12 ? tuned_ns<VPL, LPR, KT, true>(v) : tuned_ns<VPL, LPR, KT, false>(v);
}

// Base kernels of every (left sets, width) and the tuned ones for nl = 1,
// k = 4, 5.  `variants` = false instantiates only the base kernels.  The NTM
// unit has only the tuned ones (fn = NULL otherwise).
template <int VPL, int LPR, bool variants>
StepEntry entry(int nl, int k, int v, bool dbl, bool first) {
    if (first) {
        if constexpr (kNTM) return {nullptr, 0, 0, 1};
        else return first_entry<VPL, LPR>(nl, k, dbl);
    }
    if (dbl) {
        if constexpr (kNTM && !variants) return {nullptr, 0, 0, 1};
        else return dbl_entry<VPL, LPR>(k, variants ? v : 0);
    }
    if constexpr (variants) {
        if (nl == 1 && k == 4) return {tuned<VPL, LPR, 4>(v), VPL, LPR, (v & 1) ? 2 : 1};
        if (nl == 1 && k == 5) return {tuned<VPL, LPR, 5>(v), VPL, LPR, (v & 1) ? 2 : 1};
    }
    if constexpr (kNTM) {
        return {nullptr, 0, 0, 1};
    } else {
        if (nl == 2) {
            const void *f = k == 4 ? ptr<VPL, LPR, 2, 4, 1, false>()
                          : k == 5 ? ptr<VPL, LPR, 2, 5, 1, false>()
                                   : ptr<VPL, LPR, 2, 0, 1, false>();
            return {f, VPL, LPR, 1};
        }
        if constexpr (!variants) {
            if (k == 4) return {ptr<VPL, LPR, 1, 4, 1, false>(), VPL, LPR, 1};
            if (k == 5) return {ptr<VPL, LPR, 1, 5, 1, false>(), VPL, LPR, 1};
        }
        return {ptr<VPL, LPR, 1, 0, 1, false>(), VPL, LPR, 1};
    }
}

#endif  // JKPM_STEP_RV_UNIT

}  // namespace

namespace {

#ifdef JKPM_STEP_RV_UNIT
StepEntry pick_c64_rv(int rb, int nl, int k, int variant, bool dbl, bool first) {
    const int shape = first ? 0 : ((variant % 12) >> 2) & 3;
    int sh = shape;
    while (sh > 0 && (2 << sh) > rb) --sh;
    switch (rb) {
        case 2: return rv_entry<2, 1>(nl, k, variant, dbl, first);
        case 4: return sh == 0 ? rv_entry<2, 2>(nl, k, variant, dbl, first) : rv_entry<4, 1>(nl, k, variant, dbl, first);
        case 8:
            return sh == 0   ? rv_entry<2, 4>(nl, k, variant, dbl, first)
                   : sh == 1 ? rv_entry<4, 2>(nl, k, variant, dbl, first)
                             : rv_entry<8, 1>(nl, k, variant, dbl, first);
        case 16:
            return sh == 0   ? rv_entry<2, 8>(nl, k, variant, dbl, first)
                   : sh == 1 ? rv_entry<4, 4>(nl, k, variant, dbl, first)
                             : rv_entry<8, 2>(nl, k, variant, dbl, first);
        case 32:
            return sh == 0   ? rv_entry<2, 16>(nl, k, variant, dbl, first)
                   : sh == 1 ? rv_entry<4, 8>(nl, k, variant, dbl, first)
                             : rv_entry<8, 4>(nl, k, variant, dbl, first);
        default: return {nullptr, 0, 0, 1};
    }
}
#else
StepEntry pick_c64(int rb, bool wide, int nl, int k, int variant, bool dbl, bool first) {
    if (wide) {
        // 16-byte lanes and up: VPL = 2 << shape (capped by the group size)
        int shape = first ? 0 : ((variant % 12) >> 2) & 3;
        while (shape > 0 && (2 << shape) > rb) --shape;
        switch (rb) {
            case 2: return entry<2, 1, true>(nl, k, variant, dbl, first);
            case 4:
                return shape == 0 ? entry<2, 2, true>(nl, k, variant, dbl, first) : entry<4, 1, true>(nl, k, variant, dbl, first);
            case 8:
                return shape == 0   ? entry<2, 4, true>(nl, k, variant, dbl, first)
                       : shape == 1 ? entry<4, 2, true>(nl, k, variant, dbl, first)
                                    : entry<8, 1, true>(nl, k, variant, dbl, first);
            case 16:
                return shape == 0   ? entry<2, 8, true>(nl, k, variant, dbl, first)
                       : shape == 1 ? entry<4, 4, true>(nl, k, variant, dbl, first)
                                    : entry<8, 2, true>(nl, k, variant, dbl, first);
            case 32:
                return shape == 0   ? entry<2, 16, true>(nl, k, variant, dbl, first)
                       : shape == 1 ? entry<4, 8, true>(nl, k, variant, dbl, first)
                                    : entry<8, 4, true>(nl, k, variant, dbl, first);
            default: break;
        }
    }
    // 8-byte lanes (odd leading dimension or a single vector)
    switch (rb) {
        case 1: return entry<1, 1, false>(nl, k, 0, dbl, first);
        case 2: return entry<1, 2, false>(nl, k, 0, dbl, first);
        case 4: return entry<1, 4, false>(nl, k, 0, dbl, first);
        case 8: return entry<1, 8, false>(nl, k, 0, dbl, first);
        case 16: return entry<1, 16, false>(nl, k, 0, dbl, first);
        case 32: return entry<1, 32, false>(nl, k, 0, dbl, first);
        default: break;
    }
    return {nullptr, 0, 0, 1};
}
#endif  // JKPM_STEP_RV_UNIT

}  // namespace

#if defined(JKPM_STEP_RV_UNIT) && defined(JKPM_STEP_NTM_UNIT)
StepEntry pick_step_c64_rv_ntm(int rb, bool wide, int nl, int k, int variant, bool dbl) {
    // variant 0..23: the real-value twin of variant + 24
    if (!wide || variant < 0 || variant >= 24) return {nullptr, 0, 0, 1};
    return pick_c64_rv(rb, nl, k, variant, dbl, false);
}
#elif defined(JKPM_STEP_RV_UNIT)
StepEntry pick_step_c64_rv(int rb, bool wide, int nl, int k, int variant, bool dbl, bool first) {
    // 16-byte lanes and the variants without non-temporal matrix loads only
    if (!wide || variant < 0 || variant >= 24) return {nullptr, 0, 0, 1};
    // the doubling form's first launch takes the base shape (variant 0)
    if (first && dbl) return pick_c64_rv(rb, nl, k, 0, dbl, true);
    return pick_c64_rv(rb, nl, k, variant, dbl, false);
}
#elif defined(JKPM_STEP_NTM_UNIT)
StepEntry pick_step_c64_ntm(int rb, bool wide, int nl, int k, int variant, bool dbl) {
    if (!wide) return {nullptr, 0, 0, 1};
    return pick_c64(rb, wide, nl, k, variant, dbl, false);
}
#else
StepEntry pick_step_c64(int rb, bool wide, int nl, int k, int variant, bool dbl, bool first,
                        bool rv) {
    if (rv) {
        const StepEntry e = first && dbl  ? pick_step_c64_rv(rb, wide, nl, k, 0, dbl, true)
                            : variant >= 24 ? pick_step_c64_rv_ntm(rb, wide, nl, k, variant - 24, dbl)
                                            : pick_step_c64_rv(rb, wide, nl, k, variant, dbl, false);
        if (e.fn) return e;
    }
    // The doubling kernels lack the first-step accumulators: its first launch
    // takes the base kernel of the shape (the single-form kernels all have them).
    if (first && dbl) return pick_c64(rb, wide, nl, k, 0, dbl, true);
    if (variant >= 24) {
        const StepEntry e = pick_step_c64_ntm(rb, wide, nl, k, variant - 24, dbl);
        if (e.fn) return e;
        variant -= 24;
    }
    return pick_c64(rb, wide, nl, k, variant, dbl, false);
}
#endif

}  // namespace step
}  // namespace jkpm
