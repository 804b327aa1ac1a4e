// Chebyshev series application and the commutator recursion.
//
// Reference: rec_A_vec (jclsquant/recurrence_functions.py:114-143; GPU
// rec_A, jclsquant/cuda_cython/cuda/recursions.cu:267-317):
//     out = sum_{m<M} c_m T_m(A) v;
// rec_com_A_vec / rec_com_A_vec_tab (recurrence_functions.py:146-250; GPU
// rec_comm_A_vec, recursions.cu:319-388): with C_0 = 0, C_1 = V v,
//     C_m = 2 A C_{m-1} - C_{m-2} + 2 V T_{m-1},
//     out = sum_m c_m C_m   or   tab[m] = c_m C_m.
// The reference launches one SpMV plus one axpy kernel per step (two SpMVs
// and an axpy for the commutator).  Here each step is one kernel: the
// gathered T_{m-1} (and C_{m-1}) feed both recursions through one column
// index load (V lives on A's slots), and the accumulation is fused.
#include <algorithm>

#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;

template <typename Real, int N>
using vec_t = Real __attribute__((ext_vector_type(N)));

// acc modes
constexpr int kAccNone = 0;
constexpr int kAccAdd = 1;   // acc += c * y
constexpr int kAccInit = 2;  // acc = c0 * x_row + c * y   (first step of rec_A_vec)
constexpr int kAccSet = 3;   // acc = c * y                (first step of the commutator)

struct Coef {
    double tr, ti;     // c_m (acc / table factor)
    double t0r, t0i;   // c_0 for kAccInit
    double aT, bT;     // T_next = aT A T + bT T_prev
    double aHC, aVC, bC;  // C_next = aHC A C + aVC V T + bC C_prev
};

template <typename Real, int LPR, bool COMM>
__global__ __launch_bounds__(kBlock) void series_step_kernel(
    const Real *__restrict__ hval, const Real *__restrict__ vval, const int *__restrict__ col,
    int n, int k, const Real *__restrict__ x, const Real *tprev, Real *tout,
    const Real *__restrict__ cx, const Real *cprev, Real *cout, Real *acc, Real *tab, int ld,
    Coef cf, int acc_mode, int rows_per_chunk, int nchunks) {
    using C2 = vec_t<Real, 2>;
    constexpr int RPW = 64 / LPR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lv = lane % LPR, lr = lane / LPR;
    const int b = blockIdx.x;
    const int chunk = (nchunks % kXcd == 0) ? (b % kXcd) * (nchunks / kXcd) + b / kXcd : b;
    const long row_begin = (long)chunk * rows_per_chunk;
    const long row_end = min((long)n, row_begin + rows_per_chunk);
    for (long row = row_begin + wave * RPW + lr; row < row_end; row += kWaves * RPW) {
        double htr = 0, hti = 0, hcr = 0, hci = 0, vtr = 0, vti = 0;
        for (int s = 0; s < k; ++s) {
            const long e = (long)s * n + row;
            const long jo = ((long)col[e] * ld + lv) * 2;
            const C2 h = *reinterpret_cast<const C2 *>(hval + 2 * e);
            const C2 xt = *reinterpret_cast<const C2 *>(x + jo);
            htr += (double)h.x * xt.x - (double)h.y * xt.y;
            hti += (double)h.x * xt.y + (double)h.y * xt.x;
            if constexpr (COMM) {
                const C2 w = *reinterpret_cast<const C2 *>(vval + 2 * e);
                vtr += (double)w.x * xt.x - (double)w.y * xt.y;
                vti += (double)w.x * xt.y + (double)w.y * xt.x;
                if (cf.aHC != 0.0) {
                    const C2 xc = *reinterpret_cast<const C2 *>(cx + jo);
                    hcr += (double)h.x * xc.x - (double)h.y * xc.y;
                    hci += (double)h.x * xc.y + (double)h.y * xc.x;
                }
            }
        }
        const long vo = (row * ld + lv) * 2;
        double yr = cf.aT * htr, yi = cf.aT * hti;
        if (cf.bT != 0.0) {
            const C2 p = *reinterpret_cast<const C2 *>(tprev + vo);
            yr += cf.bT * p.x;
            yi += cf.bT * p.y;
        }
        C2 yt;
        yt.x = (Real)yr;
        yt.y = (Real)yi;
        *reinterpret_cast<C2 *>(tout + vo) = yt;
        C2 yv = yt;  // the series term: T_next or C_next
        if constexpr (COMM) {
            double cr = cf.aHC * hcr + cf.aVC * vtr, ci = cf.aHC * hci + cf.aVC * vti;
            if (cf.bC != 0.0) {
                const C2 p = *reinterpret_cast<const C2 *>(cprev + vo);
                cr += cf.bC * p.x;
                ci += cf.bC * p.y;
            }
            yv.x = (Real)cr;
            yv.y = (Real)ci;
            *reinterpret_cast<C2 *>(cout + vo) = yv;
        }
        const double tr = cf.tr * yv.x - cf.ti * yv.y, ti = cf.tr * yv.y + cf.ti * yv.x;
        if (tab) {
            C2 t;
            t.x = (Real)tr;
            t.y = (Real)ti;
            *reinterpret_cast<C2 *>(tab + vo) = t;
        }
        if (acc_mode != kAccNone) {
            double ar = tr, ai = ti;
            if (acc_mode == kAccAdd) {
                const C2 a = *reinterpret_cast<const C2 *>(acc + vo);
                ar += a.x;
                ai += a.y;
            } else if (acc_mode == kAccInit) {
                const C2 x0 = *reinterpret_cast<const C2 *>(x + vo);
                ar += cf.t0r * x0.x - cf.t0i * x0.y;
                ai += cf.t0r * x0.y + cf.t0i * x0.x;
            }
            C2 a;
            a.x = (Real)ar;
            a.y = (Real)ai;
            *reinterpret_cast<C2 *>(acc + vo) = a;
        }
    }
}

template <typename Real, bool COMM>
const void *pick_lpr(int lpr) {
    switch (lpr) {
        case 1: return reinterpret_cast<const void *>(&series_step_kernel<Real, 1, COMM>);
        case 2: return reinterpret_cast<const void *>(&series_step_kernel<Real, 2, COMM>);
        case 4: return reinterpret_cast<const void *>(&series_step_kernel<Real, 4, COMM>);
        case 8: return reinterpret_cast<const void *>(&series_step_kernel<Real, 8, COMM>);
        case 16: return reinterpret_cast<const void *>(&series_step_kernel<Real, 16, COMM>);
        case 32: return reinterpret_cast<const void *>(&series_step_kernel<Real, 32, COMM>);
        default: return nullptr;
    }
}

const void *pick(int vtype, bool comm, int lpr) {
    if (vtype == JKPM_C64) return comm ? pick_lpr<float, true>(lpr) : pick_lpr<float, false>(lpr);
    return comm ? pick_lpr<double, true>(lpr) : pick_lpr<double, false>(lpr);
}

// One step over all vector groups (powers of two up to 32, as the other
// step kernels; one vector per lane).
int step(const Ell &H, const Ell *V, const void *x, const void *tprev, void *tout, const void *cx,
         const void *cprev, void *cout, void *acc, void *tab, int R, const Coef &cf, int mode,
         int nchunks, hipStream_t s) {
    const size_t cb = cplx_bytes(H.vtype);
    for (int r0 = 0; r0 < R;) {
        int rb = 32;
        while (rb > R - r0) rb >>= 1;
        const void *fn = pick(H.vtype, V != nullptr, rb);
        if (!fn) return arg_error("unsupported vector group");
        const int rpb = kWaves * (64 / rb);
        long rpc = ((long)H.n + nchunks - 1) / nchunks;
        rpc = (rpc + rpb - 1) / rpb * rpb;
        const size_t off = (size_t)r0 * cb;
        auto sh = [&](const void *p) { return p ? (const void *)((const char *)p + off) : nullptr; };
        auto shm = [&](void *p) { return p ? (void *)((char *)p + off) : nullptr; };
        const void *a_h = H.val, *a_v = V ? V->val : nullptr;
        const int *a_c = H.col;
        int a_n = H.n, a_k = H.k;
        const void *a_x = sh(x), *a_tp = sh(tprev), *a_cx = sh(cx), *a_cp = sh(cprev);
        void *a_to = shm(tout), *a_co = shm(cout), *a_acc = shm(acc), *a_tab = shm(tab);
        int a_ld = R;
        Coef a_cf = cf;
        int a_mode = mode, a_rpc = (int)rpc, a_nch = nchunks;
        void *args[] = {&a_h,  &a_v,   &a_c,   &a_n,  &a_k,  &a_x,    &a_tp,  &a_to,   &a_cx,
                        &a_cp, &a_co,  &a_acc, &a_tab, &a_ld, &a_cf,  &a_mode, &a_rpc, &a_nch};
        JKPM_HIP(hipLaunchKernel(fn, dim3(nchunks), dim3(kBlock), args, 0, s));
        r0 += rb;
    }
    return JKPM_OK;
}

}  // namespace

size_t series_workspace(const Ell &H, int R, bool comm) {
    const size_t vec = align_up((size_t)H.n * R * cplx_bytes(H.vtype), 256);
    return (comm ? 4 : 2) * vec;
}

int run_series(const Ell &H, const Ell *V, const void *v, int R, int M, const double *c,
               void *out, void *tab, void *ws, int nchunks, hipStream_t s) {
    const size_t vec = align_up((size_t)H.n * R * cplx_bytes(H.vtype), 256);
    char *w = (char *)ws;
    void *Tb[2] = {w, w + vec};
    void *Cb[2] = {V ? w + 2 * vec : nullptr, V ? w + 3 * vec : nullptr};
    const bool comm = V != nullptr;
    const size_t vb = (size_t)H.n * R * cplx_bytes(H.vtype);
    if (comm && tab) {
        int rc = launch_zero_bytes(tab, vb, s);  // tab[0] = c_0 C_0 = 0
        if (rc) return rc;
    }
    auto tab_at = [&](int m) { return tab ? (void *)((char *)tab + (size_t)m * vb) : nullptr; };
    for (int m = 1; m < std::max(M, 2); ++m) {
        Coef cf{};
        const bool have_m = m < M;
        cf.tr = have_m ? c[2 * m] : 0.0;
        cf.ti = have_m ? c[2 * m + 1] : 0.0;
        cf.t0r = c[0];
        cf.t0i = c[1];
        cf.aT = m == 1 ? 1.0 : 2.0;
        cf.bT = m == 1 ? 0.0 : -1.0;
        cf.aHC = m == 1 ? 0.0 : 2.0;
        cf.aVC = m == 1 ? 1.0 : 2.0;
        cf.bC = m <= 2 ? 0.0 : -1.0;  // C_0 = 0
        const void *x = m == 1 ? v : Tb[(m - 1) & 1];
        const void *tprev = m == 1 ? nullptr : (m == 2 ? v : Tb[m & 1]);
        const void *cx = (comm && m >= 2) ? Cb[(m - 1) & 1] : nullptr;
        const void *cprev = (comm && m >= 3) ? Cb[m & 1] : nullptr;
        const int mode = out ? (m == 1 ? (comm ? kAccSet : kAccInit) : kAccAdd) : kAccNone;
        int rc = step(H, V, x, tprev, Tb[m & 1], cx, cprev, comm ? Cb[m & 1] : nullptr, out,
                      (comm && have_m) ? tab_at(m) : nullptr, R, cf, mode, nchunks, s);
        if (rc) return rc;
    }
    return JKPM_OK;
}

}  // namespace jkpm
