// Reverse-mode (adjoint) Chebyshev recursion: gradient of the moments with
// respect to the ELL values, with sqrt(M) checkpointing and recomputation.
//
// Forward:  T_0 = r, T_1 = A r, T_{n+1} = 2 A T_n - T_{n-1},
//           mu_n[r] = <r | T_n>.
// Given g_n[r] = dL/dRe(mu_n[r]) + i dL/dIm(mu_n[r]) (the gradient convention
// of torch for complex tensors), the adjoint vectors are
//           lambda_k = g_k r + 2 A^H lambda_{k+1} - lambda_{k+2},  k = M-1 .. 1,
// (lambda_M = lambda_{M+1} = 0; A^H = A for the Hermitian matrices this
// path differentiates) and the slot gradient is
//           G[i, s] = sum_r [ lambda_1 conj(T_0[j]) + 2 sum_{n=1}^{M-2} lambda_{n+1} conj(T_n[j]) ],
// j = col(i, s), i.e. G = dL/dRe(a_is) + i dL/dIm(a_is).
// There is no reference counterpart (the reference has no gradients).
//
// Kernels
//   plain_step_kernel  out = alpha A x + beta prev + coef[r] u  (no moments):
//                      forward recomputation and the adjoint step.
//   grad_fold_kernel   G[s][i] += sum_{k in segment} f_k sum_r
//                      lambda_k[i, r] conj(T_{k-1}[col(i, s), r]) for a whole
//                      segment of k at once, so G is read and written once per
//                      segment; FP64 accumulation, lane butterfly over r.
//   grad_out_kernel    slot-major G -> row-major, times the value scale.
#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;

template <typename Real, int N>
using vec_t = Real __attribute__((ext_vector_type(N)));

__device__ __forceinline__ int xcd_chunk(int b, int nchunks) {
    return (nchunks % kXcd == 0) ? (b % kXcd) * (nchunks / kXcd) + b / kXcd : b;
}

// RV: val holds one real per slot (the handle's packed real copy of a matrix
// whose imaginary parts are all 0), so each product takes two multiplies.
template <typename Real, int VPL, int LPR, bool RV = false>
__global__ __launch_bounds__(kBlock) void plain_step_kernel(
    const Real *__restrict__ val, const int *__restrict__ col, int n, int k,
    const Real *__restrict__ x, const Real *prev, Real *out, const double2 *__restrict__ coef,
    const Real *__restrict__ u, int ld, Real alpha, Real beta, int rows_per_chunk, int nchunks) {
    using V = vec_t<Real, 2 * VPL>;
    using C2 = vec_t<Real, 2>;
    constexpr int RPW = 64 / LPR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lv = lane % LPR, lr = lane / LPR;
    const int chunk = xcd_chunk(blockIdx.x, nchunks);
    const long row_begin = (long)chunk * rows_per_chunk;
    const long row_end = min((long)n, row_begin + rows_per_chunk);
    const int voff = lv * VPL * 2;
    Real cr[VPL], ci[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        cr[v] = coef ? (Real)coef[lv * VPL + v].x : Real(0);
        ci[v] = coef ? (Real)coef[lv * VPL + v].y : Real(0);
    }
    for (long row = row_begin + wave * RPW + lr; row < row_end; row += kWaves * RPW) {
        Real sr[VPL], si[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) sr[v] = si[v] = Real(0);
        for (int s = 0; s < k; ++s) {
            const long e = (long)s * n + row;
            const int j = col[e];
            const V xv = *reinterpret_cast<const V *>(x + (long)j * ld * 2 + voff);
            if constexpr (RV) {
                const Real h = val[e];
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    sr[v] += h * xv[2 * v];
                    si[v] += h * xv[2 * v + 1];
                }
            } else {
                const C2 h = *reinterpret_cast<const C2 *>(val + 2 * e);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    sr[v] += h.x * xv[2 * v] - h.y * xv[2 * v + 1];
                    si[v] += h.x * xv[2 * v + 1] + h.y * xv[2 * v];
                }
            }
        }
        const long vo = row * ld * 2 + voff;
        V y;
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            y[2 * v] = alpha * sr[v];
            y[2 * v + 1] = alpha * si[v];
        }
        if (beta != Real(0)) {
            const V p = *reinterpret_cast<const V *>(prev + vo);
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                y[2 * v] += beta * p[2 * v];
                y[2 * v + 1] += beta * p[2 * v + 1];
            }
        }
        if (coef) {
            const V q = *reinterpret_cast<const V *>(u + vo);
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                y[2 * v] += cr[v] * q[2 * v] - ci[v] * q[2 * v + 1];
                y[2 * v + 1] += cr[v] * q[2 * v + 1] + ci[v] * q[2 * v];
            }
        }
        *reinterpret_cast<V *>(out + vo) = y;
    }
}

// KT > 0 (ELL width KT): per row, each lambda_k row is read once and the KT
// gathered T rows of the same k are folded into KT register accumulators;
// KT = 0: any width, one slot at a time.
template <typename Real, int VPL, int LPR, int KT>
__global__ __launch_bounds__(kBlock) void grad_fold_kernel(
    const int *__restrict__ col, int n, int k, const Real *__restrict__ lam,
    const Real *__restrict__ tprev, long vstride, int count, int k_lo, double2 *__restrict__ G,
    int ld, int rows_per_chunk, int nchunks) {
    using V = vec_t<Real, 2 * VPL>;
    constexpr int RPW = 64 / LPR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lv = lane % LPR, lr = lane / LPR;
    const int chunk = xcd_chunk(blockIdx.x, nchunks);
    const long row_begin = (long)chunk * rows_per_chunk;
    const long row_end = min((long)n, row_begin + rows_per_chunk);
    const int voff = lv * VPL * 2;
    // Every lane of the wave runs the same number of iterations so that the
    // butterfly below always has all lanes present.
    const long span = row_end > row_begin ? row_end - row_begin : 0;
    const long iters = (span + kWaves * RPW - 1) / (kWaves * RPW);
    // lambda * conj(T), summed over this lane's vectors
    auto fold = [](const V &a, const V &t, double &pr, double &pi) {
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            const double a_r = a[2 * v], a_i = a[2 * v + 1];
            const double t_r = t[2 * v], t_i = t[2 * v + 1];
            pr += a_r * t_r + a_i * t_i;
            pi += a_i * t_r - a_r * t_i;
        }
    };
    for (long it = 0; it < iters; ++it) {
        const long row = row_begin + it * kWaves * RPW + wave * RPW + lr;
        const bool live = row < row_end;
        if constexpr (KT > 0) {
            double ar[KT], ai[KT];
            int j[KT];
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                ar[s] = ai[s] = 0.0;
                j[s] = live ? col[(long)s * n + row] : 0;
            }
            if (live) {
                // unrolled so that several moments' loads are in flight at once
#pragma unroll 4
                for (int q = 0; q < count; ++q) {
                    const double f = (k_lo + q == 1) ? 1.0 : 2.0;
                    const V lv_ = __builtin_nontemporal_load(reinterpret_cast<const V *>(
                        lam + q * vstride * 2 + row * ld * 2 + voff));
                    V tv[KT];
#pragma unroll
                    for (int s = 0; s < KT; ++s)
                        tv[s] = *reinterpret_cast<const V *>(tprev + q * vstride * 2 +
                                                             (long)j[s] * ld * 2 + voff);
#pragma unroll
                    for (int s = 0; s < KT; ++s) {
                        double pr = 0.0, pi = 0.0;
                        fold(lv_, tv[s], pr, pi);
                        ar[s] += f * pr;
                        ai[s] += f * pi;
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < KT; ++s) {
#pragma unroll
                for (int o = 1; o < LPR; o <<= 1) {
                    ar[s] += __shfl_xor(ar[s], o);
                    ai[s] += __shfl_xor(ai[s], o);
                }
                if (live && lv == 0) {
                    double2 g = G[(long)s * n + row];
                    g.x += ar[s];
                    g.y += ai[s];
                    G[(long)s * n + row] = g;
                }
            }
        } else {
            for (int s = 0; s < k; ++s) {
                double ar = 0.0, ai = 0.0;
                if (live) {
                    const int j = col[(long)s * n + row];
                    for (int q = 0; q < count; ++q) {
                        const double f = (k_lo + q == 1) ? 1.0 : 2.0;
                        const V lv_ = *reinterpret_cast<const V *>(lam + q * vstride * 2 +
                                                                   row * ld * 2 + voff);
                        const V tv = *reinterpret_cast<const V *>(tprev + q * vstride * 2 +
                                                                  (long)j * ld * 2 + voff);
                        double pr = 0.0, pi = 0.0;
                        fold(lv_, tv, pr, pi);
                        ar += f * pr;
                        ai += f * pi;
                    }
                }
#pragma unroll
                for (int o = 1; o < LPR; o <<= 1) {
                    ar += __shfl_xor(ar, o);
                    ai += __shfl_xor(ai, o);
                }
                if (live && lv == 0) {
                    double2 g = G[(long)s * n + row];
                    g.x += ar;
                    g.y += ai;
                    G[(long)s * n + row] = g;
                }
            }
        }
    }
}

__global__ void grad_out_kernel(const double2 *__restrict__ G, int n, int k, double scale,
                                double2 *__restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * k) return;
    const long i = e / k, s = e % k;
    const double2 g = G[s * n + i];
    out[e] = make_double2(g.x * scale, g.y * scale);
}

__global__ void zero_kernel(double2 *__restrict__ p, long cnt) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) p[e] = make_double2(0.0, 0.0);
}

// Clears cnt 16-byte units and, if tail8, the 8 bytes after them (a vector
// batch of an odd number of complex64 elements).
__global__ void zero_tail_kernel(double2 *__restrict__ p, long cnt, int tail8) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) p[e] = make_double2(0.0, 0.0);
    else if (e == cnt && tail8) reinterpret_cast<double *>(p + cnt)[0] = 0.0;
}

template <typename Real, int VPL, int LPR>
struct Fns {
    static const void *step() { return reinterpret_cast<const void *>(&plain_step_kernel<Real, VPL, LPR>); }
    static const void *step_rv() {
        return reinterpret_cast<const void *>(&plain_step_kernel<Real, VPL, LPR, true>);
    }
    static const void *fold(int k) {
        switch (k) {
            case 4: return reinterpret_cast<const void *>(&grad_fold_kernel<Real, VPL, LPR, 4>);
            case 5: return reinterpret_cast<const void *>(&grad_fold_kernel<Real, VPL, LPR, 5>);
            default: return reinterpret_cast<const void *>(&grad_fold_kernel<Real, VPL, LPR, 0>);
        }
    }
};

struct Pick {
    const void *step, *fold;
    int lpr;
    const void *step_rv = nullptr;  // the step on the packed real values
};

// wide (an even batch R, so rows and groups start on two-vector
// boundaries): groups of 2+ vectors run two vectors per lane (16-byte
// complex64 or 32-byte complex128 lanes); otherwise one vector per lane.
// The step kernel runs wide when it can, the fold never.
template <typename Real>
Pick pick_real(int rb, int k, bool wide) {
    if (wide) {
        switch (rb) {
            case 2: return {Fns<Real, 2, 1>::step(), Fns<Real, 2, 1>::fold(k), 1, Fns<Real, 2, 1>::step_rv()};
            case 4: return {Fns<Real, 2, 2>::step(), Fns<Real, 2, 2>::fold(k), 2, Fns<Real, 2, 2>::step_rv()};
            case 8: return {Fns<Real, 2, 4>::step(), Fns<Real, 2, 4>::fold(k), 4, Fns<Real, 2, 4>::step_rv()};
            case 16: return {Fns<Real, 2, 8>::step(), Fns<Real, 2, 8>::fold(k), 8, Fns<Real, 2, 8>::step_rv()};
            case 32: return {Fns<Real, 2, 16>::step(), Fns<Real, 2, 16>::fold(k), 16, Fns<Real, 2, 16>::step_rv()};
            default: break;
        }
    }
    switch (rb) {
        case 1: return {Fns<Real, 1, 1>::step(), Fns<Real, 1, 1>::fold(k), 1, Fns<Real, 1, 1>::step_rv()};
        case 2: return {Fns<Real, 1, 2>::step(), Fns<Real, 1, 2>::fold(k), 2, Fns<Real, 1, 2>::step_rv()};
        case 4: return {Fns<Real, 1, 4>::step(), Fns<Real, 1, 4>::fold(k), 4, Fns<Real, 1, 4>::step_rv()};
        case 8: return {Fns<Real, 1, 8>::step(), Fns<Real, 1, 8>::fold(k), 8, Fns<Real, 1, 8>::step_rv()};
        case 16: return {Fns<Real, 1, 16>::step(), Fns<Real, 1, 16>::fold(k), 16, Fns<Real, 1, 16>::step_rv()};
        case 32: return {Fns<Real, 1, 32>::step(), Fns<Real, 1, 32>::fold(k), 32, Fns<Real, 1, 32>::step_rv()};
        default: return {nullptr, nullptr, 0};
    }
}

Pick pick(int vtype, int rb, int k, bool wide) {
    return vtype == JKPM_C64 ? pick_real<float>(rb, k, wide) : pick_real<double>(rb, k, wide);
}

long rows_per_chunk(int n, int nchunks, int lpr) {
    const int rpb = kWaves * (64 / lpr);
    long rpc = ((long)n + nchunks - 1) / nchunks;
    return (rpc + rpb - 1) / rpb * rpb;
}

}  // namespace

int launch_plain_step(const Ell &H, const void *x, const void *prev, void *out,
                      const double2 *coef, const void *u, int R, double alpha, double beta,
                      int nchunks, hipStream_t s) {
    const size_t cb = cplx_bytes(H.vtype);
    for (int r0 = 0; r0 < R;) {
        int rb = 32;
        while (rb > R - r0) rb >>= 1;
        Pick p = pick(H.vtype, rb, H.k, R % 2 == 0);
        if (!p.step) return arg_error("unsupported vector group");
        // a real matrix: the step on its packed real values
        const bool rv = H.real_values && H.rval && p.step_rv;
        const size_t off = (size_t)r0 * cb;
        const void *a_val = rv ? H.rval : H.val;
        const int *a_col = H.col;
        int a_n = H.n, a_k = H.k;
        const void *a_x = (const char *)x + off;
        const void *a_prev = prev ? (const char *)prev + off : nullptr;
        void *a_out = (char *)out + off;
        const double2 *a_coef = coef ? coef + r0 : nullptr;
        const void *a_u = u ? (const char *)u + off : nullptr;
        int a_ld = R;
        int a_rpc = (int)rows_per_chunk(H.n, nchunks, p.lpr), a_nch = nchunks;
        float fa = (float)alpha, fb = (float)beta;
        double da = alpha, db = beta;
        void *args_f[] = {&a_val, &a_col, &a_n, &a_k, &a_x, &a_prev, &a_out, &a_coef,
                          &a_u,   &a_ld,  &fa,  &fb,  &a_rpc, &a_nch};
        void *args_d[] = {&a_val, &a_col, &a_n, &a_k, &a_x, &a_prev, &a_out, &a_coef,
                          &a_u,   &a_ld,  &da,  &db,  &a_rpc, &a_nch};
        JKPM_HIP(hipLaunchKernel(rv ? p.step_rv : p.step, dim3(nchunks), dim3(kBlock),
                                 H.vtype == JKPM_C64 ? args_f : args_d, 0, s));
        r0 += rb;
    }
    return JKPM_OK;
}

int launch_grad_fold(const Ell &H, const void *lam, const void *tprev, int count, int k_lo,
                     double2 *G, int R, int nchunks, hipStream_t s) {
    if (count <= 0) return JKPM_OK;
    const size_t cb = cplx_bytes(H.vtype);
    for (int r0 = 0; r0 < R;) {
        int rb = 32;
        while (rb > R - r0) rb >>= 1;
        // one vector per lane: faster for the fold (530 vs 679 us per
        // 32-moment fold, square 512^2, R = 8, complex128)
        Pick p = pick(H.vtype, rb, H.k, false);
        if (!p.fold) return arg_error("unsupported vector group");
        const size_t off = (size_t)r0 * cb;
        const int *a_col = H.col;
        int a_n = H.n, a_k = H.k;
        const void *a_lam = (const char *)lam + off;
        const void *a_t = (const char *)tprev + off;
        long a_vs = (long)H.n * R;
        int a_count = count, a_klo = k_lo;
        double2 *a_G = G;
        int a_ld = R;
        int a_rpc = (int)rows_per_chunk(H.n, nchunks, p.lpr), a_nch = nchunks;
        void *args[] = {&a_col, &a_n, &a_k, &a_lam, &a_t, &a_vs, &a_count, &a_klo,
                        &a_G,   &a_ld, &a_rpc, &a_nch};
        JKPM_HIP(hipLaunchKernel(p.fold, dim3(nchunks), dim3(kBlock), args, 0, s));
        r0 += rb;
    }
    return JKPM_OK;
}

int launch_grad_out(const double2 *G, int n, int k, double scale, double2 *out, hipStream_t s) {
    const long tot = (long)n * k;
    if (tot == 0) return JKPM_OK;
    hipLaunchKernelGGL(grad_out_kernel, dim3((unsigned)((tot + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, G, n, k, scale, out);
    JKPM_CHECK_LAUNCH("grad_out_kernel");
    return JKPM_OK;
}

int launch_zero(double2 *p, long cnt, hipStream_t s) {
    if (cnt <= 0) return JKPM_OK;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((cnt + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, p, cnt);
    JKPM_CHECK_LAUNCH("zero_kernel");
    return JKPM_OK;
}

int launch_zero_bytes(void *p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return JKPM_OK;
    if (bytes % 8 != 0 || ((uintptr_t)p & 15u) != 0) {
        set_error("launch_zero_bytes: 16-byte aligned start and a multiple of 8 bytes required");
        return JKPM_ERR_ARG;
    }
    const long cnt = (long)(bytes / 16);
    const int tail8 = (bytes % 16) != 0;
    const long threads = cnt + tail8;
    hipLaunchKernelGGL(zero_tail_kernel, dim3((unsigned)((threads + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, (double2 *)p, cnt, tail8);
    JKPM_CHECK_LAUNCH("zero_tail_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
