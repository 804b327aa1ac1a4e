// Kubo double Chebyshev moments mu_mn on MFMA, and the sigma(E) contraction.
//
// mu_mn = (1/R) sum_r <r| v T_m(H~) v T_n(H~) |r>  for Hermitian H~ and v
//       = (1/R) sum_k conj(L[m][k]) Rk[n][k],   k = (row, vector) = i*R + r,
// with the stacks  L_m = T_m(H~) (v r)  and  Rk_n = v T_n(H~) r  kept in HBM
// as complex64 [M][N*R] (one Chebyshev vector batch per stack slot).
//
// The reference has no such contraction: its nearest double expansion is
// position_operator (jclsquant/cuda_cython/cuda/obv.cu:562-644), O(M^2)
// SpMVs, and the coefficient tensor table_G_delta
// (jclsquant/cython_modules/Extra_funs.pyx:479-535), which is rank-1 per
// energy and is therefore never materialised here (kubo_sigma_kernel).
//
// Kernels
//   kubo_step_kernel   T_{n+1} = a H T_n + b T_{n-1} and, fused on the same
//                      gathered T_n, W_n = v T_n (v shares H's slot pattern).
//   kubo_gemm_kernel   split-K complex GEMM on v_mfma_f32_32x32x2f32: the
//                      MFMA's K = 2 spans two k; four MFMAs per k pair give
//                      re += Lr*Rr + Li*Ri and im += Lr*Ri - Li*Rr.
//                      Tiles of 128 x 128 complex outputs per work-group, 4
//                      waves of 64 x 64; each wave streams its operand rows
//                      from global memory into MFMA registers (no LDS, no
//                      barriers).  Work-groups of one K slice sit on one XCD
//                      so the slice is fetched from HBM once per XCD L2.
//   kubo_reduce_kernel FP64 sum of the FP32 slice partials in fixed order.
//   kubo_sigma_kernel  sigma_i = sum_mn (a_m b_n + c_m d_n) mu_mn per energy.
#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;

template <typename Real, int N>
using vec_t = Real __attribute__((ext_vector_type(N)));
using floatx16 = vec_t<float, 16>;

// ------------------------------------------------------------ stack steps

template <int VPL, int LPR, bool DUAL>
__global__ __launch_bounds__(kBlock) void kubo_step_kernel(
    const float *__restrict__ hval, const float *__restrict__ vval, const int *__restrict__ col,
    int n, int k, const float *__restrict__ x, const float *prev, float *out,
    float *__restrict__ wout, int ld, float alpha, float beta, int rows_per_chunk, int nchunks) {
    using V = vec_t<float, 2 * VPL>;
    using C2 = vec_t<float, 2>;
    constexpr int RPW = 64 / LPR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lv = lane % LPR, lr = lane / LPR;
    const int b = blockIdx.x;
    const int chunk = (nchunks % kXcd == 0) ? (b % kXcd) * (nchunks / kXcd) + b / kXcd : b;
    const long row_begin = (long)chunk * rows_per_chunk;
    const long row_end = min((long)n, row_begin + rows_per_chunk);
    const int voff = lv * VPL * 2;
    for (long row = row_begin + wave * RPW + lr; row < row_end; row += kWaves * RPW) {
        float sr[VPL], si[VPL], wr[VPL], wi[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) sr[v] = si[v] = wr[v] = wi[v] = 0.f;
        for (int s = 0; s < k; ++s) {
            const long e = (long)s * n + row;
            const int j = col[e];
            const C2 h = *reinterpret_cast<const C2 *>(hval + 2 * e);
            const V xv = *reinterpret_cast<const V *>(x + (long)j * ld * 2 + voff);
            C2 w;
            if constexpr (DUAL) w = *reinterpret_cast<const C2 *>(vval + 2 * e);
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                sr[v] += h.x * xv[2 * v] - h.y * xv[2 * v + 1];
                si[v] += h.x * xv[2 * v + 1] + h.y * xv[2 * v];
                if constexpr (DUAL) {
                    wr[v] += w.x * xv[2 * v] - w.y * xv[2 * v + 1];
                    wi[v] += w.x * xv[2 * v + 1] + w.y * xv[2 * v];
                }
            }
        }
        const long vo = row * ld * 2 + voff;
        V y;
        if (beta != 0.f) {
            const V p = *reinterpret_cast<const V *>(prev + vo);
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                y[2 * v] = alpha * sr[v] + beta * p[2 * v];
                y[2 * v + 1] = alpha * si[v] + beta * p[2 * v + 1];
            }
        } else {
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                y[2 * v] = alpha * sr[v];
                y[2 * v + 1] = alpha * si[v];
            }
        }
        *reinterpret_cast<V *>(out + vo) = y;
        if constexpr (DUAL) {
            V wv;
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                wv[2 * v] = wr[v];
                wv[2 * v + 1] = wi[v];
            }
            *reinterpret_cast<V *>(wout + vo) = wv;
        }
    }
}

template <int VPL, int LPR>
const void *step_fn(bool dual) {
    return dual ? reinterpret_cast<const void *>(&kubo_step_kernel<VPL, LPR, true>)
                : reinterpret_cast<const void *>(&kubo_step_kernel<VPL, LPR, false>);
}

struct KStep {
    const void *fn;
    int lpr;
};

KStep pick(int rb, bool wide, bool dual) {
    if (wide) {
        switch (rb) {
            case 2: return {step_fn<2, 1>(dual), 1};
            case 4: return {step_fn<2, 2>(dual), 2};
            case 8: return {step_fn<2, 4>(dual), 4};
            case 16: return {step_fn<2, 8>(dual), 8};
            case 32: return {step_fn<2, 16>(dual), 16};
            default: break;
        }
    }
    switch (rb) {
        case 1: return {step_fn<1, 1>(dual), 1};
        case 2: return {step_fn<1, 2>(dual), 2};
        case 4: return {step_fn<1, 4>(dual), 4};
        case 8: return {step_fn<1, 8>(dual), 8};
        case 16: return {step_fn<1, 16>(dual), 16};
        case 32: return {step_fn<1, 32>(dual), 32};
        default: break;
    }
    return {nullptr, 0};
}

// ------------------------------------------------------------ MFMA GEMM

constexpr int TM = 128;  // complex rows of L per work-group tile
constexpr int TN = 128;  // complex rows of Rk per work-group tile
constexpr int TK = 32;   // granularity of the split-K slices

constexpr int DK = 8;    // complex k per register batch (4 MFMA k-pairs)

// No LDS and no barriers: every wave streams its own operand rows from
// global memory (L2 serves the rows the wave pairs of a work-group share)
// straight into MFMA operand registers, with the next batch in flight while
// the current one is multiplied.  The MFMA's K = 2 spans two k: lane
// (l32, half) supplies element k0 + half * DK/2 + q of its row for the q-th
// k-pair, so each lane reads DK/2 consecutive complex numbers (two 16-byte
// loads) per row and batch.  Four MFMAs per k-pair and 32 x 32 block give
//   re += Lr Rr + Li Ri,   im += Lr Ri - Li Rr   (conj(L) Rk).
// ALIGNED (K even): the row starts are 16-byte aligned and loads go two
// complex numbers at a time.
template <bool ALIGNED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void kubo_gemm_kernel(
    const float2 *__restrict__ L, const float2 *__restrict__ Rk, long K, int M, int ntm,
    int ntn, int nslices, long k_per_slice, float2 *__restrict__ part) {
    constexpr int QP = DK / 2;  // k-pairs per batch
    const int ntiles = ntm * ntn;
    const int b = blockIdx.x;
    int slice, tile;
    if (nslices % kXcd == 0) {
        const int xcd = b % kXcd, local = b / kXcd;
        slice = (local / ntiles) * kXcd + xcd;
        tile = local % ntiles;
    } else {
        slice = b / ntiles;
        tile = b % ntiles;
    }
    const int m0 = (tile / ntn) * TM, n0 = (tile % ntn) * TN;
    const long k_begin = (long)slice * k_per_slice;
    const long k_end = min(K, k_begin + k_per_slice);

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;  // 2 x 2 waves of 64 x 64
    const int half = lane >> 5, l32 = lane & 31;

    const float2 *pa[2], *pb[2];
    bool va[2], vb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int m = m0 + wm * 64 + i * 32 + l32;
        va[i] = m < M;
        pa[i] = L + (long)(va[i] ? m : 0) * K;
        const int nn = n0 + wn * 64 + i * 32 + l32;
        vb[i] = nn < M;
        pb[i] = Rk + (long)(vb[i] ? nn : 0) * K;
    }

    floatx16 cre[2][2], cim[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) cre[i][j][r] = cim[i][j][r] = 0.f;

    // Three register batches in rotation: while batch t is multiplied,
    // batches t+1 and t+2 are in flight.
    float2 A0[2][QP], B0[2][QP], A1[2][QP], B1[2][QP], A2[2][QP], B2[2][QP];
    auto load = [&](long kc, float2 (&a)[2][QP], float2 (&bb)[2][QP]) {
        const long k0 = kc + half * QP;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if constexpr (ALIGNED) {
#pragma unroll
                for (int q = 0; q < QP; q += 2) {
                    const long kk = k0 + q;  // even, and k_end is even
                    float4 x = make_float4(0.f, 0.f, 0.f, 0.f), y = x;
                    if (kk < k_end) {
                        if (va[i]) x = *reinterpret_cast<const float4 *>(pa[i] + kk);
                        if (vb[i]) y = *reinterpret_cast<const float4 *>(pb[i] + kk);
                    }
                    a[i][q] = make_float2(x.x, x.y);
                    a[i][q + 1] = make_float2(x.z, x.w);
                    bb[i][q] = make_float2(y.x, y.y);
                    bb[i][q + 1] = make_float2(y.z, y.w);
                }
            } else {
#pragma unroll
                for (int q = 0; q < QP; ++q) {
                    const long kk = k0 + q;
                    const bool in = kk < k_end;
                    a[i][q] = (in && va[i]) ? pa[i][kk] : make_float2(0.f, 0.f);
                    bb[i][q] = (in && vb[i]) ? pb[i][kk] : make_float2(0.f, 0.f);
                }
            }
        }
    };
    // Eight independent accumulators between two updates of the same one.
    auto mul = [&](const float2 (&a)[2][QP], const float2 (&bb)[2][QP]) {
#pragma unroll
        for (int q = 0; q < QP; ++q) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const float2 x = a[i][q], c = bb[j][q];
                    cre[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.x, c.x, cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.x, c.y, cim[i][j], 0, 0, 0);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const float2 x = a[i][q], c = bb[j][q];
                    cre[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.y, c.y, cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.y, -c.x, cim[i][j], 0, 0, 0);
                }
        }
    };
    if (k_begin < k_end) load(k_begin, A0, B0);
    if (k_begin + DK < k_end) load(k_begin + DK, A1, B1);
    for (long kc = k_begin; kc < k_end; kc += 3 * DK) {
        if (kc + 2 * DK < k_end) load(kc + 2 * DK, A2, B2);
        mul(A0, B0);
        if (kc + DK >= k_end) break;
        if (kc + 3 * DK < k_end) load(kc + 3 * DK, A0, B0);
        mul(A1, B1);
        if (kc + 2 * DK >= k_end) break;
        if (kc + 4 * DK < k_end) load(kc + 4 * DK, A1, B1);
        mul(A2, B2);
    }

    // D[i][j] of a 32 x 32 block: lane l, register r -> j = l % 32,
    // i = 8 (r / 4) + 4 (l / 32) + r % 4.
    float2 *P = part + (long)slice * M * M;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * 64 + i * 32 + 8 * (r / 4) + 4 * half + (r % 4);
                const int cc = n0 + wn * 64 + j * 32 + l32;
                if (row < M && cc < M) P[(long)row * M + cc] = make_float2(cre[i][j][r], cim[i][j][r]);
            }
}

__global__ void kubo_reduce_kernel(const float2 *__restrict__ part, int nslices, long mm,
                                   double scale, double2 *__restrict__ mu) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm) return;
    double sr = 0.0, si = 0.0;
    for (int s = 0; s < nslices; ++s) {
        const float2 p = part[(long)s * mm + e];
        sr += p.x;
        si += p.y;
    }
    mu[e] = make_double2(sr * scale, si * scale);
}

// ------------------------------------------------------------ sigma(E)

// kind 0 (Kubo-Greenwood): a = b = g_m delta_m(e), c = d = 0.
// kind 1 (table_G_delta form, Extra_funs.pyx:430-535): a = G g / B,
//   b = delta g, c = -delta g, d = G g / B, with
//   G_m = -2 sin(m acos e) / s (G_0 = 0), delta_m = (2 - [m==0]) cos(m acos e)/(pi s),
//   s = sqrt(1 - e^2) (s = 1e-16 when it is 0).
__global__ __launch_bounds__(kBlock) void kubo_sigma_kernel(const double2 *__restrict__ mu, int M,
                                                             const double *__restrict__ e,
                                                             int kind, double bound,
                                                             double2 *__restrict__ sigma) {
    extern __shared__ double sh[];  // a, b, c, d: 4 * M doubles
    double *a = sh, *bq = sh + M, *c = sh + 2 * M, *d = sh + 3 * M;
    const int i = blockIdx.x;
    const double ei = e[i];
    const double th = acos(ei);
    double s = sqrt(1.0 - ei * ei);
    if (s == 0.0) s = 1e-16;
    const double q = M_PI / (M + 1.0);
    for (int m = threadIdx.x; m < M; m += kBlock) {
        const double g = ((M - m + 1.0) * cos(q * m) + sin(q * m) / tan(q)) / (M + 1.0);
        const double dl = (m == 0 ? 1.0 : 2.0) * cos(m * th) / (M_PI * s) * g;
        const double G = (m == 0 ? 0.0 : -2.0 * sin(m * th) / s) * g / bound;
        if (kind == 0) {
            a[m] = dl;
            bq[m] = dl;
            c[m] = 0.0;
            d[m] = 0.0;
        } else {
            a[m] = G;
            bq[m] = dl;
            c[m] = -dl;
            d[m] = G;
        }
    }
    __syncthreads();
    double sr = 0.0, si = 0.0;
    for (int m = threadIdx.x; m < M; m += kBlock) {
        double pr = 0.0, pi = 0.0, qr = 0.0, qi = 0.0;
        const double2 *row = mu + (long)m * M;
        for (int nn = 0; nn < M; ++nn) {
            const double2 v = row[nn];
            pr += v.x * bq[nn];
            pi += v.y * bq[nn];
            qr += v.x * d[nn];
            qi += v.y * d[nn];
        }
        sr += a[m] * pr + c[m] * qr;
        si += a[m] * pi + c[m] * qi;
    }
    __shared__ double rr[kBlock], ri[kBlock];
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            rr[threadIdx.x] += rr[threadIdx.x + o];
            ri[threadIdx.x] += ri[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) sigma[i] = make_double2(rr[0], ri[0]);
}

}  // namespace

int kubo_gemm_slices(long K, int M, const DeviceInfo &dev) {
    const int ntiles = ((M + TM - 1) / TM) * ((M + TN - 1) / TN);
    long want = ((long)dev.cu_count * 4 + ntiles - 1) / ntiles;  // ~4 work-groups per CU
    want = (want + kXcd - 1) / kXcd * kXcd;
    const long max_by_k = (K + TK - 1) / TK;
    if (want > max_by_k) want = max_by_k;
    if (want < 1) want = 1;
    return (int)want;
}

int launch_kubo_step(const Ell &H, const Ell *Vop, const void *x, const void *prev, void *out,
                     void *wout, int R, float alpha, float beta, int nchunks, hipStream_t s) {
    const bool dual = Vop != nullptr;
    for (int r0 = 0; r0 < R;) {
        int rb = 32;
        while (rb > R - r0) rb >>= 1;
        const bool wide = (R % 2 == 0) && (r0 % 2 == 0);
        KStep e = pick(rb, wide, dual);
        if (!e.fn) return arg_error("unsupported vector group");
        const int rpb = kWaves * (64 / e.lpr);
        long rpc = ((long)H.n + nchunks - 1) / nchunks;
        rpc = (rpc + rpb - 1) / rpb * rpb;
        const size_t off = (size_t)r0 * 8;
        const void *p_h = H.val;
        const void *p_v = dual ? Vop->val : nullptr;
        const int *p_c = H.col;
        int p_n = H.n, p_k = H.k;
        const void *p_x = (const char *)x + off;
        const void *p_prev = prev ? (const char *)prev + off : nullptr;
        void *p_out = (char *)out + off;
        void *p_w = wout ? (char *)wout + off : nullptr;
        int p_ld = R;
        float p_a = alpha, p_b = beta;
        int p_rpc = (int)rpc, p_nch = nchunks;
        void *args[] = {&p_h, &p_v, &p_c, &p_n, &p_k, &p_x, &p_prev, &p_out,
                        &p_w, &p_ld, &p_a, &p_b, &p_rpc, &p_nch};
        JKPM_HIP(hipLaunchKernel(e.fn, dim3(nchunks), dim3(kBlock), args, 0, s));
        r0 += rb;
    }
    return JKPM_OK;
}

int launch_kubo_gemm(const float2 *L, const float2 *Rk, long K, int M, int nslices, float2 *part,
                     hipStream_t s) {
    const int ntm = (M + TM - 1) / TM, ntn = (M + TN - 1) / TN;
    const long kps = ((K + nslices - 1) / nslices + TK - 1) / TK * TK;
    if (K % 2 == 0)
        hipLaunchKernelGGL(kubo_gemm_kernel<true>, dim3(ntm * ntn * nslices), dim3(kBlock), 0, s, L,
                           Rk, K, M, ntm, ntn, nslices, kps, part);
    else
        hipLaunchKernelGGL(kubo_gemm_kernel<false>, dim3(ntm * ntn * nslices), dim3(kBlock), 0, s, L,
                           Rk, K, M, ntm, ntn, nslices, kps, part);
    JKPM_CHECK_LAUNCH("kubo_gemm_kernel");
    return JKPM_OK;
}

int launch_kubo_reduce(const float2 *part, int nslices, int M, double scale, double2 *mu,
                       hipStream_t s) {
    const long mm = (long)M * M;
    hipLaunchKernelGGL(kubo_reduce_kernel, dim3((unsigned)((mm + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, part, nslices, mm, scale, mu);
    JKPM_CHECK_LAUNCH("kubo_reduce_kernel");
    return JKPM_OK;
}

int launch_kubo_sigma(const double2 *mu, int M, const double *e, int n_e, int kind, double bound,
                      double2 *sigma, hipStream_t s) {
    if (n_e <= 0) return JKPM_OK;
    const size_t shmem = (size_t)4 * M * sizeof(double);
    hipLaunchKernelGGL(kubo_sigma_kernel, dim3(n_e), dim3(kBlock), shmem, s, mu, M, e, kind, bound,
                       sigma);
    JKPM_CHECK_LAUNCH("kubo_sigma_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
