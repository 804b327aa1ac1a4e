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
//   kubo_gemm_kernel   split-K complex GEMM on v_mfma_f32_16x16x4f32: four
//                      MFMAs per k-set give re += Lr*Rr + Li*Ri and
//                      im += Lr*Ri - Li*Rr.  Tiles of 128 x 128 complex
//                      outputs per work-group, 4 waves of 64 x 64; each wave
//                      streams its operand rows from global memory into
//                      MFMA registers (no LDS, no barriers).  Work-groups of
//                      one K slice sit on one XCD so the slice is fetched
//                      from HBM once per XCD L2.
//   kubo_gemm3_kernel  the same GEMM with three real products per complex
//                      one (Gauss); the default.
//   kubo_reduce_kernel FP64 sum of the FP32 slice partials in fixed order.
//   kubo_sigma_kernel  sigma_i = sum_mn (a_m b_n + c_m d_n) mu_mn per energy.
#include <cstdlib>

#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;

template <typename Real, int N>
using vec_t = Real __attribute__((ext_vector_type(N)));
using floatx4 = vec_t<float, 4>;

// ------------------------------------------------------------ stack steps

// LEFT (with DUAL only): a second, independent recursion on the same column
// and H-value loads, outl = alpha_l H xl + beta_l prevl (prevl unused when
// beta_l == 0).  The Kubo driver runs the right recursion (T_n, v T_n) and the
// left one (L_m) side by side this way, so the slots are read once per step
// for both.
// HR: hval holds one real per slot (the packed real copy of a real H).
// VI: vval holds one real w per slot with v = i w (the packed imaginary parts
// of a purely imaginary v, such as the velocity of a real H).
template <int VPL, int LPR, bool DUAL, bool LEFT, bool HR = false, bool VI = false>
__global__ __launch_bounds__(kBlock) void kubo_step_kernel(
    const float *__restrict__ hval, const float *__restrict__ vval, const int *__restrict__ col,
    int n, int k, const float *__restrict__ x, const float *prev, float *out,
    float *__restrict__ wout, const float *__restrict__ xl, const float *prevl,
    float *__restrict__ outl, int ld, float alpha, float beta, float alpha_l, float beta_l,
    int rows_per_chunk, int nchunks) {
    static_assert(DUAL || !LEFT, "the left recursion rides on the dual step");
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
        float sr[VPL], si[VPL], wr[VPL], wi[VPL], qr[VPL], qi[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) sr[v] = si[v] = wr[v] = wi[v] = qr[v] = qi[v] = 0.f;
        for (int s = 0; s < k; ++s) {
            const long e = (long)s * n + row;
            const int j = col[e];
            // (h, w) as complex numbers; the packed forms load one real
            C2 h, w;
            if constexpr (HR) {
                h.x = hval[e];
                h.y = 0.f;
            } else {
                h = *reinterpret_cast<const C2 *>(hval + 2 * e);
            }
            if constexpr (DUAL) {
                if constexpr (VI) {
                    w.x = 0.f;
                    w.y = vval[e];
                } else {
                    w = *reinterpret_cast<const C2 *>(vval + 2 * e);
                }
            }
            const V xv = *reinterpret_cast<const V *>(x + (long)j * ld * 2 + voff);
            V xq;
            if constexpr (LEFT) xq = *reinterpret_cast<const V *>(xl + (long)j * ld * 2 + voff);
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                if constexpr (HR) {
                    sr[v] += h.x * xv[2 * v];
                    si[v] += h.x * xv[2 * v + 1];
                } else {
                    sr[v] += h.x * xv[2 * v] - h.y * xv[2 * v + 1];
                    si[v] += h.x * xv[2 * v + 1] + h.y * xv[2 * v];
                }
                if constexpr (DUAL) {
                    if constexpr (VI) {
                        wr[v] -= w.y * xv[2 * v + 1];
                        wi[v] += w.y * xv[2 * v];
                    } else {
                        wr[v] += w.x * xv[2 * v] - w.y * xv[2 * v + 1];
                        wi[v] += w.x * xv[2 * v + 1] + w.y * xv[2 * v];
                    }
                }
                if constexpr (LEFT) {
                    if constexpr (HR) {
                        qr[v] += h.x * xq[2 * v];
                        qi[v] += h.x * xq[2 * v + 1];
                    } else {
                        qr[v] += h.x * xq[2 * v] - h.y * xq[2 * v + 1];
                        qi[v] += h.x * xq[2 * v + 1] + h.y * xq[2 * v];
                    }
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
        if constexpr (LEFT) {
            V yl;
            if (beta_l != 0.f) {
                const V p = *reinterpret_cast<const V *>(prevl + vo);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    yl[2 * v] = alpha_l * qr[v] + beta_l * p[2 * v];
                    yl[2 * v + 1] = alpha_l * qi[v] + beta_l * p[2 * v + 1];
                }
            } else {
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    yl[2 * v] = alpha_l * qr[v];
                    yl[2 * v + 1] = alpha_l * qi[v];
                }
            }
            *reinterpret_cast<V *>(outl + vo) = yl;
        }
    }
}

template <int VPL, int LPR, bool HR, bool VI>
const void *step_fn_v(bool dual, bool left) {
    if (left) return reinterpret_cast<const void *>(&kubo_step_kernel<VPL, LPR, true, true, HR, VI>);
    return dual ? reinterpret_cast<const void *>(&kubo_step_kernel<VPL, LPR, true, false, HR, VI>)
                : reinterpret_cast<const void *>(&kubo_step_kernel<VPL, LPR, false, false, HR>);
}

// hr: H's packed real values; vi: v's packed imaginary parts (dual only)
template <int VPL, int LPR>
const void *step_fn(bool dual, bool left, bool hr, bool vi) {
    vi = vi && dual;
    if (hr) return vi ? step_fn_v<VPL, LPR, true, true>(dual, left)
                      : step_fn_v<VPL, LPR, true, false>(dual, left);
    return vi ? step_fn_v<VPL, LPR, false, true>(dual, left)
              : step_fn_v<VPL, LPR, false, false>(dual, left);
}

struct KStep {
    const void *fn;
    int lpr;
};

KStep pick(int rb, bool wide, bool dual, bool left, bool hr, bool vi) {
    if (wide) {
        switch (rb) {
            case 2: return {step_fn<2, 1>(dual, left, hr, vi), 1};
            case 4: return {step_fn<2, 2>(dual, left, hr, vi), 2};
            case 8: return {step_fn<2, 4>(dual, left, hr, vi), 4};
            case 16: return {step_fn<2, 8>(dual, left, hr, vi), 8};
            case 32: return {step_fn<2, 16>(dual, left, hr, vi), 16};
            default: break;
        }
    }
    switch (rb) {
        case 1: return {step_fn<1, 1>(dual, left, hr, vi), 1};
        case 2: return {step_fn<1, 2>(dual, left, hr, vi), 2};
        case 4: return {step_fn<1, 4>(dual, left, hr, vi), 4};
        case 8: return {step_fn<1, 8>(dual, left, hr, vi), 8};
        case 16: return {step_fn<1, 16>(dual, left, hr, vi), 16};
        case 32: return {step_fn<1, 32>(dual, left, hr, vi), 32};
        default: break;
    }
    return {nullptr, 0};
}

// ------------------------------------------------------------ MFMA GEMM

constexpr int TM = 128;  // complex rows of L per work-group tile
constexpr int TN = 128;  // complex rows of Rk per work-group tile
constexpr int TK = 32;   // granularity of the split-K slices

constexpr int DK = 8;    // complex k per register batch (two MFMA k-sets of 4)

// No LDS and no barriers: every wave streams its own operand rows from
// global memory (L2 serves the rows the wave pairs of a work-group share)
// straight into MFMA operand registers (v_mfma_f32_16x16x4f32), with the
// next two batches in flight while the current one is multiplied.
// Lane (l16, quarter) of a 16 x 16 block supplies row l16 and the quarter's
// two consecutive complex k of each DK = 8 batch, so one 16-byte load per
// row tile covers 16 rows x 64 contiguous bytes: every 64-byte sector a
// load touches is fully used by that load.  (With v_mfma_f32_32x32x2f32 a
// load touches 32 rows with 32 bytes each, every sector is requested twice,
// and the GEMM ran at half the rate.)  The two complex numbers of a lane
// are the K = 4 of two MFMAs (k-sets {0,2,4,6} and {1,3,5,7} of the batch);
// four MFMAs per k-set and block give
//   re += Lr Rr + Li Ri,   im += Lr Ri - Li Rr   (conj(L) Rk).
// Wave tile 64 x 64 complex = 4 x 4 blocks; D of a block: lane l, register
// r -> row 4 (l / 16) + r, column l % 16.
// ALIGNED (K even): the row starts are 16-byte aligned and loads go two
// complex numbers at a time.
template <bool ALIGNED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void kubo_gemm_kernel(
    const float2 *__restrict__ L, const float2 *__restrict__ Rk, long K, int M, int ntm,
    int ntn, int nslices, long k_per_slice, float2 *__restrict__ part) {
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
    const int quarter = lane >> 4, l16 = lane & 15;

    // Rows past M read row 0: their products only reach output rows /
    // columns that are never stored.
    const float2 *pa[4], *pb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + wm * 64 + i * 16 + l16;
        pa[i] = L + (long)(m < M ? m : 0) * K;
        const int nn = n0 + wn * 64 + i * 16 + l16;
        pb[i] = Rk + (long)(nn < M ? nn : 0) * K;
    }

    floatx4 cre[4][4], cim[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) cre[i][j][r] = cim[i][j][r] = 0.f;

    // Three register batches in rotation, as in kubo_gemm_kernel: plain
    // loads in the main loop (a prefetch past k_end reads the slice's last
    // valid k and is never multiplied), zero-masked loads for the at most
    // three batches that reach past k_end.
    float4 A0[4], B0[4], A1[4], B1[4], A2[4], B2[4];
    const long k_last = k_end - 2;  // last two-complex load start (ALIGNED)
    auto load = [&](long kc, float4 (&a)[4], float4 (&bb)[4], bool masked) {
        const long kk = kc + quarter * 2;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if constexpr (ALIGNED) {
                const long kc_ = kk < k_end ? kk : k_last;
                float4 x = *reinterpret_cast<const float4 *>(pa[i] + kc_);
                float4 y = *reinterpret_cast<const float4 *>(pb[i] + kc_);
                if (masked && kk >= k_end) x = y = make_float4(0.f, 0.f, 0.f, 0.f);
                a[i] = x;
                bb[i] = y;
            } else {
                const long k0 = kk < k_end ? kk : k_end - 1;
                const long k1 = kk + 1 < k_end ? kk + 1 : k_end - 1;
                float2 x0 = pa[i][k0], x1 = pa[i][k1], y0 = pb[i][k0], y1 = pb[i][k1];
                if (masked && kk >= k_end) x0 = y0 = make_float2(0.f, 0.f);
                if (masked && kk + 1 >= k_end) x1 = y1 = make_float2(0.f, 0.f);
                a[i] = make_float4(x0.x, x0.y, x1.x, x1.y);
                bb[i] = make_float4(y0.x, y0.y, y1.x, y1.y);
            }
        }
    };
    // Sixteen independent accumulator pairs between two updates of one.
    auto mul = [&](const float4 (&a)[4], const float4 (&bb)[4]) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float ar = g ? a[i].z : a[i].x, br = g ? bb[j].z : bb[j].x;
                    const float bi = g ? bb[j].w : bb[j].y;
                    cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cim[i][j], 0, 0, 0);
                }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float ai = g ? a[i].w : a[i].y, br = g ? bb[j].z : bb[j].x;
                    const float bi = g ? bb[j].w : bb[j].y;
                    cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, bi, cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, -br, cim[i][j], 0, 0, 0);
                }
        }
    };
    if (k_begin < k_end) {
        const long nb = (k_end - k_begin + DK - 1) / DK;
        const long nfull = (k_end - k_begin) / DK;
        const long n3 = nfull / 3 * 3;
        load(k_begin, A0, B0, false);
        __builtin_amdgcn_sched_barrier(0);
        load(k_begin + DK, A1, B1, false);
        __builtin_amdgcn_sched_barrier(0);
        for (long t = 0; t < n3; t += 3) {
            const long kc = k_begin + t * DK;
            load(kc + 2 * DK, A2, B2, false);
            __builtin_amdgcn_sched_barrier(0);
            mul(A0, B0);
            __builtin_amdgcn_sched_barrier(0);
            load(kc + 3 * DK, A0, B0, false);
            __builtin_amdgcn_sched_barrier(0);
            mul(A1, B1);
            __builtin_amdgcn_sched_barrier(0);
            load(kc + 4 * DK, A1, B1, false);
            __builtin_amdgcn_sched_barrier(0);
            mul(A2, B2);
            __builtin_amdgcn_sched_barrier(0);
        }
        for (long t = n3; t < nb; ++t) {
            load(k_begin + t * DK, A2, B2, true);
            mul(A2, B2);
        }
    }

    float2 *P = part + (long)slice * M * M;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + 4 * quarter + r;
                const int cc = n0 + wn * 64 + j * 16 + l16;
                if (row < M && cc < M) P[(long)row * M + cc] = make_float2(cre[i][j][r], cim[i][j][r]);
            }
}

// The same contraction with three real products per complex one (Gauss):
// for conj(a) b,
//   k1 = (ar - ai) br,   k2 = ar (bi - br),   k3 = ai (br + bi),
//   re = k1 + k3 = ar br + ai bi,   im = k1 + k2 = ar bi - ai br,
// so each block takes three MFMAs per k-set instead of four.  The third
// accumulator set (192 registers) does not fit two waves per SIMD, so this
// kernel runs one wave per SIMD, accumulators in AGPRs, with NB3 register
// batches in rotation and NB3 - 1 of them in flight (3 measured fastest of
// 3..6).  Operand layout, tiles and slice mapping are those of
// kubo_gemm_kernel.  The error against the FP64 oracle is that of the
// four-product form: both are set by the complex64 stacks.
constexpr int NB3 = 3;

// WN: 16-column blocks per wave along n.  Only 4 (wave tile 64 x 64, one
// wave per SIMD) is launched: 2 (64 x 32, whose 96 accumulators fit two
// waves per SIMD) measured 55.1 against 42.9 ms per Kubo step, its operand
// rows feeding half as many MFMAs per load.
template <bool ALIGNED, int WN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WN == 4 ? 1 : 2, WN == 4 ? 1 : 2))) void kubo_gemm3_kernel(
    const float2 *__restrict__ L, const float2 *__restrict__ Rk, long K, int M, int ntm,
    int ntn, int nslices, long k_per_slice, float2 *__restrict__ part) {
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
    constexpr int TN3 = 2 * 16 * WN;
    const int m0 = (tile / ntn) * TM, n0 = (tile % ntn) * TN3;
    const long k_begin = (long)slice * k_per_slice;
    const long k_end = min(K, k_begin + k_per_slice);

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int quarter = lane >> 4, l16 = lane & 15;

    const float2 *pa[4], *pb[WN];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + wm * 64 + i * 16 + l16;
        pa[i] = L + (long)(m < M ? m : 0) * K;
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
        const int nn = n0 + wn * 16 * WN + j * 16 + l16;
        pb[j] = Rk + (long)(nn < M ? nn : 0) * K;
    }

    floatx4 k1[4][WN], k2[4][WN], k3[4][WN];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) k1[i][j][r] = k2[i][j][r] = k3[i][j][r] = 0.f;

    float4 A[NB3][4], B[NB3][WN];
    const long k_last = k_end - 2;
    auto load = [&](long kc, float4 (&a)[4], float4 (&bb)[WN], bool masked) {
        const long kk = kc + quarter * 2;
        auto one = [&](const float2 *row) -> float4 {
            if constexpr (ALIGNED) {
                const long kc_ = kk < k_end ? kk : k_last;
                float4 x = *reinterpret_cast<const float4 *>(row + kc_);
                if (masked && kk >= k_end) x = make_float4(0.f, 0.f, 0.f, 0.f);
                return x;
            } else {
                const long k0 = kk < k_end ? kk : k_end - 1;
                const long k1_ = kk + 1 < k_end ? kk + 1 : k_end - 1;
                float2 x0 = row[k0], x1 = row[k1_];
                if (masked && kk >= k_end) x0 = make_float2(0.f, 0.f);
                if (masked && kk + 1 >= k_end) x1 = make_float2(0.f, 0.f);
                return make_float4(x0.x, x0.y, x1.x, x1.y);
            }
        };
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = one(pa[i]);
#pragma unroll
        for (int j = 0; j < WN; ++j) bb[j] = one(pb[j]);
    };
    auto mul = [&](const float4 (&a)[4], const float4 (&bb)[WN]) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            float ar[4], ai[4], d[4], br[WN], u[WN], w[WN];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ar[i] = g ? a[i].z : a[i].x;
                ai[i] = g ? a[i].w : a[i].y;
                d[i] = ar[i] - ai[i];
            }
#pragma unroll
            for (int j = 0; j < WN; ++j) {
                br[j] = g ? bb[j].z : bb[j].x;
                const float bi = g ? bb[j].w : bb[j].y;
                u[j] = bi - br[j];
                w[j] = br[j] + bi;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    k1[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(d[i], br[j], k1[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    k2[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], u[j], k2[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    k3[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], w[j], k3[i][j], 0, 0, 0);
        }
    };
    if (k_begin < k_end) {
        const long nb = (k_end - k_begin + DK - 1) / DK;
        const long nfull = (k_end - k_begin) / DK;
        const long nmain = nfull / NB3 * NB3;
        // Prefetches past the last full batch read clamped, in-range k and
        // are never multiplied; the tail reloads with zero masks.
#pragma unroll
        for (int u = 0; u < NB3 - 1; ++u) {
            load(k_begin + u * DK, A[u], B[u], false);
            __builtin_amdgcn_sched_barrier(0);
        }
        for (long t = 0; t < nmain; t += NB3) {
#pragma unroll
            for (int u = 0; u < NB3; ++u) {
                load(k_begin + (t + u + NB3 - 1) * DK, A[(u + NB3 - 1) % NB3],
                     B[(u + NB3 - 1) % NB3], false);
                __builtin_amdgcn_sched_barrier(0);
                mul(A[u], B[u]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        for (long t = nmain; t < nb; ++t) {
            load(k_begin + t * DK, A[0], B[0], true);
            mul(A[0], B[0]);
        }
    }

    float2 *P = part + (long)slice * M * M;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + 4 * quarter + r;
                const int cc = n0 + wn * 16 * WN + j * 16 + l16;
                if (row < M && cc < M)
                    P[(long)row * M + cc] =
                        make_float2(k1[i][j][r] + k3[i][j][r], k1[i][j][r] + k2[i][j][r]);
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
    // ~16 work-groups per CU over the launch (8 resident): short slices keep
    // the 16 tiles of a slice, which share its rows and run side by side on
    // one XCD, close together in k, so the rows they share stay in L2.
    long want = ((long)dev.cu_count * 16 + ntiles - 1) / ntiles;
    // JKPM_KUBO_SLICES overrides the split-K slice count (tuning)
    if (const char *env = getenv("JKPM_KUBO_SLICES")) {
        const long v = atol(env);
        if (v > 0) want = v;
    }
    want = (want + kXcd - 1) / kXcd * kXcd;
    const long max_by_k = (K + TK - 1) / TK;
    if (want > max_by_k) want = max_by_k;
    if (want < 1) want = 1;
    return (int)want;
}

int launch_kubo_step(const Ell &H, const Ell *Vop, const void *x, const void *prev, void *out,
                     void *wout, int R, float alpha, float beta, int nchunks, hipStream_t s,
                     const KuboLeft *left) {
    const bool dual = Vop != nullptr;
    if (left && !dual) return arg_error("internal: the left recursion needs the dual step");
    for (int r0 = 0; r0 < R;) {
        int rb = 32;
        while (rb > R - r0) rb >>= 1;
        const bool wide = (R % 2 == 0) && (r0 % 2 == 0);
        // packed values where the handles have them: a real H, an imaginary v
        const bool hr = H.real_values && H.rval;
        const bool vi = dual && Vop->imag_values && Vop->rval;
        KStep e = pick(rb, wide, dual, left != nullptr, hr, vi);
        if (!e.fn) return arg_error("unsupported vector group");
        const int rpb = kWaves * (64 / e.lpr);
        long rpc = ((long)H.n + nchunks - 1) / nchunks;
        rpc = (rpc + rpb - 1) / rpb * rpb;
        const size_t off = (size_t)r0 * 8;
        const void *p_h = hr ? H.rval : H.val;
        const void *p_v = dual ? (vi ? Vop->rval : Vop->val) : nullptr;
        const int *p_c = H.col;
        int p_n = H.n, p_k = H.k;
        const void *p_x = (const char *)x + off;
        const void *p_prev = prev ? (const char *)prev + off : nullptr;
        void *p_out = (char *)out + off;
        void *p_w = wout ? (char *)wout + off : nullptr;
        const void *p_xl = left ? (const char *)left->x + off : nullptr;
        const void *p_prevl = left && left->prev ? (const char *)left->prev + off : nullptr;
        void *p_outl = left ? (char *)left->out + off : nullptr;
        int p_ld = R;
        float p_a = alpha, p_b = beta;
        float p_al = left ? left->alpha : 0.f, p_bl = left ? left->beta : 0.f;
        int p_rpc = (int)rpc, p_nch = nchunks;
        void *args[] = {&p_h,  &p_v,     &p_c,    &p_n,  &p_k,   &p_x,  &p_prev,
                        &p_out, &p_w,    &p_xl,   &p_prevl, &p_outl, &p_ld, &p_a,
                        &p_b,  &p_al,    &p_bl,   &p_rpc, &p_nch};
        JKPM_HIP(hipLaunchKernel(e.fn, dim3(nchunks), dim3(kBlock), args, 0, s));
        r0 += rb;
    }
    return JKPM_OK;
}

int launch_kubo_gemm(const float2 *L, const float2 *Rk, long K, int M, int nslices, float2 *part,
                     hipStream_t s) {
    const int ntm = (M + TM - 1) / TM, ntn = (M + TN - 1) / TN;
    const long kps = ((K + nslices - 1) / nslices + TK - 1) / TK * TK;
    // JKPM_KUBO_GEMM=4 selects the four-product kernel (comparison, tuning)
    const char *env = getenv("JKPM_KUBO_GEMM");
    const bool four = env && atoi(env) == 4;
    const bool al = K % 2 == 0;
    const dim3 grid(ntm * ntn * nslices), block(kBlock);
    if (four) {
        if (al)
            hipLaunchKernelGGL(kubo_gemm_kernel<true>, grid, block, 0, s, L, Rk, K, M, ntm, ntn,
                               nslices, kps, part);
        else
            hipLaunchKernelGGL(kubo_gemm_kernel<false>, grid, block, 0, s, L, Rk, K, M, ntm, ntn,
                               nslices, kps, part);
    } else if (al) {
        hipLaunchKernelGGL((kubo_gemm3_kernel<true, 4>), grid, block, 0, s, L, Rk, K, M, ntm, ntn,
                           nslices, kps, part);
    } else {
        hipLaunchKernelGGL((kubo_gemm3_kernel<false, 4>), grid, block, 0, s, L, Rk, K, M, ntm, ntn,
                           nslices, kps, part);
    }
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
