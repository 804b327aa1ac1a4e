// Two Chebyshev steps per launch over LDS-resident patches (temporal blocking,
// the "matrix powers" scheme of Demmel, Hoemmen, Mohiyuddin and Yelick,
// "Avoiding communication in sparse matrix computations", IPDPS 2008).
//
// The rows are renumbered into patches: contiguous row ranges whose rows are
// mostly each other's neighbours (grown breadth-first on the matrix graph).
// A patch's halo is the set of rows outside it that its rows reference.  One
// work-group owns one patch and, per launch:
//   phase 1  T_{n+1} = a1 A T_n + b1 T_{n-1} on the patch AND its halo (the
//            halo rows are recomputed redundantly from T_n, T_{n-1} in HBM),
//            kept in LDS; the patch rows of T_{n+1} also go to HBM;
//   phase 2  T_{n+2} = 2 A T_{n+1} - T_n on the patch rows, gathering
//            T_{n+1} from LDS through patch-local column numbers;
//            moment partials of both steps (and of step 0 on the first
//            launch) in FP64.
// Per two steps the matrix, T_n, T_{n-1} and v_l are read once (plus the
// halo's share) instead of twice, and T_{n+1} is never gathered from HBM.
// The doubling form loads T_{n-1} and stores both outputs non-temporally
// (measured faster there; the plain form is faster with cached accesses).
// Algorithmic bytes per launch (two steps): N K (s_v + 4) + R N 5 s_x
// (T_{n-1}, T_n, v_l read; T_{n+1}, T_{n+2} written).
//
// Moments are invariant under the renumbering (<P v|P A P^T|P v> = <v|A|v>),
// so the vectors are permuted once on entry and the moments need no
// un-permute.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"

namespace jkpm {
namespace {

constexpr int kTB = 512;              // threads per work-group
constexpr int kTWaves = kTB / 64;
constexpr int kXcd = 8;
constexpr int kLdsBytes = 64 * 1024;  // LDS tile of T_{n+1} per work-group

template <typename Real, int N>
using vec_t = Real __attribute__((ext_vector_type(N)));

// ---------------------------------------------------------------- kernels

// val_t[s][i] = val[s][perm[i]], col_t[s][i] = inv[col[s][perm[i]]].
template <typename Real>
__global__ void tile_permute_kernel(const Real *__restrict__ val, const int *__restrict__ col,
                                    const int *__restrict__ perm, const int *__restrict__ inv,
                                    int n, int k, Real *__restrict__ val_t,
                                    int *__restrict__ col_t, int with_cols) {
    const long tot = (long)n * k;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += stride) {
        const long s = e / n, i = e % n;
        const long src = s * n + perm[i];
        val_t[2 * e] = val[2 * src];
        val_t[2 * e + 1] = val[2 * src + 1];
        if (with_cols) col_t[e] = inv[col[src]];
    }
}

// dst[i][r] = src[perm[i]][r] for R interleaved complex vectors.
template <typename Real>
__global__ void vec_permute_kernel(const Real *__restrict__ src, const int *__restrict__ perm,
                                   int n, int R, Real *__restrict__ dst) {
    const long tot = (long)n * R;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += stride) {
        const long i = e / R, r = e % R;
        const long o = (long)perm[i] * R + r;
        dst[2 * e] = src[2 * o];
        dst[2 * e + 1] = src[2 * o + 1];
    }
}

template <bool NT, typename V, typename Real>
__device__ __forceinline__ V load_in(const Real *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
    else
        return *reinterpret_cast<const V *>(p);
}

template <bool NT, typename V, typename Real>
__device__ __forceinline__ void store_out(const V &y, Real *p) {
    if constexpr (NT)
        __builtin_nontemporal_store(y, reinterpret_cast<V *>(p));
    else
        *reinterpret_cast<V *>(p) = y;
}

// Accumulator slots: 0 = step 0 (first launch only), 1 = step n+1, 2 = step
// n+2.  Plain form (NLL = 1): <v_l|T>.  Doubling form (NLL = 2): slot s holds
// (a_s, b_{s-1}) with a_k = <T_k|T_k>, b_k = <T_k|T_{k+1}>; slot 0 holds
// (a_0, a_0).
//
// A thread handles the same rows in both phases (row e = its lane row +
// pass * STRIDE), so the patch rows of T_n it loads in phase 1 stay in
// registers for phase 2 instead of being re-read after the caches have moved
// on.  Halo rows read their matrix slots from per-patch contiguous copies
// (hval/hcol) rather than scattered rows of the matrix.
template <typename Real, int VPL, int LPR, int KT, bool DBL>
__global__ __launch_bounds__(kTB) void tile2_kernel(
    const Real *__restrict__ val, const int *__restrict__ col, const short *__restrict__ lcol,
    const int *__restrict__ pstart, const int *__restrict__ hoff, const int *__restrict__ hrows,
    const Real *__restrict__ hval, const int *__restrict__ hcol, int htot, int n, int k_rt,
    int npatch, const Real *__restrict__ xprev, const Real *__restrict__ xcur,
    Real *__restrict__ out1, Real *__restrict__ out2, const Real *__restrict__ vl, int ld,
    double2 *__restrict__ part0, double2 *__restrict__ part1, double2 *__restrict__ part2,
    Real a1, Real b1, int first) {
    using V = vec_t<Real, 2 * VPL>;
    using C2 = vec_t<Real, 2>;
    constexpr int RPW = 64 / LPR;
    constexpr int RB = VPL * LPR;
    constexpr int STRIDE = kTWaves * RPW;
    constexpr int NLL = DBL ? 2 : 1;
    constexpr int ROWR = RB * 2;  // Reals per LDS row
    constexpr int CAP = kLdsBytes / (int)(ROWR * sizeof(Real));
    constexpr int NP = (CAP + STRIDE - 1) / STRIDE;  // passes over a full tile
    const int K = KT > 0 ? KT : k_rt;

    __shared__ __align__(16) Real tile[kLdsBytes / sizeof(Real)];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lv = lane % LPR, lr = lane / LPR;
    const int b = blockIdx.x;
    const int p = (npatch % kXcd == 0) ? (b % kXcd) * (npatch / kXcd) + b / kXcd : b;
    const int r0 = pstart[p], nown = pstart[p + 1] - r0;
    const int h0 = hoff[p], next = nown + (hoff[p + 1] - h0);
    const int voff = lv * VPL * 2;
    const int e0 = wave * RPW + lr;

    V tn[NP];  // T_n of this thread's patch rows

    // ---- phase 1: T_{n+1} on the patch and its halo -> LDS (+ HBM for the patch)
    // Row e of pass ip: its matrix slots come from the patch rows (val/col) or
    // the halo copies (hval/hcol); g is its row in the vectors.
    auto row_of = [&](int e, long &g, const Real *&mv, const int *&mc, long &ms, long &mr) {
        const bool own = e < nown;
        const long hq = own ? 0 : (long)h0 + (e - nown);
        g = own ? (long)r0 + e : (long)hrows[hq];
        mv = own ? val : hval;
        mc = own ? col : hcol;
        ms = own ? (long)n : (long)htot;
        mr = own ? g : hq;
    };
    if constexpr (KT > 0) {
        // Software pipeline: the slots of pass ip+1 load while pass ip gathers.
        int j[KT];
        C2 h[KT];
        if (e0 < next) {
            long g, ms, mr;
            const Real *mv;
            const int *mc;
            row_of(e0, g, mv, mc, ms, mr);
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const long q = (long)s * ms + mr;
                j[s] = mc[q];
                h[s] = *reinterpret_cast<const C2 *>(mv + 2 * q);
            }
        }
#pragma unroll
        for (int ip = 0; ip < NP; ++ip) {
            const int e = e0 + ip * STRIDE;
            if (e >= next) break;
            const bool own = e < nown;
            long g, ms, mr;
            const Real *mv;
            const int *mc;
            row_of(e, g, mv, mc, ms, mr);
            V x[KT];
#pragma unroll
            for (int s = 0; s < KT; ++s)
                x[s] = *reinterpret_cast<const V *>(xcur + (long)j[s] * ld * 2 + voff);
            V pv;
            if (!first) pv = load_in<DBL, V>(xprev + g * ld * 2 + voff);
            if (own) tn[ip] = *reinterpret_cast<const V *>(xcur + g * ld * 2 + voff);
            C2 hc[KT];
#pragma unroll
            for (int s = 0; s < KT; ++s) hc[s] = h[s];
            const int en = e + STRIDE;
            if (ip + 1 < NP && en < next) {
                long g2, ms2, mr2;
                const Real *mv2;
                const int *mc2;
                row_of(en, g2, mv2, mc2, ms2, mr2);
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const long q = (long)s * ms2 + mr2;
                    j[s] = mc2[q];
                    h[s] = *reinterpret_cast<const C2 *>(mv2 + 2 * q);
                }
            }
            Real sr[VPL], si[VPL];
#pragma unroll
            for (int v = 0; v < VPL; ++v) sr[v] = si[v] = Real(0);
#pragma unroll
            for (int s = 0; s < KT; ++s)
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    sr[v] += hc[s].x * x[s][2 * v] - hc[s].y * x[s][2 * v + 1];
                    si[v] += hc[s].x * x[s][2 * v + 1] + hc[s].y * x[s][2 * v];
                }
            V y;
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                y[2 * v] = first ? a1 * sr[v] : a1 * sr[v] + b1 * pv[2 * v];
                y[2 * v + 1] = first ? a1 * si[v] : a1 * si[v] + b1 * pv[2 * v + 1];
            }
            *reinterpret_cast<V *>(tile + (long)e * ROWR + voff) = y;
            if (own) store_out<DBL>(y, out1 + g * ld * 2 + voff);
        }
    } else {
#pragma unroll
        for (int ip = 0; ip < NP; ++ip) {
            const int e = e0 + ip * STRIDE;
            if (e >= next) break;
            const bool own = e < nown;
            long g, ms, mr;
            const Real *mv;
            const int *mc;
            row_of(e, g, mv, mc, ms, mr);
            Real sr[VPL], si[VPL];
#pragma unroll
            for (int v = 0; v < VPL; ++v) sr[v] = si[v] = Real(0);
            for (int s = 0; s < K; ++s) {
                const long q = (long)s * ms + mr;
                const int j = mc[q];
                const C2 h = *reinterpret_cast<const C2 *>(mv + 2 * q);
                const V x = *reinterpret_cast<const V *>(xcur + (long)j * ld * 2 + voff);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    sr[v] += h.x * x[2 * v] - h.y * x[2 * v + 1];
                    si[v] += h.x * x[2 * v + 1] + h.y * x[2 * v];
                }
            }
            if (own) tn[ip] = *reinterpret_cast<const V *>(xcur + g * ld * 2 + voff);
            V y;
            if (first) {
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    y[2 * v] = a1 * sr[v];
                    y[2 * v + 1] = a1 * si[v];
                }
            } else {
                const V pv = load_in<DBL, V>(xprev + g * ld * 2 + voff);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    y[2 * v] = a1 * sr[v] + b1 * pv[2 * v];
                    y[2 * v + 1] = a1 * si[v] + b1 * pv[2 * v + 1];
                }
            }
            *reinterpret_cast<V *>(tile + (long)e * ROWR + voff) = y;
            if (own) store_out<DBL>(y, out1 + g * ld * 2 + voff);
        }
    }
    __syncthreads();

    // ---- phase 2: T_{n+2} on the patch rows from the LDS tile; partials
    double acc[3][NLL][VPL][2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int l = 0; l < NLL; ++l)
#pragma unroll
            for (int v = 0; v < VPL; ++v) acc[a][l][v][0] = acc[a][l][v][1] = 0.0;

    // acc += conj(a) * c
    auto dot = [](double (&d)[2], const V &a, const V &c, int v) {
        const double ar = a[2 * v], ai = a[2 * v + 1], cr = c[2 * v], ci = c[2 * v + 1];
        d[0] += ar * cr + ai * ci;
        d[1] += ar * ci - ai * cr;
    };
#pragma unroll
    for (int ip = 0; ip < NP; ++ip) {
        const int e = e0 + ip * STRIDE;
        if (e >= nown) break;
        const long g = (long)r0 + e;
        Real sr[VPL], si[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) sr[v] = si[v] = Real(0);
        if constexpr (KT > 0) {
            int j[KT];
            C2 h[KT];
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const long q = (long)s * n + g;
                j[s] = lcol[q];
                h[s] = *reinterpret_cast<const C2 *>(val + 2 * q);
            }
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const V x = *reinterpret_cast<const V *>(tile + (long)j[s] * ROWR + voff);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    sr[v] += h[s].x * x[2 * v] - h[s].y * x[2 * v + 1];
                    si[v] += h[s].x * x[2 * v + 1] + h[s].y * x[2 * v];
                }
            }
        } else {
            for (int s = 0; s < K; ++s) {
                const long q = (long)s * n + g;
                const int j = lcol[q];
                const C2 h = *reinterpret_cast<const C2 *>(val + 2 * q);
                const V x = *reinterpret_cast<const V *>(tile + (long)j * ROWR + voff);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    sr[v] += h.x * x[2 * v] - h.y * x[2 * v + 1];
                    si[v] += h.x * x[2 * v + 1] + h.y * x[2 * v];
                }
            }
        }
        const long vo = g * ld * 2 + voff;
        const V t1 = *reinterpret_cast<const V *>(tile + (long)e * ROWR + voff);
        V y;
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            y[2 * v] = Real(2) * sr[v] - tn[ip][2 * v];
            y[2 * v + 1] = Real(2) * si[v] - tn[ip][2 * v + 1];
        }
        store_out<DBL>(y, out2 + vo);
        if constexpr (DBL) {
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                dot(acc[1][0][v], t1, t1, v);  // a_{n+1}
                dot(acc[1][1][v], tn[ip], t1, v);  // b_n
                dot(acc[2][0][v], y, y, v);    // a_{n+2}
                dot(acc[2][1][v], t1, y, v);   // b_{n+1}
                if (first) {
                    dot(acc[0][0][v], tn[ip], tn[ip], v);  // a_0
                    dot(acc[0][1][v], tn[ip], tn[ip], v);
                }
            }
        } else {
            const V q = *reinterpret_cast<const V *>(vl + vo);
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                dot(acc[1][0][v], q, t1, v);
                dot(acc[2][0][v], q, y, v);
                if (first) dot(acc[0][0][v], q, tn[ip], v);
            }
        }
    }

    // Butterfly over the lanes that hold the same vector index, then a fixed
    // wave order through LDS (the tile is no longer needed).
#pragma unroll
    for (int o = LPR; o < 64; o <<= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int l = 0; l < NLL; ++l)
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    acc[a][l][v][0] += __shfl_xor(acc[a][l][v][0], o);
                    acc[a][l][v][1] += __shfl_xor(acc[a][l][v][1], o);
                }
    __syncthreads();
    double2 *sh = reinterpret_cast<double2 *>(tile);  // [3][kTWaves][NLL][RB]
    if (lr == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int l = 0; l < NLL; ++l)
#pragma unroll
                for (int v = 0; v < VPL; ++v)
                    sh[((a * kTWaves + wave) * NLL + l) * RB + lv * VPL + v] =
                        make_double2(acc[a][l][v][0], acc[a][l][v][1]);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * NLL * RB; t += kTB) {
        const int a = t / (NLL * RB), l = (t / RB) % NLL, r = t % RB;
        if (a == 0 && !first) continue;
        double sr2 = 0.0, si2 = 0.0;
#pragma unroll
        for (int w = 0; w < kTWaves; ++w) {
            const double2 q = sh[((a * kTWaves + w) * NLL + l) * RB + r];
            sr2 += q.x;
            si2 += q.y;
        }
        double2 *dst = a == 0 ? part0 : (a == 1 ? part1 : part2);
        dst[((long)l * npatch + p) * ld + r] = make_double2(sr2, si2);
    }
}

// hval[s][h] = val_t[s][hrows[h]], hcol[s][h] = col_t[s][hrows[h]].
template <typename Real>
__global__ void halo_gather_kernel(const Real *__restrict__ val_t, const int *__restrict__ col_t,
                                   const int *__restrict__ hrows, int n, int k, long htot,
                                   Real *__restrict__ hval, int *__restrict__ hcol,
                                   int with_cols) {
    const long tot = htot * k;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += stride) {
        const long s = e / htot, h = e % htot;
        const long src = s * n + hrows[h];
        hval[2 * e] = val_t[2 * src];
        hval[2 * e + 1] = val_t[2 * src + 1];
        if (with_cols) hcol[e] = col_t[src];
    }
}

using TileFn = const void *;

template <typename Real, int VPL, int LPR, bool DBL>
TileFn tile_fn(int k) {
    switch (k) {
        case 4: return reinterpret_cast<TileFn>(&tile2_kernel<Real, VPL, LPR, 4, DBL>);
        case 5: return reinterpret_cast<TileFn>(&tile2_kernel<Real, VPL, LPR, 5, DBL>);
        default: return reinterpret_cast<TileFn>(&tile2_kernel<Real, VPL, LPR, 0, DBL>);
    }
}

template <typename Real, int VPL, int LPR>
TileFn tile_fn2(int k, bool dbl) {
    return dbl ? tile_fn<Real, VPL, LPR, true>(k) : tile_fn<Real, VPL, LPR, false>(k);
}

struct TileEntry {
    TileFn fn;
    int vpl, lpr;
};

TileEntry pick_tile(int vtype, int rb, int k, bool dbl, bool wide) {
    if (vtype == JKPM_C64) {
        if (wide) {
            switch (rb) {
                case 16:  // the doubling form's four dot sets fit 16 waves/CU at VPL 2
                    return dbl ? TileEntry{tile_fn2<float, 2, 8>(k, dbl), 2, 8}
                               : TileEntry{tile_fn2<float, 4, 4>(k, dbl), 4, 4};
                case 8: return {tile_fn2<float, 2, 4>(k, dbl), 2, 4};
                case 4: return {tile_fn2<float, 2, 2>(k, dbl), 2, 2};
                case 2: return {tile_fn2<float, 2, 1>(k, dbl), 2, 1};
                default: break;
            }
        }
        // 8-byte lanes (one complex64 per lane): an odd leading dimension
        // leaves a row's vectors only 8-byte aligned.
        switch (rb) {
            case 16: return {tile_fn2<float, 1, 16>(k, dbl), 1, 16};
            case 8: return {tile_fn2<float, 1, 8>(k, dbl), 1, 8};
            case 4: return {tile_fn2<float, 1, 4>(k, dbl), 1, 4};
            case 2: return {tile_fn2<float, 1, 2>(k, dbl), 1, 2};
            case 1: return {tile_fn2<float, 1, 1>(k, dbl), 1, 1};
            default: return {nullptr, 0, 0};
        }
    }
    switch (rb) {
        case 16: return {tile_fn2<double, 1, 16>(k, dbl), 1, 16};
        case 8: return {tile_fn2<double, 1, 8>(k, dbl), 1, 8};
        case 4: return {tile_fn2<double, 1, 4>(k, dbl), 1, 4};
        case 2: return {tile_fn2<double, 1, 2>(k, dbl), 1, 2};
        case 1: return {tile_fn2<double, 1, 1>(k, dbl), 1, 1};
        default: return {nullptr, 0, 0};
    }
}

int grid_for(long count, const DeviceInfo &dev) {
    const long want = (count + 255) / 256;
    return (int)std::max(1L, std::min(want, (long)dev.cu_count * 16));
}

}  // namespace

int tile_cap_rows(int vtype, int rb) {
    return kLdsBytes / (rb * (int)cplx_bytes(vtype));
}

int tile_max_group() { return 16; }

int launch_tile_permute(const Ell &H, const int *perm, const int *inv, void *val_t, int *col_t,
                        bool with_cols, hipStream_t s) {
    const long tot = (long)H.n * H.k;
    if (tot == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int grid = grid_for(tot, dev);
    if (H.vtype == JKPM_C64)
        hipLaunchKernelGGL(tile_permute_kernel<float>, dim3(grid), dim3(256), 0, s,
                           (const float *)H.val, H.col, perm, inv, H.n, H.k, (float *)val_t, col_t,
                           with_cols ? 1 : 0);
    else
        hipLaunchKernelGGL(tile_permute_kernel<double>, dim3(grid), dim3(256), 0, s,
                           (const double *)H.val, H.col, perm, inv, H.n, H.k, (double *)val_t,
                           col_t, with_cols ? 1 : 0);
    JKPM_CHECK_LAUNCH("tile_permute_kernel");
    return JKPM_OK;
}

int launch_vec_permute(int vtype, const void *src, const int *perm, int n, int R, void *dst,
                       hipStream_t s) {
    const long tot = (long)n * R;
    if (tot == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int grid = grid_for(tot, dev);
    if (vtype == JKPM_C64)
        hipLaunchKernelGGL(vec_permute_kernel<float>, dim3(grid), dim3(256), 0, s,
                           (const float *)src, perm, n, R, (float *)dst);
    else
        hipLaunchKernelGGL(vec_permute_kernel<double>, dim3(grid), dim3(256), 0, s,
                           (const double *)src, perm, n, R, (double *)dst);
    JKPM_CHECK_LAUNCH("vec_permute_kernel");
    return JKPM_OK;
}

int launch_halo_gather(const jkpm_tiles &T, bool with_cols, hipStream_t s) {
    const long tot = T.halo_rows * T.k;
    if (tot == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int grid = grid_for(tot, dev);
    if (T.vtype == JKPM_C64)
        hipLaunchKernelGGL(halo_gather_kernel<float>, dim3(grid), dim3(256), 0, s,
                           (const float *)T.val, T.col, T.hrows, T.n, T.k, (long)T.halo_rows,
                           (float *)T.hval, T.hcol, with_cols ? 1 : 0);
    else
        hipLaunchKernelGGL(halo_gather_kernel<double>, dim3(grid), dim3(256), 0, s,
                           (const double *)T.val, T.col, T.hrows, T.n, T.k, (long)T.halo_rows,
                           (double *)T.hval, T.hcol, with_cols ? 1 : 0);
    JKPM_CHECK_LAUNCH("halo_gather_kernel");
    return JKPM_OK;
}

int launch_tile2(const TileArgs &a, hipStream_t s) {
    const jkpm_tiles &T = *a.T;
    // 16-byte lanes need every row's group start 16-byte aligned: an even
    // leading dimension and an even first vector (the rule of launch_step).
    const bool wide = (a.ld % 2 == 0) && (a.r0 % 2 == 0);
    TileEntry e = pick_tile(T.vtype, a.rb, T.k, a.doubling, wide);
    if (!e.fn) return arg_error("unsupported vector group for the tiled kernel");
    if (T.max_ext > tile_cap_rows(T.vtype, a.rb))
        return arg_error("the tile plan's patches exceed the LDS tile for this vector group");
    const size_t cb = cplx_bytes(T.vtype);
    const char *xprev = a.xprev ? (const char *)a.xprev + a.r0 * cb : nullptr;
    const char *xcur = (const char *)a.xcur + a.r0 * cb;
    char *out1 = (char *)a.out1 + a.r0 * cb;
    char *out2 = (char *)a.out2 + a.r0 * cb;
    const char *vl = a.vl ? (const char *)a.vl + a.r0 * cb : nullptr;
    double2 *p0 = a.part0 ? a.part0 + a.r0 : nullptr;
    double2 *p1 = a.part1 + a.r0, *p2 = a.part2 + a.r0;
    const void *val = T.val;
    const int *col = T.col;
    const short *lcol = T.lcol;
    const int *pst = T.pstart, *hof = T.hoff, *hro = T.hrows;
    const void *hv = T.hval;
    const int *hc = T.hcol;
    int htot = (int)T.halo_rows;
    int n = T.n, k = T.k, np = T.npatch, ld = a.ld, first = a.first ? 1 : 0;
    const void *px = xprev, *pc = xcur, *pv = vl;
    void *po1 = out1, *po2 = out2;
    hipError_t err;
    if (T.vtype == JKPM_C64) {
        float a1 = (float)a.a1, b1 = (float)a.b1;
        void *args[] = {&val, &col, &lcol, &pst, &hof, &hro, &hv, &hc, &htot, &n,  &k,  &np,
                        &px,  &pc,  &po1,  &po2, &pv,  &ld,  &p0, &p1, &p2,   &a1, &b1, &first};
        err = hipLaunchKernel(e.fn, dim3(np), dim3(kTB), args, 0, s);
    } else {
        double a1 = a.a1, b1 = a.b1;
        void *args[] = {&val, &col, &lcol, &pst, &hof, &hro, &hv, &hc, &htot, &n,  &k,  &np,
                        &px,  &pc,  &po1,  &po2, &pv,  &ld,  &p0, &p1, &p2,   &a1, &b1, &first};
        err = hipLaunchKernel(e.fn, dim3(np), dim3(kTB), args, 0, s);
    }
    if (err != hipSuccess) return hip_fail(err, "tile2_kernel launch", __FILE__, __LINE__);
    JKPM_CHECK_LAUNCH("tile2_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
