// Fused ELL SpMV + three-term Chebyshev update + FP64 moment partials for
// gfx950, plus the deterministic moment reduction, ELL packing and aAxby.
//
// Reference algorithm: jclsquant/recurrence_functions.py:9-111 (rec_A_tab,
// rec_A_tab2v, rec_A_tab2v_2) over aAxby/vdot (jclsquant/cython_modules/
// blas_funs.pyx:64-91); GPU twin jclsquant/cuda_cython/cuda/recursions.cu:17-98
// with aAxby_device (blas.cu:29-43) and thrust::inner_product.
//
// Layout and mapping
//   values/cols   slot-major [K][N]: at slot s, the lanes of a wave read
//                 consecutive rows -> coalesced matrix stream.
//   vectors       row-interleaved [N][ld]: a gathered row of RB vectors is one
//                 contiguous run of RB complex numbers (RB*8 B for c64).
//   lanes         LPR lanes per row, VPL vectors per lane (16-byte loads);
//                 a 64-wide wave covers 64/LPR consecutive rows.
//   work-groups   256 threads; each owns one contiguous chunk of rows.  Chunks
//                 are assigned XCD-major: work-group b runs on XCD b % 8, so it
//                 gets chunk (b % 8) * (nchunks / 8) + b / 8 and every XCD
//                 sweeps one contiguous eighth of the lattice -- the gather
//                 window of neighbouring lattice rows stays in that XCD's L2.
//   moments       each lane accumulates conj(v_l) * T_{n+1} in FP64; a
//                 butterfly over the lanes of equal vector index, then a fixed
//                 wave order in LDS, gives one partial per (chunk, vector).
//                 reduce_kernel sums the partials in fixed chunk order, so the
//                 result is bitwise reproducible run to run.
#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;

template <typename Real, int N>
using vec_t = Real __attribute__((ext_vector_type(N)));

// Loads of data read once per step (T_{n-1}, v_l): non-temporal when NT, so
// they do not displace the gather window of T_n from the caches.
template <bool NT, typename V>
__device__ __forceinline__ V load_stream(const V *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// UNR: rows per lane per loop iteration (all their loads are issued before
// any is used, for more bytes in flight per wave).
template <typename Real, int VPL, int LPR, int NL, int KT, int UNR, bool NT>
__global__ __launch_bounds__(kBlock) void cheb_step_kernel(
    const Real *__restrict__ val, const int *__restrict__ col, int n, int k_rt,
    const Real *__restrict__ xcur, const Real *xprev, Real *xout,
    const Real *__restrict__ vl0, const Real *__restrict__ vl1, int ld,
    double2 *__restrict__ part, double2 *__restrict__ part0, int first,
    int rows_per_chunk, int nchunks, const long *__restrict__ chunk_rows) {
    using V = vec_t<Real, 2 * VPL>;
    using C2 = vec_t<Real, 2>;
    constexpr int RPW = 64 / LPR;
    constexpr int RB = VPL * LPR;
    const int K = KT > 0 ? KT : k_rt;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lv = lane % LPR;
    const int lr = lane / LPR;

    const int b = blockIdx.x;
    const int chunk = (nchunks % kXcd == 0) ? (b % kXcd) * (nchunks / kXcd) + b / kXcd : b;
    // Uniform chunks, or an explicit row table (chunks that never straddle
    // two realizations of a block-diagonal batch).
    const long row_begin = chunk_rows ? chunk_rows[chunk] : (long)chunk * rows_per_chunk;
    const long row_end =
        chunk_rows ? chunk_rows[chunk + 1] : min((long)n, row_begin + rows_per_chunk);

    double ar[NL][VPL], ai[NL][VPL];
    double a0r[NL][VPL], a0i[NL][VPL];
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
        for (int v = 0; v < VPL; ++v) ar[l][v] = ai[l][v] = a0r[l][v] = a0i[l][v] = 0.0;

    const int voff = lv * VPL * 2;
    const Real *vls[2] = {vl0, vl1};

    constexpr int STRIDE = kWaves * RPW;  // rows between a lane's consecutive rows
    for (long base = row_begin + wave * RPW + lr; base < row_end; base += (long)STRIDE * UNR) {
        // Rows past the chunk end are clamped to `base` (a valid row) for the
        // loads and their results are dropped.
        long rows[UNR];
        bool ok[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const long rw = base + (long)u * STRIDE;
            ok[u] = rw < row_end;
            rows[u] = ok[u] ? rw : base;
        }
        Real sr[UNR][VPL], si[UNR][VPL];
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
            for (int v = 0; v < VPL; ++v) sr[u][v] = si[u][v] = Real(0);

        if constexpr (KT > 0) {
            int j[UNR][KT];
            C2 h[UNR][KT];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const long e = (long)s * n + rows[u];
                    j[u][s] = col[e];
                    h[u][s] = *reinterpret_cast<const C2 *>(val + 2 * e);
                }
            V x[UNR][KT];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int s = 0; s < KT; ++s)
                    x[u][s] = *reinterpret_cast<const V *>(xcur + (long)j[u][s] * ld * 2 + voff);
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int s = 0; s < KT; ++s)
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        sr[u][v] += h[u][s].x * x[u][s][2 * v] - h[u][s].y * x[u][s][2 * v + 1];
                        si[u][v] += h[u][s].x * x[u][s][2 * v + 1] + h[u][s].y * x[u][s][2 * v];
                    }
        } else {
#pragma unroll
            for (int u = 0; u < UNR; ++u)
                for (int s = 0; s < K; ++s) {
                    const long e = (long)s * n + rows[u];
                    const int j = col[e];
                    const C2 h = *reinterpret_cast<const C2 *>(val + 2 * e);
                    const V x = *reinterpret_cast<const V *>(xcur + (long)j * ld * 2 + voff);
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        sr[u][v] += h.x * x[2 * v] - h.y * x[2 * v + 1];
                        si[u][v] += h.x * x[2 * v + 1] + h.y * x[2 * v];
                    }
                }
        }

        // T_{n-1} and v_l of every row first, then the arithmetic.
        V p[UNR], q[UNR][NL], t0[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const long vo = rows[u] * ld * 2 + voff;
            if (first)
                t0[u] = *reinterpret_cast<const V *>(xcur + vo);
            else
                p[u] = load_stream<NT>(reinterpret_cast<const V *>(xprev + vo));
#pragma unroll
            for (int l = 0; l < NL; ++l) q[u][l] = load_stream<NT>(reinterpret_cast<const V *>(vls[l] + vo));
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (!ok[u]) continue;
            const long vo = rows[u] * ld * 2 + voff;
            V y;
            if (first) {
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    y[2 * v] = sr[u][v];
                    y[2 * v + 1] = si[u][v];
                }
#pragma unroll
                for (int l = 0; l < NL; ++l)
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        const double qr = q[u][l][2 * v], qi = q[u][l][2 * v + 1];
                        const double tr = t0[u][2 * v], ti = t0[u][2 * v + 1];
                        a0r[l][v] += qr * tr + qi * ti;
                        a0i[l][v] += qr * ti - qi * tr;
                    }
            } else {
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    y[2 * v] = Real(2) * sr[u][v] - p[u][2 * v];
                    y[2 * v + 1] = Real(2) * si[u][v] - p[u][2 * v + 1];
                }
            }
            *reinterpret_cast<V *>(xout + vo) = y;
#pragma unroll
            for (int l = 0; l < NL; ++l)
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    const double qr = q[u][l][2 * v], qi = q[u][l][2 * v + 1];
                    const double yr = y[2 * v], yi = y[2 * v + 1];
                    ar[l][v] += qr * yr + qi * yi;
                    ai[l][v] += qr * yi - qi * yr;
                }
        }
    }

    // Butterfly over the lanes that hold the same vector index.
#pragma unroll
    for (int o = LPR; o < 64; o <<= 1) {
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                ar[l][v] += __shfl_xor(ar[l][v], o);
                ai[l][v] += __shfl_xor(ai[l][v], o);
                if (first) {
                    a0r[l][v] += __shfl_xor(a0r[l][v], o);
                    a0i[l][v] += __shfl_xor(a0i[l][v], o);
                }
            }
    }

    __shared__ double2 sh[2][kWaves][NL][RB];
    if (lr == 0) {
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                sh[0][wave][l][lv * VPL + v] = make_double2(ar[l][v], ai[l][v]);
                sh[1][wave][l][lv * VPL + v] = make_double2(a0r[l][v], a0i[l][v]);
            }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NL * RB; t += kBlock) {
        const int l = t / RB, r = t % RB;
        double sr2 = 0.0, si2 = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            sr2 += sh[0][w][l][r].x;
            si2 += sh[0][w][l][r].y;
        }
        part[((long)l * nchunks + chunk) * ld + r] = make_double2(sr2, si2);
        if (first) {
            double s0r = 0.0, s0i = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                s0r += sh[1][w][l][r].x;
                s0i += sh[1][w][l][r].y;
            }
            part0[((long)l * nchunks + chunk) * ld + r] = make_double2(s0r, s0i);
        }
    }
}

// One work-group per (vector, left set x group, step): fixed-order sum of the
// chunk partials of that step over the group's chunks (all chunks when there
// are no groups).
__global__ __launch_bounds__(kBlock) void reduce_kernel(const double2 *__restrict__ ring,
                                                         int ring_slots, int step_first,
                                                         int n_left, int nchunks, int ld,
                                                         double2 *__restrict__ mu, int M,
                                                         const int *__restrict__ group_chunks,
                                                         int G) {
    const int r = blockIdx.x, l = blockIdx.y / G, g = blockIdx.y % G;
    const int step = step_first + blockIdx.z;
    const double2 *p =
        ring + ((long)(step % ring_slots) * n_left + l) * (long)nchunks * ld + r;
    const int c_begin = group_chunks ? group_chunks[g] : 0;
    const int c_end = group_chunks ? group_chunks[g + 1] : nchunks;
    double sr = 0.0, si = 0.0;
    for (int c = c_begin + threadIdx.x; c < c_end; c += kBlock) {
        const double2 q = p[(long)c * ld];
        sr += q.x;
        si += q.y;
    }
    __shared__ double shr[kBlock], shi[kBlock];
    shr[threadIdx.x] = sr;
    shi[threadIdx.x] = si;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            shr[threadIdx.x] += shr[threadIdx.x + o];
            shi[threadIdx.x] += shi[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        mu[(((long)l * G + g) * M + step) * ld + r] = make_double2(shr[0], shi[0]);
}

// Row-major complex128 (device) -> slot-major value type, times `scale`.
template <typename Real>
__global__ void ell_pack_kernel(const double *__restrict__ src, const int *__restrict__ idx,
                                double scale, int n, int k, Real *__restrict__ val,
                                int *__restrict__ col) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * k) return;
    const long i = e / k, s = e % k;
    const long o = s * n + i;
    val[2 * o] = (Real)(src[2 * e] * scale);
    val[2 * o + 1] = (Real)(src[2 * e + 1] * scale);
    if (idx) col[o] = idx[e];
}

// y <- a*A*x + b*y for R interleaved vectors (one thread per element).
template <typename Real>
__global__ void aaxby_kernel(const Real *__restrict__ val, const int *__restrict__ col, int n,
                             int k, double ar, double ai, double br, double bi,
                             const Real *__restrict__ x, Real *__restrict__ y, int R) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * R) return;
    const long row = t / R, r = t % R;
    // y *= b
    Real yr = y[2 * t], yi = y[2 * t + 1];
    Real cr = (Real)br * yr - (Real)bi * yi;
    Real ci = (Real)br * yi + (Real)bi * yr;
    for (int s = 0; s < k; ++s) {
        const long e = (long)s * n + row;
        const Real hr = val[2 * e], hi = val[2 * e + 1];
        // (a * h) * x
        const Real pr = (Real)ar * hr - (Real)ai * hi;
        const Real pi = (Real)ar * hi + (Real)ai * hr;
        const long xo = (long)col[e] * R + r;
        const Real xr = x[2 * xo], xi = x[2 * xo + 1];
        cr += pr * xr - pi * xi;
        ci += pr * xi + pi * xr;
    }
    y[2 * t] = cr;
    y[2 * t + 1] = ci;
}

__global__ void copy_kernel(const int4 *__restrict__ src, int4 *__restrict__ dst, long n16) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

// ---- dispatch table for the step kernel ----
struct StepEntry {
    const void *fn;
    int vpl, lpr, unr;
};

template <typename Real, int VPL, int LPR, int NL, int KT, int UNR, bool NT>
const void *step_ptr() {
    return reinterpret_cast<const void *>(&cheb_step_kernel<Real, VPL, LPR, NL, KT, UNR, NT>);
}

// Variants (UNR, NT) exist for the single-left-set kernels of K = 4 and 5
// (the lattice workloads); every other shape has the (1, plain) kernel.
template <typename Real, int VPL, int LPR, int KT>
const void *pick_variant(int variant) {
    switch (variant) {
        case 1: return step_ptr<Real, VPL, LPR, 1, KT, 2, false>();
        case 2: return step_ptr<Real, VPL, LPR, 1, KT, 1, true>();
        case 3: return step_ptr<Real, VPL, LPR, 1, KT, 2, true>();
        default: return step_ptr<Real, VPL, LPR, 1, KT, 1, false>();
    }
}

template <typename Real, int VPL, int LPR>
const void *pick_nl(int nl, int k, int variant) {
    if (nl == 2) {
        if (k == 4) return step_ptr<Real, VPL, LPR, 2, 4, 1, false>();
        if (k == 5) return step_ptr<Real, VPL, LPR, 2, 5, 1, false>();
        return step_ptr<Real, VPL, LPR, 2, 0, 1, false>();
    }
    if (k == 4) return pick_variant<Real, VPL, LPR, 4>(variant);
    if (k == 5) return pick_variant<Real, VPL, LPR, 5>(variant);
    return step_ptr<Real, VPL, LPR, 1, 0, 1, false>();
}

int variant_unr(int nl, int k, int variant) {
    return (nl == 1 && (k == 4 || k == 5) && (variant == 1 || variant == 3)) ? 2 : 1;
}

// Kernel for a group of rb vectors.  wide = 16-byte lanes for c64 (needs
// 16-byte aligned rows: ld and r0 even).
StepEntry pick_step(int vtype, int rb, bool wide, int nl, int k, int variant) {
    const int u = variant_unr(nl, k, variant);
    if (vtype == JKPM_C64) {
        if (wide) {
            switch (rb) {
                case 2: return {pick_nl<float, 2, 1>(nl, k, variant), 2, 1, u};
                case 4: return {pick_nl<float, 2, 2>(nl, k, variant), 2, 2, u};
                case 8: return {pick_nl<float, 2, 4>(nl, k, variant), 2, 4, u};
                case 16: return {pick_nl<float, 2, 8>(nl, k, variant), 2, 8, u};
                case 32: return {pick_nl<float, 2, 16>(nl, k, variant), 2, 16, u};
                default: break;
            }
        }
        switch (rb) {
            case 1: return {pick_nl<float, 1, 1>(nl, k, variant), 1, 1, u};
            case 2: return {pick_nl<float, 1, 2>(nl, k, variant), 1, 2, u};
            case 4: return {pick_nl<float, 1, 4>(nl, k, variant), 1, 4, u};
            case 8: return {pick_nl<float, 1, 8>(nl, k, variant), 1, 8, u};
            case 16: return {pick_nl<float, 1, 16>(nl, k, variant), 1, 16, u};
            case 32: return {pick_nl<float, 1, 32>(nl, k, variant), 1, 32, u};
            default: break;
        }
    } else {
        switch (rb) {
            case 1: return {pick_nl<double, 1, 1>(nl, k, variant), 1, 1, u};
            case 2: return {pick_nl<double, 1, 2>(nl, k, variant), 1, 2, u};
            case 4: return {pick_nl<double, 1, 4>(nl, k, variant), 1, 4, u};
            case 8: return {pick_nl<double, 1, 8>(nl, k, variant), 1, 8, u};
            case 16: return {pick_nl<double, 1, 16>(nl, k, variant), 1, 16, u};
            case 32: return {pick_nl<double, 1, 32>(nl, k, variant), 1, 32, u};
            default: break;
        }
    }
    return {nullptr, 0, 0, 1};
}

}  // namespace

int max_group(int) { return 32; }

int step_chunks(const Ell &H, int, const DeviceInfo &dev) {
    // chunk_mult work-groups per CU (tunable, 8 by default), never fewer rows
    // per chunk than one wave-row sweep.
    long target = (long)dev.cu_count * (H.chunk_mult > 0 ? H.chunk_mult : 8);
    long by_rows = (H.n + 63) / 64;
    long c = target < by_rows ? target : by_rows;
    if (c >= kXcd) c = c / kXcd * kXcd;
    if (c < 1) c = 1;
    return (int)c;
}

int launch_step(const StepArgs &a, hipStream_t s) {
    const Ell &H = *a.H;
    const int nl = a.vl1 ? 2 : 1;
    const size_t cb = cplx_bytes(H.vtype);
    const bool wide = (a.ld % 2 == 0) && (a.r0 % 2 == 0);
    StepEntry e = pick_step(H.vtype, a.rb, wide, nl, H.k, H.variant);
    if (!e.fn) return arg_error("unsupported vector group size " + std::to_string(a.rb));
    const int rpb = kWaves * (64 / e.lpr) * e.unr;
    long rpc = ((long)H.n + a.nchunks - 1) / a.nchunks;
    rpc = (rpc + rpb - 1) / rpb * rpb;
    const size_t off = (size_t)a.r0 * cb;
    // Arguments in the kernel's parameter order; every pointer is passed as
    // an untyped pointer of the same size as the kernel's typed one.
    const void *p_val = H.val;
    const int *p_col = H.col;
    int p_n = H.n, p_k = H.k;
    const void *p_xcur = (const char *)a.xcur + off;
    const void *p_xprev = a.xprev ? (const char *)a.xprev + off : nullptr;
    void *p_xout = (char *)a.xout + off;
    const void *p_vl0 = (const char *)a.vl0 + off;
    const void *p_vl1 = a.vl1 ? (const char *)a.vl1 + off : nullptr;
    int p_ld = a.ld;
    double2 *p_part = a.part + a.r0;
    double2 *p_part0 = a.part0 ? a.part0 + a.r0 : nullptr;
    int p_first = a.first ? 1 : 0, p_rpc = (int)rpc, p_nch = a.nchunks;
    const long *p_rows = a.chunk_rows;
    void *args[] = {&p_val,  &p_col,  &p_n,    &p_k,     &p_xcur,  &p_xprev, &p_xout, &p_vl0,
                    &p_vl1,  &p_ld,   &p_part, &p_part0, &p_first, &p_rpc,   &p_nch,  &p_rows};
    JKPM_HIP(hipLaunchKernel(e.fn, dim3(a.nchunks), dim3(kBlock), args, 0, s));
    JKPM_CHECK_LAUNCH("cheb_step_kernel");
    return JKPM_OK;
}

int launch_reduce(const double2 *ring, int ring_slots, int step_first, int nsteps, int n_left,
                  int nchunks, int ld, double2 *mu, int M, const int *group_chunks, int G,
                  hipStream_t s) {
    if (nsteps <= 0) return JKPM_OK;
    if (G < 1) G = 1;
    hipLaunchKernelGGL(reduce_kernel, dim3(ld, n_left * G, nsteps), dim3(kBlock), 0, s, ring,
                       ring_slots, step_first, n_left, nchunks, ld, mu, M, group_chunks, G);
    JKPM_CHECK_LAUNCH("reduce_kernel");
    return JKPM_OK;
}

int launch_ell_pack(const double *src, double scale, Ell *H, const int *idx, hipStream_t s) {
    const long tot = (long)H->n * H->k;
    const int grid = (int)((tot + kBlock - 1) / kBlock);
    if (tot == 0) return JKPM_OK;
    if (H->vtype == JKPM_C64)
        hipLaunchKernelGGL(ell_pack_kernel<float>, dim3(grid), dim3(kBlock), 0, s, src, idx,
                           scale, H->n, H->k, (float *)H->val, H->col);
    else
        hipLaunchKernelGGL(ell_pack_kernel<double>, dim3(grid), dim3(kBlock), 0, s, src, idx,
                           scale, H->n, H->k, (double *)H->val, H->col);
    JKPM_CHECK_LAUNCH("ell_pack_kernel");
    return JKPM_OK;
}

int launch_aaxby(const Ell &H, double2 a, double2 b, const void *x, void *y, int R,
                 hipStream_t s) {
    const long tot = (long)H.n * R;
    if (tot == 0) return JKPM_OK;
    const int grid = (int)((tot + kBlock - 1) / kBlock);
    if (H.vtype == JKPM_C64)
        hipLaunchKernelGGL(aaxby_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                           (const float *)H.val, H.col, H.n, H.k, a.x, a.y, b.x, b.y,
                           (const float *)x, (float *)y, R);
    else
        hipLaunchKernelGGL(aaxby_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                           (const double *)H.val, H.col, H.n, H.k, a.x, a.y, b.x, b.y,
                           (const double *)x, (double *)y, R);
    JKPM_CHECK_LAUNCH("aaxby_kernel");
    return JKPM_OK;
}

int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const long n16 = (long)(bytes / 16);
    if (n16 == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    long grid = (n16 + kBlock - 1) / kBlock;
    const long cap = (long)dev.cu_count * 16;
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)grid), dim3(kBlock), 0, s,
                       (const int4 *)src, (int4 *)dst, n16);
    JKPM_CHECK_LAUNCH("copy_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
