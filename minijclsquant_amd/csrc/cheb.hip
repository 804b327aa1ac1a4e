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
//   left vectors  complex64 random-phase vectors are read as 32-bit phases
//                 (4 bytes per element) and rebuilt in registers
//                 (phase_pack_kernel writes them once per recursion).
//   moments       each lane accumulates conj(v_l) * T_{n+1} in FP64; a
//                 butterfly over the lanes of equal vector index, then a fixed
//                 wave order in LDS, gives one partial per (chunk, vector).
//                 reduce_kernel sums the partials in fixed chunk order, so the
//                 result is bitwise reproducible run to run.
//
// The step kernel template lives in cheb_step.h and is instantiated per
// precision in cheb_c64.hip / cheb_c128.hip; this file holds its launcher and
// the other small kernels.
#include <algorithm>

#include "cheb_step.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;

// One work-group per (vector, left set x group, step): fixed-order sum of the
// chunk partials of that step over the group's chunks (all chunks when there
// are no groups).  Used for batches of more than kBlock vectors.
__global__ __launch_bounds__(kBlock) void reduce_kernel(const double2 *__restrict__ ring,
                                                         int ring_slots, int step_first,
                                                         int n_left, int nchunks, int ld,
                                                         double2 *__restrict__ mu, int M,
                                                         const int *__restrict__ group_chunks,
                                                         int G, int mu_ld) {
    const int r = blockIdx.x, l = blockIdx.y / G, g = blockIdx.y % G;
    const int step = step_first + blockIdx.z;
    const double2 *p =
        ring + ((long)(step % ring_slots) * n_left + l) * (long)nchunks * ld + r;
    const int c_begin = group_chunks ? group_chunks[g] : 0;
    const int c_end = group_chunks ? group_chunks[g + 1] : nchunks;
    double sr = 0.0, si = 0.0;
    // unrolled: several partials in flight per thread, added in the same order
#pragma unroll 8
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
        mu[(((long)l * G + g) * M + step) * mu_ld + r] = make_double2(shr[0], shi[0]);
}

// The same sums with one work-group per (left set x group, step) for all
// ld <= kBlock vectors at once: thread t adds chunks t/ld, t/ld + kBlock/ld,
// ... of vector t % ld, so each pass reads kBlock/ld whole chunk rows of the
// [chunk][ld] partials contiguously; the per-vector totals are then summed
// over the kBlock/ld thread rows in a fixed order.
__global__ __launch_bounds__(kBlock) void reduce_rows_kernel(const double2 *__restrict__ ring,
                                                              int ring_slots, int step_first,
                                                              int n_left, int nchunks, int ld,
                                                              double2 *__restrict__ mu, int M,
                                                              const int *__restrict__ group_chunks,
                                                              int G, int mu_ld) {
    const int l = blockIdx.x / G, g = blockIdx.x % G;
    const int step = step_first + blockIdx.y;
    const double2 *p = ring + ((long)(step % ring_slots) * n_left + l) * (long)nchunks * ld;
    const int c_begin = group_chunks ? group_chunks[g] : 0;
    const int c_end = group_chunks ? group_chunks[g + 1] : nchunks;
    const int rows = kBlock / ld;  // chunk rows per pass
    const int t = threadIdx.x;
    const int r = t % ld, c0 = t / ld;
    double sr = 0.0, si = 0.0;
    if (c0 < rows) {
        // unrolled: several partials in flight per thread, added in the same order
#pragma unroll 8
        for (int c = c_begin + c0; c < c_end; c += rows) {
            const double2 q = p[(long)c * ld + r];
            sr += q.x;
            si += q.y;
        }
    }
    __shared__ double shr[kBlock], shi[kBlock];
    shr[t] = sr;
    shi[t] = si;
    __syncthreads();
    if (t < ld) {
        double ar = 0.0, ai = 0.0;
        for (int c = 0; c < rows; ++c) {
            ar += shr[c * ld + t];
            ai += shi[c * ld + t];
        }
        mu[(((long)l * G + g) * M + step) * mu_ld + t] = make_double2(ar, ai);
    }
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

// Moments from the doubling form's dots (see launch_doubling_finish).
__global__ void doubling_finish_kernel(const double2 *__restrict__ raw, int G, int S1, int R,
                                       int M, double2 *__restrict__ mu, int mu_ld) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)G * M * R) return;
    const int r = (int)(t % R);
    const int m = (int)((t / R) % M);
    const int g = (int)(t / ((long)R * M));
    auto at = [&](int l, int slot) { return raw[(((long)l * G + g) * S1 + slot) * R + r]; };
    double2 v;
    if (m == 0) {
        v = at(0, 0);
    } else if (m == 1) {
        v = at(1, 1);
    } else if ((m & 1) == 0) {
        const double2 a = at(0, m / 2), a0 = at(0, 0);
        v = make_double2(2.0 * a.x - a0.x, 2.0 * a.y - a0.y);
    } else {
        const double2 b = at(1, (m - 1) / 2 + 1), b0 = at(1, 1);
        v = make_double2(2.0 * b.x - b0.x, 2.0 * b.y - b0.y);
    }
    mu[((long)g * M + m) * mu_ld + r] = v;
}

// Columns c0 .. c0+cols-1 of an [n][ld] array of `u` 8-byte words per
// element, made contiguous ([n][cols]).
__global__ __launch_bounds__(kBlock) void take_columns_kernel(const unsigned long long *__restrict__ src,
                                                              long n, int ld, int c0, int cols,
                                                              int u,
                                                              unsigned long long *__restrict__ dst) {
    const long w = (long)cols * u;  // words per destination row
    const long tot = n * w;
    for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < tot; t += (long)gridDim.x * kBlock) {
        const long i = t / w, j = t - i * w;
        dst[t] = src[i * ld * u + (long)c0 * u + j];
    }
}

// Plain copy with kCopyUnroll independent 16-byte loads in flight per lane
// before any store (a lane waiting on one load at a time cannot cover the
// HBM latency).  Each work-group copies one contiguous tile per iteration.
constexpr int kCopyUnroll = 4;

// Non-temporal loads and stores: the fastest streaming form measured on
// MI355X (a copy that goes through the caches normally runs ~5 % slower).
using i32x4 = int __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void copy_kernel(const i32x4 *__restrict__ src,
                                                       i32x4 *__restrict__ dst, long n16) {
    const long tile = (long)kBlock * kCopyUnroll;
    const long ntiles = (n16 + tile - 1) / tile;
    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const long base = t * tile + threadIdx.x;
        i32x4 v[kCopyUnroll];
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const long i = base + (long)u * kBlock;
            if (i < n16) v[u] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const long i = base + (long)u * kBlock;
            if (i < n16) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

// Phases of complex64 numbers for the step kernel's left-vector stream:
// u = round(atan2(y, x) / (2 pi) * 2^32) mod 2^32, then a check that
// step::phase_cs(u) gives (x, y) back within 2^-21 per component.  An
// element off the unit circle, or a NaN, fails the check.
__global__ __launch_bounds__(kBlock) void phase_pack_kernel(const float2 *__restrict__ v, long n,
                                                             unsigned *__restrict__ ph,
                                                             int *__restrict__ bad) {
    constexpr double kTurns = 4294967296.0 / 6.283185307179586476925;  // 2^32 / (2 pi)
    constexpr float kTol = 4.76837158203125e-07f;                       // 2^-21
    bool fail = false;
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        const float2 z = v[i];
        const double t = rint(atan2((double)z.y, (double)z.x) * kTurns);  // [-2^31, 2^31]
        const unsigned u = t == t ? (unsigned)(long long)t : 0u;
        float c, sn;
        step::phase_cs(u, c, sn);
        fail |= !(fabsf(c - z.x) <= kTol && fabsf(sn - z.y) <= kTol);
        ph[i] = u;
    }
    if (__any(fail) && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}

__global__ void clear_flag_kernel(int *p) { *p = 0; }

}  // namespace

int max_group(int) { return 32; }

int step_chunks(const Ell &H, int, const DeviceInfo &dev, bool doubling) {
    // chunk_mult work-groups per CU (tunable, 8 by default), never fewer rows
    // per chunk than one wave-row sweep.
    const int mult = doubling ? H.dchunk_mult : H.chunk_mult;
    long target = (long)dev.cu_count * (mult > 0 ? mult : 8);
    long by_rows = (H.n + 63) / 64;
    long c = target < by_rows ? target : by_rows;
    if (c >= kXcd) c = c / kXcd * kXcd;
    if (c < 1) c = 1;
    return (int)c;
}

int launch_step(const StepArgs &a, hipStream_t s) {
    const Ell &H = *a.H;
    const int nl = (a.vl1 || a.doubling) ? 2 : 1;
    const size_t cb = cplx_bytes(H.vtype);
    const bool wide = (a.ld % 2 == 0) && (a.r0 % 2 == 0);
    // a real matrix: the kernels that read its packed real values, where the
    // shape has them
    const bool rv = H.real_values && H.rval;
    const step::StepEntry e =
        H.vtype == JKPM_C64
            ? step::pick_step_c64(a.rb, wide, nl, H.k, a.doubling ? H.dvariant : H.variant,
                                  a.doubling, a.first, rv)
            : step::pick_step_c128(a.rb, nl, H.k, a.doubling ? H.dvariant : H.variant, a.doubling,
                                   a.first, rv);
    if (!e.fn) return arg_error("unsupported vector group size " + std::to_string(a.rb));
    const int rpb = kWaves * (64 / e.lpr) * e.unr;
    long rpc = ((long)H.n + a.nchunks - 1) / a.nchunks;
    rpc = (rpc + rpb - 1) / rpb * rpb;
    const size_t off = (size_t)a.r0 * cb;
    // Arguments in the kernel's parameter order; every pointer is passed as
    // an untyped pointer of the same size as the kernel's typed one.
    const void *p_val = e.rv ? H.rval : H.val;
    const int *p_col = H.col;
    int p_n = H.n, p_k = H.k;
    const void *p_xcur = (const char *)a.xcur + off;
    const void *p_xprev = a.xprev ? (const char *)a.xprev + off : nullptr;
    void *p_xout = (char *)a.xout + off;
    // (the doubling form reads no left vectors; its pointers stay unused)
    const void *p_vl0 = a.vl0 ? (const char *)a.vl0 + off : nullptr;
    const void *p_vl1 = a.vl1 ? (const char *)a.vl1 + off : nullptr;
    const unsigned *p_vph = a.vph ? a.vph + a.r0 : nullptr;
    const int *p_phb = a.ph_bad;
    int p_ld = a.ld;
    double2 *p_part = a.part + a.r0;
    double2 *p_part0 = a.part0 ? a.part0 + a.r0 : nullptr;
    int p_first = a.first ? 1 : 0, p_rpc = (int)rpc, p_nch = a.nchunks;
    const long *p_rows = a.chunk_rows;
    int p_rev = a.reverse ? 1 : 0;
    void *args[] = {&p_val,   &p_col,  &p_n,   &p_k,    &p_xcur,  &p_xprev,
                    &p_xout,  &p_vl0,  &p_vl1, &p_vph,  &p_phb,   &p_ld,
                    &p_part,  &p_part0, &p_first, &p_rpc, &p_nch, &p_rows,
                    &p_rev};
    JKPM_HIP(hipLaunchKernel(e.fn, dim3(a.nchunks), dim3(kBlock), args, 0, s));
    JKPM_CHECK_LAUNCH("cheb_step_kernel");
    return JKPM_OK;
}

int launch_reduce(const double2 *ring, int ring_slots, int step_first, int nsteps, int n_left,
                  int nchunks, int ld, double2 *mu, int M, const int *group_chunks, int G,
                  hipStream_t s, int mu_ld) {
    if (nsteps <= 0) return JKPM_OK;
    if (G < 1) G = 1;
    if (mu_ld <= 0) mu_ld = ld;
    if (ld <= kBlock) {
        hipLaunchKernelGGL(reduce_rows_kernel, dim3(n_left * G, nsteps), dim3(kBlock), 0, s, ring,
                           ring_slots, step_first, n_left, nchunks, ld, mu, M, group_chunks, G,
                           mu_ld);
        JKPM_CHECK_LAUNCH("reduce_rows_kernel");
        return JKPM_OK;
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(ld, n_left * G, nsteps), dim3(kBlock), 0, s, ring,
                       ring_slots, step_first, n_left, nchunks, ld, mu, M, group_chunks, G, mu_ld);
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

int launch_doubling_finish(const double2 *raw, int G, int S1, int R, int M, double2 *mu,
                           hipStream_t s, int mu_ld) {
    const long tot = (long)G * M * R;
    if (tot == 0) return JKPM_OK;
    if (mu_ld <= 0) mu_ld = R;
    hipLaunchKernelGGL(doubling_finish_kernel, dim3((unsigned)((tot + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, raw, G, S1, R, M, mu, mu_ld);
    JKPM_CHECK_LAUNCH("doubling_finish_kernel");
    return JKPM_OK;
}

int launch_take_columns(const void *src, long n, int ld, int c0, int cols, int elem_bytes,
                        void *dst, hipStream_t s) {
    const int u = elem_bytes / 8;
    const long tot = n * cols * u;
    if (tot == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const long want = (tot + kBlock - 1) / kBlock;
    const int grid = (int)std::min(want, (long)dev.cu_count * 16);
    hipLaunchKernelGGL(take_columns_kernel, dim3(grid), dim3(kBlock), 0, s,
                       (const unsigned long long *)src, n, ld, c0, cols, u,
                       (unsigned long long *)dst);
    JKPM_CHECK_LAUNCH("take_columns_kernel");
    return JKPM_OK;
}

int launch_phase_pack(const void *v, long n, unsigned *ph, int *bad, hipStream_t s) {
    hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, s, bad);
    JKPM_CHECK_LAUNCH("clear_flag_kernel");
    if (n <= 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const long want = (n + kBlock - 1) / kBlock;
    const int grid = (int)std::min(want, (long)dev.cu_count * 8);
    hipLaunchKernelGGL(phase_pack_kernel, dim3(grid), dim3(kBlock), 0, s, (const float2 *)v, n,
                       ph, bad);
    JKPM_CHECK_LAUNCH("phase_pack_kernel");
    return JKPM_OK;
}

int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const long n16 = (long)(bytes / 16);
    if (n16 == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    long grid = (n16 + (long)kBlock * kCopyUnroll - 1) / ((long)kBlock * kCopyUnroll);
    const long cap = (long)dev.cu_count * 32;
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)grid), dim3(kBlock), 0, s,
                       (const i32x4 *)src, (i32x4 *)dst, n16);
    JKPM_CHECK_LAUNCH("copy_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
