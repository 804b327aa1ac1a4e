// Input and output checks: the Hermiticity of a handle's values, non-finite
// entries of a result array, and the allocation-failure test hook.
//
// The doubling form of the moments, the Kubo double moments and the adjoint
// recursion all rely on A^H = A (they use <T_k|T_k> for mu_2k, <r|v T_m v T_n|r>
// as L^H Rk, and A in place of A^H in the reverse recursion).  The check runs
// once per value upload over the slot pattern: for every slot (i, s) with
// column j it compares a_ij (all slots of row i with column j summed) with
// conj(a_ji) (all slots of row j with column i summed), so the padding slots
// of the ELL layout, which repeat a column with value 0, do not count twice.
// Cost O(N K^2) reads, one launch and one synchronisation per value change.
#include "common.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace jkpm {
namespace {

constexpr int kBlock = 256;

// max that keeps a NaN from either side (fmax would drop it), so a matrix
// with a NaN value reports a NaN asymmetry and is refused
__device__ __forceinline__ double max_nan(double a, double b) {
    return (b > a || b != b) ? b : a;
}

// out[0] = bits of max |a_ij - conj(a_ji)|, out[1] = bits of max |a_ij|.
// Non-negative doubles order like their bit patterns as unsigned integers,
// and every NaN pattern orders above +inf, so the maxima keep a NaN.
template <typename Real>
__global__ void __launch_bounds__(kBlock) herm_kernel(const Real *__restrict__ val,
                                                      const int *__restrict__ col, int n, int k,
                                                      unsigned long long *__restrict__ out) {
    double err = 0.0, mag = 0.0;
    const long tot = (long)n * k;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % n);
        const int j = col[t];
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        for (int s = 0; s < k; ++s) {
            const long u = (long)s * n + i;
            if (col[u] == j) {
                ar += (double)val[2 * u];
                ai += (double)val[2 * u + 1];
            }
            const long w = (long)s * n + j;
            if (col[w] == i) {
                br += (double)val[2 * w];
                bi += (double)val[2 * w + 1];
            }
        }
        // a_ij - conj(a_ji)
        const double dr = ar - br, di = ai + bi;
        err = max_nan(err, sqrt(dr * dr + di * di));
        const double vr = (double)val[2 * t], vi = (double)val[2 * t + 1];
        mag = max_nan(mag, sqrt(vr * vr + vi * vi));
    }
    for (int off = 32; off > 0; off >>= 1) {
        err = max_nan(err, __shfl_xor(err, off));
        mag = max_nan(mag, __shfl_xor(mag, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(err));
        atomicMax(&out[1], (unsigned long long)__double_as_longlong(mag));
    }
}

// out[0] = smallest index of a non-finite entry (initialised to n by the
// caller), out[1] = count of non-finite entries.
__global__ void __launch_bounds__(kBlock) nonfinite_kernel(const double *__restrict__ x, long n,
                                                           unsigned long long *__restrict__ out) {
    unsigned long long first = (unsigned long long)n, count = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        if (!isfinite(x[i])) {
            if ((unsigned long long)i < first) first = (unsigned long long)i;
            ++count;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        first = min(first, (unsigned long long)__shfl_xor(first, off));
        count += __shfl_xor(count, off);
    }
    if ((threadIdx.x & 63) == 0 && count) {
        atomicMin(&out[0], first);
        atomicAdd(&out[1], count);
    }
}

// rval[e] = Re val[e]; bit 0 of *flags is set when some Im val[e] != 0, bit
// 1 when some Re val[e] != 0 (NaN counts as nonzero for both).
template <typename Real>
__global__ void __launch_bounds__(kBlock) real_pack_kernel(const Real *__restrict__ val, long tot,
                                                           Real *__restrict__ rval,
                                                           unsigned *__restrict__ flags) {
    bool im = false, re = false;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (long)gridDim.x * blockDim.x) {
        rval[e] = val[2 * e];
        im |= !(val[2 * e + 1] == Real(0));
        re |= !(val[2 * e] == Real(0));
    }
    const unsigned bits = (__any(im) ? 1u : 0u) | (__any(re) ? 2u : 0u);
    if (bits && (threadIdx.x & 63) == 0) atomicOr(flags, bits);
}

// rval[e] = Im val[e] (a purely imaginary matrix).
template <typename Real>
__global__ void __launch_bounds__(kBlock) imag_pack_kernel(const Real *__restrict__ val, long tot,
                                                           Real *__restrict__ rval) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (long)gridDim.x * blockDim.x)
        rval[e] = val[2 * e + 1];
}

template <typename Real>
__global__ void __launch_bounds__(kBlock) fill_nan_kernel(Real *__restrict__ p, long n) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x)
        p[e] = __builtin_nan("");
}

}  // namespace

int poison_real_copy(Ell *H, hipStream_t s) {
    if (!H->rval) return JKPM_OK;
    const long tot = (long)H->n * H->k;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const long want = (tot + kBlock - 1) / kBlock;
    const int grid = (int)std::max(1L, std::min(want, (long)dev.cu_count * 8));
    if (H->vtype == JKPM_C64)
        hipLaunchKernelGGL(fill_nan_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                           (float *)H->rval, tot);
    else
        hipLaunchKernelGGL(fill_nan_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                           (double *)H->rval, tot);
    JKPM_CHECK_LAUNCH("fill_nan_kernel");
    return JKPM_OK;
}

int update_real_values(Ell *H, hipStream_t s) {
    static const bool enabled = [] {
        const char *e = getenv("JKPM_REAL_VALUES");
        return !(e && atoi(e) == 0);
    }();
    H->real_values = false;
    H->imag_values = false;
    if (!enabled) return JKPM_OK;
    const long tot = (long)H->n * H->k;
    const size_t rb = H->vtype == JKPM_C64 ? sizeof(float) : sizeof(double);
    if (!H->rval) JKPM_HIP(hipMalloc(&H->rval, (size_t)tot * rb));
    unsigned *d = nullptr;
    JKPM_HIP(hipMalloc(&d, sizeof(unsigned)));
    unsigned host = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned), s);
    if (e == hipSuccess) {
        DeviceInfo dev;
        int rc = device_info(&dev);
        if (rc) {
            hipFree(d);
            return rc;
        }
        const long want = (tot + kBlock - 1) / kBlock;
        const int grid = (int)std::max(1L, std::min(want, (long)dev.cu_count * 8));
        if (H->vtype == JKPM_C64)
            hipLaunchKernelGGL(real_pack_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                               (const float *)H->val, tot, (float *)H->rval, d);
        else
            hipLaunchKernelGGL(real_pack_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                               (const double *)H->val, tot, (double *)H->rval, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&host, d, sizeof(host), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "real-value check", __FILE__, __LINE__);
    H->real_values = (host & 1u) == 0;
    if (H->real_values) return JKPM_OK;
    if ((host & 2u) == 0) {
        // purely imaginary values: the copy holds the imaginary parts
        DeviceInfo dev;
        int rc = device_info(&dev);
        if (rc) return rc;
        const long want = (tot + kBlock - 1) / kBlock;
        const int grid = (int)std::max(1L, std::min(want, (long)dev.cu_count * 8));
        if (H->vtype == JKPM_C64)
            hipLaunchKernelGGL(imag_pack_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                               (const float *)H->val, tot, (float *)H->rval);
        else
            hipLaunchKernelGGL(imag_pack_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                               (const double *)H->val, tot, (double *)H->rval);
        JKPM_CHECK_LAUNCH("imag_pack_kernel");
        H->imag_values = true;
        return JKPM_OK;
    }
    // complex values: a graph captured while they were real or imaginary would
    // replay the packed-value kernels on this copy; NaN makes that visible
    return poison_real_copy(H, s);
}

int count_nonfinite(const double *x, long n, long long *first_bad, long long *count,
                    hipStream_t s) {
    *first_bad = -1;
    *count = 0;
    if (n <= 0) return JKPM_OK;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    JKPM_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return arg_error("jkpm_check_finite synchronises its stream: not under capture");
    unsigned long long *d = nullptr;
    JKPM_HIP(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long host[2] = {(unsigned long long)n, 0};
    hipError_t e = hipMemcpyAsync(d, host, sizeof(host), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        DeviceInfo dev;
        int rc = device_info(&dev);
        if (rc) {
            hipFree(d);
            return rc;
        }
        const long want = (n + kBlock - 1) / kBlock;
        const int grid = (int)std::max(1L, std::min(want, (long)dev.cu_count * 4));
        hipLaunchKernelGGL(nonfinite_kernel, dim3(grid), dim3(kBlock), 0, s, x, n, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, sizeof(host), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "non-finite check", __FILE__, __LINE__);
    if (host[1]) {
        *first_bad = (long long)host[0];
        *count = (long long)host[1];
    }
    return JKPM_OK;
}

bool inject_alloc_failure() {
    const char *v = getenv("JKPM_FAULT");
    return v && strcmp(v, "alloc") == 0;
}

int hermiticity(const jkpm_ell *h, hipStream_t s, double *abs_err, double *max_abs) {
    if (h->herm_err >= 0.0) {
        *abs_err = h->herm_err;
        *max_abs = h->herm_max;
        return JKPM_OK;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    JKPM_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return arg_error("the Hermiticity of this matrix is not known yet and cannot be checked "
                         "under stream capture: call the entry point (or jkpm_ell_hermiticity) "
                         "once before capturing");
    const Ell &H = h->e;
    unsigned long long *d = nullptr;
    JKPM_HIP(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long host[2] = {0, 0};
    hipError_t e = hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), s);
    if (e == hipSuccess) {
        DeviceInfo dev;
        int rc = device_info(&dev);
        if (rc) {
            hipFree(d);
            return rc;
        }
        const long tot = (long)H.n * H.k;
        const long want = (tot + kBlock - 1) / kBlock;
        const int grid = (int)std::max(1L, std::min(want, (long)dev.cu_count * 8));
        if (H.vtype == JKPM_C64)
            hipLaunchKernelGGL(herm_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                               (const float *)H.val, H.col, H.n, H.k, d);
        else
            hipLaunchKernelGGL(herm_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                               (const double *)H.val, H.col, H.n, H.k, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, sizeof(host), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "Hermiticity check", __FILE__, __LINE__);
    double err, mag;
    memcpy(&err, &host[0], sizeof(double));
    memcpy(&mag, &host[1], sizeof(double));
    h->herm_err = err;
    h->herm_max = mag;
    *abs_err = err;
    *max_abs = mag;
    return JKPM_OK;
}

int require_hermitian(const jkpm_ell *h, hipStream_t s, const char *what) {
    double err = 0.0, mag = 0.0;
    int rc = hermiticity(h, s, &err, &mag);
    if (rc) return rc;
    if (!(err <= JKPM_HERMITIAN_RTOL * mag)) {  // NaN values fail too
        char buf[320];
        snprintf(buf, sizeof(buf),
                 "%s needs a Hermitian matrix: max |A_ij - conj(A_ji)| = %.3e exceeds "
                 "%.0e * max |A_ij| = %.3e (use the plain recursion for non-Hermitian matrices)",
                 what, err, JKPM_HERMITIAN_RTOL, JKPM_HERMITIAN_RTOL * mag);
        return arg_error(buf);
    }
    return JKPM_OK;
}

}  // namespace jkpm
