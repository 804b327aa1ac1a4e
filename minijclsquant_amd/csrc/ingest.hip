// Device-side CSR -> slot-major ELL ingest.
//
// The reference converts on the host: the row width is the longest CSR row
// (jclsquant/ell_matrix.py:187-189) and data_indices_c_0
// (jclsquant/cython_modules/Extra_funs.pyx:182-219) copies row i's entries
// in CSR order to slots 0..len-1, then pads slots len..K-1 with the row's
// last column and value 0.  Here the CSR arrays are read on the device and
// written straight into the slot-major layout the step kernels read, so the
// row-major host ELL is never built.  An empty row pads with its own index
// (the reference would read the previous row's last column).
//
// Kernels
//   csr_scan_kernel    validates indptr (monotone, inside [0, nnz]) and the
//                      column indices (inside [0, n)), and takes the longest
//                      row: one pass, one atomic per wave per statistic.
//   csr_to_ell_kernel  one thread per slot-major element (s, i): coalesced
//                      writes of values and columns.
//   ell_cols_unpack_kernel  slot-major -> row-major column copy.
#include <algorithm>

#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ unsigned wave_max(unsigned v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o));
    return v;
}

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (unsigned)__shfl_xor((int)v, o);
    return v;
}

// stats[0] = longest row, stats[1] = bad indptr entries, stats[2] = bad columns.
__global__ __launch_bounds__(kBlock) void csr_scan_kernel(const int *__restrict__ indptr,
                                                          const int *__restrict__ indices, int n,
                                                          long nnz, unsigned *__restrict__ stats) {
    const long stride = (long)gridDim.x * blockDim.x;
    const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long span = max((long)n, nnz);
    // every lane runs the same number of iterations (the wave reductions
    // below need the whole wave)
    const long iters = (span + stride - 1) / stride;
    unsigned mx = 0, bad_ptr = 0, bad_col = 0;
    for (long it = 0; it < iters; ++it) {
        const long t = t0 + it * stride;
        if (t < n) {
            const long a = indptr[t], b = indptr[t + 1];
            if (a < 0 || b < a || b > nnz)
                ++bad_ptr;
            else
                mx = max(mx, (unsigned)(b - a));
        }
        if (t < nnz) {
            const int j = indices[t];
            if (j < 0 || j >= n) ++bad_col;
        }
    }
    mx = wave_max(mx);
    bad_ptr = wave_sum(bad_ptr);
    bad_col = wave_sum(bad_col);
    if ((threadIdx.x & 63) == 0) {
        if (mx) atomicMax(&stats[0], mx);
        if (bad_ptr) atomicAdd(&stats[1], bad_ptr);
        if (bad_col) atomicAdd(&stats[2], bad_col);
    }
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void csr_to_ell_kernel(
    const int *__restrict__ indptr, const int *__restrict__ indices,
    const double *__restrict__ data, double scale, int n, int k, Real *__restrict__ val,
    int *__restrict__ col) {
    const long tot = (long)n * k;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < tot; o += stride) {
        const long s = o / n, i = o - s * n;
        const long a = indptr[i], b = indptr[i + 1];
        int j;
        Real vr = Real(0), vi = Real(0);
        if (a + s < b) {
            const long e = a + s;
            j = indices[e];
            vr = (Real)(data[2 * e] * scale);
            vi = (Real)(data[2 * e + 1] * scale);
        } else {
            j = b > a ? indices[b - 1] : (int)i;
        }
        val[2 * o] = vr;
        val[2 * o + 1] = vi;
        col[o] = j;
    }
}

__global__ __launch_bounds__(kBlock) void ell_cols_unpack_kernel(const int *__restrict__ col,
                                                                 int n, int k,
                                                                 int *__restrict__ out) {
    const long tot = (long)n * k;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += stride) {
        const long i = e / k, s = e - i * k;
        out[e] = col[s * n + i];
    }
}

int grid_for(long count, const DeviceInfo &dev) {
    const long want = (count + kBlock - 1) / kBlock;
    const long cap = (long)dev.cu_count * 16;
    return (int)std::max(1L, std::min(want, cap));
}

}  // namespace

int launch_csr_scan(const int *indptr, const int *indices, int n, long nnz, unsigned *stats,
                    hipStream_t s) {
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int grid = grid_for(std::max((long)n, nnz), dev);
    hipLaunchKernelGGL(csr_scan_kernel, dim3(grid), dim3(kBlock), 0, s, indptr, indices, n, nnz,
                       stats);
    JKPM_CHECK_LAUNCH("csr_scan_kernel");
    return JKPM_OK;
}

int launch_csr_to_ell(const int *indptr, const int *indices, const double *data, double scale,
                      Ell *H, hipStream_t s) {
    const long tot = (long)H->n * H->k;
    if (tot == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int grid = grid_for(tot, dev);
    if (H->vtype == JKPM_C64)
        hipLaunchKernelGGL(csr_to_ell_kernel<float>, dim3(grid), dim3(kBlock), 0, s, indptr,
                           indices, data, scale, H->n, H->k, (float *)H->val, H->col);
    else
        hipLaunchKernelGGL(csr_to_ell_kernel<double>, dim3(grid), dim3(kBlock), 0, s, indptr,
                           indices, data, scale, H->n, H->k, (double *)H->val, H->col);
    JKPM_CHECK_LAUNCH("csr_to_ell_kernel");
    return JKPM_OK;
}

int launch_ell_cols_unpack(const Ell &H, int *out, hipStream_t s) {
    const long tot = (long)H.n * H.k;
    if (tot == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    hipLaunchKernelGGL(ell_cols_unpack_kernel, dim3(grid_for(tot, dev)), dim3(kBlock), 0, s, H.col,
                       H.n, H.k, out);
    JKPM_CHECK_LAUNCH("ell_cols_unpack_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
