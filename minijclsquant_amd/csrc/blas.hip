// Complex128 vector helpers behind the package's blas_funs module: the
// elementwise updates and dot products of jclsquant/cython_modules/
// blas_funs.pyx:15-165 (scal, axpy, axpy2, vec_mul, vec_mul2, vec_mul3, vdot,
// dot), which the reference runs as OpenMP loops on the host.
//
// Elementwise ops: one thread per element, HBM-bound (16-48 B per element).
// Dots: each work-group sums a contiguous range of elements in FP64 (lane
// partials, then a fixed wave/LDS order), writes one complex partial, and a
// single work-group adds the partials in index order, so the result does not
// depend on timing (bitwise reproducible for a given n).
#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;
constexpr int kDotGroups = 1024;  // partials of a dot (upper bound)

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// x may alias y (scal passes x = y), so no pointer here is __restrict__.
__global__ void __launch_bounds__(kBlock) zvec_kernel(int op, long n, double2 a, double2 b,
                                                      const double2 *x, double2 *y,
                                                      const double2 *w, double2 *z) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        switch (op) {
            case JKPM_VOP_AXPBY: {
                const double2 xi = x[i], yi = y[i];
                const double2 ax = cmul(a, xi), by = cmul(b, yi);
                y[i] = make_double2(ax.x + by.x, ax.y + by.y);
                break;
            }
            case JKPM_VOP_MUL:  // z = a * y * x (vec_mul)
                z[i] = cmul(cmul(a, y[i]), x[i]);
                break;
            case JKPM_VOP_MUL2:  // y = a * x * y (vec_mul2)
                y[i] = cmul(cmul(a, x[i]), y[i]);
                break;
            default:  // JKPM_VOP_MUL3: z = a * y * x * w (vec_mul3)
                z[i] = cmul(cmul(cmul(a, y[i]), x[i]), w[i]);
                break;
        }
    }
}

// part[g] = sum over the elements of group g of y_i * (conj(x_i) or x_i).
__global__ void __launch_bounds__(kBlock) zdot_partial_kernel(long n, int conj_x,
                                                              const double2 *__restrict__ x,
                                                              const double2 *__restrict__ y,
                                                              double2 *__restrict__ part) {
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * per;
    const long hi = lo + per < n ? lo + per : n;
    double sr = 0.0, si = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const double2 xi = x[i], yi = y[i];
        const double xr = xi.x, xim = conj_x ? -xi.y : xi.y;
        sr += yi.x * xr - yi.y * xim;
        si += yi.x * xim + yi.y * xr;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off);
        si += __shfl_xor(si, off);
    }
    __shared__ double sh[2][kBlock / 64];
    const int wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        sh[0][wave] = sr;
        sh[1][wave] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, ti = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) {
            tr += sh[0][q];
            ti += sh[1][q];
        }
        part[blockIdx.x] = make_double2(tr, ti);
    }
}

// out = a * sum_g part[g], summed in index order by one work-group.
__global__ void __launch_bounds__(kBlock) zdot_final_kernel(int groups, double2 a,
                                                            const double2 *__restrict__ part,
                                                            double2 *__restrict__ out) {
    __shared__ double2 sh[kBlock];
    double sr = 0.0, si = 0.0;
    for (int g = threadIdx.x; g < groups; g += blockDim.x) {
        sr += part[g].x;
        si += part[g].y;
    }
    sh[threadIdx.x] = make_double2(sr, si);
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, ti = 0.0;
        for (int q = 0; q < kBlock; ++q) {
            tr += sh[q].x;
            ti += sh[q].y;
        }
        out[0] = cmul(a, make_double2(tr, ti));
    }
}

int dot_groups(long n, const DeviceInfo &dev) {
    const long want = (n + 4095) / 4096;
    return (int)std::max(1L, std::min(want, std::min((long)kDotGroups, (long)dev.cu_count * 4)));
}

}  // namespace

int zvec(int op, long n, double2 a, double2 b, const void *x, void *y, const void *w, void *z,
         hipStream_t s) {
    if (n <= 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const long want = (n + kBlock - 1) / kBlock;
    const int grid = (int)std::max(1L, std::min(want, (long)dev.cu_count * 16));
    hipLaunchKernelGGL(zvec_kernel, dim3(grid), dim3(kBlock), 0, s, op, n, a, b,
                       (const double2 *)x, (double2 *)y, (const double2 *)w, (double2 *)z);
    JKPM_CHECK_LAUNCH("zvec_kernel");
    return JKPM_OK;
}

size_t zdot_workspace() { return (size_t)kDotGroups * sizeof(double2); }

int zdot(long n, double2 a, const void *x, const void *y, int conj_x, double2 *out,
         double2 *part, hipStream_t s) {
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int groups = n > 0 ? dot_groups(n, dev) : 0;
    if (groups > 0) {
        hipLaunchKernelGGL(zdot_partial_kernel, dim3(groups), dim3(kBlock), 0, s, n, conj_x,
                           (const double2 *)x, (const double2 *)y, part);
        JKPM_CHECK_LAUNCH("zdot_partial_kernel");
    }
    hipLaunchKernelGGL(zdot_final_kernel, dim3(1), dim3(kBlock), 0, s, groups, a,
                       (const double2 *)part, out);
    JKPM_CHECK_LAUNCH("zdot_final_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
