// Jackson-damped KPM DOS reconstruction in FP64 on the device.
//
// Reference: the host loop of jclsquant/observables.py:49-52 (kpm_dos) over
// JacksonKernel (jclsquant/kernel_and_moments.py:19-24) and moments_delta
// (jclsquant/kernel_and_moments.py:72-78):
//   dos_i = Re sum_n g_n delta_n(e_i) mu_n,
//   delta_n(e) = (2 - [n == 0]) cos(n acos e) / (pi sqrt(1 - e^2)).
// g_n and delta_n are real, so dos_i = sum_n g_n delta_n(e_i) Re(mu_n).
// Here mu_n is the mean over the R vectors of a batch.
//
// Two kernels: w_n = g_n Re(mean_r mu[n][r]) once, then one work-group per
// energy sums (2 - [n==0]) w_n cos(n theta_i) over n in a fixed order.
#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;

__global__ void dos_weights_kernel(const double2 *__restrict__ mu, int M, int R, int kernel,
                                   double *__restrict__ w) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M) return;
    double s = 0.0;
    for (int r = 0; r < R; ++r) s += mu[(long)n * R + r].x;
    s /= (double)R;
    double g = 1.0;
    if (kernel == JKPM_KERNEL_JACKSON) {
        const double q = M_PI / (M + 1.0);
        g = ((M - n + 1.0) * cos(q * n) + sin(q * n) / tan(q)) / (M + 1.0);
    }
    w[n] = g * s;
}

__global__ __launch_bounds__(kBlock) void dos_kernel(const double *__restrict__ w, int M,
                                                      const double *__restrict__ e, int n_e,
                                                      double *__restrict__ dos) {
    const int i = blockIdx.x;
    if (i >= n_e) return;
    const double ei = e[i];
    const double th = acos(ei);
    double s = 0.0;
    for (int n = threadIdx.x; n < M; n += kBlock)
        s += (n == 0 ? 1.0 : 2.0) * w[n] * cos(n * th);
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) dos[i] = sh[0] / (M_PI * sqrt(1.0 - ei * ei));
}

// Transpose of the DOS map for gradients: given w_i = dL/d dos_i,
// grad_mu[n][r] = (1/R) g_n (2 - [n==0]) sum_i w_i cos(n theta_i) / (pi s_i)
// (real; dos depends on Re mu only).  One work-group per moment n.
__global__ __launch_bounds__(kBlock) void dos_grad_kernel(const double *__restrict__ w, int M,
                                                           int R, const double *__restrict__ e,
                                                           int n_e, int kernel,
                                                           double2 *__restrict__ grad_mu) {
    const int n = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < n_e; i += kBlock) {
        const double ei = e[i];
        s += w[i] * cos(n * acos(ei)) / (M_PI * sqrt(1.0 - ei * ei));
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    double g = 1.0;
    if (kernel == JKPM_KERNEL_JACKSON) {
        const double q = M_PI / (M + 1.0);
        g = ((M - n + 1.0) * cos(q * n) + sin(q * n) / tan(q)) / (M + 1.0);
    }
    const double val = sh[0] * g * (n == 0 ? 1.0 : 2.0) / (double)R;
    for (int r = threadIdx.x; r < R; r += kBlock) grad_mu[(long)n * R + r] = make_double2(val, 0.0);
}

}  // namespace

int launch_dos_grad(const double *w, int M, int R, const double *e, int n_e, int kernel,
                    double2 *grad_mu, hipStream_t s) {
    if (M <= 0) return JKPM_OK;
    hipLaunchKernelGGL(dos_grad_kernel, dim3(M), dim3(kBlock), 0, s, w, M, R, e, n_e, kernel,
                       grad_mu);
    JKPM_CHECK_LAUNCH("dos_grad_kernel");
    return JKPM_OK;
}

int launch_dos(const double2 *mu, int M, int R, const double *e, int n_e, int kernel,
               double *dos, double *w, hipStream_t s) {
    if (M <= 0 || n_e <= 0) return JKPM_OK;
    hipLaunchKernelGGL(dos_weights_kernel, dim3((M + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       mu, M, R, kernel, w);
    JKPM_CHECK_LAUNCH("dos_weights_kernel");
    hipLaunchKernelGGL(dos_kernel, dim3(n_e), dim3(kBlock), 0, s, w, M, e, n_e, dos);
    JKPM_CHECK_LAUNCH("dos_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
