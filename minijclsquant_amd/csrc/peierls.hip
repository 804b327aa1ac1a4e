// Peierls phases on the device and the slot-major -> row-major value unpack.
//
//   peierls_kernel  dst[e] = a * src[e] * exp(i (ax dx[e] + ay dy[e])) for every
//                   slot e of one slot pattern, with the bond vector
//                   (dx, dy) = (Re b[e], Im b[e]) of a bonds handle on the
//                   same pattern and complex ax, ay (a real ax, ay is a pure
//                   phase).  One pass per time step replaces the reference's
//                   host-side copy of H followed by a rewrite of every value
//                   (jclsquant/cuda_cython/cuda/obv.cu:335-341 over
//                   modifiers.cu:25-68; CPU form Extra_funs.pyx:236-342).
//                   Bytes per slot: read src and b, write dst.
//   ell_unpack_kernel  row-major complex128 copy of a handle's values in
//                   the units they were uploaded in (stored / scale).
#include <algorithm>

#include "common.h"

namespace jkpm {
namespace {

constexpr int kBlock = 256;

template <typename T, typename B>
__global__ __launch_bounds__(kBlock) void peierls_kernel(const T *src,
                                                         const B *__restrict__ bonds, T *dst,
                                                         long count,
                                                         double ar, double ai, double axr,
                                                         double axi, double ayr, double ayi) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride) {
        const double dx = (double)bonds[2 * e], dy = (double)bonds[2 * e + 1];
        // i (ax dx + ay dy) = -(axi dx + ayi dy) + i (axr dx + ayr dy)
        const double ph = axr * dx + ayr * dy;
        const double mag = exp(-(axi * dx + ayi * dy));
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double fr = mag * cs, fi = mag * sn;
        const double gr = ar * fr - ai * fi, gi = ar * fi + ai * fr;
        const double vr = (double)src[2 * e], vi = (double)src[2 * e + 1];
        dst[2 * e] = (T)(vr * gr - vi * gi);
        dst[2 * e + 1] = (T)(vr * gi + vi * gr);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void ell_unpack_kernel(const T *__restrict__ val, int n,
                                                            int k, double inv_scale,
                                                            double *__restrict__ out) {
    const long tot = (long)n * k;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += stride) {
        const long i = e / k, s = e % k;
        const long o = s * n + i;
        out[2 * e] = (double)val[2 * o] * inv_scale;
        out[2 * e + 1] = (double)val[2 * o + 1] * inv_scale;
    }
}

int grid_for(long count, const DeviceInfo &dev) {
    const long want = (count + kBlock - 1) / kBlock;
    const long cap = (long)dev.cu_count * 16;
    return (int)std::max(1L, std::min(want, cap));
}

}  // namespace

int launch_peierls(const Ell &src, const Ell &bonds, Ell *dst, double2 a, double2 ax, double2 ay,
                   hipStream_t s) {
    const long count = (long)src.n * src.k;
    if (count == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int grid = grid_for(count, dev);
#define JKPM_PEIERLS(T, B)                                                                     \
    hipLaunchKernelGGL((peierls_kernel<T, B>), dim3(grid), dim3(kBlock), 0, s,                 \
                       (const T *)src.val, (const B *)bonds.val, (T *)dst->val, count, a.x, a.y, \
                       ax.x, ax.y, ay.x, ay.y)
    const bool s64 = src.vtype == JKPM_C64, b64 = bonds.vtype == JKPM_C64;
    if (s64 && b64)
        JKPM_PEIERLS(float, float);
    else if (s64)
        JKPM_PEIERLS(float, double);
    else if (b64)
        JKPM_PEIERLS(double, float);
    else
        JKPM_PEIERLS(double, double);
#undef JKPM_PEIERLS
    JKPM_CHECK_LAUNCH("peierls_kernel");
    return JKPM_OK;
}

int launch_ell_unpack(const Ell &H, double *out, hipStream_t s) {
    const long count = (long)H.n * H.k;
    if (count == 0) return JKPM_OK;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int grid = grid_for(count, dev);
    const double inv = 1.0 / H.scale;
    if (H.vtype == JKPM_C64)
        hipLaunchKernelGGL(ell_unpack_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                           (const float *)H.val, H.n, H.k, inv, out);
    else
        hipLaunchKernelGGL(ell_unpack_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                           (const double *)H.val, H.n, H.k, inv, out);
    JKPM_CHECK_LAUNCH("ell_unpack_kernel");
    return JKPM_OK;
}

}  // namespace jkpm
