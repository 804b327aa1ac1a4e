// Chebyshev step kernel template, shared by the per-precision instantiation
// units cheb_c64.hip and cheb_c128.hip (see cheb.hip for the design notes).
#pragma once

#include <type_traits>

#include "common.h"

namespace jkpm {
namespace step {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;

template <typename Real, int N>
using vec_t = Real __attribute__((ext_vector_type(N)));

// Loads of data read once per step (T_{n-1} and v_l when NT, the matrix
// values and columns when NTM): non-temporal, so they do not displace the
// gather window of T_n from the caches.
template <bool NT, typename V>
__device__ __forceinline__ V load_stream(const V *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Unit complex number e^{i phi} from a 32-bit phase u = phi / (2 pi) * 2^32
// (mod 2^32).  The angle is split into the nearest quarter turn q and a rest
// a in [-pi/4, pi/4); sin a and cos a are their Taylor series (truncation
// below 2e-9 at |a| = pi/4), then rotated by q quarter turns.  Each
// component is within 1.2e-7 (2^-23) of the exact value (every 7th u
// compared with double precision).
constexpr float kPhaseUnit = 1.4629180792671596e-09f;  // 2 pi / 2^32
__device__ __forceinline__ void phase_cs(unsigned u, float &c, float &s) {
    const unsigned q = (u + (1u << 29)) >> 30;
    const float a = (float)(int)(u - (q << 30)) * kPhaseUnit;
    const float a2 = a * a;
    const float sn =
        a + a * a2 * (-1.f / 6 + a2 * (1.f / 120 + a2 * (-1.f / 5040 + a2 * (1.f / 362880))));
    const float cs =
        1.f + a2 * (-0.5f + a2 * (1.f / 24 + a2 * (-1.f / 720 + a2 * (1.f / 40320 +
                                                                   a2 * (-1.f / 3628800)))));
    // rotation by q quarter turns: (cs, sn) -> (-sn, cs) per turn
    const bool sw = q & 1, neg_c = (q + 1) & 2, neg_s = q & 2;
    const float c0 = sw ? sn : cs, s0 = sw ? cs : sn;
    c = neg_c ? -c0 : c0;
    s = neg_s ? -s0 : s0;
}

// UNR: rows per lane per loop iteration (all their loads are issued before
// any is used, for more bytes in flight per wave).
// DBL (doubling form, NL = 2, Hermitian A and v_l = v_r = T_0): instead of
// dots with stored left vectors, the two partial sets are <T_{n+1}|T_{n+1}>
// and <T_n|T_{n+1}> (T_n of the row itself, a cached re-read of the gather
// source); mu_{2n} = 2<T_n|T_n> - mu_0 and mu_{2n+1} = 2<T_n|T_{n+1}> - mu_1
// then give two moments per step with no v_l stream.  The first step's
// step-0 partials are <T_0|T_0>.
// NS: the T_{n+1} stream is written with non-temporal stores.
// NTM: the matrix values and columns are loaded non-temporally (pays when
// the step's working set is far beyond the caches, costs when it is not).
// F0: the kernel can run the first step (T_1 and the step-0 partials).  The
// step-0 accumulators cost registers: every single-form kernel keeps them,
// but of the doubling-form kernels only the base kernel of each shape has
// them, and the first step of a doubling recursion runs that one.  A kernel
// without F0 treats first as 0.
// RV (real values): `val` holds one real per slot (a matrix whose imaginary
// parts are all 0, packed by the handle), so the matrix stream is 4 (8)
// bytes per slot instead of 8 (16) and each product takes two multiplies.
// Left phases (complex64, one left set, plain form): when vph is non-NULL
// and *ph_bad is 0, v_l is read as 32-bit phases (phase_cs) from vph [n][ld]
// instead of complex64 pairs, 4 bytes per element instead of 8.  The flag is
// read on the device, so a captured graph follows the vectors of each replay.
template <typename Real, int VPL, int LPR, int NL, int KT, int UNR, bool NT, bool DBL = false,
          bool NS = false, bool NTM = false, bool F0 = false, bool RV = false>
__global__ __launch_bounds__(kBlock) void cheb_step_kernel(
    const Real *__restrict__ val, const int *__restrict__ col, int n, int k_rt,
    const Real *__restrict__ xcur, const Real *xprev, Real *xout,
    const Real *__restrict__ vl0, const Real *__restrict__ vl1,
    const unsigned *__restrict__ vph, const int *__restrict__ ph_bad, int ld,
    double2 *__restrict__ part, double2 *__restrict__ part0, int first_arg,
    int rows_per_chunk, int nchunks, const long *__restrict__ chunk_rows, int reverse) {
    const bool first = F0 && first_arg;
    using V = vec_t<Real, 2 * VPL>;
    using C2 = vec_t<Real, 2>;
    using HV = std::conditional_t<RV, Real, C2>;  // one stored matrix value
    constexpr int RPW = 64 / LPR;
    constexpr int RB = VPL * LPR;
    const int K = KT > 0 ? KT : k_rt;
    constexpr bool kPH = !DBL && NL == 1 && std::is_same_v<Real, float>;
    const bool ph = kPH && vph && *ph_bad == 0;
    auto load_h = [&](long e) -> HV {
        if constexpr (RV)
            return load_stream<NTM>(val + e);
        else
            return load_stream<NTM>(reinterpret_cast<const C2 *>(val + 2 * e));
    };
    // s + h x for one vector of a lane (complex or real h)
    auto mac = [](const HV &h, Real xr, Real xi, Real &sr_, Real &si_) {
        if constexpr (RV) {
            sr_ += h * xr;
            si_ += h * xi;
        } else {
            sr_ += h.x * xr - h.y * xi;
            si_ += h.x * xi + h.y * xr;
        }
    };

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
    double a0r[F0 ? NL : 1][VPL], a0i[F0 ? NL : 1][VPL];
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
        for (int v = 0; v < VPL; ++v) ar[l][v] = ai[l][v] = 0.0;
#pragma unroll
    for (int l = 0; l < (F0 ? NL : 1); ++l)
#pragma unroll
        for (int v = 0; v < VPL; ++v) a0r[l][v] = a0i[l][v] = 0.0;

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
            // reverse: the chunk is swept from its last row to its first
            if (reverse) rows[u] = row_begin + row_end - 1 - rows[u];
        }
        Real sr[UNR][VPL], si[UNR][VPL];
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
            for (int v = 0; v < VPL; ++v) sr[u][v] = si[u][v] = Real(0);
        // doubling form: T_n of the row itself, taken from the gather of the
        // diagonal slot when the row has one (no second read of it)
        V own[UNR];
        bool have_own[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) have_own[u] = false;

        if constexpr (KT > 0) {
            int j[UNR][KT];
            HV h[UNR][KT];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const long e = (long)s * n + rows[u];
                    j[u][s] = load_stream<NTM>(col + e);
                    h[u][s] = load_h(e);
                }
            V x[UNR][KT];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int s = 0; s < KT; ++s)
                    x[u][s] = *reinterpret_cast<const V *>(xcur + (long)j[u][s] * ld * 2 + voff);
            if constexpr (DBL) {
#pragma unroll
                for (int u = 0; u < UNR; ++u)
#pragma unroll
                    for (int s = 0; s < KT; ++s)
                        if (j[u][s] == (int)rows[u]) {
                            own[u] = x[u][s];
                            have_own[u] = true;
                        }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int s = 0; s < KT; ++s)
#pragma unroll
                    for (int v = 0; v < VPL; ++v)
                        mac(h[u][s], x[u][s][2 * v], x[u][s][2 * v + 1], sr[u][v], si[u][v]);
        } else {
#pragma unroll
            for (int u = 0; u < UNR; ++u)
                for (int s = 0; s < K; ++s) {
                    const long e = (long)s * n + rows[u];
                    const int j = load_stream<NTM>(col + e);
                    const HV h = load_h(e);
                    const V x = *reinterpret_cast<const V *>(xcur + (long)j * ld * 2 + voff);
#pragma unroll
                    for (int v = 0; v < VPL; ++v) mac(h, x[2 * v], x[2 * v + 1], sr[u][v], si[u][v]);
                }
        }

        // T_{n-1} and v_l of every row first, then the arithmetic.
        V p[UNR], q[UNR][NL], t0[UNR];
        using U = vec_t<unsigned, VPL>;
        U wph[UNR];  // left phases, turned into q[u][0] after T_{n+1} is stored
        auto phase_q = [&](int u) {
            if (kPH && ph) {
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    float c, sn;
                    phase_cs(wph[u][v], c, sn);
                    q[u][0][2 * v] = c;
                    q[u][0][2 * v + 1] = sn;
                }
            }
        };
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const long vo = rows[u] * ld * 2 + voff;
            if (first)
                t0[u] = *reinterpret_cast<const V *>(xcur + vo);
            else
                p[u] = load_stream<NT>(reinterpret_cast<const V *>(xprev + vo));
            if constexpr (DBL) {
                if (!first) {
                    if (have_own[u])
                        q[u][1] = own[u];
                    else
                        q[u][1] = *reinterpret_cast<const V *>(xcur + vo);  // T_n, cached
                }
            } else {
                if (kPH && ph) {
                    wph[u] = load_stream<NT>(
                        reinterpret_cast<const U *>(vph + rows[u] * ld + lv * VPL));
                } else {
#pragma unroll
                    for (int l = 0; l < NL; ++l)
                        q[u][l] = load_stream<NT>(reinterpret_cast<const V *>(vls[l] + vo));
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (!ok[u]) continue;
            const long vo = rows[u] * ld * 2 + voff;
            V y;
            if constexpr (DBL) {
                if (first) {
                    q[u][0] = t0[u];
                    q[u][1] = t0[u];
                }
            }
            if (first) {
                phase_q(u);
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    y[2 * v] = sr[u][v];
                    y[2 * v + 1] = si[u][v];
                }
                if constexpr (F0) {
#pragma unroll
                    for (int l = 0; l < NL; ++l)
#pragma unroll
                        for (int v = 0; v < VPL; ++v) {
                            const double qr = q[u][l][2 * v], qi = q[u][l][2 * v + 1];
                            const double tr = t0[u][2 * v], ti = t0[u][2 * v + 1];
                            a0r[l][v] += qr * tr + qi * ti;
                            a0i[l][v] += qr * ti - qi * tr;
                        }
                }
            } else {
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    y[2 * v] = Real(2) * sr[u][v] - p[u][2 * v];
                    y[2 * v + 1] = Real(2) * si[u][v] - p[u][2 * v + 1];
                }
            }
            if constexpr (NS)
                __builtin_nontemporal_store(y, reinterpret_cast<V *>(xout + vo));
            else
                *reinterpret_cast<V *>(xout + vo) = y;
            if constexpr (DBL) q[u][0] = y;
            if (!first) phase_q(u);
#pragma unroll
            for (int l = 0; l < NL; ++l)
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    const double qr = q[u][l][2 * v], qi = q[u][l][2 * v + 1];
                    const double yr = y[2 * v], yi = y[2 * v + 1];
                    ar[l][v] += qr * yr + qi * yi;
                    // <T_{n+1}|T_{n+1}> of the doubling form is real: no
                    // imaginary accumulator for it
                    if (!(DBL && l == 0)) ai[l][v] += qr * yi - qi * yr;
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
                if constexpr (F0) {
                    if (first) {
                        a0r[l][v] += __shfl_xor(a0r[l][v], o);
                        a0i[l][v] += __shfl_xor(a0i[l][v], o);
                    }
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
                if constexpr (F0) sh[1][wave][l][lv * VPL + v] = make_double2(a0r[l][v], a0i[l][v]);
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
        if (F0 && first) {
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

struct StepEntry {
    const void *fn;
    int vpl, lpr, unr;
    bool rv = false;  // the kernel reads the handle's packed real values
};

// variant v = unroll | non-temporal << 1 | lane-shape << 2 (0..11), plus 12
// for non-temporal stores of T_{n+1} (12..23); the lane shape multiplies the
// vectors per lane (VPL) by 1, 2 or 4.  Variants
// apply to single-left-set kernels of ELL width 4 and 5; every other shape
// uses the base kernel.  Returns fn = nullptr for an unsupported group size.
// dbl selects the doubling form (two partial sets, see DBL above); first
// selects the base kernel of the shape with the first-step accumulators (F0).
// variant 0..47: 24..47 are 0..23 with non-temporal matrix loads (NTM).
// rv: the handle's values are real; a kernel reading the packed real copy
// is returned where one exists (entry.rv tells which was picked).
StepEntry pick_step_c64(int rb, bool wide, int nl, int k, int variant, bool dbl = false,
                        bool first = false, bool rv = false);
StepEntry pick_step_c128(int rb, int nl, int k, int variant, bool dbl = false,
                         bool first = false, bool rv = false);
// The real-value kernels of the tuned shapes (variants 0..23, widths 4 and 5;
// fn is NULL for others)  Defined in cheb_c64_rv.hip and cheb_c128_rv.hip.
StepEntry pick_step_c64_rv(int rb, bool wide, int nl, int k, int variant, bool dbl, bool first);
StepEntry pick_step_c128_rv(int rb, int nl, int k, int variant, bool dbl, bool first);
// ... and of variants 24..47 (variant - 24 given); in cheb_c64_rv_ntm.hip and
// cheb_c128_rv_ntm.hip.
StepEntry pick_step_c64_rv_ntm(int rb, bool wide, int nl, int k, int variant, bool dbl);
StepEntry pick_step_c128_rv_ntm(int rb, int nl, int k, int variant, bool dbl);
// The NTM kernels of the tuned shapes (variant 0..23 here); fn is NULL for
// shapes that have none.  Defined in their own translation units.
StepEntry pick_step_c64_ntm(int rb, bool wide, int nl, int k, int variant, bool dbl);
StepEntry pick_step_c128_ntm(int rb, int nl, int k, int variant, bool dbl);

}  // namespace step
}  // namespace jkpm
