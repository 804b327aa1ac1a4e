// Two Chebyshev steps per launch by a wavefront over the row chunks:
// temporal blocking through the L2 cache, with no halo recomputation.
//
// The one-step kernel (cheb_step.h) reads T_n, T_{n-1} (and v_l) and writes
// T_{n+1} once per step.  Here one launch makes T_{n+1} and T_{n+2}: in
// iteration it, each wave makes T_{n+1} on its rows of block it (phase 1,
// stored to out1) and T_{n+2} on its rows of block it - D (phase 2, gathering
// T_{n+1} from out1), after the progress flags of every chunk block it - D
// reads show the blocks it reads done.  Work-groups own the same row chunks
// as the one-step kernel, assigned XCD-major, so the chunks a block reads run
// on the same XCD at about the same pace and the gathered T_{n+1} rows are
// fresh in that XCD's L2.  Per two steps the matrix, T_n, T_{n-1} and v_l
// come from HBM once and T_{n+1}, T_{n+2} are written once: N K (s_v + 4) +
// R N s_x (4 or 5) bytes, against twice N K (s_v + 4) + R N s_x (3 or 4) for
// two one-step launches.
//
// Coherence.  A wave writes its flag only after its T_{n+1} stores have
// completed (s_waitcnt vmcnt(0): the L2 has them).  Flags are written and
// read in the L2 (read_flag / write_flag), which all CUs of one XCD share;
// T_{n+1} is gathered with plain loads: out1 is a buffer no one reads before
// its rows are flagged, so its lines were never in a CU's L1 before they
// were written.  Rows whose columns reach a chunk on another XCD (whose L2 is
// not coherent with this one inside a launch) are finished by
// wave2_fix_kernel after the launch; rows whose columns reach a block further
// ahead than the lag (the periodic wrap of a lattice row) by their own
// work-group after its main loop.
//
// Waits end: the plan bounds every requirement by its lag, and the kernel's
// lag D exceeds it (a wave waiting on block q has published q + D blocks and
// needs at most q + lag + 1), the launch is cooperative (every work-group
// resident), and each wait gives up after spin_limit polls, raising bit 2 of
// *err (bit 1: a work-group ran on another XCD than the chunk mapping
// assumes); once bit 2 is up, no wait of this launch or a later one waits.
// The host reads *err after the recursion and reports the results invalid.
//
// Measured on MI355X (honeycomb 1024 x 1024 cells, R = 16) this path is
// slower than two one-step launches: the step kernel is bound by the latency
// of its dependent loads, not by bytes, and without any flag traffic the
// two-phase loop already takes as long as two one-step launches (plain) or
// longer (doubling); the one-step kernel stays the default.
#include "common.h"

namespace jkpm {
namespace {

constexpr int kCW = 4;  // waves per work-group of the main kernel
constexpr int kWB = kCW * 64;
constexpr int kFB = 256;  // threads per fix-up work-group
constexpr int kFW = kFB / 64;
constexpr int kXcdW = 8;
constexpr unsigned kEpochShift = 12;     // flag = kFlagMark | epoch << 12 | blocks done
constexpr unsigned kFlagMark = 1u << 31;  // set in every flag value written

template <typename Real, int N>
using wvec_t = Real __attribute__((ext_vector_type(N)));

// XCC_ID hardware register, bits [3:0]: the XCD this wave runs on.
__device__ __forceinline__ unsigned xcd_id() {
    return (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;
}

// s_waitcnt vmcnt(0): every vector memory operation of the wave completed
// (the stores acknowledged by the L2).
__device__ __forceinline__ void wait_vm_all() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// The error word (rare reads; device scope, as its writers' atomics).
__device__ __forceinline__ unsigned load_flag(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A progress flag as the XCD's L2 holds it.  The read is an atomic OR of
// kFlagMark, which every written value already carries, so it changes
// nothing; an atomic is performed in the L2, the coherence point of all CUs
// of one XCD, and workgroup scope leaves off the device-coherence bit (sc1)
// that would send the access past the L2.  Writers and readers of a flag are
// on one XCD (the plan keeps every requirement on the chunk's own XCD).
__device__ __forceinline__ unsigned read_flag(unsigned *p) {
    return __hip_atomic_fetch_or(p, kFlagMark, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void write_flag(unsigned *p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The error word, read by one lane and shared with the wave.
__device__ __forceinline__ unsigned load_err(const unsigned *err, int lane) {
    unsigned e = 0;
    if (lane == 0) e = load_flag(err);
    return __shfl(e, 0);
}

// The rows' phase 2: T_{n+2} = 2 A T_{n+1} - T_n from the gathered x2 and
// the row's own T_n (tn); stored non-temporally; moment partials against
// qv (T_{n+1} of the row in the doubling form, v_l otherwise).
template <int RB, int KT, int VPL, bool DBL, typename V, typename C2>
__device__ __forceinline__ void wave_finish(const C2 (&h)[KT], const V (&x)[KT], const V &tn,
                                            const V &qv, float *out2, long vo,
                                            double (&ar)[DBL ? 2 : 1][VPL],
                                            double (&ai)[DBL ? 2 : 1][VPL]) {
    V y;
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        float sr = 0.f, si = 0.f;
#pragma unroll
        for (int s = 0; s < KT; ++s) {
            sr += h[s].x * x[s][2 * v] - h[s].y * x[s][2 * v + 1];
            si += h[s].x * x[s][2 * v + 1] + h[s].y * x[s][2 * v];
        }
        y[2 * v] = 2.f * sr - tn[2 * v];
        y[2 * v + 1] = 2.f * si - tn[2 * v + 1];
    }
    __builtin_nontemporal_store(y, reinterpret_cast<V *>(out2 + vo));
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        const double yr = y[2 * v], yi = y[2 * v + 1];
        const double tr = qv[2 * v], ti = qv[2 * v + 1];
        if constexpr (DBL) {
            ar[0][v] += yr * yr + yi * yi;
            ar[1][v] += tr * yr + ti * yi;
            ai[1][v] += tr * yi - ti * yr;
        } else {
            ar[0][v] += tr * yr + ti * yi;
            ai[0][v] += tr * yi - ti * yr;
        }
    }
}

// Waits until every chunk listed (lanes 0..15: chunk lane / 4's wave
// lane % 4) shows at least `need`; bounded by spin_limit polls.
__device__ __forceinline__ void wave_wait(unsigned *flags, int fchunk, bool act,
                                          unsigned need, unsigned fv, int lane, unsigned *err,
                                          int &spin_limit) {
    bool ready = !act || fv >= need;
    int spins = 0;
    while (__ballot(!ready) != 0) {
        if (++spins > spin_limit) {
            if (lane == 0) atomicOr(err, 2u);
            spin_limit = 0;
            break;
        }
        if ((spins & 255) == 0 && (load_err(err, lane) & 2u)) {
            spin_limit = 0;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (act) fv = read_flag(flags + (long)fchunk * kCW + (lane & 3));
        ready = !act || fv >= need;
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// One (chunk, step pair) per work-group of kCW waves; RB = 8 or 16 complex64
// vectors, two per lane (16-byte accesses); KT slots per row; D = blocks
// between a block's phase 1 and its phase 2 (at least the plan's lag + 1).
//
// Iteration it of a wave: the progress flags block it - D needs are loaded
// first (agent scope: the XCD's L2), then block it - D's columns, values,
// own T_n and T_{n+1} or v_l (again, from L2), then block it's columns and
// values (HBM); the phase-2 gathers go out as soon as the flags clear, with
// phase 1's, so the two phases' memory chains overlap.  After the loop the
// work-group finishes its tail rows once the chunks they read are done.
// (at least 4 waves per SIMD: 128 VGPRs, so that 4 work-groups fit a CU)
template <int RB, int KT, int D, bool DBL>
__global__ __launch_bounds__(kWB, 4) void wave2_kernel(
    const float *__restrict__ val, const int *__restrict__ col, int n,
    const float *__restrict__ xprev, const float *__restrict__ xcur, float *out1,
    float *__restrict__ out2, const float *__restrict__ vl, double2 *__restrict__ part1,
    double2 *__restrict__ part2, int nchunks, int rpc, int nblk,
    const WaveDep *__restrict__ deps, const int *__restrict__ tail_off,
    const int *__restrict__ tail_rows, const int *__restrict__ tail_dep, unsigned *flags,
    unsigned epoch, unsigned *err, int spin_limit) {
    using V = wvec_t<float, 4>;
    using C2 = wvec_t<float, 2>;
    constexpr int VPL = 2, LPR = RB / VPL, RPW = 64 / LPR, BLK = kCW * RPW, NL = DBL ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lv = lane % LPR;
    const int lr = lane / LPR;
    const int voff = lv * VPL * 2;
    const int b = blockIdx.x;
    const int chunk = (b % kXcdW) * (nchunks / kXcdW) + b / kXcdW;
    if (threadIdx.x == 0 && xcd_id() != (unsigned)(b % kXcdW)) atomicOr(err, 1u);
    const long row_begin = (long)chunk * rpc;
    const long row_end = min((long)n, row_begin + rpc);
    const unsigned ebase = kFlagMark | (epoch << kEpochShift);
    // after a wait has given up (this launch or an earlier one), nothing waits
    if (load_err(err, lane) & 2u) spin_limit = 0;

    // [step][set][v]; the doubling form's set 0 (<T|T>) is real
    double ar1[NL][VPL], ai1[NL][VPL], ar2[NL][VPL], ai2[NL][VPL];
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
        for (int v = 0; v < VPL; ++v) ar1[l][v] = ai1[l][v] = ar2[l][v] = ai2[l][v] = 0.0;

    for (int it = 0; it < nblk + D; ++it) {
        const int q = it - D;
        // flags of the chunks block q reads (lane: chunk lane / 4, wave lane % 4)
        const WaveDep *dp = deps + (long)chunk * nblk + (q >= 0 ? q : 0);
        const int dcnt = q >= 0 ? dp->cnt : 0;
        const bool fl_act = lane < 4 * dcnt;
        const int fchunk = fl_act ? dp->chunk[lane >> 2] : 0;
        const unsigned fneed = fl_act ? (ebase | dp->need[lane >> 2]) : 0u;
        unsigned fv = 0;
        if (fl_act) fv = read_flag(flags + (long)fchunk * kCW + (lane & 3));
        // block q's own data (L2): columns, values, T_n and T_{n+1} or v_l
        const int r2 = (int)row_begin + (q >= 0 ? q : 0) * BLK + wave * RPW + lr;
        const unsigned msk = q >= 0 ? dp->mask[wave] : ~0u;
        const bool ok2 = q >= 0 && r2 < row_end && !((msk >> lr) & 1u);
        const int vo2 = r2 * (RB * 2) + voff;
        int j2[KT];
        C2 h2[KT];
        V tn2, qv2;
        if (ok2) {
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const int e = s * n + r2;
                j2[s] = col[e];
                h2[s] = *reinterpret_cast<const C2 *>(val + 2 * e);
            }
            tn2 = *reinterpret_cast<const V *>(xcur + vo2);
            qv2 = *reinterpret_cast<const V *>((DBL ? (const float *)out1 : vl) + vo2);
        }
        // block it's columns and values (HBM)
        const bool p1 = it < nblk;
        const int r = (int)row_begin + (p1 ? it : 0) * BLK + wave * RPW + lr;
        const bool ok = p1 && r < row_end;
        const int rr = ok ? r : (int)row_begin;
        int j[KT];
        C2 h[KT];
        if (p1) {
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const int e = s * n + rr;
                j[s] = col[e];
                h[s] = *reinterpret_cast<const C2 *>(val + 2 * e);
            }
        }
        // phase 2 gathers once block q's rows are ready; rows that are not
        // computed here read nothing from out1 (their columns may name rows
        // that are not written yet)
        if (q >= 0) wave_wait(flags, fchunk, fl_act, fneed, fv, lane, err, spin_limit);
        V x2[KT];
        if (ok2) {
#pragma unroll
            for (int s = 0; s < KT; ++s)
                x2[s] = *reinterpret_cast<const V *>(out1 + j2[s] * (RB * 2) + voff);
        }
        // phase 1 gathers and streams
        V x[KT], tn, pv, lvv;
        const int vo = rr * (RB * 2) + voff;
        if (p1) {
#pragma unroll
            for (int s = 0; s < KT; ++s)
                x[s] = *reinterpret_cast<const V *>(xcur + j[s] * (RB * 2) + voff);
            bool have = false;
#pragma unroll
            for (int s = 0; s < KT; ++s)
                if (j[s] == rr) {
                    tn = x[s];
                    have = true;
                }
            if (!have) tn = *reinterpret_cast<const V *>(xcur + vo);
            pv = __builtin_nontemporal_load(reinterpret_cast<const V *>(xprev + vo));
            if constexpr (!DBL) lvv = *reinterpret_cast<const V *>(vl + vo);
        }
        if (ok2) wave_finish<RB, KT, VPL, DBL>(h2, x2, tn2, qv2, out2, vo2, ar2, ai2);
        if (p1) {
            V y;
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    sr += h[s].x * x[s][2 * v] - h[s].y * x[s][2 * v + 1];
                    si += h[s].x * x[s][2 * v + 1] + h[s].y * x[s][2 * v];
                }
                y[2 * v] = 2.f * sr - pv[2 * v];
                y[2 * v + 1] = 2.f * si - pv[2 * v + 1];
            }
            if (ok) {
                *reinterpret_cast<V *>(out1 + vo) = y;
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    const double yr = y[2 * v], yi = y[2 * v + 1];
                    if constexpr (DBL) {
                        ar1[0][v] += yr * yr + yi * yi;
                        const double tr = tn[2 * v], ti = tn[2 * v + 1];
                        ar1[1][v] += tr * yr + ti * yi;
                        ai1[1][v] += tr * yi - ti * yr;
                    } else {
                        const double qr = lvv[2 * v], qi = lvv[2 * v + 1];
                        ar1[0][v] += qr * yr + qi * yi;
                        ai1[0][v] += qr * yi - qi * yr;
                    }
                }
            }
            // publish: the wave's rows of blocks 0..it hold T_{n+1}
            wait_vm_all();
            if (lane == 0) write_flag(flags + (long)chunk * kCW + wave, ebase | (unsigned)(it + 1));
        }
    }

    // Tail rows: once every chunk they read has finished all its blocks.
    const int t_lo = tail_off[chunk], t_hi = tail_off[chunk + 1];
    if (t_hi > t_lo) {
        const bool act = lane < 16 && tail_dep[chunk * 4 + (lane >> 2)] >= 0;
        const int fchunk = act ? tail_dep[chunk * 4 + (lane >> 2)] : 0;
        const unsigned need = ebase | (unsigned)nblk;
        unsigned fv = 0;
        if (act) fv = read_flag(flags + (long)fchunk * kCW + (lane & 3));
        wave_wait(flags, fchunk, act, need, fv, lane, err, spin_limit);
        for (int t0 = t_lo + wave * RPW; t0 < t_hi; t0 += kCW * RPW) {
            const int t = t0 + lr;
            if (t < t_hi) {
                const long i = tail_rows[t];
                int jt[KT];
                C2 ht[KT];
                V xt[KT];
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const long e = (long)s * n + i;
                    jt[s] = col[e];
                    ht[s] = *reinterpret_cast<const C2 *>(val + 2 * e);
                }
#pragma unroll
                for (int s = 0; s < KT; ++s)
                    xt[s] = *reinterpret_cast<const V *>(out1 + (long)jt[s] * RB * 2 + voff);
                const long vot = i * RB * 2 + voff;
                const V tnt = *reinterpret_cast<const V *>(xcur + vot);
                const V qvt = *reinterpret_cast<const V *>((DBL ? (const float *)out1 : vl) + vot);
                wave_finish<RB, KT, VPL, DBL>(ht, xt, tnt, qvt, out2, vot, ar2, ai2);
            }
        }
    }

    // Butterfly over the lanes of equal vector index, then a fixed wave order.
#pragma unroll
    for (int o = LPR; o < 64; o <<= 1)
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                ar1[l][v] += __shfl_xor(ar1[l][v], o);
                ar2[l][v] += __shfl_xor(ar2[l][v], o);
                if (!(DBL && l == 0)) {  // <T|T> is real
                    ai1[l][v] += __shfl_xor(ai1[l][v], o);
                    ai2[l][v] += __shfl_xor(ai2[l][v], o);
                }
            }
    __shared__ double2 sh[2][kCW][NL][RB];
    if (lr == 0) {
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                sh[0][wave][l][lv * VPL + v] = make_double2(ar1[l][v], ai1[l][v]);
                sh[1][wave][l][lv * VPL + v] = make_double2(ar2[l][v], ai2[l][v]);
            }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * NL * RB; e += kWB) {
        const int t = e / (NL * RB), l = (e / RB) % NL, r = e % RB;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int w = 0; w < kCW; ++w) {
            sr += sh[t][w][l][r].x;
            si += sh[t][w][l][r].y;
        }
        double2 *part = t == 0 ? part1 : part2;
        part[((long)l * nchunks + chunk) * RB + r] = make_double2(sr, si);
    }
}

// Phase 2 of the rows whose columns reach another XCD (after the main launch,
// so every T_{n+1} row is visible): one work-group per group of at most
// kFixGroup rows of one chunk; each group's moment partials go to fixpart.
template <int RB, int KT, bool DBL>
__global__ __launch_bounds__(kFB) void wave2_fix_kernel(
    const float *__restrict__ val, const int *__restrict__ col, int n,
    const float *__restrict__ xcur, const float *__restrict__ out1, float *__restrict__ out2,
    const float *__restrict__ vl, const int *__restrict__ fix_group_off,
    const int *__restrict__ fix_rows, double2 *__restrict__ fixpart) {
    using V = wvec_t<float, 4>;
    using C2 = wvec_t<float, 2>;
    constexpr int VPL = 2, LPR = RB / VPL, RPW = 64 / LPR, NL = DBL ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lv = lane % LPR;
    const int lr = lane / LPR;
    const int voff = lv * VPL * 2;
    const int g = blockIdx.x;
    const int lo = fix_group_off[g], hi = fix_group_off[g + 1];
    double ar[NL][VPL], ai[NL][VPL];
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
        for (int v = 0; v < VPL; ++v) ar[l][v] = ai[l][v] = 0.0;
    for (int t0 = lo + wave * RPW; t0 < hi; t0 += kFW * RPW) {
        const int t = t0 + lr;
        if (t < hi) {
            const long i = fix_rows[t];
            int jt[KT];
            C2 ht[KT];
            V xt[KT];
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const long e = (long)s * n + i;
                jt[s] = col[e];
                ht[s] = *reinterpret_cast<const C2 *>(val + 2 * e);
            }
#pragma unroll
            for (int s = 0; s < KT; ++s)
                xt[s] = *reinterpret_cast<const V *>(out1 + (long)jt[s] * RB * 2 + voff);
            const long vot = i * RB * 2 + voff;
            const V tnt = *reinterpret_cast<const V *>(xcur + vot);
            const V qvt = *reinterpret_cast<const V *>((DBL ? out1 : vl) + vot);
            wave_finish<RB, KT, VPL, DBL>(ht, xt, tnt, qvt, out2, vot, ar, ai);
        }
    }
#pragma unroll
    for (int o = LPR; o < 64; o <<= 1)
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                ar[l][v] += __shfl_xor(ar[l][v], o);
                ai[l][v] += __shfl_xor(ai[l][v], o);
            }
    __shared__ double2 sh[kFW][NL][RB];
    if (lr == 0) {
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int v = 0; v < VPL; ++v) sh[wave][l][lv * VPL + v] = make_double2(ar[l][v], ai[l][v]);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NL * RB; e += kFB) {
        const int l = e / RB, r = e % RB;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int w = 0; w < kFW; ++w) {
            sr += sh[w][l][r].x;
            si += sh[w][l][r].y;
        }
        fixpart[((long)g * NL + l) * RB + r] = make_double2(sr, si);
    }
}

// part2[l][chunk][r] += the chunk's groups' partials, in group order.
__global__ void wave2_fold_kernel(const double2 *__restrict__ fixpart,
                                  const int *__restrict__ fold_chunk,
                                  const int *__restrict__ fold_off, int nfold, int nl, int rb,
                                  int nchunks, double2 *__restrict__ part2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfold * nl * rb) return;
    const int r = t % rb, l = (t / rb) % nl, f = t / (rb * nl);
    double2 &dst = part2[((long)l * nchunks + fold_chunk[f]) * rb + r];
    double sr = dst.x, si = dst.y;
    for (int g = fold_off[f]; g < fold_off[f + 1]; ++g) {
        const double2 p = fixpart[((long)g * nl + l) * rb + r];
        sr += p.x;
        si += p.y;
    }
    dst = make_double2(sr, si);
}

// The kernel's lag D (blocks between a block's two phases) is 2 or 3; it
// must be at least the plan's lag + 1.
template <int RB, int KT, bool DBL>
const void *pick_wave_d(int d) {
    if (d == 2) return (const void *)wave2_kernel<RB, KT, 2, DBL>;
    if (d == 3) return (const void *)wave2_kernel<RB, KT, 3, DBL>;
    return nullptr;
}

template <int RB, bool DBL>
const void *pick_wave_k(int k, int d) {
    if (k == 4) return pick_wave_d<RB, 4, DBL>(d);
    if (k == 5) return pick_wave_d<RB, 5, DBL>(d);
    return nullptr;
}

template <int RB, bool DBL>
const void *pick_fix_k(int k) {
    if (k == 4) return (const void *)wave2_fix_kernel<RB, 4, DBL>;
    if (k == 5) return (const void *)wave2_fix_kernel<RB, 5, DBL>;
    return nullptr;
}

}  // namespace

int wave_rows_per_wave(int rb) { return rb == 16 ? 8 : rb == 8 ? 16 : 0; }

const void *wave_kernel(int rb, int k, int d, bool doubling) {
    if (rb == 16) return doubling ? pick_wave_k<16, true>(k, d) : pick_wave_k<16, false>(k, d);
    if (rb == 8) return doubling ? pick_wave_k<8, true>(k, d) : pick_wave_k<8, false>(k, d);
    return nullptr;
}

const void *wave_fix_kernel(int rb, int k, bool doubling) {
    if (rb == 16) return doubling ? pick_fix_k<16, true>(k) : pick_fix_k<16, false>(k);
    if (rb == 8) return doubling ? pick_fix_k<8, true>(k) : pick_fix_k<8, false>(k);
    return nullptr;
}

int launch_wave2(const WaveArgs &a, hipStream_t s) {
    const void *fn = wave_kernel(a.rb, a.H->k, a.d, a.doubling);
    if (!fn) return arg_error("wavefront path: unsupported shape");
    const float *p_val = (const float *)a.H->val;
    const int *p_col = a.H->col;
    int p_n = a.H->n;
    const float *p_xprev = (const float *)a.xprev, *p_xcur = (const float *)a.xcur;
    float *p_out1 = (float *)a.out1, *p_out2 = (float *)a.out2;
    const float *p_vl = (const float *)a.vl;
    double2 *p_part1 = a.part1, *p_part2 = a.part2;
    int p_nch = a.nchunks, p_rpc = a.rpc, p_nblk = a.nblk;
    const WaveDep *p_deps = a.deps;
    const int *p_toff = a.tail_off, *p_trows = a.tail_rows, *p_tdep = a.tail_dep;
    unsigned *p_flags = a.flags, *p_err = a.err;
    unsigned p_epoch = a.epoch;
    int p_spin = a.spin_limit;
    void *args[] = {&p_val,  &p_col,   &p_n,    &p_xprev, &p_xcur,  &p_out1,  &p_out2,
                    &p_vl,   &p_part1, &p_part2, &p_nch,  &p_rpc,   &p_nblk,  &p_deps,
                    &p_toff, &p_trows, &p_tdep, &p_flags, &p_epoch, &p_err,   &p_spin};
    JKPM_HIP(hipLaunchCooperativeKernel(fn, dim3(a.nchunks), dim3(kWB), args, 0, s));
    JKPM_CHECK_LAUNCH("wave2_kernel");
    if (a.nfix > 0) {
        const void *fx = wave_fix_kernel(a.rb, a.H->k, a.doubling);
        if (!fx) return arg_error("wavefront path: unsupported shape");
        const int *p_go = a.fix_group_off, *p_fr = a.fix_rows;
        double2 *p_fp = a.fixpart;
        void *fargs[] = {&p_val, &p_col, &p_n, &p_xcur, &p_out1, &p_out2, &p_vl,
                         &p_go,  &p_fr,  &p_fp};
        JKPM_HIP(hipLaunchKernel(fx, dim3(a.nfix), dim3(kFB), fargs, 0, s));
        JKPM_CHECK_LAUNCH("wave2_fix_kernel");
        const int nl = a.doubling ? 2 : 1;
        const int tot = a.nfold * nl * a.rb;
        hipLaunchKernelGGL(wave2_fold_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, a.fixpart,
                           a.fold_chunk, a.fold_off, a.nfold, nl, a.rb, a.nchunks, a.part2);
        JKPM_CHECK_LAUNCH("wave2_fold_kernel");
    }
    return JKPM_OK;
}

int wave_max_blocks_per_cu(int rb, int k, int d, bool doubling, int *out) {
    const void *fn = wave_kernel(rb, k, d, doubling);
    if (!fn) return arg_error("wavefront path: unsupported shape");
    int nb = 0;
    JKPM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kWB, 0));
    *out = nb;
    return JKPM_OK;
}

}  // namespace jkpm
