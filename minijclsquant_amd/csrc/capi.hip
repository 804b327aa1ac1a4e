// extern "C" entry points of libjclkpm_hip.so (declared in include/jclkpm.h).
//
// Host orchestration of the Chebyshev recursion: the reference drives one
// SpMV launch plus one host-synchronising Thrust reduction per step and
// re-uploads the matrix on every call (jclsquant/cuda_cython/cuda/
// recursions.cu:17-98).  Here the matrix lives on the device behind a handle,
// every step is one fused kernel launch on the caller's stream, moment
// partials go to a ring of `kRing` steps that one reduction launch folds into
// mu, and nothing synchronises the host until the caller does.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <vector>

#include "common.h"

namespace jkpm {

namespace {
thread_local std::string g_err;

constexpr int kRing = 64;  // steps of partials held before one reduction launch
constexpr int kXcdCount = 8;  // XCDs of the MI355X: tables are padded to a multiple

// ---- profiling of the step kernel (events on the launch stream) ----
// A recorded run: events before its first and after its last step launch,
// and the number of steps in between.
struct ProfRun {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int steps = 0;
    int kind = JKPM_PROF_STEP;  // which kernel the run times
};
constexpr int kProfRun = 8;  // steps per recorded run (at most)

struct Prof {
    std::mutex mu;
    int every = 0;  // start a recorded run every `every` steps (0 = off)
    long long counter = 0;
    std::vector<ProfRun> pending;
    std::vector<hipEvent_t> free_events;
    double ms[JKPM_PROF_KINDS] = {};
    long long launches[JKPM_PROF_KINDS] = {};
};
Prof &prof() {
    static Prof p;
    return p;
}

hipEvent_t take_event(Prof &p) {
    if (!p.free_events.empty()) {
        hipEvent_t e = p.free_events.back();
        p.free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Folds the durations of all recorded runs into the totals (waits for them).
int drain_profile(Prof &p) {
    for (auto &pr : p.pending) {
        JKPM_HIP(hipEventSynchronize(pr.e1));
        float ms = 0.f;
        JKPM_HIP(hipEventElapsedTime(&ms, pr.e0, pr.e1));
        p.ms[pr.kind] += ms;
        p.launches[pr.kind] += pr.steps;
        p.free_events.push_back(pr.e0);
        p.free_events.push_back(pr.e1);
    }
    p.pending.clear();
    return JKPM_OK;
}

__global__ void count_bad_indices(const int *__restrict__ idx, long tot, int n,
                                  unsigned int *__restrict__ bad) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < tot) {
        const int j = idx[t];
        if (j < 0 || j >= n) atomicAdd(bad, 1u);
    }
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

struct MomentsLayout {
    size_t b0 = 0, b1 = 0, ring = 0, raw = 0, total = 0;
    size_t ph = 0, ph_flag = 0;  // left-vector phases and their flag (when has_ph)
    bool has_ph = false;
    int nchunks = 1;
};

// Steps of the doubling form for M moments (T_1 .. T_S, S = max(1, M/2)).
int doubling_steps(int M) { return std::max(1, M / 2); }

// Row-chunk table of a block-diagonal batch of G realizations sharing one
// handle: chunks never straddle two realizations, and chunks of group g are
// group_chunks[g] .. group_chunks[g+1].
struct Groups {
    const long *chunk_rows = nullptr;  // device [nchunks + 1]
    const int *group_chunks = nullptr; // device [G + 1]
    int G = 1;
    int nchunks = 0;
};

// doubling_M > 0: the doubling form for that many moments (two partial sets
// and a raw-dot buffer [2][G][S+1][R]).  A complex64 plain recursion with one
// left set also gets a buffer for the left vectors' phases (4 bytes per
// element) and its flag, unless the handle turned them off.
MomentsLayout moments_layout_once(const Ell &H, int R, int n_left, const DeviceInfo &dev,
                                  const Groups *grp = nullptr, int doubling_M = 0) {
    MomentsLayout L;
    const size_t vec = align_up((size_t)H.n * R * cplx_bytes(H.vtype), 256);
    if (doubling_M > 0) n_left = 2;
    L.nchunks = grp ? grp->nchunks : step_chunks(H, R, dev, doubling_M > 0);
    L.b0 = 0;
    L.b1 = vec;
    L.ring = 2 * vec;
    L.raw = L.ring + align_up((size_t)kRing * n_left * L.nchunks * R * sizeof(double2), 256);
    L.total = L.raw;
    if (doubling_M > 0) {
        const size_t G = grp ? grp->G : 1;
        L.total += align_up(2 * G * (doubling_steps(doubling_M) + 1) * R * sizeof(double2), 256);
    }
    if (H.left_phases && H.vtype == JKPM_C64 && doubling_M == 0 && n_left == 1) {
        L.has_ph = true;
        L.ph = L.total;
        L.ph_flag = L.ph + align_up((size_t)H.n * R * sizeof(unsigned), 256);
        L.total = L.ph_flag + 256;
    }
    return L;
}

// Vectors per recursion pass for a call with R vectors (R: all at once).
int pass_vectors(const Ell &H, int R, bool doubling) {
    const int p = doubling ? H.dpass_r : H.pass_r;
    return (p > 0 && p < R) ? p : R;
}

// Workspace of a call that runs its R vectors in passes of g < R: the pass's
// copies of v_r (and of distinct left sets, plain form) as [N][g], then the
// layout of one pass of g vectors at offset `inner`.
struct PassLayout {
    int g = 0;
    size_t vec = 0, inner = 0, total = 0;
    MomentsLayout once;
};

PassLayout pass_layout(const Ell &H, int R, int n_left, const DeviceInfo &dev, const Groups *grp,
                       int doubling_M) {
    PassLayout P;
    P.g = pass_vectors(H, R, doubling_M > 0);
    P.once = moments_layout_once(H, P.g, n_left, dev, grp, doubling_M);
    if (P.g < R) {
        P.vec = align_up((size_t)H.n * P.g * cplx_bytes(H.vtype), 256);
        P.inner = (size_t)(doubling_M > 0 ? 1 : n_left + 1) * P.vec;
    }
    P.total = P.inner + P.once.total;
    return P;
}

// Workspace bytes of a moments call: enough for the handle's passes and for
// all R vectors at once (the stack form never runs in passes).
size_t moments_bytes(const Ell &H, int R, int n_left, const DeviceInfo &dev,
                     const Groups *grp = nullptr, int doubling_M = 0) {
    return std::max(pass_layout(H, R, n_left, dev, grp, doubling_M).total,
                    moments_layout_once(H, R, n_left, dev, grp, doubling_M).total);
}

}  // namespace

void set_error(const std::string &msg) { g_err = msg; }
void clear_error() { g_err.clear(); }

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    g_err = std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
            std::to_string(line) + ")";
    return e == hipErrorOutOfMemory ? JKPM_ERR_ALLOC : JKPM_ERR_HIP;
}

int device_info(DeviceInfo *out) {
    static std::mutex m;
    static std::vector<DeviceInfo> cache;
    static std::vector<bool> have;
    int dev = 0;
    JKPM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(m);
    if ((int)cache.size() <= dev) {
        cache.resize(dev + 1);
        have.resize(dev + 1, false);
    }
    if (!have[dev]) {
        DeviceInfo d;
        int cu = 0;
        JKPM_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        d.cu_count = cu > 0 ? cu : 1;
        cache[dev] = d;
        have[dev] = true;
    }
    *out = cache[dev];
    return JKPM_OK;
}

// Runs the whole recursion for one batch; see jkpm_moments.  With grp, the
// handle is a block-diagonal batch and mu is [n_left][G][M][R].  With
// doubling (v_l = v_r, Hermitian A), S = max(1, M/2) steps produce the M
// moments through the raw dots <T_k|T_k>, <T_k|T_{k+1}>.
// stack (plain form only): T_n is written to stack + n * (N*R elements)
// instead of the two rotating buffers, so every T_1 .. T_{M-1} stays (the
// caller puts T_0 in slot 0).
// no_alloc: the caller may not allocate, so no profiling events are taken.
// mu_ld: row length of mu (R when 0), so that one pass of a call can write
// its column slice of the caller's moments.
static int run_moments_once(const Ell &H, const void *vr, const void *vl0, const void *vl1,
                            int R, int M, double2 *mu, void *ws, size_t ws_bytes, hipStream_t s,
                            const Groups *grp = nullptr, bool doubling = false,
                            void *stack = nullptr, bool no_alloc = false, int mu_ld = 0) {
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    if (mu_ld <= 0) mu_ld = R;
    const int n_left = (vl1 || doubling) ? 2 : 1;
    const MomentsLayout L = moments_layout_once(H, R, n_left, dev, grp, doubling ? M : 0);
    if (ws_bytes < L.total) {
        set_error("workspace too small: need " + std::to_string(L.total) + " bytes, got " +
                  std::to_string(ws_bytes));
        return JKPM_ERR_WORKSPACE;
    }
    char *w = (char *)ws;
    void *B[2] = {w + L.b0, w + L.b1};
    double2 *ring = (double2 *)(w + L.ring);
    const size_t slot = (size_t)n_left * L.nchunks * R;

    // Vector groups: powers of two, largest first, so 16-byte lanes stay aligned.
    std::vector<std::pair<int, int>> groups;  // (r0, rb)
    for (int r0 = 0; r0 < R;) {
        int rb = max_group(H.vtype);
        while (rb > R - r0) rb >>= 1;
        groups.push_back({r0, rb});
        r0 += rb;
    }

    // No event pairs while the stream is being captured into a graph.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    JKPM_HIP(hipStreamIsCapturing(s, &cap));
    Prof &P = prof();
    std::unique_lock<std::mutex> plk(P.mu);
    const int every = P.every;
    // (no_alloc: the caller may not allocate, and recording takes events)
    const bool profiling = every > 0 && cap == hipStreamCaptureStatusNone && !no_alloc;
    if (!profiling) plk.unlock();

    // Even steps sweep their chunks backwards: a step then starts on the rows
    // the step before wrote last, while they are still in the 256 MB
    // Infinity Cache (headline honeycomb, R = 16: 201.9 -> 193.1 us per
    // plain step at the best variant, 173.5 -> 170.5 us for the doubling
    // form).  JKPM_SWEEP_ALT=0 sweeps every step forwards.
    static const bool sweep_alt = [] {
        const char *e = getenv("JKPM_SWEEP_ALT");
        return !(e && atoi(e) == 0);
    }();
    auto step = [&](int n) -> int {
        // T_n is written to B[n & 1]; T_{n-1} is B[(n-1) & 1] (or v_r for n=1);
        // T_{n-2} is v_r for n=2, else B[n & 1] itself (in-place update).
        StepArgs a{};
        a.H = &H;
        a.first = (n == 1);
        if (stack) {
            const size_t vb = (size_t)H.n * R * cplx_bytes(H.vtype);
            char *st = (char *)stack;
            a.xcur = (n == 1) ? vr : (const void *)(st + (size_t)(n - 1) * vb);
            a.xprev = (n == 1) ? nullptr : (n == 2 ? vr : (const void *)(st + (size_t)(n - 2) * vb));
            a.xout = st + (size_t)n * vb;
        } else {
            a.xcur = (n == 1) ? vr : B[(n - 1) & 1];
            a.xprev = (n == 1) ? nullptr : (n == 2 ? vr : B[n & 1]);
            a.xout = B[n & 1];
        }
        a.vl0 = doubling ? nullptr : vl0;
        a.vl1 = doubling ? nullptr : vl1;
        if (L.has_ph) {
            a.vph = (const unsigned *)(w + L.ph);
            a.ph_bad = (const int *)(w + L.ph_flag);
        }
        a.doubling = doubling;
        a.ld = R;
        a.part = ring + (size_t)(n % kRing) * slot;
        a.part0 = (n == 1) ? ring : nullptr;  // step 0 lives in ring slot 0
        a.nchunks = L.nchunks;
        a.chunk_rows = grp ? grp->chunk_rows : nullptr;
        a.reverse = sweep_alt && (n % 2 == 0);
        for (auto &g : groups) {
            a.r0 = g.first;
            a.rb = g.second;
            int rc2 = launch_step(a, s);
            if (rc2) return rc2;
        }
        return JKPM_OK;
    };

    // Step 1 also yields the step-0 partials; M == 1 still runs it (T_1 is
    // discarded).  Plain form: steps 1 .. M-1 reduce straight into mu.
    // Doubling form: steps 1 .. S reduce into the raw dots [2][G][S+1][R].
    const int G = grp ? grp->G : 1;
    const int nsteps = doubling ? doubling_steps(M) : std::max(M, 2) - 1;
    const int last_slot = doubling ? nsteps : M - 1;
    double2 *target = doubling ? (double2 *)(w + L.raw) : mu;
    const int target_M = doubling ? nsteps + 1 : M;
    int pending_lo = 0;
    // Profiling: event pairs around runs of up to kProfRun consecutive steps,
    // one run started every `every` steps (never at the first step, whose
    // kernel may differ), each closed before a reduction launch so that no
    // reduction is timed.  A run's span over its step count is the per-launch
    // cost of the recursion, inter-kernel gaps included.
    ProfRun run;
    bool run_open = false;
    if (L.has_ph) {
        rc = launch_phase_pack(vl0, (long)H.n * R, (unsigned *)(w + L.ph), (int *)(w + L.ph_flag),
                               s);
        if (rc) return rc;
    }
    for (int n = 1; n <= nsteps; ++n) {
        if (profiling && !run_open && n > 1 && P.counter % every == 0) {
            run.e0 = take_event(P);
            run.e1 = take_event(P);
            if (!run.e0 || !run.e1)
                return hip_fail(hipErrorOutOfMemory, "hipEventCreate", __FILE__, __LINE__);
            JKPM_HIP(hipEventRecord(run.e0, s));
            run.steps = 0;
            run_open = true;
        }
        if (profiling) ++P.counter;
        rc = step(n);
        if (rc) return rc;
        const int last = std::min(n, last_slot);
        const bool reduce_now = n == nsteps || n - pending_lo + 1 == kRing;
        if (run_open && (++run.steps == kProfRun || reduce_now)) {
            JKPM_HIP(hipEventRecord(run.e1, s));
            P.pending.push_back(run);
            run_open = false;
            if (P.pending.size() >= 8192) {
                rc = drain_profile(P);
                if (rc) return rc;
            }
        }
        if (reduce_now) {
            rc = launch_reduce(ring, kRing, pending_lo, last - pending_lo + 1, n_left, L.nchunks,
                               R, target, target_M, grp ? grp->group_chunks : nullptr, G, s,
                               doubling ? R : mu_ld);
            if (rc) return rc;
            pending_lo = n + 1;
        }
    }
    if (doubling) return launch_doubling_finish(target, G, target_M, R, M, mu, s, mu_ld);
    return JKPM_OK;
}

// The whole recursion for R vectors: at once, or, when the handle sets fewer
// vectors per pass for this form (jkpm_ell_set_pass), one pass per group of
// g vectors.  A pass copies its columns of v_r (and of left sets other than
// v_r) to contiguous [N][g] arrays and writes its column slice of mu.  The
// stack form always runs at once.
static int run_moments(const Ell &H, const void *vr, const void *vl0, const void *vl1, int R,
                       int M, double2 *mu, void *ws, size_t ws_bytes, hipStream_t s,
                       const Groups *grp = nullptr, bool doubling = false,
                       void *stack = nullptr, bool no_alloc = false) {
    if (stack || pass_vectors(H, R, doubling) >= R)
        return run_moments_once(H, vr, vl0, vl1, R, M, mu, ws, ws_bytes, s, grp, doubling, stack,
                                no_alloc);
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int n_left = (vl1 || doubling) ? 2 : 1;
    const PassLayout P = pass_layout(H, R, n_left, dev, grp, doubling ? M : 0);
    if (ws_bytes < P.total) {
        set_error("workspace too small: need " + std::to_string(P.total) + " bytes, got " +
                  std::to_string(ws_bytes));
        return JKPM_ERR_WORKSPACE;
    }
    char *w = (char *)ws;
    const int cb = (int)cplx_bytes(H.vtype);
    void *pv = w, *pl0 = w + P.vec, *pl1 = w + 2 * P.vec;
    for (int p0 = 0; p0 < R; p0 += P.g) {
        const int gc = std::min(P.g, R - p0);
        rc = launch_take_columns(vr, H.n, R, p0, gc, cb, pv, s);
        const void *l0 = pv, *l1 = nullptr;
        if (!rc && !doubling && vl0 != vr) {
            rc = launch_take_columns(vl0, H.n, R, p0, gc, cb, pl0, s);
            l0 = pl0;
        }
        if (!rc && !doubling && vl1) {
            rc = launch_take_columns(vl1, H.n, R, p0, gc, cb, pl1, s);
            l1 = pl1;
        }
        if (!rc)
            rc = run_moments_once(H, pv, l0, l1, gc, M, mu + p0, w + P.inner, ws_bytes - P.inner,
                                  s, grp, doubling, nullptr, no_alloc, R);
        if (rc) return rc;
    }
    return JKPM_OK;
}

// ---- two steps per launch over LDS patches ----

struct TilesLayout {
    size_t vp = 0, buf = 0, ring = 0, raw = 0, total = 0, vec = 0;
};

TilesLayout tiles_layout(const jkpm_tiles &T, int R, int M, bool doubling) {
    TilesLayout L;
    L.vec = align_up((size_t)T.n * R * cplx_bytes(T.vtype), 256);
    L.vp = 0;
    L.buf = L.vec;
    L.ring = L.buf + 4 * L.vec;
    const int n_left = doubling ? 2 : 1;
    L.raw = L.ring + align_up((size_t)kRing * n_left * T.npatch * R * sizeof(double2), 256);
    L.total = L.raw;
    if (doubling) L.total += align_up(2 * (size_t)(doubling_steps(M) + 1) * R * sizeof(double2), 256);
    return L;
}

// Moments of v_r (v_l = v_r) through the tile plan: the vectors are permuted
// into plan order once, then every launch advances the recursion by two
// steps.  Plain form: launch k makes T_{2k+1}, T_{2k+2} and mu of both (and
// mu_0 on the first launch) into ring slots; doubling form: the raw dots
// a_k, b_k of the same steps, mapped to mu_0 .. mu_{M-1} at the end.
int run_tiles(const jkpm_tiles &T, const void *v_r, int R, int M, bool doubling, double2 *mu,
              void *ws, size_t ws_bytes, hipStream_t s) {
    const TilesLayout L = tiles_layout(T, R, M, doubling);
    if (ws_bytes < L.total) {
        set_error("workspace too small: need " + std::to_string(L.total) + " bytes, got " +
                  std::to_string(ws_bytes));
        return JKPM_ERR_WORKSPACE;
    }
    char *w = (char *)ws;
    void *vp = w + L.vp;
    void *B[4] = {w + L.buf, w + L.buf + L.vec, w + L.buf + 2 * L.vec, w + L.buf + 3 * L.vec};
    double2 *ring = (double2 *)(w + L.ring);
    const int n_left = doubling ? 2 : 1;
    const size_t slot = (size_t)n_left * T.npatch * R;
    int rc = launch_vec_permute(T.vtype, v_r, T.perm, T.n, R, vp, s);
    if (rc) return rc;

    std::vector<std::pair<int, int>> groups;  // (r0, rb), powers of two, largest first
    for (int r0 = 0; r0 < R;) {
        int rb = tile_max_group();
        while (rb > R - r0) rb >>= 1;
        groups.push_back({r0, rb});
        r0 += rb;
    }

    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    JKPM_HIP(hipStreamIsCapturing(s, &cap));
    Prof &P = prof();
    std::unique_lock<std::mutex> plk(P.mu);
    const int every = P.every;
    const bool profiling = every > 0 && cap == hipStreamCaptureStatusNone;
    if (!profiling) plk.unlock();

    // Slots needed: plain mu_0 .. mu_{M-1}; doubling raw slots 0 .. S.
    const int S = doubling_steps(M);
    const int last_slot = doubling ? S : M - 1;
    const int nlaunch = std::max(1, (last_slot + 1) / 2);  // launch k covers slots 2k+1, 2k+2
    double2 *target = doubling ? (double2 *)(w + L.raw) : mu;
    const int target_M = doubling ? S + 1 : M;
    int pending_lo = 0;
    for (int kq = 0; kq < nlaunch; ++kq) {
        const int s1 = 2 * kq + 1, s2 = 2 * kq + 2;
        TileArgs a;
        a.T = &T;
        a.first = kq == 0;
        a.doubling = doubling;
        a.xprev = kq == 0 ? nullptr : B[2 * ((kq - 1) % 2)];
        a.xcur = kq == 0 ? vp : B[2 * ((kq - 1) % 2) + 1];
        a.out1 = B[2 * (kq % 2)];
        a.out2 = B[2 * (kq % 2) + 1];
        a.vl = doubling ? nullptr : vp;
        a.ld = R;
        a.a1 = kq == 0 ? 1.0 : 2.0;
        a.b1 = kq == 0 ? 0.0 : -1.0;
        a.part0 = kq == 0 ? ring : nullptr;
        a.part1 = ring + (size_t)(s1 % kRing) * slot;
        a.part2 = ring + (size_t)(s2 % kRing) * slot;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool rec = profiling && (P.counter++ % every == 0);
        if (rec) {
            e0 = take_event(P);
            e1 = take_event(P);
            if (!e0 || !e1) return hip_fail(hipErrorOutOfMemory, "hipEventCreate", __FILE__, __LINE__);
            JKPM_HIP(hipEventRecord(e0, s));
        }
        for (auto &g : groups) {
            a.r0 = g.first;
            a.rb = g.second;
            rc = launch_tile2(a, s);
            if (rc) return rc;
        }
        if (rec) {
            JKPM_HIP(hipEventRecord(e1, s));
            P.pending.push_back({e0, e1, 1, JKPM_PROF_STEP});
            if (P.pending.size() >= 8192) {
                rc = drain_profile(P);
                if (rc) return rc;
            }
        }
        // Fold the ring before the next launch could overwrite a pending slot.
        const int last = std::min(s2, last_slot);
        const bool final_launch = kq == nlaunch - 1;
        if (final_launch || s2 + 2 - pending_lo + 1 > kRing) {
            rc = launch_reduce(ring, kRing, pending_lo, last - pending_lo + 1, n_left, T.npatch,
                               R, target, target_M, nullptr, 1, s);
            if (rc) return rc;
            pending_lo = s2 + 1;
        }
    }
    if (doubling) return launch_doubling_finish(target, 1, target_M, R, M, mu, s);
    return JKPM_OK;
}

// ---- two steps per launch by a wavefront over row chunks ----

struct WaveLayout {
    size_t vec = 0, buf = 0, ring = 0, raw = 0, total = 0;
};

WaveLayout wave_layout(const jkpm_wave &W, int M) {
    WaveLayout L;
    const int n_left = W.doubling ? 2 : 1;
    L.vec = align_up((size_t)W.h->e.n * W.R * cplx_bytes(JKPM_C64), 256);
    L.buf = 0;
    L.ring = 4 * L.vec;
    L.raw = L.ring + align_up((size_t)kRing * n_left * W.nchunks * W.R * sizeof(double2), 256);
    L.total = L.raw;
    if (W.doubling) L.total += align_up(2 * (size_t)(doubling_steps(M) + 1) * W.R * sizeof(double2), 256);
    return L;
}

// Step 1 (and a last odd step) on the one-step kernel; every other pair of
// steps in one wavefront launch.  Buffers: T_1 goes to X0; each pair reads
// (prev, cur) and writes the two buffers of X0..X3 that are neither, so no
// buffer a launch reads is written by it.  The ring and reductions are
// those of run_moments.
int run_wave(jkpm_wave &W, const void *vr, int M, double2 *mu, void *ws, size_t ws_bytes,
             hipStream_t s) {
    const WaveLayout L = wave_layout(W, M);
    if (ws_bytes < L.total) {
        set_error("workspace too small: need " + std::to_string(L.total) + " bytes, got " +
                  std::to_string(ws_bytes));
        return JKPM_ERR_WORKSPACE;
    }
    const Ell &H = W.h->e;
    const int R = W.R;
    const bool doubling = W.doubling;
    const int n_left = doubling ? 2 : 1;
    char *w = (char *)ws;
    void *X[4] = {w, w + L.vec, w + 2 * L.vec, w + 3 * L.vec};
    double2 *ring = (double2 *)(w + L.ring);
    const size_t slot = (size_t)n_left * W.nchunks * R;

    Prof &P = prof();
    std::unique_lock<std::mutex> plk(P.mu);
    const int every = P.every;
    const bool profiling = every > 0;
    if (!profiling) plk.unlock();

    auto one_step = [&](int n, const void *prev, const void *cur, void *out) -> int {
        StepArgs a{};
        a.H = &H;
        a.first = n == 1;
        a.xcur = cur;
        a.xprev = prev;
        a.xout = out;
        a.vl0 = doubling ? nullptr : vr;
        a.vl1 = nullptr;
        a.doubling = doubling;
        a.ld = R;
        a.r0 = 0;
        a.rb = R;
        a.part = ring + (size_t)(n % kRing) * slot;
        a.part0 = n == 1 ? ring : nullptr;
        a.nchunks = W.nchunks;
        return launch_step(a, s);
    };

    const int nsteps = doubling ? doubling_steps(M) : std::max(M, 2) - 1;
    const int last_slot = doubling ? nsteps : M - 1;
    double2 *target = doubling ? (double2 *)(w + L.raw) : mu;
    const int target_M = doubling ? nsteps + 1 : M;
    int pending_lo = 0;
    auto reduce_to = [&](int n) -> int {
        const int last = std::min(n, last_slot);
        if (last < pending_lo) return JKPM_OK;
        int rc2 = launch_reduce(ring, kRing, pending_lo, last - pending_lo + 1, n_left,
                                W.nchunks, R, target, target_M, nullptr, 1, s);
        pending_lo = n + 1;
        return rc2;
    };

    int rc = one_step(1, nullptr, vr, X[0]);
    if (rc) return rc;
    const void *prev = vr, *cur = X[0];
    int n = 1;
    while (n + 2 <= nsteps) {
        if (n + 2 - pending_lo + 1 > kRing && (rc = reduce_to(n))) return rc;
        void *o[2];
        int no = 0;
        for (int i = 0; i < 4 && no < 2; ++i)
            if (X[i] != prev && X[i] != cur) o[no++] = X[i];
        if (W.epoch + 1 >= (1u << 19)) {  // flags are 1 << 31 | epoch << 12 | progress
            rc = launch_zero_bytes(W.flags, (size_t)W.nchunks * 4 * sizeof(unsigned), s);
            if (rc) return rc;
            W.epoch = 0;
        }
        ++W.epoch;
        WaveArgs a;
        a.H = &H;
        a.xprev = prev;
        a.xcur = cur;
        a.out1 = o[0];
        a.out2 = o[1];
        a.vl = doubling ? nullptr : vr;
        a.rb = W.rb;
        a.d = W.d;
        a.doubling = doubling;
        a.part1 = ring + (size_t)((n + 1) % kRing) * slot;
        a.part2 = ring + (size_t)((n + 2) % kRing) * slot;
        a.nchunks = W.nchunks;
        a.rpc = W.rpc;
        a.nblk = W.nblk;
        a.deps = W.deps;
        a.flags = W.flags;
        a.err = W.err;
        a.epoch = W.epoch;
        a.tail_off = W.tail_off;
        a.tail_rows = W.tail_rows;
        a.tail_dep = W.tail_dep;
        a.nfix = W.nfix;
        a.nfold = W.nfold;
        a.fix_group_off = W.fix_group_off;
        a.fix_rows = W.fix_rows;
        a.fold_chunk = W.fold_chunk;
        a.fold_off = W.fold_off;
        a.fixpart = W.fixpart;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool rec = profiling && (P.counter++ % every == 0);
        if (rec) {
            e0 = take_event(P);
            e1 = take_event(P);
            if (!e0 || !e1) return hip_fail(hipErrorOutOfMemory, "hipEventCreate", __FILE__, __LINE__);
            JKPM_HIP(hipEventRecord(e0, s));
        }
        rc = launch_wave2(a, s);
        if (rc) return rc;
        if (rec) {
            JKPM_HIP(hipEventRecord(e1, s));
            P.pending.push_back({e0, e1, 1, JKPM_PROF_WAVE});
            if (P.pending.size() >= 8192 && (rc = drain_profile(P))) return rc;
        }
        prev = o[0];
        cur = o[1];
        n += 2;
    }
    if (n < nsteps) {
        if (n + 1 - pending_lo + 1 > kRing && (rc = reduce_to(n))) return rc;
        void *out = nullptr;
        if (prev != vr) {
            out = const_cast<void *>(prev);  // in place over T_{n-1}, as run_moments does
        } else {
            for (int i = 0; i < 4 && !out; ++i)
                if (X[i] != cur) out = X[i];
        }
        if ((rc = one_step(n + 1, prev, cur, out))) return rc;
        ++n;
    }
    if ((rc = reduce_to(n))) return rc;
    if (doubling && (rc = launch_doubling_finish(target, 1, target_M, R, M, mu, s))) return rc;
    // The kernel's error bits: a work-group off its XCD, or a wait that gave up.
    JKPM_HIP(hipStreamSynchronize(s));
    unsigned bits = 0;
    JKPM_HIP(hipMemcpy(&bits, W.err, sizeof(bits), hipMemcpyDeviceToHost));
    if (bits) {
        JKPM_HIP(hipMemset(W.err, 0, sizeof(unsigned)));
        set_error(std::string("wavefront launch invalid:") +
                  ((bits & 1u) ? " a work-group ran on another XCD than its chunk's" : "") +
                  ((bits & 2u) ? " a readiness wait gave up" : "") + "; the moments are wrong");
        return JKPM_ERR_HIP;
    }
    return JKPM_OK;
}

static int validate_vtype(int vtype) {
    if (vtype != JKPM_C64 && vtype != JKPM_C128)
        return arg_error("value_type must be JKPM_C128 (0) or JKPM_C64 (1)");
    return JKPM_OK;
}

static int upload_values(Ell *H, const double *data, double scale, const int *idx_dev,
                         int src_on_device, hipStream_t s) {
    const size_t tot = (size_t)H->n * H->k;
    H->scale = scale;
    const double *src = data;
    double *staging = nullptr;
    if (!src_on_device) {
        JKPM_HIP(hipMalloc(&staging, tot * 16));
        JKPM_HIP(hipMemcpyAsync(staging, data, tot * 16, hipMemcpyHostToDevice, s));
        src = staging;
    }
    int rc = launch_ell_pack(src, scale, H, idx_dev, s);
    if (staging) {
        hipError_t e = hipStreamSynchronize(s);
        hipFree(staging);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    return rc;
}

}  // namespace jkpm

// Scratch of jkpm_moments_noalloc: the two rotating vector batches and the
// moment-partial ring of run_moments, in device memory the library owns
// instead of a caller workspace.  Each GPU holds its own copy from module
// load until the process ends.  In host code `scratch` only names the
// variable; its device address on the current GPU comes from
// hipGetSymbolAddress (noalloc_scratch), taken once per device.
static __device__ __align__(256) unsigned char scratch[JKPM_NOALLOC_SCRATCH_BYTES];

namespace jkpm {
namespace {

// Device address of `scratch` on the current device, looked up on the first
// call per device and kept.
int noalloc_scratch(void **out) {
    static std::mutex m;
    static std::vector<void *> ptrs;
    int dev = 0;
    JKPM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(m);
    if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
    if (!ptrs[dev]) {
        void *p = nullptr;
        JKPM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned no address for the no-alloc scratch");
            return JKPM_ERR_HIP;
        }
        ptrs[dev] = p;
    }
    *out = ptrs[dev];
    return JKPM_OK;
}

}  // namespace
}  // namespace jkpm

using namespace jkpm;

extern "C" {

int jkpm_version(void) { return 100; }

const char *jkpm_last_error(void) { return g_err.c_str(); }

int jkpm_ell_create(const double *data, const int *indices, int n_rows, int len_row,
                    double scale, int value_type, int src_on_device, void *stream,
                    jkpm_ell **out) {
    clear_error();
    if (!out) return arg_error("out is NULL");
    *out = nullptr;
    if (!data || !indices) return arg_error("data/indices is NULL");
    if (n_rows <= 0 || len_row <= 0) return arg_error("n_rows and len_row must be positive");
    if ((long long)n_rows * len_row > (1LL << 31) - 1)
        return arg_error("n_rows*len_row exceeds 2^31-1 slots");
    int rc = validate_vtype(value_type);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t tot = (size_t)n_rows * len_row;

    if (!src_on_device) {
        for (size_t t = 0; t < tot; ++t)
            if (indices[t] < 0 || indices[t] >= n_rows)
                return arg_error("column index " + std::to_string(indices[t]) + " at slot " +
                                 std::to_string(t) + " is outside [0, n_rows)");
    }

    if (inject_alloc_failure()) {
        set_error("device allocation failed (injected by JKPM_FAULT=alloc)");
        return JKPM_ERR_ALLOC;
    }
    auto *h = new (std::nothrow) jkpm_ell;
    if (!h) return arg_error("host allocation failed");
    Ell &H = h->e;
    H.n = n_rows;
    H.k = len_row;
    H.vtype = value_type;
    hipGetDevice(&H.device);
    int *idx_dev = nullptr;
    unsigned int *bad = nullptr;
    auto fail = [&](int code) {
        if (idx_dev && !src_on_device) hipFree(idx_dev);
        if (bad) hipFree(bad);
        if (H.val) hipFree(H.val);
        if (H.rval) hipFree(H.rval);
        if (H.col) hipFree(H.col);
        delete h;
        return code;
    };
    hipError_t e;
    if ((e = hipMalloc(&H.val, tot * cplx_bytes(value_type))) != hipSuccess)
        return fail(hip_fail(e, "hipMalloc(values)", __FILE__, __LINE__));
    if ((e = hipMalloc(&H.col, tot * sizeof(int))) != hipSuccess)
        return fail(hip_fail(e, "hipMalloc(cols)", __FILE__, __LINE__));
    if (src_on_device) {
        idx_dev = const_cast<int *>(indices);
        if ((e = hipMalloc(&bad, sizeof(unsigned int))) != hipSuccess)
            return fail(hip_fail(e, "hipMalloc(flag)", __FILE__, __LINE__));
        if ((e = hipMemsetAsync(bad, 0, sizeof(unsigned int), s)) != hipSuccess)
            return fail(hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__));
        const int grid = (int)((tot + 255) / 256);
        hipLaunchKernelGGL(count_bad_indices, dim3(grid), dim3(256), 0, s, indices, (long)tot,
                           n_rows, bad);
        unsigned int nbad = 0;
        if ((e = hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return fail(hip_fail(e, "index validation", __FILE__, __LINE__));
        if (nbad) {
            set_error(std::to_string(nbad) + " column indices are outside [0, n_rows)");
            return fail(JKPM_ERR_ARG);
        }
    } else {
        if ((e = hipMalloc(&idx_dev, tot * sizeof(int))) != hipSuccess)
            return fail(hip_fail(e, "hipMalloc(index staging)", __FILE__, __LINE__));
        if ((e = hipMemcpyAsync(idx_dev, indices, tot * sizeof(int), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return fail(hip_fail(e, "hipMemcpyAsync(indices)", __FILE__, __LINE__));
    }
    rc = upload_values(&H, data, scale, idx_dev, src_on_device, s);
    if (rc) return fail(rc);
    if ((rc = update_real_values(&H, s))) return fail(rc);
    h->values_real = H.real_values;
    h->values_imag = H.imag_values;
    if ((e = hipStreamSynchronize(s)) != hipSuccess)
        return fail(hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__));
    if (!src_on_device) hipFree(idx_dev);
    idx_dev = nullptr;
    if (bad) hipFree(bad);
    h->col_owner = std::shared_ptr<int>(H.col, [](int *p) { hipFree(p); });
    *out = h;
    return JKPM_OK;
}

int jkpm_ell_create_csr(const int *indptr, const int *indices, const double *data, int n_rows,
                        int len_row, double scale, int value_type, int src_on_device,
                        void *stream, jkpm_ell **out) {
    clear_error();
    if (!out) return arg_error("out is NULL");
    *out = nullptr;
    if (!indptr) return arg_error("indptr is NULL");
    if (n_rows <= 0) return arg_error("n_rows must be positive");
    if (len_row < 0) return arg_error("len_row must be >= 0 (0: the longest row)");
    int rc = validate_vtype(value_type);
    if (rc) return rc;
    if (inject_alloc_failure()) {
        set_error("device allocation failed (injected by JKPM_FAULT=alloc)");
        return JKPM_ERR_ALLOC;
    }
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    // nnz = indptr[n_rows]
    int nnz_i = 0;
    if (src_on_device) {
        if ((e = hipMemcpyAsync(&nnz_i, indptr + n_rows, sizeof(int), hipMemcpyDeviceToHost, s)) !=
                hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return hip_fail(e, "read of indptr[n_rows]", __FILE__, __LINE__);
    } else {
        nnz_i = indptr[n_rows];
    }
    if (nnz_i < 0) return arg_error("indptr[n_rows] (nnz) is negative");
    const long nnz = nnz_i;
    if (nnz > 0 && (!indices || !data)) return arg_error("indices/data is NULL");

    // Device copies of the CSR arrays (staged when they live on the host).
    int *d_ptr = nullptr, *d_idx = nullptr;
    double *d_val = nullptr;
    unsigned *stats = nullptr;
    jkpm_ell *h = nullptr;
    auto cleanup = [&](int code) {
        if (!src_on_device) {
            if (d_ptr) hipFree(d_ptr);
            if (d_idx) hipFree(d_idx);
            if (d_val) hipFree(d_val);
        }
        if (stats) hipFree(stats);
        if (h) {
            if (h->e.val) hipFree(h->e.val);
            if (h->e.rval) hipFree(h->e.rval);
            if (h->e.col) hipFree(h->e.col);
            delete h;
        }
        return code;
    };
    if (src_on_device) {
        d_ptr = const_cast<int *>(indptr);
        d_idx = const_cast<int *>(indices);
        d_val = const_cast<double *>(data);
    } else {
        const size_t pb = (size_t)(n_rows + 1) * sizeof(int);
        const size_t ib = (size_t)(nnz ? nnz : 1) * sizeof(int);
        const size_t vb = (size_t)(nnz ? nnz : 1) * 16;
        if ((e = hipMalloc(&d_ptr, pb)) != hipSuccess || (e = hipMalloc(&d_idx, ib)) != hipSuccess ||
            (e = hipMalloc(&d_val, vb)) != hipSuccess)
            return cleanup(hip_fail(e, "hipMalloc(CSR staging)", __FILE__, __LINE__));
        if ((e = hipMemcpyAsync(d_ptr, indptr, pb, hipMemcpyHostToDevice, s)) != hipSuccess ||
            (nnz && (e = hipMemcpyAsync(d_idx, indices, (size_t)nnz * sizeof(int),
                                        hipMemcpyHostToDevice, s)) != hipSuccess) ||
            (nnz && (e = hipMemcpyAsync(d_val, data, (size_t)nnz * 16, hipMemcpyHostToDevice, s)) !=
                        hipSuccess))
            return cleanup(hip_fail(e, "hipMemcpyAsync(CSR)", __FILE__, __LINE__));
    }
    if ((e = hipMalloc(&stats, 3 * sizeof(unsigned))) != hipSuccess ||
        (e = hipMemsetAsync(stats, 0, 3 * sizeof(unsigned), s)) != hipSuccess)
        return cleanup(hip_fail(e, "scan statistics", __FILE__, __LINE__));
    rc = launch_csr_scan(d_ptr, d_idx, n_rows, nnz, stats, s);
    if (rc) return cleanup(rc);
    unsigned st[3] = {0, 0, 0};
    if ((e = hipMemcpyAsync(st, stats, sizeof(st), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return cleanup(hip_fail(e, "CSR scan", __FILE__, __LINE__));
    if (st[1]) {
        set_error(std::to_string(st[1]) +
                  " rows have an indptr range that decreases or leaves [0, nnz]");
        return cleanup(JKPM_ERR_ARG);
    }
    if (st[2]) {
        set_error(std::to_string(st[2]) + " column indices are outside [0, n_rows)");
        return cleanup(JKPM_ERR_ARG);
    }
    const int longest = (int)st[0];
    int K = len_row;
    if (K == 0) K = longest > 0 ? longest : 1;
    if (longest > K) {
        set_error("a row has " + std::to_string(longest) + " entries, more than len_row = " +
                  std::to_string(K));
        return cleanup(JKPM_ERR_ARG);
    }
    if ((long long)n_rows * K > (1LL << 31) - 1)
        return cleanup(arg_error("n_rows*len_row exceeds 2^31-1 slots"));

    h = new (std::nothrow) jkpm_ell;
    if (!h) return cleanup(arg_error("host allocation failed"));
    Ell &H = h->e;
    H.n = n_rows;
    H.k = K;
    H.vtype = value_type;
    H.scale = scale;
    hipGetDevice(&H.device);
    const size_t tot = (size_t)n_rows * K;
    if ((e = hipMalloc(&H.val, tot * cplx_bytes(value_type))) != hipSuccess)
        return cleanup(hip_fail(e, "hipMalloc(values)", __FILE__, __LINE__));
    if ((e = hipMalloc(&H.col, tot * sizeof(int))) != hipSuccess)
        return cleanup(hip_fail(e, "hipMalloc(cols)", __FILE__, __LINE__));
    rc = launch_csr_to_ell(d_ptr, d_idx, d_val, scale, &H, s);
    if (rc) return cleanup(rc);
    if ((rc = update_real_values(&H, s))) return cleanup(rc);
    h->values_real = H.real_values;
    h->values_imag = H.imag_values;
    if ((e = hipStreamSynchronize(s)) != hipSuccess)
        return cleanup(hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__));
    jkpm_ell *done = h;
    h = nullptr;
    cleanup(JKPM_OK);
    done->col_owner = std::shared_ptr<int>(done->e.col, [](int *p) { hipFree(p); });
    *out = done;
    return JKPM_OK;
}

int jkpm_ell_get_indices(const jkpm_ell *h, int *indices, int dst_on_device, void *stream) {
    clear_error();
    if (!h || !indices) return arg_error("handle/indices is NULL");
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)h->e.n * h->e.k * sizeof(int);
    if (dst_on_device) {
        int rc = launch_ell_cols_unpack(h->e, indices, s);
        if (rc) return rc;
        JKPM_HIP(hipStreamSynchronize(s));
        return JKPM_OK;
    }
    int *tmp = nullptr;
    JKPM_HIP(hipMalloc(&tmp, bytes ? bytes : 4));
    int rc = launch_ell_cols_unpack(h->e, tmp, s);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(indices, tmp, bytes, hipMemcpyDeviceToHost, s);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(tmp);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "copy of indices to the host", __FILE__, __LINE__);
    return JKPM_OK;
}

int jkpm_ell_create_like(const jkpm_ell *pattern, const double *data, double scale,
                         int value_type, int src_on_device, void *stream, jkpm_ell **out) {
    clear_error();
    if (!out) return arg_error("out is NULL");
    *out = nullptr;
    if (!pattern || !data) return arg_error("pattern/data is NULL");
    int rc = validate_vtype(value_type);
    if (rc) return rc;
    if (inject_alloc_failure()) {
        set_error("device allocation failed (injected by JKPM_FAULT=alloc)");
        return JKPM_ERR_ALLOC;
    }
    auto *h = new (std::nothrow) jkpm_ell;
    if (!h) return arg_error("host allocation failed");
    Ell &H = h->e;
    H.n = pattern->e.n;
    H.k = pattern->e.k;
    H.vtype = value_type;
    H.device = pattern->e.device;
    H.col = pattern->e.col;
    h->col_owner = pattern->col_owner;
    const size_t tot = (size_t)H.n * H.k;
    hipError_t e = hipMalloc(&H.val, tot * cplx_bytes(value_type));
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipMalloc(values)", __FILE__, __LINE__);
    }
    rc = upload_values(&H, data, scale, nullptr, src_on_device, (hipStream_t)stream);
    if (!rc) rc = update_real_values(&H, (hipStream_t)stream);
    h->values_real = H.real_values;
    h->values_imag = H.imag_values;
    if (!rc) {
        e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    if (rc) {
        hipFree(H.val);
        if (H.rval) hipFree(H.rval);
        delete h;
        return rc;
    }
    *out = h;
    return JKPM_OK;
}

int jkpm_ell_destroy(jkpm_ell *h) {
    clear_error();
    if (!h) return JKPM_OK;
    hipFree(h->e.val);
    if (h->e.rval) hipFree(h->e.rval);
    delete h;  // releases this handle's share of the column array
    return JKPM_OK;
}

int jkpm_ell_info(const jkpm_ell *h, int *n_rows, int *len_row, int *value_type) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (n_rows) *n_rows = h->e.n;
    if (len_row) *len_row = h->e.k;
    if (value_type) *value_type = h->e.vtype;
    return JKPM_OK;
}

int jkpm_ell_set_values(jkpm_ell *h, const double *data, double scale, int src_on_device,
                        void *stream) {
    clear_error();
    if (!h || !data) return arg_error("handle/data is NULL");
    // Values only: the slot-major column array stays as uploaded.
    h->herm_err = h->herm_max = -1.0;
    h->e.real_values = h->values_real = false;
    h->e.imag_values = h->values_imag = false;
    int rc = upload_values(&h->e, data, scale, nullptr, src_on_device, (hipStream_t)stream);
    if (rc) return rc;
    if ((rc = update_real_values(&h->e, (hipStream_t)stream))) return rc;
    h->values_real = h->e.real_values;
    h->values_imag = h->e.imag_values;
    h->e.real_values = h->values_real && h->use_real;
    h->e.imag_values = h->values_imag && h->use_real;
    JKPM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return JKPM_OK;
}

int jkpm_ell_is_real(const jkpm_ell *h, int *is_real) {
    clear_error();
    if (!h || !is_real) return arg_error("NULL argument");
    *is_real = h->values_real ? 1 : 0;
    return JKPM_OK;
}

int jkpm_ell_is_imag(const jkpm_ell *h, int *is_imag) {
    clear_error();
    if (!h || !is_imag) return arg_error("NULL argument");
    *is_imag = h->values_imag ? 1 : 0;
    return JKPM_OK;
}

int jkpm_ell_use_real_values(jkpm_ell *h, int on) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    h->use_real = on != 0;
    h->e.real_values = h->values_real && h->use_real && h->e.rval;
    h->e.imag_values = h->values_imag && h->use_real && h->e.rval;
    return JKPM_OK;
}

int jkpm_ell_use_left_phases(jkpm_ell *h, int on) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    h->e.left_phases = on != 0;
    return JKPM_OK;
}

int jkpm_phase_pack(const void *v, long long n, unsigned *phases, int *bad, void *stream) {
    clear_error();
    if (!bad || (n > 0 && (!v || !phases))) return arg_error("NULL argument");
    if (n < 0) return arg_error("n must be >= 0");
    if (!aligned16(v) || !aligned16(phases)) return arg_error("device pointers must be 16-byte aligned");
    return launch_phase_pack(v, (long)n, phases, bad, (hipStream_t)stream);
}

int jkpm_check_finite(const double *x, long long n, long long *first_bad, long long *count,
                      void *stream) {
    clear_error();
    if (!first_bad || !count || (n > 0 && !x)) return arg_error("NULL argument");
    if (n < 0) return arg_error("n must be >= 0");
    return count_nonfinite(x, (long)n, first_bad, count, (hipStream_t)stream);
}

int jkpm_ell_hermiticity(const jkpm_ell *h, double *abs_err, double *max_abs, void *stream) {
    clear_error();
    if (!h || !abs_err || !max_abs) return arg_error("NULL argument");
    return hermiticity(h, (hipStream_t)stream, abs_err, max_abs);
}

int jkpm_ell_get_values(const jkpm_ell *h, double *data, int dst_on_device, void *stream) {
    clear_error();
    if (!h || !data) return arg_error("handle/data is NULL");
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)h->e.n * h->e.k * 2 * sizeof(double);
    if (dst_on_device) {
        int rc = launch_ell_unpack(h->e, data, s);
        if (rc) return rc;
        JKPM_HIP(hipStreamSynchronize(s));
        return JKPM_OK;
    }
    double *tmp = nullptr;
    JKPM_HIP(hipMalloc(&tmp, bytes ? bytes : 16));
    int rc = launch_ell_unpack(h->e, tmp, s);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(data, tmp, bytes, hipMemcpyDeviceToHost, s);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(tmp);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "copy of values to the host", __FILE__, __LINE__);
    return JKPM_OK;
}

int jkpm_ell_peierls(jkpm_ell *dst, const jkpm_ell *src, const jkpm_ell *bonds, double a_re,
                     double a_im, double ax_re, double ax_im, double ay_re, double ay_im,
                     void *stream) {
    clear_error();
    if (!dst || !src || !bonds) return arg_error("NULL handle");
    if (src->e.col != dst->e.col || src->e.col != bonds->e.col)
        return arg_error("dst, src and bonds must share one slot pattern (jkpm_ell_create_like)");
    if (dst->e.vtype != src->e.vtype) return arg_error("dst and src must have one value type");
    dst->herm_err = dst->herm_max = -1.0;
    // Peierls phases make the values complex (see poison_real_copy)
    if (dst->values_real || dst->values_imag) {
        int rc0 = poison_real_copy(&dst->e, (hipStream_t)stream);
        if (rc0) return rc0;
    }
    dst->e.real_values = dst->values_real = false;
    dst->e.imag_values = dst->values_imag = false;
    int rc = launch_peierls(src->e, bonds->e, &dst->e, make_double2(a_re, a_im),
                            make_double2(ax_re, ax_im), make_double2(ay_re, ay_im),
                            (hipStream_t)stream);
    if (rc) return rc;
    dst->e.scale = src->e.scale;
    return JKPM_OK;
}

int jkpm_aAxby(const jkpm_ell *h, const double *a, const double *b, const void *x, void *y,
               int R, void *stream) {
    clear_error();
    if (!h || !a || !b || !x || !y) return arg_error("NULL argument");
    if (R <= 0) return arg_error("R must be positive");
    if (x == y) return arg_error("x must not alias y");
    return launch_aaxby(h->e, make_double2(a[0], a[1]), make_double2(b[0], b[1]), x, y, R,
                        (hipStream_t)stream);
}

int jkpm_zvec(int op, long long n, const double *a, const double *b, const void *x, void *y,
              const void *w, void *z, void *stream) {
    clear_error();
    if (n < 0) return arg_error("n must be >= 0");
    if (!a) return arg_error("a is NULL");
    const bool need_x = true, need_y = true;
    const bool need_z = op == JKPM_VOP_MUL || op == JKPM_VOP_MUL3;
    const bool need_w = op == JKPM_VOP_MUL3;
    if (op < JKPM_VOP_AXPBY || op > JKPM_VOP_MUL3) return arg_error("unknown vector op");
    if (op == JKPM_VOP_AXPBY && !b) return arg_error("b is NULL");
    if (n == 0) return JKPM_OK;
    if ((need_x && !x) || (need_y && !y) || (need_z && !z) || (need_w && !w))
        return arg_error("NULL vector");
    if ((x && !aligned16(x)) || (y && !aligned16(y)) || (z && !aligned16(z)) ||
        (w && !aligned16(w)))
        return arg_error("device pointers must be 16-byte aligned");
    const double2 av = make_double2(a[0], a[1]);
    const double2 bv = b ? make_double2(b[0], b[1]) : make_double2(0.0, 0.0);
    return zvec(op, (long)n, av, bv, x, y, w, z, (hipStream_t)stream);
}

int jkpm_zdot_workspace(size_t *bytes) {
    clear_error();
    if (!bytes) return arg_error("bytes is NULL");
    *bytes = zdot_workspace();
    return JKPM_OK;
}

int jkpm_zdot(long long n, const double *a, const void *x, const void *y, int conj_x,
              double *out, void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    if (n < 0) return arg_error("n must be >= 0");
    if (!a || !out || !workspace || (n > 0 && (!x || !y))) return arg_error("NULL argument");
    if (!aligned16(out) || !aligned16(workspace) || (x && !aligned16(x)) || (y && !aligned16(y)))
        return arg_error("device pointers must be 16-byte aligned");
    if (workspace_bytes < zdot_workspace()) {
        set_error("workspace too small: need " + std::to_string(zdot_workspace()) + " bytes");
        return JKPM_ERR_WORKSPACE;
    }
    return zdot((long)n, make_double2(a[0], a[1]), x, y, conj_x ? 1 : 0, (double2 *)out,
                (double2 *)workspace, (hipStream_t)stream);
}

int jkpm_moments_workspace(const jkpm_ell *h, int R, int n_left, size_t *bytes) {
    clear_error();
    if (!h || !bytes) return arg_error("NULL argument");
    if (R <= 0) return arg_error("R must be positive");
    if (n_left != 1 && n_left != 2) return arg_error("n_left must be 1 or 2");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    *bytes = moments_bytes(h->e, R, n_left, dev);
    return JKPM_OK;
}

int jkpm_moments(const jkpm_ell *h, const void *v_r, const void *v_l0, const void *v_l1, int R,
                 int M, double *mu, void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (!v_r || !v_l0 || !mu || !workspace) return arg_error("NULL vector/mu/workspace");
    if (R <= 0) return arg_error("R must be positive");
    if (M <= 0) return arg_error("M must be positive");
    if (!aligned16(v_r) || !aligned16(v_l0) || (v_l1 && !aligned16(v_l1)) ||
        !aligned16(workspace) || !aligned16(mu))
        return arg_error("device pointers must be 16-byte aligned");
    return run_moments(h->e, v_r, v_l0, v_l1, R, M, (double2 *)mu, workspace, workspace_bytes,
                       (hipStream_t)stream);
}

int jkpm_moments_noalloc_setup(void) {
    clear_error();
    void *p = nullptr;
    return noalloc_scratch(&p);
}

int jkpm_moments_noalloc(const jkpm_ell *h, const void *v_r, const void *v_l0, const void *v_l1,
                         int R, int M, double *mu, void *stream) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (!v_r || !v_l0 || !mu) return arg_error("NULL vector/mu");
    if (R <= 0) return arg_error("R must be positive");
    if (M <= 0) return arg_error("M must be positive");
    if (!aligned16(v_r) || !aligned16(v_l0) || (v_l1 && !aligned16(v_l1)) || !aligned16(mu))
        return arg_error("device pointers must be 16-byte aligned");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const size_t need = moments_bytes(h->e, R, v_l1 ? 2 : 1, dev);
    if (need > JKPM_NOALLOC_SCRATCH_BYTES) {
        set_error("jkpm_moments_noalloc: this matrix and R need " + std::to_string(need) +
                  " bytes of scratch, more than the " +
                  std::to_string((size_t)JKPM_NOALLOC_SCRATCH_BYTES) +
                  " the library holds per GPU; use jkpm_moments with a workspace");
        return JKPM_ERR_WORKSPACE;
    }
    void *scr = nullptr;
    rc = noalloc_scratch(&scr);
    if (rc) return rc;
    return run_moments(h->e, v_r, v_l0, v_l1, R, M, (double2 *)mu, scr,
                       (size_t)JKPM_NOALLOC_SCRATCH_BYTES, (hipStream_t)stream, nullptr, false,
                       nullptr, true);
}

int jkpm_moments_keep(const jkpm_ell *h, const void *v_r, int R, int M, double *mu,
                      void *stack, size_t stack_bytes, void *workspace, size_t workspace_bytes,
                      void *stream) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (!v_r || !mu || !stack || !workspace) return arg_error("NULL vector/mu/stack/workspace");
    if (R <= 0) return arg_error("R must be positive");
    if (M <= 0) return arg_error("M must be positive");
    if (!aligned16(v_r) || !aligned16(stack) || !aligned16(workspace) || !aligned16(mu))
        return arg_error("device pointers must be 16-byte aligned");
    const size_t vb = (size_t)h->e.n * R * cplx_bytes(h->e.vtype);
    const size_t need = (size_t)std::max(M, 2) * vb;
    if (stack_bytes < need)
        return arg_error("stack too small: need " + std::to_string(need) + " bytes");
    hipStream_t s = (hipStream_t)stream;
    JKPM_HIP(hipMemcpyAsync(stack, v_r, vb, hipMemcpyDeviceToDevice, s));
    return run_moments(h->e, v_r, v_r, nullptr, R, M, (double2 *)mu, workspace, workspace_bytes,
                       s, nullptr, false, stack);
}

// ---- launch shape and autotuner ----

namespace {
// unroll x non-temporal x lane shape x store mode x matrix loads (cheb_step.h)
constexpr int kVariants = JKPM_STEP_VARIANTS;
}

int jkpm_ell_set_launch(jkpm_ell *h, int variant, int chunk_mult, int flags) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (variant < 0 || variant >= kVariants) return arg_error("variant must be 0.." + std::to_string(kVariants - 1));
    if (chunk_mult < 1 || chunk_mult > 64) return arg_error("chunk_mult must be 1..64");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    if (flags & JKPM_MOMENTS_DOUBLING) {
        h->e.dvariant = variant;
        h->e.dchunk_mult = chunk_mult;
    } else {
        h->e.variant = variant;
        h->e.chunk_mult = chunk_mult;
    }
    return JKPM_OK;
}

int jkpm_ell_set_pass(jkpm_ell *h, int vectors, int flags) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (vectors < 0) return arg_error("vectors per pass must be >= 0 (0: all at once)");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    (flags & JKPM_MOMENTS_DOUBLING ? h->e.dpass_r : h->e.pass_r) = vectors;
    return JKPM_OK;
}

int jkpm_ell_get_pass(const jkpm_ell *h, int *vectors, int flags) {
    clear_error();
    if (!h || !vectors) return arg_error("NULL argument");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    *vectors = flags & JKPM_MOMENTS_DOUBLING ? h->e.dpass_r : h->e.pass_r;
    return JKPM_OK;
}

int jkpm_ell_get_launch(const jkpm_ell *h, int *variant, int *chunk_mult, int flags) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    const bool dbl = flags & JKPM_MOMENTS_DOUBLING;
    if (variant) *variant = dbl ? h->e.dvariant : h->e.variant;
    if (chunk_mult) *chunk_mult = dbl ? h->e.dchunk_mult : h->e.chunk_mult;
    return JKPM_OK;
}

namespace {
const int kTuneMult[] = {2, 4, 8, 16, 32};

// Moments one timed pass computes: `steps` fused steps in either form.
int tune_moments(int steps, bool dbl) { return dbl ? 2 * steps : steps + 1; }
}  // namespace

int jkpm_autotune_workspace(const jkpm_ell *h, int R, int steps, int flags, size_t *bytes) {
    clear_error();
    if (!h || !bytes) return arg_error("NULL argument");
    if (R <= 0 || steps <= 0) return arg_error("R and steps must be positive");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const bool dbl = flags & JKPM_MOMENTS_DOUBLING;
    const int M = tune_moments(steps, dbl);
    Ell probe = h->e;
    size_t need = 0;
    for (int m : kTuneMult) {
        probe.chunk_mult = m;
        probe.dchunk_mult = m;
        need = std::max(need, moments_bytes(probe, R, 1, dev, nullptr, dbl ? M : 0));
    }
    *bytes = need + align_up((size_t)M * R * sizeof(double2), 256);
    return JKPM_OK;
}

int jkpm_autotune(jkpm_ell *h, const void *v_r, int R, int steps, int flags, void *workspace,
                  size_t workspace_bytes, void *stream, double *best_us) {
    clear_error();
    if (!h || !v_r || !workspace) return arg_error("NULL argument");
    if (R <= 0 || steps <= 0) return arg_error("R and steps must be positive");
    size_t need = 0;
    int rc = jkpm_autotune_workspace(h, R, steps, flags, &need);
    if (rc) return rc;
    if (workspace_bytes < need) {
        set_error("workspace too small: need " + std::to_string(need) + " bytes");
        return JKPM_ERR_WORKSPACE;
    }
    const bool dbl = flags & JKPM_MOMENTS_DOUBLING;
    const int M = tune_moments(steps, dbl);
    hipStream_t s = (hipStream_t)stream;
    const size_t mu_bytes = align_up((size_t)M * R * sizeof(double2), 256);
    double2 *mu = (double2 *)workspace;
    char *ws = (char *)workspace + mu_bytes;
    const size_t ws_bytes = workspace_bytes - mu_bytes;
    int &var = dbl ? h->e.dvariant : h->e.variant;
    int &mult = dbl ? h->e.dchunk_mult : h->e.chunk_mult;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    JKPM_HIP(hipEventCreate(&e0));
    JKPM_HIP(hipEventCreate(&e1));
    const int keep_v = var, keep_m = mult;
    double best = 1e300;
    int bv = keep_v, bm = keep_m;
    for (int v = 0; v < kVariants && !rc; ++v)
        for (int m : kTuneMult) {
            var = v;
            mult = m;
            // one untimed pass, then the timed pass of `steps` fused steps
            rc = run_moments(h->e, v_r, v_r, nullptr, R, M, mu, ws, ws_bytes, s, nullptr, dbl);
            if (!rc) rc = hipEventRecord(e0, s) == hipSuccess ? 0 : JKPM_ERR_HIP;
            if (!rc) rc = run_moments(h->e, v_r, v_r, nullptr, R, M, mu, ws, ws_bytes, s, nullptr, dbl);
            if (!rc) rc = hipEventRecord(e1, s) == hipSuccess ? 0 : JKPM_ERR_HIP;
            if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? 0 : JKPM_ERR_HIP;
            if (rc) break;
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) {
                best = ms;
                bv = v;
                bm = m;
            }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    var = rc ? keep_v : bv;
    mult = rc ? keep_m : bm;
    if (rc) {
        if (g_err.empty()) set_error("autotune launch failed");
        return rc;
    }
    if (best_us) *best_us = best * 1e3 / steps;
    return JKPM_OK;
}

// ---- block-diagonal batches of realizations ----

// One chunk table of a batch: chunk row starts (device [nchunks+1]) and the
// first chunk of every realization (device [G+1]).
struct ChunkTable {
    long *chunk_rows = nullptr;
    int *group_chunks = nullptr;
    int nchunks = 0;
};

struct jkpm_batch {
    const jkpm_ell *h = nullptr;  // borrowed: must outlive the batch
    int G = 0;
    ChunkTable t[2];              // [0] plain recursion, [1] doubling form
    Groups grp(bool doubling = false) const {
        const ChunkTable &c = t[doubling ? 1 : 0];
        Groups g;
        g.chunk_rows = c.chunk_rows;
        g.group_chunks = c.group_chunks;
        g.G = G;
        g.nchunks = c.nchunks;
        return g;
    }
};

namespace {

void free_table(ChunkTable &c) {
    hipFree(c.chunk_rows);
    hipFree(c.group_chunks);
    c = ChunkTable{};
}

// About cu_count * mult chunks over the whole batch; no chunk straddles two
// realizations.  Every realization starts with one chunk and the remaining
// chunks go, one at a time, to the realization whose chunks are currently the
// largest, so the largest chunk (the launch's critical path) stays close to
// the mean even when realizations are few chunks long.
// chunk_rows (optional, per realization): a positive entry cuts that
// realization into chunks of exactly that many rows instead (the last one
// shorter), so that the work-groups of consecutive chunks, which run side by
// side on one XCD, stand one lattice stride apart and read each other's
// neighbour rows at the same time (one HBM fetch shared through L2).
int build_table(const jkpm_ell *h, int G, const long long *row_offsets, const int *chunk_rows,
                int mult, const DeviceInfo &dev, ChunkTable *out) {
    std::vector<long> rows;
    std::vector<int> gch;
    std::string perr;
    if (plan_chunks(h->e.n, G, row_offsets, chunk_rows, (long)dev.cu_count * mult, kXcdCount,
                    &rows, &gch, &perr))
        return arg_error(perr);
    ChunkTable c;
    c.nchunks = (int)rows.size() - 1;
    hipError_t e;
    if ((e = hipMalloc(&c.chunk_rows, rows.size() * sizeof(long))) != hipSuccess ||
        (e = hipMalloc(&c.group_chunks, gch.size() * sizeof(int))) != hipSuccess ||
        (e = hipMemcpy(c.chunk_rows, rows.data(), rows.size() * sizeof(long),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c.group_chunks, gch.data(), gch.size() * sizeof(int),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        free_table(c);
        return hip_fail(e, "batch tables", __FILE__, __LINE__);
    }
    *out = c;
    return JKPM_OK;
}

}  // namespace

int jkpm_batch_create(const jkpm_ell *h, int G, const long long *row_offsets, jkpm_batch **out) {
    return jkpm_batch_create_strided(h, G, row_offsets, nullptr, out);
}

int jkpm_batch_create_strided(const jkpm_ell *h, int G, const long long *row_offsets,
                              const int *chunk_rows, jkpm_batch **out) {
    clear_error();
    if (!out) return arg_error("out is NULL");
    *out = nullptr;
    if (!h || !row_offsets) return arg_error("NULL argument");
    if (chunk_rows)
        for (int g = 0; g < G; ++g)
            if (chunk_rows[g] < 0) return arg_error("chunk_rows must be >= 0");
    if (G <= 0) return arg_error("G must be positive");
    if (row_offsets[0] != 0 || row_offsets[G] != h->e.n)
        return arg_error("row_offsets must start at 0 and end at the handle's row count");
    for (int g = 0; g < G; ++g)
        if (row_offsets[g + 1] <= row_offsets[g])
            return arg_error("row_offsets must be strictly increasing");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    auto *b = new (std::nothrow) jkpm_batch;
    if (!b) return arg_error("host allocation failed");
    b->h = h;
    b->G = G;
    rc = build_table(h, G, row_offsets, chunk_rows, h->e.chunk_mult, dev, &b->t[0]);
    if (!rc) rc = build_table(h, G, row_offsets, chunk_rows, h->e.dchunk_mult, dev, &b->t[1]);
    if (rc) {
        free_table(b->t[0]);
        free_table(b->t[1]);
        delete b;
        return rc;
    }
    *out = b;
    return JKPM_OK;
}

int jkpm_batch_destroy(jkpm_batch *b) {
    clear_error();
    if (!b) return JKPM_OK;
    free_table(b->t[0]);
    free_table(b->t[1]);
    delete b;
    return JKPM_OK;
}

int jkpm_batch_workspace(const jkpm_batch *b, int R, int n_left, size_t *bytes) {
    clear_error();
    if (!b || !bytes) return arg_error("NULL argument");
    if (R <= 0) return arg_error("R must be positive");
    if (n_left != 1 && n_left != 2) return arg_error("n_left must be 1 or 2");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const Groups g = b->grp();
    *bytes = moments_bytes(b->h->e, R, n_left, dev, &g);
    return JKPM_OK;
}

int jkpm_batch_moments(const jkpm_batch *b, const void *v_r, const void *v_l0, const void *v_l1,
                       int R, int M, double *mu, void *workspace, size_t workspace_bytes,
                       void *stream) {
    clear_error();
    if (!b) return arg_error("batch is NULL");
    if (!v_r || !v_l0 || !mu || !workspace) return arg_error("NULL vector/mu/workspace");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (!aligned16(v_r) || !aligned16(v_l0) || (v_l1 && !aligned16(v_l1)) ||
        !aligned16(workspace) || !aligned16(mu))
        return arg_error("device pointers must be 16-byte aligned");
    const Groups g = b->grp();
    return run_moments(b->h->e, v_r, v_l0, v_l1, R, M, (double2 *)mu, workspace, workspace_bytes,
                       (hipStream_t)stream, &g);
}

// ---- Chebyshev series application and the commutator recursion ----

int jkpm_series_workspace(const jkpm_ell *h, int R, int comm, size_t *bytes) {
    clear_error();
    if (!h || !bytes) return arg_error("NULL argument");
    if (R <= 0) return arg_error("R must be positive");
    *bytes = series_workspace(h->e, R, comm != 0);
    return JKPM_OK;
}

int jkpm_series(const jkpm_ell *h, const jkpm_ell *v_op, const void *v, int R, int M,
                const double *c, void *out, void *tab, void *workspace, size_t workspace_bytes,
                void *stream) {
    clear_error();
    if (!h || !v || !c || !workspace) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (!out && !tab) return arg_error("out and tab are both NULL");
    if (tab && !v_op) return arg_error("the table form needs the commutator operator");
    if (v_op && (v_op->e.col != h->e.col || v_op->e.vtype != h->e.vtype))
        return arg_error("the operator must share H's slot pattern and value type "
                         "(jkpm_ell_create_like)");
    if (!aligned16(workspace) || ((uintptr_t)v & 7) || (out && ((uintptr_t)out & 7)) ||
        (tab && ((uintptr_t)tab & 15)))
        return arg_error("misaligned device pointer");
    if (workspace_bytes < series_workspace(h->e, R, v_op != nullptr)) {
        set_error("workspace too small for jkpm_series");
        return JKPM_ERR_WORKSPACE;
    }
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    return run_series(h->e, v_op ? &v_op->e : nullptr, v, R, M, c, out, tab, workspace,
                      step_chunks(h->e, R, dev), (hipStream_t)stream);
}

// ---- single-left-set moments with flags (doubling form) ----

int jkpm_moments_ex_workspace(const jkpm_ell *h, const jkpm_batch *b, int R, int M, int flags,
                              size_t *bytes) {
    clear_error();
    if ((!h && !b) || !bytes) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const bool dbl = (flags & JKPM_MOMENTS_DOUBLING) != 0;
    if (b) {
        const Groups g = b->grp(dbl);
        *bytes = moments_bytes(b->h->e, R, 1, dev, &g, dbl ? M : 0);
    } else {
        *bytes = moments_bytes(h->e, R, 1, dev, nullptr, dbl ? M : 0);
    }
    return JKPM_OK;
}

int jkpm_moments_ex(const jkpm_ell *h, const jkpm_batch *b, const void *v_r, int R, int M,
                    int flags, double *mu, void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    if (!h && !b) return arg_error("handle and batch are both NULL");
    if (!v_r || !mu || !workspace) return arg_error("NULL vector/mu/workspace");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    if (!aligned16(v_r) || !aligned16(workspace) || !aligned16(mu))
        return arg_error("device pointers must be 16-byte aligned");
    const bool dbl = (flags & JKPM_MOMENTS_DOUBLING) != 0;
    if (dbl) {
        int rc = require_hermitian(b ? b->h : h, (hipStream_t)stream, "the doubling form");
        if (rc) return rc;
    }
    if (b) {
        const Groups g = b->grp(dbl);
        return run_moments(b->h->e, v_r, v_r, nullptr, R, M, (double2 *)mu, workspace,
                           workspace_bytes, (hipStream_t)stream, &g, dbl);
    }
    return run_moments(h->e, v_r, v_r, nullptr, R, M, (double2 *)mu, workspace, workspace_bytes,
                       (hipStream_t)stream, nullptr, dbl);
}

// ---- HIP graph capture of asynchronous entry points ----

struct jkpm_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int jkpm_graph_begin(void *stream) {
    clear_error();
    if (!stream) return arg_error("capture needs a non-default stream");
    JKPM_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return JKPM_OK;
}

int jkpm_graph_end(void *stream, jkpm_graph **out) {
    clear_error();
    if (!out) return arg_error("out is NULL");
    *out = nullptr;
    hipGraph_t g = nullptr;
    JKPM_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    auto *gr = new (std::nothrow) jkpm_graph;
    if (!gr) {
        hipGraphExecDestroy(ex);
        hipGraphDestroy(g);
        return arg_error("host allocation failed");
    }
    gr->graph = g;
    gr->exec = ex;
    *out = gr;
    return JKPM_OK;
}

int jkpm_graph_launch(const jkpm_graph *g, void *stream) {
    clear_error();
    if (!g) return arg_error("graph is NULL");
    JKPM_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return JKPM_OK;
}

int jkpm_graph_destroy(jkpm_graph *g) {
    clear_error();
    if (!g) return JKPM_OK;
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    delete g;
    return JKPM_OK;
}

int jkpm_dos(const double *mu, int M, int R, const double *e, int n_e, int kernel, double *dos,
             void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    if (!mu || !e || !dos || !workspace) return arg_error("NULL argument");
    if (M <= 0 || R <= 0 || n_e < 0) return arg_error("M and R must be positive, n_e >= 0");
    if (kernel != JKPM_KERNEL_NONE && kernel != JKPM_KERNEL_JACKSON)
        return arg_error("kernel must be JKPM_KERNEL_NONE or JKPM_KERNEL_JACKSON");
    if (workspace_bytes < (size_t)M * sizeof(double)) {
        set_error("workspace too small for jkpm_dos: need 8*M bytes");
        return JKPM_ERR_WORKSPACE;
    }
    return launch_dos((const double2 *)mu, M, R, e, n_e, kernel, dos, (double *)workspace,
                      (hipStream_t)stream);
}

// ---- adjoint gradient ----

namespace {
struct GradLayout {
    size_t ckpt = 0, F0 = 0, F1 = 0, tseg = 0, lam = 0, C1 = 0, C2 = 0, G = 0, total = 0;
    size_t vec = 0;
    int S = 1, nseg = 0, nchunks = 1;
    bool keep_all = false;
};

// seg > 0: checkpoint pairs every seg steps, T recomputed per segment;
// seg == 0: the same with seg = ceil(sqrt(M-1));
// seg < 0: every T_0 .. T_{M-2} kept (tseg holds them all, no checkpoints
// and no recomputation), lambda still folded in segments of ceil(sqrt(M-1)).
GradLayout grad_layout(const Ell &H, int R, int M, int seg, const DeviceInfo &dev) {
    GradLayout L;
    L.vec = align_up((size_t)H.n * R * cplx_bytes(H.vtype), 256);
    L.keep_all = seg < 0;
    int S = seg > 0 ? seg : 1;
    if (seg <= 0)
        while ((long)S * S < (long)(M - 1)) ++S;  // ceil(sqrt(M - 1))
    L.S = S;
    L.nseg = M > 1 ? (M - 1 + S - 1) / S : 0;
    // The adjoint steps, folds and recomputed steps take a fixed wide launch
    // (up to 32 work-groups per CU of at least 64 rows), not the forward's
    // tuned chunk count: the forward tunes to anything from 2 to 32
    // work-groups per CU at near-equal times, and at 2 the backward of the
    // square 512 x 512, M = 1024, R = 8 gradient took 59.7 ms against 46.3.
    {
        long c = std::min<long>((long)dev.cu_count * 32, ((long)H.n + 63) / 64);
        if (c >= 8) c = c / 8 * 8;
        L.nchunks = (int)std::max<long>(c, 1);
    }
    size_t o = 0;
    L.ckpt = o;
    if (!L.keep_all) o += (size_t)2 * L.nseg * L.vec;
    L.F0 = o;
    if (!L.keep_all) o += L.vec;
    L.F1 = o;
    if (!L.keep_all) o += L.vec;
    L.tseg = o;
    o += (size_t)(L.keep_all ? std::max(M - 1, 1) : S) * L.vec;
    L.lam = o;
    o += (size_t)S * L.vec;
    L.C1 = o;
    o += L.vec;
    L.C2 = o;
    o += L.vec;
    L.G = o;
    o += align_up((size_t)H.n * H.k * sizeof(double2), 256);
    L.total = o;
    return L;
}
}  // namespace

namespace {
// Reverse sweep over every T_0 .. T_{M-2} kept in `stack` (vector stride
// N*R elements): segments of L.S lambda vectors, lambda_lo .. lambda_hi by
// the adjoint steps, then one fold against T_{lo-1} .. T_{hi-1}.  Uses the
// workspace regions lam, C1, C2 and G of L (G zeroed by the caller).
int reverse_sweep_kept(const Ell &H, const void *r, int R, int M, const double2 *gmu,
                       const void *stack, const GradLayout &L, char *w, hipStream_t s) {
    const size_t vb = (size_t)H.n * R * cplx_bytes(H.vtype);
    const int S = L.S;
    auto lm = [&](int q) { return (void *)(w + L.lam + (size_t)q * vb); };
    auto ts = [&](int q) { return (const void *)((const char *)stack + (size_t)q * vb); };
    void *C1 = w + L.C1, *C2 = w + L.C2;
    double2 *G = (double2 *)(w + L.G);
    // vb is a multiple of 8, not of 16, for an odd N*R complex64 batch
    int rc = launch_zero_bytes(C1, vb, s);
    if (!rc) rc = launch_zero_bytes(C2, vb, s);
    if (rc) return rc;
    for (int c = L.nseg - 1; c >= 0; --c) {
        const int lo = 1 + c * S, hi = std::min((c + 1) * S, M - 1);
        const int count = hi - lo + 1;
        for (int k = hi; k >= lo; --k) {
            const void *x = (k + 1 <= hi) ? lm(k + 1 - lo) : C1;
            const void *prev = (k + 2 <= hi) ? lm(k + 2 - lo) : (k + 2 == hi + 1 ? C1 : C2);
            rc = launch_plain_step(H, x, prev, lm(k - lo), gmu + (size_t)k * R, r, R, 2.0, -1.0,
                                   L.nchunks, s);
            if (rc) return rc;
        }
        rc = launch_grad_fold(H, lm(0), ts(lo - 1), count, lo, G, R, L.nchunks, s);
        if (rc) return rc;
        JKPM_HIP(hipMemcpyAsync(C2, count >= 2 ? lm(1) : C1, vb, hipMemcpyDeviceToDevice, s));
        JKPM_HIP(hipMemcpyAsync(C1, lm(0), vb, hipMemcpyDeviceToDevice, s));
    }
    return JKPM_OK;
}
}  // namespace

int jkpm_grad_workspace(const jkpm_ell *h, int R, int M, int seg, size_t *bytes) {
    clear_error();
    if (!h || !bytes) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    *bytes = grad_layout(h->e, R, M, seg, dev).total;
    return JKPM_OK;
}

int jkpm_moments_grad(const jkpm_ell *h, const void *r, int R, int M, const double *grad_mu,
                      int seg, double *grad_values, void *workspace, size_t workspace_bytes,
                      void *stream) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (!r || !grad_mu || !grad_values || !workspace) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (!aligned16(r) || !aligned16(grad_mu) || !aligned16(grad_values) || !aligned16(workspace))
        return arg_error("device pointers must be 16-byte aligned");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const Ell &H = h->e;
    const GradLayout L = grad_layout(H, R, M, seg, dev);
    if (workspace_bytes < L.total) {
        set_error("workspace too small: need " + std::to_string(L.total) + " bytes");
        return JKPM_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    rc = require_hermitian(h, s, "the adjoint recursion");
    if (rc) return rc;
    char *w = (char *)workspace;
    const size_t vb = (size_t)H.n * R * cplx_bytes(H.vtype);  // bytes of one vector batch
    const int S = L.S;
    auto ck = [&](int i) { return (void *)(w + L.ckpt + (size_t)i * L.vec); };
    auto ts = [&](int q) { return (void *)(w + L.tseg + (size_t)q * vb); };
    auto lm = [&](int q) { return (void *)(w + L.lam + (size_t)q * vb); };
    void *F[2] = {w + L.F0, w + L.F1};
    void *C1 = w + L.C1, *C2 = w + L.C2;
    double2 *G = (double2 *)(w + L.G);
    const double2 *gmu = (const double2 *)grad_mu;
    const long gcount = (long)H.n * H.k;

    rc = launch_zero(G, gcount, s);
    if (rc) return rc;
    if (M >= 2 && L.keep_all) {
        // Forward pass straight into the stack: T_n in ts(n), n = 0 .. M-2.
        const int last = M - 2;
        JKPM_HIP(hipMemcpyAsync(ts(0), r, vb, hipMemcpyDeviceToDevice, s));
        for (int n = 0; n < last; ++n) {
            rc = launch_plain_step(H, ts(n), n == 0 ? nullptr : ts(n - 1), ts(n + 1), nullptr,
                                   nullptr, R, n == 0 ? 1.0 : 2.0, n == 0 ? 0.0 : -1.0,
                                   L.nchunks, s);
            if (rc) return rc;
        }
        rc = reverse_sweep_kept(H, r, R, M, gmu, w + L.tseg, L, w, s);
        if (rc) return rc;
    } else if (M >= 2) {
        // Forward pass: T_0 = r, T_n (n >= 1) in F[n & 1]; keep T_{cS}, T_{cS+1}.
        const int last = M - 2;  // highest T index the gradient needs
        JKPM_HIP(hipMemcpyAsync(ck(0), r, vb, hipMemcpyDeviceToDevice, s));
        for (int n = 0; n < last; ++n) {
            const void *x = n == 0 ? r : F[n & 1];
            const void *prev = n == 0 ? nullptr : (n == 1 ? r : F[(n + 1) & 1]);
            rc = launch_plain_step(H, x, prev, F[(n + 1) & 1], nullptr, nullptr, R,
                                   n == 0 ? 1.0 : 2.0, n == 0 ? 0.0 : -1.0, L.nchunks, s);
            if (rc) return rc;
            const int m = n + 1;
            if (m % S <= 1 && 2 * (m / S) + (m % S) < 2 * L.nseg)
                JKPM_HIP(hipMemcpyAsync(ck(2 * (m / S) + (m % S)), F[m & 1], vb,
                                        hipMemcpyDeviceToDevice, s));
        }
        // Reverse sweep over segments of k = 1 .. M-1.
        rc = launch_zero_bytes(C1, vb, s);
        if (!rc) rc = launch_zero_bytes(C2, vb, s);
        if (rc) return rc;
        for (int c = L.nseg - 1; c >= 0; --c) {
            const int lo = 1 + c * S, hi = std::min((c + 1) * S, M - 1);
            const int count = hi - lo + 1;  // T_{cS} .. T_{hi-1} and lambda_lo .. lambda_hi
            JKPM_HIP(hipMemcpyAsync(ts(0), ck(2 * c), vb, hipMemcpyDeviceToDevice, s));
            if (count >= 2)
                JKPM_HIP(hipMemcpyAsync(ts(1), ck(2 * c + 1), vb, hipMemcpyDeviceToDevice, s));
            for (int q = 2; q < count; ++q) {
                rc = launch_plain_step(H, ts(q - 1), ts(q - 2), ts(q), nullptr, nullptr, R, 2.0,
                                       -1.0, L.nchunks, s);
                if (rc) return rc;
            }
            for (int k = hi; k >= lo; --k) {
                const void *x = (k + 1 <= hi) ? lm(k + 1 - lo) : C1;
                const void *prev = (k + 2 <= hi) ? lm(k + 2 - lo) : (k + 2 == hi + 1 ? C1 : C2);
                rc = launch_plain_step(H, x, prev, lm(k - lo), gmu + (size_t)k * R, r, R, 2.0,
                                       -1.0, L.nchunks, s);
                if (rc) return rc;
            }
            rc = launch_grad_fold(H, lm(0), ts(0), count, lo, G, R, L.nchunks, s);
            if (rc) return rc;
            // carry lambda_lo -> C1 and lambda_{lo+1} -> C2 for the next segment down
            JKPM_HIP(hipMemcpyAsync(C2, count >= 2 ? lm(1) : C1, vb, hipMemcpyDeviceToDevice, s));
            JKPM_HIP(hipMemcpyAsync(C1, lm(0), vb, hipMemcpyDeviceToDevice, s));
        }
    }
    // the handle holds scale * values, so dL/dvalues = scale * dL/d(handle values)
    return launch_grad_out(G, H.n, H.k, H.scale, (double2 *)grad_values, s);
}

int jkpm_grad_kept_workspace(const jkpm_ell *h, int R, int M, size_t *bytes) {
    clear_error();
    if (!h || !bytes) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const GradLayout L = grad_layout(h->e, R, M, -1, dev);
    *bytes = L.total - (L.lam - L.tseg);  // everything but the stack
    return JKPM_OK;
}

int jkpm_moments_grad_kept(const jkpm_ell *h, const void *r, int R, int M, const double *grad_mu,
                           const void *stack, double *grad_values, void *workspace,
                           size_t workspace_bytes, void *stream) {
    clear_error();
    if (!h) return arg_error("handle is NULL");
    if (!r || !grad_mu || !stack || !grad_values || !workspace) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (!aligned16(r) || !aligned16(grad_mu) || !aligned16(stack) || !aligned16(grad_values) ||
        !aligned16(workspace))
        return arg_error("device pointers must be 16-byte aligned");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const Ell &H = h->e;
    GradLayout L = grad_layout(H, R, M, -1, dev);
    // this entry's workspace holds no stack: shift the regions after it down
    const size_t drop = L.lam - L.tseg;
    L.lam -= drop;
    L.C1 -= drop;
    L.C2 -= drop;
    L.G -= drop;
    L.total -= drop;
    if (workspace_bytes < L.total) {
        set_error("workspace too small: need " + std::to_string(L.total) + " bytes");
        return JKPM_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    rc = require_hermitian(h, s, "the adjoint recursion");
    if (rc) return rc;
    char *w = (char *)workspace;
    double2 *G = (double2 *)(w + L.G);
    rc = launch_zero(G, (long)H.n * H.k, s);
    if (rc) return rc;
    if (M >= 2) {
        rc = reverse_sweep_kept(H, r, R, M, (const double2 *)grad_mu, stack, L, w, s);
        if (rc) return rc;
    }
    return launch_grad_out(G, H.n, H.k, H.scale, (double2 *)grad_values, s);
}

int jkpm_dos_grad(const double *w, int M, int R, const double *e, int n_e, int kernel,
                  double *grad_mu, void *stream) {
    clear_error();
    if (!w || !e || !grad_mu) return arg_error("NULL argument");
    if (M <= 0 || R <= 0 || n_e < 0) return arg_error("M and R must be positive, n_e >= 0");
    if (kernel != JKPM_KERNEL_NONE && kernel != JKPM_KERNEL_JACKSON)
        return arg_error("kernel must be JKPM_KERNEL_NONE or JKPM_KERNEL_JACKSON");
    return launch_dos_grad(w, M, R, e, n_e, kernel, (double2 *)grad_mu, (hipStream_t)stream);
}

// ---- Kubo double moments ----

namespace {
struct KuboLayout {
    size_t L = 0, Rk = 0, T0 = 0, T1 = 0, part = 0, total = 0;
    int nslices = 1, nchunks = 1;
};

KuboLayout kubo_layout(const Ell &H, int R, int M, const DeviceInfo &dev) {
    KuboLayout K;
    const size_t vec = (size_t)H.n * R * 8;
    const size_t stack = align_up((size_t)M * vec, 256);
    K.nslices = kubo_gemm_slices((long)H.n * R, M, dev);
    // Stack steps: 16 work-groups per CU (at least 64 rows each).  On the
    // honeycomb 512 x 512 workload (R = 4) the whole step measured 44.9-45.1
    // ms at 16, 47.3-47.8 at 8, 46.1 at 4 and 51.9 at 32.
    // JKPM_KUBO_CHUNK_MULT overrides the work-groups per CU (tuning).
    long mult = 16;
    if (const char *env = getenv("JKPM_KUBO_CHUNK_MULT")) {
        const long m = atol(env);
        if (m > 0) mult = m;
    }
    {
        long c = std::min<long>((long)dev.cu_count * mult, ((long)H.n + 63) / 64);
        if (c >= 8) c = c / 8 * 8;
        K.nchunks = (int)std::max<long>(c, 1);
    }
    K.L = 0;
    K.Rk = stack;
    K.T0 = 2 * stack;
    K.T1 = K.T0 + align_up(vec, 256);
    K.part = K.T1 + align_up(vec, 256);
    K.total = K.part + align_up((size_t)K.nslices * M * M * sizeof(float2), 256);
    return K;
}

int kubo_check(const jkpm_ell *H, const jkpm_ell *V, int R, int M) {
    if (!H || !V) return arg_error("H/V handle is NULL");
    if (H->e.vtype != JKPM_C64 || V->e.vtype != JKPM_C64)
        return arg_error("the Kubo path needs complex64 handles (JKPM_C64) for H and v");
    if (V->e.col != H->e.col)
        return arg_error("v must share H's slot pattern: create it with jkpm_ell_create_like(H, ...)");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    return JKPM_OK;
}
}  // namespace

int jkpm_kubo_workspace(const jkpm_ell *H, int R, int M, size_t *bytes) {
    clear_error();
    if (!H || !bytes) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    *bytes = kubo_layout(H->e, R, M, dev).total;
    return JKPM_OK;
}

int jkpm_kubo_mu_nm(const jkpm_ell *H, const jkpm_ell *V, const void *r, int R, int M, double *mu,
                    void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    int rc = kubo_check(H, V, R, M);
    if (rc) return rc;
    if (!r || !mu || !workspace) return arg_error("NULL vector/mu/workspace");
    if (!aligned16(r) || !aligned16(mu) || !aligned16(workspace))
        return arg_error("device pointers must be 16-byte aligned");
    DeviceInfo dev;
    rc = device_info(&dev);
    if (rc) return rc;
    const KuboLayout K = kubo_layout(H->e, R, M, dev);
    if (workspace_bytes < K.total) {
        set_error("workspace too small: need " + std::to_string(K.total) + " bytes");
        return JKPM_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    rc = require_hermitian(H, s, "the Kubo double moments (H)");
    if (!rc) rc = require_hermitian(V, s, "the Kubo double moments (v)");
    if (rc) return rc;
    char *w = (char *)workspace;
    const size_t vec = (size_t)H->e.n * R * 8;
    char *Ls = w + K.L, *Rs = w + K.Rk;
    void *Tb[2] = {w + K.T0, w + K.T1};
    // Right recursion T_n(H) r (T_0 = r, T_n in Tb[n & 1]) with the fused dual
    // output Rk_n = v T_n, and the left recursion L_m = T_m(H) (v r), L_0 =
    // Rk_0, out of place into its stack.  Step n >= 1 advances both on one
    // pass over the slots: the right one to T_{n+1}, Rk_n and the left one to
    // L_n.
    // Profiling (jkpm_profile_enable): event pairs around runs of up to
    // kProfRun stack steps (kind JKPM_PROF_KUBO_STEP, one run every `every`
    // steps, never the first) and around the GEMM (JKPM_PROF_KUBO_GEMM).
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    JKPM_HIP(hipStreamIsCapturing(s, &cap));
    Prof &P = prof();
    std::unique_lock<std::mutex> plk(P.mu);
    const int every = P.every;
    const bool profiling = every > 0 && cap == hipStreamCaptureStatusNone;
    if (!profiling) plk.unlock();
    ProfRun run;
    run.kind = JKPM_PROF_KUBO_STEP;
    bool run_open = false;
    auto open_run = [&](ProfRun &r) -> int {
        r.e0 = take_event(P);
        r.e1 = take_event(P);
        if (!r.e0 || !r.e1) return hip_fail(hipErrorOutOfMemory, "hipEventCreate", __FILE__, __LINE__);
        JKPM_HIP(hipEventRecord(r.e0, s));
        r.steps = 0;
        return JKPM_OK;
    };
    auto close_run = [&](ProfRun &r) -> int {
        JKPM_HIP(hipEventRecord(r.e1, s));
        P.pending.push_back(r);
        if (P.pending.size() >= 8192) return drain_profile(P);
        return JKPM_OK;
    };
    rc = launch_kubo_step(H->e, &V->e, r, nullptr, Tb[1], Rs, R, 1.f, 0.f, K.nchunks, s);
    if (rc) return rc;
    JKPM_HIP(hipMemcpyAsync(Ls, Rs, vec, hipMemcpyDeviceToDevice, s));
    for (int n = 1; n < M; ++n) {
        if (profiling && !run_open && P.counter % every == 0) {
            rc = open_run(run);
            if (rc) return rc;
            run_open = true;
        }
        if (profiling) ++P.counter;
        const int m = n - 1;
        KuboLeft left;
        left.x = Ls + (size_t)m * vec;
        left.prev = m == 0 ? nullptr : Ls + (size_t)(m - 1) * vec;
        left.out = Ls + (size_t)(m + 1) * vec;
        left.alpha = m == 0 ? 1.f : 2.f;
        left.beta = m == 0 ? 0.f : -1.f;
        rc = launch_kubo_step(H->e, &V->e, Tb[n & 1], n == 1 ? r : Tb[(n + 1) & 1],
                              Tb[(n + 1) & 1], Rs + (size_t)n * vec, R, 2.f, -1.f, K.nchunks, s,
                              &left);
        if (rc) return rc;
        if (run_open && (++run.steps == kProfRun || n == M - 1)) {
            rc = close_run(run);
            if (rc) return rc;
            run_open = false;
        }
    }
    ProfRun gemm;
    gemm.kind = JKPM_PROF_KUBO_GEMM;
    if (profiling) {
        rc = open_run(gemm);
        if (rc) return rc;
    }
    rc = launch_kubo_gemm((const float2 *)Ls, (const float2 *)Rs, (long)H->e.n * R, M, K.nslices,
                          (float2 *)(w + K.part), s);
    if (rc) return rc;
    if (profiling) {
        gemm.steps = 1;
        rc = close_run(gemm);
        if (rc) return rc;
    }
    return launch_kubo_reduce((const float2 *)(w + K.part), K.nslices, M, 1.0 / R, (double2 *)mu, s);
}

int jkpm_kubo_sigma(const double *mu, int M, const double *e, int n_e, int kind, double bound,
                    double *sigma, void *stream) {
    clear_error();
    if (!mu || !e || !sigma) return arg_error("NULL argument");
    if (M <= 0 || M > 2048) return arg_error("M must be in [1, 2048]");
    if (n_e < 0) return arg_error("n_e must be >= 0");
    if (kind != 0 && kind != 1) return arg_error("kind must be 0 (Kubo-Greenwood) or 1 (G-delta table)");
    if (!(bound > 0.0)) return arg_error("bound must be positive");
    return launch_kubo_sigma((const double2 *)mu, M, e, n_e, kind, bound, (double2 *)sigma,
                             (hipStream_t)stream);
}

int jkpm_profile_enable(int every) {
    clear_error();
    if (every < 0) return arg_error("every must be >= 0");
    Prof &P = prof();
    std::lock_guard<std::mutex> lk(P.mu);
    P.every = every;
    P.counter = 0;
    return JKPM_OK;
}

int jkpm_profile_read_kind(int kind, double *ms, long long *launches, int reset) {
    clear_error();
    if (kind < 0 || kind >= JKPM_PROF_KINDS) return arg_error("unknown profile kind");
    Prof &P = prof();
    std::lock_guard<std::mutex> lk(P.mu);
    int rc = drain_profile(P);
    if (rc) return rc;
    if (ms) *ms = P.ms[kind];
    if (launches) *launches = P.launches[kind];
    if (reset) {
        P.ms[kind] = 0.0;
        P.launches[kind] = 0;
    }
    return JKPM_OK;
}

int jkpm_profile_read(double *step_ms, long long *step_launches, int reset) {
    return jkpm_profile_read_kind(JKPM_PROF_STEP, step_ms, step_launches, reset);
}

int jkpm_copy(void *dst, const void *src, size_t bytes, void *stream) {
    clear_error();
    if (!dst || !src) return arg_error("NULL argument");
    if (bytes % 16 || !aligned16(dst) || !aligned16(src))
        return arg_error("jkpm_copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    return launch_copy(dst, src, bytes, (hipStream_t)stream);
}

// ---- reference drop-ins: host pointers, synchronous, complex128 ----

static int rec_host(const double *data, const int *indices, int len_row, int N, double *A_vec,
                    const double *v_l, const double *v_r, int M) {
    if (!data || !indices || !A_vec || !v_l || !v_r) return arg_error("NULL argument");
    if (N <= 0 || len_row <= 0 || M <= 0) return arg_error("N, len_row and M must be positive");
    jkpm_ell *h = nullptr;
    int rc = jkpm_ell_create(data, indices, N, len_row, 1.0, JKPM_C128, 0, nullptr, &h);
    if (rc) return rc;
    size_t ws_bytes = 0;
    rc = jkpm_moments_workspace(h, 1, 1, &ws_bytes);
    void *ws = nullptr, *vr = nullptr, *vl = nullptr, *mu = nullptr;
    auto cleanup = [&](int code) {
        hipFree(ws);
        hipFree(vr);
        hipFree(vl);
        hipFree(mu);
        jkpm_ell_destroy(h);
        return code;
    };
    if (rc) return cleanup(rc);
    hipError_t e;
    const size_t vb = (size_t)N * 16;
    if ((e = hipMalloc(&ws, ws_bytes)) != hipSuccess || (e = hipMalloc(&vr, vb)) != hipSuccess ||
        (e = hipMalloc(&vl, vb)) != hipSuccess || (e = hipMalloc(&mu, (size_t)M * 16)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMalloc", __FILE__, __LINE__));
    if ((e = hipMemcpy(vr, v_r, vb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(vl, v_l, vb, hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMemcpy H2D", __FILE__, __LINE__));
    rc = jkpm_moments(h, vr, vl, nullptr, 1, M, (double *)mu, ws, ws_bytes, nullptr);
    if (rc) return cleanup(rc);
    if ((e = hipMemcpy(A_vec, mu, (size_t)M * 16, hipMemcpyDeviceToHost)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMemcpy D2H", __FILE__, __LINE__));
    return cleanup(JKPM_OK);
}

int rec_A_tab(const double *data, const int *indices, int len_row, int N, double *A_vec,
              const double *v, int M) {
    clear_error();
    return rec_host(data, indices, len_row, N, A_vec, v, v, M);
}

int rec_A_tab_2(const double *data, const int *indices, int len_row, int N, double *A_vec,
                const double *v_l, const double *v_r, int M) {
    clear_error();
    return rec_host(data, indices, len_row, N, A_vec, v_l, v_r, M);
}

// ---- tile plans (two steps per launch) ----

int jkpm_tiles_create(const jkpm_ell *h, int R, int max_rows, jkpm_tiles **out) {
    clear_error();
    if (!out) return arg_error("out is NULL");
    *out = nullptr;
    if (!h) return arg_error("handle is NULL");
    if (R <= 0) return arg_error("R must be positive");
    const Ell &H = h->e;
    int rb = tile_max_group();
    while (rb > R) rb >>= 1;
    int cap = tile_cap_rows(H.vtype, rb);
    if (max_rows > 0) cap = std::min(cap, max_rows);
    std::vector<int> colh((size_t)H.n * H.k);
    JKPM_HIP(hipMemcpy(colh.data(), H.col, colh.size() * sizeof(int), hipMemcpyDeviceToHost));
    TilePlanHost plan;
    std::string perr;
    if (build_tile_plan(H.n, H.k, colh, cap, kXcdCount, &plan, &perr) ||
        validate_tile_plan(H.n, H.k, colh, cap, plan, &perr))
        return arg_error(perr);
    auto *t = new (std::nothrow) jkpm_tiles;
    if (!t) return arg_error("host allocation failed");
    t->h = h;
    t->n = H.n;
    t->k = H.k;
    t->vtype = H.vtype;
    t->rb = rb;
    t->cap = cap;
    t->npatch = plan.npatch;
    t->max_ext = plan.max_ext;
    t->halo_rows = (long long)plan.hrows.size();
    auto fail = [&](hipError_t e, const char *what) {
        jkpm_tiles_destroy(t);
        return hip_fail(e, what, __FILE__, __LINE__);
    };
    const size_t nk = (size_t)H.n * H.k;
    hipError_t e;
    if ((e = hipMalloc(&t->perm, (size_t)H.n * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&t->inv, (size_t)H.n * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&t->val, nk * cplx_bytes(H.vtype))) != hipSuccess ||
        (e = hipMalloc(&t->col, nk * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&t->lcol, nk * sizeof(short))) != hipSuccess ||
        (e = hipMalloc(&t->pstart, plan.pstart.size() * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&t->hoff, plan.hoff.size() * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&t->hrows, std::max<size_t>(1, plan.hrows.size()) * sizeof(int))) !=
            hipSuccess ||
        (e = hipMalloc(&t->hval, std::max<size_t>(1, plan.hrows.size() * H.k) *
                                     cplx_bytes(H.vtype))) != hipSuccess ||
        (e = hipMalloc(&t->hcol, std::max<size_t>(1, plan.hrows.size() * H.k) * sizeof(int))) !=
            hipSuccess)
        return fail(e, "hipMalloc(tile plan)");
    if ((e = hipMemcpy(t->perm, plan.perm.data(), (size_t)H.n * sizeof(int),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(t->inv, plan.inv.data(), (size_t)H.n * sizeof(int),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(t->lcol, plan.lcol.data(), nk * sizeof(short), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(t->pstart, plan.pstart.data(), plan.pstart.size() * sizeof(int),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(t->hoff, plan.hoff.data(), plan.hoff.size() * sizeof(int),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (!plan.hrows.empty() &&
         (e = hipMemcpy(t->hrows, plan.hrows.data(), plan.hrows.size() * sizeof(int),
                        hipMemcpyHostToDevice)) != hipSuccess))
        return fail(e, "hipMemcpy(tile plan)");
    int rc = launch_tile_permute(H, t->perm, t->inv, t->val, t->col, true, nullptr);
    if (!rc) rc = launch_halo_gather(*t, true, nullptr);
    if (!rc && (e = hipStreamSynchronize(nullptr)) != hipSuccess) return fail(e, "tile permute");
    if (rc) {
        jkpm_tiles_destroy(t);
        return rc;
    }
    *out = t;
    return JKPM_OK;
}

int jkpm_tiles_plan_check(int n_rows, int len_row, const int *col_slot_major, int cap,
                          int *npatch, long long *halo_rows, int *max_ext) {
    clear_error();
    if (!col_slot_major) return arg_error("columns are NULL");
    if (n_rows <= 0 || len_row <= 0) return arg_error("n_rows and len_row must be positive");
    if (cap > 32767) return arg_error("cap must fit the 16-bit patch-local columns");
    std::vector<int> colh(col_slot_major, col_slot_major + (size_t)n_rows * len_row);
    for (int j : colh)
        if (j < 0 || j >= n_rows) return arg_error("column index out of range");
    TilePlanHost plan;
    std::string perr;
    if (build_tile_plan(n_rows, len_row, colh, cap, kXcdCount, &plan, &perr) ||
        validate_tile_plan(n_rows, len_row, colh, cap, plan, &perr))
        return arg_error(perr);
    if (npatch) *npatch = plan.npatch;
    if (halo_rows) *halo_rows = (long long)plan.hrows.size();
    if (max_ext) *max_ext = plan.max_ext;
    return JKPM_OK;
}

int jkpm_tiles_destroy(jkpm_tiles *t) {
    clear_error();
    if (!t) return JKPM_OK;
    hipFree(t->perm);
    hipFree(t->inv);
    hipFree(t->val);
    hipFree(t->col);
    hipFree(t->lcol);
    hipFree(t->pstart);
    hipFree(t->hoff);
    hipFree(t->hrows);
    hipFree(t->hval);
    hipFree(t->hcol);
    delete t;
    return JKPM_OK;
}

int jkpm_tiles_info(const jkpm_tiles *t, int *npatch, long long *halo_rows, int *max_ext,
                    int *cap) {
    clear_error();
    if (!t) return arg_error("plan is NULL");
    if (npatch) *npatch = t->npatch;
    if (halo_rows) *halo_rows = t->halo_rows;
    if (max_ext) *max_ext = t->max_ext;
    if (cap) *cap = t->cap;
    return JKPM_OK;
}

int jkpm_tiles_refresh(jkpm_tiles *t, void *stream) {
    clear_error();
    if (!t) return arg_error("plan is NULL");
    int rc = launch_tile_permute(t->h->e, t->perm, t->inv, t->val, t->col, false,
                                 (hipStream_t)stream);
    if (rc) return rc;
    return launch_halo_gather(*t, false, (hipStream_t)stream);
}

int jkpm_tiles_workspace(const jkpm_tiles *t, int R, int M, int flags, size_t *bytes) {
    clear_error();
    if (!t || !bytes) return arg_error("NULL argument");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    *bytes = tiles_layout(*t, R, M, (flags & JKPM_MOMENTS_DOUBLING) != 0).total;
    return JKPM_OK;
}

int jkpm_tiles_moments(const jkpm_tiles *t, const void *v_r, int R, int M, int flags,
                       double *mu, void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    if (!t) return arg_error("plan is NULL");
    if (!v_r || !mu || !workspace) return arg_error("NULL vector/mu/workspace");
    if (R <= 0 || M <= 0) return arg_error("R and M must be positive");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    if (!aligned16(v_r) || !aligned16(workspace) || !aligned16(mu))
        return arg_error("device pointers must be 16-byte aligned");
    int rb = tile_max_group();
    while (rb > R) rb >>= 1;
    if (rb > t->rb && t->max_ext > tile_cap_rows(t->vtype, rb))
        return arg_error("the plan was built for fewer vectors per group; rebuild it for this R");
    if (flags & JKPM_MOMENTS_DOUBLING) {
        int rc = require_hermitian(t->h, (hipStream_t)stream, "the doubling form");
        if (rc) return rc;
    }
    return run_tiles(*t, v_r, R, M, (flags & JKPM_MOMENTS_DOUBLING) != 0, (double2 *)mu,
                     workspace, workspace_bytes, (hipStream_t)stream);
}

// ---- wavefront plans (two steps per launch) ----

int jkpm_wave_create(const jkpm_ell *h, int R, int flags, jkpm_wave **out) {
    clear_error();
    if (!out) return arg_error("out is NULL");
    *out = nullptr;
    if (!h) return arg_error("handle is NULL");
    if (flags & ~JKPM_MOMENTS_DOUBLING) return arg_error("unknown flags");
    const Ell &H = h->e;
    if (H.vtype != JKPM_C64) return arg_error("the wavefront path takes complex64 handles");
    const int rpw = wave_rows_per_wave(R);
    if (rpw == 0) return arg_error("the wavefront path takes R = 8 or 16");
    if (H.k != 4 && H.k != 5) return arg_error("the wavefront path takes ELL width 4 or 5");
    // the kernel indexes vectors and slots with 32-bit offsets
    if ((long)H.n * R * 2 >= (1L << 31) || (long)H.n * H.k * 2 >= (1L << 31))
        return arg_error("the wavefront path takes N * R * 2 and N * K * 2 below 2^31");
    const bool doubling = (flags & JKPM_MOMENTS_DOUBLING) != 0;
    DeviceInfo dev;
    int rc = device_info(&dev);
    if (rc) return rc;
    const int nchunks = step_chunks(H, R, dev, doubling);
    const int blk = 4 * rpw;
    long rpc = ((long)H.n + nchunks - 1) / nchunks;
    rpc = (rpc + blk - 1) / blk * blk;
    std::vector<int> colh((size_t)H.n * H.k);
    JKPM_HIP(hipMemcpy(colh.data(), H.col, colh.size() * sizeof(int), hipMemcpyDeviceToHost));
    WavePlanHost plan;
    std::string perr;
    if (build_wave_plan(H.n, H.k, colh, nchunks, (int)rpc, blk, rpw, kXcdCount, 1, kWaveFixGroup,
                        &plan, &perr) ||
        validate_wave_plan(H.n, H.k, colh, kXcdCount, 1, kWaveFixGroup, plan, &perr))
        return arg_error(perr);
    // kernel lag: the plan's + 1 at least; JKPM_WAVE_LAG (2 or 3) may ask for more
    int d = plan.lag + 1;
    if (const char *env = getenv("JKPM_WAVE_LAG")) d = std::max(d, atoi(env));
    if (d > 3) return arg_error("wavefront path: lag " + std::to_string(d) + " above 3");
    int per_cu = 0;
    if ((rc = wave_max_blocks_per_cu(R, H.k, d, doubling, &per_cu))) return rc;
    if ((long)per_cu * dev.cu_count < nchunks)
        return arg_error("wavefront path: " + std::to_string(nchunks) +
                         " work-groups cannot be co-resident (" + std::to_string(per_cu) +
                         " per CU)");
    auto *W = new (std::nothrow) jkpm_wave;
    if (!W) return arg_error("host allocation failed");
    W->h = h;
    W->R = R;
    W->rb = R;
    W->lag = plan.lag;
    W->d = d;
    W->nchunks = nchunks;
    W->rpc = plan.rpc;
    W->nblk = plan.nblk;
    W->nfix = (int)plan.fix_group_chunk.size();
    W->nfold = (int)plan.fold_chunk.size();
    W->doubling = doubling;
    W->tail = (long long)plan.tail_rows.size();
    W->deferred = (long long)plan.fix_rows.size();
    auto fail = [&](hipError_t e, const char *what) {
        jkpm_wave_destroy(W);
        return hip_fail(e, what, __FILE__, __LINE__);
    };
    auto up = [](auto **dst, const auto &v) -> hipError_t {
        using T = typename std::decay_t<decltype(v)>::value_type;
        hipError_t e = hipMalloc((void **)dst, std::max<size_t>(1, v.size()) * sizeof(T));
        if (e == hipSuccess && !v.empty())
            e = hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        return e;
    };
    const int nl = doubling ? 2 : 1;
    hipError_t e;
    if ((e = up(&W->deps, plan.deps)) != hipSuccess || (e = up(&W->tail_off, plan.tail_off)) != hipSuccess ||
        (e = up(&W->tail_rows, plan.tail_rows)) != hipSuccess ||
        (e = up(&W->tail_dep, plan.tail_dep)) != hipSuccess ||
        (e = up(&W->fix_group_off, plan.fix_group_off)) != hipSuccess ||
        (e = up(&W->fix_rows, plan.fix_rows)) != hipSuccess ||
        (e = up(&W->fold_chunk, plan.fold_chunk)) != hipSuccess ||
        (e = up(&W->fold_off, plan.fold_off)) != hipSuccess ||
        (e = hipMalloc(&W->fixpart, std::max<size_t>(1, (size_t)W->nfix * nl * R) * sizeof(double2))) !=
            hipSuccess ||
        (e = hipMalloc(&W->flags, (size_t)nchunks * 4 * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&W->err, sizeof(unsigned))) != hipSuccess ||
        (e = hipMemset(W->flags, 0, (size_t)nchunks * 4 * sizeof(unsigned))) != hipSuccess ||
        (e = hipMemset(W->err, 0, sizeof(unsigned))) != hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess)
        return fail(e, "wavefront plan upload");
    *out = W;
    return JKPM_OK;
}

int jkpm_wave_destroy(jkpm_wave *w) {
    clear_error();
    if (!w) return JKPM_OK;
    hipFree(w->deps);
    hipFree(w->tail_off);
    hipFree(w->tail_rows);
    hipFree(w->tail_dep);
    hipFree(w->fix_group_off);
    hipFree(w->fix_rows);
    hipFree(w->fold_chunk);
    hipFree(w->fold_off);
    hipFree(w->fixpart);
    hipFree(w->flags);
    hipFree(w->err);
    delete w;
    return JKPM_OK;
}

int jkpm_wave_info(const jkpm_wave *w, int *nchunks, int *lag, long long *tail_rows,
                   long long *fixup_rows) {
    clear_error();
    if (!w) return arg_error("plan is NULL");
    if (nchunks) *nchunks = w->nchunks;
    if (lag) *lag = w->d;
    if (tail_rows) *tail_rows = w->tail;
    if (fixup_rows) *fixup_rows = w->deferred;
    return JKPM_OK;
}

int jkpm_wave_plan_check(int n_rows, int len_row, const int *col_slot_major, int nchunks,
                         int rpc, int rows_per_wave, int max_lag, int *lag, long long *tail_rows,
                         long long *fixup_rows) {
    clear_error();
    if (!col_slot_major) return arg_error("columns are NULL");
    if (n_rows <= 0 || len_row <= 0) return arg_error("n_rows and len_row must be positive");
    std::vector<int> colh(col_slot_major, col_slot_major + (size_t)n_rows * len_row);
    WavePlanHost plan;
    std::string perr;
    if (build_wave_plan(n_rows, len_row, colh, nchunks, rpc, 4 * rows_per_wave, rows_per_wave,
                        kXcdCount, max_lag, kWaveFixGroup, &plan, &perr) ||
        validate_wave_plan(n_rows, len_row, colh, kXcdCount, max_lag, kWaveFixGroup, plan, &perr))
        return arg_error(perr);
    if (lag) *lag = plan.lag;
    if (tail_rows) *tail_rows = (long long)plan.tail_rows.size();
    if (fixup_rows) *fixup_rows = (long long)plan.fix_rows.size();
    return JKPM_OK;
}

int jkpm_wave_workspace(const jkpm_wave *w, int R, int M, size_t *bytes) {
    clear_error();
    if (!w || !bytes) return arg_error("NULL argument");
    if (R != w->R) return arg_error("R differs from the plan's");
    if (M <= 0) return arg_error("M must be positive");
    *bytes = wave_layout(*w, M).total;
    return JKPM_OK;
}

int jkpm_wave_moments(jkpm_wave *w, const void *v_r, int R, int M, double *mu, void *workspace,
                      size_t workspace_bytes, void *stream) {
    clear_error();
    if (!w) return arg_error("plan is NULL");
    if (!v_r || !mu || !workspace) return arg_error("NULL vector/mu/workspace");
    if (R != w->R) return arg_error("R differs from the plan's");
    if (M <= 0) return arg_error("M must be positive");
    if (!aligned16(v_r) || !aligned16(workspace) || !aligned16(mu))
        return arg_error("device pointers must be 16-byte aligned");
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    JKPM_HIP(hipStreamIsCapturing((hipStream_t)stream, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return arg_error("the wavefront path cannot be captured into a graph");
    if (w->doubling) {
        int rc = require_hermitian(w->h, (hipStream_t)stream, "the doubling form");
        if (rc) return rc;
    }
    return run_wave(*w, v_r, M, (double2 *)mu, workspace, workspace_bytes, (hipStream_t)stream);
}

}  // extern "C"
