// Internal helpers shared by the jclkpm HIP sources (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>

#include <vector>

#include "../../include/jclkpm.h"
#include "plan.h"

struct jkpm_tiles;

namespace jkpm {

// Thread-local error message behind jkpm_last_error().
void set_error(const std::string &msg);
void clear_error();

// Returns JKPM_ERR_HIP and records the message when e != hipSuccess.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define JKPM_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return ::jkpm::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define JKPM_CHECK_LAUNCH(name)                                               \
    do {                                                                      \
        hipError_t e_ = hipGetLastError();                                    \
        if (e_ != hipSuccess) return ::jkpm::hip_fail(e_, name, __FILE__, __LINE__); \
    } while (0)

inline int arg_error(const std::string &msg) {
    set_error(msg);
    return JKPM_ERR_ARG;
}

// Device properties of the current device, cached per device.
struct DeviceInfo {
    int cu_count = 256;
    int xcd_count = 8;
};
int device_info(DeviceInfo *out);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Device-resident ELL matrix: slot-major values and column indices.
struct Ell {
    int n = 0;        // rows (= columns)
    int k = 0;        // slots per row (len_row)
    int vtype = JKPM_C128;
    double scale = 1.0;    // the stored values are scale * (uploaded values)
    // Launch shape of the step kernel (jkpm_ell_set_launch / the autotuner):
    // variant 0..47 = unroll | non-temporal loads << 1 | lane shape << 2,
    // plus 12 for non-temporal stores of T_{n+1}, plus 24 for non-temporal
    // matrix loads; chunk_mult work-groups per CU.
    int variant = 0;
    int chunk_mult = 8;
    // the same for the doubling form of the recursion
    int dvariant = 0;
    int dchunk_mult = 8;
    // Vectors per recursion pass (jkpm_ell_set_pass; 0 = all R at once): a
    // call with more vectors runs the whole recursion for each group of
    // pass_r of them in turn, so one pass's vectors and the matrix can stay
    // in the 256 MB Infinity Cache.  Plain form and doubling form.
    int pass_r = 0;
    int dpass_r = 0;
    void *val = nullptr;   // [k][n] complex of vtype
    int *col = nullptr;    // [k][n]
    int device = 0;
    // Real parts of val, [k][n] reals of vtype's precision, owned by the
    // handle; valid while real_values (every stored imaginary part is exactly
    // 0, checked at each value upload), when the step kernels read these.
    // While imag_values it holds the imaginary parts instead.
    void *rval = nullptr;
    bool real_values = false;
    // every stored real part is exactly 0 and some imaginary part is not: rval
    // then holds the imaginary parts instead (read by the Kubo step for v)
    bool imag_values = false;
    // complex64 plain recursions with one left set keep the left vectors'
    // phases in the workspace and read them instead when every element is a
    // unit complex number (jkpm_ell_use_left_phases turns this off)
    bool left_phases = true;
};

// Bytes of one complex element of a value type.
inline size_t cplx_bytes(int vtype) { return vtype == JKPM_C64 ? 8 : 16; }

// Chebyshev step launch description (one launch handles a group of RB
// vectors of a batch of `ld` interleaved vectors, starting at vector r0).
struct StepArgs {
    const Ell *H;
    const void *xcur;     // T_n          [n][ld]
    const void *xprev;    // T_{n-1}      [n][ld] (unused when first)
    void *xout;           // T_{n+1}      [n][ld] (may equal xprev)
    const void *vl0;      // left vectors [n][ld]
    const void *vl1;      // second left set or nullptr
    // complex64, one left set: vl0 as 32-bit phases [n][ld] and the device
    // flag that is 0 when every phase reproduces vl0 (else vl0 is read)
    const unsigned *vph = nullptr;
    const int *ph_bad = nullptr;
    int ld;
    int r0;
    int rb;
    bool first;           // T_1 = A T_0 and also emit <vl|T_0>
    double2 *part;        // partials of this step: [n_left][nchunks][ld]
    double2 *part0;       // when first: partials of step 0 (same layout)
    int nchunks;
    const long *chunk_rows = nullptr;  // device [nchunks+1] row starts, or uniform chunks
    bool doubling = false;  // partials <T_{n+1}|T_{n+1}>, <T_n|T_{n+1}> (no left vectors)
    bool reverse = false;   // sweep each chunk from its last row to its first
};

// mu[g][m][r] from the doubling form's raw dots raw[l][g][s][r] (s < S1):
// mu_0 = a_0, mu_1 = b_0, mu_2k = 2 a_k - a_0, mu_2k+1 = 2 b_k - b_0 with
// a_k at (l=0, s=k) and b_k at (l=1, s=k+1).
// mu_ld: the row length of mu (R when 0; larger when mu is a column slice
// of a wider moment array).
int launch_doubling_finish(const double2 *raw, int G, int S1, int R, int M, double2 *mu,
                           hipStream_t s, int mu_ld = 0);

// Number of row chunks (work-groups) used for a matrix of n rows.
int step_chunks(const Ell &H, int rb, const DeviceInfo &dev, bool doubling = false);

// Largest vector group one launch handles for a value type.
int max_group(int vtype);

// Phases of n complex64 numbers (32-bit, 2 pi = 2^32, as read by the step
// kernel) into ph; *bad is cleared first and becomes nonzero when some
// element is not reproduced within 2^-21 per component by its phase.
int launch_phase_pack(const void *v, long n, unsigned *ph, int *bad, hipStream_t s);

int launch_step(const StepArgs &a, hipStream_t s);

// mu[l][g][step][r] = sum over the chunks of group g (group_chunks[g] ..
// group_chunks[g+1]; all chunks and G = 1 when group_chunks is null) of the
// ring partials, for steps [step_first, step_first + nsteps).  mu rows are
// mu_ld long (ld when 0).
int launch_reduce(const double2 *ring, int ring_slots, int step_first, int nsteps,
                  int n_left, int nchunks, int ld, double2 *mu, int M,
                  const int *group_chunks, int G, hipStream_t s, int mu_ld = 0);

// dst[i][c] = src[i][c0 + c] for i < n, c < cols: columns c0 .. c0+cols-1 of
// an [n][ld] array of elem_bytes-sized (8 or 16) elements, made contiguous.
int launch_take_columns(const void *src, long n, int ld, int c0, int cols, int elem_bytes,
                        void *dst, hipStream_t s);

int launch_aaxby(const Ell &H, double2 a, double2 b, const void *x, void *y, int R,
                 hipStream_t s);

int launch_ell_pack(const double *src_rowmajor, double scale, Ell *H, const int *idx_rowmajor,
                    hipStream_t s);

int launch_dos(const double2 *mu, int M, int R, const double *e, int n_e, int kernel,
               double *dos, double *w, hipStream_t s);

int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);

// ---- adjoint recursion (adjoint.hip) ----
// out = alpha*H*x + beta*prev + coef[r]*u (prev unused when beta == 0,
// u and coef unused when coef is null); vectors [n][R] of H's value type.
int launch_plain_step(const Ell &H, const void *x, const void *prev, void *out,
                      const double2 *coef, const void *u, int R, double alpha, double beta,
                      int nchunks, hipStream_t s);
// G[s][i] += sum_{q<count} f_{k_lo+q} sum_r lam_q[i,r] conj(tprev_q[col(i,s),r]),
// lam_q = lam + q*N*R, tprev_q = tprev + q*N*R, f_1 = 1, f_k = 2 otherwise.
int launch_grad_fold(const Ell &H, const void *lam, const void *tprev, int count, int k_lo,
                     double2 *G, int R, int nchunks, hipStream_t s);
int launch_grad_out(const double2 *G, int n, int k, double scale, double2 *out, hipStream_t s);
int launch_zero(double2 *p, long cnt, hipStream_t s);
// Clears `bytes` bytes (a multiple of 8) from a 16-byte aligned p, exactly.
int launch_zero_bytes(void *p, size_t bytes, hipStream_t s);
int launch_dos_grad(const double *w, int M, int R, const double *e, int n_e, int kernel,
                    double2 *grad_mu, hipStream_t s);

// ---- Peierls phases and value unpack (peierls.hip) ----
// dst = a * src * exp(i (ax dx + ay dy)) per slot, (dx, dy) = (Re, Im) of the
// bonds handle's values; src, bonds, dst share one slot pattern (dst may be src).
int launch_peierls(const Ell &src, const Ell &bonds, Ell *dst, double2 a, double2 ax, double2 ay,
                   hipStream_t s);
// out[i*k + s] = value(i, s) / H.scale as complex128 (device, row-major).
int launch_ell_unpack(const Ell &H, double *out, hipStream_t s);

// stats (device unsigned[3], zeroed by the caller): longest row, bad indptr
// entries, bad column indices.
int launch_csr_scan(const int *indptr, const int *indices, int n, long nnz, unsigned *stats,
                    hipStream_t s);
int launch_csr_to_ell(const int *indptr, const int *indices, const double *data, double scale,
                      Ell *H, hipStream_t s);
int launch_ell_cols_unpack(const Ell &H, int *out, hipStream_t s);

// ---- two steps per launch over LDS patches (tiles.hip) ----
// Rows of T_{n+1} the LDS tile holds for a vector group of rb (value type vtype).
int tile_cap_rows(int vtype, int rb);
int tile_max_group();
// (the patch plan itself, TilePlanHost and build_tile_plan, is in plan.h)
// Values (and columns, renumbered) of H in the plan's row order.
int launch_tile_permute(const Ell &H, const int *perm, const int *inv, void *val_t, int *col_t,
                        bool with_cols, hipStream_t s);
// dst[i][:] = src[perm[i]][:] for R interleaved complex vectors of vtype.
int launch_vec_permute(int vtype, const void *src, const int *perm, int n, int R, void *dst,
                       hipStream_t s);
struct TileArgs {
    const jkpm_tiles *T = nullptr;
    const void *xprev = nullptr;  // T_{n-1} (unused when first)
    const void *xcur = nullptr;   // T_n
    void *out1 = nullptr;         // T_{n+1}
    void *out2 = nullptr;         // T_{n+2}
    const void *vl = nullptr;     // left vectors (plain form)
    int ld = 1, r0 = 0, rb = 1;
    bool first = false, doubling = false;
    double a1 = 2.0, b1 = -1.0;   // T_{n+1} = a1 A T_n + b1 T_{n-1}
    double2 *part0 = nullptr, *part1 = nullptr, *part2 = nullptr;  // ring slots
};
int launch_tile2(const TileArgs &a, hipStream_t s);
// Per-patch contiguous copies of the halo rows' matrix slots (after the
// plan's values, and with_cols its columns, are in place).
int launch_halo_gather(const jkpm_tiles &T, bool with_cols, hipStream_t s);

// ---- two steps per launch by a wavefront over row chunks (wave.hip) ----
struct WaveArgs {
    const Ell *H = nullptr;
    const void *xprev = nullptr;  // T_{n-1}
    const void *xcur = nullptr;   // T_n
    void *out1 = nullptr;         // T_{n+1} (a buffer distinct from the inputs)
    void *out2 = nullptr;         // T_{n+2}
    const void *vl = nullptr;     // left vectors (plain form)
    int rb = 16, d = 2;           // vectors; blocks between a block's two phases
    bool doubling = false;
    double2 *part1 = nullptr, *part2 = nullptr;  // ring slots of steps n+1, n+2
    int nchunks = 0, rpc = 0, nblk = 0;
    const WaveDep *deps = nullptr;
    const int *tail_off = nullptr, *tail_rows = nullptr, *tail_dep = nullptr;
    unsigned *flags = nullptr, *err = nullptr;
    unsigned epoch = 1;
    int spin_limit = 1 << 16;
    int nfix = 0, nfold = 0;  // fix-up groups and the chunks they fold into
    const int *fix_group_off = nullptr, *fix_rows = nullptr;
    const int *fold_chunk = nullptr, *fold_off = nullptr;
    double2 *fixpart = nullptr;  // [nfix][n_left][rb]
};
// Largest number of fix-up rows one work-group of the fix-up launch takes.
constexpr int kWaveFixGroup = 64;
// Rows one wave covers per block for a complex64 group of rb vectors (0 when
// the group size is not supported).
int wave_rows_per_wave(int rb);
int launch_wave2(const WaveArgs &a, hipStream_t s);
int wave_max_blocks_per_cu(int rb, int k, int d, bool doubling, int *out);

// ---- Chebyshev series / commutator recursion (series.hip) ----
size_t series_workspace(const Ell &H, int R, bool comm);
// out (if set) = sum_m c_m X_m and tab[m] (if set, commutator only) = c_m X_m,
// X = T_m(H) v, or the commutator recursion C_m when V is set; c: host
// complex128 [M] as (re, im) pairs.
int run_series(const Ell &H, const Ell *V, const void *v, int R, int M, const double *c,
               void *out, void *tab, void *ws, int nchunks, hipStream_t s);

// ---- Kubo double moments (kubo.hip) ----
int kubo_gemm_slices(long K, int M, const DeviceInfo &dev);
// A second recursion run on the same slot loads as a dual Kubo step:
// out = alpha*H*x + beta*prev (prev unused when beta == 0).
struct KuboLeft {
    const void *x = nullptr;
    const void *prev = nullptr;
    void *out = nullptr;
    float alpha = 1.f, beta = 0.f;
};
// out = alpha*H*x + beta*prev (prev unused when beta == 0); when Vop is set,
// also wout = Vop*x on the same gathered x (Vop shares H's columns), and when
// left is set (Vop required) the left recursion as well.
int launch_kubo_step(const Ell &H, const Ell *Vop, const void *x, const void *prev, void *out,
                     void *wout, int R, float alpha, float beta, int nchunks, hipStream_t s,
                     const KuboLeft *left = nullptr);
int launch_kubo_gemm(const float2 *L, const float2 *Rk, long K, int M, int nslices, float2 *part,
                     hipStream_t s);
int launch_kubo_reduce(const float2 *part, int nslices, int M, double scale, double2 *mu,
                       hipStream_t s);
int launch_kubo_sigma(const double2 *mu, int M, const double *e, int n_e, int kind, double bound,
                      double2 *sigma, hipStream_t s);

// ---- complex128 vector helpers (blas.hip) ----
int zvec(int op, long n, double2 a, double2 b, const void *x, void *y, const void *w, void *z,
         hipStream_t s);
size_t zdot_workspace();
int zdot(long n, double2 a, const void *x, const void *y, int conj_x, double2 *out,
         double2 *part, hipStream_t s);

// ---- input checks (checks.hip) ----
// Hermiticity of h's stored values (computed once per value change, cached on
// the handle; synchronises s, refused under capture while unknown).
int hermiticity(const jkpm_ell *h, hipStream_t s, double *abs_err, double *max_abs);
// JKPM_ERR_ARG naming `what` unless max |a_ij - conj(a_ji)| <=
// JKPM_HERMITIAN_RTOL * max |a_ij|.
int require_hermitian(const jkpm_ell *h, hipStream_t s, const char *what);
// Index of the first non-finite double of x[0..n) and the count (-1, 0 when
// all are finite).  Synchronises s.
int count_nonfinite(const double *x, long n, long long *first_bad, long long *count,
                    hipStream_t s);
// JKPM_FAULT=alloc in the environment: handle creation reports an
// allocation failure (a test hook for the callers' error paths).
bool inject_alloc_failure();
// After a value upload: packs the real parts of H's values into H->rval
// (allocated on first use) and sets H->real_values when every imaginary part
// is exactly 0.  Synchronises s.  JKPM_REAL_VALUES=0 keeps every handle on
// the complex values.
int update_real_values(Ell *H, hipStream_t s);
// Fills H->rval with NaN (asynchronous on s): the values stopped being real,
// and a HIP graph captured while they were replays the real-value kernels,
// which then yield NaN moments instead of moments of the stale real parts.
int poison_real_copy(Ell *H, hipStream_t s);

}  // namespace jkpm

// The opaque handle of the C-ABI: the device matrix plus the ownership of its
// column array, shared by handles made with jkpm_ell_create_like (freed when
// the last handle that uses it is destroyed).
struct jkpm_ell {
    jkpm::Ell e;
    std::shared_ptr<int> col_owner;
    // Hermiticity of the stored values (checks.hip): max |a_ij - conj(a_ji)|
    // and max |a_ij|, or -1 while unknown (reset whenever the values change).
    mutable double herm_err = -1.0, herm_max = -1.0;
    // values_real: every stored imaginary part is 0 (checked at each value
    // upload); use_real: the caller lets the step kernels read the packed
    // real copy (jkpm_ell_use_real_values); e.real_values = both.
    bool values_real = false, use_real = true;
    bool values_imag = false;  // the values are purely imaginary (Ell::imag_values)
};

// A tile plan of a handle: rows renumbered into patches, the matrix in that
// order, and the patch tables (all device arrays owned by the plan).
struct jkpm_wave {
    const jkpm_ell *h = nullptr;  // borrowed: must outlive the plan
    int R = 0, rb = 0, lag = 1, d = 2, nchunks = 0, rpc = 0, nblk = 0, nfix = 0, nfold = 0;
    bool doubling = false;
    long long tail = 0, deferred = 0;
    unsigned epoch = 0;           // flags hold epoch << 12 | blocks done
    jkpm::WaveDep *deps = nullptr;
    int *tail_off = nullptr, *tail_rows = nullptr, *tail_dep = nullptr;
    unsigned *flags = nullptr;    // [nchunks][4] per-wave progress
    unsigned *err = nullptr;      // error bits of the launches since the last check
    int *fix_group_off = nullptr, *fix_rows = nullptr, *fold_chunk = nullptr, *fold_off = nullptr;
    double2 *fixpart = nullptr;   // [groups][n_left][R]
};

struct jkpm_tiles {
    const jkpm_ell *h = nullptr;  // borrowed: must outlive the plan
    int n = 0, k = 0, vtype = 0, rb = 1, cap = 0, npatch = 0, max_ext = 0;
    long long halo_rows = 0;
    int *perm = nullptr, *inv = nullptr;
    void *val = nullptr;
    int *col = nullptr;
    short *lcol = nullptr;
    int *pstart = nullptr, *hoff = nullptr, *hrows = nullptr;
    void *hval = nullptr;  // [k][halo_rows] halo rows' values, patch by patch
    int *hcol = nullptr;   // [k][halo_rows] their columns
};
