/*
 * jclkpm -- C-ABI of the MI355X (gfx950) Chebyshev/KPM library
 * (libjclkpm_hip.so, built from minijclsquant_amd/csrc/).
 *
 * Conventions
 *   - Every entry point returns an int status: JKPM_OK (0) or a JKPM_ERR_*
 *     code; jkpm_last_error() gives a message for the calling thread.
 *   - Complex numbers are interleaved (re, im) pairs: double for JKPM_C128,
 *     float for JKPM_C64.
 *   - Host ELL input is the reference's row-major layout: slot j of row i at
 *     [i*len_row + j] (jclsquant/cython_modules/Extra_funs.pyx:182-219).
 *     On the device the handle keeps a slot-major copy ([j*N + i]) scaled by
 *     `scale` (1/B for the spectral rescale) and cast to the value type.
 *   - Device vectors of a batch of R vectors are row-interleaved:
 *     element (row i, vector r) at [i*R + r].
 *   - Moments are complex128, laid out [n_left][M][R].
 *   - `stream` is a hipStream_t (NULL = default stream).  Device-pointer
 *     entry points are asynchronous on that stream and allocate nothing;
 *     scratch memory comes in through `workspace`.
 *   - A handle may be used by one thread at a time.
 */
#ifndef JCLKPM_H
#define JCLKPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JKPM_OK 0
#define JKPM_ERR_ARG 1         /* bad argument (size, pointer, unsupported R) */
#define JKPM_ERR_HIP 2         /* a HIP runtime call failed */
#define JKPM_ERR_ALLOC 3       /* device allocation failed */
#define JKPM_ERR_WORKSPACE 4   /* workspace smaller than required */
/* Test hook: with JKPM_FAULT=alloc in the environment, handle creation
 * (jkpm_ell_create and the entry points built on it) returns JKPM_ERR_ALLOC
 * before allocating, so callers can exercise their error paths. */

#define JKPM_C128 0
#define JKPM_C64 1

#define JKPM_KERNEL_NONE 0
#define JKPM_KERNEL_JACKSON 1

typedef struct jkpm_ell jkpm_ell;

/* Library version (major*10000 + minor*100 + patch). */
int jkpm_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *jkpm_last_error(void);

/* ---------------------------------------------------------------------
 * Reference drop-ins: host pointers in/out, synchronous, complex128.
 * ------------------------------------------------------------------- */

/* Replaces extern "C" rec_A_tab (jclsquant/cuda_cython/cuda/obv.h:20,
 * jclsquant/cuda_cython/cuda/recursions.cu:17-98): A_vec[m] = <v|T_m(A)|v>,
 * m < M, for the row-major ELL (data, indices, len_row) of N rows.
 * data must already carry the 1/B rescale, as in the reference. */
int rec_A_tab(const double *data, const int *indices, int len_row, int N,
              double *A_vec, const double *v, int M);

/* Replaces extern "C" rec_A_tab_2 (obv.h:22, recursions.cu:180-263):
 * A_vec[m] = <v_l|T_m(A)|v_r>. */
int rec_A_tab_2(const double *data, const int *indices, int len_row, int N,
                double *A_vec, const double *v_l, const double *v_r, int M);

/* ---------------------------------------------------------------------
 * Device-resident ELL handle.
 * ------------------------------------------------------------------- */

/* Uploads a row-major complex128 ELL (host or device pointers, chosen by
 * src_on_device), multiplies the values by `scale` in complex128 (as
 * jclsquant/modifiers.py:24-26 does with fl(1/B)), casts them to value_type
 * and stores them slot-major.  Replaces the per-call H2D upload of
 * recursions.cu:24-28 with one upload per matrix. */
int jkpm_ell_create(const double *data, const int *indices, int n_rows, int len_row,
                    double scale, int value_type, int src_on_device, void *stream,
                    jkpm_ell **out);

/* Device-side ingest of a CSR matrix (indptr int32 [n_rows+1], indices
 * int32 [nnz], data complex128 [nnz], nnz = indptr[n_rows]; host or device
 * pointers, chosen by src_on_device).  Row i's entries go, in CSR order, to
 * slots 0..len_i-1; slots len_i..K-1 repeat the row's last column with value
 * 0, as jclsquant/ell_matrix.py:187-200 over data_indices_c_0
 * (jclsquant/cython_modules/Extra_funs.pyx:182-219) does on the host (an
 * empty row pads with its own index).  K = len_row, or the longest row when
 * len_row is 0 (1 when every row is empty); a row longer than len_row, an
 * indptr that decreases or leaves [0, nnz], or a column outside [0, n_rows)
 * is an error.  Values are multiplied by `scale` in complex128 and cast to
 * value_type, as in jkpm_ell_create.  The host-side row-major ELL is never
 * built.  Device arrays must hold n_rows+1 and nnz entries (only indptr's
 * last entry is read before the scan).  Synchronises `stream`. */
int jkpm_ell_create_csr(const int *indptr, const int *indices, const double *data, int n_rows,
                        int len_row, double scale, int value_type, int src_on_device,
                        void *stream, jkpm_ell **out);

/* Row-major int32 copy of the column indices (ell_matrix.indices,
 * jclsquant/ell_matrix.py:119-258) to host or device memory
 * (dst_on_device).  Synchronises `stream`. */
int jkpm_ell_get_indices(const jkpm_ell *h, int *indices, int dst_on_device, void *stream);

int jkpm_ell_destroy(jkpm_ell *h);

int jkpm_ell_info(const jkpm_ell *h, int *n_rows, int *len_row, int *value_type);

/* Replaces the values (same indices) from a row-major complex128 array. */
int jkpm_ell_set_values(jkpm_ell *h, const double *data, double scale,
                        int src_on_device, void *stream);

/* Hermiticity of the stored matrix A (the handle's scaled values):
 * *abs_err = max over slots of |a_ij - conj(a_ji)| (the slots of a row that
 * share a column summed, so ELL padding counts once), *max_abs = max |a_ij|.
 * Computed on the device once per value change and cached on the handle;
 * synchronises `stream` when it computes, and is refused under stream capture
 * until computed.  The doubling form (JKPM_MOMENTS_DOUBLING), the Kubo double
 * moments and the adjoint entry points require abs_err <=
 * JKPM_HERMITIAN_RTOL * max_abs and return JKPM_ERR_ARG otherwise.  No
 * reference counterpart: the reference's H_graphene 'ELL' matrices are
 * non-Hermitian by a 1e-6 cross term (jclsquant/hams.py:160-162), which its
 * plain recursion tolerates and these forms do not. */
#define JKPM_HERMITIAN_RTOL 1e-10
int jkpm_ell_hermiticity(const jkpm_ell *h, double *abs_err, double *max_abs, void *stream);

/* *is_real = 1 when every stored imaginary part of the handle's values is
 * exactly 0 (checked on the device at each value upload; a Peierls update
 * makes the values complex).  The Chebyshev step kernels of such a matrix
 * read a packed copy of the real parts, 4 (complex64) or 8 (complex128)
 * bytes per slot instead of 8 or 16; results are those of the complex
 * arithmetic with zero imaginary parts.  JKPM_REAL_VALUES=0 in the
 * environment turns the check and the copy off (every handle then reports 0
 * here and in jkpm_ell_is_imag, and runs on its complex values).  A HIP graph captured while the values
 * were real keeps the real-value kernels: after a value upload or Peierls
 * update that makes them complex, capture it again (the stale copy is
 * overwritten with NaN, so such a replay gives NaN moments).  No reference
 * counterpart (the reference stores complex128 values always,
 * jclsquant/ell_matrix.py:187-200). */
int jkpm_ell_is_real(const jkpm_ell *h, int *is_real);
/* *is_imag = 1 when every stored real part is exactly 0 and some imaginary
 * part is not (checked at the same uploads): the velocity -i d H of a real
 * H.  The handle's packed copy then holds the imaginary parts, which the
 * Kubo stack step reads for v (4 bytes per slot instead of 8).  No reference
 * counterpart. */
int jkpm_ell_is_imag(const jkpm_ell *h, int *is_imag);
/* on = 0: the kernels of this handle read its complex values even when they
 * are all real or all imaginary (the complex128 storage of the reference, or
 * an A/B comparison); on = 1 (the default) lets them read the packed copy. */
int jkpm_ell_use_real_values(jkpm_ell *h, int on);

/* Left vectors as phases.  A complex64 handle's plain recursion with one
 * left set (jkpm_moments with v_l1 = NULL, jkpm_moments_ex, the batch and
 * autotune paths) packs v_l into 32-bit phases (2 pi = 2^32) in its
 * workspace at the start of each call, on the device.  When every element
 * is reproduced within 2^-21 per component by its phase (random-phase
 * vectors e^{i phi}, jclsquant/observables.py:19-21), the step kernels read
 * the phases, 4 bytes per element instead of 8, and rebuild e^{i phi} in
 * registers; otherwise they read v_l as given.  The choice is made on the
 * device, so a captured graph follows the vectors of each replay.  The
 * workspace sizes include the phase buffer (N * R * 4 bytes) while this is
 * on (jkpm_moments_noalloc keeps it in its scratch).  on = 0 turns it off
 * for the handle (query the workspace size again afterwards).  No reference
 * counterpart (the reference reads v_l as complex128 in every dot,
 * jclsquant/recurrence_functions.py:31-38). */
int jkpm_ell_use_left_phases(jkpm_ell *h, int on);
/* The packing step on its own: phases[i] of complex64 v[i], i < n, and *bad
 * (device int) = 0 when every element is reproduced within 2^-21 per
 * component, else nonzero.  Asynchronous on stream. */
int jkpm_phase_pack(const void *v, long long n, unsigned *phases, int *bad, void *stream);

/* Non-finite entries (NaN, +-Inf) among the n doubles of device array x
 * (e.g. the 2*M*R doubles of a moment array): *first_bad = the smallest
 * such index or -1, *count = how many.  Synchronises `stream`; refused under
 * stream capture.  No reference counterpart (the reference returns NaN
 * moments silently); growing moments usually mean a spectral bound below
 * the largest |eigenvalue|. */
int jkpm_check_finite(const double *x, long long n, long long *first_bad, long long *count,
                      void *stream);

/* Row-major complex128 copy of the values in the units they were uploaded in
 * (stored values / scale), to host or device memory (dst_on_device).  The
 * device side of ell_matrix.data (jclsquant/ell_matrix.py:119-258) after a
 * device-side modifier.  Synchronises `stream`. */
int jkpm_ell_get_values(const jkpm_ell *h, double *data, int dst_on_device, void *stream);

/* Peierls substitution on the device:
 *   dst(i,s) = a * src(i,s) * exp(i (ax dx(i,s) + ay dy(i,s)))
 * with (dx, dy) = (Re, Im) of the values of `bonds` (the bond vectors, made
 * with jkpm_ell_create_like(src, dx + i dy, ...)); a, ax, ay complex
 * (re, im); dst shares src's slot pattern and value type (dst may be src)
 * and takes src's scale.  Replaces the copy + rewrite of H per time step of
 * jclsquant/cuda_cython/cuda/obv.cu:335-341 with light_modifier_circular_
 * packed_c (modifiers.cu:25-68, alpha = i ax, beta = i ay) and the CPU
 * light_modifier_cython (jclsquant/cython_modules/Extra_funs.pyx:327-342).
 * Asynchronous on stream. */
int jkpm_ell_peierls(jkpm_ell *dst, const jkpm_ell *src, const jkpm_ell *bonds, double a_re,
                     double a_im, double ax_re, double ax_im, double ay_re, double ay_im,
                     void *stream);

/* y <- a*A*x + b*y on R interleaved device vectors of the handle's value type.
 * Replaces aAxby_device (jclsquant/cuda_cython/cuda/blas.cu:29-43) and the
 * CPU aAxby (jclsquant/cython_modules/blas_funs.pyx:64-75).  x must not alias y.
 * a and b are (re, im). */
int jkpm_aAxby(const jkpm_ell *h, const double *a, const double *b, const void *x,
               void *y, int R, void *stream);

/* ---------------------------------------------------------------------
 * Complex128 vector helpers (device vectors of n elements, asynchronous on
 * `stream`; a and b are host (re, im) pairs).  They replace the OpenMP
 * loops of jclsquant/cython_modules/blas_funs.pyx that the reference's
 * namespace exports:
 *   JKPM_VOP_AXPBY  y <- a x + b y   scal (:15-25, x = y, b = 0),
 *                                    axpy (:31-43, b = 1), axpy2 (:47-59)
 *   JKPM_VOP_MUL    z <- a y x       vec_mul  (:113-121)
 *   JKPM_VOP_MUL2   y <- a x y       vec_mul2 (:125-137)
 *   JKPM_VOP_MUL3   z <- a y x w     vec_mul3 (:145-157)
 * Unused pointers may be NULL.
 * ------------------------------------------------------------------- */
#define JKPM_VOP_AXPBY 0
#define JKPM_VOP_MUL 1
#define JKPM_VOP_MUL2 2
#define JKPM_VOP_MUL3 3
int jkpm_zvec(int op, long long n, const double *a, const double *b, const void *x, void *y,
              const void *w, void *z, void *stream);

/* out = a * sum_i conj(x_i) y_i (conj_x = 1: vdot, blas_funs.pyx:80-91) or
 * a * sum_i x_i y_i (conj_x = 0: dot, :94-105), complex128, FP64 sums in a
 * fixed order; out is a device complex128 scalar.  workspace: at least
 * jkpm_zdot_workspace bytes of device memory. */
int jkpm_zdot_workspace(size_t *bytes);
int jkpm_zdot(long long n, const double *a, const void *x, const void *y, int conj_x,
              double *out, void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------
 * Chebyshev moments (the hot path).
 * ------------------------------------------------------------------- */

/* Workspace bytes jkpm_moments needs for R vectors and n_left left sets. */
int jkpm_moments_workspace(const jkpm_ell *h, int R, int n_left, size_t *bytes);

/* mu[l][m][r] = <v_l{l}[:,r] | T_m(A) v_r[:,r]>, m < M, complex128, for
 * l < n_left (n_left = 1 if v_l1 is NULL, else 2).  The recursion runs in the
 * handle's value type; the dots accumulate in FP64 with a fixed reduction
 * order (bitwise reproducible).  Generalises rec_A_tab / rec_A_tab_2 /
 * rec_A_tab2v_2 (jclsquant/recurrence_functions.py:9-111) to R vectors that
 * share each pass over the matrix.  v_r, v_l0, v_l1: device [N][R]. */
int jkpm_moments(const jkpm_ell *h, const void *v_r, const void *v_l0, const void *v_l1,
                 int R, int M, double *mu, void *workspace, size_t workspace_bytes,
                 void *stream);

/* jkpm_moments without a workspace argument; it allocates nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its scratch (the two rotating vector batches, the
 * moment-partial ring and, for complex64, the left vectors' phases; the
 * layout of jkpm_moments_workspace) is a library-owned device array of
 * JKPM_NOALLOC_SCRATCH_BYTES (128 MiB) that every GPU the library runs on
 * holds until the process ends; a matrix and R whose layout needs more
 * (jkpm_moments_workspace) get JKPM_ERR_WORKSPACE.  Nothing in the scratch
 * is read before it is written in the same call, so it keeps the last
 * call's contents harmlessly; calls that overlap on one GPU (two streams or
 * threads) share it and race.  The step-time profiler
 * (jkpm_profile_enable) does not record these calls.
 * jkpm_moments_noalloc_setup looks up the scratch's address on the current
 * device ahead of the first call (the first call does it otherwise). */
#define JKPM_NOALLOC_SCRATCH_BYTES (128u << 20)
int jkpm_moments_noalloc_setup(void);
int jkpm_moments_noalloc(const jkpm_ell *h, const void *v_r, const void *v_l0, const void *v_l1,
                         int R, int M, double *mu, void *stream);

/* ---------------------------------------------------------------------
 * Launch shape of the Chebyshev step kernel and its autotuner (no reference
 * counterpart: the reference launches <<<ceil(N/1024), 1024>>> always,
 * jclsquant/cuda_cython/cuda/recursions.cu:19-20).
 * ------------------------------------------------------------------- */

/* variant = unroll | nontemporal << 1 | shape << 2 (0..11), plus 12 for
 * non-temporal stores of T_{n+1} (12..23), plus 24 for non-temporal loads of
 * the matrix values and columns (24..47): unroll = two rows per lane,
 * nontemporal = non-temporal loads of the read-once vector streams
 * (T_{n-1}, v_l), shape = complex64 vectors per lane 2 << shape (lanes per
 * row shrink to match; c64 groups of 4+ vectors).
 * Variants exist for ELL widths 4 and 5; other shapes use 0.  chunk_mult:
 * work-groups per CU (1..64, default 8).  flags = JKPM_MOMENTS_DOUBLING
 * addresses the doubling form's own setting; 0 the plain recursion's. */
#define JKPM_STEP_VARIANTS 48
int jkpm_ell_set_launch(jkpm_ell *h, int variant, int chunk_mult, int flags);
int jkpm_ell_get_launch(const jkpm_ell *h, int *variant, int *chunk_mult, int flags);

/* Vectors per recursion pass (0, the default: all R at once).  With
 * 0 < vectors < R, a moments call runs the whole recursion for each group of
 * `vectors` columns in turn (copied to a contiguous [N][vectors] array in the
 * workspace, the moments written to their columns of mu), so that one pass's
 * vectors and the matrix fit in the 256 MB Infinity Cache; the workspace
 * sizes cover both the passes and all R vectors at once (jkpm_moments_keep
 * never runs in passes).  flags = JKPM_MOMENTS_DOUBLING addresses the
 * doubling form, 0 the plain recursion.  autotune.tuned chooses it with the
 * launch shape.  No reference counterpart (the reference runs one vector per
 * recursion, jclsquant/recurrence_functions.py:9-38). */
int jkpm_ell_set_pass(jkpm_ell *h, int vectors, int flags);
int jkpm_ell_get_pass(const jkpm_ell *h, int *vectors, int flags);

/* Times `steps` fused steps of jkpm_moments (flags = 0) or of the doubling
 * form of jkpm_moments_ex (flags = JKPM_MOMENTS_DOUBLING) on v_r (device
 * [N][R]) for each variant x chunk_mult in {2, 4, 8, 16, 32}, keeps the
 * fastest on the handle for that form and returns its mean step time in
 * microseconds.  Synchronises `stream`. */
int jkpm_autotune_workspace(const jkpm_ell *h, int R, int steps, int flags, size_t *bytes);
int jkpm_autotune(jkpm_ell *h, const void *v_r, int R, int steps, int flags, void *workspace,
                  size_t workspace_bytes, void *stream, double *best_us);

/* ---------------------------------------------------------------------
 * Batches of disorder realizations (no reference counterpart: reference
 * callers loop over realizations and vectors serially).
 * ------------------------------------------------------------------- */

typedef struct jkpm_batch jkpm_batch;

/* Describes the handle h as a block-diagonal batch of G realizations with
 * rows [row_offsets[g], row_offsets[g+1]) (host array of G+1, strictly
 * increasing, 0 .. N); column indices of realization g stay inside its rows.
 * The rows are cut into work-group chunks proportional to each realization's
 * size that never straddle two realizations, about (CU count x the handle's
 * chunk_mult) of them; the plain recursion and the doubling form each get a
 * table from their own setting as it is at creation (tune the handle first).
 * h must outlive the batch. */
int jkpm_batch_create(const jkpm_ell *h, int G, const long long *row_offsets,
                      jkpm_batch **out);

/* As jkpm_batch_create, but realization g with chunk_rows[g] > 0 is cut into
 * chunks of exactly chunk_rows[g] rows (the last one shorter); 0 keeps the
 * proportional cut.  Giving a lattice's row stride (the index distance of
 * its vertical neighbours) makes the work-groups of consecutive chunks, which
 * run side by side on one XCD, read each other's neighbour rows at the same
 * time, so each T_n row is fetched from HBM once instead of once per use.
 * No reference counterpart (the reference runs one realization per call). */
int jkpm_batch_create_strided(const jkpm_ell *h, int G, const long long *row_offsets,
                              const int *chunk_rows, jkpm_batch **out);
int jkpm_batch_destroy(jkpm_batch *b);
int jkpm_batch_workspace(const jkpm_batch *b, int R, int n_left, size_t *bytes);

/* jkpm_moments for every realization at once, one kernel launch per
 * Chebyshev step for the whole batch: mu[l][g][m][r] (complex128). */
int jkpm_batch_moments(const jkpm_batch *b, const void *v_r, const void *v_l0,
                       const void *v_l1, int R, int M, double *mu, void *workspace,
                       size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------
 * Chebyshev series application.
 * ------------------------------------------------------------------- */

int jkpm_series_workspace(const jkpm_ell *h, int R, int comm, size_t *bytes);

/* v_op == NULL: out = sum_{m<M} c_m T_m(A) v -- replaces rec_A_vec
 *   (jclsquant/recurrence_functions.py:114-143) and the GPU rec_A
 *   (jclsquant/cuda_cython/cuda/recursions.cu:267-317).
 * v_op != NULL (made with jkpm_ell_create_like(h, ...)): the commutator
 *   recursion C_0 = 0, C_1 = V v, C_m = 2 A C_{m-1} - C_{m-2} + 2 V T_{m-1};
 *   out = sum_m c_m C_m -- rec_com_A_vec (recurrence_functions.py:146-196,
 *   GPU rec_comm_A_vec, recursions.cu:319-388) -- and/or the table
 *   tab[m] = c_m C_m -- rec_com_A_vec_tab (recurrence_functions.py:202-250).
 * c: host complex128 [M] (re, im); v, out: device [N][R]; tab: device
 * [M][N][R]; all of the handle's value type.  Asynchronous on stream. */
int jkpm_series(const jkpm_ell *h, const jkpm_ell *v_op, const void *v, int R, int M,
                const double *c, void *out, void *tab, void *workspace,
                size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------
 * Single-left-set moments mu_m = <r|T_m(A)|r> with options.
 * ------------------------------------------------------------------- */

/* Doubling form: for Hermitian A only (checked: jkpm_ell_hermiticity), the
 * M moments come from
 * S = max(1, M/2) Chebyshev steps through mu_2k = 2<T_k|T_k> - mu_0 and
 * mu_2k+1 = 2<T_k|T_{k+1}> - mu_1 (Weisse et al., Rev. Mod. Phys. 78, 275
 * (2006)); each step reads no left vectors.  Same moments as the plain
 * recursion up to rounding.  No reference counterpart: the reference runs
 * M-1 steps with a dot against the stored vector
 * (jclsquant/recurrence_functions.py:31-38). */
#define JKPM_MOMENTS_DOUBLING 1

/* Either a handle (mu [M][R]) or a batch (mu [G][M][R]); v_r device [N][R]. */
int jkpm_moments_ex_workspace(const jkpm_ell *h, const jkpm_batch *b, int R, int M, int flags,
                              size_t *bytes);
int jkpm_moments_ex(const jkpm_ell *h, const jkpm_batch *b, const void *v_r, int R, int M,
                    int flags, double *mu, void *workspace, size_t workspace_bytes,
                    void *stream);

/* ---------------------------------------------------------------------
 * Two Chebyshev steps per launch over LDS patches (no reference
 * counterpart: the reference makes one SpMV launch per step and gathers
 * every T_n from device memory, jclsquant/cuda_cython/cuda/recursions.cu:
 * 52-91).  The rows are renumbered into patches grown breadth-first on the
 * matrix graph; one work-group keeps a patch's T_{n+1}, with its halo
 * recomputed, in LDS and makes T_{n+2} from it, so a launch reads the
 * matrix, T_{n-1}, T_n and v_l once for two steps.  Moments are unchanged by
 * the renumbering.
 * ------------------------------------------------------------------- */

typedef struct jkpm_tiles jkpm_tiles;

/* Builds the plan of handle h for up to R vectors per call: patches of at
 * most max_rows rows of patch + halo (<= 0: as many as the LDS tile of
 * min(R, 16) vectors holds).  Synchronous; h must outlive the plan.  The plan
 * copies h's values: after jkpm_ell_set_values call jkpm_tiles_refresh. */
int jkpm_tiles_create(const jkpm_ell *h, int R, int max_rows, jkpm_tiles **out);
int jkpm_tiles_destroy(jkpm_tiles *t);
int jkpm_tiles_info(const jkpm_tiles *t, int *npatch, long long *halo_rows, int *max_ext,
                    int *cap);
int jkpm_tiles_refresh(jkpm_tiles *t, void *stream);

/* Host only (no device needed): builds the plan of a slot-major [K][N]
 * column array for tiles of `cap` rows (patch + halo) and checks it fully
 * (bijective renumbering, halos outside their patch, every patch-local
 * column naming the right row); returns the patch count, total halo rows
 * and the largest patch + halo. */
int jkpm_tiles_plan_check(int n_rows, int len_row, const int *col_slot_major, int cap,
                          int *npatch, long long *halo_rows, int *max_ext);

/* mu[m][r] = <r|T_m(A)|r> (complex128 [M][R]) for v_r device [N][R] in the
 * handle's value type, flags 0 or JKPM_MOMENTS_DOUBLING (Hermitian A).
 * Asynchronous on stream; allocates nothing. */
int jkpm_tiles_workspace(const jkpm_tiles *t, int R, int M, int flags, size_t *bytes);
int jkpm_tiles_moments(const jkpm_tiles *t, const void *v_r, int R, int M, int flags,
                       double *mu, void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------
 * Two Chebyshev steps per launch by a wavefront over the row chunks: the
 * work-groups of one XCD hand T_{n+1} to each other through its L2 behind
 * per-wave progress flags (bounded waits, cooperative launch), so a launch
 * reads the matrix, T_{n-1}, T_n and v_l once for two steps and recomputes
 * nothing.  Rows whose columns reach a chunk of another XCD are finished by a
 * small second launch.  No reference counterpart: the reference makes one
 * host-driven launch per step (jclsquant/cuda_cython/cuda/recursions.cu:
 * 78-92).  Complex64 handles, R = 8 or 16, ELL width 4 or 5.
 * ------------------------------------------------------------------- */

typedef struct jkpm_wave jkpm_wave;

/* Plan of handle h for R vectors in the form `flags` (0: mu = <r|T_m|r>;
 * JKPM_MOMENTS_DOUBLING: the doubling form, Hermitian A), over the row chunks
 * of the handle's launch shape for that form (jkpm_ell_set_launch).  Refused
 * with JKPM_ERR_ARG when the columns do not fit (a block of rows reading more
 * than four chunks, or blocks further ahead than the kernel's lag of one) or
 * the grid cannot be co-resident.  Synchronous; h must outlive the plan. */
int jkpm_wave_create(const jkpm_ell *h, int R, int flags, jkpm_wave **out);
int jkpm_wave_destroy(jkpm_wave *w);
/* nchunks: work-groups per launch; lag: blocks between a block's two steps;
 * tail_rows: rows each work-group finishes after its main loop (columns on
 * the same XCD but further ahead than the lag allows, e.g. the periodic wrap
 * of a lattice row); fixup_rows: rows finished by the small second launch
 * (columns on another XCD). */
int jkpm_wave_info(const jkpm_wave *w, int *nchunks, int *lag, long long *tail_rows,
                   long long *fixup_rows);

/* Host only (no device needed): builds the plan of a slot-major [K][N]
 * column array for nchunks chunks of rpc rows, blocks of 4 * rows_per_wave
 * rows on 8 XCDs and main-loop rows reading at most max_lag blocks ahead,
 * checks it independently, and returns its lag and tail and fix-up row
 * counts. */
int jkpm_wave_plan_check(int n_rows, int len_row, const int *col_slot_major, int nchunks,
                         int rpc, int rows_per_wave, int max_lag, int *lag, long long *tail_rows,
                         long long *fixup_rows);

/* mu = complex128 [M][R] moments of v_r (device [N][R] complex64) in the
 * plan's form, as jkpm_moments_ex computes them.  Runs on `stream`, then
 * synchronises it to read the kernel's error flag: JKPM_ERR_HIP when a
 * work-group ran on another XCD than the chunk mapping assumes or a wait gave
 * up (mu is then invalid).  Refused under stream capture (the flags' epoch
 * is a launch argument).  The workspace holds four vector buffers.  The
 * plan's progress flags are shared by its calls: one call at a time per
 * plan (calls that overlap on one plan race). */
int jkpm_wave_workspace(const jkpm_wave *w, int R, int M, size_t *bytes);
int jkpm_wave_moments(jkpm_wave *w, const void *v_r, int R, int M, double *mu, void *workspace,
                      size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------
 * HIP graphs: capture the launches of asynchronous entry points (moments,
 * batch moments, DOS, ...) issued on `stream` between begin and end, then
 * replay them with one launch.  Pointers and sizes are frozen at capture.
 * ------------------------------------------------------------------- */

typedef struct jkpm_graph jkpm_graph;

int jkpm_graph_begin(void *stream);
int jkpm_graph_end(void *stream, jkpm_graph **out);
int jkpm_graph_launch(const jkpm_graph *g, void *stream);
int jkpm_graph_destroy(jkpm_graph *g);

/* ---------------------------------------------------------------------
 * DOS reconstruction.
 * ------------------------------------------------------------------- */

/* dos[i] = sum_n g_n delta_n(e_i) Re(mean_r mu[m][r]) with
 * delta_n(e) = (2 - [n==0]) cos(n acos e) / (pi sqrt(1 - e^2)),
 * g = Jackson (kernel = JKPM_KERNEL_JACKSON) or 1, for rescaled energies
 * |e_i| < 1.  mu: device complex128 [M][R]; e, dos: device float64 [n_e].
 * Replaces the host loop of jclsquant/observables.py:49-52 over
 * JacksonKernel / moments_delta (jclsquant/kernel_and_moments.py:19-24,72-78).
 * workspace: >= 8*M bytes. */
int jkpm_dos(const double *mu, int M, int R, const double *e, int n_e, int kernel,
             double *dos, void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------
 * Gradients (no reference counterpart: the reference computes no
 * derivatives).
 * ------------------------------------------------------------------- */

/* Workspace bytes of jkpm_moments_grad for segment length seg (0:
 * ceil(sqrt(M-1))): 2*ceil((M-1)/seg) checkpoint vectors, 2*seg segment
 * vectors, 4 rolling vectors (batches of R) and the N*K slot gradient.
 * seg < 0 keeps every T_0 .. T_{M-2} instead (M-1 vectors, no checkpoints,
 * no recomputation) with ceil(sqrt(M-1)) lambda vectors: one forward pass
 * fewer, for (M-1) vector batches of memory (34 GB for N = 262,144, R = 8,
 * complex128, M = 1024). */
int jkpm_grad_workspace(const jkpm_ell *h, int R, int M, int seg, size_t *bytes);

/* Adjoint of jkpm_moments (one left set, v_l = v_r = r): given
 * grad_mu[m][r] = dL/dRe(mu) + i dL/dIm(mu) (device complex128 [M][R]),
 * writes grad_values = dL/dRe(data) + i dL/dIm(data) for the row-major
 * complex128 values the handle was created from (device, N*K).  Reverse
 * Chebyshev recursion lambda_k = g_k r + 2 A lambda_{k+1} - lambda_{k+2}
 * (A^H = A: Hermitian matrices) with checkpoints every seg steps and
 * recomputation of each segment (seg >= 0), or with every T kept (seg < 0).
 * r: device [N][R] of the handle's type. */
int jkpm_moments_grad(const jkpm_ell *h, const void *r, int R, int M, const double *grad_mu,
                      int seg, double *grad_values, void *workspace, size_t workspace_bytes,
                      void *stream);

/* Moments that keep their Chebyshev vectors: as jkpm_moments with one left
 * set v_l = v_r, and T_0 .. T_{M-1} left in `stack` (device, max(M,2)
 * vector batches of N*R elements of the handle's type, T_n at n*N*R).
 * T_{n+1} is written to its own slot instead of over T_{n-1}: the same HBM
 * traffic per step.  For the adjoint without recomputation
 * (jkpm_moments_grad_kept).  No reference counterpart (the reference has no
 * gradients). */
int jkpm_moments_keep(const jkpm_ell *h, const void *v_r, int R, int M, double *mu,
                      void *stack, size_t stack_bytes, void *workspace, size_t workspace_bytes,
                      void *stream);

/* Workspace bytes of jkpm_moments_grad_kept: ceil(sqrt(M-1)) lambda
 * vectors, 2 rolling vectors and the N*K slot gradient. */
int jkpm_grad_kept_workspace(const jkpm_ell *h, int R, int M, size_t *bytes);

/* jkpm_moments_grad over the vectors a jkpm_moments_keep call left in
 * `stack` (same h, r, R, M): the adjoint steps and folds only, no forward
 * pass.  No reference counterpart. */
int jkpm_moments_grad_kept(const jkpm_ell *h, const void *r, int R, int M, const double *grad_mu,
                           const void *stack, double *grad_values, void *workspace,
                           size_t workspace_bytes, void *stream);

/* Transpose of jkpm_dos: grad_mu[n][r] = (1/R) g_n (2-[n==0])
 * sum_i w[i] cos(n acos e_i) / (pi sqrt(1-e_i^2)) (imaginary part 0), for
 * w = dL/d dos.  grad_mu: device complex128 [M][R]. */
int jkpm_dos_grad(const double *w, int M, int R, const double *e, int n_e, int kernel,
                  double *grad_mu, void *stream);

/* ---------------------------------------------------------------------
 * Kubo double moments (no reference counterpart: the reference's nearest
 * double expansion is position_operator, jclsquant/cuda_cython/cuda/
 * obv.cu:562-644, and its coefficient table table_G_delta,
 * jclsquant/cython_modules/Extra_funs.pyx:479-535, is never called).
 * ------------------------------------------------------------------- */

/* A handle with new values on the slot pattern of `pattern` (shares its
 * column array): e.g. the velocity v_x = -i H .* dx of
 * jclsquant/modifiers.py:56-79, whose slots are H's slots. */
int jkpm_ell_create_like(const jkpm_ell *pattern, const double *data, double scale,
                         int value_type, int src_on_device, void *stream, jkpm_ell **out);

/* Workspace bytes of jkpm_kubo_mu_nm: two complex64 stacks of M*N*R,
 * two recursion vectors and the FP32 split-K partials. */
int jkpm_kubo_workspace(const jkpm_ell *H, int R, int M, size_t *bytes);

/* mu[m][n] = (1/R) sum_r <r| v T_m(H) v T_n(H) |r>, complex128 [M][M], for
 * Hermitian H and v (both JKPM_C64 handles, v made with jkpm_ell_create_like(H)),
 * r: device complex64 [N][R].  Stacks L_m = T_m(H)(v r), Rk_n = v T_n(H) r in
 * complex64; mu = L^H Rk on FP32 MFMA (split-K) with an FP64 combine. */
int jkpm_kubo_mu_nm(const jkpm_ell *H, const jkpm_ell *V, const void *r, int R, int M,
                    double *mu, void *workspace, size_t workspace_bytes, void *stream);

/* sigma[i] = sum_mn (a_m b_n + c_m d_n) mu[m][n] (complex128) at rescaled
 * energies e[i] (|e| < 1), M <= 2048, with g = Jackson,
 * delta_m = (2-[m==0]) cos(m acos e)/(pi s), G_m = -2 sin(m acos e)/s,
 * s = sqrt(1-e^2):
 *   kind 0 (Kubo-Greenwood T=0):  a = b = g delta, c = d = 0, i.e.
 *          Tr[v delta(e-H) v delta(e-H)] * bound^2;
 *   kind 1 (the table_G_delta contraction, Extra_funs.pyx:430-535):
 *          a = d = g G / bound, b = g delta, c = -g delta. */
int jkpm_kubo_sigma(const double *mu, int M, const double *e, int n_e, int kind,
                    double bound, double *sigma, void *stream);

/* ---------------------------------------------------------------------
 * Measurement support.
 * ------------------------------------------------------------------- */

/* every > 0: the moment calls record HIP event pairs on their stream around
 * runs of up to 8 consecutive Chebyshev step launches, one run started every
 * `every` steps (not at a recursion's first step, and closed before any
 * reduction launch); the two-step tile calls record every every-th launch
 * alone (0 = off).  jkpm_profile_read sums the recorded spans (ms) and the
 * launches they cover.  reset != 0 clears the totals. */
int jkpm_profile_enable(int every);
int jkpm_profile_read(double *step_ms, long long *step_launches, int reset);

/* The same totals per kind of timed kernel: JKPM_PROF_STEP (the Chebyshev
 * step launches above; jkpm_profile_read reads this kind),
 * JKPM_PROF_KUBO_STEP (runs of up to 8 fused Kubo stack steps, one run
 * started every `every` steps), JKPM_PROF_KUBO_GEMM (each mu_nm GEMM
 * launch alone) and JKPM_PROF_WAVE (one wavefront launch of two steps with
 * its fix-up launch, one every `every` launches).  Nothing is recorded under
 * stream capture. */
#define JKPM_PROF_STEP 0
#define JKPM_PROF_KUBO_STEP 1
#define JKPM_PROF_KUBO_GEMM 2
#define JKPM_PROF_WAVE 3
#define JKPM_PROF_KINDS 4
int jkpm_profile_read_kind(int kind, double *ms, long long *launches, int reset);

/* Plain device copy dst <- src of `bytes` bytes (a multiple of 16), used to
 * measure the achievable HBM rate and to calibrate traffic counters. */
int jkpm_copy(void *dst, const void *src, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* JCLKPM_H */
