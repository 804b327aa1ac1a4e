/*
 * CPU oracle for the jclsquant Chebyshev/KPM path: a plain C99 + OpenMP
 * restatement of the reference algorithm in complex128.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check results
 * (or the timed CPU baseline) -- never as the product path.  The product path
 * is the HIP library in minijclsquant_amd/ and fails loudly without it.
 *
 * Each function names the reference lines it follows (paths relative to the
 * reference repository root).  Arithmetic order follows the reference where it
 * is defined there (row loop, slot loop j=0..K-1, y*=b before the slot sum,
 * (y*a)*conj(x) in the dot); the order of the OpenMP reduction is thread-count
 * dependent in the reference as well.
 *
 * Complex numbers are passed as interleaved (re, im) doubles.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef double complex zc;

static int nthr(int n) { return n > 0 ? n : 1; }

/* jclsquant/cython_modules/Extra_funs.pyx:182-219 (data_indices_c_0):
 * row-major ELL, row i in slots [i*K, i*K+K); padding repeats the row's last
 * column index with value 0 and valid 0. */
void orc_csr_to_ell(int len_row, int n_rows, const int *indptr, const zc *a_data,
                    const int *a_indices, int *indices, zc *data, int *valid,
                    int n_threads) {
    int i;
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
    for (i = 0; i < n_rows; ++i) {
        int start = indptr[i], end = indptr[i + 1];
        int rem = end - start, k;
        for (k = 0; k < rem; ++k) {
            indices[(long)i * len_row + k] = a_indices[start + k];
            data[(long)i * len_row + k] = a_data[start + k];
            valid[(long)i * len_row + k] = 1;
        }
        for (k = rem; k < len_row; ++k) {
            /* reference reads row_i_indices[remaining-1]; an empty row would
             * read out of bounds there -- here it pads with the row index. */
            indices[(long)i * len_row + k] = rem > 0 ? a_indices[end - 1] : i;
            data[(long)i * len_row + k] = 0;
            valid[(long)i * len_row + k] = 0;
        }
    }
}

/* jclsquant/cython_modules/blas_funs.pyx:15-19 (scal_c): x <- a*x */
void orc_scal(const double *a_ri, zc *x, long size, int n_threads) {
    zc a = a_ri[0] + I * a_ri[1];
    long i;
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
    for (i = 0; i < size; ++i) x[i] = a * x[i];
}

/* jclsquant/cython_modules/blas_funs.pyx:64-71 (aAxby_c):
 * y_i <- b*y_i ; y_i += a*data[iK+j]*x[indices[iK+j]] for j < K. */
void orc_aAxby(const double *a_ri, const double *b_ri, const zc *x, zc *y,
               const zc *data, const int *indices, int len_row, int n_rows,
               int n_threads) {
    zc a = a_ri[0] + I * a_ri[1], b = b_ri[0] + I * b_ri[1];
    int i;
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
    for (i = 0; i < n_rows; ++i) {
        zc yi = y[i] * b;
        int j;
        for (j = 0; j < len_row; ++j) {
            long idx = (long)i * len_row + j;
            yi += a * data[idx] * x[indices[idx]];
        }
        y[i] = yi;
    }
}

/* jclsquant/cython_modules/blas_funs.pyx:83-88 (vdot_c): sum_i y_i*a*conj(x_i) */
static zc vdot_c(zc a, const zc *x, const zc *y, long size, int n_threads) {
    double sr = 0.0, si = 0.0;
    long i;
#pragma omp parallel for schedule(static) reduction(+ : sr, si) num_threads(nthr(n_threads))
    for (i = 0; i < size; ++i) {
        zc t = y[i] * a * conj(x[i]);
        sr += creal(t);
        si += cimag(t);
    }
    return sr + I * si;
}

void orc_vdot(const double *a_ri, const zc *x, const zc *y, long size, int n_threads,
              double *out) {
    zc r = vdot_c(a_ri[0] + I * a_ri[1], x, y, size, n_threads);
    out[0] = creal(r);
    out[1] = cimag(r);
}

/* jclsquant/cython_modules/blas_funs.pyx:97-102 (dot_c): sum_i y_i*a*x_i */
void orc_dot(const double *a_ri, const zc *x, const zc *y, long size, int n_threads,
             double *out) {
    zc a = a_ri[0] + I * a_ri[1];
    double sr = 0.0, si = 0.0;
    long i;
#pragma omp parallel for schedule(static) reduction(+ : sr, si) num_threads(nthr(n_threads))
    for (i = 0; i < size; ++i) {
        zc t = y[i] * a * x[i];
        sr += creal(t);
        si += cimag(t);
    }
    out[0] = sr;
    out[1] = si;
}

/* jclsquant/cython_modules/blas_funs.pyx:113-117 (vec_mul_c): z = a*y*x */
void orc_vec_mul(const double *a_ri, const zc *x, const zc *y, zc *z, long size,
                 int n_threads) {
    zc a = a_ri[0] + I * a_ri[1];
    long i;
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
    for (i = 0; i < size; ++i) z[i] = a * y[i] * x[i];
}

/* jclsquant/cython_modules/Extra_funs.pyx:352-361 (diagonal_modifier_c):
 * data_new[s] = diag[row] + data[s] on valid diagonal slots, else data[s]. */
void orc_diagonal_modifier(const zc *data, const int *indices, const int *valid,
                           zc *data_new, const zc *diag, int len_row, long len_data,
                           int n_threads) {
    long i;
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
    for (i = 0; i < len_data; ++i) {
        long row = i / len_row;
        if (row == indices[i] && valid[i] == 1)
            data_new[i] = diag[row] + data[i];
        else
            data_new[i] = data[i];
    }
}

/* Shared recursion body of jclsquant/recurrence_functions.py:9-111:
 * T_0 = v_r ; T_1 = A v_r (dot(1,0)) ; for m >= 2: T_0 <- 2 A T_1 - T_0, swap.
 * mu_l[m] = <v_l | T_m> for each of the n_left left vectors. */
static int rec_core(int M, const zc *data, const int *indices, int len_row, int n,
                    const zc *vr, int n_left, const zc *const *vl, zc *const *mu,
                    int n_threads) {
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    const double two[2] = {2.0, 0.0}, mone[2] = {-1.0, 0.0};
    zc *t0, *t1, *tmp;
    int m, l;
    if (M < 1) return 0;
    t0 = (zc *)malloc(sizeof(zc) * (size_t)n);
    t1 = (zc *)calloc((size_t)n, sizeof(zc));
    if (!t0 || !t1) { free(t0); free(t1); return -1; }
    memcpy(t0, vr, sizeof(zc) * (size_t)n);
    orc_aAxby(one, zero, vr, t1, data, indices, len_row, n, n_threads);
    for (l = 0; l < n_left; ++l) {
        mu[l][0] = vdot_c(1.0, vl[l], vr, n, n_threads);
        if (M > 1) mu[l][1] = vdot_c(1.0, vl[l], t1, n, n_threads);
    }
    for (m = 2; m < M; ++m) {
        orc_aAxby(two, mone, t1, t0, data, indices, len_row, n, n_threads);
        tmp = t0; t0 = t1; t1 = tmp;
        for (l = 0; l < n_left; ++l) mu[l][m] = vdot_c(1.0, vl[l], t1, n, n_threads);
    }
    free(t0);
    free(t1);
    return 0;
}

/* jclsquant/recurrence_functions.py:9-40 (rec_A_tab): mu_n = <v|T_n(A)|v> */
int orc_rec_A_tab(int M, const zc *data, const int *indices, int len_row, int n,
                  const zc *v, zc *mu, int n_threads) {
    const zc *vl[1] = {v};
    zc *out[1] = {mu};
    return rec_core(M, data, indices, len_row, n, v, 1, vl, out, n_threads);
}

/* jclsquant/recurrence_functions.py:44-74 (rec_A_tab2v): mu_n = <v_l|T_n(A)|v_r> */
int orc_rec_A_tab2v(int M, const zc *data, const int *indices, int len_row, int n,
                    const zc *v_l, const zc *v_r, zc *mu, int n_threads) {
    const zc *vl[1] = {v_l};
    zc *out[1] = {mu};
    return rec_core(M, data, indices, len_row, n, v_r, 1, vl, out, n_threads);
}

/* jclsquant/recurrence_functions.py:76-111 (rec_A_tab2v_2):
 * mu1_n = <v_l|T_n v_r>, mu2_n = <v_r|T_n v_r> */
int orc_rec_A_tab2v_2(int M, const zc *data, const int *indices, int len_row, int n,
                      const zc *v_l, const zc *v_r, zc *mu1, zc *mu2, int n_threads) {
    const zc *vl[2] = {v_l, v_r};
    zc *out[2] = {mu1, mu2};
    return rec_core(M, data, indices, len_row, n, v_r, 2, vl, out, n_threads);
}

/* jclsquant/kernel_and_moments.py:19-24 (JacksonKernel): g_n for n < M, no
 * halving of g_0. */
void orc_jackson(int M, double *g) {
    int n;
    double q = M_PI / (M + 1.0);
    double cot = 1.0 / tan(q);
    for (n = 0; n < M; ++n)
        g[n] = ((M - n + 1.0) * cos(q * n) + sin(q * n) * cot) / (M + 1.0);
}

/* jclsquant/kernel_and_moments.py:72-78 (moments_delta):
 * [1, 2cos(n acos E)...] / (pi sqrt(1-E^2)) */
void orc_moments_delta(double e, int M, double *out) {
    int n;
    double th = acos(e), den = M_PI * sqrt(1.0 - e * e);
    out[0] = 1.0 / den;
    for (n = 1; n < M; ++n) out[n] = 2.0 * cos(n * th) / den;
}

/* jclsquant/observables.py:39-52 (energy loop of kpm_dos):
 * dos_i = Re dot(1, moments_delta(E~_i, M) * g, mu), E~_i = E_i/h - c/h with
 * h = (b1-b0)/2, c = (b1+b0)/2.  kernel_flag 1 = Jackson, 0 = none. */
void orc_dos(const zc *mu, int M, const double *energies, int n_e, double b0, double b1,
             int kernel_flag, double *dos) {
    double *g = (double *)malloc(sizeof(double) * (size_t)M);
    double *d = (double *)malloc(sizeof(double) * (size_t)M);
    double h = 0.5 * (b1 - b0), c = 0.5 * (b1 + b0);
    int i, n;
    if (!g || !d) { free(g); free(d); return; }
    if (kernel_flag) orc_jackson(M, g);
    else for (n = 0; n < M; ++n) g[n] = 1.0;
    for (i = 0; i < n_e; ++i) {
        double es = energies[i] / h - c / h;
        double s = 0.0;
        orc_moments_delta(es, M, d);
        for (n = 0; n < M; ++n) s += creal((d[n] * g[n]) * mu[n]);
        dos[i] = s;
    }
    free(g);
    free(d);
}

/* Kubo double moments mu_mn = <r| v T_m(A) v T_n(A) |r> for Hermitian A and v
 * (no reference counterpart; the reference's nearest double expansion is
 * jclsquant/cuda_cython/cuda/obv.cu:562-644).  Built from the stacks
 * L_m = T_m(A)(v r) and Rk_n = v T_n(A) r:  mu_mn = <L_m | Rk_n>.
 * Output mu is row-major [M][M] (m rows, n columns). */
int orc_mu_nm(int M, const zc *h_data, const int *h_idx, int h_k, const zc *v_data,
              const int *v_idx, int v_k, int n, const zc *r, zc *mu, int n_threads) {
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    const double two[2] = {2.0, 0.0}, mone[2] = {-1.0, 0.0};
    zc *L = (zc *)calloc((size_t)M * n, sizeof(zc));
    zc *Rk = (zc *)calloc((size_t)M * n, sizeof(zc));
    zc *T = (zc *)calloc((size_t)M * n, sizeof(zc));
    int m, k;
    if (!L || !Rk || !T) { free(L); free(Rk); free(T); return -1; }
    /* left stack: T_m(A) (v r) */
    orc_aAxby(one, zero, r, L, v_data, v_idx, v_k, n, n_threads);
    if (M > 1) orc_aAxby(one, zero, L, L + n, h_data, h_idx, h_k, n, n_threads);
    for (m = 2; m < M; ++m) {
        memcpy(L + (size_t)m * n, L + (size_t)(m - 2) * n, sizeof(zc) * (size_t)n);
        orc_aAxby(two, mone, L + (size_t)(m - 1) * n, L + (size_t)m * n, h_data, h_idx,
                  h_k, n, n_threads);
    }
    /* right stack: v T_n(A) r */
    memcpy(T, r, sizeof(zc) * (size_t)n);
    if (M > 1) orc_aAxby(one, zero, r, T + n, h_data, h_idx, h_k, n, n_threads);
    for (m = 2; m < M; ++m) {
        memcpy(T + (size_t)m * n, T + (size_t)(m - 2) * n, sizeof(zc) * (size_t)n);
        orc_aAxby(two, mone, T + (size_t)(m - 1) * n, T + (size_t)m * n, h_data, h_idx,
                  h_k, n, n_threads);
    }
    for (m = 0; m < M; ++m)
        orc_aAxby(one, zero, T + (size_t)m * n, Rk + (size_t)m * n, v_data, v_idx, v_k, n,
                  n_threads);
    for (m = 0; m < M; ++m)
        for (k = 0; k < M; ++k)
            mu[(size_t)m * M + k] = vdot_c(1.0, L + (size_t)m * n, Rk + (size_t)k * n, n,
                                           n_threads);
    free(L);
    free(Rk);
    free(T);
    return 0;
}

/* Reverse-mode gradient of  J = Re sum_n c_n <r|T_n(A)|r>  with respect to the
 * row-major ELL values of A (no reference counterpart).  Adjoint recursion
 * lambda_k = conj(c_k) r + 2 A^H lambda_{k+1} - lambda_{k+2}  (k = M-1 .. 1),
 * evaluated with A^H = A (Hermitian A), and
 * G[i,s] = lambda_1,i conj(T_0,j) + 2 sum_{n=1}^{M-2} lambda_{n+1,i} conj(T_n,j),
 * j = indices[i*K+s], i.e. G = dJ/dRe(a) + i dJ/dIm(a).
 * This FP64 restatement keeps every T_n in memory (M*N complex). */
int orc_grad_values(int M, const zc *data, const int *indices, int len_row, int n,
                    const zc *r, const double *c_ri, zc *grad, int n_threads) {
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    const double two[2] = {2.0, 0.0}, mone[2] = {-1.0, 0.0};
    zc *T = (zc *)calloc((size_t)M * n, sizeof(zc));
    zc *la = (zc *)calloc((size_t)n, sizeof(zc));   /* lambda_{k+1} */
    zc *lb = (zc *)calloc((size_t)n, sizeof(zc));   /* lambda_{k+2} -> lambda_k */
    zc *tmp;
    int m, k;
    long s;
    if (!T || !la || !lb) { free(T); free(la); free(lb); return -1; }
    memset(grad, 0, sizeof(zc) * (size_t)n * len_row);
    if (M < 2) { free(T); free(la); free(lb); return 0; }
    memcpy(T, r, sizeof(zc) * (size_t)n);
    orc_aAxby(one, zero, r, T + n, data, indices, len_row, n, n_threads);
    for (m = 2; m < M; ++m) {
        memcpy(T + (size_t)m * n, T + (size_t)(m - 2) * n, sizeof(zc) * (size_t)n);
        orc_aAxby(two, mone, T + (size_t)(m - 1) * n, T + (size_t)m * n, data, indices,
                  len_row, n, n_threads);
    }
    /* la = lambda_{k+1}, lb = lambda_{k+2}; start k = M-1 with both zero */
    for (k = M - 1; k >= 1; --k) {
        zc ck = conj(c_ri[2 * k] + I * c_ri[2 * k + 1]);
        long i;
        /* lb <- 2 A la - lb  (= 2 A lambda_{k+1} - lambda_{k+2}) then + ck r */
        orc_aAxby(two, mone, la, lb, data, indices, len_row, n, n_threads);
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
        for (i = 0; i < n; ++i) lb[i] += ck * r[i];
        tmp = la; la = lb; lb = tmp;   /* la = lambda_k, lb = lambda_{k+1} */
        {
            const zc *tk = T + (size_t)(k - 1) * n;
            double f = (k == 1) ? 1.0 : 2.0;
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
            for (s = 0; s < (long)n * len_row; ++s) {
                long row = s / len_row;
                grad[s] += f * la[row] * conj(tk[indices[s]]);
            }
        }
    }
    free(T);
    free(la);
    free(lb);
    return 0;
}

/* jclsquant/recurrence_functions.py:114-143 (rec_A_vec):
 * out = sum_{m<M} c_m T_m(A) v, with the reference's update order
 * (A_kpm = c_0 v + c_1 T_1, then axpy2(c_m, 1, T_m, A_kpm)). */
int orc_rec_A_vec(int M, const zc *data, const int *indices, int len_row, int n,
                  const double *c_ri, const zc *v, zc *out, int n_threads) {
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    const double two[2] = {2.0, 0.0}, mone[2] = {-1.0, 0.0};
    zc *t0 = (zc *)malloc(sizeof(zc) * (size_t)n), *t1 = (zc *)calloc((size_t)n, sizeof(zc)), *tmp;
    int m;
    long i;
    if (!t0 || !t1) { free(t0); free(t1); return -1; }
    memcpy(t0, v, sizeof(zc) * (size_t)n);
    orc_aAxby(one, zero, v, t1, data, indices, len_row, n, n_threads);
    {
        zc c0 = c_ri[0] + I * c_ri[1], c1 = M > 1 ? c_ri[2] + I * c_ri[3] : 0.0;
        for (i = 0; i < n; ++i) out[i] = c0 * v[i] + (M > 1 ? c1 * t1[i] : 0.0);
    }
    for (m = 2; m < M; ++m) {
        zc cm = c_ri[2 * m] + I * c_ri[2 * m + 1];
        orc_aAxby(two, mone, t1, t0, data, indices, len_row, n, n_threads);
        tmp = t0; t0 = t1; t1 = tmp;
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
        for (i = 0; i < n; ++i) out[i] = cm * t1[i] + 1.0 * out[i];
    }
    free(t0);
    free(t1);
    return 0;
}

/* jclsquant/recurrence_functions.py:146-196 (rec_com_A_vec) and :202-250
 * (rec_com_A_vec_tab): commutator recursion C_0 = 0, C_1 = V v,
 * C_m = 2 A C_{m-1} - C_{m-2} + 2 V T_{m-1}, next to T_m(A) v.  With tab
 * (M*n) non-NULL, tab[m] = c_m C_m; with out non-NULL, out = sum c_m C_m. */
int orc_rec_com_A_vec(int M, const zc *a_data, const zc *v_data, const int *indices,
                      int len_row, int n, const double *c_ri, const zc *v, zc *out, zc *tab,
                      int n_threads) {
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    const double two[2] = {2.0, 0.0}, mone[2] = {-1.0, 0.0};
    zc *t0 = (zc *)malloc(sizeof(zc) * (size_t)n), *t1 = (zc *)calloc((size_t)n, sizeof(zc));
    zc *k0 = (zc *)calloc((size_t)n, sizeof(zc)), *k1 = (zc *)calloc((size_t)n, sizeof(zc));
    zc *tmp;
    int m;
    long i;
    if (!t0 || !t1 || !k0 || !k1) { free(t0); free(t1); free(k0); free(k1); return -1; }
    memcpy(t0, v, sizeof(zc) * (size_t)n);
    orc_aAxby(one, zero, v, t1, a_data, indices, len_row, n, n_threads);
    orc_aAxby(one, zero, v, k1, v_data, indices, len_row, n, n_threads);
    if (out) {
        zc c1 = M > 1 ? c_ri[2] + I * c_ri[3] : 0.0;
        for (i = 0; i < n; ++i) out[i] = M > 1 ? c1 * k1[i] : 0.0;
    }
    if (tab) {
        zc c0 = c_ri[0] + I * c_ri[1];
        for (i = 0; i < n; ++i) tab[i] = c0 * k0[i];
        if (M > 1) {
            zc c1 = c_ri[2] + I * c_ri[3];
            for (i = 0; i < n; ++i) tab[(size_t)n + i] = c1 * k1[i];
        }
    }
    for (m = 2; m < M; ++m) {
        zc cm = c_ri[2 * m] + I * c_ri[2 * m + 1];
        orc_aAxby(two, mone, k1, k0, a_data, indices, len_row, n, n_threads);
        {
            const double pone[2] = {1.0, 0.0};
            orc_aAxby(two, pone, t1, k0, v_data, indices, len_row, n, n_threads);
        }
        tmp = k0; k0 = k1; k1 = tmp;
        orc_aAxby(two, mone, t1, t0, a_data, indices, len_row, n, n_threads);
        tmp = t0; t0 = t1; t1 = tmp;
        if (out)
            for (i = 0; i < n; ++i) out[i] = cm * k1[i] + 1.0 * out[i];
        if (tab)
            for (i = 0; i < n; ++i) tab[(size_t)m * n + i] = cm * k1[i];
    }
    free(t0);
    free(t1);
    free(k0);
    free(k1);
    return 0;
}

/* jclsquant/cython_modules/Extra_funs.pyx:430-535 (moments_Gmin_real_c,
 * moments_Gmas_real_c, moments_delta_c, table_G_delta_c) for an explicit list
 * of rescaled energies (the reference uses linspace(-alpha, alpha, 2M)):
 *   bigmas[m][n][i] = (Gmas_m(E_i) g_m / B) * (delta_n(E_i) g_n)
 *   bigmin[m][n][i] = -(delta_m(E_i) g_m) * (Gmin_n(E_i) g_n / B)
 * Gmas_m = 2 sin(-m acos E)/s, Gmin_m = -2 sin(m acos E)/s (index 0: 0),
 * delta_m = [1, 2cos(m acos E)...]/(pi s), s = sqrt(1-E^2) (1e-16 if 0). */
void orc_table_G_delta(int M, const double *E, int n_e, double bound, double *bigmas,
                       double *bigmin) {
    double *g = (double *)malloc(sizeof(double) * (size_t)M);
    double *gmas = (double *)malloc(sizeof(double) * (size_t)M);
    double *gmin = (double *)malloc(sizeof(double) * (size_t)M);
    double *dl = (double *)malloc(sizeof(double) * (size_t)M);
    int i, m, n;
    if (!g || !gmas || !gmin || !dl) { free(g); free(gmas); free(gmin); free(dl); return; }
    orc_jackson(M, g);
    for (i = 0; i < n_e; ++i) {
        double e = E[i];
        double s = sqrt(1.0 - e * e), f;
        if (s == 0) s = 1e-16;
        f = M_PI * sqrt(1.0 - e * e);
        if (f == 0) f = 1e-16;
        gmin[0] = 0.0;
        gmas[0] = 0.0;
        dl[0] = 1.0 / f;
        for (m = 1; m < M; ++m) {
            gmin[m] = -2.0 * sin(m * acos(e)) / s;
            gmas[m] = 2.0 * sin(-m * acos(e)) / s;
            dl[m] = 2.0 * cos(m * acos(e)) / f;
        }
        for (m = 0; m < M; ++m) {
            gmin[m] = gmin[m] * g[m] / bound;
            gmas[m] = gmas[m] * g[m] / bound;
            dl[m] = dl[m] * g[m];
        }
        for (n = 0; n < M; ++n)
            for (m = 0; m < M; ++m) {
                size_t o = ((size_t)m * M + n) * (size_t)n_e + i;
                bigmas[o] = gmas[m] * dl[n];
                bigmin[o] = -dl[m] * gmin[n];
            }
    }
    free(g);
    free(gmas);
    free(gmin);
    free(dl);
}
