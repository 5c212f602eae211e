/*
 * ORACLE — test infrastructure only (tests/, the smoke check and bench.py's
 * CPU-baseline leg).  The shipped library never loads it.
 *
 * Host LAPACK driver: the reference's accuracy/throughput harness times
 * LAPACKE_zgeev('N','N') in an OpenMP parallel-for over matrices with a
 * sequential LAPACK (stream_test_with_MKL/get_eig_MKL.c:26-51,
 * compile_mkl.sh:1).  MKL is not available here; the same routine is taken
 * from the OpenBLAS build that SciPy ships (exports scipy_LAPACKE_*), opened
 * with dlopen at run time from a path the caller supplies.
 *
 *   lapack_open            dlopen + resolve symbols, force 1 BLAS thread
 *   lapack_zgeev_batched   zgeev (balanced, LAPACK default) per matrix
 *   lapack_zgeev_vectors_batched  the same with jobvl = jobvr = 'V'
 *   lapack_zhrd_hseqr_batched  zgehrd + zhseqr('E','N') without balancing:
 *                          the exact method of the GPU path, run by LAPACK
 *   lapack_zgeevx_rcond    eigenvalue condition numbers (RCONDE) of the
 *                          unbalanced matrix, for condition-scaled tolerances
 *   lapack_dgeev_batched   real dgeev (balanced, LAPACK default): wr, wi
 *   lapack_dhrd_hseqr_batched  dgehrd + dhseqr('E','N') without balancing
 *   lapack_dlanv2 / lapack_dlahqr  the LAPACK routines themselves (Fortran
 *                          interface), to pin the real restatement
 */
#include <complex.h>
#include <dlfcn.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex zc;
typedef int (*zgeev_t)(int, char, char, int, zc *, int, zc *, zc *, int, zc *, int);
typedef int (*zgehrd_t)(int, int, int, int, zc *, int, zc *);
typedef int (*zhseqr_t)(int, char, char, int, int, int, zc *, int, zc *, zc *, int);
typedef int (*zgebal_t)(int, char, int, zc *, int, int *, int *, double *);
typedef int (*zgeevx_t)(int, char, char, char, char, int, zc *, int, zc *, zc *, int, zc *, int,
                        int *, int *, double *, double *, double *, double *);
typedef int (*dgeev_t)(int, char, char, int, double *, int, double *, double *, double *, int, double *, int);
typedef int (*dgehrd_t)(int, int, int, int, double *, int, double *);
typedef int (*dhseqr_t)(int, char, char, int, int, int, double *, int, double *, double *, double *, int);
typedef void (*dlanv2_t)(double *, double *, double *, double *, double *, double *, double *, double *, double *,
                         double *);
typedef void (*dlahqr_t)(const int *, const int *, const int *, const int *, const int *, double *, const int *,
                         double *, double *, const int *, const int *, double *, const int *, int *);
typedef void (*setthr_t)(int);

static void *g_lib;
static zgeev_t p_zgeev;
static zgehrd_t p_zgehrd;
static zhseqr_t p_zhseqr;
static zgeevx_t p_zgeevx;
static zgebal_t p_zgebal;
static dgeev_t p_dgeev;
static dgehrd_t p_dgehrd;
static dhseqr_t p_dhseqr;
static dlanv2_t p_dlanv2;
static dlahqr_t p_dlahqr;

#define COLMAJOR 102

int lapack_open(const char *path)
{
    if (g_lib) return 0;
    g_lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!g_lib) return -1;
    p_zgeev = (zgeev_t)dlsym(g_lib, "scipy_LAPACKE_zgeev");
    p_zgehrd = (zgehrd_t)dlsym(g_lib, "scipy_LAPACKE_zgehrd");
    p_zhseqr = (zhseqr_t)dlsym(g_lib, "scipy_LAPACKE_zhseqr");
    p_zgeevx = (zgeevx_t)dlsym(g_lib, "scipy_LAPACKE_zgeevx");
    p_zgebal = (zgebal_t)dlsym(g_lib, "scipy_LAPACKE_zgebal");
    p_dgeev = (dgeev_t)dlsym(g_lib, "scipy_LAPACKE_dgeev");
    p_dgehrd = (dgehrd_t)dlsym(g_lib, "scipy_LAPACKE_dgehrd");
    p_dhseqr = (dhseqr_t)dlsym(g_lib, "scipy_LAPACKE_dhseqr");
    p_dlanv2 = (dlanv2_t)dlsym(g_lib, "scipy_dlanv2_");
    p_dlahqr = (dlahqr_t)dlsym(g_lib, "scipy_dlahqr_");
    setthr_t st = (setthr_t)dlsym(g_lib, "scipy_openblas_set_num_threads");
    if (st) st(1);
    if (!p_zgeev || !p_zgehrd || !p_zhseqr || !p_zgeevx || !p_zgebal) return -2;
    if (!p_dgeev || !p_dgehrd || !p_dhseqr || !p_dlanv2 || !p_dlahqr) return -2;
    return 0;
}

int lapack_max_threads(void) { return omp_get_max_threads(); }

/* eig[i + j*n]; info[j] = LAPACK INFO.  nthreads <= 0 means all. */
int lapack_zgeev_batched(int n, int batch, const zc *A, zc *eig, int *info, int nthreads)
{
    if (!p_zgeev) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        zc *a = (zc *)malloc(sizeof(zc) * (size_t)n * n);
        #pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < batch; ++b) {
            memcpy(a, A + (size_t)b * n * n, sizeof(zc) * (size_t)n * n);
            int r = p_zgeev(COLMAJOR, 'N', 'N', n, a, n, eig + (size_t)b * n, NULL, 1, NULL, 1);
            if (info) info[b] = r;
        }
        free(a);
    }
    return 0;
}

/* zgeev with eigenvectors (jobvl = jobvr = 'V', balanced as LAPACK does by
 * default): w[b*n + i], VL/VR batch x (n x n) column-major. */
int lapack_zgeev_vectors_batched(int n, int batch, const zc *A, zc *w, zc *VL, zc *VR, int *info, int nthreads)
{
    if (!p_zgeev) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        zc *a = (zc *)malloc(sizeof(zc) * (size_t)n * n);
        #pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < batch; ++b) {
            memcpy(a, A + (size_t)b * n * n, sizeof(zc) * (size_t)n * n);
            size_t off = (size_t)b * n * n;
            int r = p_zgeev(COLMAJOR, 'V', 'V', n, a, n, w + (size_t)b * n, VL + off, n, VR + off, n);
            if (info) info[b] = r;
        }
        free(a);
    }
    return 0;
}

int lapack_zhrd_hseqr_batched(int n, int batch, const zc *A, zc *eig, int *info, int nthreads)
{
    if (!p_zgehrd) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        zc *a = (zc *)malloc(sizeof(zc) * (size_t)n * n);
        zc *tau = (zc *)malloc(sizeof(zc) * (size_t)(n > 1 ? n : 1));
        zc zdummy = 0;
        #pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < batch; ++b) {
            memcpy(a, A + (size_t)b * n * n, sizeof(zc) * (size_t)n * n);
            int r = p_zgehrd(COLMAJOR, n, 1, n, a, n, tau);
            if (r == 0) r = p_zhseqr(COLMAJOR, 'E', 'N', n, 1, n, a, n, eig + (size_t)b * n, &zdummy, 1);
            if (info) info[b] = r;
        }
        free(a); free(tau);
    }
    return 0;
}

/* RCONDE of the unbalanced matrix (balanc='N'), eigenvalues in w. */
int lapack_zgeevx_rcond(int n, int batch, const zc *A, zc *w, double *rconde, int *info, int nthreads)
{
    if (!p_zgeevx) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        size_t nn = (size_t)n * n;
        zc *a = (zc *)malloc(sizeof(zc) * nn);
        zc *vl = (zc *)malloc(sizeof(zc) * nn), *vr = (zc *)malloc(sizeof(zc) * nn);
        double *scale = (double *)malloc(sizeof(double) * (size_t)n);
        double *rcondv = (double *)malloc(sizeof(double) * (size_t)n);
        #pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < batch; ++b) {
            memcpy(a, A + (size_t)b * nn, sizeof(zc) * nn);
            int ilo, ihi;
            double abnrm;
            int r = p_zgeevx(COLMAJOR, 'N', 'V', 'V', 'E', n, a, n, w + (size_t)b * n, vl, n, vr, n,
                             &ilo, &ihi, scale, &abnrm, rconde + (size_t)b * n, rcondv);
            if (info) info[b] = r;
        }
        free(a); free(vl); free(vr); free(scale); free(rcondv);
    }
    return 0;
}

/* zgebal on one column-major matrix (in place); ilo/ihi 1-based as LAPACK
 * returns them. */
int lapack_zgebal(char job, int n, zc *a, int *ilo, int *ihi, double *scale)
{
    if (!p_zgebal) return -1;
    return p_zgebal(COLMAJOR, job, n, a, n, ilo, ihi, scale);
}

/* wr/wi[i + j*n]; info[j] = LAPACK INFO. */
int lapack_dgeev_batched(int n, int batch, const double *A, double *wr, double *wi, int *info, int nthreads)
{
    if (!p_dgeev) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        double *a = (double *)malloc(sizeof(double) * (size_t)n * n);
        #pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < batch; ++b) {
            memcpy(a, A + (size_t)b * n * n, sizeof(double) * (size_t)n * n);
            int r = p_dgeev(COLMAJOR, 'N', 'N', n, a, n, wr + (size_t)b * n, wi + (size_t)b * n, NULL, 1, NULL, 1);
            if (info) info[b] = r;
        }
        free(a);
    }
    return 0;
}

int lapack_dhrd_hseqr_batched(int n, int batch, const double *A, double *wr, double *wi, int *info, int nthreads)
{
    if (!p_dgehrd) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        double *a = (double *)malloc(sizeof(double) * (size_t)n * n);
        double *tau = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n : 1));
        double zdummy = 0;
        #pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < batch; ++b) {
            memcpy(a, A + (size_t)b * n * n, sizeof(double) * (size_t)n * n);
            int r = p_dgehrd(COLMAJOR, n, 1, n, a, n, tau);
            if (r == 0)
                r = p_dhseqr(COLMAJOR, 'E', 'N', n, 1, n, a, n, wr + (size_t)b * n, wi + (size_t)b * n, &zdummy, 1);
            if (info) info[b] = r;
        }
        free(a); free(tau);
    }
    return 0;
}

/* LAPACK DLANV2 on (a, b, c, d) in place; out = rt1r, rt1i, rt2r, rt2i, cs, sn. */
int lapack_dlanv2(double *abcd, double *out)
{
    if (!p_dlanv2) return -1;
    p_dlanv2(&abcd[0], &abcd[1], &abcd[2], &abcd[3], &out[0], &out[1], &out[2], &out[3], &out[4], &out[5]);
    return 0;
}

/* LAPACK DLAHQR (WANTT = WANTZ = false, ILO = 1, IHI = n) on an upper
 * Hessenberg H (column-major, ld n, overwritten); returns INFO. */
int lapack_dlahqr(int n, double *H, double *wr, double *wi)
{
    if (!p_dlahqr) return -1;
    const int f = 0, one = 1;
    double zd = 0.0;
    int info = 0;
    p_dlahqr(&f, &f, &n, &one, &n, H, &n, wr, wi, &one, &n, &zd, &one, &info);
    return info;
}
