/*
 * ORACLE — test infrastructure only.  Built with -fsanitize=address,undefined
 * (`make -C oracle sanitize`) and run by tests/test_sanitizers.py: drives
 * every entry of the C restatements over edge sizes (n = 1, 2, 3, 64),
 * non-finite inputs, a zero sweep cap and balancing jobs, so that an out of
 * bounds access, a use of freed memory or undefined arithmetic in the
 * restatement (the checker of the GPU kernels) aborts the run.  No LAPACK is
 * loaded: the host-LAPACK driver is a thin dlopen wrapper.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

typedef double _Complex zc;

void oracle_zgeev_batched(int n, int batch, const zc *A, zc *eig, signed char *flag, int *sweeps, int itmax_override);
void oracle_zgeev_batched_dual(int n, int batch, const zc *A, zc *eig, signed char *flag, int *sweeps,
                               int itmax_override);
void oracle_zgeev_batched_bal(int n, int batch, const zc *A, zc *eig, signed char *flag, int itmax_override,
                              char job);
int oracle_zschur(int n, const zc *A, zc *T, zc *U, zc *w);
int oracle_zgeev_adjoint(int n, const zc *A, const zc *g, zc *G, zc *w);
int oracle_zgeev_vectors(int n, const zc *A, zc *w, zc *VL, zc *VR);
void oracle_dgeev_batched(int n, int batch, const double *A, double *wr, double *wi, int *info, int nthreads,
                          int itmax_override);

static double gauss(void)
{
    const double u1 = (rand() + 1.0) / (RAND_MAX + 2.0), u2 = (rand() + 1.0) / (RAND_MAX + 2.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

int main(void)
{
    srand(7);
    const int sizes[] = {1, 2, 3, 5, 10, 17, 64};
    int checked = 0;
    for (unsigned k = 0; k < sizeof sizes / sizeof sizes[0]; ++k) {
        const int n = sizes[k], batch = 3;
        const size_t nn = (size_t)n * n;
        zc *A = malloc(sizeof(zc) * nn * batch), *eig = malloc(sizeof(zc) * (size_t)n * batch);
        double *Ar = malloc(sizeof(double) * nn * batch), *wr = malloc(sizeof(double) * (size_t)n * batch),
               *wi = malloc(sizeof(double) * (size_t)n * batch);
        signed char *flag = malloc((size_t)batch);
        int *sweeps = malloc(sizeof(int) * batch), *info = malloc(sizeof(int) * batch);
        for (size_t e = 0; e < nn * batch; ++e) {
            A[e] = gauss() + I * gauss();
            Ar[e] = gauss();
        }
        A[nn + (nn - 1) / 2] = NAN;   /* matrix 1: non-finite */
        Ar[nn] = INFINITY;
        oracle_zgeev_batched(n, batch, A, eig, flag, sweeps, -1);
        oracle_zgeev_batched(n, batch, A, eig, flag, sweeps, 0);   /* sweep cap 0 */
        oracle_zgeev_batched_dual(n, batch, A, eig, flag, sweeps, -1);
        oracle_zgeev_batched_dual(n, batch, A, eig, flag, sweeps, 0);
        const char jobs[] = {'N', 'P', 'S', 'B'};
        for (int j = 0; j < 4; ++j) oracle_zgeev_batched_bal(n, batch, A, eig, flag, -1, jobs[j]);
        oracle_dgeev_batched(n, batch, Ar, wr, wi, info, 1, -1);
        oracle_dgeev_batched(n, batch, Ar, wr, wi, info, 1, 0);
        zc *T = malloc(sizeof(zc) * nn), *U = malloc(sizeof(zc) * nn), *G = malloc(sizeof(zc) * nn);
        zc *VL = malloc(sizeof(zc) * nn), *VR = malloc(sizeof(zc) * nn), *w = malloc(sizeof(zc) * (size_t)n);
        zc *g = malloc(sizeof(zc) * (size_t)n);
        for (int i = 0; i < n; ++i) g[i] = gauss() + I * gauss();
        oracle_zschur(n, A, T, U, w);
        oracle_zgeev_adjoint(n, A, g, G, w);
        oracle_zgeev_vectors(n, A, w, VL, VR);
        oracle_zgeev_vectors(n, A, w, NULL, VR);
        oracle_zgeev_vectors(n, A, w, VL, NULL);
        checked += batch;
        free(T); free(U); free(G); free(VL); free(VR); free(w); free(g);
        free(A); free(eig); free(Ar); free(wr); free(wi); free(flag); free(sweeps); free(info);
    }
    printf("sanitized oracle run: %d matrices, no report\n", checked);
    return 0;
}
