/*
 * ORACLE — test infrastructure only.  Nothing in the shipped library links,
 * loads or calls this file; only tests/, the smoke check and the CPU-baseline
 * leg of bench.py use it, and only as the checker.
 *
 * CPU restatement of the batched eigenvalue method that the reference library
 * implements on the GPU (tokura_zgeev_batched_library/src/tokura_zgeev_batched.cu:212-285:
 * Hessenberg reduction, then shifted QR, eigenvalues only, no balancing,
 * README.md:79).  The arithmetic follows LAPACK's textbook routines, because
 * parity is measured against host LAPACK zgeev:
 *
 *   oracle_zlarfg   <- LAPACK ZLARFG   (reflector with real beta); the
 *                      reference's version is tokura_zgehrd_MWB_kernel.cuh:28-133
 *   oracle_zgehd2   <- LAPACK ZGEHD2   (unblocked Householder Hessenberg);
 *                      reference tokura_zgehrd_MWB_kernel.cuh:159-287
 *   oracle_zlahqr   <- LAPACK ZLAHQR   (single-shift complex QR, Wilkinson and
 *                      exceptional shifts, Ahues-Kressner deflation, cap
 *                      30*max(10,n) sweeps); reference hseqr kernel
 *                      tokura_zhseqr_MWB_kernel.cuh:611-791 (whose real-part-only
 *                      deflation, top-left shift and n^4 cap are not reproduced)
 *   oracle_zlahqr_dual  the two-bulge QR of the GPU's eigenvalue path: two
 *                      ZLAHQR sweeps per iteration with the shifts of one
 *                      trailing 2x2 (no reference counterpart: the reference
 *                      chases one bulge per sweep)
 *   oracle_ztrevc   <- LAPACK ZTREVC   (right/left eigenvectors of the Schur form
 *                      by back-substitution), used by the adjoint checks
 *   oracle_zgeev_vectors  eigenvectors (Schur form, ZTREVC, back-transform,
 *                      ZGEEV's normalisation); no reference counterpart
 *                      (README.md:41-44: eigenvalues only)
 *   oracle_zgebal   <- LAPACK ZGEBAL   (3.11+: permutations, power-of-two
 *                      scaling), and oracle_zgeev_batched_bal = balancing then
 *                      the method; optional, the reference never balances
 *                      (README.md:79)
 *   norm_range_scale <- ZGEEV's scaling of A into [sqrt(safmin)/eps, its
 *                      inverse] before the reduction (a power of two here;
 *                      the reference does not scale, README.md:76-79)
 *
 * Flag codes per matrix (reference README.md:72 defines only 0 == success):
 *   0 converged, 1 iteration cap reached, 2 non-finite input.
 *
 * Pinned by tests/test_oracle.py against LAPACK zgehrd+zhseqr (the same
 * method, run by OpenBLAS), LAPACK zgeev (balanced, with and without
 * eigenvectors), LAPACK zgebal (bitwise), closed forms (triangular, 2x2,
 * companion of z^n - 1, Hermitian, diagonal eigenvectors), and by Hessenberg
 * invariants.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

typedef double _Complex zc;

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static inline double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }

static const double ULP = DBL_EPSILON;         /* dlamch('P') = eps*base */
static const double SAFMIN = DBL_MIN;          /* dlamch('S') */

static double dlapy3(double x, double y, double z)
{
    double xa = fabs(x), ya = fabs(y), za = fabs(z);
    double w = fmax(xa, fmax(ya, za));
    if (w == 0.0) return xa + ya + za;
    return w * sqrt((xa / w) * (xa / w) + (ya / w) * (ya / w) + (za / w) * (za / w));
}

/* Euclidean norm of m complex numbers with scaling (LAPACK DZNRM2 semantics). */
static double znrm2(int m, const zc *x)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < m; ++i) {
        double parts[2] = {creal(x[i]), cimag(x[i])};
        for (int p = 0; p < 2; ++p) {
            double a = fabs(parts[p]);
            if (a != 0.0) {
                if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
                else ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

/* ZLARFG: H^H * [alpha; x] = [beta; 0], H = I - tau*[1;v]*[1;v]^H, beta real.
 * On return *alpha = beta and x holds v. */
void oracle_zlarfg(int m, zc *alpha, zc *x, zc *tau)
{
    if (m <= 0) { *tau = 0; return; }
    double xnorm = znrm2(m - 1, x);
    double alphr = creal(*alpha), alphi = cimag(*alpha);
    if (xnorm == 0.0 && alphi == 0.0) { *tau = 0; return; }
    double beta = -copysign(dlapy3(alphr, alphi, xnorm), alphr);
    const double safmn = SAFMIN / ULP, rsafmn = 1.0 / safmn;
    int knt = 0;
    if (fabs(beta) < safmn) {
        do {
            ++knt;
            for (int i = 0; i < m - 1; ++i) x[i] *= rsafmn;
            beta *= rsafmn; alphi *= rsafmn; alphr *= rsafmn;
        } while (fabs(beta) < safmn && knt < 20);
        xnorm = znrm2(m - 1, x);
        beta = -copysign(dlapy3(alphr, alphi, xnorm), alphr);
    }
    *tau = CMPLX((beta - alphr) / beta, -alphi / beta);
    zc scal = 1.0 / (CMPLX(alphr, alphi) - beta);
    for (int i = 0; i < m - 1; ++i) x[i] *= scal;
    for (int j = 0; j < knt; ++j) beta *= safmn;
    *alpha = beta;
}

/* ZGEHD2 on an n x n column-major matrix (ilo = 1, ihi = n).  Entries below
 * the first subdiagonal are set to zero (the reflectors are returned in V/taus
 * when requested: V is n x n, column i holds the reflector of step i in rows
 * i+1..n-1 with V(i+1,i) = 1). */
void oracle_zgehd2(int n, zc *a, int lda, zc *V, zc *taus)
{
    zc *work = (zc *)malloc(sizeof(zc) * (size_t)(n > 0 ? n : 1));
    if (V) memset(V, 0, sizeof(zc) * (size_t)n * n);
    for (int i = 0; i + 1 < n; ++i) {
        /* reflector annihilating A(i+2:n-1, i) */
        int m = n - 1 - i;
        zc alpha = a[IDX(i + 1, i, lda)];
        zc tau;
        oracle_zlarfg(m, &alpha, &a[IDX(i + 2 < n ? i + 2 : n - 1, i, lda)], &tau);
        /* v = [1; A(i+2:n-1, i)] */
        zc *v = (zc *)malloc(sizeof(zc) * (size_t)m);
        v[0] = 1.0;
        for (int r = 1; r < m; ++r) v[r] = a[IDX(i + 1 + r, i, lda)];
        /* right: A(0:n-1, i+1:n-1) -= tau * (A v) v^H */
        for (int r = 0; r < n; ++r) {
            zc s = 0;
            for (int c = 0; c < m; ++c) s += a[IDX(r, i + 1 + c, lda)] * v[c];
            work[r] = s;
        }
        for (int c = 0; c < m; ++c) {
            zc f = tau * conj(v[c]);
            for (int r = 0; r < n; ++r) a[IDX(r, i + 1 + c, lda)] -= work[r] * f;
        }
        /* left: A(i+1:n-1, i+1:n-1) -= conj(tau) * v (v^H A) */
        for (int c = i + 1; c < n; ++c) {
            zc s = 0;
            for (int r = 0; r < m; ++r) s += conj(v[r]) * a[IDX(i + 1 + r, c, lda)];
            zc f = conj(tau) * s;
            for (int r = 0; r < m; ++r) a[IDX(i + 1 + r, c, lda)] -= v[r] * f;
        }
        a[IDX(i + 1, i, lda)] = alpha;
        for (int r = i + 2; r < n; ++r) a[IDX(r, i, lda)] = 0;
        if (V) for (int r = 0; r < m; ++r) V[IDX(i + 1 + r, i, n)] = v[r];
        if (taus) taus[i] = tau;
        free(v);
    }
    free(work);
}

/* Accumulate Q = H(0) H(1) ... H(n-3) from the reflectors of oracle_zgehd2. */
void oracle_zunghr(int n, const zc *V, const zc *taus, zc *Q)
{
    memset(Q, 0, sizeof(zc) * (size_t)n * n);
    for (int i = 0; i < n; ++i) Q[IDX(i, i, n)] = 1.0;
    /* Q <- Q * H(i) for i ascending: Q(:, i+1:) -= (Q v) tau v^H */
    for (int i = 0; i + 1 < n; ++i) {
        int m = n - 1 - i;
        for (int r = 0; r < n; ++r) {
            zc s = 0;
            for (int c = 0; c < m; ++c) s += Q[IDX(r, i + 1 + c, n)] * V[IDX(i + 1 + c, i, n)];
            for (int c = 0; c < m; ++c) Q[IDX(r, i + 1 + c, n)] -= s * taus[i] * conj(V[IDX(i + 1 + c, i, n)]);
        }
    }
}

/* ZLAHQR with ilo = 1, ihi = n, iloz = 1, ihiz = n.  Returns 0 on success or
 * the 1-based index of the first unconverged eigenvalue (LAPACK INFO).  When
 * `sweeps` is non-NULL it receives the total number of QR sweeps.
 * `itmax_override` < 0 means the LAPACK cap 30*max(10,n) per eigenvalue. */
int oracle_zlahqr(int wantt, int wantz, int n, zc *h, int ldh, zc *w, zc *z, int ldz,
                  int itmax_override, int *sweeps)
{
#define H(i, j) h[IDX((i) - 1, (j) - 1, ldh)]
#define Z(i, j) z[IDX((i) - 1, (j) - 1, ldz)]
#define W(i) w[(i) - 1]
    const double dat1 = 0.75;
    const int kexsh = 10;
    int ilo = 1, ihi = n, iloz = 1, ihiz = n;
    int total = 0;
    if (sweeps) *sweeps = 0;
    if (n == 0) return 0;
    if (ilo == ihi) { W(ilo) = H(ilo, ilo); return 0; }
    for (int j = ilo; j <= ihi - 3; ++j) { H(j + 2, j) = 0; H(j + 3, j) = 0; }
    if (ilo <= ihi - 2) H(ihi, ihi - 2) = 0;
    int jlo = wantt ? 1 : ilo, jhi = wantt ? n : ihi;
    for (int i = ilo + 1; i <= ihi; ++i) {
        if (cimag(H(i, i - 1)) != 0.0) {
            zc sc = H(i, i - 1) / cabs1(H(i, i - 1));
            sc = conj(sc) / cabs(sc);
            H(i, i - 1) = cabs(H(i, i - 1));
            for (int j = i; j <= jhi; ++j) H(i, j) *= sc;
            int top = (jhi < i + 1 ? jhi : i + 1);
            for (int j = jlo; j <= top; ++j) H(j, i) *= conj(sc);
            if (wantz) for (int j = iloz; j <= ihiz; ++j) Z(j, i) *= conj(sc);
        }
    }
    int nh = ihi - ilo + 1;
    double smlnum = SAFMIN * ((double)nh / ULP);
    int i1 = 1, i2 = n;
    int itmax = itmax_override >= 0 ? itmax_override : 30 * (nh > 10 ? nh : 10);
    int kdefl = 0;
    int i = ihi;
    while (i >= ilo) {
        int l = ilo, converged = 0;
        for (int its = 0; its <= itmax; ++its) {
            int k;
            for (k = i; k >= l + 1; --k) {
                if (cabs1(H(k, k - 1)) <= smlnum) break;
                double tst = cabs1(H(k - 1, k - 1)) + cabs1(H(k, k));
                if (tst == 0.0) {
                    if (k - 2 >= ilo) tst += fabs(creal(H(k - 1, k - 2)));
                    if (k + 1 <= ihi) tst += fabs(creal(H(k + 1, k)));
                }
                if (fabs(creal(H(k, k - 1))) <= ULP * tst) {
                    double ab = fmax(cabs1(H(k, k - 1)), cabs1(H(k - 1, k)));
                    double ba = fmin(cabs1(H(k, k - 1)), cabs1(H(k - 1, k)));
                    double aa = fmax(cabs1(H(k, k)), cabs1(H(k - 1, k - 1) - H(k, k)));
                    double bb = fmin(cabs1(H(k, k)), cabs1(H(k - 1, k - 1) - H(k, k)));
                    double s = aa + ab;
                    if (ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)))) break;
                }
            }
            l = k;
            if (l > ilo) H(l, l - 1) = 0;
            if (l >= i) { converged = 1; break; }
            ++kdefl;
            if (!wantt) { i1 = l; i2 = i; }
            zc t;
            if (kdefl % (2 * kexsh) == 0) {
                double s = dat1 * fabs(creal(H(i, i - 1)));
                t = s + H(i, i);
            } else if (kdefl % kexsh == 0) {
                double s = dat1 * fabs(creal(H(l + 1, l)));
                t = s + H(l, l);
            } else {
                t = H(i, i);
                zc u = csqrt(H(i - 1, i)) * csqrt(H(i, i - 1));
                double s = cabs1(u);
                if (s != 0.0) {
                    zc x = 0.5 * (H(i - 1, i - 1) - t);
                    double sx = cabs1(x);
                    s = fmax(s, cabs1(x));
                    zc xs = x / s, us = u / s;
                    zc y = s * csqrt(xs * xs + us * us);
                    if (sx > 0.0) {
                        zc xn = x / sx;
                        if (creal(xn) * creal(y) + cimag(xn) * cimag(y) < 0.0) y = -y;
                    }
                    t = t - u * (u / (x + y));
                }
            }
            /* look for two consecutive small subdiagonals */
            int m;
            zc v[2];
            for (m = i - 1; m >= l + 1; --m) {
                zc h11 = H(m, m), h22 = H(m + 1, m + 1);
                zc h11s = h11 - t;
                double h21 = creal(H(m + 1, m));
                double s = cabs1(h11s) + fabs(h21);
                h11s = h11s / s; h21 = h21 / s;
                v[0] = h11s; v[1] = h21;
                double h10 = creal(H(m, m - 1));
                if (fabs(h10) * fabs(h21) <= ULP * (cabs1(h11s) * (cabs1(h11) + cabs1(h22)))) break;
            }
            if (m == l) {
                zc h11 = H(l, l);
                zc h11s = h11 - t;
                double h21 = creal(H(l + 1, l));
                double s = cabs1(h11s) + fabs(h21);
                h11s = h11s / s; h21 = h21 / s;
                v[0] = h11s; v[1] = h21;
            }
            ++total;
            /* single-shift QR sweep */
            for (k = m; k <= i - 1; ++k) {
                if (k > m) { v[0] = H(k, k - 1); v[1] = H(k + 1, k - 1); }
                zc t1;
                oracle_zlarfg(2, &v[0], &v[1], &t1);
                if (k > m) { H(k, k - 1) = v[0]; H(k + 1, k - 1) = 0; }
                zc v2 = v[1];
                double t2 = creal(t1 * v2);
                for (int j = k; j <= i2; ++j) {
                    zc sum = conj(t1) * H(k, j) + t2 * H(k + 1, j);
                    H(k, j) = H(k, j) - sum;
                    H(k + 1, j) = H(k + 1, j) - sum * v2;
                }
                int top = (k + 2 < i ? k + 2 : i);
                for (int j = i1; j <= top; ++j) {
                    zc sum = t1 * H(j, k) + t2 * H(j, k + 1);
                    H(j, k) = H(j, k) - sum;
                    H(j, k + 1) = H(j, k + 1) - sum * conj(v2);
                }
                if (wantz) {
                    for (int j = iloz; j <= ihiz; ++j) {
                        zc sum = t1 * Z(j, k) + t2 * Z(j, k + 1);
                        Z(j, k) = Z(j, k) - sum;
                        Z(j, k + 1) = Z(j, k + 1) - sum * conj(v2);
                    }
                }
                if (k == m && m > l) {
                    zc temp = 1.0 - t1;
                    temp = temp / cabs(temp);
                    H(m + 1, m) = H(m + 1, m) * conj(temp);
                    if (m + 2 <= i) H(m + 2, m + 1) = H(m + 2, m + 1) * temp;
                    for (int j = m; j <= i; ++j) {
                        if (j != m + 1) {
                            if (i2 > j) for (int c = j + 1; c <= i2; ++c) H(j, c) *= temp;
                            for (int r = i1; r <= j - 1; ++r) H(r, j) *= conj(temp);
                            if (wantz) for (int r = iloz; r <= ihiz; ++r) Z(r, j) *= conj(temp);
                        }
                    }
                }
            }
            /* ensure H(i, i-1) is real */
            zc temp = H(i, i - 1);
            if (cimag(temp) != 0.0) {
                double rtemp = cabs(temp);
                H(i, i - 1) = rtemp;
                temp = temp / rtemp;
                if (i2 > i) for (int c = i + 1; c <= i2; ++c) H(i, c) *= conj(temp);
                for (int r = i1; r <= i - 1; ++r) H(r, i) *= temp;
                if (wantz) for (int r = iloz; r <= ihiz; ++r) Z(r, i) *= temp;
            }
        }
        if (!converged) {
            if (sweeps) *sweeps = total;
            return i;
        }
        W(i) = H(i, i);
        kdefl = 0;
        i = l - 1;
    }
    if (sweeps) *sweeps = total;
    return 0;
#undef H
#undef Z
#undef W
}

/* ZLAHQR's shift for the active block [l, i] (1-based, as oracle_zlahqr):
 * exceptional shifts when kdefl is a multiple of 20 (from the bottom) or of
 * 10 (from the top), else the Wilkinson shift. */
static zc zlahqr_shift(const zc *h, int ldh, int l, int i, int kdefl)
{
#define H(i, j) h[IDX((i) - 1, (j) - 1, ldh)]
    const double dat1 = 0.75;
    const int kexsh = 10;
    zc t;
    if (kdefl % (2 * kexsh) == 0) {
        double s = dat1 * fabs(creal(H(i, i - 1)));
        t = s + H(i, i);
    } else if (kdefl % kexsh == 0) {
        double s = dat1 * fabs(creal(H(l + 1, l)));
        t = s + H(l, l);
    } else {
        t = H(i, i);
        zc u = csqrt(H(i - 1, i)) * csqrt(H(i, i - 1));
        double s = cabs1(u);
        if (s != 0.0) {
            zc x = 0.5 * (H(i - 1, i - 1) - t);
            double sx = cabs1(x);
            s = fmax(s, cabs1(x));
            zc xs = x / s, us = u / s;
            zc y = s * csqrt(xs * xs + us * us);
            if (sx > 0.0) {
                zc xn = x / sx;
                if (creal(xn) * creal(y) + cimag(xn) * cimag(y) < 0.0) y = -y;
            }
            t = t - u * (u / (x + y));
        }
    }
    return t;
#undef H
}

/* One single-shift ZLAHQR sweep over the active block [l, i] (1-based) from
 * row l with shift sig, eigenvalues only (transformations restricted to the
 * block), ending with ZLAHQR's "make H(i, i-1) real" scaling of column i. */
static void zlahqr_sweep(zc *h, int ldh, int l, int i, zc sig)
{
#define H(i, j) h[IDX((i) - 1, (j) - 1, ldh)]
    zc v[2];
    {
        zc h11s = H(l, l) - sig;
        double h21 = creal(H(l + 1, l));
        double s = cabs1(h11s) + fabs(h21);
        v[0] = h11s / s;
        v[1] = h21 / s;
    }
    for (int k = l; k <= i - 1; ++k) {
        if (k > l) { v[0] = H(k, k - 1); v[1] = H(k + 1, k - 1); }
        zc t1;
        oracle_zlarfg(2, &v[0], &v[1], &t1);
        if (k > l) { H(k, k - 1) = v[0]; H(k + 1, k - 1) = 0; }
        zc v2 = v[1];
        double t2 = creal(t1 * v2);
        for (int j = k; j <= i; ++j) {
            zc sum = conj(t1) * H(k, j) + t2 * H(k + 1, j);
            H(k, j) = H(k, j) - sum;
            H(k + 1, j) = H(k + 1, j) - sum * v2;
        }
        int top = (k + 2 < i ? k + 2 : i);
        for (int j = l; j <= top; ++j) {
            zc sum = t1 * H(j, k) + t2 * H(j, k + 1);
            H(j, k) = H(j, k) - sum;
            H(j, k + 1) = H(j, k + 1) - sum * conj(v2);
        }
    }
    zc temp = H(i, i - 1);
    if (cimag(temp) != 0.0) {
        double rtemp = cabs(temp);
        H(i, i - 1) = rtemp;
        temp = temp / rtemp;
        for (int r = l; r <= i - 1; ++r) H(r, i) *= temp;
    }
#undef H
}

/* The two-bulge QR of the GPU's eigenvalue path, eigenvalues only (ILO = 1,
 * IHI = n): per iteration, after ZLAHQR's deflation search over [l, i], two
 * single-shift sweeps from row l with the shifts s1 (ZLAHQR's, exceptional
 * ones included) and s2 (the other eigenvalue of the trailing 2x2,
 * H(i-1,i-1) + H(i,i) - s1, or the Wilkinson shift when s1 is exceptional),
 * both taken before the first sweep; an active block of fewer than 10 rows
 * goes to oracle_zlahqr on that block, which is independent once
 * H(l, l-1) = 0.  The GPU pipelines the second sweep three positions behind
 * the first; the products it reorders commute, so the eigenvalues agree up to
 * rounding.  Returns 0 or the 1-based index of the first unconverged
 * eigenvalue; `sweeps` (optional) counts the single-shift sweeps. */
int oracle_zlahqr_dual(int n, zc *h, int ldh, zc *w, int itmax_override, int *sweeps)
{
#define H(i, j) h[IDX((i) - 1, (j) - 1, ldh)]
#define W(i) w[(i) - 1]
    const int dual_min = 10;
    int total = 0;
    if (sweeps) *sweeps = 0;
    if (n == 0) return 0;
    if (n == 1) { W(1) = H(1, 1); return 0; }
    const double smlnum = SAFMIN * ((double)n / ULP);
    const int itmax = itmax_override >= 0 ? itmax_override : 30 * (n > 10 ? n : 10);
    int kdefl = 0;
    int i = n;
    while (i >= 2) {
        int l = 1, converged = 0;
        for (int its = 0; its <= itmax; ++its) {
            int k;
            for (k = i; k >= l + 1; --k) {
                if (cabs1(H(k, k - 1)) <= smlnum) break;
                double tst = cabs1(H(k - 1, k - 1)) + cabs1(H(k, k));
                if (tst == 0.0) {
                    if (k - 2 >= 1) tst += fabs(creal(H(k - 1, k - 2)));
                    if (k + 1 <= n) tst += fabs(creal(H(k + 1, k)));
                }
                if (fabs(creal(H(k, k - 1))) <= ULP * tst) {
                    double ab = fmax(cabs1(H(k, k - 1)), cabs1(H(k - 1, k)));
                    double ba = fmin(cabs1(H(k, k - 1)), cabs1(H(k - 1, k)));
                    double aa = fmax(cabs1(H(k, k)), cabs1(H(k - 1, k - 1) - H(k, k)));
                    double bb = fmin(cabs1(H(k, k)), cabs1(H(k - 1, k - 1) - H(k, k)));
                    double s = aa + ab;
                    if (ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)))) break;
                }
            }
            l = k;
            if (l > 1) H(l, l - 1) = 0;
            if (l >= i) { converged = 1; break; }
            if (i - l + 1 < dual_min) {
                int sw = 0;
                const int info = oracle_zlahqr(0, 0, i - l + 1, &H(l, l), ldh, &W(l), NULL, 1, itmax_override, &sw);
                total += sw;
                if (info != 0) {
                    if (sweeps) *sweeps = total;
                    return l - 1 + info;
                }
                converged = 2;   /* eigenvalues l..i stored by oracle_zlahqr */
                break;
            }
            ++kdefl;
            const zc s1 = zlahqr_shift(h, ldh, l, i, kdefl);
            const zc s2 = (kdefl % 10 == 0) ? zlahqr_shift(h, ldh, l, i, 1) : (H(i - 1, i - 1) + H(i, i)) - s1;
            zlahqr_sweep(h, ldh, l, i, s1);
            zlahqr_sweep(h, ldh, l, i, s2);
            total += 2;
        }
        if (!converged) {
            if (sweeps) *sweeps = total;
            return i;
        }
        if (converged == 1) W(i) = H(i, i);
        kdefl = 0;
        i = l - 1;
    }
    if (i == 1) W(1) = H(1, 1);
    if (sweeps) *sweeps = total;
    return 0;
#undef H
#undef W
}

static int all_finite(int n, const zc *a)
{
    for (size_t k = 0; k < (size_t)n * n; ++k)
        if (!isfinite(creal(a[k])) || !isfinite(cimag(a[k]))) return 0;
    return 1;
}

/* ZGEEV's norm-range scaling (LAPACK zgeev scales A when its largest entry
 * lies outside [SMLNUM, BIGNUM], SMLNUM = sqrt(safmin)/eps = 2^-459, BIGNUM =
 * 1/SMLNUM) as the GPU path applies it: the test on m = max(|Re|, |Im|) over
 * the entries, the factor the power of two 2^e that brings m into
 * [2^-458, 2^-457) or [2^458, 2^459) (exact), where LAPACK multiplies by
 * SMLNUM/ANRM or BIGNUM/ANRM (ZLASCL).  Scales a in place and returns e (0:
 * in range, nothing done); the eigenvalues are multiplied by 2^-e. */
static int norm_range_scale(int n, zc *a)
{
    double m = 0.0;
    for (size_t k = 0; k < (size_t)n * n; ++k) m = fmax(m, fmax(fabs(creal(a[k])), fabs(cimag(a[k]))));
    if (!(m > 0.0) || (m >= 0x1p-459 && m <= 0x1p459)) return 0;
    const int k = ilogb(m);
    const int e = m < 0x1p-459 ? -458 - k : 458 - k;
    const double f = ldexp(1.0, e);
    for (size_t j = 0; j < (size_t)n * n; ++j) a[j] = CMPLX(creal(a[j]) * f, cimag(a[j]) * f);
    return e;
}

static void norm_range_unscale(int n, zc *w, int e)
{
    if (e == 0) return;
    const double f = ldexp(1.0, -e);
    for (int i = 0; i < n; ++i) w[i] = CMPLX(creal(w[i]) * f, cimag(w[i]) * f);
}

/* The batched method: Hessenberg reduction then eigenvalues-only QR, per
 * matrix, as the GPU path computes it (after the norm-range scaling).  A: batch x (n x n) column-major,
 * contiguous; eig[i + j*n]; flag[j].  sweeps (optional) per matrix. */
void oracle_zgeev_batched(int n, int batch, const zc *A, zc *eig, signed char *flag,
                          int *sweeps, int itmax_override)
{
    #pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < batch; ++b) {
        zc *h = (zc *)malloc(sizeof(zc) * (size_t)n * n);
        memcpy(h, A + (size_t)b * n * n, sizeof(zc) * (size_t)n * n);
        zc *w = eig + (size_t)b * n;
        int sw = 0;
        if (!all_finite(n, h)) {
            for (int i = 0; i < n; ++i) w[i] = CMPLX(NAN, NAN);
            flag[b] = 2;
        } else {
            const int e = norm_range_scale(n, h);
            oracle_zgehd2(n, h, n, NULL, NULL);
            for (int i = 0; i < n; ++i) w[i] = h[IDX(i, i, n)];
            int info = oracle_zlahqr(0, 0, n, h, n, w, NULL, n, itmax_override, &sw);
            if (info != 0) {
                /* unconverged eigenvalues: report the current diagonal */
                for (int i = 0; i < info; ++i) w[i] = h[IDX(i, i, n)];
            }
            norm_range_unscale(n, w, e);
            flag[b] = info != 0 ? 1 : 0;
        }
        if (sweeps) sweeps[b] = sw;
        free(h);
    }
}

/* oracle_zgeev_batched with the two-bulge QR (oracle_zlahqr_dual). */
void oracle_zgeev_batched_dual(int n, int batch, const zc *A, zc *eig, signed char *flag, int *sweeps,
                               int itmax_override)
{
    #pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < batch; ++b) {
        zc *h = (zc *)malloc(sizeof(zc) * (size_t)n * n);
        memcpy(h, A + (size_t)b * n * n, sizeof(zc) * (size_t)n * n);
        zc *w = eig + (size_t)b * n;
        int sw = 0;
        if (!all_finite(n, h)) {
            for (int i = 0; i < n; ++i) w[i] = CMPLX(NAN, NAN);
            flag[b] = 2;
        } else {
            const int e = norm_range_scale(n, h);
            oracle_zgehd2(n, h, n, NULL, NULL);
            for (int i = 0; i < n; ++i) w[i] = h[IDX(i, i, n)];
            int info = oracle_zlahqr_dual(n, h, n, w, itmax_override, &sw);
            if (info != 0)
                for (int i = 0; i < info; ++i) w[i] = h[IDX(i, i, n)];
            norm_range_unscale(n, w, e);
            flag[b] = info != 0 ? 1 : 0;
        }
        if (sweeps) sweeps[b] = sw;
        free(h);
    }
}

/* Full Schur form A = U T U^H (T upper triangular) with the same method:
 * zgehd2 + reflector accumulation + zlahqr(wantt, wantz).  Returns INFO. */
int oracle_zschur(int n, const zc *A, zc *T, zc *U, zc *w)
{
    zc *V = (zc *)malloc(sizeof(zc) * (size_t)n * n);
    zc *taus = (zc *)malloc(sizeof(zc) * (size_t)(n > 1 ? n : 1));
    memcpy(T, A, sizeof(zc) * (size_t)n * n);
    oracle_zgehd2(n, T, n, V, taus);
    oracle_zunghr(n, V, taus, U);
    int info = oracle_zlahqr(1, 1, n, T, n, w, U, n, -1, NULL);
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) T[IDX(i, j, n)] = 0;
    free(V); free(taus);
    return info;
}

/* ZTREVC-style eigenvectors of an upper-triangular T (no rescaling to avoid
 * overflow, which n <= 64 with O(1) entries does not need).  XR(:,i) solves
 * (T - t_ii) x = 0 with x_i = 1, x_{>i} = 0; YL(:,i) solves
 * y^H (T - t_ii) = 0 with y_i = 1, y_{<i} = 0.  Near-singular pivots are
 * replaced by smin = max(ulp*|t_ii|, smlnum) as ZTREVC does. */
void oracle_ztrevc(int n, const zc *T, zc *XR, zc *YL)
{
    double smlnum = SAFMIN * ((double)n / ULP);
    memset(XR, 0, sizeof(zc) * (size_t)n * n);
    memset(YL, 0, sizeof(zc) * (size_t)n * n);
    for (int i = 0; i < n; ++i) {
        zc lam = T[IDX(i, i, n)];
        double smin = fmax(ULP * cabs1(lam), smlnum);
        zc *x = XR + (size_t)i * n;
        x[i] = 1.0;
        for (int r = i - 1; r >= 0; --r) {
            zc s = 0;
            for (int c = r + 1; c <= i; ++c) s += T[IDX(r, c, n)] * x[c];
            zc d = T[IDX(r, r, n)] - lam;
            if (cabs1(d) < smin) d = smin;
            x[r] = -s / d;
        }
        zc *y = YL + (size_t)i * n;
        y[i] = 1.0;
        /* y^H T = lam y^H  <=>  T^H y = conj(lam) y: forward solve for r > i */
        for (int r = i + 1; r < n; ++r) {
            zc s = 0;
            for (int c = i; c < r; ++c) s += conj(T[IDX(c, r, n)]) * y[c];
            zc d = conj(T[IDX(r, r, n)] - lam);
            if (cabs1(d) < smin) d = smin;
            y[r] = -s / d;
        }
    }
}

/* dL/dA for L = Re sum_i conj(g_i) lambda_i (torch convention):
 *   G = sum_i g_i y_i x_i^H / conj(y_i^H x_i)
 * with x_i, y_i the right/left eigenvectors of A.  Returns INFO of the Schur
 * step.  Matrices column-major n x n; w receives the eigenvalues in the Schur
 * diagonal order (g is matched to that order). */
int oracle_zgeev_adjoint(int n, const zc *A, const zc *g, zc *G, zc *w)
{
    size_t nn = (size_t)n * n;
    zc *T = (zc *)malloc(sizeof(zc) * nn), *U = (zc *)malloc(sizeof(zc) * nn);
    zc *xr = (zc *)malloc(sizeof(zc) * nn), *yl = (zc *)malloc(sizeof(zc) * nn);
    zc *X = (zc *)malloc(sizeof(zc) * nn), *Y = (zc *)malloc(sizeof(zc) * nn);
    int info = oracle_zschur(n, A, T, U, w);
    oracle_ztrevc(n, T, xr, yl);
    /* X = U xr, Y = U yl */
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < n; ++r) {
            zc sx = 0, sy = 0;
            for (int c = 0; c < n; ++c) {
                sx += U[IDX(r, c, n)] * xr[IDX(c, i, n)];
                sy += U[IDX(r, c, n)] * yl[IDX(c, i, n)];
            }
            X[IDX(r, i, n)] = sx; Y[IDX(r, i, n)] = sy;
        }
    memset(G, 0, sizeof(zc) * nn);
    for (int i = 0; i < n; ++i) {
        zc d = 0;
        for (int r = 0; r < n; ++r) d += conj(Y[IDX(r, i, n)]) * X[IDX(r, i, n)];
        zc coef = g[i] / conj(d);
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r)
                G[IDX(r, c, n)] += coef * Y[IDX(r, i, n)] * conj(X[IDX(c, i, n)]);
    }
    free(T); free(U); free(xr); free(yl); free(X); free(Y);
    return info;
}

/* ZGEEV's eigenvector normalisation (LAPACK zgeev.f after ZTREVC3 and the
 * back-transform): unit Euclidean norm (DZNRM2, scaled), then the component
 * of largest |re|^2 + |im|^2 (first one on ties, IDAMAX) is rotated to be
 * real and non-negative, its imaginary part set to exactly zero. */
static void zgeev_normalise(int n, zc *v)
{
    double scl = 1.0 / znrm2(n, v);
    for (int r = 0; r < n; ++r) v[r] *= scl;
    int k = 0;
    double best = -1.0;
    for (int r = 0; r < n; ++r) {
        double m = creal(v[r]) * creal(v[r]) + cimag(v[r]) * cimag(v[r]);
        if (m > best) { best = m; k = r; }
    }
    zc tmp = conj(v[k]) / sqrt(best);
    for (int r = 0; r < n; ++r) v[r] *= tmp;
    v[k] = creal(v[k]);
}

/* Eigenvalues and eigenvectors by the method of the GPU path: Schur form
 * (oracle_zschur), eigenvectors of T (oracle_ztrevc), back-transform by U,
 * ZGEEV normalisation.  VR(:,i) right (A v = w_i v), VL(:,i) left
 * (u^H A = w_i u^H); either may be NULL.  Column-major n x n.  Returns the
 * INFO of the Schur step (vectors are not formed when it is non-zero). */
int oracle_zgeev_vectors(int n, const zc *A, zc *w, zc *VL, zc *VR)
{
    size_t nn = (size_t)n * n;
    zc *T = (zc *)malloc(sizeof(zc) * nn), *U = (zc *)malloc(sizeof(zc) * nn);
    zc *xr = (zc *)malloc(sizeof(zc) * nn), *yl = (zc *)malloc(sizeof(zc) * nn);
    int info = oracle_zschur(n, A, T, U, w);
    if (info == 0) {
        oracle_ztrevc(n, T, xr, yl);
        for (int side = 0; side < 2; ++side) {
            zc *V = side == 0 ? VR : VL;
            const zc *E = side == 0 ? xr : yl;
            if (!V) continue;
            for (int i = 0; i < n; ++i) {
                for (int r = 0; r < n; ++r) {
                    zc s = 0;
                    for (int c = 0; c < n; ++c) s += U[IDX(r, c, n)] * E[IDX(c, i, n)];
                    V[IDX(r, i, n)] = s;
                }
                zgeev_normalise(n, V + (size_t)i * n);
            }
        }
    }
    free(T); free(U); free(xr); free(yl);
    return info;
}

/* ZGEBAL (LAPACK 3.11+ form): balance the n x n column-major a.
 * job: 'N' nothing, 'P' permute, 'S' scale, 'B' both.  On return rows/cols
 * outside [ilo, ihi] (0-based) hold isolated eigenvalues; scale[j] is the
 * permutation index (j outside [ilo, ihi]) or the scale factor (inside).
 * Returns 0, or -3 for a NaN met while scaling. */
int oracle_zgebal(char job, int n, zc *a, int lda, int *ilo_out, int *ihi_out, double *scale)
{
#define A(i, j) a[IDX(i, j, lda)]
    const double sclfac = 2.0, factor = 0.95;
    int k = 0, l = n - 1;
    if (n == 0) { *ilo_out = 0; *ihi_out = -1; return 0; }
    if (job == 'N') {
        for (int i = 0; i < n; ++i) scale[i] = 1.0;
        *ilo_out = 0; *ihi_out = n - 1;
        return 0;
    }
    if (job != 'S') {
        /* rows isolating an eigenvalue: push down */
        int noconv = 1;
        while (noconv) {
            noconv = 0;
            for (int i = l; i >= 0; --i) {
                int canswap = 1;
                for (int j = 0; j <= l; ++j)
                    if (i != j && (creal(A(i, j)) != 0.0 || cimag(A(i, j)) != 0.0)) { canswap = 0; break; }
                if (canswap) {
                    scale[l] = i;
                    if (i != l) {
                        for (int r = 0; r <= l; ++r) { zc t = A(r, i); A(r, i) = A(r, l); A(r, l) = t; }
                        for (int c = k; c < n; ++c) { zc t = A(i, c); A(i, c) = A(l, c); A(l, c) = t; }
                    }
                    noconv = 1;
                    if (l == 0) { *ilo_out = 0; *ihi_out = 0; return 0; }
                    --l;
                }
            }
        }
        /* columns isolating an eigenvalue: push left */
        noconv = 1;
        while (noconv) {
            noconv = 0;
            for (int j = k; j <= l; ++j) {
                int canswap = 1;
                for (int i = k; i <= l; ++i)
                    if (i != j && (creal(A(i, j)) != 0.0 || cimag(A(i, j)) != 0.0)) { canswap = 0; break; }
                if (canswap) {
                    scale[k] = j;
                    if (j != k) {
                        for (int r = 0; r <= l; ++r) { zc t = A(r, j); A(r, j) = A(r, k); A(r, k) = t; }
                        for (int c = k; c < n; ++c) { zc t = A(j, c); A(j, c) = A(k, c); A(k, c) = t; }
                    }
                    noconv = 1;
                    ++k;
                }
            }
        }
    }
    for (int i = k; i <= l; ++i) scale[i] = 1.0;
    *ilo_out = k; *ihi_out = l;
    if (job == 'P') return 0;
    /* iterative norm reduction with powers of two */
    const double sfmin1 = SAFMIN / ULP, sfmax1 = 1.0 / sfmin1;
    const double sfmin2 = sfmin1 * sclfac, sfmax2 = 1.0 / sfmin2;
    int noconv = 1;
    while (noconv) {
        noconv = 0;
        for (int i = k; i <= l; ++i) {
            double c = znrm2(l - k + 1, &A(k, i));
            zc rowv[64];
            for (int j = k; j <= l; ++j) rowv[j - k] = A(i, j);
            double r = znrm2(l - k + 1, rowv);
            /* IZAMAX picks the first entry of largest |re| + |im|; CA, RA are its modulus */
            int ica = 0, ira = k;
            for (int q = 1; q <= l; ++q) if (cabs1(A(q, i)) > cabs1(A(ica, i))) ica = q;
            for (int q = k + 1; q < n; ++q) if (cabs1(A(i, q)) > cabs1(A(i, ira))) ira = q;
            double ca = cabs(A(ica, i)), ra = cabs(A(i, ira));
            if (c == 0.0 || r == 0.0) continue;
            if (isnan(c + ca + r + ra)) return -3;
            double g = r / sclfac, f = 1.0, s = c + r;
            while (c < g && fmax(f, fmax(c, ca)) < sfmax2 && fmin(r, fmin(g, ra)) > sfmin2) {
                f *= sclfac; c *= sclfac; ca *= sclfac; r /= sclfac; g /= sclfac; ra /= sclfac;
            }
            g = c / sclfac;
            while (g >= r && fmax(r, ra) < sfmax2 && fmin(fmin(f, c), fmin(g, ca)) > sfmin2) {
                f /= sclfac; c /= sclfac; g /= sclfac; ca /= sclfac; r *= sclfac; ra *= sclfac;
            }
            if (c + r >= factor * s) continue;
            if (f < 1.0 && scale[i] < 1.0 && f * scale[i] <= sfmin1) continue;
            if (f > 1.0 && scale[i] > 1.0 && scale[i] >= sfmax1 / f) continue;
            g = 1.0 / f;
            scale[i] *= f;
            noconv = 1;
            for (int q = k; q < n; ++q) A(i, q) *= g;
            for (int q = 0; q <= l; ++q) A(q, i) *= f;
        }
    }
    return 0;
#undef A
}

/* The batched method with balancing first (ZGEBAL job 'N', 'P', 'S' or 'B'),
 * then zgehd2 + eigenvalues-only zlahqr on the whole balanced matrix, as the
 * GPU path does. */
void oracle_zgeev_batched_bal(int n, int batch, const zc *A, zc *eig, signed char *flag, int itmax_override,
                              char job)
{
    #pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < batch; ++b) {
        zc *h = (zc *)malloc(sizeof(zc) * (size_t)n * n);
        double *scale = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        memcpy(h, A + (size_t)b * n * n, sizeof(zc) * (size_t)n * n);
        zc *w = eig + (size_t)b * n;
        if (!all_finite(n, h)) {
            for (int i = 0; i < n; ++i) w[i] = CMPLX(NAN, NAN);
            flag[b] = 2;
        } else {
            int ilo, ihi;
            const int e = norm_range_scale(n, h);
            oracle_zgebal(job, n, h, n, &ilo, &ihi, scale);
            oracle_zgehd2(n, h, n, NULL, NULL);
            for (int i = 0; i < n; ++i) w[i] = h[IDX(i, i, n)];
            int info = oracle_zlahqr(0, 0, n, h, n, w, NULL, n, itmax_override, NULL);
            if (info != 0)
                for (int i = 0; i < info; ++i) w[i] = h[IDX(i, i, n)];
            norm_range_unscale(n, w, e);
            flag[b] = info != 0 ? 1 : 0;
        }
        free(h);
        free(scale);
    }
}
