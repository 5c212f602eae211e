/*
 * ORACLE — test infrastructure only (tests/, the smoke check and bench.py's
 * CPU-baseline leg).  The shipped library never loads it.
 *
 * CPU restatement of the real-matrix path (DGEEV analogue, eigenvalues only,
 * no balancing), the method the GPU's real kernels run:
 *   oracle_dgehd2   Householder reduction to upper Hessenberg form
 *                   (LAPACK DGEHD2 with DLARFG; entries below the
 *                   sub-diagonal are zeroed on return)
 *   oracle_dlahqr   Francis double-shift QR with deflation, eigenvalues only
 *                   (LAPACK DLAHQR, WANTT = WANTZ = false: Ahues-Kressner
 *                   deflation test, exceptional shifts every 10th / 20th
 *                   sweep without a deflation, cap 30 max(10, n) sweeps),
 *                   converged 2x2 blocks standardised by
 *   oracle_dlanv2   LAPACK DLANV2 (eigenvalues of a real 2x2, complex pairs
 *                   with the positive imaginary part first)
 *   oracle_dgeev_batched  the three over a batch (OpenMP over matrices)
 * The reference has no real path at all; these restate LAPACK's published
 * algorithms, which the GPU kernels follow step for step.
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define HH(r, c) H[(r) + (size_t)(c) * ld]

static double dlapy2(double x, double y)
{
    const double xa = fabs(x), ya = fabs(y);
    const double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0 || w > DBL_MAX) return w;
    const double q = z / w;
    return w * sqrt(1.0 + q * q);
}

/* Euclidean norm with scaling (DNRM2's result up to rounding). */
static double dnrm2(int n, const double *x)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; ++i) {
        if (x[i] != 0.0) {
            const double a = fabs(x[i]);
            if (scale < a) {
                const double q = scale / a;
                ssq = 1.0 + ssq * q * q;
                scale = a;
            } else {
                const double q = a / scale;
                ssq += q * q;
            }
        }
    }
    return scale * sqrt(ssq);
}

/* DLARFG: H = I - tau (1, v)(1, v)^T with H (alpha, x) = (beta, 0).  The
 * safmin rescaling loop for |beta| < safmin/eps is not reproduced (the GPU
 * kernels do not reproduce it either). */
static void dlarfg(int n, double *alpha, double *x, double *tau)
{
    if (n <= 1) { *tau = 0.0; return; }
    const double xnorm = dnrm2(n - 1, x);
    if (xnorm == 0.0) { *tau = 0.0; return; }
    const double beta = -copysign(dlapy2(*alpha, xnorm), *alpha);
    *tau = (beta - *alpha) / beta;
    const double sc = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[i] *= sc;
    *alpha = beta;
}

void oracle_dgehd2(int n, double *H)
{
    const int ld = n;
    for (int i = 0; i + 2 <= n; ++i) {
        const int m = n - 1 - i;   /* rows i+1 .. n-1 */
        double tau;
        dlarfg(m, &HH(i + 1, i), &HH(i + 2 < n ? i + 2 : i + 1, i), &tau);
        if (tau == 0.0) continue;
        const double beta = HH(i + 1, i);
        HH(i + 1, i) = 1.0;
        const double *v = &HH(i + 1, i);
        /* right: H(0:n, i+1:n) -= tau (H v) v^T */
        for (int r = 0; r < n; ++r) {
            double u = 0.0;
            for (int c = 0; c < m; ++c) u += HH(r, i + 1 + c) * v[c];
            u *= tau;
            for (int c = 0; c < m; ++c) HH(r, i + 1 + c) -= u * v[c];
        }
        /* left: H(i+1:n, i+1:n) -= tau v (v^T H) */
        for (int c = i + 1; c < n; ++c) {
            double s = 0.0;
            for (int r = 0; r < m; ++r) s += v[r] * HH(i + 1 + r, c);
            s *= tau;
            for (int r = 0; r < m; ++r) HH(i + 1 + r, c) -= v[r] * s;
        }
        HH(i + 1, i) = beta;
    }
    for (int c = 0; c + 2 < n; ++c)
        for (int r = c + 2; r < n; ++r) HH(r, c) = 0.0;
}

void oracle_dlanv2(double *a, double *b, double *c, double *d, double *rt1r, double *rt1i, double *rt2r,
                   double *rt2i, double *cs, double *sn)
{
    const double multpl = 4.0, eps = DBL_EPSILON;   /* DLAMCH('P') = 2^-52 */
    const double safmin = DBL_MIN;
    /* SAFMN2 = base ** int(log(safmin / eps) / log(base) / 2) */
    const double safmn2 = pow(2.0, (int)(log(safmin / eps) / log(2.0) / 2.0));
    const double safmx2 = 1.0 / safmn2;
    double A = *a, B = *b, C = *c, D = *d;
    double CS, SN;
    if (C == 0.0) {
        CS = 1.0; SN = 0.0;
    } else if (B == 0.0) {
        CS = 0.0; SN = 1.0;
        const double temp = D;
        D = A; A = temp; B = -C; C = 0.0;
    } else if ((A - D) == 0.0 && copysign(1.0, B) != copysign(1.0, C)) {
        CS = 1.0; SN = 0.0;
    } else {
        double temp = A - D;
        double p = 0.5 * temp;
        const double bcmax = fmax(fabs(B), fabs(C));
        const double bcmis = fmin(fabs(B), fabs(C)) * copysign(1.0, B) * copysign(1.0, C);
        double scale = fmax(fabs(p), bcmax);
        double z = (p / scale) * p + (bcmax / scale) * bcmis;
        if (z >= multpl * eps) {
            /* real eigenvalues */
            z = p + copysign(sqrt(scale) * sqrt(z), p);
            A = D + z;
            D = D - (bcmax / z) * bcmis;
            const double tau = dlapy2(C, z);
            CS = z / tau;
            SN = C / tau;
            B = B - C;
            C = 0.0;
        } else {
            /* complex or almost equal real eigenvalues: equal diagonal */
            int count = 0;
            double sigma = B + C;
            for (;;) {
                ++count;
                scale = fmax(fabs(temp), fabs(sigma));
                if (scale >= safmx2) {
                    sigma *= safmn2; temp *= safmn2;
                    if (count <= 20) continue;
                }
                if (scale <= safmn2) {
                    sigma *= safmx2; temp *= safmx2;
                    if (count <= 20) continue;
                }
                break;
            }
            p = 0.5 * temp;
            double tau = dlapy2(sigma, temp);
            CS = sqrt(0.5 * (1.0 + fabs(sigma) / tau));
            SN = -(p / (tau * CS)) * copysign(1.0, sigma);
            const double aa = A * CS + B * SN, bb = -A * SN + B * CS;
            const double cc = C * CS + D * SN, dd = -C * SN + D * CS;
            A = aa * CS + cc * SN;
            B = bb * CS + dd * SN;
            C = -aa * SN + cc * CS;
            D = -bb * SN + dd * CS;
            temp = 0.5 * (A + D);
            A = temp;
            D = temp;
            if (C != 0.0) {
                if (B != 0.0) {
                    if (copysign(1.0, B) == copysign(1.0, C)) {
                        /* real eigenvalues: upper triangular */
                        const double sab = sqrt(fabs(B)), sac = sqrt(fabs(C));
                        p = copysign(sab * sac, C);
                        tau = 1.0 / sqrt(fabs(B + C));
                        A = temp + p;
                        D = temp - p;
                        B = B - C;
                        C = 0.0;
                        const double cs1 = sab * tau, sn1 = sac * tau;
                        temp = CS * cs1 - SN * sn1;
                        SN = CS * sn1 + SN * cs1;
                        CS = temp;
                    }
                } else {
                    B = -C;
                    C = 0.0;
                    temp = CS;
                    CS = -SN;
                    SN = temp;
                }
            }
        }
    }
    *rt1r = A;
    *rt2r = D;
    if (C == 0.0) {
        *rt1i = 0.0;
        *rt2i = 0.0;
    } else {
        *rt1i = sqrt(fabs(B)) * sqrt(fabs(C));
        *rt2i = -*rt1i;
    }
    *a = A; *b = B; *c = C; *d = D;
    *cs = CS; *sn = SN;
}

/* DLAHQR, eigenvalues only, on the whole matrix (ILO = 1, IHI = n).  H is
 * n x n column-major with leading dimension ld and is overwritten.  Returns
 * 0 or the LAPACK INFO (1-based row of the unconverged block's end); on
 * failure the unconverged eigenvalues are set to the diagonal of H. */
int oracle_dlahqr(int n, double *H, int ld, double *wr, double *wi, int itmax_override)
{
    if (n <= 0) return 0;
    if (n == 1) { wr[0] = HH(0, 0); wi[0] = 0.0; return 0; }
    for (int j = 0; j + 3 < n; ++j) { HH(j + 2, j) = 0.0; HH(j + 3, j) = 0.0; }
    if (n >= 3) HH(n - 1, n - 3) = 0.0;
    const double safmin = DBL_MIN, ulp = DBL_EPSILON;   /* DLAMCH('P') = eps * base: 2^-52 */
    const double smlnum = safmin * ((double)n / ulp);
    const double dat1 = 0.75, dat2 = -0.4375;
    const int kexsh = 10;
    const int itmax = itmax_override >= 0 ? itmax_override : 30 * (n > 10 ? n : 10);
    int kdefl = 0;
    int i = n - 1;
    while (i >= 0) {
        int l = 0;
        int conv = 0;
        for (int its = 0; its <= itmax; ++its) {
            int k;
            for (k = i; k > l; --k) {
                if (fabs(HH(k, k - 1)) <= smlnum) break;
                double tst = fabs(HH(k - 1, k - 1)) + fabs(HH(k, k));
                if (tst == 0.0) {
                    if (k - 2 >= 0) tst += fabs(HH(k - 1, k - 2));
                    if (k + 1 <= n - 1) tst += fabs(HH(k + 1, k));
                }
                if (fabs(HH(k, k - 1)) <= ulp * tst) {
                    const double ab = fmax(fabs(HH(k, k - 1)), fabs(HH(k - 1, k)));
                    const double ba = fmin(fabs(HH(k, k - 1)), fabs(HH(k - 1, k)));
                    const double aa = fmax(fabs(HH(k, k)), fabs(HH(k - 1, k - 1) - HH(k, k)));
                    const double bb = fmin(fabs(HH(k, k)), fabs(HH(k - 1, k - 1) - HH(k, k)));
                    const double s = aa + ab;
                    if (ba * (ab / s) <= fmax(smlnum, ulp * (bb * (aa / s)))) break;
                }
            }
            l = k;
            if (l > 0) HH(l, l - 1) = 0.0;
            if (l >= i - 1) { conv = 1; break; }
            ++kdefl;
            double h11, h12, h21, h22;
            if (kdefl % (2 * kexsh) == 0) {
                const double s = fabs(HH(i, i - 1)) + fabs(HH(i - 1, i - 2));
                h11 = dat1 * s + HH(i, i); h12 = dat2 * s; h21 = s; h22 = h11;
            } else if (kdefl % kexsh == 0) {
                const double s = fabs(HH(l + 1, l)) + fabs(HH(l + 2, l + 1));
                h11 = dat1 * s + HH(l, l); h12 = dat2 * s; h21 = s; h22 = h11;
            } else {
                h11 = HH(i - 1, i - 1); h21 = HH(i, i - 1); h12 = HH(i - 1, i); h22 = HH(i, i);
            }
            double rt1r, rt1i, rt2r, rt2i;
            double s = fabs(h11) + fabs(h12) + fabs(h21) + fabs(h22);
            if (s == 0.0) {
                rt1r = rt1i = rt2r = rt2i = 0.0;
            } else {
                h11 /= s; h21 /= s; h12 /= s; h22 /= s;
                const double tr = (h11 + h22) / 2.0;
                const double det = (h11 - tr) * (h22 - tr) - h12 * h21;
                const double rtdisc = sqrt(fabs(det));
                if (det >= 0.0) {
                    rt1r = tr * s; rt2r = rt1r; rt1i = rtdisc * s; rt2i = -rt1i;
                } else {
                    rt1r = tr + rtdisc; rt2r = tr - rtdisc;
                    if (fabs(rt1r - h22) <= fabs(rt2r - h22)) { rt1r *= s; rt2r = rt1r; }
                    else { rt2r *= s; rt1r = rt2r; }
                    rt1i = rt2i = 0.0;
                }
            }
            /* two consecutive small sub-diagonal elements */
            double v[3];
            int m;
            for (m = i - 2; m >= l; --m) {
                double h21s = HH(m + 1, m);
                s = fabs(HH(m, m) - rt2r) + fabs(rt2i) + fabs(h21s);
                h21s = HH(m + 1, m) / s;
                v[0] = h21s * HH(m, m + 1) + (HH(m, m) - rt1r) * ((HH(m, m) - rt2r) / s) - rt1i * (rt2i / s);
                v[1] = h21s * (HH(m, m) + HH(m + 1, m + 1) - rt1r - rt2r);
                v[2] = h21s * HH(m + 2, m + 1);
                s = fabs(v[0]) + fabs(v[1]) + fabs(v[2]);
                v[0] /= s; v[1] /= s; v[2] /= s;
                if (m == l) break;
                const double h00 = fabs(HH(m, m - 1)) * (fabs(v[1]) + fabs(v[2]));
                const double h01 = fabs(v[0]) * (fabs(HH(m - 1, m - 1)) + fabs(HH(m, m)) + fabs(HH(m + 1, m + 1)));
                if (h00 <= ulp * h01) break;
            }
            /* double-shift sweep */
            for (int kk = m; kk <= i - 1; ++kk) {
                const int nr = (i - kk + 1) < 3 ? (i - kk + 1) : 3;
                if (kk > m)
                    for (int q = 0; q < nr; ++q) v[q] = HH(kk + q, kk - 1);
                double t1;
                dlarfg(nr, &v[0], &v[1], &t1);
                if (kk > m) {
                    HH(kk, kk - 1) = v[0];
                    HH(kk + 1, kk - 1) = 0.0;
                    if (kk < i - 1) HH(kk + 2, kk - 1) = 0.0;
                } else if (m > l) {
                    HH(kk, kk - 1) = HH(kk, kk - 1) * (1.0 - t1);
                }
                const double v2 = v[1], t2 = t1 * v2;
                if (nr == 3) {
                    const double v3 = v[2], t3 = t1 * v3;
                    for (int j = kk; j <= i; ++j) {
                        const double sum = HH(kk, j) + v2 * HH(kk + 1, j) + v3 * HH(kk + 2, j);
                        HH(kk, j) -= sum * t1;
                        HH(kk + 1, j) -= sum * t2;
                        HH(kk + 2, j) -= sum * t3;
                    }
                    const int jmax = kk + 3 < i ? kk + 3 : i;
                    for (int j = l; j <= jmax; ++j) {
                        const double sum = HH(j, kk) + v2 * HH(j, kk + 1) + v3 * HH(j, kk + 2);
                        HH(j, kk) -= sum * t1;
                        HH(j, kk + 1) -= sum * t2;
                        HH(j, kk + 2) -= sum * t3;
                    }
                } else if (nr == 2) {
                    for (int j = kk; j <= i; ++j) {
                        const double sum = HH(kk, j) + v2 * HH(kk + 1, j);
                        HH(kk, j) -= sum * t1;
                        HH(kk + 1, j) -= sum * t2;
                    }
                    for (int j = l; j <= i; ++j) {
                        const double sum = HH(j, kk) + v2 * HH(j, kk + 1);
                        HH(j, kk) -= sum * t1;
                        HH(j, kk + 1) -= sum * t2;
                    }
                }
            }
        }
        if (!conv) {
            for (int r = 0; r <= i; ++r) { wr[r] = HH(r, r); wi[r] = 0.0; }
            return i + 1;
        }
        if (l == i) {
            wr[i] = HH(i, i);
            wi[i] = 0.0;
        } else {   /* l == i - 1: a 2x2 block */
            double cs, sn;
            oracle_dlanv2(&HH(i - 1, i - 1), &HH(i - 1, i), &HH(i, i - 1), &HH(i, i), &wr[i - 1], &wi[i - 1], &wr[i],
                          &wi[i], &cs, &sn);
        }
        kdefl = 0;
        i = l - 1;
    }
    return 0;
}

/* wr/wi[i + j*n]; info[j] = 0, the DLAHQR INFO, or -1 for a non-finite
 * input (then wr = wi = NaN). */
void oracle_dgeev_batched(int n, int batch, const double *A, double *wr, double *wi, int *info, int nthreads,
                          int itmax_override)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        double *H = (double *)malloc(sizeof(double) * (size_t)n * n);
        #pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < batch; ++b) {
            const size_t nn = (size_t)n * n;
            memcpy(H, A + (size_t)b * nn, sizeof(double) * nn);
            int finite = 1;
            for (size_t e = 0; e < nn; ++e) finite &= isfinite(H[e]) ? 1 : 0;
            double *w1 = wr + (size_t)b * n, *w2 = wi + (size_t)b * n;
            if (!finite) {
                for (int r = 0; r < n; ++r) { w1[r] = NAN; w2[r] = NAN; }
                info[b] = -1;
                continue;
            }
            oracle_dgehd2(n, H);
            info[b] = oracle_dlahqr(n, H, n, w1, w2, itmax_override);
        }
        free(H);
    }
}
