// Device building blocks of the real path (LAPACK DGEEV analogue,
// eigenvalues only, no balancing): DGEHD2 Householder Hessenberg reduction
// on a lane group's LDS matrix, and DLAHQR's Francis double-shift QR with
// DLANV2 for the converged 2x2 blocks.  No counterpart in the reference,
// which solves complex matrices only; the steps are LAPACK's, restated.
#pragma once
#include <float.h>

#include "zgeev_device.hpp"

namespace tkz {

// Upper Hessenberg storage with room for the double-shift bulge: column c
// holds rows 0..c+3 (the sub-diagonal and the two entries below it that a
// 3x3 reflector's bulge occupies for one chase step).
struct BulgeLayout {
    __device__ __forceinline__ int cb(int c) const { return (c * (c + 7)) >> 1; }
    __device__ __forceinline__ int operator()(int r, int c) const { return ((c * (c + 7)) >> 1) + r; }
};
__host__ __device__ constexpr int bulge_size(int n) { return (n * (n + 7)) / 2; }

// sqrt(x^2 + y^2) without destructive overflow (LAPACK DLAPY2; also DNRM2 of
// a 2-vector, to the bit).
__device__ __forceinline__ double dlapy2_d(double x, double y)
{
    const double xa = fabs(x), ya = fabs(y);
    const double w = fmax(xa, ya), z = fmin(xa, ya);
    if (z == 0.0) return w;
    const double q = z / w;
    return w * sqrt(1.0 + q * q);
}

// DLARFG for a 2- or 3-vector (alpha, x1[, x2]): tau = 0 leaves H = I;
// otherwise alpha becomes beta and (x1, x2) the scaled reflector tail.  In
// the normal range (the chase step's serial chain) one v_rsq and one v_rcp
// with Newton steps replace DLAPY2's sqrt and the three divisions; outside
// it the scaled LAPACK formulas run.
__device__ __forceinline__ void dlarfg_small(int nr, double &alpha, double &x1, double &x2, double &tau)
{
    if (nr == 2) x2 = 0.0;
    const double xx = x1 * x1 + x2 * x2;
    const double ss = alpha * alpha + xx;
    if (xx > 1e-290 && ss < 1e290) {
        const double r = rsq_nr(ss);                 // 1 / |(alpha, x)|
        const double beta = -copysign(ss * r, alpha);
        tau = (beta - alpha) * -copysign(r, alpha);  // (beta - alpha) / beta
        const double sc = rcp_nr(alpha - beta);      // |alpha - beta| >= |beta|: no cancellation
        x1 *= sc;
        x2 *= sc;
        alpha = beta;
        return;
    }
    const double xnorm = nr == 3 ? dlapy2_d(x1, x2) : fabs(x1);
    if (xnorm == 0.0) {
        tau = 0.0;
        return;
    }
    const double beta = -copysign(dlapy2_d(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    const double sc = 1.0 / (alpha - beta);
    x1 *= sc;
    x2 *= sc;
    alpha = beta;
}

// DGEHD2 by a group of LPM lanes on the n x n matrix A in LDS: rows are split
// over the lanes for the right update, columns for the left update.  Entries
// below the sub-diagonal are zero on return.
template <int LPM>
__device__ void dhessenberg(double *A, const FullLayout lay, int n, int t)
{
    for (int i = 0; i + 2 <= n; ++i) {
        const int m = n - 1 - i;   // reflector rows i+1 .. n-1
        double ssq = 0.0, amx = 0.0;
        for (int r = i + 2 + t; r < n; r += LPM) {
            const double v = A[lay(r, i)];
            ssq += v * v;
            amx = fmax(amx, fabs(v));
        }
        ssq = group_sum<LPM>(ssq);
        amx = group_max<LPM>(amx);
        const double alpha = A[lay(i + 1, i)];
        double xnorm = 0.0;
        if (amx > 0.0) {
            if (ssq > 1e-290 && ssq < 1e290) {
                xnorm = sqrt(ssq);
            } else {
                const double ia = 1.0 / amx;
                double s2 = 0.0;
                for (int r = i + 2 + t; r < n; r += LPM) {
                    const double v = A[lay(r, i)] * ia;
                    s2 += v * v;
                }
                xnorm = amx * sqrt(group_sum<LPM>(s2));
            }
        }
        if (xnorm == 0.0) continue;   // tau = 0: H(i) = I
        const double beta = -copysign(dlapy2_d(alpha, xnorm), alpha);
        const double tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        for (int r = i + 2 + t; r < n; r += LPM) A[lay(r, i)] *= scal;
        lds_order();
        const double *vcol = A + lay(i + 1, i);   // v(c) = vcol[c], c >= 1, v(0) = 1
        const int ld = lay.ld;
        // right: A(0:n, i+1:n) -= tau (A v) v^T
        for (int r = t; r < n; r += LPM) {
            double *row = A + r;
            double u = row[(i + 1) * ld];
            for (int c = 1; c < m; ++c) u += row[(i + 1 + c) * ld] * vcol[c];
            u *= tau;
            row[(i + 1) * ld] -= u;
            for (int c = 1; c < m; ++c) row[(i + 1 + c) * ld] -= u * vcol[c];
        }
        lds_order();
        // left: A(i+1:n, i+1:n) -= tau v (v^T A)
        for (int c = i + 1 + t; c < n; c += LPM) {
            double *col = A + lay(i + 1, c);
            double s = col[0];
            for (int r = 1; r < m; ++r) s += vcol[r] * col[r];
            s *= tau;
            col[0] -= s;
            for (int r = 1; r < m; ++r) col[r] -= vcol[r] * s;
        }
        lds_order();
        if (t == 0) A[lay(i + 1, i)] = beta;
        for (int r = i + 2 + t; r < n; r += LPM) A[lay(r, i)] = 0.0;
        lds_order();
    }
}

// DLANV2: eigenvalues (rt1r + i rt1i, rt2r + i rt2i) of the real 2x2
// [a b; c d] in standard form, a complex pair with the positive imaginary
// part first.  Only the eigenvalues are kept (the rotation is not needed
// without Schur vectors).
__device__ __forceinline__ void dlanv2_d(double A, double B, double C, double D, double &rt1r, double &rt1i,
                                         double &rt2r, double &rt2i)
{
    constexpr double multpl = 4.0, eps = DBL_EPSILON;   // DLAMCH('P')
    constexpr double safmn2 = 0x1p-485;                  // 2^int(log2(safmin / eps) / 2)
    constexpr double safmx2 = 0x1p485;
    if (C == 0.0) {
    } else if (B == 0.0) {
        const double temp = D;
        D = A;
        A = temp;
        B = -C;
        C = 0.0;
    } else if ((A - D) == 0.0 && copysign(1.0, B) != copysign(1.0, C)) {
    } else {
        double temp = A - D;
        double p = 0.5 * temp;
        const double bcmax = fmax(fabs(B), fabs(C));
        const double bcmis = fmin(fabs(B), fabs(C)) * copysign(1.0, B) * copysign(1.0, C);
        double scale = fmax(fabs(p), bcmax);
        double z = (p / scale) * p + (bcmax / scale) * bcmis;
        if (z >= multpl * eps) {   // real eigenvalues
            z = p + copysign(sqrt(scale) * sqrt(z), p);
            A = D + z;
            D = D - (bcmax / z) * bcmis;
            B = B - C;
            C = 0.0;
        } else {                   // complex or nearly equal real eigenvalues
            double sigma = B + C;
            for (int count = 1; count <= 21; ++count) {
                scale = fmax(fabs(temp), fabs(sigma));
                if (scale >= safmx2) {
                    sigma *= safmn2;
                    temp *= safmn2;
                    if (count <= 20) continue;
                }
                if (scale <= safmn2) {
                    sigma *= safmx2;
                    temp *= safmx2;
                    if (count <= 20) continue;
                }
                break;
            }
            p = 0.5 * temp;
            const double tau = dlapy2_d(sigma, temp);
            const double cs = sqrt(0.5 * (1.0 + fabs(sigma) / tau));
            const double sn = -(p / (tau * cs)) * copysign(1.0, sigma);
            const double aa = A * cs + B * sn, bb = -A * sn + B * cs;
            const double cc = C * cs + D * sn, dd = -C * sn + D * cs;
            A = aa * cs + cc * sn;
            B = bb * cs + dd * sn;
            C = -aa * sn + cc * cs;
            D = -bb * sn + dd * cs;
            temp = 0.5 * (A + D);
            A = temp;
            D = temp;
            if (C != 0.0) {
                if (B != 0.0) {
                    if (copysign(1.0, B) == copysign(1.0, C)) {   // real: upper triangular
                        const double sab = sqrt(fabs(B)), sac = sqrt(fabs(C));
                        p = copysign(sab * sac, C);
                        A = temp + p;
                        D = temp - p;
                        B = B - C;
                        C = 0.0;
                    }
                } else {
                    B = -C;
                    C = 0.0;
                }
            }
        }
    }
    rt1r = A;
    rt2r = D;
    if (C == 0.0) {
        rt1i = 0.0;
        rt2i = 0.0;
    } else {
        rt1i = sqrt(fabs(B)) * sqrt(fabs(C));
        rt2i = -rt1i;
    }
}

// DLAHQR, eigenvalues only, by a group of LPM >= n lanes on the Hessenberg
// matrix H in LDS (BulgeLayout, entries below the sub-diagonal zero).  The
// serial parts (shifts, the reflector of each chase step) are computed by
// every lane; lane p applies each reflector to column p (from the left) and
// row p (from the right).  Lane 0 writes wr/wi.  Returns 0 or the LAPACK INFO
// (1-based end of the unconverged block; its eigenvalues are then the
// diagonal of H).
template <int LPM>
__device__ int dlahqr_group(double *H, const BulgeLayout lay, int n, int t, int itmax_override,
                            double *__restrict__ wr, double *__restrict__ wi)
{
    if (n == 1) {
        if (t == 0) {
            wr[0] = H[lay(0, 0)];
            wi[0] = 0.0;
        }
        return 0;
    }
    const double ulp = DBL_EPSILON, smlnum = DBL_MIN * ((double)n / ulp);
    constexpr double dat1 = 0.75, dat2 = -0.4375;
    constexpr int kexsh = 10;
    const int itmax = itmax_override >= 0 ? itmax_override : 30 * (n > 10 ? n : 10);
    int kdefl = 0;
    int i = n - 1;
    while (i >= 0) {
        int l = 0;
        bool conv = false;
        for (int its = 0; its <= itmax; ++its) {
            // a single negligible sub-diagonal element (Ahues-Kressner test)
            const int k = group_last_true<LPM, 1>(t, l + 1, i, l, [&](int q) {
                const double hq = fabs(H[lay(q, q - 1)]);
                if (hq <= smlnum) return true;
                double tst = fabs(H[lay(q - 1, q - 1)]) + fabs(H[lay(q, q)]);
                if (tst == 0.0) {
                    if (q - 2 >= 0) tst += fabs(H[lay(q - 1, q - 2)]);
                    if (q + 1 <= n - 1) tst += fabs(H[lay(q + 1, q)]);
                }
                if (hq > ulp * tst) return false;
                const double up = fabs(H[lay(q - 1, q)]);
                const double dq = fabs(H[lay(q, q)]), df = fabs(H[lay(q - 1, q - 1)] - H[lay(q, q)]);
                const double ab = fmax(hq, up), ba = fmin(hq, up);
                const double aa = fmax(dq, df), bb = fmin(dq, df);
                const double s = aa + ab;
                return ba * (ab / s) <= fmax(smlnum, ulp * (bb * (aa / s)));
            });
            l = k;
            if (l > 0 && t == 0) H[lay(l, l - 1)] = 0.0;
            lds_order();
            if (l >= i - 1) {
                conv = true;
                break;
            }
            ++kdefl;
            // shifts: Francis double shift or an exceptional pair
            double h11, h12, h21, h22;
            if (kdefl % (2 * kexsh) == 0) {
                const double s = fabs(H[lay(i, i - 1)]) + fabs(H[lay(i - 1, i - 2)]);
                h11 = dat1 * s + H[lay(i, i)];
                h12 = dat2 * s;
                h21 = s;
                h22 = h11;
            } else if (kdefl % kexsh == 0) {
                const double s = fabs(H[lay(l + 1, l)]) + fabs(H[lay(l + 2, l + 1)]);
                h11 = dat1 * s + H[lay(l, l)];
                h12 = dat2 * s;
                h21 = s;
                h22 = h11;
            } else {
                h11 = H[lay(i - 1, i - 1)];
                h21 = H[lay(i, i - 1)];
                h12 = H[lay(i - 1, i)];
                h22 = H[lay(i, i)];
            }
            double rt1r, rt1i, rt2r, rt2i;
            {
                const double s = fabs(h11) + fabs(h12) + fabs(h21) + fabs(h22);
                if (s == 0.0) {
                    rt1r = rt1i = rt2r = rt2i = 0.0;
                } else {
                    h11 /= s;
                    h21 /= s;
                    h12 /= s;
                    h22 /= s;
                    const double tr = (h11 + h22) / 2.0;
                    const double det = (h11 - tr) * (h22 - tr) - h12 * h21;
                    const double rtdisc = sqrt(fabs(det));
                    if (det >= 0.0) {
                        rt1r = tr * s;
                        rt2r = rt1r;
                        rt1i = rtdisc * s;
                        rt2i = -rt1i;
                    } else {
                        rt1r = tr + rtdisc;
                        rt2r = tr - rtdisc;
                        if (fabs(rt1r - h22) <= fabs(rt2r - h22)) {
                            rt1r *= s;
                            rt2r = rt1r;
                        } else {
                            rt2r *= s;
                            rt1r = rt2r;
                        }
                        rt1i = rt2i = 0.0;
                    }
                }
            }
            // first column of the double-shift polynomial at row q, scaled
            auto first_col = [&](int q, double &v0, double &v1, double &v2) {
                double h21s = H[lay(q + 1, q)];
                double s = fabs(H[lay(q, q)] - rt2r) + fabs(rt2i) + fabs(h21s);
                h21s = H[lay(q + 1, q)] / s;
                v0 = h21s * H[lay(q, q + 1)] + (H[lay(q, q)] - rt1r) * ((H[lay(q, q)] - rt2r) / s) - rt1i * (rt2i / s);
                v1 = h21s * (H[lay(q, q)] + H[lay(q + 1, q + 1)] - rt1r - rt2r);
                v2 = h21s * H[lay(q + 2, q + 1)];
                s = fabs(v0) + fabs(v1) + fabs(v2);
                v0 /= s;
                v1 /= s;
                v2 /= s;
            };
            // two consecutive small sub-diagonal elements: the start row m
            const int m = group_last_true<LPM, 1>(t, l + 1, i - 2, l, [&](int q) {
                double v0, v1, v2;
                first_col(q, v0, v1, v2);
                const double h00 = fabs(H[lay(q, q - 1)]) * (fabs(v1) + fabs(v2));
                const double h01 = fabs(v0) * (fabs(H[lay(q - 1, q - 1)]) + fabs(H[lay(q, q)]) + fabs(H[lay(q + 1, q + 1)]));
                return h00 <= ulp * h01;
            });
            double v0, v1, v2;
            first_col(m, v0, v1, v2);
            // double-shift sweep
            for (int kk = m; kk <= i - 1; ++kk) {
                const int nr = (i - kk + 1) < 3 ? (i - kk + 1) : 3;
                if (kk > m) {
                    v0 = H[lay(kk, kk - 1)];
                    v1 = H[lay(kk + 1, kk - 1)];
                    v2 = nr == 3 ? H[lay(kk + 2, kk - 1)] : 0.0;
                }
                double t1;
                dlarfg_small(nr, v0, v1, v2, t1);
                if (t == 0) {
                    if (kk > m) {
                        H[lay(kk, kk - 1)] = v0;
                        H[lay(kk + 1, kk - 1)] = 0.0;
                        if (kk < i - 1) H[lay(kk + 2, kk - 1)] = 0.0;
                    } else if (m > l) {
                        H[lay(kk, kk - 1)] *= (1.0 - t1);
                    }
                }
                lds_order();
                const double t2 = t1 * v1, t3 = t1 * v2;
                const int p = t;
                if (nr == 3) {
                    if (p >= kk && p <= i) {   // left: column p, rows kk..kk+2
                        const int b = lay.cb(p) + kk;
                        const double x0 = H[b], x1 = H[b + 1], x2 = H[b + 2];
                        const double sum = x0 + v1 * x1 + v2 * x2;
                        H[b] = x0 - sum * t1;
                        H[b + 1] = x1 - sum * t2;
                        H[b + 2] = x2 - sum * t3;
                    }
                    lds_order();
                    if (p >= l && p <= (kk + 3 < i ? kk + 3 : i)) {   // right: row p, columns kk..kk+2
                        const int a0 = lay(p, kk), a1 = lay(p, kk + 1), a2 = lay(p, kk + 2);
                        const double x0 = H[a0], x1 = H[a1], x2 = H[a2];
                        const double sum = x0 + v1 * x1 + v2 * x2;
                        H[a0] = x0 - sum * t1;
                        H[a1] = x1 - sum * t2;
                        H[a2] = x2 - sum * t3;
                    }
                } else {
                    if (p >= kk && p <= i) {
                        const int b = lay.cb(p) + kk;
                        const double x0 = H[b], x1 = H[b + 1];
                        const double sum = x0 + v1 * x1;
                        H[b] = x0 - sum * t1;
                        H[b + 1] = x1 - sum * t2;
                    }
                    lds_order();
                    if (p >= l && p <= i) {
                        const int a0 = lay(p, kk), a1 = lay(p, kk + 1);
                        const double x0 = H[a0], x1 = H[a1];
                        const double sum = x0 + v1 * x1;
                        H[a0] = x0 - sum * t1;
                        H[a1] = x1 - sum * t2;
                    }
                }
                lds_order();
            }
        }
        if (!conv) {
            for (int r = t; r <= i; r += LPM) {
                wr[r] = H[lay(r, r)];
                wi[r] = 0.0;
            }
            return i + 1;
        }
        if (t == 0) {
            if (l == i) {
                wr[i] = H[lay(i, i)];
                wi[i] = 0.0;
            } else {
                double a, b, c, d;
                dlanv2_d(H[lay(i - 1, i - 1)], H[lay(i - 1, i)], H[lay(i, i - 1)], H[lay(i, i)], a, b, c, d);
                wr[i - 1] = a;
                wi[i - 1] = b;
                wr[i] = c;
                wi[i] = d;
            }
        }
        kdefl = 0;
        i = l - 1;
    }
    return 0;
}

}  // namespace tkz
