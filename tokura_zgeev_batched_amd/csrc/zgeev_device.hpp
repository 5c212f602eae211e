// Device building blocks of the batched complex eigenvalue solver for CDNA4
// (gfx950): complex FP64 helpers, lane-group reductions, Householder
// Hessenberg reduction and single-shift complex QR, all on one matrix held in
// LDS by a group of LPM lanes of one wavefront.
//
// Method (reference: tokura_zgeev_batched_library/src/tokura_zgeev_batched.cu:266-278,
// Hessenberg tokura_zgehrd_MWB_kernel.cuh:159-287, QR
// tokura_zhseqr_MWB_kernel.cuh:611-791), with the arithmetic of LAPACK's
// ZGEHD2 / ZLARFG / ZLAHQR (eigenvalues only, no balancing):
//  * Hessenberg: for i = 0..n-2 a reflector from column i, applied from the
//    right to rows 0..n-1 and from the left to rows/cols i+1..n-1.
//  * QR: Ahues-Kressner deflation test, Wilkinson shift with exceptional
//    shifts every 10th/20th sweep without deflation, the two-consecutive-
//    small-subdiagonals start, one 2-element reflector per chase step, a cap
//    of 30*max(10,n) sweeps per eigenvalue.
//
// Work split of a chase step k (rows/cols k, k+1 of the active block [l, i]):
//  * every lane of the group computes the reflector and the 2x2 "window"
//    H(k:k+1, k:k+1) plus H(k+2, k+1) redundantly - this is the serial chain
//    that carries the bulge to the next step;
//  * lane position p >= k+2 applies the reflector from the left to column p
//    (rows k, k+1), position p <= k-1 applies it from the right to row p
//    (cols k, k+1).  Each position keeps the element it will touch again at
//    step k+1 in a register ("carry"), so a step costs one LDS read and one
//    or two LDS writes per position.
// Only LDS traffic crosses lanes; a wavefront executes its LDS operations in
// order, so ordering a write before a read of another lane of the same wave
// needs only a compiler barrier (lds_order()).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <type_traits>

#include "zgeev_launch.hpp"

namespace tkz {

struct __align__(16) zd {
    double re, im;
};

__device__ __forceinline__ zd mk(double r, double i) { zd z; z.re = r; z.im = i; return z; }
__device__ __forceinline__ zd zadd(zd a, zd b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ zd zsub(zd a, zd b) { return mk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ zd zmul(zd a, zd b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
// conj(a) * b
__device__ __forceinline__ zd zmulc(zd a, zd b) { return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re); }
__device__ __forceinline__ zd zconj(zd a) { return mk(a.re, -a.im); }
__device__ __forceinline__ zd zscale(double s, zd a) { return mk(s * a.re, s * a.im); }
__device__ __forceinline__ double cabs1(zd a) { return fabs(a.re) + fabs(a.im); }
__device__ __forceinline__ double zabs(zd a) { return hypot(a.re, a.im); }
__device__ __forceinline__ bool zfinite(zd a) { return isfinite(a.re) && isfinite(a.im); }

// Robust complex division x / y (Smith's method with Stewart's ordering).
__device__ __forceinline__ zd zdiv(zd x, zd y)
{
    if (fabs(y.re) >= fabs(y.im)) {
        double r = y.im / y.re, den = y.re + y.im * r;
        return mk((x.re + x.im * r) / den, (x.im - x.re * r) / den);
    } else {
        double r = y.re / y.im, den = y.im + y.re * r;
        return mk((x.re * r + x.im) / den, (x.im * r - x.re) / den);
    }
}

// Principal complex square root.
__device__ __forceinline__ zd zsqrt(zd z)
{
    if (z.re == 0.0 && z.im == 0.0) return mk(0.0, z.im);
    double r = zabs(z);
    if (z.re >= 0.0) {
        double s = sqrt(0.5 * (r + z.re));
        return mk(s, z.im / (2.0 * s));
    } else {
        double s = sqrt(0.5 * (r - z.re));
        return mk(fabs(z.im) / (2.0 * s), copysign(s, z.im));
    }
}

// Compiler-level ordering of LDS accesses between lanes of one wavefront.
__device__ __forceinline__ void lds_order() { __asm__ __volatile__("" ::: "memory"); }

// ---- lane groups: LPM consecutive lanes of a wavefront own one matrix ----
template <int LPM>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int o = LPM / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPM);
    return v;
}
template <int LPM>
__device__ __forceinline__ double group_max(double v)
{
#pragma unroll
    for (int o = LPM / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, LPM));
    return v;
}
// Bits of `p` for the lanes of the caller's group, bit t = group lane t.
template <int LPM>
__device__ __forceinline__ uint64_t group_ballot(bool p)
{
    uint64_t b = __ballot(p);
    if constexpr (LPM == 64) {
        return b;
    } else {
        const int g = (int)(threadIdx.x & 63) / LPM;
        return (b >> (g * LPM)) & ((1ull << LPM) - 1ull);
    }
}
template <int LPM>
__device__ __forceinline__ bool group_any(bool p) { return group_ballot<LPM>(p) != 0ull; }

// ---- matrix layouts in LDS: lay(r, c) = cb(c) + r ----
struct FullLayout {   // column-major with odd leading dimension (bank spread)
    int ld;
    __device__ __forceinline__ int cb(int c) const { return c * ld; }
    __device__ __forceinline__ int operator()(int r, int c) const { return r + c * ld; }
};
struct PackedLayout { // upper Hessenberg, column c holds rows 0..c+1
    __device__ __forceinline__ int cb(int c) const { return (c * (c + 3)) >> 1; }
    __device__ __forceinline__ int operator()(int r, int c) const { return ((c * (c + 3)) >> 1) + r; }
};
__host__ __device__ constexpr int packed_size(int n) { return n <= 0 ? 0 : ((n - 1) * (n + 2)) / 2 + n; }

// ---- fast FP64 reciprocal and reciprocal square root ----
// The hardware estimates (v_rcp_f64, v_rsq_f64) are good to about 2^-24
// relative (5.2e-8 measured over 2^20 arguments on gfx950), so one
// third-order correction leaves a truncation error near 1e-22, below the
// final rounding: with e = 1 - x r (e = 1 - x y^2),
//   1/x       = r (1 + e + e^2 + ...)
//   1/sqrt(x) = y (1 + e/2 + 3e^2/8 + ...).
// Valid for normal arguments well inside the exponent range (callers
// guarantee 1e-290 < x < 1e290), accurate to about half an ulp.
__device__ __forceinline__ double rcp_nr(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    return fma(r, fma(e, e, e), r);
}
__device__ __forceinline__ double rsq_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// ---- ZLARFG for a 2-vector (x1, x2), x2 real: H^H [x1; x2] = [beta; 0] ----
// In the single-shift sweep the second entry is always real (it is a bulge
// created from a real sub-diagonal), as ZLAHQR relies on.  On return
// t1 = tau, v2 = scaled x2, t2 = Re(tau v2) = -x2/beta, beta; ident = true
// when tau = 0 (then LAPACK leaves alpha = x1 and x2 untouched).
// beta = -sign(Re x1) * |(x1, x2)|, tau = (beta - x1)/beta, v2 = x2/(x1 - beta).
// In the normal range the square root and both reciprocals come from one
// rsq and one rcp estimate (|x1 - beta|^2 = (|Re x1| + |beta|)^2 + Im(x1)^2 is
// within [ss, 4 ss]); outside it a scaled, exactly rounded path is used.  The
// rescaling loop of ZLARFG for |beta| < safmin/eps (~1e-292) is not reproduced.
__device__ __forceinline__ void zlarfg2(zd x1, double x2, zd &t1, zd &v2, double &t2, double &beta, bool &ident)
{
    ident = (x2 == 0.0 && x1.im == 0.0);
    const double ss = fma(x1.re, x1.re, fma(x1.im, x1.im, x2 * x2));
    double ib;
    zd rz;
    if (ss > 1e-290 && ss < 1e290) {
        const double y = rsq_nr(ss);
        double nrm = ss * y;
        nrm = fma(fma(-nrm, nrm, ss), 0.5 * y, nrm);
        beta = -copysign(nrm, x1.re);
        ib = -copysign(y, x1.re);
        const double zr = x1.re - beta;
        const double iz = rcp_nr(zr * zr + x1.im * x1.im);
        rz = mk(zr * iz, -x1.im * iz);
    } else {
        double w = fmax(fmax(fabs(x1.re), fabs(x1.im)), fabs(x2));
        double iw = w > 0.0 ? 1.0 / w : 0.0;
        double a = x1.re * iw, b = x1.im * iw, c = x2 * iw;
        beta = -copysign(w * sqrt(a * a + b * b + c * c), x1.re);
        ib = 1.0 / beta;
        rz = zdiv(mk(1.0, 0.0), mk(x1.re - beta, x1.im));
    }
    t1 = mk((beta - x1.re) * ib, -x1.im * ib);
    v2 = mk(x2 * rz.re, x2 * rz.im);
    t2 = -x2 * ib;
    if (ident) { t1 = mk(0.0, 0.0); v2 = mk(x2, 0.0); t2 = 0.0; beta = x1.re; }
}

// zlarfg2 for the bulge chase (one or two bulges; the m > l restart of a
// single-bulge sweep recomputes its first reflector with it too): the same
// arithmetic, with ZLARFG's tau = 0
// (H = I for x2 = Im x1 = 0) kept off the fast range only, where it also
// covers x = 0.  In the fast range such an x gets beta = -x1, tau = 2 and
// v2 = 0: the reflection diag(-1, 1), an exact sign change of row and column
// k, a unitary similarity like the identity, and the hot path loses the
// select of the tau = 0 case.  Later steps see the same values up to signs;
// ZLAHQR's shift takes sqrt(H(i-1,i)) sqrt(H(i,i-1)), and the complex square
// root of a negated number rounds differently, so such a matrix can get
// eigenvalues that differ from the tau = 0 choice in the last bits (about
// one N(0,1) matrix in 65,536 at n = 64).  H(k, k-1) = beta + 0i in both.
__device__ __forceinline__ void zlarfg2c(zd x1, double x2, zd &t1, zd &v2, double &t2, double &beta)
{
    const double ss = fma(x1.re, x1.re, fma(x1.im, x1.im, x2 * x2));
    if (ss > 1e-290 && ss < 1e290) {
        const double y = rsq_nr(ss);
        double nrm = ss * y;
        nrm = fma(fma(-nrm, nrm, ss), 0.5 * y, nrm);
        beta = -copysign(nrm, x1.re);
        const double ib = -copysign(y, x1.re);
        const double zr = x1.re - beta;
        const double iz = rcp_nr(zr * zr + x1.im * x1.im);
        t1 = mk((beta - x1.re) * ib, -x1.im * ib);
        v2 = mk(x2 * (zr * iz), x2 * (-x1.im * iz));
        t2 = -x2 * ib;
    } else {
        bool ident;
        zlarfg2(x1, x2, t1, v2, t2, beta, ident);
    }
}

// Application of a 2-element reflector to one element pair (x, y):
//   s = c1 x + t2 y,  x <- x - s,  y <- y - s v
// with (c1, v) = (conj(tau), v2) from the left (rows k, k+1) and (tau, conj(v2))
// from the right (columns k, k+1); t2 = Re(tau v2) is real.  12 FP64 ops.
__device__ __forceinline__ void pair_op(zd &x, zd &y, const zd c1, const double t2, const zd v)
{
    const double sr = fma(c1.re, x.re, fma(-c1.im, x.im, t2 * y.re));
    const double si = fma(c1.re, x.im, fma(c1.im, x.re, t2 * y.im));
    x.re -= sr;
    x.im -= si;
    y.re = fma(-sr, v.re, fma(si, v.im, y.re));
    y.im = fma(-sr, v.im, fma(-si, v.re, y.im));
}

// pair_op with the parts of c1 and v as separate values, so that a caller can
// pass negated parts (conjugates, per-lane signs) that the compiler folds into
// the FMAs' sign modifiers instead of forming new register pairs; the same
// roundings as pair_op for the same values.
__device__ __forceinline__ void pair_op_s(zd &x, zd &y, const double cr, const double ci, const double t2,
                                          const double vr, const double vi)
{
    const double sr = fma(cr, x.re, fma(-ci, x.im, t2 * y.re));
    const double si = fma(cr, x.im, fma(ci, x.re, t2 * y.im));
    x.re -= sr;
    x.im -= si;
    y.re = fma(-sr, vr, fma(si, vi, y.re));
    y.im = fma(-sr, vi, fma(-si, vr, y.im));
}

// ---- ZGEHD2 on an n x n matrix in LDS (full layout), LPM lanes ----
// Entries below the sub-diagonal are zeroed (LAPACK stores the reflectors
// there; nothing later reads them).
// With KEEPV the reflector vectors stay below the sub-diagonal (LAPACK's
// storage, v_0 = 1 implied) and taus[i] receives tau of step i, so that
// form_q can accumulate Q = H(0) H(1) ... H(n-2).
template <int LPM, bool KEEPV = false>
__device__ void hessenberg(zd *A, const FullLayout lay, int n, int t, zd *taus = nullptr)
{
    for (int i = 0; i + 2 <= n; ++i) {
        const int m = n - 1 - i;      // reflector length, rows i+1..n-1
        double ssq = 0.0, amx = 0.0;
        for (int r = i + 2 + t; r < n; r += LPM) {
            zd v = A[lay(r, i)];
            ssq += v.re * v.re + v.im * v.im;
            amx = fmax(amx, fmax(fabs(v.re), fabs(v.im)));
        }
        ssq = group_sum<LPM>(ssq);
        amx = group_max<LPM>(amx);
        zd alpha = A[lay(i + 1, i)];
        double xnorm = 0.0;
        if (amx > 0.0) {
            if (ssq > 1e-290 && ssq < 1e290) {
                xnorm = sqrt(ssq);
            } else {
                double ia = 1.0 / amx, s2 = 0.0;
                for (int r = i + 2 + t; r < n; r += LPM) {
                    zd v = A[lay(r, i)];
                    s2 += (v.re * ia) * (v.re * ia) + (v.im * ia) * (v.im * ia);
                }
                xnorm = amx * sqrt(group_sum<LPM>(s2));
            }
        }
        if (xnorm == 0.0 && alpha.im == 0.0) {   // tau = 0, H(i) = I
            if constexpr (KEEPV) {
                if (t == 0) taus[i] = mk(0.0, 0.0);
            }
            continue;
        }
        // beta = -sign(alpha_r) * |(alpha, xnorm)|  (DLAPY3 with scaling)
        double w = fmax(fmax(fabs(alpha.re), fabs(alpha.im)), xnorm);
        double iw = 1.0 / w;
        double ar = alpha.re * iw, ai = alpha.im * iw, xn = xnorm * iw;
        double beta = -copysign(w * sqrt(ar * ar + ai * ai + xn * xn), alpha.re);
        zd tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
        zd scal = zdiv(mk(1.0, 0.0), mk(alpha.re - beta, alpha.im));
        for (int r = i + 2 + t; r < n; r += LPM) A[lay(r, i)] = zmul(A[lay(r, i)], scal);
        lds_order();
        // right: A(0:n-1, i+1:n-1) -= tau (A v) v^H   (v_0 = 1 at row i+1).
        // Two accumulators and unrolled loops keep several LDS loads in flight.
        const zd *vcol = A + lay(i + 1, i);   // v(c) = vcol[c], c >= 1
        for (int r = t; r < n; r += LPM) {
            zd *row = A + r;
            const int ld = lay.ld;
            zd u0 = row[(i + 1) * ld], u1 = mk(0.0, 0.0);
            int c = 1;
#pragma unroll 2
            for (; c + 1 < m; c += 2) {
                const zd a0 = row[(i + 1 + c) * ld], v0 = vcol[c];
                const zd a1 = row[(i + 2 + c) * ld], v1 = vcol[c + 1];
                u0 = zadd(u0, zmul(a0, v0));
                u1 = zadd(u1, zmul(a1, v1));
            }
            if (c < m) u0 = zadd(u0, zmul(row[(i + 1 + c) * ld], vcol[c]));
            const zd u = zmul(zadd(u0, u1), tau);
            row[(i + 1) * ld] = zsub(row[(i + 1) * ld], u);
#pragma unroll 4
            for (c = 1; c < m; ++c) {
                const zd v = vcol[c];
                row[(i + 1 + c) * ld] = zsub(row[(i + 1 + c) * ld], zmul(u, zconj(v)));
            }
        }
        lds_order();
        // left: A(i+1:n-1, i+1:n-1) -= conj(tau) v (v^H A)
        for (int c = i + 1 + t; c < n; c += LPM) {
            zd *col = A + lay(i + 1, c);
            zd s0 = col[0], s1 = mk(0.0, 0.0);
            int r = 1;
#pragma unroll 2
            for (; r + 1 < m; r += 2) {
                s0 = zadd(s0, zmulc(vcol[r], col[r]));
                s1 = zadd(s1, zmulc(vcol[r + 1], col[r + 1]));
            }
            if (r < m) s0 = zadd(s0, zmulc(vcol[r], col[r]));
            const zd sv = zmulc(tau, zadd(s0, s1));
            col[0] = zsub(col[0], sv);
#pragma unroll 4
            for (r = 1; r < m; ++r) col[r] = zsub(col[r], zmul(vcol[r], sv));
        }
        lds_order();
        if (t == 0) A[lay(i + 1, i)] = mk(beta, 0.0);
        if constexpr (KEEPV) {
            if (t == 0) taus[i] = tau;
        } else {
            for (int r = i + 2 + t; r < n; r += LPM) A[lay(r, i)] = mk(0.0, 0.0);
        }
        lds_order();
    }
}

// ---- ZGEHD2 by a whole workgroup of NT threads on one LDS matrix ----
// Same algorithm as hessenberg<>(); each row/column dot product is split over
// TPR consecutive threads and combined with shuffles.  Three workgroup
// barriers per column: after the reflector vector is scaled, after the
// right update, and after the left update.  The threads that update column
// i+1 in the left update also form the sum of squares and max of its tail
// (the next reflector's x), so no separate norm pass is needed; entries
// below the sub-diagonal are zeroed once at the end.  `red` is LDS scratch
// of at least 2 doubles.  With KEEPV the reflector vectors stay below the
// sub-diagonal (v_0 = 1 implied) and taus[i] (LDS) receives tau of step i,
// as hessenberg<LPM, true>, for form_q_wg.
template <int NT, int TPR, bool KEEPV = false>
__device__ void hessenberg_wg(zd *A, const FullLayout lay, int n, double *red, zd *taus = nullptr)
{
    const int tid = threadIdx.x;
    const int q = tid % TPR, row = tid / TPR;     // lane within the row team, row/column slot
    const int ld = lay.ld;
    // tail norm of column 0 (rows 2..n-1) by the first team
    if (row == 0) {
        double ssq = 0.0, amx = 0.0;
        for (int r = 2 + q; r < n; r += TPR) {
            const zd v = A[lay(r, 0)];
            ssq += v.re * v.re + v.im * v.im;
            amx = fmax(amx, fmax(fabs(v.re), fabs(v.im)));
        }
#pragma unroll
        for (int o = 1; o < TPR; o <<= 1) {
            ssq += __shfl_xor(ssq, o, TPR);
            amx = fmax(amx, __shfl_xor(amx, o, TPR));
        }
        if (q == 0) { red[0] = ssq; red[1] = amx; }
    }
    __syncthreads();
    for (int i = 0; i + 2 <= n; ++i) {
        const int m = n - 1 - i;
        const double ssq = red[0], amx = red[1];
        const zd alpha = A[lay(i + 1, i)];
        double xnorm = 0.0;
        if (amx > 0.0) {
            if (ssq > 1e-290 && ssq < 1e290) {
                xnorm = sqrt(ssq);
            } else {                      // rare: rescaled second pass
                const double ia = 1.0 / amx;
                double s2 = 0.0;
                if (row == 0)
                    for (int r = i + 2 + q; r < n; r += TPR) {
                        const zd v = A[lay(r, i)];
                        s2 += (v.re * ia) * (v.re * ia) + (v.im * ia) * (v.im * ia);
                    }
#pragma unroll
                for (int o = 1; o < TPR; o <<= 1) s2 += __shfl_xor(s2, o, TPR);
                __syncthreads();          // everyone has read red[]
                if (tid == 0) red[0] = s2;
                __syncthreads();
                xnorm = amx * sqrt(red[0]);
            }
        }
        if (xnorm == 0.0 && alpha.im == 0.0) {   // tau = 0, H(i) = I; next tail norm
            if (KEEPV && tid == 0) taus[i] = mk(0.0, 0.0);
            __syncthreads();                      // everyone has read red[] and alpha
            if (row == 0) {
                double s = 0.0, mx = 0.0;
                for (int r = i + 3 + q; r < n; r += TPR) {
                    const zd v = A[lay(r, i + 1)];
                    s += v.re * v.re + v.im * v.im;
                    mx = fmax(mx, fmax(fabs(v.re), fabs(v.im)));
                }
#pragma unroll
                for (int o = 1; o < TPR; o <<= 1) {
                    s += __shfl_xor(s, o, TPR);
                    mx = fmax(mx, __shfl_xor(mx, o, TPR));
                }
                if (q == 0) { red[0] = s; red[1] = mx; }
            }
            __syncthreads();
            continue;
        }
        double w = fmax(fmax(fabs(alpha.re), fabs(alpha.im)), xnorm);
        double iw = 1.0 / w;
        double ar = alpha.re * iw, ai = alpha.im * iw, xn = xnorm * iw;
        const double beta = -copysign(w * sqrt(ar * ar + ai * ai + xn * xn), alpha.re);
        const zd tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
        const zd scal = zdiv(mk(1.0, 0.0), mk(alpha.re - beta, alpha.im));
        if (KEEPV && tid == 0) taus[i] = tau;
        for (int r = i + 2 + tid; r < n; r += NT) A[lay(r, i)] = zmul(A[lay(r, i)], scal);
        __syncthreads();                 // v is scaled; red[] and alpha have been read
        const zd *vcol = A + lay(i + 1, i);   // v(c) = vcol[c], c >= 1, v(0) = 1
        // right: rows r = row (+ NT/TPR ...), team lanes split the columns
        for (int r = row; r < n; r += NT / TPR) {
            zd *rp = A + r;
            zd u = q == 0 ? rp[(i + 1) * ld] : mk(0.0, 0.0);
            for (int c = (q == 0 ? TPR : q); c < m; c += TPR) u = zadd(u, zmul(rp[(i + 1 + c) * ld], vcol[c]));
#pragma unroll
            for (int o = 1; o < TPR; o <<= 1) {
                u.re += __shfl_xor(u.re, o, TPR);
                u.im += __shfl_xor(u.im, o, TPR);
            }
            u = zmul(u, tau);
            if (q == 0) rp[(i + 1) * ld] = zsub(rp[(i + 1) * ld], u);
            for (int c = (q == 0 ? TPR : q); c < m; c += TPR)
                rp[(i + 1 + c) * ld] = zsub(rp[(i + 1 + c) * ld], zmul(u, zconj(vcol[c])));
        }
        __syncthreads();
        // left: columns c = i+1+row (...), team lanes split the rows; the team
        // of column i+1 also accumulates its tail (rows i+3..) for the next step
        for (int c = i + 1 + row; c < n; c += NT / TPR) {
            zd *col = A + lay(i + 1, c);
            zd sv = q == 0 ? col[0] : mk(0.0, 0.0);
            for (int r = (q == 0 ? TPR : q); r < m; r += TPR) sv = zadd(sv, zmulc(vcol[r], col[r]));
#pragma unroll
            for (int o = 1; o < TPR; o <<= 1) {
                sv.re += __shfl_xor(sv.re, o, TPR);
                sv.im += __shfl_xor(sv.im, o, TPR);
            }
            sv = zmulc(tau, sv);
            if (q == 0) col[0] = zsub(col[0], sv);
            double s = 0.0, mx = 0.0;
            for (int r = (q == 0 ? TPR : q); r < m; r += TPR) {
                const zd x = zsub(col[r], zmul(vcol[r], sv));
                col[r] = x;
                if (r >= 2) {
                    s += x.re * x.re + x.im * x.im;
                    mx = fmax(mx, fmax(fabs(x.re), fabs(x.im)));
                }
            }
            if (c == i + 1) {             // the whole team takes this branch together
#pragma unroll
                for (int o = 1; o < TPR; o <<= 1) {
                    s += __shfl_xor(s, o, TPR);
                    mx = fmax(mx, __shfl_xor(mx, o, TPR));
                }
                if (q == 0) { red[0] = s; red[1] = mx; }
            }
        }
        if (tid == 0) A[lay(i + 1, i)] = mk(beta, 0.0);
        __syncthreads();
    }
    if constexpr (!KEEPV) {   // the reflector vectors below the sub-diagonal are not kept
        for (int c = 0; c + 2 < n; ++c)
            for (int r = c + 2 + tid; r < n; r += NT) A[lay(r, c)] = mk(0.0, 0.0);
        __syncthreads();
    }
}

// ---- balancing (LAPACK ZGEBAL, the 3.11+ algorithm) ----
// job bit 0: permute rows/columns to isolate eigenvalues; bit 1: scale rows
// and columns by powers of two until their norms are comparable.  The matrix
// is held by a group of LPM lanes (positions p = t + s*LPM).  Only the
// balanced matrix is produced: the Hessenberg reduction and QR that follow
// run on the whole matrix, and the block triangular structure the
// permutations create survives both (reflectors of columns whose tail is
// already zero are the identity; the isolated eigenvalues deflate at once).
template <int LPM, int S>
__device__ void zgebal_group(zd *A, const FullLayout lay, int n, int t, int job)
{
    int k = 0, l = n - 1;
    // exchange column i with j over rows 0..l, then row i with j over columns k..n-1
    auto swap_rc = [&](int i, int j) {
        for (int r = t; r <= l; r += LPM) {
            const zd x = A[lay(r, i)];
            A[lay(r, i)] = A[lay(r, j)];
            A[lay(r, j)] = x;
        }
        lds_order();
        for (int c = k + t; c < n; c += LPM) {
            const zd x = A[lay(i, c)];
            A[lay(i, c)] = A[lay(j, c)];
            A[lay(j, c)] = x;
        }
        lds_order();
    };
    auto nonzero = [](const zd z) { return z.re != 0.0 || z.im != 0.0; };
    if (job & BAL_PERMUTE) {
        // rows with no off-diagonal nonzero in columns 0..l: push down
        bool noconv = true;
        while (noconv) {
            noconv = false;
            for (int i = l; i >= 0; --i) {
                bool nz = false;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int j = t + s * LPM;
                    if (j <= l && j != i) nz |= nonzero(A[lay(i, j)]);
                }
                if (!group_any<LPM>(nz)) {
                    if (i != l) swap_rc(i, l);
                    noconv = true;
                    if (l == 0) return;
                    --l;
                }
            }
        }
        // columns with no off-diagonal nonzero in rows k..l: push left
        noconv = true;
        while (noconv) {
            noconv = false;
            for (int j = k; j <= l; ++j) {
                bool nz = false;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int i = t + s * LPM;
                    if (i >= k && i <= l && i != j) nz |= nonzero(A[lay(i, j)]);
                }
                if (!group_any<LPM>(nz)) {
                    if (j != k) swap_rc(j, k);
                    noconv = true;
                    ++k;
                }
            }
        }
    }
    if (!(job & BAL_SCALE)) return;
    const double sclfac = 2.0, factor = 0.95;
    const double sfmin1 = DBL_MIN / DBL_EPSILON, sfmax1 = 1.0 / sfmin1;
    const double sfmin2 = sfmin1 * sclfac, sfmax2 = 1.0 / sfmin2;
    double sc[S];                                  // scale factor of position p
#pragma unroll
    for (int s = 0; s < S; ++s) sc[s] = 1.0;
    // first index of the largest |re| + |im| over the group; returns its modulus
    auto izamax_abs = [&](double v1, int idx, zd z) {
        double best = v1;
        int bi = idx;
#pragma unroll
        for (int o = LPM / 2; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, LPM);
            const int oi = __shfl_xor(bi, o, LPM);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        // the owner lane of bi holds the element in z when bi is its position
        double m = (idx == bi) ? zabs(z) : 0.0;
        return group_max<LPM>(m);
    };
    bool noconv = true;
    for (int sweep = 0; noconv && sweep < 100; ++sweep) {   // LAPACK converges in a few sweeps
        noconv = false;
        for (int i = k; i <= l; ++i) {
            // column i: rows k..l for the norm, rows 0..l for IZAMAX;
            // row i: columns k..l for the norm, columns k..n-1 for IZAMAX
            double cm = 0.0, rm = 0.0;
            double ckey = -1.0, rkey = -1.0;
            int cidx = 0x7fffffff, ridx = 0x7fffffff;
            zd cz = mk(0.0, 0.0), rz = cz;
            zd colv[S], rowv[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int q = t + s * LPM;
                colv[s] = (q <= l) ? A[lay(q, i)] : mk(0.0, 0.0);
                rowv[s] = (q >= k && q < n) ? A[lay(i, q)] : mk(0.0, 0.0);
                if (q >= k && q <= l) {
                    cm = fmax(cm, fmax(fabs(colv[s].re), fabs(colv[s].im)));
                    rm = fmax(rm, fmax(fabs(rowv[s].re), fabs(rowv[s].im)));
                }
                if (q <= l && cabs1(colv[s]) > ckey) { ckey = cabs1(colv[s]); cidx = q; cz = colv[s]; }
                if (q >= k && q < n && cabs1(rowv[s]) > rkey) { rkey = cabs1(rowv[s]); ridx = q; rz = rowv[s]; }
            }
            cm = group_max<LPM>(cm);
            rm = group_max<LPM>(rm);
            double cs = 0.0, rs = 0.0;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int q = t + s * LPM;
                if (q >= k && q <= l) {
                    if (cm > 0.0) {
                        const double a = colv[s].re / cm, b = colv[s].im / cm;
                        cs += a * a + b * b;
                    }
                    if (rm > 0.0) {
                        const double a = rowv[s].re / rm, b = rowv[s].im / rm;
                        rs += a * a + b * b;
                    }
                }
            }
            double c = cm * sqrt(group_sum<LPM>(cs));
            double r = rm * sqrt(group_sum<LPM>(rs));
            double ca = izamax_abs(ckey, cidx, cz);
            double ra = izamax_abs(rkey, ridx, rz);
            if (c == 0.0 || r == 0.0) continue;
            double g = r / sclfac, f = 1.0;
            const double ssum = c + r;
            while (c < g && fmax(f, fmax(c, ca)) < sfmax2 && fmin(r, fmin(g, ra)) > sfmin2) {
                f *= sclfac; c *= sclfac; ca *= sclfac; r /= sclfac; g /= sclfac; ra /= sclfac;
            }
            g = c / sclfac;
            while (g >= r && fmax(r, ra) < sfmax2 && fmin(fmin(f, c), fmin(g, ca)) > sfmin2) {
                f /= sclfac; c /= sclfac; g /= sclfac; ca /= sclfac; r *= sclfac; ra *= sclfac;
            }
            if (c + r >= factor * ssum) continue;
            double sci = 0.0;                       // scale factor of position i
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (t + s * LPM == i) sci = sc[s];
            sci = group_max<LPM>(sci);
            if (f < 1.0 && sci < 1.0 && f * sci <= sfmin1) continue;
            if (f > 1.0 && sci > 1.0 && sci >= sfmax1 / f) continue;
            g = 1.0 / f;
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (t + s * LPM == i) sc[s] = sci * f;
            noconv = true;
            for (int q = k + t; q < n; q += LPM) A[lay(i, q)] = zscale(g, A[lay(i, q)]);
            lds_order();
            for (int q = t; q <= l; q += LPM) A[lay(q, i)] = zscale(f, A[lay(q, i)]);
            lds_order();
        }
    }
}

// Q = H(0) H(1) ... H(n-2) from the reflectors kept by hessenberg<LPM, true>
// (LAPACK ZUNGHR), accumulated backwards: Q <- H(i) Q acts on rows and
// columns i+1..n-1 only.  Afterwards the reflector storage in A is zeroed.
template <int LPM>
__device__ void form_q(zd *Q, const FullLayout ql, zd *A, const FullLayout lay, const zd *taus, int n, int t)
{
    for (int c = t; c < n; c += LPM)
        for (int r = 0; r < n; ++r) Q[ql(r, c)] = mk(r == c ? 1.0 : 0.0, 0.0);
    lds_order();
    for (int i = n - 2; i >= 0; --i) {
        const zd tau = taus[i];
        if (tau.re == 0.0 && tau.im == 0.0) continue;
        const int m = n - 1 - i;
        for (int c = i + 1 + t; c < n; c += LPM) {
            zd s = Q[ql(i + 1, c)];
            for (int r = 1; r < m; ++r) s = zadd(s, zmulc(A[lay(i + 1 + r, i)], Q[ql(i + 1 + r, c)]));
            s = zmul(tau, s);
            Q[ql(i + 1, c)] = zsub(Q[ql(i + 1, c)], s);
            for (int r = 1; r < m; ++r) Q[ql(i + 1 + r, c)] = zsub(Q[ql(i + 1 + r, c)], zmul(A[lay(i + 1 + r, i)], s));
        }
        lds_order();
    }
    for (int c = 0; c + 2 < n; ++c)
        for (int r = c + 2 + t; r < n; r += LPM) A[lay(r, c)] = mk(0.0, 0.0);
    lds_order();
}

// Coalesced load of matrix `src` (n*n column-major) into LDS; returns
// whether all entries are finite (group-uniform).  With mx, also the largest
// |Re| or |Im| of the matrix (group-uniform).
template <int LPM>
__device__ __forceinline__ bool load_full(zd *H, const FullLayout lay, const zd *__restrict__ src, int n, int t,
                                          double *mx = nullptr)
{
    const int nn = n * n;
    bool bad = false;
    double m = 0.0;
    int r = t % n, c = t / n;
    const int dr = LPM % n, dc = LPM / n;
    for (int e = t; e < nn; e += LPM) {
        zd v = src[e];
        bad |= !zfinite(v);
        m = fmax(m, fmax(fabs(v.re), fabs(v.im)));
        H[lay(r, c)] = v;
        r += dr;
        c += dc;
        if (r >= n) { r -= n; ++c; }
    }
    if (mx) *mx = group_max<LPM>(m);
    return !group_any<LPM>(bad);
}

// ZGEEV's norm-range scaling: LAPACK scales A when its largest entry lies
// outside [SMLNUM, BIGNUM], SMLNUM = sqrt(safmin)/eps = 2^-459, BIGNUM =
// 1/SMLNUM, so that the QR's deflation threshold (safmin n / ulp) and its
// products of two entries stay clear of underflow and overflow.  Here the
// test is on m = max(|Re|, |Im|) over the entries and the factor is the power
// of two 2^e (exact) that brings m into [2^-458, 2^-457) or [2^458, 2^459);
// e = 0 when m is in range (then nothing is multiplied).  The eigenvalues
// are multiplied by 2^-e at the end.
__host__ __device__ inline int zgeev_scale_exp(double m)
{
    if (!(m > 0.0) || (m >= 0x1p-459 && m <= 0x1p459)) return 0;
    const int k = ilogb(m);
    return m < 0x1p-459 ? -458 - k : 458 - k;
}

// Multiplies the n x n LDS matrix by 2^e (positions t, t+LPM, ... of the
// column-major order).
template <int LPM>
__device__ __forceinline__ void scale_full(zd *H, const FullLayout lay, int n, int t, int e)
{
    const double f = ldexp(1.0, e);
    for (int k = t; k < n * n; k += LPM) {
        zd &v = H[lay(k % n, k / n)];
        v = zscale(f, v);
    }
}

// Store the upper Hessenberg part of an LDS matrix in packed order
// (column c: rows 0..min(c+1, n-1)); consecutive lanes write consecutive rows.
template <int LPM>
__device__ __forceinline__ void store_packed(const zd *H, const FullLayout lay, zd *__restrict__ dst, int n, int t)
{
    const PackedLayout pk;
    for (int c = 0; c < n; ++c) {
        const int rows = (c + 2 < n ? c + 2 : n);
        for (int r = t; r < rows; r += LPM) dst[pk(r, c)] = H[lay(r, c)];
    }
}

// Ahues-Kressner negligibility test of H(k, k-1) (ZLAHQR loop 40); lo, hi:
// the ILO, IHI of the call.
template <class Lay>
__device__ __forceinline__ bool subdiag_negligible(const zd *H, const Lay &lay, int k, int lo, int hi,
                                                   double smlnum, double ulp)
{
    zd h = H[lay(k, k - 1)];
    if (cabs1(h) <= smlnum) return true;
    zd hkm = H[lay(k - 1, k - 1)], hkk = H[lay(k, k)];
    double tst = cabs1(hkm) + cabs1(hkk);
    if (tst == 0.0) {
        if (k - 2 >= lo) tst += fabs(H[lay(k - 1, k - 2)].re);
        if (k + 1 <= hi) tst += fabs(H[lay(k + 1, k)].re);
    }
    if (fabs(h.re) <= ulp * tst) {
        zd hup = H[lay(k - 1, k)];
        double ab = fmax(cabs1(h), cabs1(hup)), ba = fmin(cabs1(h), cabs1(hup));
        zd dd = zsub(hkm, hkk);
        double aa = fmax(cabs1(hkk), cabs1(dd)), bb = fmin(cabs1(hkk), cabs1(dd));
        double s = aa + ab;
        if (ba * (ab / s) <= fmax(smlnum, ulp * (bb * (aa / s)))) return true;
    }
    return false;
}

// Largest position in [lo, hi] (positions p = t + s*LPM) whose predicate is
// true, or `none`.  pred(p) is evaluated only for p in [lo, hi].
template <int LPM, int S, class F>
__device__ __forceinline__ int group_last_true(int t, int lo, int hi, int none, F pred)
{
    int found = none;
#pragma unroll
    for (int s = S - 1; s >= 0; --s) {
        int p = t + s * LPM;
        bool q = (p >= lo && p <= hi) ? pred(p) : false;
        uint64_t bm = group_ballot<LPM>(q);
        if (found == none && bm != 0ull) found = s * LPM + (63 - __clzll((long long)bm));
    }
    return found;
}

// ZLAHQR's shift for the active block [l, i]: exceptional shifts on every
// 10th (from the top) and 20th (from the bottom) sweep without a deflation,
// else the Wilkinson shift (the eigenvalue of the trailing 2x2 closer to
// H(i,i)).
template <class Lay>
__device__ __forceinline__ zd zlahqr_shift(const zd *H, const Lay &lay, int l, int i, int kdefl)
{
    zd sh;
    if (kdefl % 20 == 0) {
        double s = 0.75 * fabs(H[lay(i, i - 1)].re);
        sh = zadd(H[lay(i, i)], mk(s, 0.0));
    } else if (kdefl % 10 == 0) {
        double s = 0.75 * fabs(H[lay(l + 1, l)].re);
        sh = zadd(H[lay(l, l)], mk(s, 0.0));
    } else {
        sh = H[lay(i, i)];
        zd u = zmul(zsqrt(H[lay(i - 1, i)]), zsqrt(H[lay(i, i - 1)]));
        double s = cabs1(u);
        if (s != 0.0) {
            zd x = zscale(0.5, zsub(H[lay(i - 1, i - 1)], sh));
            double sx = cabs1(x);
            s = fmax(s, sx);
            double is = 1.0 / s;
            zd xs = zscale(is, x), us = zscale(is, u);
            zd y = zscale(s, zsqrt(zadd(zmul(xs, xs), zmul(us, us))));
            if (sx > 0.0) {
                zd xn = zscale(1.0 / sx, x);
                if (xn.re * y.re + xn.im * y.im < 0.0) y = mk(-y.re, -y.im);
            }
            sh = zsub(sh, zmul(u, zdiv(u, zadd(x, y))));
        }
    }
    return sh;
}

// Shift snapping (a QR whose eigenvalues are already known, such as the
// recompute backward's, which is given the forward's eigenvalues): lane t of
// the group holds known eigenvalue t (`have`); returns the known eigenvalue
// nearest to mu (squared distance, ties to the lower lane), or mu when the
// group holds none (non-finite values never win).  With exact shifts a sweep
// deflates one eigenvalue, so the QR takes about half the chase steps of
// Wilkinson shifts on random matrices.  Only the first kSnapSweeps sweeps
// after a deflation are snapped: if the known values are not the matrix's
// eigenvalues, the Wilkinson shifts that follow still converge (about 1.3x
// the sweeps of an unsnapped QR, where snapping every sweep can stall).
constexpr int kSnapSweeps = 2;
template <int LPM>
__device__ __forceinline__ zd snap_shift(zd mu, zd kev, bool have, int t)
{
    const double dr = kev.re - mu.re, di = kev.im - mu.im;
    double d = dr * dr + di * di;
    if (!have || !(d <= __builtin_huge_val())) d = __builtin_huge_val();
    int b = t;
#pragma unroll
    for (int o = LPM / 2; o > 0; o >>= 1) {
        const double od = __shfl_xor(d, o, LPM);
        const int ob = __shfl_xor(b, o, LPM);
        if (od < d || (od == d && ob < b)) {
            d = od;
            b = ob;
        }
    }
    const zd k = mk(__shfl(kev.re, b, LPM), __shfl(kev.im, b, LPM));
    return d < __builtin_huge_val() ? k : mu;
}

// ---- ZLAHQR with ilo = 0, ihi = n-1 ----
// WANTT = false: eigenvalues only (transformations restricted to the active
// block [l, i], as ZLAHQR with WANTT = WANTZ = false).  WANTT = true: the full
// Schur form T (rows 0..min(k+2,i), columns k..n-1 per step) and, when Z is
// non-null, Z <- Z * G for every reflector (WANTZ).  Returns 0 or 1 + the
// 0-based index of the first eigenvalue that did not converge (LAPACK INFO).
// On return the eigenvalues are H(i,i).
// ZH ("Z on the upper half-wave", requires S == 1 and n <= LPM/2): lanes
// t < LPM/2 are the H positions, lanes t >= LPM/2 own row t - LPM/2 of Z and
// apply the reflectors to it in the same instructions that update H.
// RANGE: only the eigenvalues ilo..ihi are computed (ZLAHQR's ILO/IHI with
// H(ilo, ilo-1) = 0); with WANTT the transformations still span the whole of
// T and Z, so the routine finishes a block that a caller has split off.
// SNAP: the shift of the first kSnapSweeps sweeps after each deflation is
// replaced by the nearest known eigenvalue (snap_shift; lane t holds known
// eigenvalue t in kev when kh).
struct NoZ {
    __device__ __forceinline__ int cb(int) const { return 0; }
    __device__ __forceinline__ int operator()(int, int) const { return 0; }
};

template <int LPM, int S, bool WANTT, bool ZH = false, class Lay, class ZLay = NoZ, bool RANGE = false,
          bool SNAP = false>
__device__ int zlahqr_core(zd *H, const Lay lay, int n, int t, int itmax_override, zd *Z = nullptr,
                           const ZLay zlay = ZLay(), int ihi = 0, int ilo = 0, zd kev = zd{}, bool kh = false)
{
    static_assert(!ZH || (S == 1 && WANTT), "half-wave Z needs one position per lane");
    static_assert(!SNAP || (S == 1 && !ZH), "snapped shifts: one known eigenvalue per lane");
    constexpr int HALF = LPM / 2;
    const bool zlane = ZH && t >= HALF;            // this lane holds a row of Z
    const int zrow = ZH ? (t >= HALF ? t - HALF : 0) : 0;
    const bool zact = zlane && zrow < n;
    const int tp = zlane ? (1 << 20) : t;          // H position base (none for Z lanes)
    const double ulp = DBL_EPSILON, safmin = DBL_MIN;
    const int nh = RANGE ? ihi - ilo + 1 : n;
    const double smlnum = safmin * ((double)nh / ulp);
    const int dlo = RANGE ? ilo : 0, dhi = RANGE ? ihi : n - 1;   // bounds of the deflation test
    const int itmax = itmax_override >= 0 ? itmax_override : 30 * (nh > 10 ? nh : 10);
    const bool wantz = WANTT && Z != nullptr;
    const int lfloor = RANGE ? ilo : 0;
    int kdefl = 0;
    int i = RANGE ? ihi : n - 1;
    while (i >= lfloor + 1) {
        int l = lfloor;
        bool conv = false;
        for (int its = 0; its <= itmax; ++its) {
            // ---- look for a single small sub-diagonal element
            const int k = group_last_true<LPM, S>(t, l + 1, i, l, [&](int p) {
                return subdiag_negligible(H, lay, p, dlo, dhi, smlnum, ulp);
            });
            l = k;
            if (l > 0 && t == 0) H[lay(l, l - 1)] = mk(0.0, 0.0);
            lds_order();
            if (l >= i) { conv = true; break; }
            ++kdefl;
            const int I1 = WANTT ? 0 : l, I2 = WANTT ? n - 1 : i;
            // ---- shift
            zd sh = zlahqr_shift(H, lay, l, i, kdefl);
            if constexpr (SNAP) {
                if (kdefl <= kSnapSweeps) sh = snap_shift<LPM>(sh, kev, kh, t);
            }
            // ---- look for two consecutive small sub-diagonal elements
            const int m = group_last_true<LPM, S>(t, l + 1, i - 1, l, [&](int p) {
                zd h11 = H[lay(p, p)], h22 = H[lay(p + 1, p + 1)];
                zd h11s = zsub(h11, sh);
                double h21 = H[lay(p + 1, p)].re;
                double s = cabs1(h11s) + fabs(h21);
                h11s = zscale(1.0 / s, h11s);
                h21 = h21 / s;
                double h10 = H[lay(p, p - 1)].re;
                return fabs(h10) * fabs(h21) <= ulp * (cabs1(h11s) * (cabs1(h11) + cabs1(h22)));
            });
            zd x1;
            double x2;
            {
                zd h11s = zsub(H[lay(m, m)], sh);
                double h21 = H[lay(m + 1, m)].re;
                double s = cabs1(h11s) + fabs(h21);
                x1 = mk(h11s.re / s, h11s.im / s);
                x2 = h21 / s;
            }
            // ---- single-shift sweep k = m .. i-1; H rows I1.., columns ..I2
            // State between steps: reflector input (x1, x2), window
            // a = H(k,k), b = H(k,k+1), c = H(k+1,k), d = H(k+1,k+1),
            // e = H(k+2,k+1); per position the carried element and the
            // prefetched partner (ypre), and for Z the carried Z(p,k) and
            // Z(p,k+1), the latter read two steps ahead into zyE (steps of
            // even k - m) or zyO (odd) so that no register move waits for a
            // load in flight.
            zd carry[S], ypre[S], zc[S], zyE[S], zyO[S];
            int cbp[S];   // column base of the position's own column
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int p = tp + s * LPM;
                cbp[s] = lay.cb(p < n ? p : n - 1);
                carry[s] = mk(0.0, 0.0);
                ypre[s] = mk(0.0, 0.0);
                if (p >= m + 2 && p <= I2) { carry[s] = H[cbp[s] + m]; ypre[s] = H[cbp[s] + m + 1]; }
                else if (p >= I1 && p <= m - 1) { carry[s] = H[lay(p, m)]; ypre[s] = H[lay(p, m + 1)]; }
                if constexpr (ZH) {
                    if (zact) { carry[s] = Z[zlay(zrow, m)]; ypre[s] = Z[zlay(zrow, m + 1)]; }
                }
                if constexpr (WANTT && !ZH) {
                    zc[s] = mk(0.0, 0.0);
                    zyE[s] = mk(0.0, 0.0);
                    zyO[s] = mk(0.0, 0.0);
                    if (wantz && p < n) {
                        zc[s] = Z[zlay(p, m)];
                        zyE[s] = Z[zlay(p, m + 1)];
                        if (m + 2 <= i) zyO[s] = Z[zlay(p, m + 2)];
                    }
                }
            }
            zd a = H[lay(m, m)], b = H[lay(m, m + 1)], c = H[lay(m + 1, m)], d = H[lay(m + 1, m + 1)];
            zd e = (m + 2 <= i) ? H[lay(m + 2, m + 1)] : mk(0.0, 0.0);

            // one chase step kk; returns the new bulge H(kk+2, kk)
            auto step = [&](auto PAR, const int kk, const bool first, const bool has_e) -> double {
                constexpr int par = decltype(PAR)::value;   // (kk - m) & 1: the Z read-ahead register
                zd t1, v2;
                double t2, beta;
                zlarfg2c(x1, x2, t1, v2, t2, beta);
                const int cbk = lay.cb(kk);
                // bulk: left on columns kk+2 <= p <= I2, right on rows I1 <= p <= kk-1
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int p = tp + s * LPM;
                    const bool rowop = (p >= kk + 2 && p <= I2);
                    const bool colop = (p >= I1 && p <= kk - 1) || zact;
                    if (rowop || colop) {
                        // -1 from the left (conj(tau), v2), +1 from the right (tau, conj(v2))
                        const double sg = rowop ? -1.0 : 1.0;
                        zd x = carry[s], y = ypre[s];
                        pair_op_s(x, y, t1.re, sg * t1.im, t2, v2.re, -(sg * v2.im));
                        if (rowop) {
                            H[cbp[s] + kk] = x;
                            H[cbp[s] + kk + 1] = y;
                        } else if (ZH && zlane) {
                            Z[zlay.cb(kk) + zrow] = x;
                        } else {
                            H[cbk + p] = x;
                        }
                        carry[s] = y;
                    }
                    if constexpr (WANTT && !ZH) {
                        if (wantz && p < n) {
                            zd &zy = par ? zyO[s] : zyE[s];
                            zd x = zc[s], y = zy;
                            pair_op_s(x, y, t1.re, t1.im, t2, v2.re, -v2.im);
                            Z[zlay(p, kk)] = x;
                            zc[s] = y;
                            if (kk + 3 <= i) zy = Z[zlay(p, kk + 3)];   // for step kk + 2
                        }
                    }
                }
                // window: rows/cols kk, kk+1 and row kk+2
                if (!first && t == 0) H[lay(kk, kk - 1)] = mk(beta, 0.0);
                pair_op_s(a, c, t1.re, -t1.im, t2, v2.re, v2.im);   // left: column kk
                pair_op_s(b, d, t1.re, -t1.im, t2, v2.re, v2.im);   // left: column kk+1
                pair_op_s(a, b, t1.re, t1.im, t2, v2.re, -v2.im);   // right: row kk
                pair_op_s(c, d, t1.re, t1.im, t2, v2.re, -v2.im);   // right: row kk+1
                double f = 0.0;
                if (has_e) {                  // right: row kk+2, H(kk+2,kk) = 0 before, e real
                    const double sr = t2 * e.re;
                    f = -sr;
                    e.re = fma(-sr, v2.re, e.re);
                    e.im = sr * v2.im;
                }
                if (t == 0) H[cbk + kk] = a;
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (tp + s * LPM == kk) carry[s] = b;
                lds_order();
                return f;
            };
            // move the state from step kk to step k1 = kk+1 (< i)
            auto advance = [&](const int kk, const double f) {
                const int k1 = kk + 1;
                const int cb1 = lay.cb(k1 + 1);
                x1 = c;
                x2 = f;
                a = d;
                c = e;
                b = H[cb1 + k1];
                d = H[cb1 + k1 + 1];
                e = H[cb1 + (k1 + 2 <= i ? k1 + 2 : k1 + 1)];   // unused at the last step
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int p = tp + s * LPM;
                    const bool rowop = (p >= k1 + 2 && p <= I2);
                    const int pc = p < k1 ? p : k1;
                    if constexpr (ZH) {
                        const zd *src = zlane ? Z + zlay.cb(k1 + 1) + (zrow < n ? zrow : n - 1)
                                              : H + (rowop ? cbp[s] + k1 + 1 : cb1 + pc);
                        ypre[s] = *src;
                    } else {
                        ypre[s] = H[rowop ? cbp[s] + k1 + 1 : cb1 + pc];
                    }

                }
            };

            using P0 = std::integral_constant<int, 0>;
            using P1 = std::integral_constant<int, 1>;
            double f = step(P0(), m, true, m + 2 <= i);
            if (m > l) {
                // ZLAHQR keeps H(m, m-1) real after a start at m > l:
                // flush the step's state, scale in LDS, reload.
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int p = tp + s * LPM;
                    if (p >= I1 && p <= m - 1) H[lay(p, m + 1)] = carry[s];
                    if constexpr (ZH) {
                        if (zact) Z[zlay(zrow, m + 1)] = carry[s];
                    }
                    if constexpr (WANTT && !ZH) {
                        if (wantz && p < n) Z[zlay(p, m + 1)] = zc[s];
                    }
                }
                if (t == 0) {
                    H[lay(m, m + 1)] = b;
                    H[lay(m + 1, m)] = c;
                    H[lay(m + 1, m + 1)] = d;
                    if (m + 2 <= i) H[lay(m + 2, m + 1)] = e;
                }
                lds_order();
                zd t1, v2;
                double t2r, beta;
                zlarfg2c(x1, x2, t1, v2, t2r, beta);   // the reflector of step m again (as the step)
                zd temp = zsub(mk(1.0, 0.0), t1);
                temp = zscale(1.0 / zabs(temp), temp);
                const zd tempc = zconj(temp);
                if (t == 0) {
                    H[lay(m + 1, m)] = zmul(H[lay(m + 1, m)], tempc);
                    if (m + 2 <= i) H[lay(m + 2, m + 1)] = zmul(H[lay(m + 2, m + 1)], temp);
                }
                lds_order();
                for (int j = m + t; j <= i; j += LPM)
                    if (j != m + 1)
                        for (int cc = j + 1; cc <= I2; ++cc) H[lay(j, cc)] = zmul(H[lay(j, cc)], temp);
                lds_order();
                for (int j = m + t; j <= i; j += LPM)
                    if (j != m + 1)
                        for (int r = I1; r <= j - 1; ++r) H[lay(r, j)] = zmul(H[lay(r, j)], tempc);
                if constexpr (WANTT) {
                    if (wantz)
                        for (int j = m; j <= i; ++j)
                            if (j != m + 1)
                                for (int r = t; r < n; r += LPM) Z[zlay(r, j)] = zmul(Z[zlay(r, j)], tempc);
                }
                lds_order();
                const int k1 = m + 1;
                if (k1 < i) {
                    x1 = H[lay(k1, m)];
                    x2 = f;
                    a = H[lay(k1, k1)];
                    b = H[lay(k1, k1 + 1)];
                    c = H[lay(k1 + 1, k1)];
                    d = H[lay(k1 + 1, k1 + 1)];
                    e = (k1 + 2 <= i) ? H[lay(k1 + 2, k1 + 1)] : mk(0.0, 0.0);
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int p = tp + s * LPM;
                        if (p >= k1 + 2 && p <= I2) { carry[s] = H[lay(k1, p)]; ypre[s] = H[lay(k1 + 1, p)]; }
                        else if (p >= I1 && p <= k1 - 1) { carry[s] = H[lay(p, k1)]; ypre[s] = H[lay(p, k1 + 1)]; }
                        if constexpr (ZH) {
                            if (zact) { carry[s] = Z[zlay(zrow, k1)]; ypre[s] = Z[zlay(zrow, k1 + 1)]; }
                        }
                        if constexpr (WANTT && !ZH) {   // k1 - m = 1: step k1 reads zyO
                            if (wantz && p < n) {
                                zc[s] = Z[zlay(p, k1)];
                                zyO[s] = Z[zlay(p, k1 + 1)];
                                if (k1 + 2 <= i) zyE[s] = Z[zlay(p, k1 + 2)];
                            }
                        }
                    }
                } else {
                    c = H[lay(i, i - 1)];
                    d = H[lay(i, i)];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int p = tp + s * LPM;
                        if (p >= I1 && p <= i - 1) carry[s] = H[lay(p, i)];
                        else if (p >= i + 1 && p <= I2) carry[s] = H[lay(i, p)];
                        if constexpr (ZH) {
                            if (zact) carry[s] = Z[zlay(zrow, i)];
                        }
                        if constexpr (WANTT && !ZH) {
                            if (wantz && p < n) zc[s] = Z[zlay(p, i)];
                        }
                    }
                }
            } else if (m + 1 < i) {
                advance(m, f);
            }
            // main chase: every step has a row kk+2 and a successor
            if constexpr (WANTT && !ZH) {
                int kk = m + 1;
                for (; kk + 1 < i - 1; kk += 2) {
                    f = step(P1(), kk, false, true);
                    advance(kk, f);
                    f = step(P0(), kk + 1, false, true);
                    advance(kk + 1, f);
                }
                if (kk < i - 1) {
                    f = step(P1(), kk, false, true);
                    advance(kk, f);
                }
                if (i - 1 > m) {
                    if ((i - 1 - m) & 1) f = step(P1(), i - 1, false, false);
                    else f = step(P0(), i - 1, false, false);
                }
            } else {
                for (int kk = m + 1; kk < i - 1; ++kk) {
                    f = step(P0(), kk, false, true);
                    advance(kk, f);
                }
                if (i - 1 > m) f = step(P0(), i - 1, false, false);
            }
            // ---- end of sweep: c = H(i,i-1), d = H(i,i); make H(i,i-1) real
            zd hsub = c;
            bool rot = (c.im != 0.0);
            zd temp = mk(1.0, 0.0);
            if (rot) {
                double r = zabs(c);
                hsub = mk(r, 0.0);
                temp = zscale(1.0 / r, c);
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int p = tp + s * LPM;
                if (p >= I1 && p <= i - 1) H[lay(p, i)] = rot ? zmul(carry[s], temp) : carry[s];
                if constexpr (WANTT) {
                    if (p >= i + 1 && p <= I2) H[lay(i, p)] = rot ? zmul(carry[s], zconj(temp)) : carry[s];
                }
                if constexpr (ZH) {
                    if (zact) Z[zlay(zrow, i)] = rot ? zmul(carry[s], temp) : carry[s];
                }
                if constexpr (WANTT && !ZH) {
                    if (wantz && p < n) Z[zlay(p, i)] = rot ? zmul(zc[s], temp) : zc[s];
                }
            }
            if (t == 0) {
                H[lay(i, i - 1)] = hsub;
                H[lay(i, i)] = d;
            }
            lds_order();
        }
        if (!conv) return i + 1;
        kdefl = 0;
        i = l - 1;
    }
    return 0;
}

// View of the trailing part of a layout from (o, o): element (r, c) of the
// view is element (o + r, o + c) of the base.
template <class Lay>
struct SubLayout {
    Lay base;
    int o;
    __device__ __forceinline__ int cb(int c) const { return base.cb(c + o) + o; }
    __device__ __forceinline__ int operator()(int r, int c) const { return cb(c) + r; }
};

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    union { double d; int w[2]; } u;
    u.d = v;
    u.w[0] = __builtin_amdgcn_readlane(u.w[0], lane);
    u.w[1] = __builtin_amdgcn_readlane(u.w[1], lane);
    return u.d;
}

// Broadcast of the two half-waves: lo receives lanes 0..31's value, hi lanes
// 32..63's, in every lane (v_permlane32_swap with both operands = v; the
// halves are uniform where this is used).
__device__ __forceinline__ void bcast_halves(double v, double &lo, double &hi)
{
    union { double d; unsigned w[2]; } u, a, b;
    u.d = v;
    const auto r0 = __builtin_amdgcn_permlane32_swap(u.w[0], u.w[0], false, false);
    const auto r1 = __builtin_amdgcn_permlane32_swap(u.w[1], u.w[1], false, false);
    a.w[0] = r0[0];
    a.w[1] = r1[0];
    b.w[0] = r0[1];
    b.w[1] = r1[1];
    lo = a.d;
    hi = b.d;
}

// The same within each 32-lane group of a wavefront (v_permlane16_swap): lo
// receives the group's lanes 0..15 value, hi its lanes 16..31 value, in every
// lane of the group (the 16-lane halves are uniform where this is used).
__device__ __forceinline__ void bcast_halves16(double v, double &lo, double &hi)
{
    union { double d; unsigned w[2]; } u, a, b;
    u.d = v;
    const auto r0 = __builtin_amdgcn_permlane16_swap(u.w[0], u.w[0], false, false);
    const auto r1 = __builtin_amdgcn_permlane16_swap(u.w[1], u.w[1], false, false);
    a.w[0] = r0[0];
    a.w[1] = r1[0];
    b.w[0] = r0[1];
    b.w[1] = r1[1];
    lo = a.d;
    hi = b.d;
}

// Halves of a lane group of LPM (32 or 64) lanes, broadcast to the group.
template <int LPM>
__device__ __forceinline__ void group_halves(double v, double &lo, double &hi)
{
    if constexpr (LPM == 64) bcast_halves(v, lo, hi);
    else bcast_halves16(v, lo, hi);
}

// ---- QR with two bulges per sweep (LPM = 64: one matrix per wavefront;
// LPM = 32: two, n <= 32) ----
// A "dual sweep" over an active block [l, i] is two consecutive single-shift
// ZLAHQR sweeps with the shifts s1 (the Wilkinson shift) and s2 (the other
// eigenvalue of the trailing 2x2), pipelined: bulge A (s1) starts at step 0,
// bulge B (s2) three positions behind it.  The first half of the group's
// lanes computes A's reflector chain and window, the second half B's, in the
// same instructions; every lane p then applies both reflectors at its
// position (B's first, then A's).
//  * The two bulges act on disjoint rows/columns except the 2x2 block rows
//    {kB, kB+1} x columns {kA, kA+1}, which B changes from the left and A
//    from the right; those products commute, and the phases run in order
//    with every bulk update reading and writing LDS, so the result is that of
//    the two sweeps run one after the other (implicit-Q argument), up to
//    rounding.
//  * Windows: A's window (rows/cols kA..kA+2) is untouched by B, B's by A;
//    B reads H(kA, kA-1) as its next window element after A wrote it.
//  * Each bulge ends with ZLAHQR's "make H(i,i-1) real" scaling of column i.
//  * WANTT: the Schur form T, ZLAHQR's WANTT ranges (left updates to column
//    n-1, right updates from row 0); the extra elements of the two bulges'
//    updates stay disjoint but for the same 2x2 block.  With Z (global
//    memory, column-major with leading dimension zld, lane p owns row p)
//    every reflector is also applied from the right to Z.  Per lane, A reads
//    its Z columns from memory two steps ahead, hands each column it
//    finishes to B through registers (B needs column kA-1 one step after A
//    finished it), and B stores the columns it finishes: one load and one
//    store per lane and step.
// Blocks smaller than DUAL_MIN and the last rows go to the single-shift code
// (zlahqr_core) on the sub-block, which is independent once H(l, l-1) = 0
// (with WANTT on the whole of T and Z, limited to the block's eigenvalues).
// Returns 0 or 1 + the 0-based index of the first unconverged eigenvalue.
//  * Packed layout only: column c starts c + 2 elements after column c - 1,
//    so every address of a step is an increment of the previous step's.
//  * DENSE (one matrix per wavefront): every lane loads, updates and stores
//    its bulk operands with no branch, idle lanes at H[packed_size(n)], one
//    element past the matrix that the caller allocates for them (the QR 5.9%
//    faster at n = 64, 2.9% at n = 56, 2.2% at n = 48, bit-identical).
//  * Hook (eigenvalues only): after every deflation, hook->moved(H, idle, i,
//    t) may move the leading packed_size(i + 1) elements (columns 0..i, all
//    that the rest of the QR reads) elsewhere and repoint H and the idle
//    lanes' element; rows > i are final at that point.
//  * SNAP: both shifts of the first kSnapSweeps sweeps after a deflation are
//    snapped to the nearest known eigenvalue (snap_shift; lane t holds known
//    eigenvalue t).
struct NoMove {
    __device__ __forceinline__ void moved(zd *&, int &, int, int) {}
};
template <int LPM = 64, bool WANTT = false, bool DENSE = false, class Hook = NoMove, bool SNAP = false>
__device__ __forceinline__ int zlahqr_dual(zd *H, const PackedLayout lay, int n, int t, int itmax_override,
                                           zd *Z = nullptr, int zld = 0, Hook *hook = nullptr,
                                           int idle0 = -1, zd kev = zd{}, bool kh = false)
{
    static_assert(LPM == 32 || LPM == 64, "two-bulge QR: 32 or 64 lanes per matrix");
    constexpr int HALF = LPM / 2;
#ifndef TKZ_DUAL_MIN
#define TKZ_DUAL_MIN 10
#endif
    constexpr int DUAL_MIN = TKZ_DUAL_MIN;
    static_assert(DUAL_MIN >= 8, "the pipelined sweep needs at least 8 rows");
    const double ulp = DBL_EPSILON, safmin = DBL_MIN;
    const double smlnum = safmin * ((double)n / ulp);
    const int itmax = itmax_override >= 0 ? itmax_override : 30 * (n > 10 ? n : 10);
    const int half = t >> (LPM == 64 ? 5 : 4);   // 0: bulge A's chain, 1: bulge B's
    const bool leader = (t & (HALF - 1)) == 0;   // the first lane of each half stores its window
    const int p = t;                   // bulk position of this lane
    const int cbp = lay.cb(p < n ? p : n - 1);
    int idle_slot = idle0 >= 0 ? idle0 : packed_size(n);   // DENSE: the idle lanes' element
    const bool zp = WANTT && Z != nullptr && p < n;   // this lane updates row p of Z
    zd *const zr = zp ? Z + p : Z;     // element (p, c) of Z at zr[c * zld]
    int kdefl = 0;
    int i = n - 1;
    while (i >= 1) {
        int l = 0;
        bool conv = false;
        for (int its = 0; its <= itmax; ++its) {
            const int kf = group_last_true<LPM, 1>(t, l + 1, i, l, [&](int q) {
                return subdiag_negligible(H, lay, q, 0, n - 1, smlnum, ulp);
            });
            l = kf;
            if (l > 0 && t == 0) H[lay(l, l - 1)] = mk(0.0, 0.0);
            lds_order();
            if (l >= i) { conv = true; break; }
            if (i - l + 1 < DUAL_MIN) {
                // the whole block [l, i] by single-shift sweeps
                if constexpr (WANTT) {
                    int info;
                    [[clang::always_inline]] info =
                        zlahqr_core<LPM, 1, true, false, PackedLayout, FullLayout, true, SNAP>(
                            H, lay, n, t, itmax_override, Z, FullLayout{zld}, i, l, kev, kh);
                    if (info != 0) return info;
                } else {
                    // inlined: as a call, its ABI would raise the kernel's VGPRs
                    // from 123 to 216 and add stack
                    int info;
                    [[clang::always_inline]] info = zlahqr_core<LPM, 1, false>(
                        H, SubLayout<PackedLayout>{lay, l}, i - l + 1, t, itmax_override);
                    if (info != 0) return l + info;
                }
                conv = true;
                break;
            }
            ++kdefl;
            // ---- shifts: s1 as ZLAHQR (incl. exceptional), s2 the other
            // eigenvalue of the trailing 2x2 (trace - Wilkinson shift), or the
            // Wilkinson shift when s1 is exceptional
            zd s1 = zlahqr_shift(H, lay, l, i, kdefl);
            zd s2 = (kdefl % 10 == 0) ? zlahqr_shift(H, lay, l, i, 1)
                                       : zsub(zadd(H[lay(i - 1, i - 1)], H[lay(i, i)]), s1);
            if constexpr (SNAP) {   // the two eigenvalues of the trailing 2x2, each snapped
                if (kdefl <= kSnapSweeps) {
                    s1 = snap_shift<LPM>(s1, kev, kh, t);
                    s2 = snap_shift<LPM>(s2, kev, kh, t);
                }
            }
            const zd sig = half ? s2 : s1;
            const int nsteps = i - l;          // steps of one bulge: k = l .. i-1 (>= DUAL_MIN - 1)
            const bool pin = p >= l && p <= i;
            // per-half chain state: window, reflector input, and the bulge's
            // position kk with pw = &H(kk, kk).  A half without a bulge keeps
            // x1 = 1 (inside the fast reflector's range); its results are
            // never stored.
            zd x1 = mk(1.0, 0.0), a = mk(0.0, 0.0), b = a, c = a, d = a, e = a;
            double x2 = 0.0;
            int kk = l;
            zd *pw = H + lay.cb(l) + l;
            // Z state of this lane's row (WANTT with Z): A's carried column
            // kA; the column A's step tau needs next, read two steps ahead
            // into zyE (even tau) or zyO (odd tau) so that no register move
            // waits for a load in flight; the two columns A finished last
            // (zA0 newest) and the first one; B's carried and next column
            zd zcA = mk(0.0, 0.0), zyE = zcA, zyO = zcA, zA0 = zcA, zA1 = zcA, zF = zcA, zcB = zcA, zyB = zcA;
            if constexpr (WANTT) {
                if (zp) {
                    zcA = zr[l * zld];
                    zyE = zr[(l + 1) * zld];
                    zyO = zr[(l + 2) * zld];
                }
            }
            auto start = [&]() {
                kk = l;
                pw = H + lay.cb(l) + l;
                a = pw[0];
                c = pw[1];
                b = pw[l + 2];                // column l + 1 starts at cb(l) + l + 2
                d = pw[l + 3];
                e = pw[l + 4];                // l + 2 <= i: the block has >= DUAL_MIN rows
                const zd h11s = zsub(a, sig);
                const double h21 = c.re;
                const double sc = cabs1(h11s) + fabs(h21);
                x1 = mk(h11s.re / sc, h11s.im / sc);
                x2 = h21 / sc;
            };
            // bulk operands of the bulge at k: from the left on rows k, k+1 of
            // columns p >= k+2 (elements cb(p) + k, + 1), from the right on
            // columns k, k+1 of rows l <= p <= k-1 (cb(k) + p, cb(k+1) + p).
            // Bulge A's k advances by one per step, so both candidates are
            // running pointers: the row form advances by one element, the
            // column form by k + 2 (cb(k+1) = cb(k) + k + 2), and the second
            // operand of each form is the first one of the next step.
            zd *pnr = H + cbp + l, *pnc = H + lay.cb(l) + p;
            zd *const pidle = H + idle_slot;   // DENSE: the idle lanes' element
            // LPM = 64: t is the lane id, and a step's lane sets are bit masks
            // formed on the scalar unit
            const int wb = i - l + 1;
            const uint64_t pinm = (wb >= 64 ? ~0ull : (1ull << wb) - 1ull) << l;   // lanes l..i
            const uint64_t nm = n >= 64 ? ~0ull : (1ull << n) - 1ull;              // lanes 0..n-1
            auto phase_addr = [&](const int k, bool &row, bool &col, zd *&ax, zd *&ay) {
                if constexpr (LPM == 64) {    // k <= i - 1 <= 62
                    const uint64_t ge = k + 2 < 64 ? ~0ull << (k + 2) : 0ull;   // lanes >= k+2
                    const uint64_t le = (1ull << k) - 1ull;                     // lanes <= k-1
                    if constexpr (WANTT) {    // columns k+2..n-1, rows 0..k-1
                        row = __builtin_amdgcn_inverse_ballot_w64(ge & nm);
                        col = __builtin_amdgcn_inverse_ballot_w64(le);
                    } else {
                        row = __builtin_amdgcn_inverse_ballot_w64(ge & pinm);
                        col = __builtin_amdgcn_inverse_ballot_w64(le & pinm);
                    }
                } else if constexpr (WANTT) {
                    row = p >= k + 2 && p < n;
                    col = p <= k - 1;
                } else {
                    row = pin && p >= k + 2;
                    col = pin && p <= k - 1;
                }
                zd *const nr = pnr + 1, *const nc = pnc + (k + 2);
                ax = row ? pnr : pnc;
                ay = row ? nr : nc;
                pnr = nr;
                pnc = nc;
                if constexpr (DENSE) {
                    if (!(row || col)) ax = ay = pidle;
                }
            };
            // sg = -1 from the left (c1 = conj(tau), v = v2), +1 from the
            // right (c1 = tau, v = conj(v2)): sign changes are exact, so the
            // roundings are those of pair_op with the conjugated values
            auto phase_apply = [&](const bool row, const bool col, zd *const ax, zd *const ay, zd x, zd y,
                                   const zd tr, const zd vr, const double t2r, const double sg) {
                if (DENSE || row || col) {
                    pair_op_s(x, y, tr.re, sg * tr.im, t2r, vr.re, -(sg * vr.im));
                    *ax = x;
                    *ay = y;
                }
            };
            // bulk operand addresses and sign of bulge A's last three steps:
            // bulge B, three positions behind in the same sweep (H does not
            // move within one), takes them from slot 0
            bool rrow[3] = {false, false, false}, rcol[3] = {false, false, false};
            zd *rax[3] = {pidle, pidle, pidle}, *ray[3] = {pidle, pidle, pidle};
            double rsg[3] = {1.0, 1.0, 1.0};
            // one time step: bulge A at kA = l + tau, bulge B at kA - 3.
            // HA/HB: bulge A/B active; ADV bit 0: A advances, bit 1: B
            // advances (compile-time: the steady state has no per-half branch)
            // RS: the ring slot of this step (tau mod 3, steady state
            // only), or -1: B takes slot 0 and the ring shifts by one
            auto dstep = [&](auto HA, auto HB, auto ADV, auto PAR, auto RS, const int tau) {
                constexpr bool hasA = decltype(HA)::value, hasB = decltype(HB)::value;
                constexpr int adv = decltype(ADV)::value;
                constexpr int par = decltype(PAR)::value;   // tau & 1 (Z read-ahead register)
                constexpr int rs = decltype(RS)::value;
                constexpr int rb = rs < 0 ? 0 : rs, ra = rs < 0 ? 2 : rs;
                const int kA = l + tau, kB = kA - 3;
                // phase-1 operands (bulge B): independent of this step's chain
                bool rowB = false, colB = false;
                zd *axB = pidle, *ayB = pidle;
                double sgB = 1.0;
                zd xB, yB;                    // used only where rowB || colB
                if constexpr (hasB) {
                    rowB = rrow[rb];          // phase_addr(kB, ...) of step tau - 3
                    colB = rcol[rb];
                    axB = rax[rb];
                    ayB = ray[rb];
                    sgB = rsg[rb];
                    if (DENSE || rowB || colB) {
                        xB = *axB;
                        yB = *ayB;
                    }
                }
                // chain: reflector and window of this half's bulge; after a
                // bulge's last step e = 0 and the row kk+2 update is a no-op
                zd t1, v2;
                double t2, beta;
                zlarfg2c(x1, x2, t1, v2, t2, beta);
                const zd xsub = mk(beta, 0.0);
                pair_op_s(a, c, t1.re, -t1.im, t2, v2.re, v2.im);   // left: column kk
                pair_op_s(b, d, t1.re, -t1.im, t2, v2.re, v2.im);   // left: column kk+1
                pair_op_s(a, b, t1.re, t1.im, t2, v2.re, -v2.im);   // right: row kk
                pair_op_s(c, d, t1.re, t1.im, t2, v2.re, -v2.im);   // right: row kk+1
                const double sr = t2 * e.re;  // right: row kk+2 (H(kk+2,kk) = 0, e real)
                const double f = -sr;
                e.re = fma(-sr, v2.re, e.re);
                e.im = sr * v2.im;
                // window stores: H(kk,kk-1) at pw - kk - 1, H(kk,kk) at pw,
                // H(kk,kk+1) at pw + kk + 2
                if (leader && (hasA && hasB ? true : (hasA ? half == 0 : half == 1))) {
                    if (kk > l) pw[-(kk + 1)] = xsub;
                    pw[0] = a;
                    pw[kk + 2] = b;
                }
                // both reflectors to every lane of the group (one matrix per
                // wavefront: wave-uniform values read from lanes 0 and 32)
                zd tA, vA, tB, vB;
                double t2A, t2B;
                if constexpr (LPM == 64) {
                    tA = mk(readlane_d(t1.re, 0), readlane_d(t1.im, 0));
                    vA = mk(readlane_d(v2.re, 0), readlane_d(v2.im, 0));
                    t2A = readlane_d(t2, 0);
                    tB = mk(readlane_d(t1.re, 32), readlane_d(t1.im, 32));
                    vB = mk(readlane_d(v2.re, 32), readlane_d(v2.im, 32));
                    t2B = readlane_d(t2, 32);
                } else {
                    group_halves<LPM>(t1.re, tA.re, tB.re);
                    group_halves<LPM>(t1.im, tA.im, tB.im);
                    group_halves<LPM>(v2.re, vA.re, vB.re);
                    group_halves<LPM>(v2.im, vA.im, vB.im);
                    group_halves<LPM>(t2, t2A, t2B);
                }
                if constexpr (hasB) phase_apply(rowB, colB, axB, ayB, xB, yB, tB, vB, t2B, sgB);
                if constexpr (WANTT && hasB) {   // Z(p, kB), Z(p, kB+1): column kB is final
                    if (zp) {
                        zd x = zcB, y = zyB;
                        pair_op_s(x, y, tB.re, tB.im, t2B, vB.re, -vB.im);
                        zr[kB * zld] = x;
                        zcB = y;
                    }
                }
                lds_order();
                if constexpr (rs < 0) {
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        rrow[s] = rrow[s + 1];
                        rcol[s] = rcol[s + 1];
                        rax[s] = rax[s + 1];
                        ray[s] = ray[s + 1];
                        rsg[s] = rsg[s + 1];
                    }
                }
                if constexpr (hasA) {         // phase 2 reads after phase 1's writes
                    bool rowA, colA;
                    zd *axA, *ayA;
                    phase_addr(kA, rowA, colA, axA, ayA);
                    const double sgA = rowA ? -1.0 : 1.0;
                    rrow[ra] = rowA;
                    rcol[ra] = colA;
                    rax[ra] = axA;
                    ray[ra] = ayA;
                    rsg[ra] = sgA;
                    zd xA, yA;                // used only where rowA || colA
                    if (DENSE || rowA || colA) {
                        xA = *axA;
                        yA = *ayA;
                    }
                    phase_apply(rowA, colA, axA, ayA, xA, yA, tA, vA, t2A, sgA);
                    if constexpr (WANTT) {       // Z(p, kA), Z(p, kA+1): column kA goes to B
                        zd &zy = par ? zyO : zyE;
                        zd x = zcA, y = zy;
                        pair_op_s(x, y, tA.re, tA.im, t2A, vA.re, -vA.im);
                        if constexpr (!hasB) {
                            if (tau == 0) zF = x;
                        }
                        zA1 = zA0;
                        zA0 = x;
                        zcA = y;
                        if (zp && kA + 3 <= i) zy = zr[(kA + 3) * zld];   // for step tau + 2
                    }
                }
                lds_order();
                // advance to kk+1: b, d, e are rows kk+1 .. kk+3 of column kk+2,
                // which starts at cb(kk+2) = cb(kk) + 2kk + 5; H(kk+1, kk+1) is
                // at cb(kk+1) + kk + 1 = pw + kk + 3
                if constexpr (adv != 0) {
                    if (adv == 3 || (adv == 1 ? half == 0 : half == 1)) {
                        const zd *const base = pw + (2 * kk + 6);
                        x1 = c;
                        x2 = f;
                        a = d;
                        c = e;
                        b = base[0];
                        d = base[1];
                        e = (kk + 3 <= i) ? base[2] : mk(0.0, 0.0);
                        pw += kk + 3;
                        kk += 1;
                    }
                }
                if constexpr (WANTT) {
                    // B's next column kB+2 = kA-1: the column A finished last step
                    if constexpr ((adv & 2) != 0) zyB = zA1;
                    if constexpr (!hasA) zA1 = zA0;   // after A's end: column i next
                }
            };
            // end of a bulge's sweep (ZLAHQR): H(i,i-1) made real, column i
            // rows l..i-1 scaled; c, d of half `hb` hold H(i,i-1), H(i,i)
            auto finish = [&](const int hb) {
                double cre, cim, dre, dim, o;
                if (hb == 0) {
                    group_halves<LPM>(c.re, cre, o);
                    group_halves<LPM>(c.im, cim, o);
                    group_halves<LPM>(d.re, dre, o);
                    group_halves<LPM>(d.im, dim, o);
                } else {
                    group_halves<LPM>(c.re, o, cre);
                    group_halves<LPM>(c.im, o, cim);
                    group_halves<LPM>(d.re, o, dre);
                    group_halves<LPM>(d.im, o, dim);
                }
                const zd ce = mk(cre, cim), de = mk(dre, dim);
                const bool rot = ce.im != 0.0;
                zd temp = mk(1.0, 0.0);
                if (rot) {
                    const double r = zabs(ce);
                    temp = zscale(1.0 / r, ce);
                    if (p >= (WANTT ? 0 : l) && p <= i - 1) H[lay(p, i)] = zmul(H[lay(p, i)], temp);
                    if constexpr (WANTT) {
                        if (p >= i + 1 && p < n) H[lay(i, p)] = zmul(H[lay(i, p)], zconj(temp));
                    }
                    if (t == 0) H[lay(i, i - 1)] = mk(r, 0.0);
                } else if (t == 0) {
                    H[lay(i, i - 1)] = ce;
                }
                if constexpr (WANTT) {
                    // column i of Z: A's goes to B as the column after i-1,
                    // B's is final
                    if (hb == 0) {
                        zA1 = zA0;
                        zA0 = rot ? zmul(zcA, temp) : zcA;
                    } else if (zp) {
                        zr[i * zld] = rot ? zmul(zcB, temp) : zcB;
                    }
                }
                if (t == 0) H[lay(i, i)] = de;
                lds_order();
            };
            using T_ = std::true_type;
            using F_ = std::false_type;
            using P0 = std::integral_constant<int, 0>;
            using P1 = std::integral_constant<int, 1>;
            using A1 = std::integral_constant<int, 1>;
            using A3 = std::integral_constant<int, 3>;
            using A2 = std::integral_constant<int, 2>;
            // bulge A alone for three steps, then B starts three positions behind
            using RSH = std::integral_constant<int, -1>;
            if (half == 0) start();
            if constexpr (WANTT) {            // the Z read-ahead register follows tau's parity
                dstep(T_(), F_(), A1(), P0(), RSH(), 0);
                dstep(T_(), F_(), A1(), P1(), RSH(), 1);
                dstep(T_(), F_(), A1(), P0(), RSH(), 2);
            } else {
                for (int tau = 0; tau < 3; ++tau) dstep(T_(), F_(), A1(), P0(), RSH(), tau);
            }
            if (half == 1) start();
            if constexpr (WANTT) {            // B's columns l, l+1: A's first two
                zcB = zF;
                zyB = zA1;
                int tau = 3;
                for (; tau + 1 < nsteps - 1; tau += 2) {
                    dstep(T_(), T_(), A3(), P1(), RSH(), tau);
                    dstep(T_(), T_(), A3(), P0(), RSH(), tau + 1);
                }
                if (tau < nsteps - 1) dstep(T_(), T_(), A3(), P1(), RSH(), tau);
                if ((nsteps - 1) & 1) dstep(T_(), T_(), A2(), P1(), RSH(), nsteps - 1);   // A's last step
                else dstep(T_(), T_(), A2(), P0(), RSH(), nsteps - 1);
            } else {
                // steady state by threes, ring slots fixed at compile time (no
                // register moves); after each three the ring is in shift order
                // again (slot 0 the oldest)
                using RS0 = std::integral_constant<int, 0>;
                using RS1 = std::integral_constant<int, 1>;
                using RS2 = std::integral_constant<int, 2>;
                int tau = 3;
                for (; tau + 3 < nsteps; tau += 3) {
                    dstep(T_(), T_(), A3(), P0(), RS0(), tau);
                    dstep(T_(), T_(), A3(), P0(), RS1(), tau + 1);
                    dstep(T_(), T_(), A3(), P0(), RS2(), tau + 2);
                }
                for (; tau < nsteps - 1; ++tau) dstep(T_(), T_(), A3(), P0(), RSH(), tau);
                dstep(T_(), T_(), A2(), P0(), RSH(), nsteps - 1);   // A's last step
            }
            finish(0);
            dstep(F_(), T_(), A2(), P0(), RSH(), nsteps);
            dstep(F_(), T_(), A2(), P0(), RSH(), nsteps + 1);
            dstep(F_(), T_(), std::integral_constant<int, 0>(), P0(), RSH(), nsteps + 2);   // B's last step
            finish(1);
        }
        if (!conv) return i + 1;
        kdefl = 0;
        i = l - 1;
        if constexpr (!WANTT && !std::is_same<Hook, NoMove>::value) {
            if (i >= 1) hook->moved(H, idle_slot, i, t);
        }
    }
    return 0;
}

template <int LPM, int S, class Lay>
__device__ __forceinline__ int zlahqr_eigs(zd *H, const Lay lay, int n, int t, int itmax_override)
{
    return zlahqr_core<LPM, S, false>(H, lay, n, t, itmax_override);
}

}  // namespace tkz
