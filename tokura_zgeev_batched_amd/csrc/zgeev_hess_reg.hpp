// Householder Hessenberg reduction (LAPACK ZGEHD2 / ZLARFG arithmetic, as
// hessenberg_wg) by one 256-thread workgroup with the matrix in registers.
//
// Reference: the unblocked reduction tokura_zgehrd_MWB_kernel.cuh:159-287,
// which keeps the matrix in shared memory; here only the reflector vector and
// the partial row sums of the right update pass through LDS.
//
// Distribution: the 256 threads form a 16 x 16 grid.  Thread (tr, tc) holds
// A(tr + 16a, tc + 16b) for a, b < NBLK (n <= 16 NBLK), in R[a][b].  Lane
// tr + 16 (tc % 4) of wavefront tc / 4: the 16 lanes of a DPP row hold 16
// consecutive rows of the same columns (coalesced column-major loads), the
// four DPP rows of a wavefront four columns, the four wavefronts the rest.
//
// Step i (reflector v from column i, v(i+1) = 1, v(r) = 0 for r <= i):
//  * the wavefront holding column i computes beta, tau and v after its own
//    update of that column in step i-1 and writes v and tau to LDS;
//  * right update A(:, i+1:) -= tau (A v) v^H: the row sums A v are reduced
//    over a thread's columns, over the wavefront's four DPP rows with
//    v_permlane16_swap / v_permlane32_swap (each lane keeps one row's sum),
//    then over the four wavefronts through LDS;
//  * left update A(i+1:, i+1:) -= conj(tau) v (v^H A): the column sums are
//    reduced over a thread's rows and the 16 lanes of its DPP row with DPP
//    moves, no LDS and no barrier.
// Two workgroup barriers per column (v ready, row sums ready); blocks of 16
// rows or columns that lie entirely in the finished part are skipped.  The
// arithmetic is that of hessenberg_wg; only the order of the sums differs.
#pragma once
#include "zgeev_device.hpp"

namespace tkz {
namespace hreg {

template <int CTRL>
__device__ __forceinline__ double dpp(double x)
{
    union { double d; int w[2]; } u;
    u.d = x;
    u.w[0] = __builtin_amdgcn_mov_dpp(u.w[0], CTRL, 0xF, 0xF, false);   // every lane has a source lane
    u.w[1] = __builtin_amdgcn_mov_dpp(u.w[1], CTRL, 0xF, 0xF, false);
    return u.d;
}

// Sum and max over the 16 lanes of a DPP row with the same bits in every lane
// (partners: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror; each step adds two values in either order).
__device__ __forceinline__ double row_sum(double x)
{
    x += dpp<0xB1>(x);
    x += dpp<0x4E>(x);
    x += dpp<0x141>(x);
    x += dpp<0x140>(x);
    return x;
}
__device__ __forceinline__ double row_max(double x)
{
    x = fmax(x, dpp<0xB1>(x));
    x = fmax(x, dpp<0x4E>(x));
    x = fmax(x, dpp<0x141>(x));
    x = fmax(x, dpp<0x140>(x));
    return x;
}

// v_permlane16_swap: DPP rows (x0, x1, x2, x3), (y0, y1, y2, y3) become
// (x0, y0, x2, y2), (x1, y1, x3, y3).
__device__ __forceinline__ void swap16(double &x, double &y)
{
    union { double d; unsigned w[2]; } ux, uy;
    ux.d = x;
    uy.d = y;
    const auto r0 = __builtin_amdgcn_permlane16_swap(ux.w[0], uy.w[0], false, false);
    const auto r1 = __builtin_amdgcn_permlane16_swap(ux.w[1], uy.w[1], false, false);
    ux.w[0] = r0[0];
    uy.w[0] = r0[1];
    ux.w[1] = r1[0];
    uy.w[1] = r1[1];
    x = ux.d;
    y = uy.d;
}
// v_permlane32_swap: (x0, x1, x2, x3), (y0, y1, y2, y3) become
// (x0, x1, y0, y1), (x2, x3, y2, y3).
__device__ __forceinline__ void swap32(double &x, double &y)
{
    union { double d; unsigned w[2]; } ux, uy;
    ux.d = x;
    uy.d = y;
    const auto r0 = __builtin_amdgcn_permlane32_swap(ux.w[0], uy.w[0], false, false);
    const auto r1 = __builtin_amdgcn_permlane32_swap(ux.w[1], uy.w[1], false, false);
    ux.w[0] = r0[0];
    uy.w[0] = r0[1];
    ux.w[1] = r1[0];
    uy.w[1] = r1[1];
    x = ux.d;
    y = uy.d;
}

// p_a (a = 0..3) in every lane; returns, in DPP row q, the sum of p_q over
// the four DPP rows (same bits in the 16 lanes of a column of rows).
__device__ __forceinline__ double rows_reduce_scatter(double p0, double p1, double p2, double p3)
{
    swap16(p0, p1);
    swap16(p2, p3);
    double s01 = p0 + p1, s23 = p2 + p3;
    swap32(s01, s23);
    return s01 + s23;
}
// u in DPP row q; g_a receives DPP row a's value, in every lane.
__device__ __forceinline__ void rows_all_gather(double u, double &g0, double &g1, double &g2, double &g3)
{
    double x = u, y = u;
    swap16(x, y);
    double x2 = x, y2 = y;
    swap32(x, x2);
    swap32(y, y2);
    g0 = x;
    g1 = y;
    g2 = x2;
    g3 = y2;
}

// Complex multiply-accumulate as four FMAs.
__device__ __forceinline__ zd fma_add(zd a, zd b, zd c)   // c + a b
{
    return mk(fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im)));
}
__device__ __forceinline__ zd fma_sub(zd a, zd b, zd c)   // c - a b
{
    return mk(fma(-a.re, b.re, fma(a.im, b.im, c.re)), fma(-a.re, b.im, fma(-a.im, b.re, c.im)));
}
__device__ __forceinline__ zd fma_subc(zd a, zd b, zd c)  // c - a conj(b)
{
    return mk(fma(-a.re, b.re, fma(-a.im, b.im, c.re)), fma(-a.im, b.re, fma(a.re, b.im, c.im)));
}
__device__ __forceinline__ zd fma_addc(zd a, zd b, zd c)  // c + conj(a) b
{
    return mk(fma(a.re, b.re, fma(a.im, b.im, c.re)), fma(a.re, b.im, fma(-a.im, b.re, c.im)));
}

}  // namespace hreg

struct HregSmem {
    zd xv[2][64];    // reflector vector of step i in xv[i & 1]
    zd up[4][64];    // per-wavefront row sums of the right update
    zd tau[2];
    double beta[64];  // H(j+1, j) of the reduced matrix
    double red[4];
};

// Reflector of column j by the wavefront holding it, column block B = j / 16
// (wave index wv uniform): tau (and v, unless tau = 0) to LDS buffer j & 1,
// H(j+1, j) to S.beta[j].  R's column j is left as it is: no later step reads
// it (the right updates multiply it by v(c) = 0, the left updates skip
// columns <= i), and only its rows <= j are part of the result.
template <int NBLK, int B>
__device__ __forceinline__ void hreg_reflect_b(const zd (&R)[NBLK][NBLK], HregSmem &S, int j, int q, int tr)
{
    const int qo = (j & 15) & 3;
    double ssq = 0.0, amx = 0.0;
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
        if (tr + 16 * a >= j + 2) {
            ssq += R[a][B].re * R[a][B].re + R[a][B].im * R[a][B].im;
            amx = fmax(amx, fmax(fabs(R[a][B].re), fabs(R[a][B].im)));
        }
    const int l0 = qo * 16;
    ssq = readlane_d(hreg::row_sum(ssq), l0);
    amx = readlane_d(hreg::row_max(amx), l0);
    const int aa = (j + 1) >> 4, la = l0 + ((j + 1) & 15);
    zd ca = mk(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
        if (a == aa) ca = R[a][B];
    const zd alpha = mk(readlane_d(ca.re, la), readlane_d(ca.im, la));
    double xnorm = 0.0;
    if (amx > 0.0) {
        if (ssq > 1e-290 && ssq < 1e290) {
            xnorm = sqrt(ssq);
        } else {   // rare: rescaled second pass
            const double ia = 1.0 / amx;
            double s2 = 0.0;
#pragma unroll
            for (int a = 0; a < NBLK; ++a)
                if (tr + 16 * a >= j + 2)
                    s2 += (R[a][B].re * ia) * (R[a][B].re * ia) + (R[a][B].im * ia) * (R[a][B].im * ia);
            xnorm = amx * sqrt(readlane_d(hreg::row_sum(s2), l0));
        }
    }
    const int buf = j & 1;
    double beta;
    zd tau, scal;
    if (xnorm == 0.0 && alpha.im == 0.0) {   // tau = 0, H(j) = I: the updates of step j add zeros
        beta = alpha.re;
        tau = mk(0.0, 0.0);
        scal = mk(0.0, 0.0);
    } else {
        const double wm = fmax(fmax(fabs(alpha.re), fabs(alpha.im)), xnorm);
        const double iw = 1.0 / wm;
        const double ar = alpha.re * iw, ai = alpha.im * iw, xn = xnorm * iw;
        beta = -copysign(wm * sqrt(ar * ar + ai * ai + xn * xn), alpha.re);
        tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
        scal = zdiv(mk(1.0, 0.0), mk(alpha.re - beta, alpha.im));
    }
    if (q == qo) {
#pragma unroll
        for (int a = 0; a < NBLK; ++a) {
            const int r = tr + 16 * a;
            S.xv[buf][r] = r == j + 1 ? mk(1.0, 0.0) : (r > j + 1 ? zmul(R[a][B], scal) : mk(0.0, 0.0));
        }
    }
    S.tau[buf] = tau;   // the same value from every lane
    S.beta[j] = beta;
}

template <int NBLK, int B = 0>
__device__ __forceinline__ void hreg_reflect(const zd (&R)[NBLK][NBLK], HregSmem &S, int j, int wv, int q, int tr)
{
    if constexpr (B < NBLK) {
        if ((j >> 4) == B) {
            if (wv == ((j & 15) >> 2)) hreg_reflect_b<NBLK, B>(R, S, j, q, tr);
        } else {
            hreg_reflect<NBLK, B + 1>(R, S, j, wv, q, tr);
        }
    }
}

// Right and left update of step i; K = (i + 1) / 16 is the first block of
// rows / columns with an entry >= i + 1.
template <int NBLK, int K>
__device__ __forceinline__ void hreg_step(zd (&R)[NBLK][NBLK], HregSmem &S, int i, int wv, int lane, int tr, int tc)
{
    const int buf = i & 1;
    const zd tau = S.tau[buf];
    const zd *v = S.xv[buf];
    zd vc[NBLK];
#pragma unroll
    for (int b = K; b < NBLK; ++b) vc[b] = v[tc + 16 * b];
    zd p[4] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int b = K; b < NBLK; ++b) p[a] = hreg::fma_add(R[a][b], vc[b], p[a]);
    zd t;
    t.re = hreg::rows_reduce_scatter(p[0].re, p[1].re, p[2].re, p[3].re);
    t.im = hreg::rows_reduce_scatter(p[0].im, p[1].im, p[2].im, p[3].im);
    S.up[wv][lane] = t;   // row lane's sum over this wavefront's columns
    __syncthreads();
    const zd u = zmul(zadd(zadd(zadd(S.up[0][lane], S.up[1][lane]), S.up[2][lane]), S.up[3][lane]), tau);
    double gr[4], gi[4];
    hreg::rows_all_gather(u.re, gr[0], gr[1], gr[2], gr[3]);
    hreg::rows_all_gather(u.im, gi[0], gi[1], gi[2], gi[3]);
#pragma unroll
    for (int b = K; b < NBLK; ++b) {
        const zd vb = v[tc + 16 * b];   // read again: vc is not kept across the barrier
#pragma unroll
        for (int a = 0; a < NBLK; ++a) R[a][b] = hreg::fma_subc(mk(gr[a], gi[a]), vb, R[a][b]);
    }
    lds_order();   // the loads of vr below stay after the rank-1 update
    zd vr[NBLK];
#pragma unroll
    for (int a = K; a < NBLK; ++a) vr[a] = v[tr + 16 * a];
#pragma unroll
    for (int b = K; b < NBLK; ++b) {
        zd s = mk(0.0, 0.0);
#pragma unroll
        for (int a = K; a < NBLK; ++a) s = hreg::fma_addc(vr[a], R[a][b], s);
        s.re = hreg::row_sum(s.re);
        s.im = hreg::row_sum(s.im);
        if (tc + 16 * b <= i) s = mk(0.0, 0.0);   // columns 0..i are not touched from the left
        s = zmulc(tau, s);
#pragma unroll
        for (int a = K; a < NBLK; ++a) R[a][b] = hreg::fma_sub(vr[a], s, R[a][b]);
    }
}

// Steps i0 <= i <= G - 2 of the grid (G = 16 NBLK) with (i + 1) / 16 == K,
// then the next phase: one loop per K, so the loop body has a single path
// and R stays in place.
template <int NBLK, int K = 0>
__device__ __forceinline__ void hreg_phases(zd (&R)[NBLK][NBLK], HregSmem &S, int i0, int wv, int lane, int q,
                                            int tr, int tc)
{
    constexpr int G = 16 * NBLK;
    if constexpr (K < NBLK) {
        const int i1 = 16 * K + 14 < G - 2 ? 16 * K + 14 : G - 2;
        const int ib = K == 0 ? 0 : 16 * K - 1;
        for (int i = ib > i0 ? ib : i0; i <= i1; ++i) {
            __syncthreads();   // v and tau of step i are in LDS
            hreg_step<NBLK, K>(R, S, i, wv, lane, tr, tc);
            if (i + 3 <= G) hreg_reflect<NBLK>(R, S, i + 1, wv, q, tr);
        }
        hreg_phases<NBLK, K + 1>(R, S, i0, wv, lane, q, tr, tc);
    }
}

// One matrix (n <= 16 NBLK, column-major at src) by the whole workgroup:
// NaN/Inf check, ZGEEV norm-range scaling (factor 2^-se to eig0[0]), the
// reduction, the packed Hessenberg matrix to dst and the flag (0, or 2 with
// the unreduced matrix in dst).  Ends with a barrier, so S can serve the
// next matrix at once.
// The matrix sits at the bottom right of the G x G grid (G = 16 NBLK):
// element (r, c) at grid (r + p, c + p), p = G - n, the rows and columns
// above and left of it zero.  The reduction then runs grid steps p .. G - 2,
// the last n - 1 steps of a G x G reduction, whose active blocks shrink
// with the step: with the padding at the bottom right instead, every step
// of an n = 49 matrix would update all four 16-row blocks.  At n = G the
// two placements are the same.
// OFF = false: the matrix at the top left (p = 0; for n = 16 NBLK, where
// the placements agree, with fewer registers).
// Wavefronts per SIMD of the kernels around it: four 256-thread workgroups
// per CU at NBLK = 3 (108 VGPRs); three at NBLK = 4, where the 168-VGPR
// budget holds the 4 x 4 blocks without spills (at four workgroups the
// 128-VGPR budget spilled 15 VGPRs, 88 with the offset, and ran 4% slower
// at n = 64).
template <int NBLK>
constexpr int kHregWaves = NBLK == 4 ? 3 : 4;
template <int NBLK, bool OFF = true>
__device__ __forceinline__ void hess_reg(HregSmem &S, const zd *__restrict__ src, int n, zd *__restrict__ dst,
                                         int8_t *__restrict__ flagp, zd *__restrict__ eig0)
{
    const int tid = (int)threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63, q = lane >> 4, tr = lane & 15, tc = 4 * wv + q;
    const int p = OFF ? 16 * NBLK - n : 0;   // grid offset of the matrix
    zd R[NBLK][NBLK];
    double chk = 0.0, mx = 0.0;
#pragma unroll
    for (int b = 0; b < NBLK; ++b)
#pragma unroll
        for (int a = 0; a < NBLK; ++a) {
            const int r = tr + 16 * a - p, c = tc + 16 * b - p;
            const zd x = src[max(c, 0) * n + max(r, 0)];   // in bounds; zero outside n x n
            R[a][b] = (r >= 0 && c >= 0) ? x : mk(0.0, 0.0);
            chk += (R[a][b].re - R[a][b].re) + (R[a][b].im - R[a][b].im);   // NaN iff an entry is not finite
            mx = fmax(mx, fmax(fabs(R[a][b].re), fabs(R[a][b].im)));
        }
    mx = group_max<64>(mx);
    if (lane == 0) S.red[wv] = mx;
    const int ok = __builtin_amdgcn_readfirstlane(__syncthreads_or(!(chk == 0.0)) == 0);
    mx = fmax(fmax(S.red[0], S.red[1]), fmax(S.red[2], S.red[3]));
    const int se = ok ? zgeev_scale_exp(mx) : 0;
    if (se) {
        const double f = ldexp(1.0, se);
#pragma unroll
        for (int a = 0; a < NBLK; ++a)
#pragma unroll
            for (int b = 0; b < NBLK; ++b) R[a][b] = zscale(f, R[a][b]);
    }
    if (ok && tid == 0) eig0[0] = mk(ldexp(1.0, -se), 0.0);
    if (ok) {
        // column p < 16 (n > 16 (NBLK - 1)) is in block 0
        if (n >= 2 && wv == ((p & 15) >> 2)) hreg_reflect_b<NBLK, 0>(R, S, p, q, tr);
        hreg_phases<NBLK>(R, S, p, wv, lane, q, tr, tc);
    }
    const PackedLayout pk;
    int tr_o = tr, tc_o = tc;
    // opaque copies: keeps the compiler from computing the store addresses
    // before the reduction and holding them in registers through it
    __asm__ volatile("" : "+v"(tr_o), "+v"(tc_o));
#pragma unroll
    for (int b = 0; b < NBLK; ++b)
#pragma unroll
        for (int a = 0; a < NBLK; ++a) {
            const int r = tr_o + 16 * a - p, c = tc_o + 16 * b - p;
            if (r >= 0 && c >= 0 && r <= c + 1) dst[pk(r, c)] = (r == c + 1 && ok) ? mk(S.beta[c + p], 0.0) : R[a][b];
        }
    if (tid == 0) *flagp = ok ? 0 : 2;
    __syncthreads();
}

}  // namespace tkz
