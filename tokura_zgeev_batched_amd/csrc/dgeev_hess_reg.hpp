// Householder Hessenberg reduction of a real matrix (LAPACK DGEHD2 / DLARFG
// arithmetic, as dhessenberg) by one 256-thread workgroup with the matrix
// in registers: the real counterpart of hess_reg (zgeev_hess_reg.hpp), with
// the same distribution, the same two barriers per column and the same
// placement of n < 16 NBLK matrices at the bottom right of the grid.  No
// counterpart in the reference, which solves complex matrices only.
//
// Distribution: the 256 threads form a 16 x 16 grid.  Thread (tr, tc) holds
// grid element (tr + 16a, tc + 16b) for a, b < NBLK in R[a][b]; lane
// tr + 16 (tc % 4) of wavefront tc / 4.  The matrix element (r, c) sits at
// grid (r + p, c + p), p = 16 NBLK - n, with zeros above and left of it, so
// the reduction runs grid steps p .. 16 NBLK - 2.
// Step i (reflector v from column i, v(i+1) = 1, v(r) = 0 for r <= i):
//  * the wavefront holding column i computes beta, tau and v after its own
//    update of that column in step i-1 and writes v and tau to LDS;
//  * right update A(:, i+1:) -= tau (A v) v^T: row sums over a thread's
//    columns, over the wavefront's four DPP rows (v_permlane16_swap /
//    v_permlane32_swap), then over the four wavefronts through LDS;
//  * left update A(i+1:, i+1:) -= tau v (v^T A): column sums over a thread's
//    rows and the 16 lanes of its DPP row (DPP moves), no LDS, no barrier.
#pragma once
#include "dgeev_device.hpp"
#include "zgeev_hess_reg.hpp"

namespace tkz {

struct DHregSmem {
    double xv[2][64];   // reflector vector of step i in xv[i & 1]
    double up[4][64];   // per-wavefront row sums of the right update
    double tau[2];
    double beta[64];    // H(j+1, j) of the reduced matrix (grid column j)
};

// Reflector of grid column j by the wavefront holding it (column block B =
// j / 16): tau (and v, unless tau = 0) to LDS buffer j & 1, H(j+1, j) to
// S.beta[j].  R's column j is left as it is: no later step reads it.
template <int NBLK, int B>
__device__ __forceinline__ void dhreg_reflect_b(const double (&R)[NBLK][NBLK], DHregSmem &S, int j, int q, int tr)
{
    const int qo = (j & 15) & 3;
    double ssq = 0.0, amx = 0.0;
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
        if (tr + 16 * a >= j + 2) {
            ssq += R[a][B] * R[a][B];
            amx = fmax(amx, fabs(R[a][B]));
        }
    const int l0 = qo * 16;
    ssq = readlane_d(hreg::row_sum(ssq), l0);
    amx = readlane_d(hreg::row_max(amx), l0);
    const int aa = (j + 1) >> 4, la = l0 + ((j + 1) & 15);
    double ca = 0.0;
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
        if (a == aa) ca = R[a][B];
    const double alpha = readlane_d(ca, la);
    double xnorm = 0.0;
    if (amx > 0.0) {
        if (ssq > 1e-290 && ssq < 1e290) {
            xnorm = sqrt(ssq);
        } else {   // rare: rescaled second pass
            const double ia = 1.0 / amx;
            double s2 = 0.0;
#pragma unroll
            for (int a = 0; a < NBLK; ++a)
                if (tr + 16 * a >= j + 2) s2 += (R[a][B] * ia) * (R[a][B] * ia);
            xnorm = amx * sqrt(readlane_d(hreg::row_sum(s2), l0));
        }
    }
    const int buf = j & 1;
    double beta, tau, scal;
    if (xnorm == 0.0) {   // tau = 0, H(j) = I: the updates of step j add zeros
        beta = alpha;
        tau = 0.0;
        scal = 0.0;
    } else {
        beta = -copysign(dlapy2_d(alpha, xnorm), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    if (q == qo) {
#pragma unroll
        for (int a = 0; a < NBLK; ++a) {
            const int r = tr + 16 * a;
            S.xv[buf][r] = r == j + 1 ? 1.0 : (r > j + 1 ? R[a][B] * scal : 0.0);
        }
    }
    S.tau[buf] = tau;   // the same value from every lane
    S.beta[j] = beta;
}

template <int NBLK, int B = 0>
__device__ __forceinline__ void dhreg_reflect(const double (&R)[NBLK][NBLK], DHregSmem &S, int j, int wv, int q,
                                              int tr)
{
    if constexpr (B < NBLK) {
        if ((j >> 4) == B) {
            if (wv == ((j & 15) >> 2)) dhreg_reflect_b<NBLK, B>(R, S, j, q, tr);
        } else {
            dhreg_reflect<NBLK, B + 1>(R, S, j, wv, q, tr);
        }
    }
}

// Right and left update of step i; K = (i + 1) / 16 is the first block of
// rows / columns with an entry >= i + 1.
template <int NBLK, int K>
__device__ __forceinline__ void dhreg_step(double (&R)[NBLK][NBLK], DHregSmem &S, int i, int wv, int lane, int tr,
                                           int tc)
{
    const int buf = i & 1;
    const double tau = S.tau[buf];
    const double *v = S.xv[buf];
    double vc[NBLK];
#pragma unroll
    for (int b = K; b < NBLK; ++b) vc[b] = v[tc + 16 * b];
    double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int b = K; b < NBLK; ++b) p[a] = fma(R[a][b], vc[b], p[a]);
    S.up[wv][lane] = hreg::rows_reduce_scatter(p[0], p[1], p[2], p[3]);   // row lane's sum over these columns
    __syncthreads();
    const double u = (((S.up[0][lane] + S.up[1][lane]) + S.up[2][lane]) + S.up[3][lane]) * tau;
    double g[4];
    hreg::rows_all_gather(u, g[0], g[1], g[2], g[3]);
#pragma unroll
    for (int b = K; b < NBLK; ++b) {
        const double vb = v[tc + 16 * b];   // read again: vc is not kept across the barrier
#pragma unroll
        for (int a = 0; a < NBLK; ++a) R[a][b] = fma(-g[a], vb, R[a][b]);
    }
    lds_order();   // the loads of vr below stay after the rank-1 update
    double vr[NBLK];
#pragma unroll
    for (int a = K; a < NBLK; ++a) vr[a] = v[tr + 16 * a];
#pragma unroll
    for (int b = K; b < NBLK; ++b) {
        double s = 0.0;
#pragma unroll
        for (int a = K; a < NBLK; ++a) s = fma(vr[a], R[a][b], s);
        s = hreg::row_sum(s);
        if (tc + 16 * b <= i) s = 0.0;   // columns 0..i are not touched from the left
        s *= tau;
#pragma unroll
        for (int a = K; a < NBLK; ++a) R[a][b] = fma(-vr[a], s, R[a][b]);
    }
}

// Grid steps i0 <= i <= G - 2 (G = 16 NBLK) with (i + 1) / 16 == K, then the
// next phase: one loop per K, so the loop body has a single path.
template <int NBLK, int K = 0>
__device__ __forceinline__ void dhreg_phases(double (&R)[NBLK][NBLK], DHregSmem &S, int i0, int wv, int lane, int q,
                                             int tr, int tc)
{
    constexpr int G = 16 * NBLK;
    if constexpr (K < NBLK) {
        const int i1 = 16 * K + 14 < G - 2 ? 16 * K + 14 : G - 2;
        const int ib = K == 0 ? 0 : 16 * K - 1;
        for (int i = ib > i0 ? ib : i0; i <= i1; ++i) {
            __syncthreads();   // v and tau of step i are in LDS
            dhreg_step<NBLK, K>(R, S, i, wv, lane, tr, tc);
            if (i + 3 <= G) dhreg_reflect<NBLK>(R, S, i + 1, wv, q, tr);
        }
        dhreg_phases<NBLK, K + 1>(R, S, i0, wv, lane, q, tr, tc);
    }
}

// One real matrix (16 (NBLK - 1) < n <= 16 NBLK, column-major at src) by the
// whole workgroup: finiteness check, the reduction, the packed Hessenberg
// matrix to dst and the flag (0, or 2 with the unreduced matrix in dst).
// OFF = false places the matrix at the top left (p = 0; for n = 16 NBLK).
template <int NBLK, bool OFF>
__device__ __forceinline__ void dhess_reg(DHregSmem &S, const double *__restrict__ src, int n,
                                          double *__restrict__ dst, int8_t *__restrict__ flagp)
{
    const int tid = (int)threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63, q = lane >> 4, tr = lane & 15, tc = 4 * wv + q;
    const int p = OFF ? 16 * NBLK - n : 0;   // grid offset of the matrix
    double R[NBLK][NBLK];
    bool bad = false;
#pragma unroll
    for (int b = 0; b < NBLK; ++b)
#pragma unroll
        for (int a = 0; a < NBLK; ++a) {
            const int r = tr + 16 * a - p, c = tc + 16 * b - p;
            const double x = src[max(c, 0) * n + max(r, 0)];   // in bounds; zero outside n x n
            R[a][b] = (r >= 0 && c >= 0) ? x : 0.0;
            bad |= !isfinite(R[a][b]);
        }
    const int ok = __builtin_amdgcn_readfirstlane(__syncthreads_or(bad) == 0);
    if (ok) {
        // column p < 16 (n > 16 (NBLK - 1)) is in block 0
        if (n >= 2 && wv == ((p & 15) >> 2)) dhreg_reflect_b<NBLK, 0>(R, S, p, q, tr);
        dhreg_phases<NBLK>(R, S, p, wv, lane, q, tr, tc);
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
            if (r >= 0 && c >= 0 && r <= c + 1) dst[pk(r, c)] = (r == c + 1 && ok) ? S.beta[c + p] : R[a][b];
        }
    if (tid == 0) *flagp = ok ? 0 : 2;
}

}  // namespace tkz
