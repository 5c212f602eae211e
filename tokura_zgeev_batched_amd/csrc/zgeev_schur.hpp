// Per-matrix building blocks of the Schur-form paths (forward with Schur
// vectors, the adjoint, eigenvectors), shared by the same-size kernels and
// the ragged ones.
#pragma once
#include "zgeev_device.hpp"

// threads per workgroup of the stored-Schur backward for n > 32 (adjoint_wg)
#ifndef TKZ_ADJ_WG_NT
#define TKZ_ADJ_WG_NT 256
#endif

namespace tkz {

template <int NB>
struct AdjGeom {
    static constexpr int LDP = NB | 1;
    static constexpr int MAT = NB * LDP;
    static constexpr int PER = 2 * MAT + 2 * NB;   // buf1, buf2, taus, g
    static constexpr int PER_GZ = MAT + 2 * NB;    // buf1, taus, g (Z in global memory)
};

// Coalesced store of an n x n LDS matrix to column-major global memory.
template <int LPM>
__device__ __forceinline__ void store_full(const zd *H, const FullLayout lay, zd *__restrict__ dst, int n, int t)
{
    const int nn = n * n;
    int r = t % n, c = t / n;
    const int dr = LPM % n, dc = LPM / n;
    for (int e = t; e < nn; e += LPM) {
        dst[e] = H[lay(r, c)];
        r += dr;
        c += dc;
        if (r >= n) { r -= n; ++c; }
    }
}

template <int LPM>
__device__ __forceinline__ void fill_nan(zd *__restrict__ dst, int count, int t)
{
    for (int e = t; e < count; e += LPM) dst[e] = mk(__builtin_nan(""), __builtin_nan(""));
}

// Q = H(0) H(1) ... H(n-2) formed in place of the reflectors in A (LAPACK
// ZUNG2R's backward accumulation, the arithmetic of form_q): before H(i) is
// applied, row and column i+1 of Q are set to e_{i+1}.  Q(i+1:n, i+1:n)
// then covers only columns whose reflectors were applied already, while
// H(i)'s vector still sits in column i below row i+1.  The Hessenberg part
// of A is overwritten; the caller saves it first.
template <int LPM>
__device__ __forceinline__ void form_q_inplace(zd *A, const FullLayout lay, const zd *taus, int n, int t)
{
    for (int i = n - 2; i >= -1; --i) {
        const int j = i + 1;
        for (int r = j + t; r < n; r += LPM) {
            A[lay(r, j)] = mk(r == j ? 1.0 : 0.0, 0.0);
            if (r != j) A[lay(j, r)] = mk(0.0, 0.0);
        }
        lds_order();
        if (i < 0) break;
        const zd tau = taus[i];
        if (tau.re == 0.0 && tau.im == 0.0) continue;
        const int m = n - 1 - i;
        for (int c = i + 1 + t; c < n; c += LPM) {
            zd s = A[lay(i + 1, c)];
            for (int r = 1; r < m; ++r) s = zadd(s, zmulc(A[lay(i + 1 + r, i)], A[lay(i + 1 + r, c)]));
            s = zmul(tau, s);
            A[lay(i + 1, c)] = zsub(A[lay(i + 1, c)], s);
            for (int r = 1; r < m; ++r) A[lay(i + 1 + r, c)] = zsub(A[lay(i + 1 + r, c)], zmul(A[lay(i + 1 + r, i)], s));
        }
        lds_order();
    }
}

// The Schur-form QR: the two-bulge sweeps (zlahqr_dual with WANTT, one
// matrix per wavefront) for n > 32, the single-shift ZLAHQR below (two or
// more matrices per wavefront, where the single chain serves them all), in
// every Schur path (same-size, ragged, eigenvectors, the recompute backward),
// so that all of them give the same T and U bits, and the eigenvalue order of
// the matching eigenvalue-only kernels (for n > 32 the two-bulge QR after the
// one-wavefront Hessenberg).
__host__ __device__ constexpr bool schur_dual(int n) { return n > 32; }

// In-place conversion of the Hessenberg part of a full LDS layout (ld) to
// the packed layout (column c: rows 0..c+1 from offset c(c+3)/2) and back.
// Packed column c ends where full column c+1 starts or before, so columns
// move in ascending order (descending to unpack); within a column every lane
// reads its element before any lane writes (n <= LPM).
template <int LPM>
__device__ __forceinline__ void pack_hessenberg(zd *H, const FullLayout lay, int n, int t)
{
    const PackedLayout pk;
    for (int c = 0; c < n; ++c) {
        const int rows = c + 2 < n ? c + 2 : n;
        zd v = mk(0.0, 0.0);
        if (t < rows) v = H[lay(t, c)];
        lds_order();
        if (t < rows) H[pk(t, c)] = v;
        lds_order();
    }
}

template <int LPM>
__device__ __forceinline__ void unpack_hessenberg(zd *H, const FullLayout lay, int n, int t)
{
    const PackedLayout pk;
    for (int c = n - 1; c >= 0; --c) {
        const int rows = c + 2 < n ? c + 2 : n;
        zd v = mk(0.0, 0.0);
        if (t < rows) v = H[pk(t, c)];
        lds_order();
        if (t < rows) H[lay(t, c)] = v;
        lds_order();
    }
}

// form_q_inplace by a workgroup of NT threads (teams of TPR threads share a
// column; each column's dot product is split over the team's lanes and
// combined with shuffles, as hessenberg_wg's left update), for the
// reflectors and taus of hessenberg_wg<NT, TPR, true>.
template <int NT, int TPR>
__device__ __forceinline__ void form_q_wg(zd *A, const FullLayout lay, const zd *taus, int n)
{
    const int tid = (int)threadIdx.x, q = tid % TPR, team = tid / TPR;
    for (int i = n - 2; i >= -1; --i) {
        const int j = i + 1;
        for (int r = j + tid; r < n; r += NT) {
            A[lay(r, j)] = mk(r == j ? 1.0 : 0.0, 0.0);
            if (r != j) A[lay(j, r)] = mk(0.0, 0.0);
        }
        __syncthreads();
        if (i < 0) break;
        const zd tau = taus[i];
        if (tau.re == 0.0 && tau.im == 0.0) continue;
        const int m = n - 1 - i;
        const zd *v = A + lay(i + 1, i);   // v(r) = v[r], r >= 1, v(0) = 1
        for (int c = i + 1 + team; c < n; c += NT / TPR) {
            zd *col = A + lay(i + 1, c);
            zd sv = q == 0 ? col[0] : mk(0.0, 0.0);
            for (int r = (q == 0 ? TPR : q); r < m; r += TPR) sv = zadd(sv, zmulc(v[r], col[r]));
#pragma unroll
            for (int o = 1; o < TPR; o <<= 1) {
                sv.re += __shfl_xor(sv.re, o, TPR);
                sv.im += __shfl_xor(sv.im, o, TPR);
            }
            sv = zmul(tau, sv);
            if (q == 0) col[0] = zsub(col[0], sv);
            for (int r = (q == 0 ? TPR : q); r < m; r += TPR) col[r] = zsub(col[r], zmul(v[r], sv));
        }
        __syncthreads();
    }
}

// First half of the Schur form for n > 32 by a workgroup of NT threads
// (hessenberg_wg with kept reflectors, form_q_wg): the Hessenberg matrix
// (zeros below the sub-diagonal) to Th and Q to Ug, n x n column-major.
// Returns false for a non-finite input (nothing written).  H holds n x (n|1)
// complex, taus n, red 8 doubles (LDS).
template <int NT, int TPR>
__device__ __forceinline__ bool schur_hq_wg(zd *H, zd *taus, double *red, const zd *__restrict__ src, int n, zd *Th,
                                            zd *Ug)
{
    const int tid = (int)threadIdx.x;
    const FullLayout lay{n | 1};
    const int nn = n * n;
    bool bad = false;
    for (int e = tid; e < nn; e += NT) {
        const zd v = src[e];
        bad |= !zfinite(v);
        H[lay(e % n, e / n)] = v;
    }
    if (__syncthreads_or(bad)) return false;
    hessenberg_wg<NT, TPR, true>(H, lay, n, red, taus);
    for (int e = tid; e < nn; e += NT) {
        const int r = e % n, c = e / n;
        Th[e] = r > c + 1 ? mk(0.0, 0.0) : H[lay(r, c)];
    }
    __syncthreads();
    form_q_wg<NT, TPR>(H, lay, taus, n);
    for (int e = tid; e < nn; e += NT) Ug[e] = H[lay(e % n, e / n)];
    return true;
}

// Schur form of the matrix at src: T into H (upper triangular), U into Z.
// Returns -1 for a non-finite input, else the QR INFO.
// GZ: Z is an n x n column-major global buffer (ld n), so that the LDS holds
// only H (twice the matrices per CU at n = 64).  The Hessenberg matrix is
// parked in Z's home while Q is formed in its LDS buffer (form_q_inplace),
// then the two are swapped.  Every global access here and in the QR's Z
// updates is by the lane that owns the row (lane p: rows p, p+LPM, ...),
// coalesced along columns, so no fence is needed; the results are bitwise
// those of the all-LDS path.
template <int NB, int LPM, bool ZH = false, bool GZ = false>
__device__ __forceinline__ int schur_in_lds(zd *H, zd *Z, zd *taus, const FullLayout lay, const zd *__restrict__ src, int n, int t,
                                            int itmax)
{
    static_assert(!(GZ && ZH), "global Z keeps one row per lane");
    constexpr int S = (NB + LPM - 1) / LPM;
    const FullLayout zl{GZ ? n : lay.ld};
    const bool ok = load_full<LPM>(H, lay, src, n, t);
    lds_order();
    if (!ok) return -1;
    hessenberg<LPM, true>(H, lay, n, t, taus);
    if constexpr (GZ) {
        for (int r = t; r < n; r += LPM)
            for (int c = 0; c < n; ++c) Z[r + c * n] = r > c + 1 ? mk(0.0, 0.0) : H[lay(r, c)];
        lds_order();
        form_q_inplace<LPM>(H, lay, taus, n, t);
        for (int r = t; r < n; r += LPM)
            for (int c = 0; c < n; ++c) {
                const zd h = Z[r + c * n];
                Z[r + c * n] = H[lay(r, c)];
                H[lay(r, c)] = h;
            }
        lds_order();
    } else {
        form_q<LPM>(Z, zl, H, lay, taus, n, t);
    }
    int info = 0;
    if constexpr (LPM == 64) {
        if (schur_dual(n)) {   // n <= NB <= LPM: one position and one row of Z per lane
            pack_hessenberg<LPM>(H, lay, n, t);
            info = zlahqr_dual<LPM, true>(H, PackedLayout{}, n, t, itmax, Z, zl.ld);
            unpack_hessenberg<LPM>(H, lay, n, t);
        } else if (n > 1) {
            info = zlahqr_core<LPM, S, true, ZH>(H, lay, n, t, itmax, Z, zl);
        }
    } else {
        if (n > 1) info = zlahqr_core<LPM, S, true, ZH>(H, lay, n, t, itmax, Z, zl);
    }
    for (int c = 0; c < n; ++c)
        for (int r = c + 1 + t; r < n; r += LPM) H[lay(r, c)] = mk(0.0, 0.0);
    lds_order();
    return info;
}

// The Schur form in two kernels, so that the QR, which dominates, runs on
// the packed Hessenberg layout (34 KB of LDS at n = 64, four matrices per CU;
// 8.7 KB at n = 32) with U in global memory.  Same arithmetic as
// schur_in_lds, hence the same bits.
//
// First half (full LDS layout, LPM lanes per matrix): Hessenberg with kept
// reflectors; the Hessenberg matrix (zeros below the subdiagonal) to Th and
// Q (ZUNGHR, formed in place of the reflectors) to Ug, both n x n
// column-major in global memory and written row by row by the owning lane.
// Returns false for a non-finite input (nothing written).
template <int LPM>
__device__ __forceinline__ bool schur_hq_one(zd *H, zd *taus, const zd *__restrict__ src, int n, int t, zd *Th,
                                             zd *Ug)
{
    const FullLayout lay{n | 1};
    const bool ok = load_full<LPM>(H, lay, src, n, t);
    lds_order();
    if (!ok) return false;
    hessenberg<LPM, true>(H, lay, n, t, taus);
    for (int r = t; r < n; r += LPM)
        for (int c = 0; c < n; ++c) Th[r + c * n] = r > c + 1 ? mk(0.0, 0.0) : H[lay(r, c)];
    lds_order();
    form_q_inplace<LPM>(H, lay, taus, n, t);
    for (int r = t; r < n; r += LPM)
        for (int c = 0; c < n; ++c) Ug[r + c * n] = H[lay(r, c)];
    lds_order();
    return true;
}

// Second half (packed LDS layout, packed_size(n) complex; n <= LPM): the
// Schur-form QR (two bulges per sweep for n > 32, else ZLAHQR) with
// WANTT/WANTZ on the Hessenberg matrix from Th, Z = Ug updated in place
// (lane p owns row p, as it wrote it), then the eigenvalues to out and T
// (zeros below the diagonal) back to Th, row by row by the lane that read
// it.  Returns the QR INFO.  SNAP: `known` holds the matrix's eigenvalues
// (n, any order) and the QR's shifts are snapped to them (snap_shift): the
// same Schur form up to rounding in about half the sweeps.
template <int LPM, bool DENSE = false, bool SNAP = false>
__device__ __forceinline__ int schur_qr_one(zd *H, int n, int t, zd *Th, zd *Ug, zd *__restrict__ out, int itmax,
                                            const zd *__restrict__ known = nullptr)
{
    const PackedLayout lay{};
    const int r = t;
    if (r < n)
        for (int c = r > 0 ? r - 1 : 0; c < n; ++c) H[lay(r, c)] = Th[r + c * n];
    const bool kh = SNAP && r < n;
    const zd kev = kh ? known[r] : zd{};
    lds_order();
    int info = 0;
    if constexpr (LPM == 64) {
        // DENSE: H has a spare element past the matrix for the idle lanes
        if (schur_dual(n))
            info = zlahqr_dual<LPM, true, DENSE, NoMove, SNAP>(H, lay, n, t, itmax, Ug, n, nullptr, -1, kev, kh);
        else if (n > 1)   // inlined: as a call, its ABI would raise the kernel's VGPRs to 274 and add stack
            [[clang::always_inline]] info = zlahqr_core<LPM, 1, true, false, PackedLayout, FullLayout, false, SNAP>(
                H, lay, n, t, itmax, Ug, FullLayout{n}, 0, 0, kev, kh);
    } else {
        if (n > 1)
            info = zlahqr_core<LPM, 1, true, false, PackedLayout, FullLayout, false, SNAP>(
                H, lay, n, t, itmax, Ug, FullLayout{n}, 0, 0, kev, kh);
    }
    lds_order();
    if (r < n) {
        out[r] = H[lay(r, r)];
        for (int c = 0; c < n; ++c) Th[r + c * n] = r > c ? mk(0.0, 0.0) : H[lay(r, c)];
    }
    return info;
}

// Left eigenvectors of the upper triangular T held in the upper part of B
// (ZTREVC, SIDE = 'L', without rescaling): y_i^H (T - t_ii) = 0 with
// y_i(i) = 1 and y_i(r) = 0 for r < i; y_i(r), r > i, is written to the
// strictly lower part of B at (r, i).  Near-singular pivots become
// smin = max(ulp |t_ii|, smlnum) as in ZTREVC.
template <int LPM, int S>
__device__ void trevc_left_lower(zd *B, const FullLayout lay, int n, int t)
{
    const double ulp = DBL_EPSILON, smlnum = DBL_MIN * ((double)n / ulp);
    for (int r = 1; r < n; ++r) {
        const zd trr = B[lay(r, r)];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = t + s * LPM;
            if (i < r) {
                zd acc = zconj(B[lay(i, r)]);   // c = i, y_i(i) = 1
                for (int c = i + 1; c < r; ++c) acc = zadd(acc, zmulc(B[lay(c, r)], B[lay(c, i)]));
                const zd lam = B[lay(i, i)];
                const double smin = fmax(ulp * cabs1(lam), smlnum);
                zd dd = zconj(zsub(trr, lam));
                if (cabs1(dd) < smin) dd = mk(smin, 0.0);
                zd q = zdiv(acc, dd);
                B[lay(r, i)] = mk(-q.re, -q.im);
            }
        }
        lds_order();
    }
}

// Right eigenvectors of the upper triangular T in the upper part of B
// (ZTREVC, SIDE = 'R', without rescaling): (T - t_ii) x_i = 0 with
// x_i(i) = 1 and x_i(r) = 0 for r > i; x_i(r), r < i, is written TRANSPOSED
// to the strictly lower part of B at (i, r), so T stays intact.
template <int LPM, int S>
__device__ void trevc_right_lower_t(zd *B, const FullLayout lay, int n, int t)
{
    const double ulp = DBL_EPSILON, smlnum = DBL_MIN * ((double)n / ulp);
    for (int r = n - 2; r >= 0; --r) {
        const zd trr = B[lay(r, r)];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = t + s * LPM;
            if (i > r && i < n) {
                zd acc = mk(0.0, 0.0);
                for (int c = r + 1; c < i; ++c) acc = zadd(acc, zmul(B[lay(r, c)], B[lay(i, c)]));
                acc = zadd(acc, B[lay(r, i)]);   // c = i, x_i(i) = 1
                const zd lam = B[lay(i, i)];
                const double smin = fmax(ulp * cabs1(lam), smlnum);
                zd dd = zsub(trr, lam);
                if (cabs1(dd) < smin) dd = mk(smin, 0.0);
                zd q = zdiv(acc, dd);
                B[lay(i, r)] = mk(-q.re, -q.im);
            }
        }
        lds_order();
    }
}

// ZGEEV's normalisation of one eigenvector held one component per lane
// (S == 1): unit Euclidean norm with DZNRM2-style scaling, then the component
// of largest |re|^2 + |im|^2 (the first on ties, as IDAMAX) is rotated to be
// real and non-negative with an imaginary part of exactly zero.  Lanes with
// `live` false hold no component and return their input.
template <int LPM>
__device__ __forceinline__ zd zgeev_normalise(zd v, bool live, int r)
{
    const double amx = group_max<LPM>(live ? fmax(fabs(v.re), fabs(v.im)) : 0.0);
    if (amx == 0.0) return v;
    const double ia = 1.0 / amx;
    const double sr = v.re * ia, si = v.im * ia;
    const double nrm = amx * sqrt(group_sum<LPM>(live ? sr * sr + si * si : 0.0));
    const double scl = 1.0 / nrm;
    v = zscale(scl, v);
    double best = live ? v.re * v.re + v.im * v.im : -1.0;
    int bi = live ? r : 0x7fffffff;
#pragma unroll
    for (int o = LPM / 2; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, LPM);
        const int oi = __shfl_xor(bi, o, LPM);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const zd vk = mk(__shfl(v.re, bi % LPM, LPM), __shfl(v.im, bi % LPM, LPM));
    const double inv = 1.0 / sqrt(best);
    v = zmul(v, mk(vk.re * inv, -vk.im * inv));
    if (r == bi) v.im = 0.0;
    return v;
}

// Eigenvectors from a Schur form held as T in LDS (B1, full layout `lay`) and
// U (layout `ul`, in LDS or global memory): ZTREVC, back-transform by U,
// ZGEEV normalisation, into the matrix's VL and/or VR (n x n column-major;
// null = not asked for).  INPLACE: U lives in VR's home (VL's when VR is
// null), so the columns are formed in an order that reads each column of U
// before overwriting it: VL ascending (column i needs U's columns >= i), VR
// descending (columns <= i), VL first.
template <int LPM, bool INPLACE>
__device__ __forceinline__ void vectors_from_schur(zd *B1, const FullLayout lay, const zd *U, const FullLayout ul,
                                                   int n, int t, zd *VLm, zd *VRm)
{
    const bool live = t < n;
    const int r = t;
    // v_i = U x_i = U(:, i) + sum_{c < i} U(:, c) x_i(c); x_i(c) at B1(i, c)
    auto right = [&](int i) {
        zd v = mk(0.0, 0.0);
        if (live) {
            for (int c = 0; c < i; ++c) v = zadd(v, zmul(U[ul(r, c)], B1[lay(i, c)]));
            v = zadd(v, U[ul(r, i)]);
        }
        v = zgeev_normalise<LPM>(v, live, r);
        if (live) VRm[r + (long long)i * n] = v;
    };
    // u_i = U y_i = U(:, i) + sum_{c > i} U(:, c) y_i(c); y_i(c) at B1(c, i)
    auto left = [&](int i) {
        zd v = mk(0.0, 0.0);
        if (live) {
            v = U[ul(r, i)];
            for (int c = i + 1; c < n; ++c) v = zadd(v, zmul(U[ul(r, c)], B1[lay(c, i)]));
        }
        v = zgeev_normalise<LPM>(v, live, r);
        if (live) VLm[r + (long long)i * n] = v;
    };
    if constexpr (INPLACE) {
        if (VLm) {
            trevc_left_lower<LPM, 1>(B1, lay, n, t);
            for (int i = 0; i < n; ++i) left(i);
            lds_order();
        }
        if (VRm) {
            trevc_right_lower_t<LPM, 1>(B1, lay, n, t);
            for (int i = n - 1; i >= 0; --i) right(i);
        }
    } else {
        if (VRm) {
            trevc_right_lower_t<LPM, 1>(B1, lay, n, t);
            for (int i = 0; i < n; ++i) right(i);
            lds_order();
        }
        if (VLm) {
            trevc_left_lower<LPM, 1>(B1, lay, n, t);
            for (int i = 0; i < n; ++i) left(i);
        }
    }
}

// X (unit upper triangular, into B2) of the right eigenvectors of the upper
// triangular T in B1: (T - t_ii) x_i = 0 for rows r < i (ZTREVC without
// rescaling; near-equal eigenvalues perturbed to ulp*|t_ii|).  Lane i owns
// column i; only B1's upper part and diagonal are read.
template <int LPM, int S>
__device__ __forceinline__ void trevc_right_upper(const zd *B1, zd *B2, const FullLayout lay, int n, int t)
{
    const double ulp = DBL_EPSILON, smlnum = DBL_MIN * ((double)n / ulp);
    for (int c = t; c < n; c += LPM)
        for (int r = 0; r < n; ++r) B2[lay(r, c)] = mk(r == c ? 1.0 : 0.0, 0.0);
    lds_order();
    for (int r = n - 2; r >= 0; --r) {
        const zd trr = B1[lay(r, r)];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = t + s * LPM;
            if (i > r && i < n) {
                zd acc = mk(0.0, 0.0);
                for (int c = r + 1; c <= i; ++c) acc = zadd(acc, zmul(B1[lay(r, c)], B2[lay(c, i)]));
                const zd lam = B1[lay(i, i)];
                const double smin = fmax(ulp * cabs1(lam), smlnum);
                zd dd = zsub(trr, lam);
                if (cabs1(dd) < smin) dd = mk(smin, 0.0);
                zd q = zdiv(acc, dd);
                B2[lay(r, i)] = mk(-q.re, -q.im);
            }
        }
        lds_order();
    }
}

// Pairs the eigenvalues of a Schur form (T's diagonal, this function's
// order) with those of a forward pass (ef[0..n), any order): perm[i] = the
// index j of ef closest to T(i,i) in |Re| + |Im|.  Each position i finds its
// nearest j; when two positions pick the same j (near-equal eigenvalues), the
// group's first lane assigns greedily in position order instead, every j
// used once.  Returns false when a pair lies further apart than 1e-6 times
// the largest |T(i,i)| (then the eigenvalues are not those of this matrix).
// Positions i = t + s*LPM (S per lane); perm and ef are LDS, all of the group
// reads them after the call.
template <int LPM, int S>
__device__ __forceinline__ bool match_eigenvalues(const zd *B1, const FullLayout lay, const zd *ef, int *perm, int n,
                                                  int t)
{
    double scale = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = t + s * LPM;
        if (i < n) scale = fmax(scale, cabs1(B1[lay(i, i)]));
    }
    scale = group_max<LPM>(scale);
    const double tol = 1e-6 * scale;
    bool clash = false, far = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = t + s * LPM;
        if (i < n) {
            const zd w = B1[lay(i, i)];
            int best = 0;
            double bd = cabs1(zsub(w, ef[0]));
            for (int j = 1; j < n; ++j) {
                const double d = cabs1(zsub(w, ef[j]));
                if (d < bd) { bd = d; best = j; }
            }
            perm[i] = best;
            far |= !(bd <= tol);
        }
    }
    lds_order();
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = t + s * LPM;
        if (i < n)
            for (int k = 0; k < n; ++k) clash |= (k != i && perm[k] == perm[i]);
    }
    if (group_any<LPM>(clash)) {
        lds_order();
        far = false;
        if (t == 0) {
            uint64_t used = 0;
            for (int i = 0; i < n; ++i) {
                const zd w = B1[lay(i, i)];
                int best = -1;
                double bd = 0.0;
                for (int j = 0; j < n; ++j) {
                    if ((used >> j) & 1ull) continue;
                    const double d = cabs1(zsub(w, ef[j]));
                    if (best < 0 || d < bd) { bd = d; best = j; }
                }
                used |= 1ull << best;
                perm[i] = best;
                far |= !(bd <= tol);
            }
        }
    }
    lds_order();
    return !group_any<LPM>(far);
}

// dL/dA of one matrix into dst (n x n column-major) from its Schur form, T
// at Tg and U at Ug (n x n column-major each), for cotangents gbar_m, with
// the buffers B1, B2, taus, gv of AdjGeom<NB>.  f is the forward's flag; a
// matrix with f != 0 gets dA = NaN.  perm (LDS, may be null): cotangent i is
// gbar_m[perm[i]].  dst may be Ug: every read of Ug precedes, in the same
// lane, the first write of dst.
template <int NB, int LPM>
__device__ __forceinline__ void adjoint_one(zd *B1, zd *B2, zd *gv, int n, int t, const zd *__restrict__ Tg,
                                            const zd *Ug, const zd *__restrict__ gbar_m, const int *perm, zd *dst,
                                            int f, bool loaded = false)
{
    constexpr int S = (NB + LPM - 1) / LPM;
    const FullLayout lay{n | 1};
    const long long nn = (long long)n * n;
    if (f != 0) {
        fill_nan<LPM>(dst, (int)nn, t);
        return;
    }
    if (!loaded) {
        load_full<LPM>(B1, lay, Tg, n, t);
        lds_order();
    }
    for (int i = t; i < n; i += LPM) gv[i] = gbar_m[perm ? perm[i] : i];
    trevc_right_upper<LPM, S>(B1, B2, lay, n, t);
    // ---- Y (buf1 strictly lower, y_i(i) = 1): y_i^H (T - t_ii) = 0, rows r > i
    trevc_left_lower<LPM, S>(B1, lay, n, t);
    // ---- W = Y diag(g) X^H (lower triangular) into buf1, one row at a time
    for (int a = 0; a < n; ++a) {
        zd w[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int b = t + s * LPM;
            w[s] = mk(0.0, 0.0);
            if (b <= a) {
                zd acc = mk(0.0, 0.0);
                for (int i = b; i <= a; ++i) {
                    const zd y = (i == a) ? mk(1.0, 0.0) : B1[lay(a, i)];
                    const zd x = (i == b) ? mk(1.0, 0.0) : B2[lay(b, i)];
                    acc = zadd(acc, zmul(zmul(y, gv[i]), zconj(x)));
                }
                w[s] = acc;
            }
        }
        lds_order();
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int b = t + s * LPM;
            if (b <= a) B1[lay(a, b)] = w[s];
        }
        lds_order();
    }
    // ---- P = W U^H into buf2 (X is consumed); U read from global once
    for (int c = t; c < n; c += LPM)
        for (int a = 0; a < n; ++a) B2[lay(a, c)] = mk(0.0, 0.0);
    lds_order();
    for (int b = 0; b < n; ++b) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int c = t + s * LPM;
            if (c < n) {
                const zd u = zconj(Ug[c + (long long)b * n]);
                for (int a = b; a < n; ++a) B2[lay(a, c)] = zadd(B2[lay(a, c)], zmul(B1[lay(a, b)], u));
            }
        }
    }
    lds_order();
    // ---- dA = U P: U into buf1, coalesced, then rows of dA per lane
    load_full<LPM>(B1, lay, Ug, n, t);
    lds_order();
    for (int c = 0; c < n; ++c) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int r = t + s * LPM;
            if (r < n) {
                zd acc = mk(0.0, 0.0);
                for (int a = 0; a < n; ++a) acc = zadd(acc, zmul(B1[lay(r, a)], B2[lay(a, c)]));
                dst[r + (long long)c * n] = acc;
            }
        }
    }
}

// adjoint_wg's phases after T (upper triangular, full layout) is in B1 and
// the cotangent in gv: U is read once from Ug (global memory), dA written to
// dst (which may be where U was parked: every read of Ug precedes a
// workgroup barrier before the first write of dst).
template <int NT>
__device__ __forceinline__ void adjoint_wg_phases(zd *B1, zd *B2, const zd *gv, int n, const zd *Ug, zd *dst)
{
    static_assert(NT % 64 == 0 && NT >= 128, "X and Y need one wavefront each");
    constexpr int NW = NT / 64;
    const int tid = (int)threadIdx.x, w = tid >> 6, t = tid & 63;
    const FullLayout lay{n | 1};
    const int nn = n * n;
    if (w == 0) trevc_right_upper<64, 1>(B1, B2, lay, n, t);
    else if (w == 1) trevc_left_lower<64, 1>(B1, lay, n, t);
    __syncthreads();
    // W = Y diag(g) X^H (lower triangular), row a: lanes b <= a
    for (int a = w; a < n; a += NW) {
        const int b = t;
        zd acc = mk(0.0, 0.0);
        if (b <= a) {
            for (int i = b; i <= a; ++i) {
                const zd y = (i == a) ? mk(1.0, 0.0) : B1[lay(a, i)];
                const zd x = (i == b) ? mk(1.0, 0.0) : B2[lay(b, i)];
                acc = zadd(acc, zmul(zmul(y, gv[i]), zconj(x)));
            }
        }
        lds_order();
        if (b <= a) B1[lay(a, b)] = acc;
    }
    __syncthreads();
    for (int e = tid; e < nn; e += NT) B2[lay(e % n, e / n)] = Ug[e];
    __syncthreads();
    // P = W U^H, row a over W's row a: lane c
    for (int a = w; a < n; a += NW) {
        const int c = t;
        zd acc = mk(0.0, 0.0);
        if (c < n)
            for (int b = 0; b <= a; ++b) acc = zadd(acc, zmul(B1[lay(a, b)], zconj(B2[lay(c, b)])));
        lds_order();
        if (c < n) B1[lay(a, c)] = acc;
    }
    __syncthreads();
    // dA = U P, column c: lane r
    for (int c = w; c < n; c += NW) {
        const int r = t;
        if (r < n) {
            zd acc = mk(0.0, 0.0);
            for (int a = 0; a < n; ++a) acc = zadd(acc, zmul(B2[lay(r, a)], B1[lay(a, c)]));
            dst[r + c * n] = acc;
        }
    }
}

// dL/dA of one matrix (n <= 64) from its Schur form (T at Tg, U at Ug, n x n
// column-major each) by a workgroup of NT threads (NT/64 wavefronts):
// adjoint_one's quantities with the same per-element arithmetic, so bitwise
// its results, spread over the wavefronts.  X (wavefront 0) and Y (wavefront
// 1) are solved side by side; then wavefront w forms rows a = w (mod NT/64)
// of W, rows of P = W U^H over W's rows (a row of P needs only its own row
// of W, so no other wavefront reads what is overwritten), and columns
// c = w (mod NT/64) of dA = U P.  Buffers (AdjGeom<NB>): B1 = T/Y -> W -> P,
// B2 = X -> U, gv = the cotangent.  `f` is the matrix's forward flag
// (uniform over the workgroup).  With ef (the forward's eigenvalues, global)
// the cotangents are matched to T's diagonal first (match_eigenvalues by
// wavefront 0, perm and efl: n ints and n complex of LDS); a failed match
// returns 3 with dA = NaN.  Returns the flag the matrix ends with.  dst may be
// Ug (every read of Ug precedes a workgroup barrier before the first write
// of dst).
template <int NT>
__device__ __forceinline__ int adjoint_wg(zd *B1, zd *B2, zd *gv, int n, const zd *__restrict__ Tg, const zd *Ug,
                                          const zd *__restrict__ gbar_m, zd *dst, int f,
                                          const zd *__restrict__ ef = nullptr, zd *efl = nullptr,
                                          int *perm = nullptr, int *ishared = nullptr)
{
    const int tid = (int)threadIdx.x;
    const FullLayout lay{n | 1};
    const int nn = n * n;
    if (f != 0) {
        fill_nan<NT>(dst, nn, tid);
        return f;
    }
    for (int e = tid; e < nn; e += NT) B1[lay(e % n, e / n)] = Tg[e];
    if (ef) {
        for (int i = tid; i < n; i += NT) efl[i] = ef[i];
        __syncthreads();
        if (tid < 64) {
            const bool ok = match_eigenvalues<64, 1>(B1, lay, efl, perm, n, tid);
            if (tid == 0) *ishared = ok ? 0 : 3;
        }
        __syncthreads();
        if (*ishared != 0) {
            fill_nan<NT>(dst, nn, tid);
            return 3;
        }
    }
    for (int i = tid; i < n; i += NT) gv[i] = gbar_m[perm && ef ? perm[i] : i];
    __syncthreads();
    adjoint_wg_phases<NT>(B1, B2, gv, n, Ug, dst);
    return 0;
}

// The Schur form of one matrix for n > 32 in LDS by a workgroup of NT
// threads, with the arithmetic of the two-launch forward (schur_hq_wg's
// Hessenberg and Q, the two-bulge QR by the first wavefront): T into H (upper
// triangular, full layout), U into Z, an n x n column-major buffer with
// leading dimension zld in LDS or in global memory (the Hessenberg matrix is
// parked there while Q is formed in H).  taus: n complex, red: 8 doubles,
// ishared: one int, all LDS.  Returns -1 for a non-finite input, else the QR
// INFO (uniform over the workgroup).
template <int NT>
__device__ __forceinline__ int schur_wg(zd *H, zd *Z, int zld, zd *taus, double *red, int *ishared,
                                        const zd *__restrict__ src, int n, int itmax)
{
    const int tid = (int)threadIdx.x, w = tid >> 6, t = tid & 63;
    const FullLayout lay{n | 1};
    const int nn = n * n;
    bool bad = false;
    for (int e = tid; e < nn; e += NT) {
        const zd v = src[e];
        bad |= !zfinite(v);
        H[lay(e % n, e / n)] = v;
    }
    if (__syncthreads_or(bad)) return -1;
    hessenberg_wg<NT, NT / 64, true>(H, lay, n, red, taus);
    // park the Hessenberg matrix in Z's home, form Q in place, swap (each
    // element by the thread that parked it)
    for (int e = tid; e < nn; e += NT) {
        const int r = e % n, c = e / n;
        Z[r + c * zld] = r > c + 1 ? mk(0.0, 0.0) : H[lay(r, c)];
    }
    __syncthreads();
    form_q_wg<NT, NT / 64>(H, lay, taus, n);
    for (int e = tid; e < nn; e += NT) {
        const int r = e % n, c = e / n;
        const zd h = Z[r + c * zld];
        Z[r + c * zld] = H[lay(r, c)];
        H[lay(r, c)] = h;
    }
    __syncthreads();
    if (w == 0) {
        pack_hessenberg<64>(H, lay, n, t);
        // DENSE: the spare element past the packed matrix lies inside the full-layout buffer
        const int info = zlahqr_dual<64, true, true>(H, PackedLayout{}, n, t, itmax, Z, zld);
        unpack_hessenberg<64>(H, lay, n, t);
        for (int c = 0; c < n; ++c)
            for (int r = c + 1 + t; r < n; r += 64) H[lay(r, c)] = mk(0.0, 0.0);
        if (t == 0) *ishared = info;
    }
    __syncthreads();
    return *ishared;
}

}  // namespace tkz
