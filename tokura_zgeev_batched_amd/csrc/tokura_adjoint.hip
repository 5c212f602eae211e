// Differentiation of the eigenvalues: forward with Schur vectors and the
// adjoint (backward) kernel; eigenvectors (ZGEEV with JOBVL/JOBVR = 'V') from
// the same Schur form.  No counterpart in the reference, which has no
// gradient path (README.md:41-44: "No eigenvectors are computed").
//
// Forward (k_zgeev_schur): A = U T U^H by the same method as the eigenvalue
// path, in Schur-form mode: Hessenberg with kept reflectors, Q accumulation
// (ZUNGHR), QR with WANTT/WANTZ (ZLAHQR), T and U stored per matrix.
//
// Backward (k_zgeev_adjoint): for L = Re sum_i conj(g_i) lambda_i,
//   dL/dA = sum_i g_i Y_i X_i^H / conj(Y_i^H X_i)   (torch's convention)
// with right/left eigenvectors X_i = U x_i, Y_i = U y_i, where x_i, y_i are
// eigenvectors of T by back-substitution (ZTREVC) normalised to
// x_i(i) = y_i(i) = 1.  x_i lives on rows <= i and y_i on rows >= i, so
// y_i^H x_i = 1 exactly and
//   dL/dA = U W U^H,  W = sum_i g_i y_i x_i^H  (lower triangular).
// Per matrix two LDS buffers suffice for every n <= 64:
//   buf1: T (upper) + Y (strictly lower) -> W (lower) -> U
//   buf2: Z = U during QR -> X (upper)   -> P = W U^H
// then dA = U P.  In recompute mode (no stored Schur form) U is parked in dA
// between the QR and the final product.
#include "tokura_internal.hpp"
#include "zgeev_device.hpp"

namespace tkz {

template <int NB>
struct AdjGeom {
    static constexpr int LDP = NB | 1;
    static constexpr int MAT = NB * LDP;
    static constexpr int PER = 2 * MAT + 2 * NB;   // buf1, buf2, taus, g
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

// Schur form of the matrix at src: T into H (upper triangular), U into Z.
// Returns -1 for a non-finite input, else the QR INFO.
template <int NB, int LPM, bool ZH = false>
__device__ __forceinline__ int schur_in_lds(zd *H, zd *Z, zd *taus, const FullLayout lay, const zd *__restrict__ src, int n, int t,
                            int itmax)
{
    constexpr int S = (NB + LPM - 1) / LPM;
    const bool ok = load_full<LPM>(H, lay, src, n, t);
    lds_order();
    if (!ok) return -1;
    hessenberg<LPM, true>(H, lay, n, t, taus);
    form_q<LPM>(Z, lay, H, lay, taus, n, t);
    int info = 0;
    if (n > 1) info = zlahqr_core<LPM, S, true, ZH>(H, lay, n, t, itmax, Z, lay);
    for (int c = 0; c < n; ++c)
        for (int r = c + 1 + t; r < n; r += LPM) H[lay(r, c)] = mk(0.0, 0.0);
    lds_order();
    return info;
}

// ZH: the upper half of each lane group applies the reflectors to Z in the
// same instructions that update H (n <= LPM/2), doubling the matrices per CU
// compared with one group of n lanes doing both.
template <int NB, int LPM, bool ZH>
__global__ __launch_bounds__(64) void k_zgeev_schur(int n, int batch, const zd *__restrict__ A, zd *__restrict__ eig,
                                                    zd *__restrict__ schur, int8_t *__restrict__ flag, int itmax)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM;
    __shared__ zd lds[GPW * G::PER];
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    zd *H = lds + g * G::PER, *Z = H + G::MAT, *taus = Z + G::MAT;
    const FullLayout lay{n | 1};
    const long long nn = (long long)n * n;
    const int info = schur_in_lds<NB, LPM, ZH>(H, Z, taus, lay, A + mat * nn, n, t, itmax);
    zd *out = eig + mat * n;
    zd *T = schur + mat * 2 * nn, *U = T + nn;
    if (info < 0) {
        fill_nan<LPM>(out, n, t);
        fill_nan<LPM>(T, (int)(2 * nn), t);
        if (t == 0) flag[mat] = 2;
        return;
    }
    for (int r = t; r < n; r += LPM) out[r] = H[lay(r, r)];
    store_full<LPM>(H, lay, T, n, t);
    store_full<LPM>(Z, lay, U, n, t);
    if (t == 0) flag[mat] = info != 0 ? 1 : 0;
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

// Eigenvalues and right and/or left eigenvectors (LAPACK ZGEEV with
// JOBVR/JOBVL = 'V' where VR/VL is non-null, no balancing): the Schur form
// of the forward kernel, ZTREVC on T, back-transform by U, ZGEEV
// normalisation.  LDS per matrix: T (+ the triangular eigenvector factor in
// its strictly lower part) and U.
template <int NB, int LPM, bool ZH>
__global__ __launch_bounds__(64) void k_zgeev_vectors(int n, int batch, const zd *__restrict__ A, zd *__restrict__ eig,
                                                      zd *__restrict__ VL, zd *__restrict__ VR,
                                                      int8_t *__restrict__ flag, int itmax)
{
    using G = AdjGeom<NB>;
    static_assert(NB <= LPM, "one eigenvector component per lane");
    constexpr int GPW = 64 / LPM;
    __shared__ zd lds[GPW * G::PER];
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    zd *B1 = lds + g * G::PER, *U = B1 + G::MAT, *taus = U + G::MAT;
    const FullLayout lay{n | 1};
    const long long nn = (long long)n * n;
    const int info = schur_in_lds<NB, LPM, ZH>(B1, U, taus, lay, A + mat * nn, n, t, itmax);
    zd *out = eig + mat * n;
    if (info < 0) fill_nan<LPM>(out, n, t);
    else
        for (int r = t; r < n; r += LPM) out[r] = B1[lay(r, r)];
    if (t == 0) flag[mat] = info < 0 ? 2 : (info != 0 ? 1 : 0);
    if (info != 0) {
        if (VR) fill_nan<LPM>(VR + mat * nn, (int)nn, t);
        if (VL) fill_nan<LPM>(VL + mat * nn, (int)nn, t);
        return;
    }
    const bool live = t < n;
    const int r = t;
    if (VR) {
        // v_i = U x_i = U(:, i) + sum_{c < i} U(:, c) x_i(c); x_i(c) at B1(i, c)
        trevc_right_lower_t<LPM, 1>(B1, lay, n, t);
        zd *dst = VR + mat * nn;
        for (int i = 0; i < n; ++i) {
            zd v = mk(0.0, 0.0);
            if (live) {
                for (int c = 0; c < i; ++c) v = zadd(v, zmul(U[lay(r, c)], B1[lay(i, c)]));
                v = zadd(v, U[lay(r, i)]);
            }
            v = zgeev_normalise<LPM>(v, live, r);
            if (live) dst[r + (long long)i * n] = v;
        }
        lds_order();
    }
    if (VL) {
        // u_i = U y_i = U(:, i) + sum_{c > i} U(:, c) y_i(c); y_i(c) at B1(c, i)
        trevc_left_lower<LPM, 1>(B1, lay, n, t);
        zd *dst = VL + mat * nn;
        for (int i = 0; i < n; ++i) {
            zd v = mk(0.0, 0.0);
            if (live) {
                v = U[lay(r, i)];
                for (int c = i + 1; c < n; ++c) v = zadd(v, zmul(U[lay(r, c)], B1[lay(c, i)]));
            }
            v = zgeev_normalise<LPM>(v, live, r);
            if (live) dst[r + (long long)i * n] = v;
        }
    }
}

template <int NB, int LPM, bool RECOMPUTE>
__global__ __launch_bounds__(64) void k_zgeev_adjoint(int n, int batch, const zd *__restrict__ A,
                                                      const zd *__restrict__ schur, const zd *__restrict__ gbar,
                                                      zd *dA, int8_t *__restrict__ flag, int itmax)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM, S = (NB + LPM - 1) / LPM;
    __shared__ zd lds[GPW * G::PER];
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    zd *B1 = lds + g * G::PER, *B2 = B1 + G::MAT, *taus = B2 + G::MAT, *gv = taus + NB;
    const FullLayout lay{n | 1};
    const long long nn = (long long)n * n;
    zd *dst = dA + mat * nn;
    const zd *Ug;   // global copy of U
    int f;
    if constexpr (RECOMPUTE) {
        const int info = schur_in_lds<NB, LPM>(B1, B2, taus, lay, A + mat * nn, n, t, itmax);
        f = info < 0 ? 2 : (info != 0 ? 1 : 0);
        if (t == 0) flag[mat] = (int8_t)f;
        if (f == 0) {
            store_full<LPM>(B2, lay, dst, n, t);   // park U in dA
            __threadfence();
        }
        Ug = dst;
    } else {
        f = flag[mat];
        if (f == 0) load_full<LPM>(B1, lay, schur + mat * 2 * nn, n, t);
        Ug = schur + mat * 2 * nn + nn;
    }
    if (f != 0) {
        fill_nan<LPM>(dst, (int)nn, t);
        return;
    }
    for (int i = t; i < n; i += LPM) gv[i] = gbar[mat * n + i];
    const double ulp = DBL_EPSILON, smlnum = DBL_MIN * ((double)n / ulp);
    // ---- X (buf2, upper, unit diagonal): (T - t_ii) x_i = 0, rows r < i
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

#define TKZ_ADJ_NB(X) X(16, 16) X(32, 32) X(64, 64)

static int adj_nb(int n) { return n <= 16 ? 16 : (n <= 32 ? 32 : 64); }

// Schur forward configurations (NB, lanes per matrix, Z on the upper half-wave)
#define TKZ_SCHUR_CFG(X) X(16, 32, true) X(32, 64, true) X(64, 64, false)

static hipError_t launch_schur(int n, int batch, const zd *A, zd *eig, zd *schur, int8_t *flag, int itmax,
                               hipStream_t s)
{
    switch (adj_nb(n)) {
#define X(nb, lpm, zh)                                                                                       \
    case nb:                                                                                                 \
        hipLaunchKernelGGL((k_zgeev_schur<nb, lpm, zh>), dim3((batch + 64 / lpm - 1) / (64 / lpm)), dim3(64), 0, \
                           s, n, batch, A, eig, schur, flag, itmax);                                         \
        return hipGetLastError();
        TKZ_SCHUR_CFG(X)
#undef X
    }
    return hipErrorInvalidValue;
}

static hipError_t launch_vectors(int n, int batch, const zd *A, zd *eig, zd *VL, zd *VR, int8_t *flag, int itmax,
                                 hipStream_t s)
{
    switch (adj_nb(n)) {
#define X(nb, lpm, zh)                                                                                          \
    case nb:                                                                                                    \
        hipLaunchKernelGGL((k_zgeev_vectors<nb, lpm, zh>), dim3((batch + 64 / lpm - 1) / (64 / lpm)), dim3(64), 0, \
                           s, n, batch, A, eig, VL, VR, flag, itmax);                                           \
        return hipGetLastError();
        TKZ_SCHUR_CFG(X)
#undef X
    }
    return hipErrorInvalidValue;
}

static hipError_t launch_adjoint(int n, int batch, const zd *A, const zd *schur, const zd *gbar, zd *dA,
                                 int8_t *flag, int itmax, hipStream_t s)
{
    switch (adj_nb(n)) {
#define X(nb, lpm)                                                                                               \
    case nb:                                                                                                     \
        if (schur)                                                                                               \
            hipLaunchKernelGGL((k_zgeev_adjoint<nb, lpm, false>), dim3((batch + 64 / lpm - 1) / (64 / lpm)),      \
                               dim3(64), 0, s, n, batch, A, schur, gbar, dA, flag, itmax);                       \
        else                                                                                                     \
            hipLaunchKernelGGL((k_zgeev_adjoint<nb, lpm, true>), dim3((batch + 64 / lpm - 1) / (64 / lpm)),       \
                               dim3(64), 0, s, n, batch, A, schur, gbar, dA, flag, itmax);                       \
        return hipGetLastError();
        TKZ_ADJ_NB(X)
#undef X
    }
    return hipErrorInvalidValue;
}

}  // namespace tkz

using namespace tkz;

extern "C" size_t tokura_get_zgeev_schur_worksize(int n, int batch_count)
{
    if (!(0 < n && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return (size_t)-1;
    if (!(0 < batch_count)) return (size_t)-2;
    return sizeof(hipDoubleComplex) * 2 * (size_t)n * (size_t)n * (size_t)batch_count;
}

extern "C" int tokura_zgeev_batched_schur_gpu(tokurablas_t *handle, int n, int batch_count,
                                              const hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                                              hipDoubleComplex *schur, char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (schur == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    hipError_t e = launch_schur(n, batch_count, (const zd *)A, (zd *)eigenvalues, (zd *)schur, (int8_t *)flag,
                                handle->max_sweeps, stream);
    return e == hipSuccess ? 0 : 2;
}

extern "C" int tokura_zgeev_batched_backward_gpu(tokurablas_t *handle, int n, int batch_count,
                                                 const hipDoubleComplex *A, const hipDoubleComplex *schur,
                                                 const hipDoubleComplex *gbar, hipDoubleComplex *dA, char *flag,
                                                 hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr && schur == nullptr) return -3;
    if (gbar == nullptr) return -4;
    if (dA == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    hipError_t e = launch_adjoint(n, batch_count, (const zd *)A, (const zd *)schur, (const zd *)gbar, (zd *)dA,
                                  (int8_t *)flag, handle->max_sweeps, stream);
    return e == hipSuccess ? 0 : 2;
}

extern "C" int tokura_zgeev_batched_vectors_gpu(tokurablas_t *handle, int n, int batch_count,
                                                const hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                                                hipDoubleComplex *VL, hipDoubleComplex *VR, char *flag,
                                                hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (VL == nullptr && VR == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    hipError_t e = launch_vectors(n, batch_count, (const zd *)A, (zd *)eigenvalues, (zd *)VL, (zd *)VR,
                                  (int8_t *)flag, handle->max_sweeps, stream);
    return e == hipSuccess ? 0 : 2;
}
