// Differentiation of the eigenvalues: forward with Schur vectors and the
// adjoint (backward) kernel; eigenvectors (ZGEEV with JOBVL/JOBVR = 'V') from
// the same Schur form.  No counterpart in the reference, which has no
// gradient path (README.md:41-44: "No eigenvectors are computed").
//
// Forward (k_schur_hq + k_schur_qr): A = U T U^H by the same method as the
// eigenvalue path, in Schur-form mode: Hessenberg with kept reflectors, Q
// accumulation (ZUNGHR), QR with WANTT/WANTZ (ZLAHQR), T and U stored per
// matrix.
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
#include "zgeev_schur.hpp"

namespace tkz {

// Eigenvalues and right and/or left eigenvectors (LAPACK ZGEEV with
// JOBVR/JOBVL = 'V' where VR/VL is non-null, no balancing): the Schur form
// of the forward kernel, ZTREVC on T, back-transform by U, ZGEEV
// normalisation.  LDS per matrix: T (+ the triangular eigenvector factor in
// its strictly lower part) and U.  GZ: U is accumulated in the home of VR
// (VL when only left vectors are asked for) and the LDS holds only T.  The
// vectors are then formed in place: VL ascending (column i of VL needs U's
// columns >= i), VR descending (columns <= i), VL first when both are asked
// for.
template <int NB, int LPM, bool ZH, bool GZ>
__global__ __launch_bounds__(64) void k_zgeev_vectors(int n, int batch, const zd *__restrict__ A, zd *__restrict__ eig,
                                                      zd *VL, zd *VR, int8_t *__restrict__ flag, int itmax)
{
    using G = AdjGeom<NB>;
    static_assert(NB <= LPM, "one eigenvector component per lane");
    constexpr int GPW = 64 / LPM, PER = GZ ? G::PER_GZ : G::PER;
    __shared__ zd lds[GPW * PER];
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    const long long nn = (long long)n * n;
    zd *B1 = lds + g * PER, *taus = B1 + (GZ ? 1 : 2) * G::MAT;
    zd *U = GZ ? (VR ? VR : VL) + mat * nn : B1 + G::MAT;
    const FullLayout lay{n | 1};
    const FullLayout ul{GZ ? n : lay.ld};
    const int info = schur_in_lds<NB, LPM, ZH, GZ>(B1, U, taus, lay, A + mat * nn, n, t, itmax);
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
    vectors_from_schur<LPM, GZ>(B1, lay, U, ul, n, t, VL ? VL + mat * nn : nullptr, VR ? VR + mat * nn : nullptr);
}

// Eigenvectors for n > 32 without a workspace, one matrix per workgroup of NT
// threads: schur_wg with U accumulated in the home of VR (VL when only left
// vectors are asked for) and T in LDS, then vectors_from_schur by the first
// wavefront (the arithmetic of the two-launch path and k_vectors_from_schur).
template <int NT>
__global__ __launch_bounds__(NT) void k_zgeev_vectors_wg(int n, int batch, const zd *__restrict__ A,
                                                        zd *__restrict__ eig, zd *VL, zd *VR,
                                                        int8_t *__restrict__ flag, int itmax)
{
    __shared__ zd lds[64 * 65 + 64];
    __shared__ double red[8];
    __shared__ int ish;
    const long long mat = blockIdx.x, nn = (long long)n * n;
    const int tid = (int)threadIdx.x;
    zd *B1 = lds, *taus = lds + 64 * 65;
    zd *U = (VR ? VR : VL) + mat * nn;
    const FullLayout lay{n | 1};
    const int info = schur_wg<NT>(B1, U, n, taus, red, &ish, A + mat * nn, n, itmax);
    zd *out = eig + mat * n;
    if (info < 0) fill_nan<NT>(out, n, tid);
    else
        for (int r = tid; r < n; r += NT) out[r] = B1[lay(r, r)];
    if (tid == 0) flag[mat] = info < 0 ? 2 : (info != 0 ? 1 : 0);
    if (info != 0) {
        if (VR) fill_nan<NT>(VR + mat * nn, (int)nn, tid);
        if (VL) fill_nan<NT>(VL + mat * nn, (int)nn, tid);
        return;
    }
    if (tid < 64) vectors_from_schur<64, true>(B1, lay, U, FullLayout{n}, n, tid, VL ? VL + mat * nn : nullptr,
                                               VR ? VR + mat * nn : nullptr);
}

// Schur form in two launches (schur_hq_one, schur_qr_one), LPM lanes per
// matrix.  The matrix's T and U slots are Tout + mat*tstride and
// Uout + mat*ustride: the Schur-form output (T then U per matrix) or, for
// eigenvectors, VL's and VR's slots.  First half: full LDS layout; writes
// flag 0, or 2 with NaN outputs for a non-finite input.
template <int NB, int LPM>
__global__ __launch_bounds__(64) void k_schur_hq(int n, int batch, const zd *__restrict__ A, zd *__restrict__ eig,
                                                 zd *Tout, long long tstride, zd *Uout, long long ustride,
                                                 int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM, PER = G::MAT + NB;
    __shared__ zd lds[GPW * PER];
    const int lane = (int)threadIdx.x, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g, nn = (long long)n * n;
    if (mat >= batch) return;
    zd *Tm = Tout + mat * tstride, *Um = Uout + mat * ustride, *H = lds + g * PER;
    if (schur_hq_one<LPM>(H, H + G::MAT, A + mat * nn, n, t, Tm, Um)) {
        if (t == 0) flag[mat] = 0;
    } else {
        fill_nan<LPM>(eig + mat * n, n, t);
        fill_nan<LPM>(Tm, (int)nn, t);
        fill_nan<LPM>(Um, (int)nn, t);
        if (t == 0) flag[mat] = 2;
    }
}

// First half for n > 32 by a workgroup of NT threads per matrix
// (schur_hq_wg): the 256-thread Hessenberg of the eigenvalue path with kept
// reflectors, then Q, LDS sized for NB.
template <int NB, int NT>
__global__ __launch_bounds__(NT) void k_schur_hq_wg(int n, int batch, const zd *__restrict__ A, zd *__restrict__ eig,
                                                    zd *Tout, long long tstride, zd *Uout, long long ustride,
                                                    int8_t *__restrict__ flag)
{
    __shared__ zd lds[NB * (NB | 1) + NB];
    __shared__ double red[8];
    const long long mat = blockIdx.x, nn = (long long)n * n;
    const int tid = (int)threadIdx.x;
    zd *Tm = Tout + mat * tstride, *Um = Uout + mat * ustride;
    if (schur_hq_wg<NT, NT / 64>(lds, lds + NB * (NB | 1), red, A + mat * nn, n, Tm, Um)) {
        if (tid == 0) flag[mat] = 0;
    } else {
        fill_nan<NT>(eig + mat * n, n, tid);
        fill_nan<NT>(Tm, (int)nn, tid);
        fill_nan<NT>(Um, (int)nn, tid);
        if (tid == 0) flag[mat] = 2;
    }
}

// Second half: the QR on the packed layout (four matrices per CU at n = 64);
// eigenvalues, T, U in place, flag 0/1 (matrices flagged 2 are skipped).
// SNAP: shifts snapped to the matrix's known eigenvalues, `known` + mat * n.
template <int NB, int LPM, bool SNAP = false>
__global__ __launch_bounds__(64) void k_schur_qr(int n, int batch, zd *__restrict__ eig, zd *Tout, long long tstride,
                                                 zd *Uout, long long ustride, int8_t *flag, int itmax,
                                                 const zd *__restrict__ known = nullptr)
{
    constexpr int GPW = 64 / LPM, PER = packed_size(NB) + 1;   // + the idle lanes' element
    __shared__ zd lds[GPW * PER];
    const int lane = (int)threadIdx.x, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch || flag[mat] == 2) return;
    const int info = schur_qr_one<LPM, (LPM == 64), SNAP>(lds + g * PER, n, t, Tout + mat * tstride,
                                                          Uout + mat * ustride, eig + mat * n, itmax,
                                                          SNAP ? known + mat * n : nullptr);
    if (t == 0) flag[mat] = info != 0 ? 1 : 0;
}

// Eigenvectors after k_schur_hq / k_schur_qr left T in its slot (VL's when
// both sides are asked for, else the caller's workspace) and U in VR's (VL's
// when VR is NULL): T into LDS (each lane its own rows, so the later writes
// of VL by the same lanes need no fence), then vectors_from_schur in place.
// Flagged matrices get NaN vectors.
template <int NB, int LPM>
__global__ __launch_bounds__(64) void k_vectors_from_schur(int n, int batch, const zd *Tslot, zd *VL, zd *VR,
                                                           const int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM;
    __shared__ zd lds[GPW * G::MAT];
    const int lane = (int)threadIdx.x, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g, nn = (long long)n * n;
    if (mat >= batch) return;
    zd *VLm = VL ? VL + mat * nn : nullptr, *VRm = VR ? VR + mat * nn : nullptr, *B1 = lds + g * G::MAT;
    if (flag[mat] != 0) {
        if (VLm) fill_nan<LPM>(VLm, (int)nn, t);
        if (VRm) fill_nan<LPM>(VRm, (int)nn, t);
        return;
    }
    const FullLayout lay{n | 1};
    const zd *Tm = Tslot + mat * nn;
    for (int r = t; r < n; r += LPM)
        for (int c = 0; c < n; ++c) B1[lay(r, c)] = Tm[r + c * n];
    lds_order();
    vectors_from_schur<LPM, true>(B1, lay, VRm ? VRm : VLm, FullLayout{n}, n, t, VLm, VRm);
}

// Stored-Schur backward for n <= 16: LPM lanes per matrix, 64/LPM matrices
// per wavefront.
template <int NB, int LPM>
__global__ __launch_bounds__(64) void k_zgeev_adjoint(int n, int batch, const zd *__restrict__ schur,
                                                      const zd *__restrict__ gbar, zd *dA,
                                                      const int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM;
    __shared__ zd lds[GPW * G::PER];
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    zd *B1 = lds + g * G::PER, *B2 = B1 + G::MAT, *gv = B2 + G::MAT + NB;
    const long long nn = (long long)n * n;
    const zd *Tm = schur + mat * 2 * nn;
    adjoint_one<NB, LPM>(B1, B2, gv, n, t, Tm, Tm + nn, gbar + mat * n, nullptr, dA + mat * nn, flag[mat]);
}

// Stored-Schur backward for n > 16: one matrix per workgroup of
// TKZ_ADJ_WG_NT threads, LDS for NB (at NB = 64 the two buffers allow one
// matrix per CU, so the extra wavefronts share its work instead of idling the
// CU; NB = 32 fits four workgroups).
template <int NT, int NB>
__global__ __launch_bounds__(NT) void k_zgeev_adjoint_wg(int n, const zd *__restrict__ schur,
                                                         const zd *__restrict__ gbar, zd *__restrict__ dA,
                                                         const int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    __shared__ zd lds[G::PER];
    const long long mat = blockIdx.x, nn = (long long)n * n;
    zd *B1 = lds, *B2 = lds + G::MAT, *gv = B2 + G::MAT;
    const zd *Tm = schur + mat * 2 * nn;
    (void)adjoint_wg<NT>(B1, B2, gv, n, Tm, Tm + nn, gbar + mat * n, dA + mat * nn, flag[mat]);
}

// Recompute backward, second launch: the Schur form of the first launch has
// T in Tslot and U in dA (n x n each per matrix) and its eigenvalues in eo;
// the forward's eigenvalues ef (any order) are paired with T's diagonal
// (match_eigenvalues), so cotangent i of the forward drives eigenvalue perm^-1
// of the Schur form; dA overwrites U in place.  Flag: the forward's 0/1/2,
// or 3 when the eigenvalues do not match (then dA = NaN).
template <int NB, int LPM>
__global__ __launch_bounds__(64) void k_adjoint_matched(int n, int batch, const zd *__restrict__ Tslot,
                                                        const zd *__restrict__ ef, const zd *__restrict__ gbar,
                                                        zd *dA, int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM, S = (NB + LPM - 1) / LPM;
    __shared__ zd lds[GPW * G::PER];
    __shared__ int perm_s[GPW * NB];
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    zd *B1 = lds + g * G::PER, *B2 = B1 + G::MAT, *efl = B2 + G::MAT, *gv = efl + NB;
    int *perm = perm_s + g * NB;
    const long long nn = (long long)n * n;
    const zd *Tm = Tslot + mat * nn;
    zd *Um = dA + mat * nn;
    const FullLayout lay{n | 1};
    int f = flag[mat];
    if (f == 0) {
        load_full<LPM>(B1, lay, Tm, n, t);
        for (int i = t; i < n; i += LPM) efl[i] = ef[mat * n + i];
        lds_order();
        if (!match_eigenvalues<LPM, S>(B1, lay, efl, perm, n, t)) {
            f = 3;
            if (t == 0) flag[mat] = 3;
        }
    }
    adjoint_one<NB, LPM>(B1, B2, gv, n, t, Tm, Um, gbar + mat * n, perm, Um, f, true);
}

template <int NT, int NB>
__global__ __launch_bounds__(NT) void k_adjoint_matched_wg(int n, const zd *__restrict__ Tslot,
                                                           const zd *__restrict__ ef, const zd *__restrict__ gbar,
                                                           zd *dA, int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    __shared__ zd lds[G::PER];
    __shared__ int perm[NB];
    __shared__ int ish;
    const long long mat = blockIdx.x, nn = (long long)n * n;
    zd *B1 = lds, *B2 = lds + G::MAT, *efl = B2 + G::MAT, *gv = efl + NB;
    zd *Um = dA + mat * nn;
    const int f = adjoint_wg<NT>(B1, B2, gv, n, Tslot + mat * nn, Um, gbar + mat * n, Um, flag[mat], ef + mat * n,
                                 efl, perm, &ish);
    if (threadIdx.x == 0 && f == 3) flag[mat] = 3;
}

static int adj_nb(int n) { return n <= 16 ? 16 : (n <= 32 ? 32 : 64); }

// One-kernel Schur configurations (NB, lanes per matrix, Z on the upper
// half-wave, Z in global memory), used by one-sided eigenvector requests
// (there is no second slot to park T in); everything else goes through
// k_schur_hq / k_schur_qr with NB lanes per matrix.
#define TKZ_SCHUR_CFG(X) X(16, 32, true, false) X(32, 64, true, false)   // n > 32: k_zgeev_vectors_wg
// (NB, lanes per matrix) of the two-launch Schur form: for n > 32 the LDS is
// sized in steps (the size classes of the ragged paths), so that smaller
// matrices keep more wavefronts resident in the QR (packed n = 64: 34.3 KB,
// four per CU; n = 56: six; n = 48: eight)
#define TKZ_SPLIT_NB(X) X(16, 16) X(32, 32) X(40, 64) X(48, 64) X(56, 64) X(60, 64) X(64, 64)

static int split_nb(int n)
{
    return n <= 16 ? 16 : (n <= 32 ? 32 : (n <= 40 ? 40 : (n <= 48 ? 48 : (n <= 56 ? 56 : (n <= 60 ? 60 : 64)))));
}

// Tslot receives the Hessenberg matrix, then T; Uslot receives Q, then U.
// With VEC, k_vectors_from_schur follows (eigenvectors into VLv / VRv).
template <int NB, int LPM>
static hipError_t launch_schur_split_nb(int n, int batch, const zd *A, zd *eig, zd *Tout, long long tstride,
                                        zd *Uout, long long ustride, int8_t *flag, int itmax, hipStream_t s, bool vec,
                                        zd *VLv, zd *VRv, const zd *known)
{
    const dim3 grid((batch + 64 / LPM - 1) / (64 / LPM));
    if constexpr (NB > 32)
        hipLaunchKernelGGL((k_schur_hq_wg<NB, TKZ_ADJ_WG_NT>), dim3(batch), dim3(TKZ_ADJ_WG_NT), 0, s, n, batch, A, eig,
                           Tout, tstride, Uout, ustride, flag);
    else
        hipLaunchKernelGGL((k_schur_hq<NB, LPM>), grid, dim3(64), 0, s, n, batch, A, eig, Tout, tstride, Uout, ustride,
                           flag);
    if (known)
        hipLaunchKernelGGL((k_schur_qr<NB, LPM, true>), grid, dim3(64), 0, s, n, batch, eig, Tout, tstride, Uout,
                           ustride, flag, itmax, known);
    else
        hipLaunchKernelGGL((k_schur_qr<NB, LPM>), grid, dim3(64), 0, s, n, batch, eig, Tout, tstride, Uout, ustride,
                           flag, itmax, nullptr);
    if (vec)
        hipLaunchKernelGGL((k_vectors_from_schur<NB, LPM>), grid, dim3(64), 0, s, n, batch, Tout, VLv, VRv, flag);
    return hipGetLastError();
}

static hipError_t launch_schur_split(int n, int batch, const zd *A, zd *eig, zd *Tout, long long tstride,
                                     zd *Uout, long long ustride, int8_t *flag, int itmax, hipStream_t s,
                                     bool vec = false, zd *VLv = nullptr, zd *VRv = nullptr,
                                     const zd *known = nullptr)
{
    switch (split_nb(n)) {
#define X(nb, lpm)                                                                                             \
    case nb:                                                                                                   \
        return launch_schur_split_nb<nb, lpm>(n, batch, A, eig, Tout, tstride, Uout, ustride, flag, itmax, s, vec, \
                                              VLv, VRv, known);
        TKZ_SPLIT_NB(X)
#undef X
    }
    return hipErrorInvalidValue;
}

static hipError_t launch_schur(int n, int batch, const zd *A, zd *eig, zd *schur, int8_t *flag, int itmax,
                               hipStream_t s)
{
    const long long nn = (long long)n * n;
    return launch_schur_split(n, batch, A, eig, schur, 2 * nn, schur + nn, 2 * nn, flag, itmax, s);
}

// Two sides: T parks in VL's slot.  One side: T parks in `work` when given,
// else the one-kernel form runs (T in LDS).
static hipError_t launch_vectors(int n, int batch, const zd *A, zd *eig, zd *VL, zd *VR, zd *work, int8_t *flag,
                                 int itmax, hipStream_t s)
{
    const long long nn = (long long)n * n;
    zd *Tslot = (VL && VR) ? VL : work;
    if (Tslot) return launch_schur_split(n, batch, A, eig, Tslot, nn, VR ? VR : VL, nn, flag, itmax, s, true, VL, VR);
    if (n > 32) {
        hipLaunchKernelGGL((k_zgeev_vectors_wg<TKZ_ADJ_WG_NT>), dim3(batch), dim3(TKZ_ADJ_WG_NT), 0, s, n, batch, A, eig,
                           VL, VR, flag, itmax);
        return hipGetLastError();
    }
    switch (adj_nb(n)) {
#define X(nb, lpm, zh, gz)                                                                                        \
    case nb:                                                                                                      \
        hipLaunchKernelGGL((k_zgeev_vectors<nb, lpm, zh, gz>), dim3((batch + 64 / lpm - 1) / (64 / lpm)), dim3(64), 0, \
                           s, n, batch, A, eig, VL, VR, flag, itmax);                                             \
        return hipGetLastError();
        TKZ_SCHUR_CFG(X)
#undef X
    }
    return hipErrorInvalidValue;
}

// n > 16: one 256-thread workgroup per matrix, LDS sized in steps (two per CU
// for n <= 48, three for n <= 40, four for n <= 32)
static int adj_wg_nb(int n) { return n <= 32 ? 32 : (n <= 40 ? 40 : (n <= 48 ? 48 : 64)); }

static hipError_t launch_adjoint(int n, int batch, const zd *schur, const zd *gbar, zd *dA, const int8_t *flag,
                                 hipStream_t s)
{
    if (n > 16) {
        switch (adj_wg_nb(n)) {
#define X(b)                                                                                                   \
    case b:                                                                                                    \
        hipLaunchKernelGGL((k_zgeev_adjoint_wg<TKZ_ADJ_WG_NT, b>), dim3(batch), dim3(TKZ_ADJ_WG_NT), 0, s, n, schur, \
                           gbar, dA, flag);                                                                    \
        break;
            X(32) X(40) X(48) X(64)
#undef X
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_zgeev_adjoint<16, 16>), dim3((batch + 3) / 4), dim3(64), 0, s, n, batch, schur, gbar, dA,
                       flag);
    return hipGetLastError();
}

// Recompute backward: the two-launch Schur forward with T parked in `work`
// (n^2 complex per matrix, then the recomputed eigenvalues, n per matrix:
// the reference's workspace size) and U in dA, its QR shifts snapped to the
// forward's eigenvalues (about half the sweeps), then the matched adjoint.
static hipError_t launch_adjoint_recompute(int n, int batch, const zd *A, const zd *ef, const zd *gbar, zd *dA,
                                           zd *work, int8_t *flag, int itmax, hipStream_t s)
{
    const long long nn = (long long)n * n;
    zd *eo = work + nn * batch;
    hipError_t e = launch_schur_split(n, batch, A, eo, work, nn, dA, nn, flag, itmax, s, false, nullptr, nullptr, ef);
    if (e != hipSuccess) return e;
    if (n > 16) {
        switch (adj_wg_nb(n)) {
#define X(b)                                                                                                    \
    case b:                                                                                                     \
        hipLaunchKernelGGL((k_adjoint_matched_wg<TKZ_ADJ_WG_NT, b>), dim3(batch), dim3(TKZ_ADJ_WG_NT), 0, s, n, work, \
                           ef, gbar, dA, flag);                                                                 \
        break;
            X(32) X(40) X(48) X(64)
#undef X
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_adjoint_matched<16, 16>), dim3((batch + 3) / 4), dim3(64), 0, s, n, batch, work, ef, gbar,
                       dA, flag);
    return hipGetLastError();
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
    (void)A;   // the stored Schur form alone defines the backward
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (schur == nullptr) return -3;
    if (gbar == nullptr) return -4;
    if (dA == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    hipError_t e = launch_adjoint(n, batch_count, (const zd *)schur, (const zd *)gbar, (zd *)dA,
                                  (const int8_t *)flag, stream);
    return e == hipSuccess ? 0 : 2;
}

extern "C" int tokura_zgeev_batched_backward_recompute_gpu(tokurablas_t *handle, int n, int batch_count,
                                                           const hipDoubleComplex *A,
                                                           const hipDoubleComplex *eigenvalues,
                                                           const hipDoubleComplex *gbar, hipDoubleComplex *dA,
                                                           hipDoubleComplex *work, char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr || gbar == nullptr) return -4;
    if (dA == nullptr || work == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    hipError_t e = launch_adjoint_recompute(n, batch_count, (const zd *)A, (const zd *)eigenvalues, (const zd *)gbar,
                                            (zd *)dA, (zd *)work, (int8_t *)flag, handle->max_sweeps, stream);
    return e == hipSuccess ? 0 : 2;
}

extern "C" size_t tokura_get_zgeev_vectors_worksize(int n, int batch_count)
{
    if (!(0 < n && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return (size_t)-1;
    if (!(0 < batch_count)) return (size_t)-2;
    return sizeof(hipDoubleComplex) * (size_t)n * (size_t)n * (size_t)batch_count;
}

extern "C" int tokura_zgeev_batched_vectors_gpu(tokurablas_t *handle, int n, int batch_count,
                                                const hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                                                hipDoubleComplex *VL, hipDoubleComplex *VR, hipDoubleComplex *work,
                                                char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (VL == nullptr && VR == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    hipError_t e = launch_vectors(n, batch_count, (const zd *)A, (zd *)eigenvalues, (zd *)VL, (zd *)VR, (zd *)work,
                                  (int8_t *)flag, handle->max_sweeps, stream);
    return e == hipSuccess ? 0 : 2;
}
