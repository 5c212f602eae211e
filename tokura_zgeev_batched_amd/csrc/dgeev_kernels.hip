// Real matrices: batched eigenvalues of real n x n matrices (n <= 64), the
// analogue of LAPACK DGEEV with JOBVL = JOBVR = 'N' and no balancing.  No
// counterpart in the reference, which solves complex matrices only.
//
// Two launches, as the complex split path:
//  1. k_dhess (lane groups, n <= 32) loads each matrix into LDS (full
//              layout), checks it is finite, reduces it to upper Hessenberg
//              form (DGEHD2) and writes the packed Hessenberg matrix (column
//              c: rows 0..c+1) to `work`; for n > 32 k_dhess_reg does the
//              same with the matrix in the registers of a 256-thread
//              workgroup (dgeev_hess_reg.hpp);
//  2. k_dqr    loads the packed matrix into a BulgeLayout LDS buffer (two
//              spare rows per column for the double-shift bulge) and runs
//              DLAHQR (Francis double shift, DLANV2 for converged 2x2
//              blocks), writing wr / wi and the flag.
// Lanes per matrix = NB (16, 32 or 64): 4, 2 or 1 matrices per wavefront.
#include "tokura_internal.hpp"
#include "dgeev_device.hpp"
#include "dgeev_hess_reg.hpp"

namespace tkz {

template <int NB>
__global__ __launch_bounds__(64) void k_dhess(int n, int batch, const double *__restrict__ A,
                                              double *__restrict__ work, int8_t *__restrict__ flag)
{
    constexpr int LPM = NB, GPW = 64 / LPM, LDP = NB | 1;
    __shared__ double lds[GPW * NB * LDP];
    const int lane = (int)threadIdx.x, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    const FullLayout lay{n | 1};
    double *H = lds + g * NB * LDP;
    const int nn = n * n;
    const double *src = A + mat * nn;
    bool bad = false;
    for (int e = t; e < nn; e += LPM) {
        const double v = src[e];
        bad |= !isfinite(v);
        H[lay(e % n, e / n)] = v;
    }
    const bool ok = !group_any<LPM>(bad);
    lds_order();
    if (ok) dhessenberg<LPM>(H, lay, n, t);
    double *dst = work + mat * (long long)packed_size(n);
    const PackedLayout pk;
    for (int c = 0; c < n; ++c) {
        const int rows = c + 2 < n ? c + 2 : n;
        for (int r = t; r < rows; r += LPM) dst[pk(r, c)] = H[lay(r, c)];
    }
    if (t == 0) flag[mat] = ok ? 0 : 2;
}

// Stage 1 for n > 32 with the matrix in registers (dhess_reg): one
// 256-thread workgroup per matrix, 16 (NBLK - 1) < n <= 16 NBLK.
template <int NBLK, bool OFF>
__global__ __launch_bounds__(256) void k_dhess_reg(int n, int /*batch: one workgroup per matrix*/,
                                                   const double *__restrict__ A, double *__restrict__ work,
                                                   int8_t *__restrict__ flag)
{
    __shared__ DHregSmem S;
    const long long mat = blockIdx.x;
    dhess_reg<NBLK, OFF>(S, A + mat * (long long)n * n, n, work + mat * (long long)packed_size(n), flag + mat);
}

template <int NB>
__global__ __launch_bounds__(64) void k_dqr(int n, int batch, const double *__restrict__ work,
                                            double *__restrict__ wr, double *__restrict__ wi,
                                            int8_t *__restrict__ flag, int itmax)
{
    constexpr int LPM = NB, GPW = 64 / LPM, PER = bulge_size(NB);
    __shared__ double lds[GPW * PER];
    const int lane = (int)threadIdx.x, g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    double *outr = wr + mat * n, *outi = wi + mat * n;
    if (flag[mat] == 2) {
        for (int r = t; r < n; r += LPM) {
            outr[r] = __builtin_nan("");
            outi[r] = __builtin_nan("");
        }
        return;
    }
    double *H = lds + g * PER;
    const BulgeLayout bl;
    const PackedLayout pk;
    const double *src = work + mat * (long long)packed_size(n);
    for (int c = 0; c < n; ++c) {
        const int rows = c + 2 < n ? c + 2 : n;
        for (int r = t; r < c + 4; r += LPM) H[bl(r, c)] = r < rows ? src[pk(r, c)] : 0.0;
    }
    lds_order();
    const int info = dlahqr_group<LPM>(H, bl, n, t, itmax, outr, outi);
    if (t == 0) flag[mat] = info != 0 ? 1 : 0;
}

#define TKZ_DGEEV_NB(X) X(16) X(32) X(64)

static int dgeev_nb(int n) { return n <= 16 ? 16 : (n <= 32 ? 32 : 64); }

// The register Hessenberg for 32 < n <= 64 (the offset placement below the
// block multiples 48 and 64).
static void launch_dhess_reg(int n, int batch, const double *A, double *work, int8_t *flag, hipStream_t s)
{
    const dim3 grid(batch), block(256);
    if (n < 48) hipLaunchKernelGGL((k_dhess_reg<3, true>), grid, block, 0, s, n, batch, A, work, flag);
    else if (n == 48) hipLaunchKernelGGL((k_dhess_reg<3, false>), grid, block, 0, s, n, batch, A, work, flag);
    else if (n < 64) hipLaunchKernelGGL((k_dhess_reg<4, true>), grid, block, 0, s, n, batch, A, work, flag);
    else hipLaunchKernelGGL((k_dhess_reg<4, false>), grid, block, 0, s, n, batch, A, work, flag);
}

static hipError_t launch_dgeev(int n, int batch, const double *A, double *wr, double *wi, double *work,
                               int8_t *flag, int itmax, hipStream_t s, int *stage)
{
    switch (dgeev_nb(n)) {
#define X(nb)                                                                                              \
    case nb: {                                                                                             \
        const dim3 grid((batch + 64 / nb - 1) / (64 / nb));                                                \
        *stage = 1;                                                                                        \
        if (nb == 64)                                                                                      \
            launch_dhess_reg(n, batch, A, work, flag, s);                                                  \
        else                                                                                               \
            hipLaunchKernelGGL((k_dhess<nb>), grid, dim3(64), 0, s, n, batch, A, work, flag);              \
        hipError_t e = hipGetLastError();                                                                  \
        if (e != hipSuccess) return e;                                                                     \
        *stage = 2;                                                                                        \
        hipLaunchKernelGGL((k_dqr<nb>), grid, dim3(64), 0, s, n, batch, work, wr, wi, flag, itmax);        \
        return hipGetLastError();                                                                          \
    }
        TKZ_DGEEV_NB(X)
#undef X
    }
    return hipErrorInvalidValue;
}

}  // namespace tkz

using namespace tkz;

extern "C" size_t tokura_get_dgeev_worksize(int n, int batch_count)
{
    if (!(0 < n && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return (size_t)-1;
    if (!(0 < batch_count)) return (size_t)-2;
    return sizeof(double) * (size_t)packed_size(n) * (size_t)batch_count;
}

extern "C" int tokura_dgeev_batched_gpu(tokurablas_t *handle, int n, int batch_count, const double *A, double *wr,
                                        double *wi, double *work, char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr) return -3;
    if (wr == nullptr || wi == nullptr) return -4;
    if (work == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    int stage = 1;
    const hipError_t e = launch_dgeev(n, batch_count, A, wr, wi, work, (int8_t *)flag, handle->max_sweeps, stream,
                                      &stage);
    return e == hipSuccess ? 0 : stage;
}
