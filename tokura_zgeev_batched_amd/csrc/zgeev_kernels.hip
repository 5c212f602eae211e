// HIP kernels of the same-size batched path (tokura_zgeev_batched_gpu) and
// their launchers.
//
// Replaces the reference's three stages (Hessenberg kernel
// tokura_zgehrd_MWB_kernel.cuh:159-287, QR kernel
// tokura_zhseqr_MWB_kernel.cuh:611-791, and the missing layout transposes
// declared at tokura_blas.h:142,148) with kernels that read the caller's
// column-major, matrix-contiguous layout directly: a group of LPM lanes of
// one wavefront streams its matrix into LDS with coalesced 16-byte loads, so
// no interleaved copy of the batch is ever made.
//
// Two kernel families:
//  * fused  — load, Hessenberg, QR and the eigenvalue store in one kernel,
//             matrix in LDS with leading dimension n|1 (NB = LDS capacity).
//  * split  — k_hessenberg_packed reduces in a full LDS copy and writes the
//             upper Hessenberg matrix in packed form (column c: rows 0..c+1,
//             (n-1)(n+2)/2+n elements) to the caller's workspace; k_qr_packed
//             runs QR on the packed copy, which needs about half the LDS, so
//             twice as many matrices are resident per CU at n = 64.
#include <atomic>
#include "zgeev_device.hpp"
#include "zgeev_hess_reg.hpp"
#include "zgeev_launch.hpp"
#include "zgeev_pool.hpp"

namespace tkz {

template <int NB, int LPM>
struct Geometry {
    static constexpr int GPW = 64 / LPM;          // matrices per wavefront
    static constexpr int LDP = NB | 1;             // padded leading dimension
    static constexpr int S = (NB + LPM - 1) / LPM; // positions per lane
};

template <int NB, int LPM>
__global__ __launch_bounds__(64) void k_zgeev_fused(int n, int batch, const zd *__restrict__ A,
                                                    zd *__restrict__ eig, int8_t *__restrict__ flag,
                                                    int itmax, int bal)
{
    using G = Geometry<NB, LPM>;
    __shared__ zd lds[G::GPW * NB * G::LDP];
    const int lane = threadIdx.x & 63;
    const int g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * G::GPW + g;
    if (mat >= batch) return;
    zd *H = lds + g * NB * G::LDP;
    const FullLayout lay{n | 1};
    double mx;
    const bool ok = load_full<LPM>(H, lay, A + mat * (long long)n * n, n, t, &mx);
    lds_order();
    zd *out = eig + mat * n;
    if (!ok) {
        for (int r = t; r < n; r += LPM) out[r] = mk(__builtin_nan(""), __builtin_nan(""));
        if (t == 0) flag[mat] = 2;
        return;
    }
    const int se = zgeev_scale_exp(mx);
    if (se) {
        scale_full<LPM>(H, lay, n, t, se);
        lds_order();
    }
    if (bal) zgebal_group<LPM, G::S>(H, lay, n, t, bal);
    hessenberg<LPM>(H, lay, n, t);
    int info = 0;
    if (n > 1) info = zlahqr_eigs<LPM, G::S>(H, lay, n, t, itmax);
    const double us = ldexp(1.0, -se);
    for (int r = t; r < n; r += LPM) out[r] = se ? zscale(us, H[lay(r, r)]) : H[lay(r, r)];
    if (t == 0) flag[mat] = info != 0 ? 1 : 0;
}

// Stage 1 of the split path: Hessenberg in LDS, packed store to `work`.
// The norm-range factor 2^-e by which the QR multiplies the eigenvalues is
// passed to it in the matrix's first eigenvalue slot, which the QR reads
// before it writes the eigenvalues.
template <int NB, int LPM>
__global__ __launch_bounds__(64) void k_hessenberg_packed(int n, int batch, const zd *__restrict__ A,
                                                          zd *__restrict__ work, int8_t *__restrict__ flag,
                                                          int bal, zd *__restrict__ eig)
{
    using G = Geometry<NB, LPM>;
    __shared__ zd lds[G::GPW * NB * G::LDP];
    const int lane = threadIdx.x & 63;
    const int g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * G::GPW + g;
    if (mat >= batch) return;
    zd *H = lds + g * NB * G::LDP;
    const FullLayout lay{n | 1};
    double mx;
    const bool ok = load_full<LPM>(H, lay, A + mat * (long long)n * n, n, t, &mx);
    lds_order();
    const int se = ok ? zgeev_scale_exp(mx) : 0;
    if (se) {
        scale_full<LPM>(H, lay, n, t, se);
        lds_order();
    }
    if (ok && t == 0) eig[mat * n] = mk(ldexp(1.0, -se), 0.0);
    if (ok && bal) zgebal_group<LPM, G::S>(H, lay, n, t, bal);
    if (ok) hessenberg<LPM>(H, lay, n, t);
    store_packed<LPM>(H, lay, work + mat * (long long)packed_size(n), n, t);
    if (t == 0) flag[mat] = ok ? 0 : 2;
}

// Workgroup size and team size (threads sharing a row or column) of the
// large-matrix Hessenberg kernel.
#ifndef TKZ_HESS_NT
#define TKZ_HESS_NT 256
#endif
#ifndef TKZ_HESS_TPR
#define TKZ_HESS_TPR (TKZ_HESS_NT / 64)
#endif

// Stage 1 of the split path with a whole workgroup of NT threads per matrix
// (teams of NT/64 threads share each row/column): at NB = 64 two workgroups
// of 256 threads fit per CU, two waves per SIMD.
// qctr (may be null): the QR's work counter, zeroed here for the pool
// variants' QR kernel that follows on the same stream.
// The norm-range factor goes to the matrix's first eigenvalue slot, as in
// k_hessenberg_packed.
template <int NB, int NT>
__global__ __launch_bounds__(NT) void k_hessenberg_packed_wg(int n, int batch, const zd *__restrict__ A,
                                                             zd *__restrict__ work, int8_t *__restrict__ flag,
                                                             int bal, zd *__restrict__ eig, int *qctr = nullptr)
{
    constexpr int LDP = NB | 1;
    if (qctr && blockIdx.x == 0 && threadIdx.x == 0) *qctr = 0;
    __shared__ zd lds[NB * LDP];
    __shared__ double red[2 * NT / 64];
    const long long mat = blockIdx.x;
    const int tid = threadIdx.x;
    const FullLayout lay{n | 1};
    const long long nn = (long long)n * n;
    const zd *src = A + mat * nn;
    bool bad = false;
    double mx = 0.0;
    for (int e = tid; e < nn; e += NT) {
        const zd v = src[e];
        bad |= !zfinite(v);
        mx = fmax(mx, fmax(fabs(v.re), fabs(v.im)));
        lds[lay(e % n, e / n)] = v;
    }
    mx = group_max<64>(mx);
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    const int ok = __syncthreads_or(bad) == 0;
    for (int w = 0; w < NT / 64; ++w) mx = fmax(mx, red[w]);
    const int se = ok ? zgeev_scale_exp(mx) : 0;
    if (se) {
        const double f = ldexp(1.0, se);
        for (int e = tid; e < nn; e += NT) lds[lay(e % n, e / n)] = zscale(f, lds[lay(e % n, e / n)]);
    }
    __syncthreads();   // red is reused by the reduction below; the scaled matrix is complete
    if (ok && tid == 0) eig[mat * n] = mk(ldexp(1.0, -se), 0.0);
    if (ok && bal) {                 // balancing by the first wavefront
        if (tid < 64) zgebal_group<64, (NB + 63) / 64>(lds, lay, n, tid, bal);
        __syncthreads();
    }
    if (ok) hessenberg_wg<NT, TKZ_HESS_TPR>(lds, lay, n, red);
    zd *dst = work + mat * (long long)packed_size(n);
    const PackedLayout pk;
    for (int c = 0; c < n; ++c) {
        const int rows = (c + 2 < n ? c + 2 : n);
        for (int r = tid; r < rows; r += NT) dst[pk(r, c)] = lds[lay(r, c)];
    }
    if (tid == 0) flag[mat] = ok ? 0 : 2;
}

// Stage 1 with the matrix in registers (hess_reg, n <= 16 NBLK, no
// balancing): 256 threads per matrix and 6 KiB of LDS, so the register file
// (about 128 VGPRs per thread, four workgroups per CU) bounds the occupancy.
template <int NBLK, bool OFF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kHregWaves<NBLK>, kHregWaves<NBLK>))) void k_hessenberg_reg(int n, int batch, const zd *__restrict__ A,
                                                        zd *__restrict__ work, int8_t *__restrict__ flag,
                                                        zd *__restrict__ eig, int *qctr = nullptr)
{
    if (qctr && blockIdx.x == 0 && threadIdx.x == 0) *qctr = 0;
    __shared__ HregSmem S;
    const long long mat = blockIdx.x;
    hess_reg<NBLK, OFF>(S, A + mat * (long long)n * n, n, work + mat * (long long)packed_size(n), flag + mat,
                        eig + mat * n);
}

// Stage 2 of the split path: QR on the packed Hessenberg matrix (DUAL: two
// bulges per sweep, LPM == 64, or LPM == 32 with n <= 32).
template <int NB, int LPM, bool DUAL = false>
__global__ __launch_bounds__(64) void k_qr_packed(int n, int batch, const zd *__restrict__ work,
                                                  zd *__restrict__ eig, int8_t *__restrict__ flag,
                                                  int itmax)
{
    constexpr int GPW = 64 / LPM, S = (NB + LPM - 1) / LPM, PS = packed_size(NB);
    __shared__ zd lds[GPW * PS + (DUAL && LPM == 64)];   // + the two-bulge QR's idle-lane element
    const int lane = threadIdx.x & 63;
    const int g = lane / LPM, t = lane % LPM;
    const long long mat = (long long)blockIdx.x * GPW + g;
    if (mat >= batch) return;
    zd *out = eig + mat * n;
    if (flag[mat] == 2) {
        for (int r = t; r < n; r += LPM) out[r] = mk(__builtin_nan(""), __builtin_nan(""));
        return;
    }
    zd *H = lds + g * PS;
    const int ps = packed_size(n);
    const zd *src = work + mat * (long long)ps;
    const double us = out[0].re;   // the Hessenberg kernel's norm-range factor
    for (int e = t; e < ps; e += LPM) H[e] = src[e];
    lds_order();
    const PackedLayout pk;
    int info = 0;
    if constexpr (DUAL) {
        static_assert(LPM == 64 || (LPM == 32 && NB <= 32), "two-bulge QR: a lane per position");
        if (n > 1) info = zlahqr_dual<LPM, false, (LPM == 64)>(H, pk, n, t, itmax);
    } else {
        if (n > 1) info = zlahqr_eigs<LPM, S>(H, pk, n, t, itmax);
    }
    for (int r = t; r < n; r += LPM) out[r] = us != 1.0 ? zscale(us, H[pk(r, r)]) : H[pk(r, r)];
    if (t == 0) flag[mat] = info != 0 ? 1 : 0;
}

// Stage 2 of the pool variants: one persistent workgroup of W wavefronts per
// CU whose LDS is one arena of slots (zgeev_pool.hpp).  Each wavefront
// claims the next matrix from the counter *qctr (zeroed by the Hessenberg
// kernel), waits for a slot that holds it and runs the two-bulge QR, moving
// to smaller slots as the matrix deflates; eigenvalues are those of
// k_qr_packed<NB, 64, true> bit for bit.  Every wavefront leaves the loop
// when the counter passes the batch.
template <int W>
__global__ __launch_bounds__(64 * W) void k_qr_pool(int n, int batch, const zd *__restrict__ work,
                                                    zd *__restrict__ eig, int8_t *__restrict__ flag, int itmax,
                                                    int *qctr, PoolCfg P)
{
    __shared__ zd arena[kArena];
    __shared__ unsigned free_mask;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) free_mask = (1u << P.nslots) - 1u;
    __syncthreads();
    const int ps = packed_size(n);
    const unsigned fit = pool_fit(P, ps + 1);
    if (fit == 0) return;   // the host builds layouts whose first slot holds the matrix
    for (;;) {
        // lane 0 adds 1, the others 0 (see zgeev_pool.hpp on the control flow)
        const int j = __builtin_amdgcn_readfirstlane(atomicAdd(qctr, lane == 0 ? 1 : 0));
        if (j >= batch) break;
        zd *out = eig + (long long)j * n;
        if (__builtin_amdgcn_readfirstlane(flag[j]) == 2) {
            if (lane < n) out[lane] = mk(__builtin_nan(""), __builtin_nan(""));
            continue;
        }
        const double us = out[0].re;   // the Hessenberg kernel's norm-range factor
        const int s = pool_acquire(&free_mask, fit, lane);
        const int info = pool_qr_one(P, &free_mask, arena, s, work + (long long)j * ps, n, out, us, itmax, lane);
        flag[j] = info != 0 ? 1 : 0;   // every lane stores the same byte
    }
}

// Stage 2 of the chunk variants: one persistent workgroup of W wavefronts per
// CU whose LDS is a chunked arena (ChunkPool).  Each wavefront claims the
// next matrix from *qctr (zeroed by the Hessenberg kernel), waits for the
// lowest run of chunks that holds it and runs the two-bulge QR; after every
// deflation the matrix moves to the lowest run that holds its active part
// (ChunkMover<0>), so the arena stays compacted and more matrices are
// resident than in fixed slots.  Eigenvalues are those of
// k_qr_packed<NB, 64, true> bit for bit.  A wavefront waits only while it
// holds no chunks; every wavefront leaves the loop when the counter passes
// the batch.
template <int W>
__global__ __launch_bounds__(64 * W) void k_qr_chunk(int n, int batch, const zd *__restrict__ work,
                                                     zd *__restrict__ eig, int8_t *__restrict__ flag, int itmax,
                                                     int *qctr)
{
    __shared__ zd arena[kChunks * kChunk];
    __shared__ unsigned long long bits[kBitWords];
    __shared__ unsigned lock, freed;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < kBitWords) bits[threadIdx.x] = run_mask((int)threadIdx.x, 0, kChunks);
    if (threadIdx.x == 0) lock = freed = 0u;
    __syncthreads();
    const ChunkPool P{bits, &lock, &freed, arena, lane};
    const int ps = packed_size(n);
    const int k0 = chunks_for(ps + 1);   // + the idle lanes' element
    for (;;) {
        // lane 0 adds 1, the others 0 (see zgeev_pool.hpp on the control flow)
        const int j = __builtin_amdgcn_readfirstlane(atomicAdd(qctr, lane == 0 ? 1 : 0));
        if (j >= batch) break;
        zd *out = eig + (long long)j * n;
        if (__builtin_amdgcn_readfirstlane(flag[j]) == 2) {
            if (lane < n) out[lane] = mk(__builtin_nan(""), __builtin_nan(""));
            continue;
        }
        const double us = out[0].re;   // the Hessenberg kernel's norm-range factor
        const int s0 = P.alloc_wait(k0);
        zd *H = arena + s0 * kChunk;
        const zd *src = work + (long long)j * ps;
        for (int e = lane; e < ps; e += 64) H[e] = src[e];
        lds_order();
        ChunkMover<0> mv{&P, out, s0, k0, n, H, us};
        int info = 0;
        if (n > 1)
            info = zlahqr_dual<64, false, true, ChunkMover<0>>(H, PackedLayout{}, n, lane, itmax, nullptr, 0, &mv,
                                                              k0 * kChunk - 1);
        const PackedLayout pk;
        if (lane < mv.whi) out[lane] = mv.eig(mv.H[pk(lane, lane)]);   // whi <= n <= 64
        P.give(mv.s, mv.k);
        flag[j] = info != 0 ? 1 : 0;   // every lane stores the same byte
    }
}

static const PoolLayout g_pool_layouts[] = {
#define X(id, w, a0, a1, a2, a3, a4, a5) {w, {a0, a1, a2, a3, a4, a5}},
    TKZ_POOL_LAYOUTS(X)
#undef X
};

static int device_cus()
{
    // per device, filled on first use (threads that race store the same value)
    static std::atomic<int> cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int c = cus[dev].load(std::memory_order_relaxed);
    if (c == 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1) c = 256;
        cus[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

// The 256-thread Hessenberg with LDS for the smallest of 40, 48, 56, 60, 64
// that holds n (the arithmetic does not depend on the LDS size).
static void launch_hess_wg(int n, int batch, const zd *A, zd *work, int8_t *flag, int bal, zd *eig, int *qctr,
                           hipStream_t s)
{
    if (bal == 0 && n > 32) {   // the register-resident reduction
        if (n < 48)
            hipLaunchKernelGGL((k_hessenberg_reg<3, true>), dim3(batch), dim3(256), 0, s, n, batch, A, work, flag, eig, qctr);
        else if (n == 48)
            hipLaunchKernelGGL((k_hessenberg_reg<3, false>), dim3(batch), dim3(256), 0, s, n, batch, A, work, flag, eig, qctr);
        else if (n < 64)
            hipLaunchKernelGGL((k_hessenberg_reg<4, true>), dim3(batch), dim3(256), 0, s, n, batch, A, work, flag, eig, qctr);
        else
            hipLaunchKernelGGL((k_hessenberg_reg<4, false>), dim3(batch), dim3(256), 0, s, n, batch, A, work, flag, eig, qctr);
        return;
    }
    const int nb = n <= 40 ? 40 : (n <= 48 ? 48 : (n <= 56 ? 56 : (n <= 60 ? 60 : 64)));
    switch (nb) {
#define X(b)                                                                                                    \
    case b:                                                                                                     \
        hipLaunchKernelGGL((k_hessenberg_packed_wg<b, TKZ_HESS_NT>), dim3(batch), dim3(TKZ_HESS_NT), 0, s, n, batch, \
                           A, work, flag, bal, eig, qctr);                                                      \
        break;
        X(40) X(48) X(56) X(60) X(64)
#undef X
    }
}

// Pool variant: the 256-thread Hessenberg, then k_qr_pool with layout `pl`
// on one workgroup per CU.  The QR's counter sits in `work` after the packed
// matrices (inside the reference's workspace size, which leaves at least
// 16 (n^2 + n - packed_size(n)) bytes per matrix there).
static hipError_t launch_pool(int pl, int n, int batch, const zd *A, zd *eig, int8_t *flag, zd *work, int itmax,
                              int bal, hipStream_t s, int *stage, hipEvent_t mid)
{
    PoolCfg P;
    if (!build_pool_cfg(n, g_pool_layouts[pl], P) || P.cap[0] < packed_size(n) + 1) return hipErrorInvalidValue;
    int *qctr = (int *)(work + (long long)packed_size(n) * batch);
    *stage = 1;
    launch_hess_wg(n, batch, A, work, flag, bal, eig, qctr, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (mid) (void)hipEventRecord(mid, s);   // stage-timing marker: a failed record only spoils that reading
    *stage = 2;
    const dim3 grid(device_cus());
    switch (g_pool_layouts[pl].waves) {
#define X(w)                                                                                                     \
    case w:                                                                                                      \
        hipLaunchKernelGGL((k_qr_pool<w>), grid, dim3(64 * w), 0, s, n, batch, work, eig, flag, itmax, qctr, P); \
        break;
        X(8) X(12) X(16)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Chunk variant: the 256-thread Hessenberg, then k_qr_chunk<w> on one
// workgroup per CU; the QR's counter sits after the packed matrices, as for
// the pool variants.
static hipError_t launch_chunk(int w, int n, int batch, const zd *A, zd *eig, int8_t *flag, zd *work, int itmax,
                               int bal, hipStream_t s, int *stage, hipEvent_t mid)
{
    int *qctr = (int *)(work + (long long)packed_size(n) * batch);
    *stage = 1;
    launch_hess_wg(n, batch, A, work, flag, bal, eig, qctr, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (mid) (void)hipEventRecord(mid, s);   // stage-timing marker: a failed record only spoils that reading
    *stage = 2;
    const dim3 grid(device_cus());
    switch (w) {
#define X(wv)                                                                                                  \
    case wv:                                                                                                   \
        hipLaunchKernelGGL((k_qr_chunk<wv>), grid, dim3(64 * wv), 0, s, n, batch, work, eig, flag, itmax, qctr); \
        break;
        X(8) X(10) X(12)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Variant table and launcher
// ---------------------------------------------------------------------------
#define TKZ_VARIANTS(X)                 \
    X(0, FUSED, 4, 4)                   \
    X(1, FUSED, 8, 4)                   \
    X(2, FUSED, 8, 8)                   \
    X(3, FUSED, 16, 4)                  \
    X(4, FUSED, 16, 8)                  \
    X(5, FUSED, 16, 16)                 \
    X(6, FUSED, 32, 8)                  \
    X(7, FUSED, 32, 16)                 \
    X(8, FUSED, 32, 32)                 \
    X(9, FUSED, 48, 32)                 \
    X(10, FUSED, 48, 64)                \
    X(11, FUSED, 64, 32)                \
    X(12, FUSED, 64, 64)                \
    X(13, SPLIT, 32, 16)                \
    X(14, SPLIT, 32, 32)                \
    X(15, SPLIT, 48, 32)                \
    X(16, SPLIT, 48, 64)                \
    X(17, SPLIT, 64, 32)                \
    X(18, SPLIT, 64, 64)                \
    X(19, DUAL, 48, 64)                 \
    X(20, DUAL, 64, 64)                 \
    X(21, DUAL, 32, 32)                 \
    X(22, DUAL_1W, 64, 64)              \
    X(23, DUAL, 56, 64)                 \
    X(24, DUAL, 40, 64)                 \
    X(25, DUAL, 60, 64)                 \
    X(26, SPLIT, 16, 16)                \
    X(27, SPLIT, 16, 8)                 \
    X(28, SPLIT, 8, 4)                  \
    X(29, POOL, 64, 64)                 \
    X(30, POOL, 64, 64)                 \
    X(31, POOL, 64, 64)                 \
    X(32, POOL, 64, 64)                 \
    X(33, POOL, 64, 64)                 \
    X(34, POOL, 64, 64)                 \
    X(35, POOL, 64, 64)                 \
    X(36, POOL, 64, 64)                 \
    X(37, CHUNK, 64, 64)                \
    X(38, CHUNK, 64, 64)                \
    X(39, CHUNK, 64, 64)
// pool variants: layout id - kFirstPool of TKZ_POOL_LAYOUTS
constexpr int kFirstPool = 29;

static const VariantInfo g_variants[] = {
#define X(id, kind, nb, lpm) {id, VK_##kind, nb, lpm, 64 / lpm},
    TKZ_VARIANTS(X)
#undef X
};

int num_variants() { return (int)(sizeof(g_variants) / sizeof(g_variants[0])); }
const VariantInfo *variant_info(int v) { return (v >= 0 && v < num_variants()) ? &g_variants[v] : nullptr; }

template <int NB, int LPM>
static hipError_t launch_fused(int n, int batch, const zd *A, zd *eig, int8_t *flag, int itmax, int bal,
                               hipStream_t s)
{
    constexpr int GPW = 64 / LPM;
    const int blocks = (batch + GPW - 1) / GPW;
    hipLaunchKernelGGL((k_zgeev_fused<NB, LPM>), dim3(blocks), dim3(64), 0, s, n, batch, A, eig, flag, itmax, bal);
    return hipGetLastError();
}

template <int NB, int LPM, bool DUAL = false, bool ONEWAVE = false>
static hipError_t launch_split(int n, int batch, const zd *A, zd *eig, int8_t *flag, zd *work, int itmax,
                               int bal, hipStream_t s, int *stage, hipEvent_t mid)
{
    constexpr int GPW = 64 / LPM;
    const int blocks = (batch + GPW - 1) / GPW;
    *stage = 1;
    if constexpr (NB >= 48 && !ONEWAVE)   // large matrices: one workgroup per matrix
        launch_hess_wg(n, batch, A, work, flag, bal, eig, nullptr, s);
    else
        hipLaunchKernelGGL((k_hessenberg_packed<NB, LPM>), dim3(blocks), dim3(64), 0, s, n, batch, A, work, flag, bal,
                           eig);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (mid) (void)hipEventRecord(mid, s);   // stage-timing marker: a failed record only spoils that reading
    *stage = 2;
    hipLaunchKernelGGL((k_qr_packed<NB, LPM, DUAL>), dim3(blocks), dim3(64), 0, s, n, batch, work, eig, flag,
                       itmax);
    return hipGetLastError();
}

template <int KIND, int NB, int LPM>
static hipError_t launch_kind(int id, int n, int batch, const zd *A, zd *eig, int8_t *flag, zd *work, int itmax,
                              int bal, hipStream_t s, int *stage, hipEvent_t mid)
{
    if constexpr (KIND == VK_POOL)
        return launch_pool(id - kFirstPool, n, batch, A, eig, flag, work, itmax, bal, s, stage, mid);
    else if constexpr (KIND == VK_CHUNK)   // variant 37: 8 wavefronts per CU, 38: 12, 39: 10
        return launch_chunk(id == 37 ? 8 : (id == 38 ? 12 : 10), n, batch, A, eig, flag, work, itmax, bal, s, stage,
                            mid);
    else if constexpr (KIND == VK_FUSED)
        return launch_fused<NB, LPM>(n, batch, A, eig, flag, itmax, bal, s);
    else
        return launch_split<NB, LPM, KIND == VK_DUAL || KIND == VK_DUAL_1W, KIND == VK_DUAL_1W>(
            n, batch, A, eig, flag, work, itmax, bal, s, stage, mid);
}

hipError_t launch_plain(int variant, int n, int batch, const void *A, void *eig, char *flag, void *work,
                        int itmax, int bal, hipStream_t s, int *stage, hipEvent_t mid)
{
    const zd *a = (const zd *)A;
    zd *w = (zd *)eig;
    int8_t *f = (int8_t *)flag;
    zd *wk = (zd *)work;
    *stage = 1;
    switch (variant) {
#define X(id, kind, nb, lpm) \
    case id: return launch_kind<VK_##kind, nb, lpm>(id, n, batch, a, w, f, wk, itmax, bal, s, stage, mid);
        TKZ_VARIANTS(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace tkz
