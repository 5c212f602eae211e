// Ragged batches: n varies per matrix (1..64).  No counterpart in the
// reference, which requires one n per call (README.md:5, tokura_blas.h:12).
//
// Launches of one eigenvalue call (tokura_zgeev_ragged_graph captures them
// together):
//  1. k_ragged_reset     checks that `work` holds this handle's schedule for
//                        this batch and zeroes the queue head (a fill kernel,
//                        graph-safe).
//  2. Householder Hessenberg of every matrix, packed result to `work`:
//     k_ragged_hess_reg  one 256-thread workgroup per matrix of n > 32 with
//                        the matrix in registers (two launches: n 49..64 and
//                        n 33..48); with balancing on, k_ragged_hess_wg, the
//                        LDS-resident reduction, one launch per LDS size
//                        (n <= 64, 60, 56, 48, 40),
//     k_ragged_hess      one wavefront per tile of small matrices (n <= 32).
//  3. k_ragged_qr_pool   ONE persistent QR kernel for the whole batch: a
//                        workgroup of W wavefronts per CU whose LDS is one
//                        arena of 32-element chunks (ChunkPool).  Each
//                        wavefront pulls the next tile from one atomic queue
//                        (tiles in decreasing size, longest first), waits for
//                        a run of chunks that holds it and runs its QR; a
//                        matrix of n > 32 (one per wavefront, two bulges per
//                        sweep) moves to a smaller run as it deflates, so
//                        matrices of all sizes share each CU's LDS and no
//                        size class waits for another to drain.
// A tile is the work of one wavefront: one matrix of n > 32, or 64/LPM
// matrices of a small class, B (17..32) or C (1..16), whose lanes per matrix
// form, with W, the "class configuration" that the autotuner picks.
//
// `work` layout (built on the host by tokura_zgeev_ragged_prepare):
//   RaggedHeader | tiles int4[ntiles] | order int[B] | n int[B] |
//   moff int64[B] | eoff int64[B] | hoff int64[B] | packed H (complex)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "tokura_internal.hpp"
#include "zgeev_hess_reg.hpp"
#include "zgeev_pool.hpp"
#include "zgeev_schur.hpp"

namespace tkz {

// Size classes, largest n first (the order of order[], tiles and launches).
// Classes 0..4 (n > 32) run one matrix per wavefront with LDS for cls_hi(c).
constexpr int kNumCls = 7, kClsB = 5, kClsC = 6;
__host__ __device__ constexpr int cls_hi(int c)
{
    return c == 0 ? 64 : (c == 1 ? 60 : (c == 2 ? 56 : (c == 3 ? 48 : (c == 4 ? 40 : (c == 5 ? 32 : 16)))));
}
__host__ __device__ constexpr int cls_of(int n)
{
    return n > 60 ? 0 : (n > 56 ? 1 : (n > 48 ? 2 : (n > 40 ? 3 : (n > 32 ? 4 : (n > 16 ? 5 : 6)))));
}

constexpr int kRaggedMagic = 0x52474b54;   // "TKGR"

struct RaggedHeader {
    int qctr[8];       // queue heads: [0] the QR of the eigenvalue path, one per size class for the Schur QR
    int cls_end[8];    // class c's matrices: order[cls_end[c-1] .. cls_end[c]) (order[0 ..) for c = 0)
    int tile_cls[8];   // class c's tiles: [tile_cls[c], tile_cls[c+1]); tile_cls[kNumCls] = ntiles
    int ntiles;
    int batch;
    int cfg;
    int tile_small0;   // first tile of classes B/C
    int magic;         // kRaggedMagic
    int gen;           // the handle's prepare count when this schedule was written
    long long off_tiles, off_order, off_n, off_moff, off_eoff, off_hoff, off_H;
    long long total_bytes;
};
static_assert(sizeof(RaggedHeader) == 184, "header size");

// class configurations: lanes per matrix of the classes n > 48, 33..48 (one
// matrix per wavefront), B (17..32), C (1..16); wavefronts per workgroup of
// the persistent QR; the deflating matrices' mover (ChunkMover): 0 compacts
// the arena at every deflation, 4..7 moves a matrix to the highest free run
// when its active part needs at most mv/8 of its run; cap: when nonzero, only
// the first `cap` wavefronts take tiles of n 49..64 from the front of the
// queue while the front is there, the others take tiles from its back (at
// n = 64 eight wavefronts per CU are faster than twelve, for the smaller
// sizes twelve are)
#define TKZ_RAGGED_CFGS(X)          \
    X(0, 64, 64, 16, 8, 12, 0, 0)   \
    X(1, 64, 64, 32, 8, 12, 0, 0)   \
    X(2, 64, 64, 32, 16, 12, 0, 0)  \
    X(3, 64, 64, 32, 32, 12, 0, 0)  \
    X(4, 64, 64, 32, 16, 8, 0, 0)   \
    X(5, 64, 64, 32, 8, 8, 5, 0)    \
    X(6, 64, 64, 32, 8, 8, 0, 0)    \
    X(7, 64, 64, 32, 16, 8, 6, 0)   \
    X(8, 64, 64, 32, 16, 12, 5, 0)  \
    X(9, 64, 64, 32, 32, 8, 5, 0)   \
    X(10, 64, 64, 32, 32, 12, 6, 0) \
    X(11, 64, 64, 32, 8, 8, 6, 0)   \
    X(12, 64, 64, 32, 8, 12, 5, 0)  \
    X(13, 64, 64, 32, 32, 12, 0, 8) \
    X(14, 64, 64, 32, 16, 12, 0, 8) \
    X(15, 64, 64, 32, 8, 12, 0, 8)
static const int g_cfg_lpm[][7] = {
#define X(id, a1, a2, b, c, w, mv, cap) {a1, a2, b, c, w, mv, cap},
    TKZ_RAGGED_CFGS(X)
#undef X
};
static constexpr int kNumRaggedCfg = 16;

static int cfg_lpm(int cfg, int cls) { return cls < kClsB ? 64 : g_cfg_lpm[cfg][cls == kClsB ? 2 : 3]; }

struct RaggedView {
    RaggedHeader *h;
    const int4 *tiles;
    const int *order, *nsz;
    const long long *moff, *eoff, *hoff;
    zd *H;
    __device__ RaggedView(char *w)
    {
        h = (RaggedHeader *)w;
        tiles = (const int4 *)(w + h->off_tiles);
        order = (const int *)(w + h->off_order);
        nsz = (const int *)(w + h->off_n);
        moff = (const long long *)(w + h->off_moff);
        eoff = (const long long *)(w + h->off_eoff);
        hoff = (const long long *)(w + h->off_hoff);
        H = (zd *)(w + h->off_H);
    }
};

// Zeroes the queue heads when `work` holds the schedule that this handle's
// prepare wrote for this batch (magic, generation and batch agree).
// Otherwise (a buffer freed and reallocated at the same address, or written
// over since) every matrix gets flag 4 and the header is emptied, so that no
// kernel of the call finds work in it.
__global__ void k_ragged_reset(char *work, int gen, int batch, int8_t *flag)
{
    RaggedHeader *h = (RaggedHeader *)work;
    const bool valid = h->magic == kRaggedMagic && h->gen == gen && h->batch == batch;
    if (valid) {
        if (blockIdx.x == 0 && threadIdx.x < 8) h->qctr[threadIdx.x] = 0;
        return;
    }
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < batch; b += gridDim.x * blockDim.x) flag[b] = 4;
    if (blockIdx.x == 0 && threadIdx.x < 8) {
        h->cls_end[threadIdx.x] = 0;
        h->tile_cls[threadIdx.x] = 0;
        if (threadIdx.x == 0) {
            h->ntiles = 0;
            h->tile_small0 = 0;
        }
    }
}

static hipError_t launch_reset(char *work, int gen, int batch, int8_t *flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_ragged_reset, dim3((batch + 255) / 256), dim3(256), 0, s, work, gen, batch, flag);
    return hipGetLastError();
}

// matrices of size class c: order[lo .. hi)
__device__ __forceinline__ void cls_range(const RaggedView &V, int c, int &lo, int &hi)
{
    lo = c == 0 ? 0 : V.h->cls_end[c - 1];
    hi = V.h->cls_end[c];
}

// The norm-range factor of the eigenvalues goes to the matrix's first
// eigenvalue slot, which the QR reads before writing the eigenvalues.
template <int NB, int LPM>
__device__ __forceinline__ void hess_tile(zd *lds, const RaggedView &V, int4 td, const zd *__restrict__ A,
                                          int8_t *__restrict__ flag, int bal, zd *__restrict__ eig)
{
    constexpr int LDP = NB | 1;
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    if (g >= td.y) return;
    const int b = V.order[td.z + g];
    const int n = V.nsz[b];
    zd *H = lds + g * NB * LDP;
    const FullLayout lay{n | 1};
    double mx;
    const bool ok = load_full<LPM>(H, lay, A + V.moff[b], n, t, &mx);
    lds_order();
    const int se = ok ? zgeev_scale_exp(mx) : 0;
    if (se) {
        scale_full<LPM>(H, lay, n, t, se);
        lds_order();
    }
    if (ok && t == 0) eig[V.eoff[b]] = mk(ldexp(1.0, -se), 0.0);
    if (ok && bal) zgebal_group<LPM, (NB + LPM - 1) / LPM>(H, lay, n, t, bal);
    if (ok) hessenberg<LPM>(H, lay, n, t);
    store_packed<LPM>(H, lay, V.H + V.hoff[b], n, t);
    if (t == 0) flag[b] = ok ? 0 : 2;
}

// The QR of the small matrices of one tile (class B or C: 64/LPM matrices
// of n <= NB per wavefront, LPM lanes each) from the LDS run at `base`,
// `stride` elements per matrix; eigenvalues times the Hessenberg's
// norm-range factor.
template <int NB, int LPM>
__device__ __forceinline__ void small_tile(zd *base, int stride, const RaggedView &V, int4 td,
                                           zd *__restrict__ eig, int8_t *__restrict__ flag, int itmax)
{
    constexpr int S = (NB + LPM - 1) / LPM;
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    if (g < td.y) {
        const int b = V.order[td.z + g];
        const int n = V.nsz[b];
        zd *out = eig + V.eoff[b];
        if (flag[b] == 2) {
            for (int r = t; r < n; r += LPM) out[r] = mk(__builtin_nan(""), __builtin_nan(""));
        } else {
            zd *H = base + g * stride;
            const zd *src = V.H + V.hoff[b];
            const int ps = packed_size(n);
            const double us = out[0].re;   // the Hessenberg's norm-range factor
            for (int e = t; e < ps; e += LPM) H[e] = src[e];
            lds_order();
            const PackedLayout pk;
            int info = 0;
            if (n > 1) info = zlahqr_eigs<LPM, S>(H, pk, n, t, itmax);
            for (int r = t; r < n; r += LPM) out[r] = us != 1.0 ? zscale(us, H[pk(r, r)]) : H[pk(r, r)];
            if (t == 0) flag[b] = info != 0 ? 1 : 0;
        }
    }
    lds_order();
}

// Both ends of the tile queue in one 64-bit counter (qctr[0..1]): the low
// half counts tiles taken from the front (largest first), the high half
// tiles taken from the back (smallest first).  A claim adds 1 to its half
// and is valid while front + back < ntiles, so no tile is taken twice.
__device__ __forceinline__ unsigned long long *queue_ends(const RaggedView &V)
{
    return (unsigned long long *)&V.h->qctr[0];
}
__device__ __forceinline__ void claim(const RaggedView &V, bool back, int lane, int nt, int &tile)
{
    const unsigned long long inc = lane == 0 ? (back ? 1ull << 32 : 1ull) : 0ull;
    const unsigned long long o = atomicAdd(queue_ends(V), inc);
    const unsigned f = __builtin_amdgcn_readfirstlane((unsigned)o);
    const unsigned b = __builtin_amdgcn_readfirstlane((unsigned)(o >> 32));
    tile = (long long)f + (long long)b < (long long)nt ? (back ? nt - 1 - (int)b : (int)f) : -1;
}

// The QR of every tile of the batch in one persistent launch: one workgroup
// of W wavefronts per CU whose LDS is a chunked arena (ChunkPool).  A
// wavefront takes the next tile from the front of the queue (tiles in
// decreasing size), and the lowest run of chunks that holds it; while none
// is free it takes tiles from the back of the queue (the smallest) and runs
// those in whatever room there is, so the small matrices fill the gaps that
// the large ones leave.  A tile is one matrix of n > 32 (the two-bulge
// sweeps, moving up to a smaller run as the matrix deflates: ChunkMover) or
// several small matrices.  With CAP > 0, wavefronts CAP.. take tiles from
// the back first while the front tile is of n 49..64 (a peek at the front
// counter), so that at most CAP wavefronts work on those matrices while
// they lead the queue.  A wavefront waits only while it holds no chunks, so
// every wait ends; every wavefront leaves when the two ends of the queue
// meet.
template <int W, int LB, int LC, int MV, int CAP>
__global__ __launch_bounds__(64 * W) void k_ragged_qr_pool(char *work, zd *__restrict__ eig,
                                                           int8_t *__restrict__ flag, int itmax)
{
    __shared__ zd arena[kChunks * kChunk];
    __shared__ unsigned long long bits[kBitWords];
    __shared__ unsigned lock, freed;
    const RaggedView V(work);
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < kBitWords) bits[threadIdx.x] = run_mask((int)threadIdx.x, 0, kChunks);
    if (threadIdx.x == 0) lock = freed = 0u;
    __syncthreads();
    const ChunkPool P{bits, &lock, &freed, arena, lane};
    const int nt = V.h->ntiles;
    // a wavefront past the first CAP takes back tiles while the front is in n 49..64
    const bool back_first = CAP > 0 && __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) >= CAP;
    // tile t: its descriptor (uniform) and the chunks its QR needs
    auto desc = [&](int t) {
        const int4 tv = V.tiles[t];
        return make_int4(__builtin_amdgcn_readfirstlane(tv.x), __builtin_amdgcn_readfirstlane(tv.y),
                         __builtin_amdgcn_readfirstlane(tv.z), __builtin_amdgcn_readfirstlane(tv.w));
    };
    auto need = [&](int4 td) {
        return td.x < kClsB ? chunks_for(packed_size(td.w) + 1) : chunks_for(td.y * packed_size(td.w));
    };
    // the QR of tile td from chunk run s0 (k chunks), which it frees
    auto run = [&](int4 td, int s0, int k) {
        if (td.x < kClsB) {   // one matrix of n > 32
            const int b = __builtin_amdgcn_readfirstlane(V.order[td.z]);
            const int n = __builtin_amdgcn_readfirstlane(V.nsz[b]);
            zd *out = eig + V.eoff[b];
            if (__builtin_amdgcn_readfirstlane(flag[b]) == 2) {
                if (lane < n) out[lane] = mk(__builtin_nan(""), __builtin_nan(""));
                P.give(s0, k);
                return;
            }
            const double us = out[0].re;   // the Hessenberg's norm-range factor
            zd *H = arena + s0 * kChunk;
            const zd *src = V.H + V.hoff[b];
            for (int e = lane; e < packed_size(n); e += 64) H[e] = src[e];
            lds_order();
            ChunkMover<MV> mv{&P, out, s0, k, n, H, us};
            int info = 0;
            if (n > 1)
                info = zlahqr_dual<64, false, true, ChunkMover<MV>>(H, PackedLayout{}, n, lane, itmax, nullptr, 0, &mv,
                                                                k * kChunk - 1);
            const PackedLayout pk;
            if (lane < mv.whi) out[lane] = mv.eig(mv.H[pk(lane, lane)]);   // whi <= n <= 64
            P.give(mv.s, mv.k);
            flag[b] = info != 0 ? 1 : 0;   // every lane stores the same byte
        } else {              // a tile of small matrices
            const int stride = packed_size(td.w);
            if (td.x == kClsB) small_tile<32, LB>(arena + s0 * kChunk, stride, V, td, eig, flag, itmax);
            else small_tile<16, LC>(arena + s0 * kChunk, stride, V, td, eig, flag, itmax);
            P.give(s0, k);
        }
    };
    // pend: the front tile this wavefront has taken but not yet placed
    int pend = -1, pk = 0;
    int4 pd = make_int4(0, 0, 0, 0);
    bool back_open = true;
    for (;;) {
        int4 td = make_int4(0, 0, 0, 0);
        int k = 0, s0 = -1;
        if (back_first && pend < 0 && back_open) {
            const unsigned long long qe = __hip_atomic_load(queue_ends(V), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int f = (int)__builtin_amdgcn_readfirstlane((unsigned)qe);   // tiles taken from the front
            if (f < nt && desc(f).x <= 2) {   // classes 0-2: n 49..64
                int bt = -1;
                claim(V, true, lane, nt, bt);
                if (bt >= 0) {
                    td = desc(bt);
                    k = need(td);
                    s0 = P.alloc_wait(k);
                } else {
                    back_open = false;
                }
            }
        }
        if (s0 < 0) {
            if (pend < 0) {
                claim(V, false, lane, nt, pend);
                if (pend < 0) break;   // the two ends met: no tile is left anywhere
                pd = desc(pend);
                pk = need(pd);
            }
            const unsigned seen = P.frees();
            s0 = P.try_alloc(pk, false);
            if (s0 >= 0) {
                td = pd;
                k = pk;
                pend = -1;
            } else {
                int bt = -1;
                if (back_open) {
                    claim(V, true, lane, nt, bt);
                    back_open = bt >= 0;
                }
                if (bt < 0) {
                    P.wait_free(seen);
                    continue;
                }
                td = desc(bt);
                k = need(td);
                s0 = P.alloc_wait(k);
            }
        }
        run(td, s0, k);   // the only call: one copy of the QR code
    }
}

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Hessenberg of the small classes (B, C): one wavefront per tile.
template <int LB, int LC>
__global__ __launch_bounds__(64) void k_ragged_hess(const zd *__restrict__ A, char *work, int8_t *__restrict__ flag,
                                                    int bal, zd *__restrict__ eig)
{
    constexpr int SZ = cmax((64 / LB) * 32 * 33, (64 / LC) * 16 * 17);
    __shared__ zd lds[SZ];
    const RaggedView V(work);
    const int nt = V.h->ntiles;
    for (int tile = V.h->tile_small0 + blockIdx.x; tile < nt; tile += gridDim.x) {
        const int4 td = V.tiles[tile];
        if (td.x == kClsB) hess_tile<32, LB>(lds, V, td, A, flag, bal, eig);
        else hess_tile<16, LC>(lds, V, td, A, flag, bal, eig);
        lds_order();
    }
}

// Hessenberg of a large class (n > 32): one 256-thread workgroup per
// matrix, quads of threads sharing each row/column, LDS sized for the class.
template <int NB, int CLS>
__global__ __launch_bounds__(256) void k_ragged_hess_wg(const zd *__restrict__ A, char *work,
                                                        int8_t *__restrict__ flag, int bal, zd *__restrict__ eig)
{
    __shared__ zd lds[NB * (NB | 1)];
    __shared__ double red[8];
    const RaggedView V(work);
    int lo, hi;
    cls_range(V, CLS, lo, hi);
    const int tid = threadIdx.x;
    for (int j = lo + (int)blockIdx.x; j < hi; j += gridDim.x) {
        const int b = V.order[j];
        const int n = V.nsz[b];
        const FullLayout lay{n | 1};
        const zd *src = A + V.moff[b];
        bool bad = false;
        double mx = 0.0;
        for (int e = tid; e < n * n; e += 256) {
            const zd v = src[e];
            bad |= !zfinite(v);
            mx = fmax(mx, fmax(fabs(v.re), fabs(v.im)));
            lds[lay(e % n, e / n)] = v;
        }
        mx = group_max<64>(mx);
        if ((tid & 63) == 0) red[tid >> 6] = mx;
        const int ok = __syncthreads_or(bad) == 0;
        for (int w = 0; w < 4; ++w) mx = fmax(mx, red[w]);
        const int se = ok ? zgeev_scale_exp(mx) : 0;
        if (se) {
            const double f = ldexp(1.0, se);
            for (int e = tid; e < n * n; e += 256) lds[lay(e % n, e / n)] = zscale(f, lds[lay(e % n, e / n)]);
        }
        __syncthreads();   // red is reused below; the scaled matrix is complete
        if (ok && tid == 0) eig[V.eoff[b]] = mk(ldexp(1.0, -se), 0.0);
        if (ok && bal) {             // balancing by the first wavefront
            if (tid < 64) zgebal_group<64, 1>(lds, lay, n, tid, bal);
            __syncthreads();
        }
        if (ok) hessenberg_wg<256, 4>(lds, lay, n, red);
        zd *dst = V.H + V.hoff[b];
        const PackedLayout pk;
        for (int c = 0; c < n; ++c) {
            const int rows = (c + 2 < n ? c + 2 : n);
            for (int r = tid; r < rows; r += 256) dst[pk(r, c)] = lds[lay(r, c)];
        }
        if (tid == 0) flag[b] = ok ? 0 : 2;
        __syncthreads();
    }
}

// Hessenberg of the classes c_lo..c_hi (n > 32, n <= 16 NBLK) with the
// matrix in registers (hess_reg; no balancing): one 256-thread workgroup per
// matrix, the grid sized on the host from the schedule; the range itself
// comes from the header (empty for a stale schedule).
template <int NBLK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kHregWaves<NBLK>, kHregWaves<NBLK>))) void k_ragged_hess_reg(
    const zd *__restrict__ A, char *work, int8_t *__restrict__ flag, zd *__restrict__ eig, int c_lo, int c_hi)
{
    __shared__ HregSmem S;
    const RaggedView V(work);
    int lo, hi, lo2, hi2;
    cls_range(V, c_lo, lo, hi);
    cls_range(V, c_hi, lo2, hi2);
    const int j = lo + (int)blockIdx.x;
    if (j >= hi2) return;
    const int b = V.order[j];
    hess_reg<NBLK>(S, A + V.moff[b], V.nsz[b], V.H + V.hoff[b], flag + b, eig + V.eoff[b]);
}

// The classes n > 32: (class, NB)
#define TKZ_BIG_CLS(X) X(0, 64) X(1, 60) X(2, 56) X(3, 48) X(4, 40)

// Branches for launches that touch disjoint matrices (the Hessenberg
// groups of the eigenvalue path, the size classes of the Schur forward and
// of the backward): side streams fork from the caller's stream after the
// schedule check and join it again, so the launches' tails overlap.
// Without side streams everything runs in order on the caller's stream.
struct RaggedFork {
    int nside = 0;
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    bool on() const { return nside > 0; }
    // the stream of branch i (0: the caller's)
    hipStream_t branch(int i, hipStream_t s) const { return (i == 0 || !on()) ? s : side[(i - 1) % nside]; }
    // the side streams wait for what is queued on s so far
    hipError_t begin(hipStream_t s) const
    {
        if (!on()) return hipSuccess;
        if (hipEventRecord(fork, s) != hipSuccess) return hipErrorUnknown;
        for (int i = 0; i < nside; ++i)
            if (hipStreamWaitEvent(side[i], fork, 0) != hipSuccess) return hipErrorUnknown;
        return hipSuccess;
    }
    // s waits for what is queued on the side streams
    hipError_t end(hipStream_t s) const
    {
        if (!on()) return hipSuccess;
        for (int i = 0; i < nside; ++i)
            if (hipEventRecord(join[i], side[i]) != hipSuccess || hipStreamWaitEvent(s, join[i], 0) != hipSuccess)
                return hipErrorUnknown;
        return hipSuccess;
    }
};

// The handle's side streams and events (created on first use, default
// priority); an empty fork (serial launches) if they cannot be created.
static RaggedFork handle_fork(tokurablas_t *h)
{
    RaggedFork fk;
    bool ok = true;
    for (auto &x : h->ragged_side)
        ok = ok && (x != nullptr || hipStreamCreateWithFlags(&x, hipStreamNonBlocking) == hipSuccess);
    for (auto &x : h->ragged_fork_ev)
        ok = ok && (x != nullptr || hipEventCreateWithFlags(&x, hipEventDisableTiming) == hipSuccess);
    if (!ok) return fk;
    fk.nside = 3;
    for (int i = 0; i < 3; ++i) {
        fk.side[i] = h->ragged_side[i];
        fk.join[i] = h->ragged_fork_ev[1 + i];
    }
    fk.fork = h->ragged_fork_ev[0];
    return fk;
}

template <int LA1, int LA2, int LB, int LC, int W, int MV, int CAP>
static hipError_t launch_ragged_cfg(const zd *A, char *work, zd *eig, int8_t *flag, int itmax, int bal, int cus,
                                    const RaggedPrepared &pe, hipStream_t s, const RaggedFork &fk)
{
    static_assert(LA1 == 64 && LA2 == 64, "classes n > 32: one matrix per wavefront");
    hipError_t e = launch_reset(work, pe.gen, pe.batch, flag, s);
    if (e != hipSuccess) return e;
    hipStream_t s1 = s, s2 = s;   // streams of the n 33..48 and the small-class Hessenberg
    const bool fork = fk.on() && bal == 0;
    if (fork) {
        if (fk.begin(s) != hipSuccess) return hipErrorUnknown;
        s1 = fk.branch(1, s);
        s2 = fk.branch(2, s);
    }
    // resident workgroups per CU, queried once (outside any graph capture the
    // first call is made from tokura_zgeev_ragged_graph's capture stream, where
    // occupancy queries are not stream operations and are allowed)
    if (bal == 0) {   // classes 0-2 (n 49..64) and 3-4 (n 33..48) with the matrix in registers
        if (pe.n49_64 > 0)
            hipLaunchKernelGGL((k_ragged_hess_reg<4>), dim3(pe.n49_64), dim3(256), 0, s, A, work, flag, eig, 0, 2);
        if (pe.n33_48 > 0)
            hipLaunchKernelGGL((k_ragged_hess_reg<3>), dim3(pe.n33_48), dim3(256), 0, s1, A, work, flag, eig, 3, 4);
    } else {
#define X(c, nb)                                                                                          \
    {                                                                                                      \
        static const int occ = resident_blocks(k_ragged_hess_wg<nb, c>, 256);                              \
        hipLaunchKernelGGL((k_ragged_hess_wg<nb, c>), dim3(cus * occ * 4), dim3(256), 0, s, A, work, flag, bal, eig); \
    }
        TKZ_BIG_CLS(X)
#undef X
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    static const int occ_h = resident_blocks(k_ragged_hess<LB, LC>);
    hipLaunchKernelGGL((k_ragged_hess<LB, LC>), dim3(cus * occ_h * pe.hmul), dim3(64), 0, s2, A, work, flag, bal, eig);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (fork && fk.end(s) != hipSuccess) return hipErrorUnknown;   // join the branches
    // one workgroup per CU: the arena is the CU's whole LDS
    hipLaunchKernelGGL((k_ragged_qr_pool<W, LB, LC, MV, CAP>), dim3(cus), dim3(64 * W), 0, s, work, eig, flag, itmax);
    return hipGetLastError();
}

static hipError_t launch_ragged(const RaggedPrepared &pe, const zd *A, char *work, zd *eig, int8_t *flag, int itmax,
                                int bal, int cus, hipStream_t s, const RaggedFork &fk = RaggedFork())
{
    switch (pe.cfg) {
#define X(id, a1, a2, b, c, w, mv, cap) \
    case id: return launch_ragged_cfg<a1, a2, b, c, w, mv, cap>(A, work, eig, flag, itmax, bal, cus, pe, s, fk);
        TKZ_RAGGED_CFGS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}
// ---- differentiation of a ragged batch ------------------------------------
// One launch per size class, each with the LDS footprint and lanes-per-matrix
// of the same-size Schur / adjoint kernels of that class, so every matrix gets
// the same arithmetic as in a same-size call: the Schur form in the size
// classes of the eigenvalue path, the backward in n > 32, 17..32, <= 16.
// order[] lists each class's matrices contiguously (n descending); the Schur
// form of matrix b is at 2*moff[b] (T then U), dA at moff[b], the eigenvalues
// and cotangents at eoff[b].
__device__ __forceinline__ void class_range(const RaggedView &V, int cls, int &lo, int &hi)
{
    lo = cls == 0 ? 0 : V.h->cls_end[kClsB - 2 + cls];
    hi = V.h->cls_end[kClsB - 1 + cls];
}

// The Schur forward of size class CLS in two launches, as the same-size path:
// Hessenberg + Q (full LDS layout), then the QR on the packed layout with U
// in the output; LPM lanes per matrix, 64/LPM matrices per wavefront.
template <int NB, int LPM, int CLS>
__global__ __launch_bounds__(64) void k_ragged_schur_hq(char *work, const zd *__restrict__ A, zd *__restrict__ eig,
                                                        zd *schur, int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM, PER = G::MAT + NB;
    __shared__ zd lds[GPW * PER];
    const RaggedView V(work);
    int lo, hi;
    cls_range(V, CLS, lo, hi);
    const int lane = (int)threadIdx.x, g = lane / LPM, t = lane % LPM;
    zd *H = lds + g * PER;
    for (int base = lo + (int)blockIdx.x * GPW; base < hi; base += (int)gridDim.x * GPW) {
        const int j = base + g;
        if (j < hi) {
            const int b = V.order[j], n = V.nsz[b];
            const long long nn = (long long)n * n;
            zd *Tm = schur + 2 * V.moff[b], *Um = Tm + nn;
            if (schur_hq_one<LPM>(H, H + G::MAT, A + V.moff[b], n, t, Tm, Um)) {
                if (t == 0) flag[b] = 0;
            } else {
                fill_nan<LPM>(eig + V.eoff[b], n, t);
                fill_nan<LPM>(Tm, (int)(2 * nn), t);
                if (t == 0) flag[b] = 2;
            }
        }
        lds_order();
    }
}

// The first half for the classes n > 32: one matrix at a time per workgroup
// of TKZ_ADJ_WG_NT threads (schur_hq_wg, the arithmetic of k_schur_hq_wg).
template <int NB, int CLS>
__global__ __launch_bounds__(TKZ_ADJ_WG_NT) void k_ragged_schur_hq_wg(char *work, const zd *__restrict__ A,
                                                                     zd *__restrict__ eig, zd *schur,
                                                                     int8_t *__restrict__ flag)
{
    constexpr int NT = TKZ_ADJ_WG_NT;
    __shared__ zd lds[NB * (NB | 1) + NB];
    __shared__ double red[8];
    const RaggedView V(work);
    int lo, hi;
    cls_range(V, CLS, lo, hi);
    const int tid = (int)threadIdx.x;
    for (int j = lo + (int)blockIdx.x; j < hi; j += (int)gridDim.x) {
        const int b = V.order[j], n = V.nsz[b];
        const long long nn = (long long)n * n;
        zd *Tm = schur + 2 * V.moff[b], *Um = Tm + nn;
        if (schur_hq_wg<NT, NT / 64>(lds, lds + NB * (NB | 1), red, A + V.moff[b], n, Tm, Um)) {
            if (tid == 0) flag[b] = 0;
        } else {
            fill_nan<NT>(eig + V.eoff[b], n, tid);
            fill_nan<NT>(Tm, (int)(2 * nn), tid);
            if (tid == 0) flag[b] = 2;
        }
        __syncthreads();
    }
}

// The QR of a Schur class as a persistent kernel: every wavefront pulls the
// class's next GPW matrices from the class's queue (qctr[CLS], zeroed by
// k_ragged_reset) until the class is done.
template <int NB, int LPM, int CLS>
__global__ __launch_bounds__(64) void k_ragged_schur_qr(char *work, zd *__restrict__ eig, zd *schur, int8_t *flag,
                                                        int itmax)
{
    constexpr int GPW = 64 / LPM, PER = packed_size(NB) + 1;   // + the idle lanes' element
    __shared__ zd lds[GPW * PER];
    const RaggedView V(work);
    int lo, hi;
    cls_range(V, CLS, lo, hi);
    const int lane = (int)threadIdx.x, g = lane / LPM, t = lane % LPM;
    for (;;) {
        // every lane takes part in the atomic (lane 0 adds 1, the others 0),
        // so no branch on the lane index precedes the loop's back edge
        const int base = lo + __builtin_amdgcn_readfirstlane(atomicAdd(&V.h->qctr[CLS], lane == 0 ? 1 : 0)) * GPW;
        if (base >= hi) break;
        const int j = base + g;
        if (j < hi) {
            const int b = V.order[j], n = V.nsz[b];
            if (flag[b] != 2) {
                zd *Tm = schur + 2 * V.moff[b];
                const int info = schur_qr_one<LPM, (LPM == 64)>(lds + g * PER, n, t, Tm, Tm + (long long)n * n,
                                                                eig + V.eoff[b], itmax);
                if (t == 0) flag[b] = info != 0 ? 1 : 0;
            }
        }
        lds_order();
    }
}

template <int NB, int LPM, int CLS>
__global__ __launch_bounds__(64) void k_ragged_adjoint(char *work, const zd *__restrict__ schur,
                                                       const zd *__restrict__ gbar, zd *dA,
                                                       int8_t *__restrict__ flag, int itmax)
{
    using G = AdjGeom<NB>;
    constexpr int GPW = 64 / LPM;
    __shared__ zd lds[GPW * G::PER];
    const RaggedView V(work);
    int lo, hi;
    class_range(V, CLS, lo, hi);
    const int lane = threadIdx.x & 63, g = lane / LPM, t = lane % LPM;
    zd *B1 = lds + g * G::PER, *B2 = B1 + G::MAT, *gv = B2 + G::MAT + NB;
    for (int base = lo + (int)blockIdx.x * GPW; base < hi; base += (int)gridDim.x * GPW) {
        const int j = base + g;
        if (j < hi) {
            const int b = V.order[j], n = V.nsz[b];
            const zd *Tm = schur + 2 * V.moff[b];
            adjoint_one<NB, LPM>(B1, B2, gv, n, t, Tm, Tm + (long long)n * n, gbar + V.eoff[b], nullptr,
                                 dA + V.moff[b], flag[b]);
        }
        lds_order();
    }
}

// The backward of size classes C0..C1 (n > 16): one matrix at a time per
// workgroup of NT threads (adjoint_wg), LDS sized for NB
template <int NT, int NB, int C0, int C1>
__global__ __launch_bounds__(NT) void k_ragged_adjoint_wg(char *work, const zd *__restrict__ schur,
                                                          const zd *__restrict__ gbar, zd *dA,
                                                          const int8_t *__restrict__ flag)
{
    using G = AdjGeom<NB>;
    __shared__ zd lds[G::PER];
    const RaggedView V(work);
    const int lo = C0 == 0 ? 0 : V.h->cls_end[C0 - 1], hi = V.h->cls_end[C1];
    zd *B1 = lds, *B2 = lds + G::MAT, *gv = B2 + G::MAT;
    for (int j = lo + (int)blockIdx.x; j < hi; j += (int)gridDim.x) {
        const int b = V.order[j], n = V.nsz[b];
        const zd *Tm = schur + 2 * V.moff[b];
        (void)adjoint_wg<NT>(B1, B2, gv, n, Tm, Tm + (long long)n * n, gbar + V.eoff[b], dA + V.moff[b], flag[b]);
        __syncthreads();
    }
}



// the two launches of the Schur forward for size class CLS (LDS for NB)
template <int NB, int LPM, int CLS>
static void launch_ragged_schur_cls(const zd *A, char *work, zd *eig, zd *schur, int8_t *flag, int itmax, int cus,
                                    hipStream_t s)
{
    if constexpr (NB > 32) {
        static const int oh = resident_blocks(k_ragged_schur_hq_wg<NB, CLS>, TKZ_ADJ_WG_NT);
        hipLaunchKernelGGL((k_ragged_schur_hq_wg<NB, CLS>), dim3(cus * oh * 4), dim3(TKZ_ADJ_WG_NT), 0, s, work, A, eig,
                           schur, flag);
    } else {
        static const int oh = resident_blocks(k_ragged_schur_hq<NB, LPM, CLS>);
        hipLaunchKernelGGL((k_ragged_schur_hq<NB, LPM, CLS>), dim3(cus * oh), dim3(64), 0, s, work, A, eig, schur, flag);
    }
    static const int oq = resident_blocks(k_ragged_schur_qr<NB, LPM, CLS>);
    hipLaunchKernelGGL((k_ragged_schur_qr<NB, LPM, CLS>), dim3(cus * oq), dim3(64), 0, s, work, eig, schur, flag, itmax);
}

// The size classes' launch pairs (Hessenberg + Q, then the QR) on four
// branches; each class keeps its two launches in order on its branch and
// its own queue head.
static hipError_t launch_ragged_schur(const zd *A, char *work, zd *eig, zd *schur, int8_t *flag, int itmax, int cus,
                                      int gen, int batch, hipStream_t s, const RaggedFork &fk = RaggedFork())
{
    hipError_t e = launch_reset(work, gen, batch, flag, s);
    if (e != hipSuccess) return e;
    if (fk.begin(s) != hipSuccess) return hipErrorUnknown;
    launch_ragged_schur_cls<64, 64, 0>(A, work, eig, schur, flag, itmax, cus, fk.branch(0, s));
    launch_ragged_schur_cls<60, 64, 1>(A, work, eig, schur, flag, itmax, cus, fk.branch(1, s));
    launch_ragged_schur_cls<56, 64, 2>(A, work, eig, schur, flag, itmax, cus, fk.branch(2, s));
    launch_ragged_schur_cls<48, 64, 3>(A, work, eig, schur, flag, itmax, cus, fk.branch(3, s));
    launch_ragged_schur_cls<40, 64, 4>(A, work, eig, schur, flag, itmax, cus, fk.branch(2, s));
    launch_ragged_schur_cls<32, 32, kClsB>(A, work, eig, schur, flag, itmax, cus, fk.branch(1, s));
    launch_ragged_schur_cls<16, 16, kClsC>(A, work, eig, schur, flag, itmax, cus, fk.branch(0, s));
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return fk.end(s);
}

// The backward's five launches stay in order on the caller's stream: run as
// branches they took 38.7 ms against 35.5 ms in series on the hard-spectra
// batch (the n 49..64 launch, one workgroup per CU, lost its slots to the
// smaller ones).
static hipError_t launch_ragged_adjoint(const zd *schur, const zd *gbar, zd *dA, char *work, int8_t *flag, int itmax,
                                        int cus, int gen, int batch, hipStream_t s)
{
    // the schedule check (a stale `work` gives flag 4 and no work)
    hipError_t e = launch_reset(work, gen, batch, flag, s);
    if (e != hipSuccess) return e;
    // n > 16 by 256-thread workgroups; the LDS is sized in steps where a
    // smaller footprint keeps more workgroups resident (n <= 48: two per CU,
    // n <= 40: three, n <= 32: four; the classes above 48 share the n = 64 size)
#define X(nb, c0, c1)                                                                                      \
    {                                                                                                       \
        static const int o = resident_blocks(k_ragged_adjoint_wg<TKZ_ADJ_WG_NT, nb, c0, c1>, TKZ_ADJ_WG_NT); \
        hipLaunchKernelGGL((k_ragged_adjoint_wg<TKZ_ADJ_WG_NT, nb, c0, c1>), dim3(cus * o), dim3(TKZ_ADJ_WG_NT), 0, s, \
                           work, schur, gbar, dA, flag);                                                   \
    }
    X(64, 0, 2)
    X(48, 3, 3)
    X(40, 4, 4)
    X(32, kClsB, kClsB)
#undef X
    static const int o2 = resident_blocks(k_ragged_adjoint<16, 16, 2>);
    hipLaunchKernelGGL((k_ragged_adjoint<16, 16, 2>), dim3(cus * o2), dim3(64), 0, s, work, schur, gbar, dA, flag, itmax);
    return hipGetLastError();
}


// ---- host-side schedule ---------------------------------------------------
struct Schedule {
    RaggedHeader hdr;
    std::vector<int> tiles;   // 4 ints per tile
    std::vector<int> order, nsz;
    std::vector<long long> moff, eoff, hoff;
    long long nH = 0;
};

static long long align256(long long x) { return (x + 255) & ~255LL; }

static int build_schedule(int batch, const int64_t *offsets, int cfg, Schedule &S)
{
    S.nsz.resize(batch);
    S.moff.resize(batch);
    S.eoff.resize(batch);
    S.hoff.resize(batch);
    long long e = 0;
    for (int b = 0; b < batch; ++b) {
        long long cnt = (long long)offsets[b + 1] - (long long)offsets[b];
        long long n = (long long)std::llround(std::sqrt((double)cnt));
        if (offsets[b] < 0 || cnt < 1 || n * n != cnt || n > TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE) return -3;
        S.nsz[b] = (int)n;
        S.moff[b] = offsets[b];
        S.eoff[b] = e;
        e += n;
    }
    // order: by size class, then n descending (largest work first)
    S.order.resize(batch);
    std::iota(S.order.begin(), S.order.end(), 0);
    std::stable_sort(S.order.begin(), S.order.end(), [&](int a, int b) { return S.nsz[a] > S.nsz[b]; });
    long long h = 0;
    for (int i = 0; i < batch; ++i) {
        int b = S.order[i];
        S.hoff[b] = h;
        h += packed_size(S.nsz[b]);
    }
    S.nH = h;
    S.tiles.clear();
    int i = 0;
    while (i < batch) {
        const int c = cls_of(S.nsz[S.order[i]]);
        const int per = 64 / cfg_lpm(cfg, c);
        int j = i;
        while (j < batch && j - i < per && cls_of(S.nsz[S.order[j]]) == c) ++j;
        S.tiles.insert(S.tiles.end(), {c, j - i, i, S.nsz[S.order[i]]});
        i = j;
    }
    const int ntiles = (int)(S.tiles.size() / 4);
    std::memset(&S.hdr, 0, sizeof S.hdr);
    for (int c = 0, k = 0; c < kNumCls; ++c) {   // order[] is in class order
        while (k < batch && cls_of(S.nsz[S.order[k]]) <= c) ++k;
        S.hdr.cls_end[c] = k;
    }
    S.hdr.tile_small0 = 0;
    while (S.hdr.tile_small0 < ntiles && S.tiles[4 * S.hdr.tile_small0] < kClsB) ++S.hdr.tile_small0;
    for (int c = 0, k = 0; c <= kNumCls; ++c) {   // tiles are in class order
        while (k < ntiles && S.tiles[4 * k] < c) ++k;
        S.hdr.tile_cls[c] = c == kNumCls ? ntiles : k;
    }
    S.hdr.ntiles = ntiles;
    S.hdr.batch = batch;
    S.hdr.cfg = cfg;
    long long o = align256(sizeof(RaggedHeader));
    S.hdr.off_tiles = o;  o = align256(o + 16LL * ntiles);
    S.hdr.off_order = o;  o = align256(o + 4LL * batch);
    S.hdr.off_n = o;      o = align256(o + 4LL * batch);
    S.hdr.off_moff = o;   o = align256(o + 8LL * batch);
    S.hdr.off_eoff = o;   o = align256(o + 8LL * batch);
    S.hdr.off_hoff = o;   o = align256(o + 8LL * batch);
    S.hdr.off_H = o;      o = align256(o + 16LL * S.nH);
    S.hdr.total_bytes = o;
    return 0;
}

}  // namespace tkz

using namespace tkz;

void ragged_release(tokurablas_t *h)
{
    if (h->graph.exec) (void)hipGraphExecDestroy(h->graph.exec);   // teardown: nothing to undo on failure
    if (h->graph.graph) (void)hipGraphDestroy(h->graph.graph);
    h->graph = RaggedGraphCache();
    if (h->ragged_host_stage) (void)hipHostFree(h->ragged_host_stage);
    h->ragged_host_stage = nullptr;
    h->ragged_host_bytes = 0;
    if (h->ragged_stage_event) (void)hipEventDestroy(h->ragged_stage_event);
    h->ragged_stage_event = nullptr;
    for (auto &x : h->ragged_side)
        if (x) (void)hipStreamDestroy(x);
    for (auto &x : h->ragged_fork_ev)
        if (x) (void)hipEventDestroy(x);
    for (auto &x : h->ragged_side) x = nullptr;
    for (auto &x : h->ragged_fork_ev) x = nullptr;
    h->prepared.clear();
}

int ragged_num_configs() { return kNumRaggedCfg; }

extern "C" size_t tokura_get_zgeev_ragged_worksize(int batch_count, const int64_t *offsets_host)
{
    if (batch_count <= 0) return (size_t)-2;
    if (offsets_host == nullptr) return (size_t)-3;
    Schedule S;
    // the byte count does not depend on the class configuration except for
    // the tile array; size it for the configuration with the most tiles
    long long worst = 0;
    for (int c = 0; c < kNumRaggedCfg; ++c) {
        if (build_schedule(batch_count, offsets_host, c, S) != 0) return (size_t)-3;
        worst = std::max(worst, S.hdr.total_bytes);
    }
    return (size_t)worst;
}

// The size signature of a batch, the key of the ragged tuning table: each
// size class's share of the batch in sixteenths (5 bits per class) and
// floor(log2(batch)) (bits 35..39).  Entries that do not describe a
// matrix are skipped (the schedule builder rejects such offsets).
static unsigned long long ragged_signature(int batch, const int64_t *offsets)
{
    long long cnt[kNumCls] = {0};
    for (int b = 0; b < batch; ++b) {
        const long long d = (long long)offsets[b + 1] - (long long)offsets[b];
        const long long n = d > 0 ? (long long)std::llround(std::sqrt((double)d)) : 0;
        if (n >= 1 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE && n * n == d) ++cnt[cls_of((int)n)];
    }
    unsigned long long sig = 0;
    for (int c = 0; c < kNumCls; ++c) sig |= (unsigned long long)((16 * cnt[c] + batch / 2) / batch) << (5 * c);
    int lb = 0;
    while ((2LL << lb) <= (long long)batch) ++lb;
    return sig | (unsigned long long)lb << 35;
}

// The configuration for this batch: the tuned entry of its signature when
// the handle has a tuning table (tokuraAutotuneRagged or a cache file) and
// no configuration was set by hand, else the handle's current one.
static void ragged_choose(const tokurablas_t *h, int batch, const int64_t *offsets, int &cfg, int &hmul)
{
    cfg = h->ragged_cfg;
    hmul = h->ragged_hmul;
    if (!h->ragged_auto) return;
    const unsigned long long sig = ragged_signature(batch, offsets);
    for (const auto &t : h->ragged_tuned)
        if (t.sig == sig) {
            cfg = t.cfg;
            hmul = t.hmul;
        }
}

extern "C" int tokura_zgeev_ragged_prepare(tokurablas_t *handle, int batch_count, const int64_t *offsets_host,
                                           void *work, size_t work_bytes, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (batch_count <= 0) return -6;
    if (offsets_host == nullptr) return -3;
    if (work == nullptr) return -5;
    Schedule S;
    int cfg, hmul;
    ragged_choose(handle, batch_count, offsets_host, cfg, hmul);
    if (build_schedule(batch_count, offsets_host, cfg, S) != 0) return -3;
    if ((long long)work_bytes < S.hdr.total_bytes) return -5;
    const size_t meta = (size_t)S.hdr.off_H;
    // the staging buffer may still feed an earlier copy: wait for that copy
    // (an event of the handle, recorded after it on the caller's stream)
    if (handle->ragged_stage_event == nullptr &&
        hipEventCreateWithFlags(&handle->ragged_stage_event, hipEventDisableTiming) != hipSuccess) {
        handle->ragged_stage_event = nullptr;
        return 3;
    }
    if (handle->ragged_staged && hipEventSynchronize(handle->ragged_stage_event) != hipSuccess) return 3;
    handle->ragged_staged = false;
    S.hdr.magic = kRaggedMagic;
    S.hdr.gen = ++handle->ragged_gen;
    if (handle->ragged_host_bytes < meta) {
        if (handle->ragged_host_stage) (void)hipHostFree(handle->ragged_host_stage);
        handle->ragged_host_stage = nullptr;
        handle->ragged_host_bytes = 0;
        if (hipHostMalloc(&handle->ragged_host_stage, meta, hipHostMallocDefault) != hipSuccess) return 3;
        handle->ragged_host_bytes = meta;
    }
    char *st = (char *)handle->ragged_host_stage;
    std::memset(st, 0, meta);
    std::memcpy(st, &S.hdr, sizeof S.hdr);
    std::memcpy(st + S.hdr.off_tiles, S.tiles.data(), 4 * S.tiles.size());
    std::memcpy(st + S.hdr.off_order, S.order.data(), 4LL * batch_count);
    std::memcpy(st + S.hdr.off_n, S.nsz.data(), 4LL * batch_count);
    std::memcpy(st + S.hdr.off_moff, S.moff.data(), 8LL * batch_count);
    std::memcpy(st + S.hdr.off_eoff, S.eoff.data(), 8LL * batch_count);
    std::memcpy(st + S.hdr.off_hoff, S.hoff.data(), 8LL * batch_count);
    if (hipMemcpyAsync(work, st, meta, hipMemcpyHostToDevice, stream) != hipSuccess) return 3;
    if (hipEventRecord(handle->ragged_stage_event, stream) != hipSuccess) {
        // the copy may still read the staging buffer: wait for it here
        if (hipStreamSynchronize(stream) != hipSuccess) return 3;
    } else {
        handle->ragged_staged = true;
    }
    // remember which configuration and generation this work buffer holds
    const RaggedPrepared rp{work, cfg, hmul, batch_count, S.hdr.gen, S.hdr.cls_end[2],
                            S.hdr.cls_end[4] - S.hdr.cls_end[2]};
    bool found = false;
    for (auto &p : handle->prepared)
        if (p.work == work) { p = rp; found = true; }
    if (!found) handle->prepared.push_back(rp);
    return 0;
}

// The entry prepared for (work, batch), or null.
static const RaggedPrepared *prepared_entry(tokurablas_t *h, void *work, int batch)
{
    for (auto &p : h->prepared)
        if (p.work == work && p.batch == batch) return &p;
    return nullptr;
}

extern "C" int tokura_zgeev_ragged_gpu(tokurablas_t *handle, int batch_count, const hipDoubleComplex *A,
                                       hipDoubleComplex *eigenvalues, void *work, char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (work == nullptr) return -5;
    if (batch_count <= 0) return -6;
    if (flag == nullptr) return -7;
    const RaggedPrepared *pe = prepared_entry(handle, work, batch_count);
    if (pe == nullptr) return -8;
    hipError_t e = launch_ragged(*pe, (const zd *)A, (char *)work, (zd *)eigenvalues, (int8_t *)flag,
                                 handle->max_sweeps, handle->balance, handle->cu_count > 0 ? handle->cu_count : 1,
                                 stream);
    return e == hipSuccess ? 0 : 2;
}

extern "C" int tokura_zgeev_ragged_graph(tokurablas_t *handle, int batch_count, const hipDoubleComplex *A,
                                         hipDoubleComplex *eigenvalues, void *work, char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (work == nullptr) return -5;
    if (batch_count <= 0) return -6;
    if (flag == nullptr) return -7;
    const RaggedPrepared *pe = prepared_entry(handle, work, batch_count);
    if (pe == nullptr) return -8;
    const int cfg = pe->cfg, gen = pe->gen;
    RaggedGraphCache &G = handle->graph;
    const bool hit = G.exec && G.A == (const void *)A && G.eig == (void *)eigenvalues && G.work == work &&
                     G.flag == flag && G.batch == batch_count && G.cfg == cfg && G.max_sweeps == handle->max_sweeps &&
                     G.balance == handle->balance && G.gen == gen;
    if (!hit) {
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        if (G.graph) (void)hipGraphDestroy(G.graph);
        G = RaggedGraphCache();
        hipStream_t cs;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return 3;
        // the capture stream and two side streams (default priority) with
        // the fork/join events of the Hessenberg branches
        RaggedFork fk;
        fk.nside = 2;
        bool fk_ok = true;
        for (int i = 0; i < 2; ++i)
            fk_ok = fk_ok && hipStreamCreateWithFlags(&fk.side[i], hipStreamNonBlocking) == hipSuccess &&
                    hipEventCreateWithFlags(&fk.join[i], hipEventDisableTiming) == hipSuccess;
        fk_ok = fk_ok && hipEventCreateWithFlags(&fk.fork, hipEventDisableTiming) == hipSuccess;
        hipError_t e = fk_ok ? hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) : hipErrorOutOfMemory;
        if (e == hipSuccess) {
            hipError_t el = launch_ragged(*pe, (const zd *)A, (char *)work, (zd *)eigenvalues, (int8_t *)flag,
                                          handle->max_sweeps, handle->balance,
                                          handle->cu_count > 0 ? handle->cu_count : 1, cs, fk);
            e = hipStreamEndCapture(cs, &G.graph);
            if (el != hipSuccess) e = el;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cs);
        for (auto &x : fk.side)
            if (x) (void)hipStreamDestroy(x);
        for (auto &x : fk.join)
            if (x) (void)hipEventDestroy(x);
        if (fk.fork) (void)hipEventDestroy(fk.fork);
        if (e != hipSuccess) {
            if (G.graph) (void)hipGraphDestroy(G.graph);
            G = RaggedGraphCache();
            return 2;
        }
        G.A = A; G.eig = eigenvalues; G.work = work; G.flag = flag; G.batch = batch_count; G.cfg = cfg;
        G.max_sweeps = handle->max_sweeps;
        G.balance = handle->balance;
        G.gen = gen;
    }
    return hipGraphLaunch(G.exec, stream) == hipSuccess ? 0 : 2;
}

extern "C" int tokuraSetRaggedConfig(tokurablas_t *handle, int cfg)
{
    if (handle == nullptr) return -1;
    if (cfg < 0 || cfg >= kNumRaggedCfg) return -2;
    handle->ragged_cfg = cfg;
    handle->ragged_auto = false;   // a configuration set by hand wins over the tuning table
    return 0;
}

extern "C" int tokuraGetRaggedConfig(tokurablas_t *handle, int *cfg, int *lpm_a1, int *lpm_a2, int *lpm_b,
                                     int *lpm_c)
{
    if (handle == nullptr) return -1;
    const int c = handle->ragged_cfg;
    if (cfg) *cfg = c;
    if (lpm_a1) *lpm_a1 = g_cfg_lpm[c][0];
    if (lpm_a2) *lpm_a2 = g_cfg_lpm[c][1];
    if (lpm_b) *lpm_b = g_cfg_lpm[c][2];
    if (lpm_c) *lpm_c = g_cfg_lpm[c][3];
    return 0;
}

extern "C" int tokuraGetRaggedWaves(tokurablas_t *handle, int *waves)
{
    if (handle == nullptr) return -1;
    if (waves) *waves = g_cfg_lpm[handle->ragged_cfg][4];
    return 0;
}

extern "C" int tokuraGetRaggedPrepared(tokurablas_t *handle, void *work, int batch_count, int *config,
                                       int *multiplier)
{
    if (handle == nullptr) return -1;
    const RaggedPrepared *pe = prepared_entry(handle, work, batch_count);
    if (pe == nullptr) return -8;
    if (config) *config = pe->cfg;
    if (multiplier) *multiplier = pe->hmul;
    return 0;
}

extern "C" int tokuraGetRaggedHessGrid(tokurablas_t *handle, int *multiplier)
{
    if (handle == nullptr) return -1;
    if (multiplier) *multiplier = handle->ragged_hmul;
    return 0;
}

// Milliseconds of one ragged call (the second of two) with the handle's
// current configuration, re-preparing `work` first; -1 on an error (rc set).
static float time_ragged(tokurablas_t *h, int batch, const int64_t *offs, const hipDoubleComplex *A,
                         hipDoubleComplex *eig, void *work, size_t work_bytes, char *flag, hipStream_t s,
                         hipEvent_t e0, hipEvent_t e1, int &rc)
{
    rc = tokura_zgeev_ragged_prepare(h, batch, offs, work, work_bytes, s);
    if (rc) return -1.f;
    float ms = 0.f;
    for (int rep = 0; rep < 2 && rc == 0; ++rep) {
        if (hipEventRecord(e0, s) != hipSuccess) { rc = 3; break; }
        rc = tokura_zgeev_ragged_gpu(h, batch, A, eig, work, flag, s);
        if (rc == 0 && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = 3;
    }
    return rc ? -1.f : ms;
}

// Two passes over the caller's batch: every class configuration at the
// current Hessenberg grid multiplier, then the multipliers 1, 2, 4, 8 of the
// small-class Hessenberg grid with the fastest configuration.  The winner
// becomes the handle's configuration and the tuning-table entry of the
// batch's size signature (tokuraSaveTuning writes the table).
extern "C" int tokuraAutotuneRagged(tokurablas_t *handle, int batch_count, const int64_t *offsets_host,
                                    const hipDoubleComplex *A, hipDoubleComplex *eigenvalues, void *work,
                                    size_t work_bytes, char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (batch_count <= 0) return -6;
    if (offsets_host == nullptr) return -3;
    const int keep = handle->ragged_cfg, keep_h = handle->ragged_hmul;
    const bool keep_auto = handle->ragged_auto;
    handle->ragged_auto = false;   // time the configurations set below
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best = 1e30f;
    int bestc = keep, besth = keep_h, rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = 3;
    for (int c = 0; c < kNumRaggedCfg && rc == 0; ++c) {
        handle->ragged_cfg = c;
        const float ms = time_ragged(handle, batch_count, offsets_host, A, eigenvalues, work, work_bytes, flag,
                                     stream, e0, e1, rc);
        if (rc == 0 && ms < best) { best = ms; bestc = c; }
    }
    handle->ragged_cfg = bestc;
    for (int hm = 1; hm <= 8 && rc == 0; hm *= 2) {
        if (hm == keep_h) continue;   // timed above
        handle->ragged_hmul = hm;
        const float ms = time_ragged(handle, batch_count, offsets_host, A, eigenvalues, work, work_bytes, flag,
                                     stream, e0, e1, rc);
        if (rc == 0 && ms < best) { best = ms; besth = hm; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == 0) {
        handle->ragged_cfg = bestc;
        handle->ragged_hmul = besth;
        const unsigned long long sig = ragged_signature(batch_count, offsets_host);
        bool found = false;
        for (auto &t : handle->ragged_tuned)
            if (t.sig == sig) { t.cfg = bestc; t.hmul = besth; found = true; }
        if (!found) handle->ragged_tuned.push_back(RaggedTune{sig, bestc, besth});
        handle->ragged_auto = true;
    } else {
        handle->ragged_cfg = keep;
        handle->ragged_hmul = keep_h;
        handle->ragged_auto = keep_auto;
    }
    int rc2 = tokura_zgeev_ragged_prepare(handle, batch_count, offsets_host, work, work_bytes, stream);
    if (rc2 == 0 && hipStreamSynchronize(stream) != hipSuccess) rc2 = 3;
    return rc ? rc : rc2;
}

void ragged_write_tuning(const tokurablas_t *h, FILE *f)
{
    for (const auto &t : h->ragged_tuned)
        std::fprintf(f, "ragged %s %d %llx %d %d\n", tuning_arch(h), h->cu_count, t.sig, t.cfg, t.hmul);
}

int ragged_parse_tuning(tokurablas_t *h, const char *line)
{
    char arch[64];
    int cus, cfg, hmul;
    unsigned long long sig;
    if (std::sscanf(line, "ragged %63s %d %llx %d %d", arch, &cus, &sig, &cfg, &hmul) != 5) return 0;
    if (std::strcmp(arch, tuning_arch(h)) != 0 || cus != h->cu_count) return 0;
    if (cfg < 0 || cfg >= kNumRaggedCfg || !(hmul == 1 || hmul == 2 || hmul == 4 || hmul == 8)) return 0;
    for (auto &t : h->ragged_tuned)
        if (t.sig == sig) {
            t.cfg = cfg;
            t.hmul = hmul;
            h->ragged_auto = true;
            return 1;
        }
    h->ragged_tuned.push_back(RaggedTune{sig, cfg, hmul});
    h->ragged_auto = true;
    return 1;
}

extern "C" size_t tokura_get_zgeev_ragged_schur_worksize(int batch_count, const int64_t *offsets_host)
{
    if (batch_count <= 0) return (size_t)-2;
    if (offsets_host == nullptr) return (size_t)-1;
    long long top = 0;
    for (int b = 0; b < batch_count; ++b) {
        const long long cnt = (long long)offsets_host[b + 1] - (long long)offsets_host[b];
        const long long n = (long long)std::llround(std::sqrt((double)cnt));
        if (offsets_host[b] < 0 || cnt < 1 || n * n != cnt || n > TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)
            return (size_t)-1;
        top = std::max(top, (long long)offsets_host[b] + cnt);
    }
    return sizeof(hipDoubleComplex) * 2 * (size_t)top;
}

extern "C" int tokura_zgeev_ragged_schur_gpu(tokurablas_t *handle, int batch_count, const hipDoubleComplex *A,
                                             hipDoubleComplex *eigenvalues, hipDoubleComplex *schur, void *work,
                                             char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (work == nullptr || schur == nullptr) return -5;
    if (batch_count <= 0) return -6;
    if (flag == nullptr) return -7;
    const RaggedPrepared *pe = prepared_entry(handle, work, batch_count);
    if (pe == nullptr) return -8;
    hipError_t e = launch_ragged_schur((const zd *)A, (char *)work, (zd *)eigenvalues, (zd *)schur, (int8_t *)flag,
                                       handle->max_sweeps, handle->cu_count > 0 ? handle->cu_count : 1, pe->gen,
                                       batch_count, stream, handle_fork(handle));
    return e == hipSuccess ? 0 : 2;
}

extern "C" int tokura_zgeev_ragged_backward_gpu(tokurablas_t *handle, int batch_count,
                                                const hipDoubleComplex *schur, const hipDoubleComplex *gbar,
                                                hipDoubleComplex *dA, void *work, char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (schur == nullptr) return -3;
    if (gbar == nullptr) return -4;
    if (dA == nullptr || work == nullptr) return -5;
    if (batch_count <= 0) return -6;
    if (flag == nullptr) return -7;
    const RaggedPrepared *pe = prepared_entry(handle, work, batch_count);
    if (pe == nullptr) return -8;
    hipError_t e = launch_ragged_adjoint((const zd *)schur, (const zd *)gbar, (zd *)dA, (char *)work, (int8_t *)flag,
                                         handle->max_sweeps, handle->cu_count > 0 ? handle->cu_count : 1, pe->gen,
                                         batch_count, stream);
    return e == hipSuccess ? 0 : 2;
}
