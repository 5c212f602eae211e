// An LDS arena shared by the wavefronts of one persistent workgroup per CU,
// for the eigenvalue-only QR of matrices held one per wavefront.
//
// The QR of an n x n Hessenberg matrix needs packed_size(n) elements of LDS
// at its start (34.3 KB at n = 64: four matrices per CU when every
// wavefront has its own fixed buffer), but after each deflation only the
// leading columns 0..i are ever read again, a prefix of the packed layout.
// Averaged over the QR's time that is about half of the initial size.  The
// arena is cut into slots of a few sizes (a "layout"); a wavefront takes a
// slot that fits its work item, and when the active part of its matrix fits
// a smaller free slot it copies the prefix there and frees the larger one
// for the next matrix.  More matrices are then resident per CU, so the SIMDs
// interleave more serial chase chains.  The arithmetic of the QR does not
// change (the eigenvalues are bit-identical to the fixed-buffer kernels').
//
// Slots are handed out through a bit mask in LDS (one bit per slot, set =
// free).  A wavefront never waits while holding a slot, so every wait ends
// when a working wavefront finishes its matrix (the QR is bounded by its
// sweep cap).
//
// Control flow: every lane of the wavefront executes each atomic, lane 0 with
// the real operand and the others with the operation's identity (add 0, and
// with all ones, or with 0), and every decision is taken on a value read
// with readfirstlane, so no branch on the lane index sits between the
// wave-wide reads of these loops.  (A branch "if (lane == 0) atomic" before
// the loop's back edge let the compiler send lane 0 down that branch alone
// while the other lanes started the next claim without it.)
#pragma once
#include "zgeev_device.hpp"

namespace tkz {

constexpr int kPoolMaxSlots = 16;
// complex elements of the arena: 160 KB of LDS minus 256 bytes of control
constexpr int kArena = (160 * 1024 - 256) / 16;

// A slot layout, slots sorted by capacity (complex elements), largest first.
struct PoolCfg {
    int nslots;
    int cap[kPoolMaxSlots];
    int off[kPoolMaxSlots];
};

// Slots whose capacity is at least `need`: the leading bits, since the
// capacities are sorted in decreasing order.
__device__ __forceinline__ unsigned pool_fit(const PoolCfg &P, int need)
{
    int k = 0;
    while (k < P.nslots && P.cap[k] >= need) ++k;
    return (1u << k) - 1u;
}

// One attempt to take a free slot among `cand` (the smallest-capacity one);
// returns the slot or -1.  Called by the whole wavefront (uniform result).
__device__ __forceinline__ int pool_try(unsigned *free_mask, unsigned cand, int lane)
{
    unsigned m = __builtin_amdgcn_readfirstlane(__hip_atomic_load(free_mask, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_WORKGROUP) & cand);
    while (m) {
        const int s = 31 - __builtin_clz(m);
        const unsigned old = __builtin_amdgcn_readfirstlane(atomicAnd(free_mask, lane == 0 ? ~(1u << s) : ~0u));
        if (old & (1u << s)) return s;
        m = old & cand;   // taken meanwhile: the others that are still free
    }
    return -1;
}

// Waits for a free slot among `cand` (non-empty).
__device__ __forceinline__ int pool_acquire(unsigned *free_mask, unsigned cand, int lane)
{
    for (;;) {
        const int s = pool_try(free_mask, cand, lane);
        if (s >= 0) return s;
        __builtin_amdgcn_s_sleep(2);
    }
}

// Frees slot s after every LDS access of the wavefront to it (the LDS
// serves one wavefront's operations in order; the fence keeps the compiler
// from moving accesses past the atomic).
__device__ __forceinline__ void pool_release(unsigned *free_mask, int s, int lane)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    atomicOr(free_mask, lane == 0 ? 1u << s : 0u);
}

// zlahqr_dual's hook: after a deflation that leaves columns 0..i active,
// move them to a smaller free slot when one fits (the idle lanes' spare
// element included), writing the eigenvalues of the rows that became final
// (i+1 .. whi-1) from the old slot first.
struct PoolMover {
    const PoolCfg *P;
    unsigned *free_mask;
    zd *arena;
    zd *out;      // eigenvalues of the matrix (global)
    int slot;     // current slot
    int whi;      // rows whi.. of out are written
    zd *H;        // current base
    int lane;
    double us;    // norm-range factor of the eigenvalues (1 when none)

    __device__ __forceinline__ zd eig(zd h) const { return us != 1.0 ? zscale(us, h) : h; }

    __device__ __forceinline__ void moved(zd *&Hq, int &idle, int i, int t)
    {
        const int ps = packed_size(i + 1);
        const unsigned cand = pool_fit(*P, ps + 1) & ~pool_fit(*P, P->cap[slot]);
        if (!cand) return;
        const int s2 = pool_try(free_mask, cand, lane);
        if (s2 < 0) return;
        const PackedLayout pk;
        if (i + 1 + t < whi) out[i + 1 + t] = eig(Hq[pk(i + 1 + t, i + 1 + t)]);   // whi <= n <= 64
        whi = i + 1;
        zd *H2 = arena + P->off[s2];
        for (int e = t; e < ps; e += 64) H2[e] = Hq[e];
        pool_release(free_mask, slot, lane);
        lds_order();
        slot = s2;
        Hq = H2;
        H = H2;
        idle = P->cap[s2] - 1;
    }
};

// The eigenvalue-only QR of one packed Hessenberg matrix (n x n at src, the
// packed layout) by the calling wavefront, from slot `slot` of the arena
// (capacity >= packed_size(n) + 1), moving to smaller slots as it deflates;
// eigenvalues times us to out, returns the QR INFO.  The slot held at the end
// is released.
__device__ __forceinline__ int pool_qr_one(const PoolCfg &P, unsigned *free_mask, zd *arena, int slot,
                                           const zd *__restrict__ src, int n, zd *__restrict__ out, double us,
                                           int itmax, int lane)
{
    zd *H = arena + P.off[slot];
    const int ps = packed_size(n);
    for (int e = lane; e < ps; e += 64) H[e] = src[e];
    lds_order();
    PoolMover mv{&P, free_mask, arena, out, slot, n, H, lane, us};
    int info = 0;
    if (n > 1) info = zlahqr_dual<64, false, true, PoolMover>(H, PackedLayout{}, n, lane, itmax, nullptr, 0, &mv,
                                                              P.cap[slot] - 1);
    const PackedLayout pk;
    if (lane < mv.whi) out[lane] = mv.eig(mv.H[pk(lane, lane)]);   // whi <= n <= 64
    pool_release(free_mask, mv.slot, lane);
    return info;
}

}  // namespace tkz

namespace tkz {

// Slot layouts for the same-size pool variants: wavefronts per workgroup and
// the share of the arena given to slots of six tiers, for active sizes
// ceil(f n) with f = 1, 7/8, 3/4, 5/8, 1/2, 1/4 (each slot one element
// larger than the packed matrix, for the idle lanes).  A matrix spends about
// the fraction (s_k^2 - s_{k+1}^2) / n^2 of its QR time with an active size
// in (s_{k+1}, s_k] (a sweep costs steps in proportion to the active size),
// and a new matrix needs a tier-0 slot, so the tier-0 count bounds the rate
// at which matrices enter; the autotuner picks among these layouts.
constexpr int kPoolTiers = 6;
struct PoolLayout {
    int waves;
    float share[kPoolTiers];
};
static constexpr int kTierEighths[kPoolTiers] = {8, 7, 6, 5, 4, 2};
#define TKZ_POOL_LAYOUTS(X)                                       \
    X(0, 12, 0.63f, 0.00f, 0.24f, 0.00f, 0.11f, 0.015f)           \
    X(1, 16, 0.63f, 0.00f, 0.24f, 0.00f, 0.11f, 0.015f)           \
    X(2, 12, 0.63f, 0.17f, 0.12f, 0.00f, 0.055f, 0.03f)           \
    X(3, 12, 0.63f, 0.17f, 0.00f, 0.09f, 0.055f, 0.03f)           \
    X(4, 12, 0.84f, 0.00f, 0.00f, 0.00f, 0.11f, 0.03f)            \
    X(5, 12, 0.42f, 0.33f, 0.12f, 0.00f, 0.11f, 0.015f)           \
    X(6, 8, 0.63f, 0.00f, 0.24f, 0.00f, 0.11f, 0.015f)            \
    X(7, 12, 0.63f, 0.00f, 0.12f, 0.09f, 0.11f, 0.03f)

// Builds the slot table of a layout for matrices of size n; returns false
// when it cannot hold one matrix.  Tiers that coincide for small n merge.
inline bool build_pool_cfg(int n, const PoolLayout &L, PoolCfg &P)
{
    int used = 0;
    P.nslots = 0;
    int last_cap = -1;
    for (int k = 0; k < kPoolTiers; ++k) {
        const int tn = (kTierEighths[k] * n + 7) / 8;
        const int cap = packed_size(tn) + 1;
        int cnt = (int)(L.share[k] * (float)kArena) / cap;
        if (k == 0 && cnt < 1) cnt = 1;
        if (cap == last_cap) cnt = 0;   // the same size as the tier above
        if (P.nslots + cnt > kPoolMaxSlots) cnt = kPoolMaxSlots - P.nslots;
        if (used + cnt * cap > kArena) cnt = (kArena - used) / cap;
        for (int c = 0; c < cnt; ++c) {
            P.cap[P.nslots] = cap;
            used += cap;
            ++P.nslots;
        }
        if (cnt > 0) last_cap = cap;
    }
    // the shares round down: fill what is left with slots of the largest
    // tier that still fits and that the layout uses (capacities stay sorted)
    for (int k = 0; k < kPoolTiers && P.nslots < kPoolMaxSlots; ++k) {
        if (L.share[k] <= 0.f) continue;
        const int cap = packed_size((kTierEighths[k] * n + 7) / 8) + 1;
        if (used + cap > kArena) continue;
        int pos = 0;   // insert after the slots of capacity >= cap
        while (pos < P.nslots && P.cap[pos] >= cap) ++pos;
        for (int s = P.nslots; s > pos; --s) P.cap[s] = P.cap[s - 1];
        P.cap[pos] = cap;
        ++P.nslots;
        used += cap;
        --k;   // the same tier may take another slot
    }
    used = 0;
    for (int s = 0; s < P.nslots; ++s) {
        P.off[s] = used;
        used += P.cap[s];
    }
    for (int s = P.nslots; s < kPoolMaxSlots; ++s) P.cap[s] = P.off[s] = 0;
    return P.nslots > 0 && P.cap[0] == packed_size(n) + 1 && used <= kArena;
}

}  // namespace tkz

namespace tkz {

// ---- chunked arena for work items of any size (the ragged path and the
// same-size chunk QR) ----------
// The arena is cut into kChunks chunks of kChunk complex elements, with one
// bit per chunk (set = free) in kBitWords 64-bit LDS words and a lock word.
// A new work item takes the lowest run of free chunks that holds it; a
// deflating matrix either compacts down (the lowest run that holds its
// active part, its own run counting as free), which keeps the free chunks
// together at the top, or moves to the highest free run once its active
// part needs a fraction of its run (ChunkMover below), so that the
// regions that items start in stay at the bottom, where the next (LPT
// order: equal or smaller) item fits the hole it leaves.  Claims run
// under the lock (one wavefront at a time, a few dozen instructions, no
// wait inside) after a search of a snapshot without it, and nothing waits
// for the lock beyond a few tries; frees need no lock (they only set bits
// the wavefront owns).  Every lane runs each atomic, the lanes that have no word
// to change with the identity operand.
constexpr int kChunk = 32;
constexpr int kChunks = (160 * 1024 - 512) / (16 * kChunk);   // 319
constexpr int kBitWords = (kChunks + 63) / 64;                 // 5

__device__ __forceinline__ int chunks_for(int elems) { return (elems + kChunk - 1) / kChunk; }

// Bits [s, s+k) of word q (64 bits each) as a mask; 0 when disjoint.
__device__ __forceinline__ unsigned long long run_mask(int q, int s, int k)
{
    const int lo = s > 64 * q ? s - 64 * q : 0, hi = s + k < 64 * (q + 1) ? s + k - 64 * q : 64;
    if (hi <= lo) return 0ull;
    const unsigned long long top = hi == 64 ? ~0ull : ((1ull << hi) - 1ull);
    return top & ~((1ull << lo) - 1ull);
}

struct ChunkPool {
    unsigned long long *bits;
    unsigned *lock;
    unsigned *freed;   // count of frees: a waiting wavefront searches again only after one
    zd *arena;
    int lane;

    __device__ __forceinline__ bool try_lock() const
    {
        return (__builtin_amdgcn_readfirstlane(atomicOr(lock, lane == 0 ? 1u : 0u)) & 1u) == 0u;
    }
    __device__ __forceinline__ void unlock() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        atomicAnd(lock, lane == 0 ? ~1u : ~0u);
    }
    // Start of a run of k free chunks (1 <= k <= kChunks): the lowest
    // (top = false) or the highest; -1 if none.  Chunks [xs, xs + xk) count
    // as free (the caller's own run, when it looks for a place to compact
    // into); a lowest-run search looks at starts below 64 (jmax + 1) only.
    // Lane l tests the start 64 j + l of start word j: the 64 bits from
    // there (one funnel shift of two bit words; the word index is uniform)
    // hold a run of at least k when their count of trailing ones is >= k,
    // and for k > 64 the next 64 bits are tested the same way.
    __device__ __forceinline__ int find(int k, bool top, int xs = 0, int xk = 0, int jmax = kBitWords - 1) const
    {
        unsigned long long w[kBitWords + 2];
#pragma unroll
        for (int q = 0; q < kBitWords; ++q)
            w[q] = __hip_atomic_load(&bits[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) | run_mask(q, xs, xk);
        w[kBitWords] = w[kBitWords + 1] = 0ull;   // past the arena: not free
        int found = -1;
#pragma unroll
        for (int jj = 0; jj < kBitWords; ++jj) {
            const int j = top ? kBitWords - 1 - jj : jj;
            if (found >= 0 || (!top && j > jmax)) break;
            auto bits64 = [&](int q) {   // 64 bits from chunk 64 q + lane
                return lane == 0 ? w[q] : (w[q] >> lane) | (w[q + 1] << (64 - lane));
            };
            // trailing ones of x (64 when all are ones; the count's argument is
            // never 0: bit 63 set in it changes no count below 64)
            auto ones = [](unsigned long long x) {
                return x == ~0ull ? 64 : (int)__builtin_ctzll(~x | (1ull << 63));
            };
            const unsigned long long x0 = bits64(j);
            bool ok;
            if (k <= 64) {
                ok = ones(x0) >= k;
            } else {
                const unsigned long long x1 = bits64(j + 1);
                ok = x0 == ~0ull && ones(x1) >= k - 64;
            }
            const unsigned long long b = __ballot(ok);
            if (b != 0ull) found = 64 * j + (top ? 63 - __clzll((long long)b) : __ffsll((long long)b) - 1);
        }
        return __builtin_amdgcn_readfirstlane(found);
    }
    // Whether any chunk below chunk s is free (a snapshot).
    __device__ __forceinline__ bool free_below(int s) const
    {
        bool any = false;
#pragma unroll
        for (int q = 0; q < kBitWords; ++q) {
            const unsigned long long w = __hip_atomic_load(&bits[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            any = any || (w & run_mask(q, 0, s)) != 0ull;
        }
        return __builtin_amdgcn_readfirstlane(any ? 1 : 0) != 0;
    }
    // Marks chunks [s, s+k) used (lanes 0..kBitWords-1 each change one word).
    __device__ __forceinline__ void take(int s, int k) const
    {
        const int q = lane % kBitWords;
        atomicAnd(&bits[q], lane < kBitWords ? ~run_mask(q, s, k) : ~0ull);
    }
    __device__ __forceinline__ void give(int s, int k) const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        const int q = lane % kBitWords;
        atomicOr(&bits[q], lane < kBitWords ? run_mask(q, s, k) : 0ull);
        atomicAdd(freed, lane == 0 ? 1u : 0u);
    }
    __device__ __forceinline__ unsigned frees() const
    {
        return __builtin_amdgcn_readfirstlane(__hip_atomic_load(freed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
    }
    // Sleeps until a run is freed after `seen` frees, or for kPolls polls
    // (a search can also fail for want of the lock): a waiting wavefront
    // issues one LDS read per ~500 cycles instead of a search, leaving the
    // SIMD to the wavefronts that hold matrices.
    static constexpr int kPolls = 16;
    __device__ __forceinline__ void wait_free(unsigned seen) const
    {
        for (int p = 0; p < kPolls && frees() == seen; ++p) __builtin_amdgcn_s_sleep(8);
    }
    // The highest (top) or lowest run of k free chunks, claimed, or -1.  A
    // search of a snapshot of the bit words comes first, without the lock,
    // so wavefronts that find no room never take it (a mover that wants a
    // smaller run then seldom finds the lock held by waiting wavefronts);
    // then the lock is tried at most kLockTries times (its holder runs one
    // search and claim) for a second search and the claim.
    static constexpr int kLockTries = 8;
    __device__ __forceinline__ int try_alloc(int k, bool top) const
    {
        if (find(k, top) < 0) return -1;
        for (int r = 0; r < kLockTries; ++r) {
            if (try_lock()) {
                const int s = find(k, top);
                if (s >= 0) take(s, k);
                unlock();
                return s;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return -1;
    }
    // The lowest run of k chunks, waiting until one is free.
    __device__ __forceinline__ int alloc_wait(int k) const
    {
        for (;;) {
            const unsigned seen = frees();
            const int s = try_alloc(k, false);
            if (s >= 0) return s;
            wait_free(seen);
        }
    }
    __device__ __forceinline__ int try_alloc_top(int k) const { return try_alloc(k, true); }
    // The lowest run of `need` chunks when the caller's run [s, s + k) counts
    // as free, claimed (the caller's own chunks in it stay claimed), or -1
    // when that is the caller's run itself with need == k or the lock was
    // not had in kLockTries tries.  A run less than min_shift chunks below s
    // is replaced by s itself (shrinking in place).  The result is <= s (the
    // caller's run holds `need` chunks); the caller copies its data down and
    // gives back the chunks of its run that the new one does not cover.
    __device__ __forceinline__ int compact(int need, int s, int k, int min_shift) const
    {
        const int jmax = s / 64;   // the result is <= s
        auto useful = [&](int r) { return r >= 0 && (r <= s - min_shift || (r == s && need < k)); };
        int s0 = find(need, false, s, k, jmax);
        if (s0 > s - min_shift) s0 = s;   // a shorter move down is not worth its copy: shrink in place
        if (!useful(s0)) return -1;
        for (int r = 0; r < kLockTries; ++r) {
            if (try_lock()) {
                int s2 = find(need, false, s, k, jmax);
                if (s2 > s - min_shift) s2 = s;
                const bool move = useful(s2);
                if (move) take(s2, need);
                unlock();
                return move ? s2 : -1;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return -1;
    }
};

// zlahqr_dual's hook for a matrix in a chunk run, after every deflation
// that leaves columns 0..i (and the idle lanes' element) active.  The
// eigenvalues of the rows that became final (i+1 .. whi-1) are written from
// the old run before any of it is given back.
//  MV = 0 (compaction): the active part moves to the lowest run of free
//    chunks that holds it, the matrix's own run counting as free
//    (ChunkPool::compact): in place when nothing below is free (the run's
//    tail chunks are given back, no lock), else copied down.  The arena
//    stays compacted at its low end, where new tiles start (a simulation
//    of one CU at n = 64: 7.3 resident matrices against 5.8 for MV = 5).
//  MV = 4..7: when the active part needs at most MV/8 of the run, it moves
//    to the highest free run that holds it, if any (cheaper per deflation;
//    faster where the CU's wavefronts, not its LDS, are the limit).
template <int MV>
struct ChunkMover {
    static constexpr int kMinShift = 1;   // any lower run (4 was 1-2% slower at n = 64)
    const ChunkPool *P;
    zd *out;
    int s, k;     // current run
    int whi;      // rows whi.. of out are written
    zd *H;
    double us;    // norm-range factor of the eigenvalues

    __device__ __forceinline__ zd eig(zd h) const { return us != 1.0 ? zscale(us, h) : h; }

    __device__ __forceinline__ void moved(zd *&Hq, int &idle, int i, int t)
    {
        const int ps = packed_size(i + 1);
        const int need = chunks_for(ps + 1);
        int s2;
        if constexpr (MV == 0) {
            // nothing free below the run: shrink in place (giving back the
            // run's own tail needs no lock); else the lowest run that fits
            const bool below = P->free_below(s);
            if (!below && need == k) return;
            s2 = below ? P->compact(need, s, k, kMinShift) : s;
        } else {
            if (8 * need > MV * k) return;
            s2 = P->try_alloc_top(need);
        }
        if (s2 < 0) return;
        const PackedLayout pk;
        if (i + 1 + t < whi) out[i + 1 + t] = eig(Hq[pk(i + 1 + t, i + 1 + t)]);   // whi <= n <= 64
        whi = i + 1;
        zd *H2 = P->arena + s2 * kChunk;
        if constexpr (MV == 0) {
            if (s2 != s) {
                // s2 < s: copy down, lowest elements first, in blocks whose
                // loads all precede their stores (a block writes only
                // elements that this or an earlier block has read; the
                // wavefront's LDS operations run in order)
                // (loads past the prefix are clamped to its last element; the
                // values are held as doubles, which stay in registers across
                // the compiler barriers)
                for (int e0 = 0; e0 < ps; e0 += 4 * 64) {
                    const int e1 = e0 + t, e2 = e1 + 64, e3 = e1 + 128, e4 = e1 + 192;
                    const zd *q1 = Hq + (e1 < ps ? e1 : ps - 1), *q2 = Hq + (e2 < ps ? e2 : ps - 1);
                    const zd *q3 = Hq + (e3 < ps ? e3 : ps - 1), *q4 = Hq + (e4 < ps ? e4 : ps - 1);
                    const double r1 = q1->re, i1 = q1->im, r2 = q2->re, i2 = q2->im;
                    const double r3 = q3->re, i3 = q3->im, r4 = q4->re, i4 = q4->im;
                    lds_order();
                    if (e1 < ps) H2[e1] = mk(r1, i1);
                    if (e2 < ps) H2[e2] = mk(r2, i2);
                    if (e3 < ps) H2[e3] = mk(r3, i3);
                    if (e4 < ps) H2[e4] = mk(r4, i4);
                    lds_order();
                }
            }
            // the chunks of the old run outside the new one
            const int lo = s2 + need > s ? s2 + need : s;
            if (lo < s + k) P->give(lo, s + k - lo);
        } else {   // a separate run: copy, then give the whole old run
            for (int e = t; e < ps; e += 64) H2[e] = Hq[e];
            P->give(s, k);
        }
        lds_order();
        s = s2;
        k = need;
        Hq = H2;
        H = H2;
        idle = need * kChunk - 1;
    }
};

}  // namespace tkz
