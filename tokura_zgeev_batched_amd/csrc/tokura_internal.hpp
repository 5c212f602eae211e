// Internal definition of the opaque handle shared by the host-side sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <array>
#include <vector>

#include "../../include/tokura_blas.h"

struct RaggedGraphCache {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    const void *A = nullptr;
    void *eig = nullptr, *work = nullptr;
    char *flag = nullptr;
    int batch = -1;
    int cfg = -1;
    int max_sweeps = -1;
    int balance = 0;
    int gen = 0;
};

struct RaggedPrepared {
    void *work;
    int cfg;
    int hmul;     // small-class Hessenberg grid: CUs x resident blocks x hmul workgroups
    int batch;
    int gen;      // the prepare that wrote the schedule (also in its header)
    int n49_64;   // matrices with 49 <= n <= 64 (size classes 0-2)
    int n33_48;   // matrices with 33 <= n <= 48 (size classes 3-4)
};

// A tuned ragged configuration for batches with one size signature
// (ragged_signature: the batch's share of each size class in sixteenths and
// log2 of its size).
struct RaggedTune {
    unsigned long long sig;
    int cfg;
    int hmul;
};

struct tokurablas_t {
    tokurablas_t();
    int device;
    int cu_count;
    char arch[64];
    int max_sweeps;                                           // < 0: LAPACK cap
    int balance = 0;                                          // ZGEBAL job bits (1 permute, 2 scale)
    int variant_for_n[TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE + 1];  // same-size path
    int ragged_cfg = 0;                                       // ragged class configuration
    int ragged_hmul = 4;                                      // small-class Hessenberg grid multiplier
    bool ragged_auto = false;                                 // prepare looks up ragged_tuned first
    std::vector<RaggedTune> ragged_tuned;                     // from tokuraAutotuneRagged / a cache
    RaggedGraphCache graph;
    void *ragged_host_stage = nullptr;                        // pinned staging of the schedule
    size_t ragged_host_bytes = 0;
    hipEvent_t ragged_stage_event = nullptr;                  // after the last staging copy
    hipStream_t ragged_side[3] = {nullptr, nullptr, nullptr}; // branches of the per-class launches
    hipEvent_t ragged_fork_ev[4] = {nullptr, nullptr, nullptr, nullptr};  // fork, then one join per branch
    bool ragged_staged = false;                               // that copy may still be in flight
    int ragged_gen = 0;                                       // prepares so far
    std::vector<RaggedPrepared> prepared;
    // stage timing: events (start, between stages, end) per timed call
    bool stage_timing = false;
    std::vector<std::array<hipEvent_t, 3>> timed_calls;
    std::vector<std::array<hipEvent_t, 3>> event_pool;
    int timed_split = 0;
    std::array<hipEvent_t, 3> *timing_begin(hipStream_t s);
    void timing_end(std::array<hipEvent_t, 3> *ev, hipStream_t s);
};

// Resident workgroups of kernel `k` per CU at `threads` threads per block
// (1 when the query fails or reports 0), for sizing persistent grids.
template <class K>
int resident_blocks(K k, int threads = 64)
{
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, threads, 0) != hipSuccess) return 1;
    return occ < 1 ? 1 : occ;
}

void ragged_release(tokurablas_t *h);
// The device key of a tuning cache's lines: the handle's architecture name
// ("none" for a handle made without a device) and its CU count.
inline const char *tuning_arch(const tokurablas_t *h) { return h->arch[0] ? h->arch : "none"; }
// Writes the ragged lines of a tuning cache ("ragged arch cus sig cfg hmul");
// parses one such line into the handle (returns 1 if it was taken).
void ragged_write_tuning(const tokurablas_t *h, FILE *f);
int ragged_parse_tuning(tokurablas_t *h, const char *line);
// Looks up the no-allocation entry's scratch array on device `dev` (current).
void noalloc_setup(int dev);
int ragged_num_configs();
