// Host side of the C ABI declared in include/tokura_blas.h.
//
// Reference counterparts: handle life cycle tokura_zgeev_batched.cu:64-190,
// workspace sizing :192-210, argument validation and stage dispatch
// :212-285, per-n launch tables loaded from generated headers :13-60 and the
// compile-time tuner TOKURA_ZHSEQR_BATCHED_GEHRD_TUNING_main.cu:89-242.
// Here the launch table is a runtime object in the handle: defaults from a
// size heuristic, optionally replaced by tokuraAutotune / a cache file keyed
// by the device (gfx arch name and CU count).
#include <hip/hip_runtime.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>
#include <unistd.h>

#include "../../include/tokura_blas.h"
#include "tokura_internal.hpp"
#include "zgeev_launch.hpp"

using namespace tkz;

namespace {

// Default variant per n before any tuning (ids of zgeev_kernels.hip's table).
int default_variant(int n)
{
    // what tokuraAutotune picks on an MI355X at batches of 20,000 and 50,000
    if (n <= 4) return 0;    // fused, 4 lanes/matrix, 16 matrices/wave
    if (n <= 8) return 28;   // split 8 x 4: packed Hessenberg to the workspace, then QR
    if (n <= 16) return 27;  // split 16 x 8
    if (n <= 32) return 14;  // split 32 x 32
    if (n <= 40) return 24;  // split, one matrix per wavefront, two bulges per sweep, LDS for n = 40
    if (n <= 48) return 19;  // the same with LDS for n = 48
    if (n <= 56) return 23;  // n = 56
    // n > 56: the chunk variant with 8 wavefronts (one persistent workgroup
    // per CU whose LDS the deflating matrices keep compacted): 5.6% faster
    // than the best pool layout at n = 64 x 100,000; the fixed 34 KB buffers
    // of variant 20 hold one matrix per SIMD at n = 64
    return 37;
}

bool variant_fits(int v, int n)
{
    const VariantInfo *vi = variant_info(v);
    return vi && n <= vi->nb;
}

}  // namespace

tokurablas_t::tokurablas_t()
{
    device = 0;
    cu_count = 0;
    arch[0] = '\0';
    max_sweeps = -1;
    for (int n = 0; n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE; ++n) variant_for_n[n] = default_variant(n < 1 ? 1 : n);
}

extern "C" int tokuraCreate(tokurablas_t **handle)
{
    if (handle == nullptr) return -1;
    tokurablas_t *h = new (std::nothrow) tokurablas_t();
    if (h == nullptr) return -1;
    *handle = h;
    if (hipGetDevice(&h->device) != hipSuccess) return 1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) return 1;
    h->cu_count = prop.multiProcessorCount;
    noalloc_setup(h->device);   // the scratch array's address on this device (no allocation)
    std::snprintf(h->arch, sizeof(h->arch), "%s", prop.gcnArchName);
    if (char *p = std::strchr(h->arch, ':')) *p = '\0';
    if (const char *cache = std::getenv("TOKURA_TUNE_CACHE")) tokuraLoadTuning(h, cache);
    return 0;
}

extern "C" int tokuraDestroy(tokurablas_t *handle)
{
    if (handle == nullptr) return 0;
    ragged_release(handle);
    for (auto &ev : handle->timed_calls)
        for (auto &x : ev) (void)hipEventDestroy(x);   // teardown: nothing to undo on failure
    for (auto &ev : handle->event_pool)
        for (auto &x : ev) (void)hipEventDestroy(x);
    delete handle;
    return 0;
}

extern "C" size_t tokura_get_zgeeveigenvaluesgetworspacesize(int n, int batch_count)
{
    if (!(0 < n && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return (size_t)-1;
    if (!(0 < batch_count)) return (size_t)-2;
    return sizeof(hipDoubleComplex) * (size_t)n * (size_t)n * (size_t)batch_count +
           sizeof(hipDoubleComplex) * (size_t)n * (size_t)batch_count;
}

extern "C" int tokura_zgeev_batched_gpu(tokurablas_t *handle, int n, int batch_count, hipDoubleComplex *A,
                                        hipDoubleComplex *eigenvalues, hipDoubleComplex *work, char *flag,
                                        hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (work == nullptr) return -5;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    int v = handle->variant_for_n[n];
    if (!variant_fits(v, n)) v = default_variant(n);
    int stage = 1;
    std::array<hipEvent_t, 3> *ev = handle->stage_timing ? handle->timing_begin(stream) : nullptr;
    hipError_t e = launch_plain(v, n, batch_count, A, eigenvalues, flag, work, handle->max_sweeps, handle->balance,
                                stream, &stage, ev ? (*ev)[1] : nullptr);
    if (ev) {   // stage timing is diagnostic: a failed record only spoils that reading
        if (variant_info(v)->kind == VK_FUSED) (void)hipEventRecord((*ev)[1], stream);
        handle->timing_end(ev, stream);
    }
    if (e != hipSuccess) return stage;
    return 0;
}

// Events for one timed call (nullptr, and the call goes untimed, when they
// cannot be created or recorded).
std::array<hipEvent_t, 3> *tokurablas_t::timing_begin(hipStream_t s)
{
    std::array<hipEvent_t, 3> ev{};
    if (!event_pool.empty()) {
        ev = event_pool.back();
        event_pool.pop_back();
    } else {
        for (int k = 0; k < 3; ++k)
            if (hipEventCreate(&ev[k]) != hipSuccess) {
                for (int j = 0; j < k; ++j) (void)hipEventDestroy(ev[j]);
                return nullptr;
            }
    }
    if (hipEventRecord(ev[0], s) != hipSuccess) {
        event_pool.push_back(ev);
        return nullptr;
    }
    timed_calls.push_back(ev);
    return &timed_calls.back();
}

void tokurablas_t::timing_end(std::array<hipEvent_t, 3> *ev, hipStream_t s) { (void)hipEventRecord((*ev)[2], s); }

extern "C" int tokuraSetStageTiming(tokurablas_t *handle, int enable)
{
    if (handle == nullptr) return -1;
    handle->stage_timing = enable != 0;
    return 0;
}

extern "C" int tokuraGetStageTimes(tokurablas_t *handle, int *calls, float *hessenberg_ms, float *qr_ms)
{
    if (handle == nullptr) return -1;
    float s1 = 0.f, s2 = 0.f;
    int rc = 0;
    for (auto &ev : handle->timed_calls) {
        if (hipEventSynchronize(ev[2]) != hipSuccess) rc = 1;
        float a = 0.f, b = 0.f;
        if (hipEventElapsedTime(&a, ev[0], ev[1]) != hipSuccess || hipEventElapsedTime(&b, ev[1], ev[2]) != hipSuccess)
            rc = 1;
        s1 += a;
        s2 += b;
        handle->event_pool.push_back(ev);
    }
    if (calls) *calls = (int)handle->timed_calls.size();
    if (hessenberg_ms) *hessenberg_ms = s1;
    if (qr_ms) *qr_ms = s2;
    handle->timed_calls.clear();
    return rc;
}

extern "C" int tokuraSetMaxSweeps(tokurablas_t *handle, int max_sweeps)
{
    if (handle == nullptr) return -1;
    handle->max_sweeps = max_sweeps < 0 ? -1 : max_sweeps;
    return 0;
}

extern "C" int tokuraSetBalancing(tokurablas_t *handle, char job)
{
    if (handle == nullptr) return -1;
    switch (job) {
    case 'N': case 'n': handle->balance = 0; return 0;
    case 'P': case 'p': handle->balance = BAL_PERMUTE; return 0;
    case 'S': case 's': handle->balance = BAL_SCALE; return 0;
    case 'B': case 'b': handle->balance = BAL_PERMUTE | BAL_SCALE; return 0;
    default: return -2;
    }
}

extern "C" int tokuraGetLaunchConfig(tokurablas_t *handle, int n, int *variant, int *lanes_per_matrix,
                                     int *matrices_per_block)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    int v = handle->variant_for_n[n];
    if (!variant_fits(v, n)) v = default_variant(n);
    const VariantInfo *vi = variant_info(v);
    if (variant) *variant = v;
    if (lanes_per_matrix) *lanes_per_matrix = vi->lpm;
    if (matrices_per_block) *matrices_per_block = vi->mpb;
    return 0;
}

extern "C" int tokuraSetVariant(tokurablas_t *handle, int n, int variant)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (!variant_fits(variant, n)) return -2;
    handle->variant_for_n[n] = variant;
    return 0;
}

extern "C" int tokuraLoadTuning(tokurablas_t *handle, const char *cache_path)
{
    if (handle == nullptr) return -1;
    if (cache_path == nullptr) return -3;
    FILE *f = std::fopen(cache_path, "r");
    if (!f) return -3;
    char line[256];
    int count = 0;
    while (std::fgets(line, sizeof line, f)) {
        char arch[64];
        int cus, n, v;
        if (line[0] == '#') continue;
        if (std::strncmp(line, "ragged ", 7) == 0) {   // a ragged tuning-table entry
            count += ragged_parse_tuning(handle, line);
            continue;
        }
        if (std::sscanf(line, "%63s %d %d %d", arch, &cus, &n, &v) != 4) continue;
        if (std::strcmp(arch, tuning_arch(handle)) != 0 || cus != handle->cu_count) continue;
        if (n < 1 || n > TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE || !variant_fits(v, n)) continue;
        handle->variant_for_n[n] = v;
        ++count;
    }
    std::fclose(f);
    return count;
}

extern "C" int tokuraAutotune(tokurablas_t *handle, int n_lo, int n_hi, int batch_count, const char *cache_path,
                              hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (n_lo < 1) n_lo = 1;
    if (n_hi > TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE) n_hi = TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE;
    if (n_lo > n_hi || batch_count <= 0) return -2;
    const size_t maxn = (size_t)n_hi;
    const size_t nA = maxn * maxn * (size_t)batch_count;
    hipDoubleComplex *dA = nullptr, *dE = nullptr, *dW = nullptr;
    char *dF = nullptr;
    if (hipMalloc(&dA, nA * sizeof(hipDoubleComplex)) != hipSuccess) return 3;
    if (hipMalloc(&dE, maxn * batch_count * sizeof(hipDoubleComplex)) != hipSuccess ||
        hipMalloc(&dW, tokura_get_zgeeveigenvaluesgetworspacesize((int)maxn, batch_count)) != hipSuccess ||
        hipMalloc(&dF, (size_t)batch_count) != hipSuccess) {
        (void)hipFree(dA); (void)hipFree(dE); (void)hipFree(dW); (void)hipFree(dF);
        return 3;
    }
    // seeded N(0,1) entries, the distribution of the benchmark workloads
    std::vector<hipDoubleComplex> hA(nA);
    std::mt19937_64 rng(12345);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (auto &z : hA) z = make_hipDoubleComplex(nd(rng), nd(rng));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMemcpyAsync(dA, hA.data(), nA * sizeof(hipDoubleComplex), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = 4;
    for (int n = n_lo; n <= n_hi && rc == 0; ++n) {
        float best = 1e30f;
        int bestv = default_variant(n);
        for (int v = 0; v < num_variants(); ++v) {
            if (!variant_fits(v, n)) continue;
            // skip variants whose LDS capacity is more than twice what n needs
            // (they are never faster than a smaller-NB variant of the same lpm)
            if (variant_info(v)->nb > 2 * ((n + 3) & ~3) && variant_info(v)->nb > 8) continue;
            int stage;
            float tms = 0.f;
            for (int rep = 0; rep < 3; ++rep) {
                float ms = 0.f;
                if (hipEventRecord(e0, stream) != hipSuccess ||
                    launch_plain(v, n, batch_count, dA, dE, dF, dW, -1, handle->balance, stream, &stage) != hipSuccess ||
                    hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = 4;
                    break;
                }
                if (rep > 0) tms += ms;   // the first run warms up
            }
            if (rc) break;
            if (tms < best) { best = tms; bestv = v; }
        }
        handle->variant_for_n[n] = bestv;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dA); (void)hipFree(dE); (void)hipFree(dW); (void)hipFree(dF);
    if (rc == 0 && cache_path != nullptr && tokuraSaveTuning(handle, cache_path) != 0) return 5;
    return rc;
}

// The file keeps the lines of other devices (another architecture or CU
// count, same-size or ragged) and gets this handle's lines in place of its
// old ones; it is written to a temporary file beside it and renamed over it,
// so a reader never sees a half-written cache.
extern "C" int tokuraSaveTuning(tokurablas_t *handle, const char *cache_path)
{
    if (handle == nullptr) return -1;
    if (cache_path == nullptr) return -3;
    const char *key = tuning_arch(handle);
    std::vector<std::string> others;
    if (FILE *old = std::fopen(cache_path, "r")) {
        char line[256];
        while (std::fgets(line, sizeof line, old)) {
            if (line[0] == '#' || std::strchr(line, '\n') == nullptr) continue;   // comments, over-long lines
            const char *p = std::strncmp(line, "ragged ", 7) == 0 ? line + 7 : line;
            char arch[64];
            int cus;
            if (std::sscanf(p, "%63s %d", arch, &cus) != 2) continue;
            if (std::strcmp(arch, key) == 0 && cus == handle->cu_count) continue;   // replaced below
            others.emplace_back(line);
        }
        std::fclose(old);
    }
    const std::string tmp = std::string(cache_path) + ".tmp" + std::to_string((long long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return 5;
    std::fprintf(f, "# arch cus n variant  (tokura_zgeev_batched_amd autotune cache)\n");
    std::fprintf(f, "# ragged arch cus signature config hess_grid_multiplier\n");
    for (const auto &ln : others) std::fputs(ln.c_str(), f);
    for (int n = 1; n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE; ++n)
        std::fprintf(f, "%s %d %d %d\n", key, handle->cu_count, n, handle->variant_for_n[n]);
    ragged_write_tuning(handle, f);
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), cache_path) != 0) {
        std::remove(tmp.c_str());
        return 5;
    }
    return 0;
}
