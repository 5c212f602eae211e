// The same-size eigenvalue call for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc): tokura_zgeev_batched_noalloc_gpu takes no `work`; the
// kernels' scratch (the packed Hessenberg matrices, the pool variants' work
// counter) lives in `scratch`, a fixed array in device memory that each GPU
// holds from the library's load until the process ends.
//
// Host code never passes `scratch` itself to a kernel (in host code the name
// is a host-side stand-in, not a device address): the device address of the
// array on each device comes from hipGetSymbolAddress, looked up once per
// device (tokuraCreate does it for its device, the call itself for a device
// that no handle was created on) and kept in g_scratch_dev.
//
// Calls on one GPU share its array: two calls that overlap in time on one GPU
// (on different streams, or from different host threads) race on it.  Nothing
// in it has to start at zero: the Hessenberg kernel of every call writes the
// packed matrices and zeroes the pool counter before the QR kernel reads them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/tokura_blas.h"
#include "tokura_internal.hpp"
#include "zgeev_device.hpp"
#include "zgeev_launch.hpp"

namespace {

// 128 MiB: the packed Hessenberg matrices of batch matrices of size n take
// 16 * packed_size(n) bytes each (34,320 at n = 64, 2,416 at n = 16), then the
// counter (256 bytes reserved).
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The device address of `scratch` on `dev` (the current device), looked up on
// the first use; nullptr when the lookup fails.
void *scratch_on(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

}  // namespace

// Called by tokuraCreate with the handle's device current.
void noalloc_setup(int dev) { (void)scratch_on(dev); }

extern "C" int tokura_zgeev_batched_noalloc_gpu(tokurablas_t *handle, int n, int batch_count,
                                                const hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                                                char *flag, hipStream_t stream)
{
    if (handle == nullptr) return -1;
    if (!(n > 0 && n <= TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE)) return -2;
    if (A == nullptr) return -3;
    if (eigenvalues == nullptr) return -4;
    if (!(batch_count > 0)) return -6;
    if (flag == nullptr) return -7;
    const size_t need = sizeof(hipDoubleComplex) * (size_t)tkz::packed_size(n) * (size_t)batch_count + 256;
    if (need > kScratchBytes) return -9;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    void *work = scratch_on(dev);
    if (work == nullptr) return 1;
    int v = handle->variant_for_n[n];
    const tkz::VariantInfo *vi = tkz::variant_info(v);
    if (vi == nullptr || n > vi->nb) return -2;
    int stage = 1;
    const hipError_t e = tkz::launch_plain(v, n, batch_count, A, eigenvalues, flag, work, handle->max_sweeps,
                                           handle->balance, stream, &stage);
    return e == hipSuccess ? 0 : stage;
}
