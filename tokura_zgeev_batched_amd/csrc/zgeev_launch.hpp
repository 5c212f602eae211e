// Internal interface between the HIP kernels and the host-side C ABI layer.
#pragma once
#include <hip/hip_runtime.h>

namespace tkz {

// VK_DUAL: split path whose QR runs two bulges per sweep (a 256-thread
// workgroup Hessenberg for NB >= 48); VK_DUAL_1W: the same QR after the
// one-wavefront Hessenberg (the arithmetic of the Schur-form kernels, so its
// eigenvalues come in their order)
// VK_POOL: the 256-thread Hessenberg, then the two-bulge QR on a persistent
// workgroup per CU whose LDS is an arena of slots that matrices move down
// through as they deflate (its eigenvalues are VK_DUAL's bit for bit)
// VK_CHUNK: the same two stages with the QR workgroup's LDS cut into chunks
// (the ragged path's arena) that the deflating matrices compact at every
// deflation (VK_DUAL's eigenvalues bit for bit)
enum VariantKind { VK_FUSED = 0, VK_SPLIT = 1, VK_DUAL = 2, VK_DUAL_1W = 3, VK_POOL = 4, VK_CHUNK = 5 };

// ZGEBAL job bits of tokuraSetBalancing ('P' = permute, 'S' = scale, 'B' = both)
enum BalanceBits { BAL_PERMUTE = 1, BAL_SCALE = 2 };

// One compiled launch configuration of the same-size path.
//   nb   largest n the LDS allocation holds
//   lpm  lanes per matrix (a wavefront holds 64/lpm matrices)
//   mpb  matrices per workgroup (one wavefront per workgroup)
struct VariantInfo {
    int id, kind, nb, lpm, mpb;
};

int num_variants();
const VariantInfo *variant_info(int v);

// Launches variant `variant` on `s`; *stage reports 1 (Hessenberg) or 2 (QR)
// for the launch that failed.  bal: ZGEBAL job bits (1 permute, 2 scale),
// 0 for none.
// When `mid` is non-null it is recorded on `s` between the Hessenberg and the
// QR kernel of a split variant (fused variants record nothing).
hipError_t launch_plain(int variant, int n, int batch, const void *A, void *eig, char *flag, void *work,
                        int itmax, int bal, hipStream_t s, int *stage, hipEvent_t mid = nullptr);

}  // namespace tkz
