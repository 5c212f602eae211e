/*
 * tokura_blas.h — C ABI of the MI355X-native batched complex eigenvalue solver.
 *
 * Drop-in for the reference library's public interface
 * (tokura_zgeev_batched_library/include/tokura_blas.h:129-151, documented in
 * README.md:25-79).  All four reference entry points keep their names,
 * argument order, argument meaning, memory layouts and return codes:
 *
 *   matrices  A[k*n*n + r + c*n]  (column-major, one matrix after another)
 *   output    eigenvalues[i + j*n]
 *   flag[j]   0 when the eigenvalues of matrix j converged, non-zero otherwise
 *   stream    NULL selects the default stream; nothing synchronises the host
 *
 * Types: hipDoubleComplex has the layout of the reference's cuDoubleComplex
 * ({double x, y}, 16-byte aligned); hipStream_t replaces cudaStream_t.
 *
 * The library allocates no device memory on the forward paths: callers own
 * A, eigenvalues, work and flag.  A is read, not clobbered (the reference
 * clobbers it; callers that rely on README.md:69 "do not use A" are unaffected).
 *
 * Flag codes (the reference defines only 0 / non-zero, README.md:72):
 *   0  converged
 *   1  iteration cap reached (30*max(10,n) QR sweeps per eigenvalue, LAPACK's
 *      cap; the reference uses n^4 total)
 *   2  non-finite entry in the input matrix (eigenvalues are NaN)
 *   3  (recompute backward only) the forward's eigenvalues are not those of
 *      the matrix, so no cotangent can be paired (dA is NaN)
 *   4  (ragged entries only) `work` does not hold the schedule prepared for
 *      the call (nothing computed)
 *
 * One handle must not be used by two host threads at once (the reference
 * keeps per-call state in the handle, tokura_zgeev_batched.cu:224-235).
 */
#ifndef TOKURA_BLAS_MI355X_H
#define TOKURA_BLAS_MI355X_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>
#include <hip/hip_complex.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOKURA_ZGEEV_BATCHED_MAX_MATRIX_SIZE 64

/* Opaque handle.  The reference exposes its struct (tokura_blas.h:59-100) but
 * no caller outside its build-time tuner reads a field. */
typedef struct tokurablas_t tokurablas_t;

/* ---------------------------------------------------------------------------
 * Reference entry points
 * ------------------------------------------------------------------------- */

/* Replaces tokuraCreate (tokura_zgeev_batched.cu:64-160).  Allocates the host
 * handle, queries the current HIP device and loads the autotuned launch table
 * for it.  Returns 0 on success, -1 when the handle cannot be allocated, 1
 * when no HIP device can be queried (as the reference returns 1 when
 * cudaGetDeviceProperties fails; *handle is still set and must be destroyed). */
int tokuraCreate(tokurablas_t **handle);

/* Replaces tokuraDestroy (tokura_zgeev_batched.cu:161-190).  NULL is accepted. */
int tokuraDestroy(tokurablas_t *handle);

/* Replaces tokura_get_zgeeveigenvaluesgetworspacesize
 * (tokura_zgeev_batched.cu:192-210): 16*(n*n*batch + n*batch) bytes,
 * (size_t)-1 when n is outside [1,64], (size_t)-2 when batch_count <= 0. */
size_t tokura_get_zgeeveigenvaluesgetworspacesize(int n, int batch_count);

/* Replaces tokura_zgeev_batched_gpu (tokura_zgeev_batched.cu:212-285).
 * Eigenvalues of batch_count n x n complex matrices, n in [1,64].
 * Returns 0 on success; -1 handle NULL, -2 n out of range, -3 A NULL,
 * -4 eigenvalues NULL, -5 work NULL, -6 batch_count <= 0, -7 flag NULL (the
 * reference's validation order); 1 Hessenberg stage launch failure, 2 QR stage
 * launch failure.  Asynchronous on `stream`. */
int tokura_zgeev_batched_gpu(tokurablas_t *handle, int n, int batch_count,
                             hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                             hipDoubleComplex *work, char *flag, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Extensions (no counterpart in the reference)
 * ------------------------------------------------------------------------- */

/* No counterpart in the reference.  tokura_zgeev_batched_gpu without `work`,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * kernels, with their scratch in a fixed array of 128 MiB that the library
 * keeps in each GPU's memory until the process ends; this call allocates
 * nothing.  Limit: batch_count * 8 * ((n-1)(n+2) + 2n) + 256 <= 134217728
 * bytes (3,914 matrices at n = 64, 55,553 at n = 16), else -9.  Results,
 * flags and the other return codes as tokura_zgeev_batched_gpu.  Calls that
 * overlap in time on one GPU (other streams or host threads) share that
 * array and race. */
int tokura_zgeev_batched_noalloc_gpu(tokurablas_t *handle, int n, int batch_count,
                                     const hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                                     char *flag, hipStream_t stream);

/* Debug knob: cap on QR sweeps per eigenvalue (< 0 restores LAPACK's
 * 30*max(10,n)).  Lets tests force flag 1 from the host. */
int tokuraSetMaxSweeps(tokurablas_t *handle, int max_sweeps);

/* Launch configuration chosen for size n (variant id, lanes per matrix,
 * matrices per workgroup); for diagnostics and the autotuner's tests. */
int tokuraGetLaunchConfig(tokurablas_t *handle, int n, int *variant, int *lanes_per_matrix,
                          int *matrices_per_block);

/* Stage timing for profiling: while enabled, every same-size call records
 * HIP events on its stream before the first kernel, between the Hessenberg
 * and the QR kernel (split variants) and after the last kernel.
 * tokuraGetStageTimes waits for those events and returns the number of timed
 * calls and the summed milliseconds of each stage (a fused variant reports
 * its whole time as stage 2), then clears the record. */
int tokuraSetStageTiming(tokurablas_t *handle, int enable);
int tokuraGetStageTimes(tokurablas_t *handle, int *calls, float *hessenberg_ms, float *qr_ms);

/* Runtime autotuner (replaces the reference's compile-time tuner,
 * TOKURA_ZHSEQR_BATCHED_GEHRD_TUNING_main.cu:89-242, which rebuilds the .so).
 * For every n in [n_lo, n_hi] it times each eligible kernel variant on
 * `batch_count` seeded random matrices on `stream` (device buffers are
 * allocated and freed inside this call), keeps the fastest in the handle and,
 * when cache_path is non-NULL, writes the table there.  Returns 0 on success. */
int tokuraAutotune(tokurablas_t *handle, int n_lo, int n_hi, int batch_count,
                   const char *cache_path, hipStream_t stream);

/* Loads a table written by tokuraAutotune / tokuraSaveTuning (entries whose
 * device key — arch and CU count — differ are ignored): the same-size
 * variants and the ragged tuning table.  Returns the number of entries read
 * or a negative value on error.  tokuraCreate already loads the file named
 * by the environment variable TOKURA_TUNE_CACHE when it is set. */
int tokuraLoadTuning(tokurablas_t *handle, const char *cache_path);

/* Writes the handle's tuning (same-size variant per n, and the ragged table
 * that tokuraAutotuneRagged fills: one line per batch size signature) to
 * cache_path, keyed by the device (architecture name, "none" for a handle
 * made without a device, and CU count).  Lines of other devices already in
 * the file are kept; this device's old lines are replaced.  The file is
 * written beside cache_path and renamed over it.  Returns 0, -1 handle, -3
 * path NULL, 5 I/O error.  No counterpart in the reference, whose tables are
 * compiled in. */
int tokuraSaveTuning(tokurablas_t *handle, const char *cache_path);

/* Sets the variant used for size n by hand (a variant id as reported by
 * tokuraGetLaunchConfig).  Returns 0 or -2 for an invalid n / variant. */
int tokuraSetVariant(tokurablas_t *handle, int n, int variant);

/* ---- ragged batches: n varies per matrix --------------------------------
 * Matrix b is stored column-major at A + offsets[b] and is n_b x n_b with
 * n_b*n_b = offsets[b+1] - offsets[b], 1 <= n_b <= 64 (offsets are int64,
 * counted in complex elements, batch_count+1 entries, so matrix b+1 starts
 * where matrix b ends).  Eigenvalues of matrix b are written to
 * eigenvalues[e_b + i], e_b = n_0 + ... + n_{b-1}; flag[b] as above.
 *
 * tokura_get_zgeev_ragged_worksize: bytes of `work` for this batch ((size_t)-2
 * for batch_count <= 0, (size_t)-3 for invalid offsets).
 * tokura_zgeev_ragged_prepare reads the host copy of offsets, builds the
 * schedule (size-class tiles in decreasing size, eigenvalue offsets, packed
 * Hessenberg offsets) and copies it into `work` with an asynchronous copy on
 * `stream`.  The schedule stays valid for any number of solves of batches
 * with the same offsets.  It uses the class configuration that the handle's
 * ragged tuning table holds for the batch's size signature (after
 * tokuraAutotuneRagged or tokuraLoadTuning), else the handle's current one
 * (tokuraSetRaggedConfig).
 * tokura_zgeev_ragged_gpu launches, ordered after earlier work on `stream`
 * and before later work on it: a check of the schedule (one fill kernel),
 * the Hessenberg pass (two launches for n > 32, a workgroup per matrix with
 * the matrix in registers, or with balancing on five, one per LDS size; one
 * launch for n <= 32; without balancing the three run on the handle's side
 * streams, forked from and joined back to `stream` with events), and ONE
 * persistent QR kernel for the whole batch
 * (a workgroup per CU whose LDS is shared by matrices of every size; its
 * wavefronts pull tiles, largest first, from one atomic queue, and a matrix
 * moves to a smaller LDS region as it deflates); tokura_zgeev_ragged_graph
 * does the same through a hipGraph that the handle captures on first use and
 * replays while pointers, batch, class configuration, sweep cap and
 * schedule are unchanged.
 * Return codes as tokura_zgeev_batched_gpu, plus -8 when `work` has not been
 * prepared for this batch_count on this handle.  When `work` no longer holds
 * the schedule that prepare wrote (the buffer was freed and reallocated at
 * the same address, or written over), every matrix gets flag 4 and nothing
 * is computed. */
size_t tokura_get_zgeev_ragged_worksize(int batch_count, const int64_t *offsets_host);
int tokura_zgeev_ragged_prepare(tokurablas_t *handle, int batch_count, const int64_t *offsets_host,
                                void *work, size_t work_bytes, hipStream_t stream);
int tokura_zgeev_ragged_gpu(tokurablas_t *handle, int batch_count, const hipDoubleComplex *A,
                            hipDoubleComplex *eigenvalues, void *work, char *flag,
                            hipStream_t stream);
int tokura_zgeev_ragged_graph(tokurablas_t *handle, int batch_count, const hipDoubleComplex *A,
                              hipDoubleComplex *eigenvalues, void *work, char *flag,
                              hipStream_t stream);

/* Class configuration of the ragged path (0..15): lanes per matrix of class
 * A1 (n 49..64), A2 (33..48), B (17..32), C (1..16) (a wavefront holds
 * 64/lanes matrices of a class), wavefronts per workgroup of the persistent
 * QR (tokuraGetRaggedWaves), the deflating matrices' mover and, in 13..15,
 * a cap of eight wavefronts on n 49..64 while those lead the queue (the
 * other four take small tiles meanwhile).  tokuraAutotuneRagged times every
 * configuration on the caller's batch (buffers as for
 * tokura_zgeev_ragged_gpu; the stream is synchronised), then the small-class
 * Hessenberg grid (1, 2, 4 or 8 workgroups per resident slot,
 * tokuraGetRaggedHessGrid) with the fastest, keeps the winner, records it in
 * the handle's tuning table under the batch's size signature (each size
 * class's share of the batch and log2 of its size) and re-prepares `work`.
 * From then on, and after tokuraLoadTuning of a cache with such entries,
 * tokura_zgeev_ragged_prepare takes the configuration of a batch from the
 * table when its signature is there; tokuraSetRaggedConfig switches the
 * table off for the handle (a configuration set by hand wins). */
int tokuraSetRaggedConfig(tokurablas_t *handle, int config);
int tokuraGetRaggedConfig(tokurablas_t *handle, int *config, int *lanes_a1, int *lanes_a2, int *lanes_b,
                          int *lanes_c);
int tokuraGetRaggedWaves(tokurablas_t *handle, int *waves);
int tokuraGetRaggedHessGrid(tokurablas_t *handle, int *multiplier);
/* The configuration and Hessenberg grid multiplier of the schedule that the
 * last prepare wrote to `work` for batch_count matrices (-8 if none). */
int tokuraGetRaggedPrepared(tokurablas_t *handle, void *work, int batch_count, int *config, int *multiplier);
int tokuraAutotuneRagged(tokurablas_t *handle, int batch_count, const int64_t *offsets_host,
                         const hipDoubleComplex *A, hipDoubleComplex *eigenvalues, void *work,
                         size_t work_bytes, char *flag, hipStream_t stream);

/* ---- balancing ------------------------------------------------------------
 * No counterpart in the reference, which never balances (README.md:79).
 * Sets the LAPACK ZGEBAL job applied before the Hessenberg reduction by the
 * eigenvalue paths of this handle (same-size and ragged): 'N' none (the
 * default, as the reference), 'P' permute to isolate eigenvalues, 'S' scale
 * rows/columns by powers of two, 'B' both (LAPACK zgeev's choice).  The
 * eigenvalues are those of the balanced matrix, a similarity transform of A;
 * the Schur-form, backward and eigenvector entry points do not balance.
 * Returns 0, -1 (NULL handle) or -2 (unknown job). */
int tokuraSetBalancing(tokurablas_t *handle, char job);

/* ---- differentiation ----------------------------------------------------
 * Forward with Schur vectors: besides eigenvalues[i + j*n] and flag, writes
 * the Schur factorisation A_j = U_j T_j U_j^H of every matrix into `schur`
 * (T_j then U_j, each n*n column-major, 2*n*n complex per matrix; size from
 * tokura_get_zgeev_schur_worksize).  Eigenvalue i is T_j(i,i).
 *
 * Backward: for L = Re sum_ij conj(gbar[i + j*n]) * eigenvalues[i + j*n]
 * writes dA[j] = dL/dA_j = sum_i gbar_ij * y_i x_i^H / conj(y_i^H x_i)
 * (the convention of torch.linalg.eigvals' autograd), with x_i, y_i the right
 * and left eigenvectors, obtained by back-substitution on T_j and back-
 * transformed by U_j.  gbar is in the order of the forward's eigenvalues
 * (the Schur diagonal).  `schur` must not be NULL (-3); `A` is not read.
 * flag is the forward's; a matrix with flag != 0 gets dA = NaN.
 *
 * Recompute backward (memory-light mode, nothing kept from the forward but
 * its eigenvalues): tokura_zgeev_batched_backward_recompute_gpu rebuilds the
 * Schur form from A (which must be the forward's unmodified input) with
 * `work` of tokura_get_zgeeveigenvaluesgetworspacesize(n, batch) bytes as
 * scratch (T and the recomputed eigenvalues; U is parked in dA), its QR
 * shifts snapped to the nearest of `eigenvalues` (about half the sweeps of
 * the forward's QR; the Schur form, hence dA, equals the stored-Schur one up
 * to rounding), pairs each recomputed eigenvalue with the nearest of
 * `eigenvalues` (the forward's output, from any same-size variant, in its
 * order), applies the cotangent gbar of that forward index, and writes dA
 * as above.  flag: 0/1/2 of the
 * recomputed forward, or 3 when the eigenvalues of matrix j are not those
 * of A_j (a pair further apart than 1e-6 * max |lambda|; dA = NaN).
 * Returns 0, -1 handle, -2 n, -3 A NULL, -4 eigenvalues or gbar NULL, -5 dA
 * or work NULL, -6 batch_count <= 0, -7 flag NULL, 2 launch failure. */
size_t tokura_get_zgeev_schur_worksize(int n, int batch_count);
int tokura_zgeev_batched_schur_gpu(tokurablas_t *handle, int n, int batch_count,
                                   const hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                                   hipDoubleComplex *schur, char *flag, hipStream_t stream);
int tokura_zgeev_batched_backward_gpu(tokurablas_t *handle, int n, int batch_count,
                                      const hipDoubleComplex *A, const hipDoubleComplex *schur,
                                      const hipDoubleComplex *gbar, hipDoubleComplex *dA,
                                      char *flag, hipStream_t stream);
int tokura_zgeev_batched_backward_recompute_gpu(tokurablas_t *handle, int n, int batch_count,
                                                const hipDoubleComplex *A, const hipDoubleComplex *eigenvalues,
                                                const hipDoubleComplex *gbar, hipDoubleComplex *dA,
                                                hipDoubleComplex *work, char *flag, hipStream_t stream);

/* Ragged differentiation (no counterpart in the reference): the forward
 * with Schur vectors and the backward for a ragged batch prepared with
 * tokura_zgeev_ragged_prepare (same `work`, same offsets).  The Schur form of
 * matrix b is stored at schur + 2*offsets[b] (T then U, n_b^2 each), so
 * `schur` holds tokura_get_zgeev_ragged_schur_worksize bytes; eigenvalues and
 * cotangents are at the ragged eigenvalue offsets (prefix sums of n_b), dA at
 * dA + offsets[b].  Every matrix gets the arithmetic of the same-size entries.
 * Returns as tokura_zgeev_ragged_gpu (-8: work not prepared). */
size_t tokura_get_zgeev_ragged_schur_worksize(int batch_count, const int64_t *offsets_host);
int tokura_zgeev_ragged_schur_gpu(tokurablas_t *handle, int batch_count, const hipDoubleComplex *A,
                                  hipDoubleComplex *eigenvalues, hipDoubleComplex *schur, void *work,
                                  char *flag, hipStream_t stream);
int tokura_zgeev_ragged_backward_gpu(tokurablas_t *handle, int batch_count, const hipDoubleComplex *schur,
                                     const hipDoubleComplex *gbar, hipDoubleComplex *dA, void *work,
                                     char *flag, hipStream_t stream);

/* ---- eigenvectors -------------------------------------------------------
 * No counterpart in the reference, which computes eigenvalues only
 * (README.md:41-44); the analogue of LAPACK ZGEEV with JOBVL/JOBVR = 'V'
 * (without balancing).  For every matrix j writes eigenvalues[i + j*n] and,
 * where the pointer is non-NULL, right eigenvectors VR_j(:, i)
 * (A_j v = lambda_i v) and left eigenvectors VL_j(:, i)
 * (u^H A_j = lambda_i u^H), each n*n column-major per matrix, normalised as
 * ZGEEV does: unit Euclidean norm, the component of largest modulus real and
 * non-negative.  Eigenvalues are in the order of the Schur diagonal.  VL and
 * VR may not both be NULL (-5).  flag as in tokura_zgeev_batched_gpu; a
 * matrix with flag != 0 gets NaN vectors.  A is not modified.
 * `work` (may be NULL) is used only when one side is asked for: with
 * tokura_get_zgeev_vectors_worksize(n, batch) bytes the Schur factor T is
 * parked there and the faster split kernels run (both sides park T in VL);
 * without it a slower one-kernel form keeps T in LDS.  Results are the same
 * bits either way. */
size_t tokura_get_zgeev_vectors_worksize(int n, int batch_count);
int tokura_zgeev_batched_vectors_gpu(tokurablas_t *handle, int n, int batch_count,
                                     const hipDoubleComplex *A, hipDoubleComplex *eigenvalues,
                                     hipDoubleComplex *VL, hipDoubleComplex *VR, hipDoubleComplex *work,
                                     char *flag, hipStream_t stream);

/* ---- real matrices (LAPACK DGEEV analogue, eigenvalues only) ------------
 * No counterpart in the reference, which solves complex matrices only.  A
 * holds batch_count real n x n matrices (double), column-major, matrix after
 * matrix, and is not modified.  Eigenvalue i of matrix j is
 * wr[i + j*n] + i*wi[i + j*n]; complex conjugate pairs are consecutive,
 * positive imaginary part first, with equal real parts (as DGEEV).  Method:
 * Householder Hessenberg (DGEHD2) + Francis double-shift QR (DLAHQR, DLANV2
 * for 2x2 blocks), no balancing.  work: tokura_get_dgeev_worksize(n, batch)
 * bytes; flag and return codes as tokura_zgeev_batched_gpu (-4: wr or wi
 * NULL).  The sweep cap of tokuraSetMaxSweeps applies. */
size_t tokura_get_dgeev_worksize(int n, int batch_count);
int tokura_dgeev_batched_gpu(tokurablas_t *handle, int n, int batch_count, const double *A, double *wr, double *wi,
                             double *work, char *flag, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TOKURA_BLAS_MI355X_H */
