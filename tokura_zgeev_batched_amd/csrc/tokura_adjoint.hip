// Differentiation: forward with Schur vectors and the adjoint kernel.
#include "tokura_internal.hpp"
#include "zgeev_device.hpp"

extern "C" size_t tokura_get_zgeev_schur_worksize(int n, int batch_count) { return (size_t)2 * n * n * batch_count * 16; }
extern "C" int tokura_zgeev_batched_schur_gpu(tokurablas_t *, int, int, const hipDoubleComplex *, hipDoubleComplex *, hipDoubleComplex *, char *, hipStream_t) { return -100; }
extern "C" int tokura_zgeev_batched_backward_gpu(tokurablas_t *, int, int, const hipDoubleComplex *, const hipDoubleComplex *, const hipDoubleComplex *, hipDoubleComplex *, char *, hipStream_t) { return -100; }
