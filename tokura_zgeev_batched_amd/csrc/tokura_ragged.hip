// Ragged batches (n varies per matrix): persistent work-queue kernel.
#include "tokura_internal.hpp"
#include "zgeev_device.hpp"

void ragged_release(tokurablas_t *h)
{
    if (h->graph.exec) hipGraphExecDestroy(h->graph.exec);
    if (h->graph.graph) hipGraphDestroy(h->graph.graph);
    h->graph = RaggedGraphCache();
    if (h->ragged_host_stage) hipHostFree(h->ragged_host_stage);
    h->ragged_host_stage = nullptr;
    h->ragged_host_bytes = 0;
}

extern "C" size_t tokura_get_zgeev_ragged_worksize(int batch_count) { return batch_count > 0 ? 0 : (size_t)-2; }
extern "C" int tokura_zgeev_ragged_prepare(tokurablas_t *, int, const int64_t *, void *, size_t, hipStream_t) { return -100; }
extern "C" int tokura_zgeev_ragged_gpu(tokurablas_t *, int, const hipDoubleComplex *, hipDoubleComplex *, void *, char *, hipStream_t) { return -100; }
extern "C" int tokura_zgeev_ragged_graph(tokurablas_t *, int, const hipDoubleComplex *, hipDoubleComplex *, void *, char *, hipStream_t) { return -100; }
