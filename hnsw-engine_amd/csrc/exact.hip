// Exact k-nearest-neighbour search (the reference's oracle.py:35-126) on the GPU.
#include <hip/hip_runtime.h>

#include "../../include/hnsw_mi355x.h"
#include "index.hpp"

using namespace hm;

extern "C" int hm_exact(const float *data, int64_t n, int32_t dim, int32_t metric,
                        const float *queries, int64_t nq, int32_t k, const uint8_t *allowed,
                        int device, int64_t *ids, double *dists) {
    (void)data; (void)n; (void)dim; (void)metric; (void)queries; (void)nq; (void)k;
    (void)allowed; (void)device; (void)ids; (void)dists;
    return invalid("hm_exact: the matrix-core exact search is not in this build yet");
}
