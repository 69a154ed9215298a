// gen_gpu.hip -- GPU RMAT generator (same draws as gen.cpp, for large scales).
#include "gen_common.hpp"

namespace dcg {
__global__ void k_rmat_draw(RmatParams p, uint64_t begin, uint64_t count, uint64_t *keys) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        keys[i] = rmat_draw(p, begin + i);
}
}  // namespace dcg
