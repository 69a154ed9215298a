// mlp.hip -- learned cache policy (placeholder until the MLP kernels land).
#include "mlp.hpp"

namespace dcg {

struct MlpState {
    int dummy = 0;
};

void MlpDeleter::operator()(MlpState *m) const { delete m; }

void mlp_create(Engine &e) { e.cache.mlp.reset(new MlpState()); }
void mlp_observe(Engine &, const std::vector<float> &) {}
std::vector<uint64_t> mlp_plan(Engine &e, uint64_t, uint64_t, uint64_t, uint64_t) {
    std::vector<uint64_t> keep;
    for (uint64_t f = 0; f < e.cache.frames; ++f)
        if (e.cache.frame_chunk[f] >= 0) keep.push_back((uint64_t)e.cache.frame_chunk[f]);
    return keep;
}

}  // namespace dcg
