// mlp.hpp -- learned chunk-cache admission/prefetch policy (MLP 8->64->64->2).
#pragma once
#include <vector>

#include "engine.hpp"

namespace dcg {

void mlp_create(Engine &e);
// Online training step on the reads observed in the snapshot that just ended
// (the engine's per-chunk read counters, still on the device).
void mlp_observe(Engine &e);
// Start the pending training launch behind the work enqueued so far (called
// once the delete mutation is enqueued).
void mlp_release_training(Engine &e);
// Chunks to keep resident for the staged batch about to run.
std::vector<uint64_t> mlp_plan(Engine &e, uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc);
// Wait for the training launches (they run on the engine stream).
void mlp_sync(Engine &e);
// Decision counters and the judge's last comparison (dcg_cache_stats fields).
void mlp_stats(const Engine &e, dcg_cache_stats *st);

}  // namespace dcg
