// mlp.hpp -- learned chunk-cache admission/prefetch policy (MLP 8->64->64->2).
#pragma once
#include <vector>

#include "engine.hpp"

namespace dcg {

void mlp_create(Engine &e);
// Online training step on the reads observed in the snapshot that just ended
// (the engine's per-chunk read counters, still on the device).
void mlp_observe(Engine &e);
// Start the pending training launch on the training stream behind the work
// enqueued so far (called when a delete stage starts, so it runs beside the
// delete mutation).
void mlp_release_training(Engine &e);
// Make the engine stream wait for everything issued on the training stream
// (before every persistent closure and stage graph: no training kernel may run
// beside a closure).
void mlp_join(Engine &e);
// Chunks to keep resident for the staged batch about to run.
std::vector<uint64_t> mlp_plan(Engine &e, uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc);
// Wait for the training stream.
void mlp_sync(Engine &e);
// Decision counters and the judge's last comparison (dcg_cache_stats fields).
void mlp_stats(const Engine &e, dcg_cache_stats *st);

}  // namespace dcg
