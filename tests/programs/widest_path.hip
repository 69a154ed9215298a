// widest_path.hip -- a vertex program that is not one of the engine's built-in
// programs, built as a plugin: the widest (maximum-bottleneck) path from the
// query's source, with the engine's edge weights as capacities.
//
// Widths are kept in the engine's min-lattice as value = ~width (smaller is
// wider): the source has width +inf (value 0), an edge offers
// max(value_src, ~capacity) = ~min(width_src, capacity), and unreached vertices
// keep IDENTITY_ELEMENT (width 0).  An edge can offer exactly its source's value
// (a bottleneck further back), so the program needs the rooted deletion witness,
// as CC does.  A second program in the same set, hop-limited BFS (levels capped
// at a fixed depth), exercises several programs in one plugin.
#include "dcg_program_plugin.hpp"

struct WidestPath : dcg::api::AppBase<uint32_t, uint32_t, uint32_t> {
    static constexpr bool kSupportsGpuDeletionRepair = true;
    static constexpr bool kRootedDeletionWitness = true;
    __host__ __device__ explicit WidestPath(const dcg::api::Env &e) : AppBase(e) {}
    __host__ __device__ __forceinline__ uint32_t TraversalEdgeWeight(uint32_t s, uint32_t d) const {
        return m_env.weight(s, d);
    }
    __host__ __device__ __forceinline__ uint32_t AccumulateBuffer(uint32_t, uint32_t, uint32_t capacity,
                                                                  uint32_t buffer) const {
        const uint32_t edge = ~capacity;   // capacities are >= 1, so this is below IDENTITY
        return buffer > edge ? buffer : edge;
    }
};

// BFS levels up to kMaxHops; vertices farther away stay unreached.
struct HopLimitedBfs : dcg::api::AppBase<uint32_t, uint32_t, uint32_t> {
    static constexpr uint32_t kMaxHops = 5;
    static constexpr bool kSupportsGpuDeletionRepair = true;
    __host__ __device__ explicit HopLimitedBfs(const dcg::api::Env &e) : AppBase(e) {}
    __host__ __device__ __forceinline__ uint32_t TraversalEdgeWeight(uint32_t, uint32_t) const { return 1u; }
    __host__ __device__ __forceinline__ uint32_t DeletionEdgeWeight(uint32_t, uint32_t) const { return 1u; }
    __host__ __device__ __forceinline__ uint32_t AccumulateBuffer(uint32_t, uint32_t, uint32_t w,
                                                                  uint32_t buffer) const {
        return buffer + w > kMaxHops ? dcg::api::IDENTITY_ELEMENT : buffer + w;
    }
    __host__ __device__ static __forceinline__ uint32_t OrderedBucketWidth(uint32_t quarters, uint32_t) {
        return quarters;
    }
};

DCG_DEFINE_PROGRAM(dcg_widest_path_programs, WidestPath, HopLimitedBfs)
