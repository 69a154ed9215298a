/*
 * dcg_program.hpp -- vertex-program API of the MI355X incremental streaming-graph
 * engine (header-only C++/HIP).
 *
 * The counterpart of sepgraph::api::AppBase (include/framework/variants/api.cuh:22-174):
 * a program is a struct deriving from dcg::api::AppBase<TValue, TBuffer, TWeight>
 * with the same hook names and static traits.  Hooks are resolved at compile time:
 * the engine's repair kernels are templates over a set of programs (no virtual
 * dispatch on the device), instantiated for the built-in programs below inside
 * libdcg.so and for a user's program in a plugin built with dcg_program_plugin.hpp.
 *
 * Min-lattice programs (everything but kSignedResidual) keep one 32-bit value per
 * (query, vertex); smaller is better and IDENTITY_ELEMENT (0xFFFFFFFF) is
 * "unreached".  The engine packs value and parent (or value and hop depth for
 * rooted-witness programs) into one 64-bit word, so the reference's pair of
 * AccumulateBuffer (atomicMin of the buffer plus a parent CAS loop,
 * hybrid_sssp.cu:126-145) and CombineValueBuffer (value = min(value, buffer)) is
 * one 64-bit atomicMin done by the engine.  A program therefore states only what
 * an edge offers:
 *
 *   TBuffer AccumulateBuffer(src, dst, weight, buffer)   candidate for dst given the
 *                                                        source's value `buffer`;
 *                                                        must be >= buffer (monotone)
 *   TBuffer GetInitBuffer(node)                          seed value (IDENTITY: none)
 *   TWeight TraversalEdgeWeight(src, dst)                weight of an edge
 *   TWeight DeletionEdgeWeight(src, dst)                 weight the dependency test uses
 *                                                        (the same as the traversal's
 *                                                        for every program here)
 *   static bool IsDeletionDependency(src, dst, source, destination, candidate,
 *                                    parent, parent_tight)
 *                                                        the reference's predicate
 *                                                        (api.cuh:43-49)
 *
 * Traits (api.cuh:31-37): kSupportsGpuDeletionRepair, kComponentLabels,
 * kVertexSeeds (every vertex is seeded: its frontier starts at all vertices),
 * kRootedDeletionWitness (the engine keeps (value, depth) with parents apart,
 * the depth counting hops inside one value class (an edge that raises the value
 * offers depth 1), so parents form a tree rooted at the seeds and a deletion
 * invalidates only the part that lost its path to a root; needed whenever an edge
 * can offer a value equal to its source's, as labels and bottlenecks do),
 * kSignedResidual (PageRank's residual push).  A program without kVertexSeeds
 * seeds exactly its query's source.
 *
 * Differences from the reference's signatures: the weight hooks are const member
 * functions (a program instance carries its query's Env: source and weight
 * parameters, so seeded weights need no global state), and AccumulateBuffer is the
 * pure candidate function described above.
 */
#ifndef DCG_PROGRAM_HPP
#define DCG_PROGRAM_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dcg.h"

namespace dcg {
namespace api {

typedef uint32_t index_t;
constexpr uint32_t IDENTITY_ELEMENT = 0xFFFFFFFFu;

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// What one program instance knows of its query: the source and the engine's
// weight parameters (dcg_config weight_mode / weight_max / weight_seed).
struct Env {
    uint32_t source;
    uint32_t weight_mode;   // DCG_W_REFERENCE or DCG_W_SEEDED
    uint32_t weight_max;
    uint64_t weight_seed;
    // The engine's edge weight: (u+v)%128+1 (api.cuh:26-28) or 1 + splitmix64(seed ^ (u<<32|v)) % W.
    __host__ __device__ __forceinline__ uint32_t weight(index_t u, index_t v) const {
        if (weight_mode == DCG_W_SEEDED)
            return 1u + (uint32_t)(splitmix64(weight_seed ^ (((uint64_t)u << 32) | v)) % weight_max);
        return (uint32_t)(((uint64_t)u + v) % 128u + 1u);
    }
};

template <typename TValue, typename TBuffer, typename TWeight>
struct AppBase {
    typedef TValue value_type;
    typedef TBuffer buffer_type;
    typedef TWeight weight_type;

    static constexpr bool kSupportsGpuDeletionRepair = false;
    static constexpr bool kComponentLabels = false;
    static constexpr bool kVertexSeeds = false;
    static constexpr bool kRootedDeletionWitness = false;
    static constexpr bool kSignedResidual = false;

    Env m_env;
    __host__ __device__ explicit AppBase(const Env &env) : m_env(env) {}

    // api.cuh:26-28 (the historical default); programs override.
    __host__ __device__ __forceinline__ TWeight TraversalEdgeWeight(index_t src, index_t dst) const {
        return static_cast<TWeight>((src + dst) % 128 + 1);
    }
    __host__ __device__ __forceinline__ TWeight DeletionEdgeWeight(index_t src, index_t dst) const {
        return TraversalEdgeWeight(src, dst);
    }
    // api.cuh:43-49: the parent edge, or a tight witness while the parent is not tight.
    __host__ __device__ static __forceinline__ bool IsDeletionDependency(index_t src, index_t, TValue source,
                                                                         TValue destination, uint64_t candidate,
                                                                         index_t parent, bool parent_tight) {
        return parent == src ||
               (source != static_cast<TValue>(UINT32_MAX) && candidate == static_cast<uint64_t>(destination) &&
                !parent_tight);
    }
    __host__ __device__ __forceinline__ TValue GetInitValue(index_t) const {
        return static_cast<TValue>(IDENTITY_ELEMENT);
    }
    __host__ __device__ __forceinline__ TBuffer GetInitBuffer(index_t node) const {
        return node == m_env.source ? TBuffer(0) : static_cast<TBuffer>(IDENTITY_ELEMENT);
    }
    __host__ __device__ __forceinline__ TBuffer GetIdentityElement() const {
        return static_cast<TBuffer>(IDENTITY_ELEMENT);
    }
    // The reference's min-lattice combine (hybrid_sssp.cu:109-121): lower the
    // value to the buffer and schedule the vertex if it dropped.  The engine's
    // packed state applies it inside its atomicMin; kept for completeness.
    __host__ __device__ __forceinline__ bool CombineValueBuffer(index_t, TValue *p_value, TBuffer *p_buffer) const {
        if (*p_value > *p_buffer) {
            *p_value = *p_buffer;
            return true;
        }
        return false;
    }
    __host__ __device__ __forceinline__ bool IsActiveNode(index_t, TBuffer buffer, TValue value) const {
        return buffer < value;
    }
    // Bucket width of the ordered repair (ordered_gpu_repair.cuh:116-160) for the
    // engine's width setting `quarters` (dcg_set_buckets) and weight cap;
    // IDENTITY_ELEMENT keeps the program out of the ordered closure.
    __host__ __device__ static __forceinline__ uint32_t OrderedBucketWidth(uint32_t, uint32_t) {
        return IDENTITY_ELEMENT;
    }
};

}  // namespace api

// ---- the reference's four programs ---------------------------------------------
namespace apps {

// samples/hybrid_sssp/hybrid_sssp.cu:51-204
struct SSSP : api::AppBase<uint32_t, uint32_t, uint32_t> {
    static constexpr bool kSupportsGpuDeletionRepair = true;
    __host__ __device__ explicit SSSP(const api::Env &e) : AppBase(e) {}
    __host__ __device__ __forceinline__ uint32_t TraversalEdgeWeight(api::index_t s, api::index_t d) const {
        return m_env.weight(s, d);
    }
    __host__ __device__ __forceinline__ uint32_t DeletionEdgeWeight(api::index_t s, api::index_t d) const {
        return m_env.weight(s, d);
    }
    // buffer + weight (hybrid_sssp.cu:135-136); sums at or above IDENTITY offer nothing
    __host__ __device__ __forceinline__ uint32_t AccumulateBuffer(api::index_t, api::index_t, uint32_t weight,
                                                                  uint32_t buffer) const {
        const uint64_t c = (uint64_t)buffer + weight;
        return c >= api::IDENTITY_ELEMENT ? api::IDENTITY_ELEMENT : (uint32_t)c;
    }
    __host__ __device__ static __forceinline__ uint32_t OrderedBucketWidth(uint32_t quarters, uint32_t wmax) {
        const uint64_t w = (uint64_t)wmax * quarters / 4;
        return w < 1 ? 1u : (uint32_t)w;
    }
};

// samples/hybrid_bfs/hybrid_bfs.cu:51-204: unit weights for traversal and deletion
struct BFS : api::AppBase<uint32_t, uint32_t, uint32_t> {
    static constexpr bool kSupportsGpuDeletionRepair = true;
    __host__ __device__ explicit BFS(const api::Env &e) : AppBase(e) {}
    __host__ __device__ __forceinline__ uint32_t TraversalEdgeWeight(api::index_t, api::index_t) const { return 1u; }
    __host__ __device__ __forceinline__ uint32_t DeletionEdgeWeight(api::index_t, api::index_t) const { return 1u; }
    __host__ __device__ __forceinline__ uint32_t AccumulateBuffer(api::index_t, api::index_t, uint32_t weight,
                                                                  uint32_t buffer) const {
        const uint64_t c = (uint64_t)buffer + weight;
        return c >= api::IDENTITY_ELEMENT ? api::IDENTITY_ELEMENT : (uint32_t)c;
    }
    __host__ __device__ static __forceinline__ uint32_t OrderedBucketWidth(uint32_t quarters, uint32_t) {
        return quarters;
    }
};

// samples/hybrid_cc/hybrid_cc.cu:51-201: every vertex seeded with its id, labels
// forwarded unchanged (weight 0), rooted deletion witness
// (include/framework/directed_label_witness.h).
struct CC : api::AppBase<uint32_t, uint32_t, uint32_t> {
    static constexpr bool kSupportsGpuDeletionRepair = true;
    static constexpr bool kComponentLabels = true;
    static constexpr bool kVertexSeeds = true;
    static constexpr bool kRootedDeletionWitness = true;
    __host__ __device__ explicit CC(const api::Env &e) : AppBase(e) {}
    __host__ __device__ __forceinline__ uint32_t TraversalEdgeWeight(api::index_t, api::index_t) const { return 0u; }
    __host__ __device__ __forceinline__ uint32_t DeletionEdgeWeight(api::index_t, api::index_t) const { return 0u; }
    __host__ __device__ __forceinline__ uint32_t GetInitBuffer(api::index_t node) const { return node; }
    __host__ __device__ __forceinline__ uint32_t AccumulateBuffer(api::index_t, api::index_t, uint32_t,
                                                                  uint32_t buffer) const {
        return buffer;
    }
};

// samples/hybrid_pr/hybrid_pr.cu:38-128: signed-residual push (pr_residual.cuh).
// Value = rank, buffer = residual (FP32); the engine's PageRank kernels call
// these hooks.  eps is the residual threshold (--error).
struct PageRank {
    static constexpr bool kSupportsGpuDeletionRepair = false;
    static constexpr bool kComponentLabels = false;
    static constexpr bool kVertexSeeds = true;
    static constexpr bool kRootedDeletionWitness = false;
    static constexpr bool kSignedResidual = true;
    static constexpr float ALPHA = 0.85f;
    float m_error;
    float m_base;   // 1 - ALPHA on the reference's scale; base/n on the probability scale
    __host__ __device__ PageRank(float error, float base) : m_error(error), m_base(base) {}
    __host__ __device__ __forceinline__ float GetInitValue(api::index_t) const { return 0.f; }
    __host__ __device__ __forceinline__ float GetInitBuffer(api::index_t) const { return m_base; }
    __host__ __device__ __forceinline__ float GetIdentityElement() const { return 0.f; }
    // Fold the residual into the rank; returns the folded residual (the engine
    // spreads ALPHA * r / live degree over the out-row, hybrid_pr.cu:74-80).
    __device__ __forceinline__ float CombineValueBuffer(api::index_t, float *value, float *residual) const {
        const float r = *residual;
        *value += r;
        *residual = 0.f;
        return r;
    }
    // hybrid_pr.cu:89-94; returns the residual before the add
    __device__ __forceinline__ float AccumulateBuffer(float *residual, float delta) const {
        return atomicAdd(residual, delta);
    }
    __host__ __device__ __forceinline__ bool IsActiveNode(api::index_t, float buffer, float) const {
        return fabsf(buffer) > m_error;
    }
};

}  // namespace apps
}  // namespace dcg

#endif
