// mlp.hip -- learned chunk-cache admission/prefetch policy.
//
// An MLP 8 -> 64 -> 64 -> 2 scores every chunk before a batch runs:
//   features  log1p of the chunk's edge reads in each of the last 4 snapshots,
//             log1p of rows expanded in the last snapshot (frontier overlap),
//             log1p of the chunk's adjacency entries (out-degree sum),
//             log1p of update records of the staged batch touching the chunk,
//             recency exp(-(snapshots since last read)/4)
//   outputs   o0 = logit(chunk is read in the coming snapshot),
//             o1 = log1p(edge reads in the coming snapshot)
//   score     sigmoid(o0) * max(o1, 0)
// Weights are kept in FP16 for the device passes (FP32 masters for Adam);
// every product accumulates in FP32.  The 64x64 layer and both of its
// backward products run on MFMA (v_mfma_f32_32x32x8_f16), one wave per
// 32-row tile.  After each snapshot the observed reads label the features the
// decision used, and a few Adam steps train the network online.  The
// hotness/LRU ranking stays as the fallback and the judge: when the network's
// top-`frames` set captured fewer of the observed reads than the hotness set
// would have, the next decision uses hotness.  The policy decides placement
// only; values never depend on it.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <numeric>


#include "mlp.hpp"
#include "primitives.hpp"

namespace dcg {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr int kIn = 8, kHid = 64, kOut = 2, kTile = 32;
constexpr int kW1 = kIn * kHid, kW2 = kHid * kHid, kW3 = kHid * kOut;
// parameter block layout (FP32 masters and gradients use the same layout)
constexpr int P_W1 = 0, P_B1 = P_W1 + kW1, P_W2 = P_B1 + kHid, P_B2 = P_W2 + kW2, P_W3 = P_B2 + kHid,
              P_B3 = P_W3 + kW3, P_NUM = P_B3 + kOut;

constexpr int kReplay = 4;   // decisions kept for training (ring)

struct MlpState {
    // Training and the judge run on their own stream (default priority)
    // beside the delete mutation, whose row passes wait on the host link:
    // training forks from the engine stream when a delete stage starts
    // (`released`).  No kernel of it may run beside a persistent repair
    // closure (it could hold CU slots that a closure workgroup needs, and the
    // closures' grid barrier requires every workgroup resident), so the engine
    // stream waits on `trained`, the end of the last launch on this stream,
    // before every closure launch and stage graph (mlp_join), before a pass
    // reads the weights, and before a decision is written into the ring.
    hipStream_t train = nullptr;
    hipEvent_t trained = nullptr, released = nullptr, labelled = nullptr;
    bool train_due = false;   // a decision was made; its training launch waits for the next delete stage
    bool joined = true;       // the engine stream has waited for everything issued on `train`
    bool judge_pending = false;   // the last judge's sums are on their way to h_judge
    double *h_judge = nullptr;    // pinned: network top-k, hotness top-k, all reads
    ~MlpState() {
        if (train) {
            (void)hipStreamSynchronize(train);
            (void)hipStreamDestroy(train);
        }
        if (trained) (void)hipEventDestroy(trained);
        if (released) (void)hipEventDestroy(released);
        if (labelled) (void)hipEventDestroy(labelled);
        if (h_judge) (void)hipHostFree(h_judge);
    }
    DevBuf<float> master, grad, m1, m2;   // FP32 masters, gradients, Adam moments
    DevBuf<float> part;                   // per-workgroup gradient sums of a pass
    DevBuf<_Float16> w16;                 // FP16 weights used by the device passes
    DevBuf<float> X, Y, out, degsum, lastuse;   // X/Y hold kReplay labelled decisions of C rows each
    DevBuf<float> Xcur;                   // features of the decision awaiting its labels
    uint64_t rows = 0;                    // C (rows per decision)
    int slot = 0, filled = 0;             // ring slot the next labelled decision goes to
    int steps = 0;
    bool have_decision = false, fallback = true;
    double mlp_capture = 0, hot_capture = 0;
    DevBuf<float> mscore, hot_dec;        // the decision's network and hotness scores (judge)
    DevBuf<float> plan;                   // the decision's network score as predicted reads
    DevBuf<uint32_t> sort_k;              // judge: descending-order keys of a ranking, then sorted keys
    DevBuf<float> sort_v;                 // judge: reads in ranking order
    DevBuf<double> judge;                 // judge sums: network top-k, hotness top-k, all reads
    DevBuf<float> judge_reads;            // the snapshot's reads, kept for the judge
    DevBuf<uint8_t> judge_tmp;            // radix-sort scratch of the judge
    int decisions = 0, fallbacks = 0;
    int rejected = 0;   // judged decisions in a row, after warm-up, that kept the hotness ranking
    int next_probe = 0;   // while quiet: the next decision that runs the network
};

// While the judge keeps rejecting the trained network (kQuiet judged decisions
// in a row), the network runs (features, forward pass, judge and a training
// step) only at probe decisions, kProbe decisions apart after the first quiet
// rejection and twice as far after each further one, up to kProbeMax; the
// other decisions take the hotness plan at no extra cost.  One network decision
// accepted by the judge restores every decision.
constexpr int kQuiet = 2, kProbe = 4, kProbeMax = 64;

void MlpDeleter::operator()(MlpState *m) const { delete m; }

// ---------------------------------------------------------------- device part
// C(32x32 tile) += A(32xK) * B(Kx32) on one wave.  Operand layout of
// v_mfma_f32_32x32x8_f16: lane l supplies A[l%32][k0+4*(l/32)+j] and
// B[k0+4*(l/32)+j][l%32], j=0..3; accumulator register r of lane l holds
// C[8*(r/4) + 4*(l/32) + r%4][l%32].
template <typename FA, typename FB>
__device__ __forceinline__ float16_t mfma_tile(FA a, FB b, int K, float16_t acc) {
    const int l = (int)lane_id();
    const int i = l & 31, kq = 4 * (l >> 5);
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 8) {
        half4_t av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            av[j] = a(i, k0 + kq + j);
            bv[j] = b(k0 + kq + j, i);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ int acc_row(int r) { return 8 * (r >> 2) + 4 * ((int)lane_id() >> 5) + (r & 3); }

struct TileLds {
    _Float16 w2[kHid][kHid];      // W2[i][j] (in i, out j)
    _Float16 h1[kTile][kHid];     // layer-1 activations
    _Float16 dz2[kTile][kHid];    // layer-2 pre-activation gradient
    float z2[kTile][kHid];        // layer-2 pre-activations (+bias)
    float x[kTile][kIn];
    float dout[kTile][kOut];
    float w1[kIn * kHid];
    float b1[kHid], b2[kHid], w3[kHid * kOut], b3[kOut];
};

// Forward (and, with Y != null, backward) over 32-row tiles; one wave per
// workgroup walks tiles blockIdx.x, +gridDim.x, ...  Gradient contributions stay
// in registers across the workgroup's tiles (the dW2 tiles in the MFMA
// accumulators) and are stored once per workgroup as its row of `part`
// (unscaled sums); k_grad_reduce adds the rows in workgroup order, so the
// gradient is the same on every run.
__global__ __launch_bounds__(64) void k_mlp_tile(const float *X, const float *Y, uint64_t rows, const _Float16 *w16,
                                                 const float *master, float *out, float *part) {
    __shared__ TileLds L;
    const int l = (int)lane_id();
    // stage parameters: W2 in FP16, the small layers in FP32 from the FP16 copy
    for (int t = l; t < kW2; t += 64) L.w2[t / kHid][t % kHid] = w16[P_W2 + t];
    for (int t = l; t < kW1; t += 64) L.w1[t] = (float)w16[P_W1 + t];
    for (int t = l; t < kW3; t += 64) L.w3[t] = (float)w16[P_W3 + t];
    L.b1[l] = master[P_B1 + l];
    L.b2[l] = master[P_B2 + l];
    if (l < kOut) L.b3[l] = master[P_B3 + l];
    // gradient accumulators (per lane): dW2 as four 32x32 MFMA tiles, lane-owned rest
    float16_t gw2[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) gw2[a][b] = float16_t{};
    float gw1[kIn] = {0, 0, 0, 0, 0, 0, 0, 0}, gb1 = 0.f, gb2 = 0.f, gw3a = 0.f, gw3b = 0.f, gb3 = 0.f;
    const uint64_t tiles = (rows + kTile - 1) / kTile;
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint64_t r0 = tile * kTile;
        __syncthreads();
        for (int t = l; t < kTile * kIn; t += 64) {
            const uint64_t r = r0 + t / kIn;
            L.x[t / kIn][t % kIn] = r < rows ? X[r * kIn + t % kIn] : 0.f;
        }
        __syncthreads();
        // layer 1 (FP32 FMA): lane owns hidden unit l for all 32 rows
        #pragma unroll 4
        for (int i = 0; i < kTile; ++i) {
            float a = L.b1[l];
#pragma unroll
            for (int k = 0; k < kIn; ++k) a += L.x[i][k] * L.w1[k * kHid + l];
            L.h1[i][l] = (_Float16)(a > 0.f ? a : 0.f);
        }
        __syncthreads();
        // layer 2 on MFMA: Z2 = H1 * W2 (two 32x32 output tiles)
        for (int nt = 0; nt < 2; ++nt) {
            float16_t acc = {};
            acc = mfma_tile([&](int i, int k) { return L.h1[i][k]; },
                            [&](int k, int j) { return L.w2[k][nt * 32 + j]; }, kHid, acc);
#pragma unroll
            for (int r = 0; r < 16; ++r) L.z2[acc_row(r)][nt * 32 + (l & 31)] = acc[r] + L.b2[nt * 32 + (l & 31)];
        }
        __syncthreads();
        // layer 3 (FP32): lanes 0..63 -> (row = l/2, output = l%2)
        {
            const int i = l >> 1, o = l & 1;
            float a = L.b3[o];
            #pragma unroll 4
            for (int k = 0; k < kHid; ++k) {
                const float h = L.z2[i][k] > 0.f ? L.z2[i][k] : 0.f;
                a += h * L.w3[k * kOut + o];
            }
            const uint64_t r = r0 + i;
            if (r < rows && out) out[r * kOut + o] = a;
            if (Y) {
                // unscaled per-row gradient (keeps the FP16 MFMA operands in
                // range); the 1/rows of the mean is applied at the flush
                float g = 0.f;
                if (r < rows) {
                    const float t = Y[r * kOut + o];
                    g = o == 0 ? (1.f / (1.f + __expf(-a)) - t) : (a - t);
                }
                L.dout[i][o] = g;
            }
        }
        if (!Y) continue;
        __syncthreads();
        // dW3, db3, dZ2 = (dout * W3^T) .* relu'(Z2)
        #pragma unroll 4
        for (int i = 0; i < kTile; ++i) {
            const float h = L.z2[i][l] > 0.f ? L.z2[i][l] : 0.f;
            gw3a += h * L.dout[i][0];
            gw3b += h * L.dout[i][1];
            const float dh = L.dout[i][0] * L.w3[l * kOut + 0] + L.dout[i][1] * L.w3[l * kOut + 1];
            L.dz2[i][l] = (_Float16)(L.z2[i][l] > 0.f ? dh : 0.f);
        }
        if (l < kOut)
            for (int i = 0; i < kTile; ++i) gb3 += L.dout[i][l];
        __syncthreads();
        for (int i = 0; i < kTile; ++i) gb2 += (float)L.dz2[i][l];
        // dW2 += H1^T * dZ2 on MFMA (M=64 in two tiles, N=64 in two tiles, K=32 rows)
        for (int mt = 0; mt < 2; ++mt)
            for (int nt = 0; nt < 2; ++nt)
                gw2[mt][nt] = mfma_tile([&](int i, int k) { return L.h1[k][mt * 32 + i]; },
                                        [&](int k, int j) { return L.dz2[k][nt * 32 + j]; }, kTile, gw2[mt][nt]);
        // dH1 = dZ2 * W2^T on MFMA; reuse z2 as dH1 storage
        __syncthreads();
        for (int nt = 0; nt < 2; ++nt) {
            float16_t acc = {};
            acc = mfma_tile([&](int i, int k) { return L.dz2[i][k]; },
                            [&](int k, int j) { return L.w2[nt * 32 + j][k]; }, kHid, acc);
#pragma unroll
            for (int r = 0; r < 16; ++r) L.z2[acc_row(r)][nt * 32 + (l & 31)] = acc[r];
        }
        __syncthreads();
        // dZ1 = dH1 .* relu'(Z1); lane owns hidden unit l of layer 1
        #pragma unroll 4
        for (int i = 0; i < kTile; ++i) {
            const float dz = (float)L.h1[i][l] > 0.f ? L.z2[i][l] : 0.f;
            gb1 += dz;
#pragma unroll
            for (int k = 0; k < kIn; ++k) gw1[k] += L.x[i][k] * dz;
        }
    }
    if (!Y || blockIdx.x >= tiles) return;
    float *g = part + (uint64_t)blockIdx.x * P_NUM;
    g[P_W3 + l * kOut + 0] = gw3a;
    g[P_W3 + l * kOut + 1] = gw3b;
    if (l < kOut) g[P_B3 + l] = gb3;
    g[P_B2 + l] = gb2;
    for (int mt = 0; mt < 2; ++mt)
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int r = 0; r < 16; ++r) g[P_W2 + (mt * 32 + acc_row(r)) * kHid + nt * 32 + (l & 31)] = gw2[mt][nt][r];
    g[P_B1 + l] = gb1;
#pragma unroll
    for (int k = 0; k < kIn; ++k) g[P_W1 + k * kHid + l] = gw1[k];
}

// grad = (sum of the first nb workgroup rows of part) / rows (mean loss).  A
// workgroup takes 64 parameters; its 16 waves sum interleaved rows, then one
// wave adds the 16 partial sums in order (the same result on every run).
__global__ __launch_bounds__(1024) void k_grad_reduce(const float *part, uint32_t nb, float inv, float *grad) {
    __shared__ float acc[16][64];
    const uint32_t c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    const uint32_t p = blockIdx.x * 64 + c;
    float a[4] = {0.f, 0.f, 0.f, 0.f};   // four chains keep four loads in flight
    if (p < P_NUM) {
        uint32_t b = r0;
        for (; b + 48 < nb; b += 64)
            for (int u = 0; u < 4; ++u) a[u] += part[(uint64_t)(b + 16 * u) * P_NUM + p];
        for (; b < nb; b += 16) a[0] += part[(uint64_t)b * P_NUM + p];
    }
    acc[r0][c] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (threadIdx.x < 64 && p < P_NUM) {
        float sum = 0.f;
        for (int r = 0; r < 16; ++r) sum += acc[r][c];
        grad[p] = sum * inv;
    }
}

constexpr uint32_t kMlpGrid = 1024;   // workgroups (one wave each) of a pass: four per CU

static uint32_t mlp_grid(uint64_t rows) {
    const uint64_t tiles = (rows + kTile - 1) / kTile;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, kMlpGrid));
}

// Gradient of the mean loss over `rows` rows: the tile pass, then the reduction
// of its per-workgroup rows.
static void mlp_gradient(hipStream_t st, const float *X, const float *Y, uint64_t rows, const _Float16 *w16,
                         const float *master, float *part, float *grad) {
    const uint32_t nb = mlp_grid(rows);
    hipLaunchKernelGGL(k_mlp_tile, dim3(nb), dim3(64), 0, st, X, Y, rows, w16, master, (float *)nullptr, part);
    hipLaunchKernelGGL(k_grad_reduce, dim3((P_NUM + 63) / 64), dim3(1024), 0, st, part, nb, 1.f / (float)rows, grad);
}

__global__ void k_adam(float *master, const float *grad, float *m1, float *m2, _Float16 *w16, int n, float lr,
                       float b1c, float b2c) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float a = 0.9f * m1[i] + 0.1f * g;
        const float b = 0.999f * m2[i] + 0.001f * g * g;
        m1[i] = a;
        m2[i] = b;
        const float w = master[i] - lr * (a / b1c) / (sqrtf(b / b2c) + 1e-8f);
        master[i] = w;
        w16[i] = (_Float16)w;
    }
}

__global__ void k_to_half(const float *a, _Float16 *b, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (_Float16)a[i];
}

// Update records of the staged batch per chunk (rows of the touched out- and in-rows).
__global__ void k_touch(const RowDesc *out, const RowDesc *in, const dcg_edge *adds, uint64_t na, const dcg_edge *dels,
                        uint64_t nd, uint32_t n, float *touch) {
    const uint64_t total = na + nd;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const dcg_edge e = i < na ? adds[i] : dels[i - na];
        if (e.u >= n || e.v >= n) continue;
        const RowDesc a = out[e.u], b = in[e.v];
        if (a.meta) atomicAdd(&touch[a.off >> kChunkLog2], 1.f);
        if (b.meta) atomicAdd(&touch[b.off >> kChunkLog2], 1.f);
    }
}

__global__ void k_degsum(const RowDesc *out, const RowDesc *in, uint32_t n, float *degsum) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < 2ull * n; v += (uint64_t)gridDim.x * blockDim.x) {
        const RowDesc d = v < n ? out[v] : in[v - n];
        if (!d.ext) continue;
        uint64_t a = d.off, e = d.off + d.ext;
        while (a < e) {
            const uint64_t c = a >> kChunkLog2, ce = (c + 1) << kChunkLog2, b = e < ce ? e : ce;
            atomicAdd(&degsum[c], (float)(b - a));
            a = b;
        }
    }
}

__global__ void k_features(const float *hist, const float *rows_last, const float *degsum, const float *touch,
                           const float *lastuse, float now, uint64_t C, float *X) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < C; c += (uint64_t)gridDim.x * blockDim.x) {
        const float4 h = reinterpret_cast<const float4 *>(hist)[c];
        float *x = X + c * kIn;
        x[0] = log1pf(h.x);
        x[1] = log1pf(h.y);
        x[2] = log1pf(h.z);
        x[3] = log1pf(h.w);
        x[4] = log1pf(rows_last[c]);
        x[5] = log1pf(degsum[c]) * 0.25f;
        x[6] = log1pf(touch[c]);
        x[7] = lastuse[c] > 0.f ? __expf(-(now - lastuse[c]) * 0.25f) : 0.f;
    }
}

__global__ void k_labels(const float *reads, uint64_t C, float *Y) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < C; c += (uint64_t)gridDim.x * blockDim.x) {
        Y[2 * c] = reads[c] > 0.f ? 1.f : 0.f;
        Y[2 * c + 1] = log1pf(reads[c]);
    }
}

// Decision scores from the network outputs (P(read) * predicted log reads) and
// the hotness score (reads over the four-snapshot window).  `mlp` holds the
// network score as predicted reads (expm1) for the admission plan; `raw`, when
// given, the score itself (the judge ranks by it).
__global__ void k_plan_scores(const float *out, const float *hist, uint64_t C, float *mlp, float *raw, float *hot) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < C; c += (uint64_t)gridDim.x * blockDim.x) {
        const float p = 1.f / (1.f + __expf(-out[2 * c]));
        const float sc = p * fmaxf(out[2 * c + 1], 0.f);
        mlp[c] = expm1f(fminf(sc, 30.f));
        if (raw) raw[c] = sc;
        const float4 h = reinterpret_cast<const float4 *>(hist)[c];
        hot[c] = h.x + h.y + h.z + h.w;
    }
}

// One workgroup: sums[slot] = reads held by the first k chunks of a ranking
// (sorted[0, k)), sums[2] = all reads; FP64 accumulation in a fixed order.
__global__ __launch_bounds__(1024) void k_judge_sum(const float *sorted, uint64_t C, uint64_t k, double *sums,
                                                   int slot) {
    __shared__ double part[2][1024];
    double top = 0, all = 0;
    for (uint64_t i = threadIdx.x; i < C; i += blockDim.x) {
        const double r = sorted[i];
        all += r;
        if (i < k) top += r;
    }
    part[0][threadIdx.x] = top;
    part[1][threadIdx.x] = all;
    __syncthreads();
    for (uint32_t w = 512; w >= 1; w >>= 1) {
        if (threadIdx.x < w) {
            part[0][threadIdx.x] += part[0][threadIdx.x + w];
            part[1][threadIdx.x] += part[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[slot] = part[0][0];
        sums[2] = part[1][0];
    }
}

// ------------------------------------------------------------------ host part
// `steps` Adam steps over every labelled decision in the ring, on the training
// stream (the caller has made the ring and weights ready on the engine stream
// and synchronised it).
static void mlp_train_steps(MlpState &m, int steps) {
    const uint64_t rows = m.rows * (uint64_t)m.filled;
    if (!rows) return;
    for (int s = 0; s < steps; ++s) {
        mlp_gradient(m.train, m.X.p, m.Y.p, rows, m.w16.p, m.master.p, m.part.p, m.grad.p);
        m.steps++;
        const float b1c = 1.f - std::pow(0.9f, (float)m.steps), b2c = 1.f - std::pow(0.999f, (float)m.steps);
        hipLaunchKernelGGL(k_adam, dim3(grid_for(P_NUM)), dim3(256), 0, m.train, m.master.p, m.grad.p, m.m1.p,
                           m.m2.p, m.w16.p, P_NUM, 3e-3f, b1c, b2c);
    }
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipEventRecord(m.trained, m.train));
    m.joined = false;
}

void mlp_create(Engine &e) {
    MlpState *m = new MlpState();
    e.cache.mlp.reset(m);
    const uint64_t C = e.cache.chunks;
    m->master.alloc(P_NUM);
    m->grad.alloc(P_NUM);
    m->part.alloc((uint64_t)kMlpGrid * P_NUM);
    m->mscore.alloc(C);
    m->plan.alloc(C);
    m->hot_dec.alloc(C);
    m->sort_k.alloc(2 * C);
    m->sort_v.alloc(C);
    m->judge.alloc(3);
    m->m1.alloc(P_NUM);
    m->m2.alloc(P_NUM);
    m->w16.alloc(P_NUM);
    DCG_HIP(hipStreamCreateWithFlags(&m->train, hipStreamNonBlocking));
    DCG_HIP(hipEventCreateWithFlags(&m->trained, hipEventDisableTiming));
    DCG_HIP(hipEventCreateWithFlags(&m->released, hipEventDisableTiming));
    DCG_HIP(hipEventCreateWithFlags(&m->labelled, hipEventDisableTiming));
    DCG_HIP(hipHostMalloc((void **)&m->h_judge, 3 * sizeof(double), hipHostMallocDefault));
    m->judge_reads.alloc(C);
    m->X.alloc(kReplay * C * kIn);
    m->Xcur.alloc(C * kIn);
    m->Y.alloc(kReplay * C * kOut);
    m->out.alloc(C * kOut);
    m->degsum.alloc(C);
    m->lastuse.alloc(C);
    m->rows = C;
    // He-style init from a fixed seed (splitmix64 -> uniform)
    std::vector<float> w(P_NUM, 0.f);
    uint64_t st = 0x5eed5eedull;
    auto uni = [&]() { return (float)((mix64(st++) >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f; };
    for (int i = 0; i < kW1; ++i) w[P_W1 + i] = uni() * std::sqrt(6.f / kIn);
    for (int i = 0; i < kW2; ++i) w[P_W2 + i] = uni() * std::sqrt(6.f / kHid);
    for (int i = 0; i < kW3; ++i) w[P_W3 + i] = uni() * std::sqrt(6.f / kHid);
    DCG_HIP(hipMemcpyAsync(m->master.p, w.data(), P_NUM * 4, hipMemcpyHostToDevice, e.stream));
    fill<float>(e.stream, m->m1.p, P_NUM, 0.f);
    fill<float>(e.stream, m->m2.p, P_NUM, 0.f);
    hipLaunchKernelGGL(k_to_half, dim3(grid_for(P_NUM)), dim3(256), 0, e.stream, m->master.p, m->w16.p, P_NUM);
    fill<float>(e.stream, m->degsum.p, C, 0.f);
    hipLaunchKernelGGL(k_degsum, dim3(grid_for(2ull * e.n, 256, 8192)), dim3(256), 0, e.stream, e.d_out.p, e.d_in.p,
                       e.n, m->degsum.p);
    DCG_HIP(hipGetLastError());
    e.sync();
}

// Judge: reads of the snapshot held by each ranking's top `frames` chunks.
// Rankings sort descending and stably (ties keep chunk order, lowest id
// first); the reads travel with the scores.  Runs after the training launches
// on a copy of the reads; judge_resolve reads the sums at the next decision.
static void judge_launch(Engine &e, MlpState &m) {
    const uint64_t C = e.cache.chunks, k = e.cache.frames;
    DCG_HIP(hipMemcpyAsync(m.judge_reads.p, e.cache.d_reads.p, C * 4, hipMemcpyDeviceToDevice, e.stream));
    DCG_HIP(hipEventRecord(m.labelled, e.stream));
    DCG_HIP(hipStreamWaitEvent(m.train, m.labelled, 0));
    const size_t tb = radix_sort_temp_bytes<uint32_t>(C, true);
    m.judge_tmp.ensure(tb);
    const float *ranks[2] = {m.mscore.p, m.hot_dec.p};
    for (int r = 0; r < 2; ++r) {
        // stable descending sort of the scores, the reads (as bits) carried along
        float_desc_keys(m.train, ranks[r], m.sort_k.p, C);
        radix_sort<uint32_t>(m.train, m.sort_k.p, m.sort_k.p + C, reinterpret_cast<const uint32_t *>(m.judge_reads.p),
                             reinterpret_cast<uint32_t *>(m.sort_v.p), C, 0, 32, m.judge_tmp.p, tb);
        hipLaunchKernelGGL(k_judge_sum, dim3(1), dim3(1024), 0, m.train, m.sort_v.p, C, k, m.judge.p, r);
    }
    DCG_HIP(hipGetLastError());
    DCG_HIP(hipMemcpyAsync(m.h_judge, m.judge.p, 3 * sizeof(double), hipMemcpyDeviceToHost, m.train));
    DCG_HIP(hipEventRecord(m.trained, m.train));
    m.joined = false;
    m.judge_pending = true;
}

// The network's ranking is used only once trained and clearly ahead (call
// after the engine stream waited for `trained` and synchronised).
static void judge_resolve(MlpState &m) {
    if (!m.judge_pending) return;
    const double *sums = m.h_judge;
    m.mlp_capture = sums[2] > 0 ? sums[0] / sums[2] : 1.0;
    m.hot_capture = sums[2] > 0 ? sums[1] / sums[2] : 1.0;
    m.fallback = m.steps < 8 || !(m.mlp_capture > 1.02 * m.hot_capture);
    if (m.steps >= 8) m.rejected = m.fallback ? m.rejected + 1 : 0;
    m.judge_pending = false;
}

void mlp_observe(Engine &e) {
    MlpState &m = *e.cache.mlp;
    if (!m.have_decision) return;
    const uint64_t C = e.cache.chunks;
    // the ring is read by a training launch that may still run
    mlp_join(e);
    // label the decision's features with what the snapshot actually read
    DCG_HIP(hipMemcpyAsync(m.X.p + (uint64_t)m.slot * C * kIn, m.Xcur.p, C * kIn * 4, hipMemcpyDeviceToDevice,
                           e.stream));
    hipLaunchKernelGGL(k_labels, dim3(grid_for(C)), dim3(256), 0, e.stream, e.cache.d_reads.p, C,
                       m.Y.p + (uint64_t)m.slot * C * kOut);
    // judge: which ranking would have captured more of the observed reads
    judge_launch(e, m);
    m.filled = std::min(kReplay, m.filled + 1);
    m.slot = (m.slot + 1) % kReplay;
    m.have_decision = false;
}

void mlp_release_training(Engine &e) {
    if (!e.cache.mlp) return;
    MlpState &m = *e.cache.mlp;
    if (!m.train_due) return;
    DCG_HIP(hipEventRecord(m.released, e.stream));
    DCG_HIP(hipStreamWaitEvent(m.train, m.released, 0));
    mlp_train_steps(m, 4);
    m.train_due = false;
}

void mlp_join(Engine &e) {
    if (!e.cache.mlp || e.cache.mlp->joined) return;
    DCG_HIP(hipStreamWaitEvent(e.stream, e.cache.mlp->trained, 0));
    e.cache.mlp->joined = true;
}

void mlp_sync(Engine &e) {
    if (e.cache.mlp && e.cache.mlp->train) DCG_HIP(hipStreamSynchronize(e.cache.mlp->train));
}

void mlp_stats(const Engine &e, dcg_cache_stats *st) {
    const MlpState &m = *e.cache.mlp;
    st->learned_decisions = (uint64_t)(m.decisions - m.fallbacks);
    st->fallback_decisions = (uint64_t)m.fallbacks;
    st->mlp_capture = m.mlp_capture;
    st->hot_capture = m.hot_capture;
}

uint64_t Engine::mlp_bytes() const {
    if (!cache.mlp) return 0;
    const MlpState &m = *cache.mlp;
    return m.master.bytes() + m.grad.bytes() + m.m1.bytes() + m.m2.bytes() + m.part.bytes() + m.w16.bytes() +
           m.X.bytes() + m.Y.bytes() + m.out.bytes() + m.degsum.bytes() + m.lastuse.bytes() + m.Xcur.bytes() +
           m.mscore.bytes() + m.hot_dec.bytes() + m.plan.bytes() + m.sort_k.bytes() + m.sort_v.bytes() +
           m.judge.bytes() + m.judge_reads.bytes() + m.judge_tmp.bytes();
}

std::vector<uint64_t> mlp_plan(Engine &e, uint64_t ab, uint64_t ac, uint64_t db, uint64_t dc) {
    MlpState &m = *e.cache.mlp;
    const uint64_t C = e.cache.chunks;
    if (m.rejected >= kQuiet && m.decisions < m.next_probe && e.cache.hot_score.size() == C && !m.train_due) {
        m.decisions++;
        m.fallbacks++;
        return e.plan_admission(e.cache.hot_score, 4.0 * kLoadLines, 1.5);
    }
    if (m.train_due) {   // no delete stage ran since the last decision
        mlp_train_steps(m, 4);
        m.train_due = false;
    }
    mlp_join(e);   // the forward pass reads the trained weights
    fill<float>(e.stream, e.cache.d_touch.p, C, 0.f);
    const dcg_edge *adds = e.batch_records(Engine::kAdds, ab, ac), *dels = e.batch_records(Engine::kDels, db, dc);
    hipLaunchKernelGGL(k_touch, dim3(grid_for(ac + dc, 256, 8192)), dim3(256), 0, e.stream, e.d_out.p, e.d_in.p,
                       adds, ac, dels, dc, e.n, e.cache.d_touch.p);
    e.mark_batches_used();
    std::vector<float> lu(C);
    for (uint64_t c = 0; c < C; ++c) lu[c] = (float)e.cache.last_use[c];
    DCG_HIP(hipMemcpyAsync(m.lastuse.p, lu.data(), C * 4, hipMemcpyHostToDevice, e.stream));
    float *X = m.Xcur.p;
    hipLaunchKernelGGL(k_features, dim3(grid_for(C)), dim3(256), 0, e.stream, e.cache.d_hist.p,
                       e.cache.d_rows_last.p, m.degsum.p, e.cache.d_touch.p, m.lastuse.p, (float)e.snapshots, C, X);
    hipLaunchKernelGGL(k_mlp_tile, dim3(mlp_grid(C)), dim3(64), 0, e.stream, X, (const float *)nullptr, C, m.w16.p,
                       m.master.p, m.out.p, (float *)nullptr);
    // both rankings of this decision stay on the device for the judge; the
    // plan reads back the one it follows
    hipLaunchKernelGGL(k_plan_scores, dim3(grid_for(C)), dim3(256), 0, e.stream, m.out.p, e.cache.d_hist.p, C,
                       m.plan.p, m.mscore.p, m.hot_dec.p);
    DCG_HIP(hipGetLastError());
    std::vector<float> hot(C), net(C);
    DCG_HIP(hipMemcpyAsync(hot.data(), m.hot_dec.p, C * 4, hipMemcpyDeviceToHost, e.stream));
    DCG_HIP(hipMemcpyAsync(net.data(), m.plan.p, C * 4, hipMemcpyDeviceToHost, e.stream));
    e.sync();
    judge_resolve(m);   // the previous decision's judge (the stream waited for it above)
    if (m.rejected >= kQuiet)
        m.next_probe = m.decisions + 1 + std::min(kProbeMax, kProbe << std::min(m.rejected - kQuiet, 4));
    // online training on the labelled decisions so far overlaps the next
    // delete mutation (mlp_release_training)
    m.train_due = true;
    m.decisions++;
    m.fallbacks += m.fallback ? 1 : 0;
    m.have_decision = true;
    if (m.fallback) return e.plan_admission(hot, 4.0 * kLoadLines, 1.5);
    // the network predicts one snapshot's line reads: admit above a load's cost
    return e.plan_admission(net, kLoadLines, 1.5);
}

}  // namespace dcg

namespace dcg {

// Workspace of dcg_mlp_eval_device, which allocates nothing: the per-workgroup
// gradient rows of one pass (kMlpGrid x P_NUM floats, 19.5 MiB) followed by the
// FP16 copy of the parameters.  Every GPU the library runs on holds its own
// copy until the process ends.  Host code never passes the array's name to a
// kernel: it uses the device address hipGetSymbolAddress returns for the
// current device (scratch_ptr).
constexpr size_t kScratchPart = (size_t)kMlpGrid * P_NUM;
constexpr size_t kScratchFloats = kScratchPart + (P_NUM + 1) / 2;
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr uint32_t kNoAllocMaxRows = 1u << 24;   // rows of one dcg_mlp_eval_device call
constexpr int kMaxDevices = 64;

// Device address of `scratch` on `device` (the current device), looked up once
// per device.
static float *scratch_ptr(int device) {
    static std::mutex mu;
    static float *ptr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) throw Error(DCG_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        void *p = nullptr;
        DCG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[device] = static_cast<float *>(p);
    }
    return ptr[device];
}

}  // namespace dcg

// The same passes as dcg_mlp_eval on caller device buffers, with every
// intermediate in the library's static workspace: no allocation of any kind.
// Runs on the null stream and returns after the passes finished.  Calls that
// overlap on one GPU share its workspace and race.
extern "C" int dcg_mlp_eval_device(int32_t device, uint32_t rows, const float *d_x, const float *d_y,
                                   const float *d_params, float *d_out, float *d_grad) {
    using namespace dcg;
    if (!d_x || !d_params || !d_out || rows == 0) return DCG_ERR_ARG;
    if (rows > kNoAllocMaxRows) return DCG_ERR_CAPACITY;
    if (d_grad && !d_y) return DCG_ERR_ARG;
    try {
        DCG_HIP(hipSetDevice(device));
        float *ws = scratch_ptr(device);
        float *part = ws;
        _Float16 *w16 = reinterpret_cast<_Float16 *>(ws + kScratchPart);
        const hipStream_t s = nullptr;
        hipLaunchKernelGGL(k_to_half, dim3(grid_for(P_NUM)), dim3(256), 0, s, d_params, w16, P_NUM);
        hipLaunchKernelGGL(k_mlp_tile, dim3(mlp_grid(rows)), dim3(64), 0, s, d_x, (const float *)nullptr,
                           (uint64_t)rows, w16, d_params, d_out, (float *)nullptr);
        if (d_y && d_grad) {
            // mlp_grid(rows) <= kMlpGrid workgroup rows: within the workspace
            mlp_gradient(s, d_x, d_y, rows, w16, d_params, part, d_grad);
        }
        DCG_HIP(hipGetLastError());
        DCG_HIP(hipStreamSynchronize(s));
    } catch (const Error &e) {
        return e.status;
    }
    return DCG_OK;
}

// Evaluate the predictor on caller-provided rows and FP32 parameters (rounded to
// FP16 exactly as the policy does): outputs, and with y != NULL the gradient of
// the mean loss.  Host pointers.  Exposed so the device passes can be checked
// against an FP64 reference; the policy itself never calls it.
extern "C" int dcg_mlp_eval(int32_t device, uint32_t rows, const float *x, const float *y, const float *params,
                            float *out, float *grad) {
    using namespace dcg;
    if (!x || !params || !out || rows == 0) return DCG_ERR_ARG;
    try {
        DCG_HIP(hipSetDevice(device));
        hipStream_t s;
        DCG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DevBuf<float> dx, dy, dp, dout, dg;
        DevBuf<_Float16> w16;
        dx.alloc((uint64_t)rows * kIn);
        dp.alloc(P_NUM);
        dout.alloc((uint64_t)rows * kOut);
        dg.alloc(P_NUM);
        w16.alloc(P_NUM);
        DCG_HIP(hipMemcpyAsync(dx.p, x, (uint64_t)rows * kIn * 4, hipMemcpyHostToDevice, s));
        DCG_HIP(hipMemcpyAsync(dp.p, params, P_NUM * 4, hipMemcpyHostToDevice, s));
        if (y) {
            dy.alloc((uint64_t)rows * kOut);
            DCG_HIP(hipMemcpyAsync(dy.p, y, (uint64_t)rows * kOut * 4, hipMemcpyHostToDevice, s));
        }
        hipLaunchKernelGGL(k_to_half, dim3(grid_for(P_NUM)), dim3(256), 0, s, dp.p, w16.p, P_NUM);
        hipLaunchKernelGGL(k_mlp_tile, dim3(mlp_grid(rows)), dim3(64), 0, s, dx.p, (const float *)nullptr,
                           (uint64_t)rows, w16.p, dp.p, dout.p, (float *)nullptr);
        if (y) {
            DevBuf<float> part;
            part.alloc((uint64_t)kMlpGrid * P_NUM);
            mlp_gradient(s, dx.p, dy.p, rows, w16.p, dp.p, part.p, dg.p);
            DCG_HIP(hipStreamSynchronize(s));
        }
        DCG_HIP(hipGetLastError());
        DCG_HIP(hipMemcpyAsync(out, dout.p, (uint64_t)rows * kOut * 4, hipMemcpyDeviceToHost, s));
        if (y && grad) DCG_HIP(hipMemcpyAsync(grad, dg.p, P_NUM * 4, hipMemcpyDeviceToHost, s));
        DCG_HIP(hipStreamSynchronize(s));
        DCG_HIP(hipStreamDestroy(s));
    } catch (const Error &e) {
        return e.status;
    }
    return DCG_OK;
}
