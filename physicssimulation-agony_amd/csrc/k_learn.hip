// Learned warm start inside the step: contact features, the network's initial impulses, training
// samples from the solver's own accumulators, and the residual gate.
//
// The reference warm-starts every contact point from the accumulators of the matching point of
// the previous step (NarrowPhaseCollisionDetector.cpp:753-765, applied by applyWarmStarting,
// BodyCollisionSolver.cpp:416-490). Here the network predicts a correction to that cached value:
//   jn = max(0, jn_cached + S * y0),  jt = jt_cached + S * y1,
// with S = normalMass * |g| * dt, the impulse that cancels one step of gravity along the normal.
// An all-zero network therefore reproduces the cached warm start bit for bit, and the same
// corrections, (jn_final - jn_cached) / S and (jt_final - jt_cached) / S, are the training targets.
//
// Features of contact point p of a manifold (all dimensionless, clamped to [-64, 64]):
//   0, 1   cached jn, jt of the point / S          2, 3   cached jn, jt of the other point / S
//   4      1 if the point had a cached impulse      5      1 if the manifold has two points
//   6, 7   relative normal, tangent velocity / (|g| dt) before the warm start
//   8      1 if the point has accumulators from two steps ago
//   9      manifold depth / (|g| dt^2)              10, 11 normal along and across "up" (-g)
//   12-15  invMassA, invMassB, rAn^2 invInertiaA, rBn^2 invInertiaB, each times normalMass
//   16     tangentMass / normalMass
//   17, 18 change of the cached jn, jt over the previous step / S (0 without history)
//   19, 20 1 if body A, B is static                 21, 22 contacts of body A, B / 8
//   23     relative normal velocity of the other point / (|g| dt)
// Each manifold also records the cached accumulators it started from (warmHist), which the next
// step's narrow phase carries over as the history of features 8, 17 and 18.
#include "agx_internal.h"
#include "mlp_internal.h"

namespace agx {

namespace {

constexpr float FEATURE_CLAMP = 64.0f;

struct Scale {
    float gdt;        // |g| dt
    float gdt2;       // |g| dt^2
    float upX, upY;   // -g / |g|
};

__device__ inline float clampFeature(float v) {
    if (!(v == v)) return 0.0f;
    return fminf(fmaxf(v, -FEATURE_CLAMP), FEATURE_CLAMP);
}

// Normal mass of point p at a solve slot (0: a point that takes no impulse).
__device__ inline float normalMassOf(const Constraints& cs, uint32_t slot, uint32_t p) {
    const float4 k0 = cs.k0[slot];
    return p == 0 ? k0.x : k0.w;
}

__global__ void kLearnRows(Counters* ctr, uint32_t predict, uint32_t gated, uint32_t collect) {
    const uint32_t use = predict ? (gated ? ctr->gate.use : 1u) : 0u;
    const uint32_t rows = 2u * ctr->manifolds;
    ctr->learnUsed = use;
    ctr->learnRows = (use || collect) ? rows : 0u;
    ctr->learnPredictRows = use ? rows : 0u;
}

// Features of contact point p of manifold m (see the list above); false for a point that takes no
// impulse (the row stays all zero).
__device__ inline bool pointFeatures(const Constraints& cs, uint32_t slot, uint32_t p, const float4* __restrict__ vel,
                                     const uint32_t* __restrict__ degree, float depth, float4 hist, const Scale& sc,
                                     float (&f)[MLP_INP]) {
    for (uint32_t k = 0; k < MLP_INP; k++) f[k] = 0.0f;
    const uint4 id = cs.idx[slot];
    const float4 k0 = cs.k0[slot];
    const float nMass[2] = { k0.x, k0.w };
    const float S = nMass[p] * sc.gdt;
    if (!(p < id.z && S > 0.0f)) return false;
    const float4 j = cs.imp[slot];
    const float4 nr = cs.nrm[slot], iv = cs.inv[slot], r0 = cs.r0[slot], r1 = cs.r1[slot];
    const float4 k1 = cs.k1[slot];
    const float4 vA = vel[id.x], vB = vel[id.y];
    const float hn[2] = { hist.x, hist.z }, ht[2] = { hist.y, hist.w };
    const float nx = nr.x, ny = nr.y, tx = -ny, ty = nx;
    const float dvx = vB.x - vA.x, dvy = vB.y - vA.y;
    const float dvn = dvx * nx + dvy * ny, dvt = dvx * tx + dvy * ty;
    const float vn[2] = { dvn + r0.y * vB.z - r0.x * vA.z, dvn + r1.y * vB.z - r1.x * vA.z };
    const float vt[2] = { dvt + r0.w * vB.z - r0.z * vA.z, dvt + r1.w * vB.z - r1.z * vA.z };
    const float tMass[2] = { k0.y, k1.x };
    const float rAn[2] = { r0.x, r1.x }, rBn[2] = { r0.y, r1.y };
    const float jn[2] = { j.x, j.z }, jt[2] = { j.y, j.w };
    const float degA = iv.x != 0.0f ? (float)degree[id.x] : 0.0f;
    const float degB = iv.y != 0.0f ? (float)degree[id.y] : 0.0f;
    const float invG = 1.0f / sc.gdt;
    const uint32_t o = 1u - p;
    const float invS = 1.0f / S;
    const bool two = id.z > 1;
    f[0] = jn[p] * invS;
    f[1] = jt[p] * invS;
    f[2] = two ? jn[o] * invS : 0.0f;
    f[3] = two ? jt[o] * invS : 0.0f;
    f[4] = (jn[p] != 0.0f || jt[p] != 0.0f) ? 1.0f : 0.0f;
    f[5] = two ? 1.0f : 0.0f;
    f[6] = vn[p] * invG;
    f[7] = vt[p] * invG;
    const bool hasHist = hn[p] == hn[p] && ht[p] == ht[p];
    f[8] = hasHist ? 1.0f : 0.0f;
    f[9] = depth / sc.gdt2;
    f[10] = nx * sc.upX + ny * sc.upY;
    f[11] = nx * sc.upY - ny * sc.upX;
    f[12] = iv.x * nMass[p];
    f[13] = iv.y * nMass[p];
    f[14] = rAn[p] * rAn[p] * iv.z * nMass[p];
    f[15] = rBn[p] * rBn[p] * iv.w * nMass[p];
    f[16] = tMass[p] / nMass[p];
    f[17] = hasHist ? (jn[p] - hn[p]) * invS : 0.0f;
    f[18] = hasHist ? (jt[p] - ht[p]) * invS : 0.0f;
    f[19] = iv.x == 0.0f ? 1.0f : 0.0f;
    f[20] = iv.y == 0.0f ? 1.0f : 0.0f;
    f[21] = degA * 0.125f;
    f[22] = degB * 0.125f;
    f[23] = two ? vn[o] * invG : 0.0f;
    return true;
}

__device__ inline void storeFeatureRow(uint16_t* __restrict__ x, uint64_t row, const float (&f)[MLP_INP]) {
    uint32_t packed[MLP_INP / 2];
    for (uint32_t k = 0; k < MLP_INP / 2; k++)
        packed[k] = (uint32_t)bf16Bits(clampFeature(f[2 * k])) | ((uint32_t)bf16Bits(clampFeature(f[2 * k + 1])) << 16);
    uint4* dst = reinterpret_cast<uint4*>(x + row * MLP_INP);
    for (uint32_t k = 0; k < MLP_INP / 8; k++)
        dst[k] = make_uint4(packed[4 * k], packed[4 * k + 1], packed[4 * k + 2], packed[4 * k + 3]);
}

// Wave-wide exclusive prefix of `count` plus one atomic claim of the wave's total on `counter`:
// the first of this lane's rows. Every lane of the wave must call it.
__device__ inline uint32_t claimRows(uint32_t count, uint32_t* counter) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    uint32_t prefix = count;   // inclusive prefix over the wave
    for (uint32_t off = 1; off < WAVE; off <<= 1) {
        const uint32_t v = __shfl_up(prefix, off, WAVE);
        if (lane >= off) prefix += v;
    }
    const uint32_t total = __shfl(prefix, WAVE - 1, WAVE);
    uint32_t row0 = 0;
    if (lane == WAVE - 1 && total) row0 = atomicAdd(counter, total);
    return __shfl(row0, WAVE - 1, WAVE) + prefix - count;
}

// Block-wide claim of rows on `counter`: the first of this thread's `count` rows. Every thread of
// the (FEATURE_THREADS-thread) block must call it; one atomic per block, since updates of one
// address from many blocks queue behind each other.
constexpr uint32_t FEATURE_THREADS = 1024;
__device__ inline uint32_t claimRowsBlock(uint32_t count, uint32_t* counter) {
    __shared__ uint32_t waveTotal[FEATURE_THREADS / WAVE];
    __shared__ uint32_t blockBase;
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    uint32_t prefix = count;   // inclusive prefix over the wave
    for (uint32_t off = 1; off < WAVE; off <<= 1) {
        const uint32_t v = __shfl_up(prefix, off, WAVE);
        if (lane >= off) prefix += v;
    }
    if (lane == WAVE - 1) waveTotal[wid] = prefix;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t k = 0; k < FEATURE_THREADS / WAVE; k++) total += waveTotal[k];
        blockBase = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    uint32_t base = blockBase;
    for (uint32_t k = 0; k < wid; k++) base += waveTotal[k];
    __syncthreads();   // waveTotal and blockBase are reused by the next call
    return base + prefix - count;
}

// Records every manifold's starting accumulators (the next step's history) and, when the network
// runs or rows are collected, the feature rows: two per manifold at rows 2m and 2m + 1, or with
// newOnly (no history kept) one per contact point without a cached impulse, in claimed rows with
// map[row] = 2m + p.
// Runs over solve positions, so the constraint planes are read in order (holes are skipped).
__global__ __launch_bounds__(FEATURE_THREADS) void kLearnFeatures(Counters* __restrict__ ctr, const agx_manifold* __restrict__ man,
                                                      const uint32_t* __restrict__ degree, const float4* __restrict__ vel,
                                                      Constraints cs, Scale sc, const float4* __restrict__ histOf,
                                                      float4* __restrict__ warmHist, uint32_t newOnly,
                                                      uint16_t* __restrict__ x, uint32_t* __restrict__ map) {
    const uint32_t P = ctr->colourStart[MAX_COLOURS];
    const bool rows = ctr->learnRows != 0;
    // block-uniform rounds, so that the row claims of newOnly stay block-wide
    for (uint32_t base = blockIdx.x * blockDim.x; base < P; base += gridDim.x * blockDim.x) {
        const uint32_t c = base + threadIdx.x;
        uint4 id = make_uint4(0u, 0u, 0u, 0u);
        float4 j = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (c < P) {
            id = cs.idx[c];
            if (id.z != 0) {
                j = cs.imp[c];
                if (!newOnly) warmHist[id.w] = j;   // new-contacts mode keeps no history
            }
        }
        if (!rows) continue;
        const uint32_t m = id.w;
        float f[MLP_INP];
        if (!newOnly) {
            if (id.z == 0) continue;
            const float depth = man[m].depth;
            const float4 hist = histOf[m];
            for (uint32_t p = 0; p < 2; p++) {
                pointFeatures(cs, c, p, vel, degree, depth, hist, sc, f);
                storeFeatureRow(x, 2ull * m + p, f);
            }
            continue;
        }
        uint32_t keep = 0;
        if (id.z != 0) {
            const uint32_t count = min(id.z, 2u);
            const float4 k0 = cs.k0[c];
            const float jn[2] = { j.x, j.z }, jt[2] = { j.y, j.w }, nm[2] = { k0.x, k0.w };
            for (uint32_t p = 0; p < count; p++)
                if (jn[p] == 0.0f && jt[p] == 0.0f && nm[p] > 0.0f) keep |= 1u << p;
        }
        uint32_t row = claimRowsBlock(__builtin_popcount(keep), &ctr->learnCompactRows);
        if (!keep) continue;
        const float depth = man[m].depth;
        const float none = __builtin_nanf("");
        const float4 hist = make_float4(none, none, none, none);
        for (uint32_t p = 0; p < 2; p++) {
            if (!(keep >> p & 1u)) continue;
            pointFeatures(cs, c, p, vel, degree, depth, hist, sc, f);
            storeFeatureRow(x, row, f);
            map[row] = 2u * m + p;
            row++;
        }
    }
}

// New-contacts mode: the network runs on the claimed rows only, on steps the gate gave to it.
__global__ void kLearnCompactRows(Counters* ctr) {
    ctr->learnPredictRows = ctr->learnUsed ? ctr->learnCompactRows : 0u;
}

// Initial accumulators from the network's corrections (only on steps the gate gave to it).
__global__ __launch_bounds__(256) void kLearnApply(const Counters* __restrict__ ctr, const uint32_t* __restrict__ posOf,
                                                   Constraints cs, const float* __restrict__ y, float gdt) {
    if (!ctr->learnUsed) return;
    const uint32_t M = ctr->learnPredictRows / 2;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint32_t slot = posOf[m];
        const uint32_t count = cs.idx[slot].z;
        float4 j = cs.imp[slot];
        float jn[2] = { j.x, j.z }, jt[2] = { j.y, j.w };
        for (uint32_t p = 0; p < count && p < 2; p++) {
            const float S = normalMassOf(cs, slot, p) * gdt;
            const float y0 = y[(size_t)(2 * m + p) * MLP_OUT], y1 = y[(size_t)(2 * m + p) * MLP_OUT + 1];
            if (!(S > 0.0f) || !isfinite(y0) || !isfinite(y1)) continue;
            // corrections beyond the range of the training targets are not trusted
            jn[p] = fmaxf(0.0f, jn[p] + S * fminf(fmaxf(y0, -FEATURE_CLAMP), FEATURE_CLAMP));
            jt[p] = jt[p] + S * fminf(fmaxf(y1, -FEATURE_CLAMP), FEATURE_CLAMP);
        }
        cs.imp[slot] = make_float4(jn[0], jt[0], jn[1], jt[1]);
    }
}

// New-contacts mode: the correction of each claimed row's point. Each row owns one half of its
// slot's accumulators (jn, jt of one point), so rows of the same manifold never share a store.
__global__ __launch_bounds__(256) void kLearnApplyRows(const Counters* __restrict__ ctr, const uint32_t* __restrict__ posOf,
                                                       const uint32_t* __restrict__ map, Constraints cs,
                                                       const float* __restrict__ y, float gdt) {
    const uint32_t n = ctr->learnPredictRows;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const uint32_t pt = map[r], m = pt >> 1, p = pt & 1u;
        const uint32_t slot = posOf[m];
        const float S = normalMassOf(cs, slot, p) * gdt;
        const float y0 = y[(size_t)r * MLP_OUT], y1 = y[(size_t)r * MLP_OUT + 1];
        if (!(S > 0.0f) || !isfinite(y0) || !isfinite(y1)) continue;
        float2* half = reinterpret_cast<float2*>(cs.imp + slot) + p;
        const float2 j = *half;
        *half = make_float2(fmaxf(0.0f, j.x + S * fminf(fmaxf(y0, -FEATURE_CLAMP), FEATURE_CLAMP)),
                            j.y + S * fminf(fmaxf(y1, -FEATURE_CLAMP), FEATURE_CLAMP));
    }
}

// Training rows from this step: features of every contact point and, as targets, the corrections
// that would have turned the cached accumulators into the solver's final ones. Runs after the
// sweeps and before the accumulators are written back to the manifolds (which still hold the
// cached values). Rows are claimed per wave; rows past the buffer's capacity are dropped.
__global__ __launch_bounds__(256) void kLearnCollect(Counters* __restrict__ ctr, const agx_manifold* __restrict__ man,
                                                     const uint32_t* __restrict__ posOf, Constraints cs,
                                                     const uint16_t* __restrict__ x, const uint32_t* __restrict__ map,
                                                     float gdt, uint32_t newOnly, uint16_t* __restrict__ sx,
                                                     float* __restrict__ st, uint32_t* __restrict__ sampleCount,
                                                     uint64_t cap) {
    // newOnly: one row per claimed feature row (points without a cached impulse); else the points
    // of every manifold, features at rows 2m + p
    const uint32_t n = newOnly ? ctr->learnCompactRows : ctr->learnRows / 2;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane of a wave runs the same number of rounds so the wave-wide claim stays uniform
    for (uint32_t base = first - lane; base < n; base += stride) {
        const uint32_t k = base + lane;
        uint32_t m = 0, slot = 0, keep = 0;   // bit p: point p of manifold m gives a row
        if (k < n) {
            if (newOnly) {
                m = map[k] >> 1;
                keep = 1u << (map[k] & 1u);
                slot = posOf[m];
            } else {
                m = k;
                slot = posOf[m];
                keep = (1u << min(cs.idx[slot].z, 2u)) - 1u;
            }
        }
        const uint32_t count = __builtin_popcount(keep);
        const uint32_t row0 = claimRows(count, sampleCount);
        if (count == 0) continue;
        const agx_manifold d = man[m];
        const float4 j = cs.imp[slot];
        const float jnC[2] = { d.jn0, d.jn1 }, jtC[2] = { d.jt0, d.jt1 };
        const float jnF[2] = { j.x, j.z }, jtF[2] = { j.y, j.w };
        uint64_t row = row0;
        for (uint32_t p = 0; p < 2; p++) {
            if (!(keep >> p & 1u)) continue;
            if (row >= cap) break;
            const float S = normalMassOf(cs, slot, p) * gdt;
            const float invS = S > 0.0f ? 1.0f / S : 0.0f;
            const uint64_t srcRow = newOnly ? k : 2ull * m + p;
            const uint4* src = reinterpret_cast<const uint4*>(x + srcRow * MLP_INP);
            uint4* dst = reinterpret_cast<uint4*>(sx + row * MLP_INP);
            for (uint32_t q = 0; q < MLP_INP / 8; q++) dst[q] = src[q];
            st[row * MLP_OUT] = clampFeature((jnF[p] - jnC[p]) * invS);
            st[row * MLP_OUT + 1] = clampFeature((jtF[p] - jtC[p]) * invS);
            row++;
        }
    }
}

// The step's cost for the gate: log10 of the relative residual after velocity sweep GATE_SWEEP (or
// after the last sweep when fewer ran); a better warm start leaves less for the early sweeps to
// change, and the early residual is far less noisy than the number of sweeps a step needs, which a
// collapsing stack somewhere in a batch solved as one can drive to the cap whatever the warm start.
// A batch solved scene by scene to a residual target is judged by what the warm start is for: the
// mean over dynamic bodies of the sweeps their scene needed, which one collapsing scene moves little.
constexpr uint32_t GATE_SWEEP = 8;
__global__ void kGateUpdate(Counters* ctr, uint32_t nDynBodies, uint32_t toTarget) {
    const uint32_t sweeps = min(ctr->sweeps, MAX_SWEEPS);
    if (sweeps == 0) return;
    float cost;
    if (ctr->perScene && toTarget) {
        cost = (float)((double)ctr->bodySweeps / (double)max(nDynBodies, 1u));
    } else {
        const uint32_t s = min(sweeps, GATE_SWEEP) - 1;
        const double num = ctr->residualNum[s], den = ctr->residualDen[s];
        const float res = (float)(sqrt(num) / fmax(sqrt(den), 1e-12));
        cost = log10f(fmaxf(res, 1e-12f));
    }
    LearnGate g = ctr->gate;
    gateUpdate(g, ctr->learnUsed, cost);
    ctr->gate = g;
}

inline uint32_t gridOf(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)); }

Scale scaleOf(float gx, float gy, float dt) {
    const float g = std::sqrt(gx * gx + gy * gy);
    Scale s;
    const float gs = std::max(g, 1.0f);
    s.gdt = gs * dt;
    s.gdt2 = gs * dt * dt;
    s.upX = g > 0.0f ? -gx / g : 0.0f;
    s.upY = g > 0.0f ? -gy / g : 1.0f;
    return s;
}

}  // namespace

hipError_t launchLearnBeforeWarm(hipStream_t st, DeviceWorld& w, const MlpDevice* mlp, uint32_t predict, uint32_t gated,
                                 uint32_t collect, uint32_t newOnly, float gx, float gy, float dt) {
    hipLaunchKernelGGL(kLearnRows, dim3(1), dim3(1), 0, st, w.ctr, predict && mlp ? 1u : 0u, gated, collect);
    const Scale sc = scaleOf(gx, gy, dt);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((w.manCap + FEATURE_THREADS - 1) / FEATURE_THREADS, 4096));
    hipLaunchKernelGGL(kLearnFeatures, dim3(grid), dim3(FEATURE_THREADS), 0, st, w.ctr, w.man[w.cur], w.degree, w.vel, w.cs, sc,
                       w.histOf, w.warmHist[w.cur], newOnly, w.learnX, w.learnMap);
    if (newOnly) hipLaunchKernelGGL(kLearnCompactRows, dim3(1), dim3(1), 0, st, w.ctr);
    if (predict && mlp) {
        const uint32_t rows = (uint32_t)std::min<uint64_t>(2 * w.manCap, 0xFFFFFFF0u);
        hipError_t e = launchMlpForward(st, *mlp, w.learnX, rows, w.learnY, nullptr, nullptr, &w.ctr->learnPredictRows);
        if (e != hipSuccess) return e;
        if (newOnly)
            hipLaunchKernelGGL(kLearnApplyRows, dim3(grid), dim3(256), 0, st, w.ctr, w.posOf, w.learnMap, w.cs, w.learnY, sc.gdt);
        else
            hipLaunchKernelGGL(kLearnApply, dim3(grid), dim3(256), 0, st, w.ctr, w.posOf, w.cs, w.learnY, sc.gdt);
    }
    return hipGetLastError();
}

hipError_t launchLearnCollect(hipStream_t st, DeviceWorld& w, const MlpSamples& smp, uint32_t newOnly, float gx,
                              float gy, float dt) {
    const Scale sc = scaleOf(gx, gy, dt);
    const uint32_t grid = gridOf(std::min<uint64_t>(w.manCap, 1u << 22));
    hipLaunchKernelGGL(kLearnCollect, dim3(grid), dim3(256), 0, st, w.ctr, w.man[w.cur], w.posOf, w.cs, w.learnX, w.learnMap,
                       sc.gdt, newOnly, smp.x, smp.t, smp.count, smp.cap);
    return hipGetLastError();
}

hipError_t launchGateUpdate(hipStream_t st, DeviceWorld& w, uint32_t toTarget) {
    hipLaunchKernelGGL(kGateUpdate, dim3(1), dim3(1), 0, st, w.ctr, w.nDynBodies, toTarget);
    return hipGetLastError();
}

}  // namespace agx
