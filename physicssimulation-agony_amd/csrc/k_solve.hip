// Contact solver (Solving/BodyCollisionSolver.cpp) and islands (IslandManager.cpp).
//
// The reference solves its manifolds one after another (Gauss-Seidel) in the order of its planner
// waves (Solving/SolvingPlanner.cpp:141-310). Here the order is colour-major: constraints of one
// colour share no dynamic body, so every order inside a colour gives the same bits. The device
// runs that sequential order as a dataflow instead of colour by colour behind grid barriers:
// every dynamic body keeps records whose sequence counts the operations finished on it, and a
// constraint runs as soon as the records of both its bodies show that every constraint before it
// (by sweep, then by solve position) touching them is done. The result equals a sequential sweep
// in the exported order exactly, and later colours and sweeps start as soon as their own inputs
// are ready.
//
// The colouring is a dataflow too: greedy first fit in a fixed pseudo-random priority order
// (each manifold takes the lowest colour not taken at its bodies by manifolds of higher
// priority). It is deterministic and its dependency chains are short.
#include "agx_internal.h"

namespace agx {

namespace {

constexpr uint64_t STALL_TICKS = 200000000ull;   // 2 s of the 100 MHz wall clock
// A waiting constraint whose body counter is d short of its target cannot run before d - 1 more
// constraints finish there, each at least one coherent memory round trip (about 1 us): it polls
// again only after that long, which keeps the polling traffic off the critical path.
constexpr uint64_t BACKOFF_TICKS_PER_HOP = 100;   // 1 us
constexpr uint32_t FINISH_BLOCKS = SOLVE_FINISH_BLOCKS;

inline uint32_t blocksFor(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 8192u)); }
inline uint32_t blocksForManifolds(const DeviceWorld& w) {
    return blocksFor((uint32_t)std::min<uint64_t>(w.manCap, 1u << 22));
}

// Bijective on 32 bits, so priorities of distinct manifolds never tie.
__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

__device__ inline bool isDynamic(const float2* inv, uint32_t b) { return inv[b].x != 0.0f; }

// ---------------------------------------------------------------- incidence lists
__device__ inline unsigned long long packWord(uint32_t seq, uint32_t bits) {
    return ((unsigned long long)seq << 32) | bits;
}
__device__ inline uint32_t seqOf(unsigned long long w) { return (uint32_t)(w >> 32); }
__device__ inline uint32_t bitsOf(unsigned long long w) { return (uint32_t)w; }

__global__ __launch_bounds__(256) void kFlowClear(uint32_t n, uint32_t* __restrict__ degree, uint32_t* __restrict__ fill,
                                                  unsigned long long* __restrict__ colRec) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        degree[i] = 0; fill[i] = 0;
        for (uint32_t k = 0; k < 4; k++) colRec[4 * i + k] = 0;
    }
}

__global__ __launch_bounds__(256) void kDegree(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                               const uint32_t* __restrict__ colBody, const float2* __restrict__ inv,
                                               uint2* __restrict__ manBodies, uint32_t* __restrict__ degree) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
        manBodies[m] = make_uint2(a, b);
        if (isDynamic(inv, a)) atomicAdd(degree + a, 1u);
        if (isDynamic(inv, b)) atomicAdd(degree + b, 1u);
    }
}

__global__ __launch_bounds__(256) void kIncident(const Counters* __restrict__ ctr, const uint2* __restrict__ manBodies,
                                                 const float2* __restrict__ inv, const uint32_t* __restrict__ incStart,
                                                 uint32_t* __restrict__ fill, uint32_t* __restrict__ inc) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint2 mb = manBodies[m];
        if (isDynamic(inv, mb.x)) inc[incStart[mb.x] + atomicAdd(fill + mb.x, 1u)] = 2 * m;
        if (isDynamic(inv, mb.y)) inc[incStart[mb.y] + atomicAdd(fill + mb.y, 1u)] = 2 * m + 1;
    }
}

// Rank of every list entry among its body's entries by a key: the manifold priority
// (BY_POSITION false, for the colouring) or the solve position (BY_POSITION true, which also
// writes the lists sorted by position). One thread per list entry.
template <bool BY_POSITION>
__global__ __launch_bounds__(256) void kRankEntries(const Counters* __restrict__ ctr, const uint2* __restrict__ manBodies,
                                                    const uint32_t* __restrict__ degree,
                                                    const uint32_t* __restrict__ incStart, const uint32_t* __restrict__ inc,
                                                    const uint32_t* __restrict__ posOf, uint32_t* __restrict__ rank,
                                                    uint32_t* __restrict__ incSorted) {
    const uint32_t total = ctr->scratch[4];   // list entries (scan total of the degrees)
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x) {
        const uint32_t e = inc[k];
        const uint32_t m = e >> 1;
        const uint2 mb = manBodies[m];
        const uint32_t b = (e & 1u) ? mb.y : mb.x;
        const uint32_t s = incStart[b], d = degree[b];
        const uint32_t key = BY_POSITION ? posOf[m] : mix32(m);
        uint32_t r = 0;
        for (uint32_t j = 0; j < d; j++) {
            const uint32_t o = inc[s + j] >> 1;
            r += (BY_POSITION ? posOf[o] : mix32(o)) < key ? 1u : 0u;
        }
        rank[e] = r;
        if (BY_POSITION) incSorted[s + r] = e;
    }
}

// ---------------------------------------------------------------- dataflow colouring
// Lane t holds manifolds t, t + T, ... (T = grid threads) and takes them in priority order, so
// the manifold of lowest priority still open anywhere is always the next one of its lane: with
// every block resident the flow cannot deadlock. A manifold is ready when the colour records of
// both its bodies carry its rank there as sequence; it takes the first colour free at both and
// writes the records back with the next sequence.
__device__ inline bool readColours(const unsigned long long* rec, uint32_t expect, uint32_t* bits, uint32_t& dist) {
    unsigned long long w[4];
    for (uint32_t i = 0; i < 4; i++) w[i] = loadCoherent64(rec + i);
    bool ok = true;
    uint32_t lo = expect;
    for (uint32_t i = 0; i < 4; i++) {
        bits[i] = bitsOf(w[i]);
        ok = ok && seqOf(w[i]) == expect;
        lo = min(lo, seqOf(w[i]));
    }
    dist = max(dist, expect - lo);
    return ok;
}

__global__ __launch_bounds__(256) void kColourFlow(Counters* __restrict__ ctr, const uint2* __restrict__ manBodies,
                                                   const float2* __restrict__ inv, const uint32_t* __restrict__ rank,
                                                   unsigned long long* colRec, uint16_t* __restrict__ colourId) {
    const uint32_t M = ctr->manifolds;
    const uint32_t T = gridDim.x * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if ((uint64_t)M > (uint64_t)MAX_FLOW_ITEMS * T) {
        if (tid == 0) atomicOr(&ctr->overflow, OVERFLOW_ITEMS);
        return;
    }
    unsigned long long pending = 0;
    for (uint32_t k = 0; k < MAX_FLOW_ITEMS && (uint64_t)tid + (uint64_t)k * T < M; k++) pending |= 1ull << k;
    auto pick = [&]() -> uint32_t {
        uint32_t best = 0xFFFFFFFFu, bestK = MAX_FLOW_ITEMS;
        for (unsigned long long q = pending; q; q &= q - 1) {
            const uint32_t k = (uint32_t)__ffsll(q) - 1;
            const uint32_t pr = mix32(tid + k * T);
            if (pr < best) { best = pr; bestK = k; }
        }
        return bestK;
    };
    uint32_t k = pick();
    uint32_t m = 0, a = 0, b = 0, ra = 0, rb = 0;
    bool da = false, db = false;
    auto load = [&]() {
        m = tid + k * T;
        const uint2 mb = manBodies[m];
        a = mb.x; b = mb.y;
        da = isDynamic(inv, a); db = isDynamic(inv, b);
        ra = da ? rank[2 * m] : 0u;
        rb = db ? rank[2 * m + 1] : 0u;
    };
    if (k < MAX_FLOW_ITEMS) load();
    const uint64_t deadline = wall_clock64() + STALL_TICKS;
    uint64_t wake = 0;
    while (true) {
        bool progressed = false;
        const uint64_t now = wall_clock64();
        if (k < MAX_FLOW_ITEMS && now >= wake) {
            uint32_t ba[4] = { 0, 0, 0, 0 }, bb[4] = { 0, 0, 0, 0 };
            uint32_t dist = 1;
            bool ok = true;
            if (da) ok = readColours(colRec + 4 * (size_t)a, ra, ba, dist);
            if (db) ok = readColours(colRec + 4 * (size_t)b, rb, bb, dist) && ok;
            if (!ok) {
                wake = now + (uint64_t)(dist - 1) * BACKOFF_TICKS_PER_HOP;
            } else {
                uint32_t c = FIRST_FIT_COLOURS;
                for (uint32_t i = 0; i < 4; i++) {
                    const uint32_t free = ~(ba[i] | bb[i]);
                    if (free) { c = 32 * i + (uint32_t)__ffs((int)free) - 1; break; }
                }
                colourId[m] = (uint16_t)c;
                if (c == FIRST_FIT_COLOURS) atomicOr(&ctr->overflow, OVERFLOW_COLOURING);
                for (uint32_t i = 0; i < 4; i++) {
                    const uint32_t bit = (c < FIRST_FIT_COLOURS && (c >> 5) == i) ? (1u << (c & 31u)) : 0u;
                    if (da) storeCoherent64(colRec + 4 * (size_t)a + i, packWord(ra + 1, ba[i] | bit));
                    if (db) storeCoherent64(colRec + 4 * (size_t)b + i, packWord(rb + 1, bb[i] | bit));
                }
                pending &= ~(1ull << k);
                k = pick();
                if (k < MAX_FLOW_ITEMS) load();
                progressed = true;
            }
        }
        if (__all(k >= MAX_FLOW_ITEMS)) break;
        if (!__any(progressed)) {
            if (now > deadline) { atomicOr(&ctr->overflow, OVERFLOW_STALL); break; }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// ---------------------------------------------------------------- lanes of a scene
constexpr uint32_t SCENE_GROUPS = 4;       // lane groups of a block for the smaller scenes
constexpr uint32_t WORK_SORT_MAX = 4096;   // scenes the per-step work order sorts in LDS

// A value every lane of the wave holds, in a scalar register.
__device__ inline uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// The lanes that work on one scene: the whole block (block barriers) for the largest scenes, else
// a group of the block's lanes, whose waves meet at an arrival counter in LDS. Every lane of the
// group must call sync() the same number of times.
struct SceneLanes {
    uint32_t t;        // lane in the group
    uint32_t T;        // lanes in the group
    uint32_t g;        // group (0 for the whole block)
    uint32_t waves;    // waves in the group
    bool whole;
    uint32_t* bar;     // the group's arrival counter
    uint32_t* abort;   // the block's flag: a wait timed out, no sync waits any more
    uint32_t target;   // arrivals the next sync waits for
    Counters* ctr;

    __device__ void sync() {
        if (whole) { __syncthreads(); return; }
        target += waves;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if ((threadIdx.x & (WAVE - 1)) == 0) __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        // bounded: after about 2 s of 64-cycle sleeps the step is flagged invalid and every later
        // sync of the block returns at once, so the kernel ends
        for (uint32_t spins = 0; __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target;) {
            if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
            __builtin_amdgcn_s_sleep(1);
            if (++spins == (1u << 26)) {
                __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                atomicOr(&ctr->overflow, OVERFLOW_STALL);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

// ---------------------------------------------------------------- per-scene colouring
// The same greedy first fit in priority order as kColourFlow, one workgroup per scene, as rounds
// in LDS: a manifold takes its colour once it has the smallest priority among the uncoloured
// manifolds at each of its dynamic bodies. Then every higher-priority manifold at those bodies is
// coloured and no lower-priority one is, so the colours free at both bodies are exactly the
// sequential first fit's, and at most one manifold per body takes a colour in a round.
constexpr uint32_t COLOUR_THREADS = 1024;   // one block per CU
constexpr uint16_t UNCOLOURED = 0xFFFFu;

// Empties the scenes' manifold ranges (S) and their colour ranges (S x MAX_COLOURS) before the
// atomic min/max fill them.
__global__ __launch_bounds__(256) void kSceneRangesClear(uint2* __restrict__ manRange, uint2* __restrict__ colourRange,
                                                         uint32_t S) {
    const uint32_t n = S * (MAX_COLOURS + 1);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < S) manRange[i] = make_uint2(0xFFFFFFFFu, 0u);
        else colourRange[i - S] = make_uint2(0xFFFFFFFFu, 0u);
    }
}

// Ranges of keys over an index space: range[key] = (smallest index, one past the largest) holding
// that key. The indices come in runs of equal keys, so only a run's first and last index update
// the range (an atomic min and max per run, not per index); keyAt(i) is NO_KEY beyond the end.
// Every lane of the block's waves must call it with base uniform over the wave.
constexpr uint32_t NO_KEY = 0xFFFFFFFFu;
template <class KeyAt>
__device__ inline void runRanges(uint32_t i, KeyAt keyAt, uint2* __restrict__ range) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t key = keyAt(i);
    uint32_t prev = __shfl_up(key, 1, WAVE), next = __shfl_down(key, 1, WAVE);
    if (lane == 0) prev = i > 0 ? keyAt(i - 1) : NO_KEY;
    if (lane == WAVE - 1) next = keyAt(i + 1);
    if (key == NO_KEY) return;
    if (prev != key) atomicMin(&range[key].x, i);
    if (next != key) atomicMax(&range[key].y, i + 1);
}

// First manifold and one past the last of every scene (manifolds come grouped by scene up to a
// pair wave that straddles two scenes). Also clears manBodies of every manifold before the scenes
// fill in their own: a scene's manifold range can hold a few of a neighbouring scene's manifolds,
// which it recognises by bodies outside its own, whether their owner has written them yet or not.
__global__ __launch_bounds__(256) void kSceneManRanges(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                                       const uint32_t* __restrict__ colScene, uint2* __restrict__ range,
                                                       uint2* __restrict__ manBodies) {
    const uint32_t M = ctr->manifolds;
    auto keyAt = [&](uint32_t m) { return m < M ? colScene[man[m].a] : NO_KEY; };
    for (uint32_t base = blockIdx.x * blockDim.x; base < M; base += gridDim.x * blockDim.x) {
        const uint32_t m = base + threadIdx.x;
        if (m < M) manBodies[m] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
        runRanges(m, keyAt, range);
    }
}

// Also writes manBodies (the manifold's bodies) and degree (contacts per dynamic body) for the
// scene's manifolds and bodies, as kDegree does for the dataflow. The rounds work on a copy of the
// scene's manifolds (bodies packed with their dynamic flags, colour) in LDS when it fits after the
// body records, else on the same copy in global scratch (`spill`, `spillColour`).
__device__ inline uint32_t packManifold(uint32_t la, bool da, uint32_t lb, bool db) {
    return (la | (da ? 0x8000u : 0u)) | ((lb | (db ? 0x8000u : 0u)) << 16);
}

// The colouring of scene s by the lanes L, in the LDS at `lds` laid out as layout =
// (maxBodies << 16) | capM: body records for up to maxBodies bodies, then up to capM manifolds.
__device__ inline void colourScene(SceneLanes& L, uint32_t s, const agx_manifold* __restrict__ man, Counters* __restrict__ ctr,
                                   const uint32_t* __restrict__ colBody, const float2* __restrict__ inv,
                                   const uint32_t* __restrict__ sceneBodyStart, const uint2* __restrict__ manRange,
                                   uint32_t ldsManifolds, uint2* __restrict__ manBodies, uint32_t* __restrict__ degree,
                                   uint32_t* __restrict__ spill, uint16_t* __restrict__ colourId, uint32_t* lds, int* left) {
    const uint32_t t = L.t, T = L.T;
    const uint32_t b0 = sceneBodyStart[s], b1 = sceneBodyStart[s + 1], nb = b1 - b0;
    const uint32_t maxBodies = ldsManifolds >> 16;   // LDS layout: body records for the largest scene first
    const uint32_t capM = ldsManifolds & 0xFFFFu;
    // nb: the round's smallest priority of an uncoloured manifold at the body, tagged with the
    // round (~round in the high word, so an entry of an earlier round loses to any of this round's
    // and needs no clearing)
    unsigned long long* minPrio = reinterpret_cast<unsigned long long*>(lds);
    uint32_t* mask = lds + 2 * maxBodies;            // nb x 4: colours taken at the body
    uint32_t* packedL = lds + 6 * maxBodies;         // capM: the scene's manifolds (bodies, flags)
    uint16_t* colourL = reinterpret_cast<uint16_t*>(packedL + capM);   // capM colours
    const uint2 r = manRange[s];
    const uint32_t nm = r.y > r.x ? r.y - r.x : 0u;
    const bool inLds = nm <= capM;
    uint32_t* packed = inLds ? packedL : spill + r.x;
    uint16_t* col = inLds ? colourL : colourId + r.x;
    // a manifold of this scene: marked in LDS, or in the spill case recognised by its bodies
    // (manBodies was cleared before, and only a manifold's own scene writes its entry)
    auto own = [&](uint32_t i) -> bool {
        return inLds ? packed[i] != 0xFFFFFFFFu : manBodies[r.x + i].x - b0 < nb;
    };
    for (uint32_t i = t; i < nb; i += T) {
        minPrio[i] = ~0ull;
        for (uint32_t k = 0; k < 4; k++) mask[4 * i + k] = 0u;
    }
    if (t == 0) { left[0] = 0; left[1] = 0; }
    L.sync();
    // the scene's manifolds: bodies, degrees (counted in the first mask word), uncoloured
    for (uint32_t i = t; i < nm; i += T) {
        const uint32_t m = r.x + i;
        const agx_manifold d = man[m];
        const uint32_t a = colBody[d.a], b = colBody[d.b];
        if (a < b0 || a >= b1) {   // another scene's manifold (its global slot is its owner's)
            if (inLds) packed[i] = 0xFFFFFFFFu;
            continue;
        }
        manBodies[m] = make_uint2(a, b);
        const bool da = inv[a].x != 0.0f, db = inv[b].x != 0.0f;
        packed[i] = packManifold(a - b0, da, b - b0, db);
        col[i] = UNCOLOURED;
        if (da) atomicAdd(&mask[4 * (a - b0)], 1u);
        if (db) atomicAdd(&mask[4 * (b - b0)], 1u);
    }
    L.sync();
    for (uint32_t i = t; i < nb; i += T) {
        degree[b0 + i] = mask[4 * i];   // 0 for a static body: none of its contacts counted it
        mask[4 * i] = 0u;
    }
    L.sync();
    // round q: every uncoloured manifold posts its priority at its dynamic bodies and flags
    // left[q & 1]; after a sync the ones holding both minima take their colours. left[(q + 1) & 1]
    // is cleared during round q, which reads only left[q & 1].
    for (uint32_t q = 0;; q++) {
        const unsigned long long tag = (unsigned long long)(0xFFFFFFFFu - q) << 32;
        for (uint32_t i = t; i < nm; i += T) {
            if (!own(i)) continue;
            const uint32_t pk = packed[i];
            if (col[i] != UNCOLOURED) continue;
            const unsigned long long pr = tag | mix32(r.x + i);
            if (pk & 0x8000u) atomicMin(&minPrio[pk & 0x7FFFu], pr);
            if (pk & 0x80000000u) atomicMin(&minPrio[(pk >> 16) & 0x7FFFu], pr);
            left[q & 1] = 1;
        }
        if (t == 0) left[(q + 1) & 1] = 0;
        L.sync();
        if (!left[q & 1]) break;
        for (uint32_t i = t; i < nm; i += T) {
            if (!own(i)) continue;
            const uint32_t pk = packed[i];
            if (col[i] != UNCOLOURED) continue;
            const unsigned long long pr = tag | mix32(r.x + i);
            const uint32_t la = pk & 0x7FFFu, lb = (pk >> 16) & 0x7FFFu;
            const bool da = (pk & 0x8000u) != 0, db = (pk & 0x80000000u) != 0;
            if ((da && minPrio[la] != pr) || (db && minPrio[lb] != pr)) continue;
            uint32_t c = FIRST_FIT_COLOURS;
            for (uint32_t k = 0; k < 4; k++) {
                const uint32_t used = (da ? mask[4 * la + k] : 0u) | (db ? mask[4 * lb + k] : 0u);
                if (~used) { c = 32 * k + (uint32_t)__ffs((int)~used) - 1; break; }
            }
            col[i] = (uint16_t)c;
            if (c == FIRST_FIT_COLOURS) { atomicOr(&ctr->overflow, OVERFLOW_COLOURING); continue; }
            // no other manifold of this round reads these bodies' masks (it would need the
            // smallest priority here), so the bits can be set at once
            if (da) mask[4 * la + (c >> 5)] |= 1u << (c & 31u);
            if (db) mask[4 * lb + (c >> 5)] |= 1u << (c & 31u);
        }
        L.sync();
    }
    if (inLds)
        for (uint32_t i = t; i < nm; i += T)
            if (own(i)) colourId[r.x + i] = col[i];
}

// Blocks (one per CU) take scenes largest first from a cursor: the first nWhole whole, the rest
// by `groups` lane groups of groupWords words of LDS each (layoutGroup); a scene's colours do not
// depend on which block or group colours it.
__global__ __launch_bounds__(COLOUR_THREADS) void kSceneColour(const agx_manifold* __restrict__ man, Counters* __restrict__ ctr,
                                                               const uint32_t* __restrict__ colBody, const float2* __restrict__ inv,
                                                               const uint32_t* __restrict__ sceneBodyStart,
                                                               const uint32_t* __restrict__ sceneOrder,
                                                               const uint2* __restrict__ manRange, uint32_t S, uint32_t nWhole,
                                                               uint32_t groups, uint32_t layoutWhole, uint32_t layoutGroup,
                                                               uint32_t groupWords, uint2* __restrict__ manBodies,
                                                               uint32_t* __restrict__ degree, uint32_t* __restrict__ spill,
                                                               uint16_t* __restrict__ colourId) {
    // 16-byte aligned: the static arrays above it total 68 bytes, and the round-tagged minima are
    // 64-bit LDS atomics, which need 8-byte alignment
    extern __shared__ uint4 ldsBase[];
    uint32_t* lds = reinterpret_cast<uint32_t*>(ldsBase);
    __shared__ int left[SCENE_GROUPS][2];
    __shared__ uint32_t taken[SCENE_GROUPS];
    __shared__ uint32_t bar[SCENE_GROUPS];
    __shared__ uint32_t abort;
    if (threadIdx.x < SCENE_GROUPS) bar[threadIdx.x] = 0u;
    if (threadIdx.x == 0) abort = 0u;
    SceneLanes L{ threadIdx.x, COLOUR_THREADS, 0u, COLOUR_THREADS / WAVE, true, nullptr, &abort, 0u, ctr };
    uint32_t* region = lds;
    uint32_t layout = layoutWhole;
    bool grouped = false;
    while (true) {
        if (L.t == 0) taken[L.g] = grouped ? nWhole + atomicAdd(&ctr->colourCursor2, 1u) : atomicAdd(&ctr->colourCursor, 1u);
        L.sync();
        const uint32_t k0 = taken[L.g];
        if (!grouped && k0 >= nWhole) {
            __syncthreads();   // every wave has read taken[0] before the groups reuse it
            if (nWhole >= S) break;
            const uint32_t T = COLOUR_THREADS / groups;
            L.g = threadIdx.x / T;
            L.t = threadIdx.x % T;
            L.T = T;
            L.waves = T / WAVE;
            L.whole = groups == 1;
            L.bar = &bar[L.g];
            region = lds + (size_t)L.g * groupWords;
            layout = layoutGroup;
            grouped = true;
            continue;
        }
        if (k0 >= S) break;
        colourScene(L, sceneOrder[k0], man, ctr, colBody, inv, sceneBodyStart, manRange, layout, manBodies, degree, spill,
                    colourId, region, left[L.g]);
    }
}

// ---------------------------------------------------------------- solve order
// Block histograms over contiguous manifold chunks, one column scan, then a scatter that
// reserves each block's slots from its scanned base.
__device__ inline void chunkOf(uint32_t M, uint32_t& lo, uint32_t& hi) {
    const uint32_t chunk = (M + gridDim.x - 1) / gridDim.x;
    lo = blockIdx.x * chunk;
    hi = min(M, lo + chunk);
}

__global__ __launch_bounds__(256) void kColourHist(const Counters* __restrict__ ctr, const uint16_t* __restrict__ colourId,
                                                   uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[MAX_COLOURS];
    for (uint32_t i = threadIdx.x; i < MAX_COLOURS; i += blockDim.x) h[i] = 0;
    __syncthreads();
    uint32_t lo, hi;
    chunkOf(ctr->manifolds, lo, hi);
    for (uint32_t m = lo + threadIdx.x; m < hi; m += blockDim.x) atomicAdd(&h[colourId[m]], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < MAX_COLOURS; i += blockDim.x) hist[blockIdx.x * MAX_COLOURS + i] = h[i];
}

// One block per colour: exclusive scan of that colour's count over the histogram blocks (each
// block's first slot within the colour) and the colour's total.
__global__ __launch_bounds__(256) void kColourColumnScan(Counters* __restrict__ ctr, uint32_t* __restrict__ hist,
                                                         uint32_t blocks) {
    __shared__ uint32_t waveSum[4];
    const uint32_t c = blockIdx.x;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < blocks; base += blockDim.x) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < blocks ? hist[j * MAX_COLOURS + c] : 0u;
        // inclusive scan inside the wave, then across the four waves
        uint32_t x = v;
        for (uint32_t off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, WAVE);
            if ((threadIdx.x & (WAVE - 1)) >= off) x += y;
        }
        if ((threadIdx.x & (WAVE - 1)) == WAVE - 1) waveSum[threadIdx.x / WAVE] = x;
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t k = 0; k < threadIdx.x / WAVE; k++) before += waveSum[k];
        if (j < blocks) hist[j * MAX_COLOURS + c] = before + x - v;
        carry += waveSum[0] + waveSum[1] + waveSum[2] + waveSum[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) ctr->colourCount[c] = carry;
}

// Colour starts: every colour's slots start on a wave boundary (COLOUR_ALIGN), so a wave of the
// solve never mixes colours; the slots past a colour's count are holes.
__global__ void kColourStarts(Counters* __restrict__ ctr) {
    if (threadIdx.x != 0) return;
    uint32_t s = 0, K = 0;
    for (uint32_t i = 0; i < MAX_COLOURS; i++) {
        const uint32_t n = ctr->colourCount[i];
        ctr->colourStart[i] = s;
        s += (n + COLOUR_ALIGN - 1) / COLOUR_ALIGN * COLOUR_ALIGN;
        if (n) K = i + 1;
    }
    ctr->colourStart[MAX_COLOURS] = s;
    ctr->colours = K;
}

// Marks the padding slots at the end of every colour: no manifold (count 0), never waited on.
__global__ __launch_bounds__(256) void kMarkHoles(const Counters* __restrict__ ctr, uint32_t* __restrict__ order,
                                                  Constraints cs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = i / COLOUR_ALIGN, j = i % COLOUR_ALIGN;
    if (k >= MAX_COLOURS) return;
    const uint32_t p = ctr->colourStart[k] + ctr->colourCount[k] + j;
    if (p >= ctr->colourStart[k + 1]) return;
    order[p] = 0xFFFFFFFFu;
    cs.idx[p] = make_uint4(0u, 0u, 0u, 0xFFFFFFFFu);
    cs.dep[p] = make_uint4(0u, 0u, 0u, 0u);
}

// Stable: a colour's slots follow manifold order (blocks own contiguous chunks in order, a block
// takes its chunk in rounds of 256, waves in order, lanes in order), so the manifolds of one scene
// occupy a contiguous run of each colour's slots up to the few manifolds of a pair wave that
// straddles two scenes.
__global__ __launch_bounds__(256) void kColourScatter(const Counters* __restrict__ ctr, const uint16_t* __restrict__ colourId,
                                                      const uint32_t* __restrict__ hist, uint32_t* __restrict__ order,
                                                      uint32_t* __restrict__ posOf) {
    constexpr uint32_t COLOUR_BITS = 8;
    static_assert(MAX_COLOURS <= (1u << COLOUR_BITS), "colour ids fit the matched bits");
    __shared__ uint32_t cursor[MAX_COLOURS];
    __shared__ uint32_t waveCount[4][MAX_COLOURS];
    for (uint32_t i = threadIdx.x; i < MAX_COLOURS; i += blockDim.x)
        cursor[i] = ctr->colourStart[i] + hist[blockIdx.x * MAX_COLOURS + i];
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const uint64_t ltMask = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    uint32_t lo, hi;
    chunkOf(ctr->manifolds, lo, hi);
    for (uint32_t base = lo; base < hi; base += blockDim.x) {   // block-uniform rounds
        for (uint32_t i = threadIdx.x; i < 4 * MAX_COLOURS; i += blockDim.x) (&waveCount[0][0])[i] = 0;
        __syncthreads();
        const uint32_t m = base + threadIdx.x;
        const bool valid = m < hi;
        const uint32_t c = valid ? colourId[m] : 0u;
        // lanes of this wave with the same colour
        uint64_t same = __ballot(valid);
        for (uint32_t b = 0; b < COLOUR_BITS; b++) {
            const uint64_t ones = __ballot(valid && ((c >> b) & 1u));
            same &= ((c >> b) & 1u) ? ones : ~ones;
        }
        const uint32_t rank = (uint32_t)__popcll(same & ltMask);
        if (valid && (same >> lane) == 1ull) waveCount[wid][c] = (uint32_t)__popcll(same);   // the group's last lane
        __syncthreads();
        if (valid) {
            uint32_t p = cursor[c] + rank;
            for (uint32_t k = 0; k < wid; k++) p += waveCount[k][c];
            order[p] = m;
            posOf[m] = p;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < MAX_COLOURS; i += blockDim.x)
            cursor[i] += waveCount[0][i] + waveCount[1][i] + waveCount[2][i] + waveCount[3][i];
        __syncthreads();   // before the next round clears waveCount
    }
}

// 1/det of a two-point contact's 2x2 normal block, or NaN when the block is too ill-conditioned for
// the direct solve (BodyCollisionSolver.cpp block solver: maxK^2 < 1000 det).
__device__ inline float blockInvDet(float K00, float K11, float K01) {
    const float det = K00 * K11 - K01 * K01;
    const float maxK = fmaxf(K00, K11);
    return maxK * maxK < 1000.0f * det ? 1.0f / det : __builtin_nanf("");
}

// Predecessor of a constraint at one of its dynamic bodies in the solve order: the body's previous
// constraint in the same sweep, or, for the body's first constraint, its last constraint of the
// previous sweep (flagged PRED_PREV_SWEEP). NO_PRED: a static body.
constexpr uint32_t NO_PRED = 0xFFFFFFFFu;
constexpr uint32_t PRED_PREV_SWEEP = 0x80000000u;

// computeConstraintData (BodyCollisionSolver.cpp:229-390), written at the solve position.
__device__ inline uint32_t predecessorOf(uint32_t body, uint32_t r, const uint32_t* degree, const uint32_t* incStart,
                                         const uint32_t* incSorted, const uint32_t* posOf) {
    const uint32_t s = incStart[body];
    if (r > 0) return posOf[incSorted[s + r - 1] >> 1];
    return posOf[incSorted[s + degree[body] - 1] >> 1] | PRED_PREV_SWEEP;
}

// PRED: the predecessor slots the dataflow solve waits on (the per-scene solve needs none).
template <bool PRED>
__global__ __launch_bounds__(256) void kPrep(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                             const uint2* __restrict__ manBodies, const uint32_t* __restrict__ posOf,
                                             const uint32_t* __restrict__ degree, const uint32_t* __restrict__ rank,
                                             const uint32_t* __restrict__ incStart, const uint32_t* __restrict__ incSorted,
                                             const float2* __restrict__ inv, const float4* __restrict__ pos,
                                             const float4* __restrict__ vel, const uint4* __restrict__ shape,
                                             const float4* __restrict__ materials, Constraints cs) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint32_t slot = posOf[m];
        const agx_manifold d = man[m];
        const uint2 mb = manBodies[m];
        const uint32_t bA = mb.x, bB = mb.y;
        const float4 pA = pos[bA], pB = pos[bB];
        const float2 iA = inv[bA], iB = inv[bB];
        const float nx = d.normal_x, ny = d.normal_y;
        const float penBase = d.depth + ((pB.x - pA.x) * nx + (pB.y - pA.y) * ny) - 0.001f;
        const float4 mA = materials[shape[d.a].w], mB = materials[shape[d.b].w];
        const float e = (mA.x + mB.x) * 0.5f;
        const float sf = sqrtf(fmaxf(0.0f, mA.y * mB.y));
        const float df = sqrtf(fmaxf(0.0f, mA.z * mB.z));
        const float effMass = iA.x + iB.x;
        const float tx = -ny, ty = nx;
        const float4 vA = vel[bA], vB = vel[bB];
        float pr[2][7] = { { 0, 0, 0, 0, 0, 0, 0 }, { 0, 0, 0, 0, 0, 0, 0 } };
        for (uint32_t i = 0; i < d.count; i++) {
            const float cx = i == 0 ? d.p0_x : d.p1_x, cy = i == 0 ? d.p0_y : d.p1_y;
            const float rAx = cx - pA.x, rAy = cy - pA.y, rBx = cx - pB.x, rBy = cy - pB.y;
            const float rApx = -rAy, rApy = rAx, rBpx = -rBy, rBpy = rBx;
            const float rAn = rApx * nx + rApy * ny, rBn = rBpx * nx + rBpy * ny;
            const float rAt = rApx * tx + rApy * ty, rBt = rBpx * tx + rBpy * ty;
            const float nd = effMass + rAn * rAn * iA.y + rBn * rBn * iB.y;
            const float td = effMass + rAt * rAt * iA.y + rBt * rBt * iB.y;
            const float negVn = -(((vB.x - vA.x) * nx + (vB.y - vA.y) * ny) + rBn * vB.z - rAn * vA.z);
            pr[i][0] = rAn; pr[i][1] = rBn; pr[i][2] = rAt; pr[i][3] = rBt;
            pr[i][4] = nd > 0.0f ? 1.0f / nd : 0.0f;
            pr[i][5] = td > 0.0f ? 1.0f / td : 0.0f;
            pr[i][6] = negVn > 0.0f ? e * negVn : 0.0f;
        }
        const bool dA = iA.x != 0.0f, dB = iB.x != 0.0f;
        cs.idx[slot] = make_uint4(bA, bB, d.count, m);
        cs.dep[slot] = make_uint4(PRED && dA ? predecessorOf(bA, rank[2 * m], degree, incStart, incSorted, posOf) : NO_PRED,
                                  PRED && dB ? predecessorOf(bB, rank[2 * m + 1], degree, incStart, incSorted, posOf) : NO_PRED,
                                  dA ? 1u : 0u, dB ? 1u : 0u);
        cs.nrm[slot] = make_float4(nx, ny, sf, df);
        cs.inv[slot] = make_float4(iA.x, iB.x, iA.y, iB.y);
        cs.r0[slot] = make_float4(pr[0][0], pr[0][1], pr[0][2], pr[0][3]);
        cs.r1[slot] = make_float4(pr[1][0], pr[1][1], pr[1][2], pr[1][3]);
        cs.k0[slot] = make_float4(pr[0][4], pr[0][5], pr[0][6], pr[1][4]);
        cs.k1[slot] = make_float4(pr[1][5], pr[1][6], penBase, 0.0f);
        cs.imp[slot] = make_float4(d.jn0, d.jt0, d.count > 1 ? d.jn1 : 0.0f, d.count > 1 ? d.jt1 : 0.0f);
        float4 blk = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (d.count > 1) {
            // the operations of the block solve in velocityStep, on the same values: the same bits
            const float K00 = effMass + pr[0][0] * pr[0][0] * iA.y + pr[0][1] * pr[0][1] * iB.y;
            const float K11 = effMass + pr[1][0] * pr[1][0] * iA.y + pr[1][1] * pr[1][1] * iB.y;
            const float K01 = effMass + pr[0][0] * pr[1][0] * iA.y + pr[0][1] * pr[1][1] * iB.y;
            blk = make_float4(K00, K11, K01, blockInvDet(K00, K11, K01));
        }
        cs.blk[slot] = blk;
    }
}

// applyWarmStarting (BodyCollisionSolver.cpp:416-490). A warm-start impulse depends on the
// constraint alone, so each body adds its constraints' impulses in solve order: the same
// operations, in the same order per body, as a sequential pass.
__global__ __launch_bounds__(256) void kWarm(const uint32_t* __restrict__ dynBodies, uint32_t nDyn,
                                             const uint32_t* __restrict__ degree, const uint32_t* __restrict__ incStart,
                                             const uint32_t* __restrict__ incSorted, const uint32_t* __restrict__ posOf,
                                             Constraints cs, float4* __restrict__ vel) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nDyn; k += gridDim.x * blockDim.x) {
        const uint32_t b = dynBodies[k], s = incStart[b], d = degree[b];
        if (d == 0) continue;
        float4 v = vel[b];
        for (uint32_t r = 0; r < d; r++) {
            const uint32_t e = incSorted[s + r];
            const uint32_t c = posOf[e >> 1];
            const float4 nr = cs.nrm[c], iv = cs.inv[c], r0 = cs.r0[c], r1 = cs.r1[c], j = cs.imp[c];
            const float nx = nr.x, ny = nr.y, tx = -ny, ty = nx;
            const float jn0 = j.x, jt0 = j.y, jn1 = j.z, jt1 = j.w;
            const float sn = jn0 + jn1, st = jt0 + jt1;
            const float sx = sn * nx + st * tx, sy = sn * ny + st * ty;
            if ((e & 1u) == 0) {
                v.x = v.x - sx * iv.x; v.y = v.y - sy * iv.x;
                v.z -= (jn0 * r0.x + jt0 * r0.z + jn1 * r1.x + jt1 * r1.z) * iv.z;
            } else {
                v.x = v.x + sx * iv.y; v.y = v.y + sy * iv.y;
                v.z += (jn0 * r0.y + jt0 * r0.w + jn1 * r1.y + jt1 * r1.w) * iv.w;
            }
        }
        vel[b] = v;
    }
}

// ---------------------------------------------------------------- dataflow sweeps
struct Item {
    uint4 id;     // bodyA, bodyB, count, manifold
    uint4 dep;    // predecessor at A, predecessor at B, A dynamic, B dynamic (dataflow)
    float4 nrm, iv, r0, r1, k0, k1, imp;
    float4 blk;   // the precomputed 2x2 block (per-scene solve)
};

__device__ inline void loadItem(const Constraints& cs, uint32_t c, Item& it) {
    it.id = cs.idx[c]; it.dep = cs.dep[c];
    it.nrm = cs.nrm[c]; it.iv = cs.inv[c]; it.r0 = cs.r0[c]; it.r1 = cs.r1[c];
    it.k0 = cs.k0[c]; it.k1 = cs.k1[c]; it.imp = cs.imp[c];
}

// The per-scene solve's record: the block instead of the predecessors (a body is dynamic iff its
// inverse mass is not zero, which is what the dynamic flags of dep hold), and no inverse masses
// (the solve keeps the bodies' own in LDS: the same values).
__device__ inline void loadSceneItem(const Constraints& cs, uint32_t c, Item& it) {
    it.id = cs.idx[c]; it.blk = cs.blk[c];
    it.nrm = cs.nrm[c]; it.r0 = cs.r0[c]; it.r1 = cs.r1[c];
    it.k0 = cs.k0[c]; it.k1 = cs.k1[c]; it.imp = cs.imp[c];
}

// solveVelocityConstraints (BodyCollisionSolver.cpp:492-928): one point = clamped normal step
// (skipped while separating with zero impulse) + iterative friction; two points = 2x2 block LCP
// (4 cases, iterative fallback) + iterative friction per point. Adds the squared accumulator
// changes and the squared accumulators to dsq and lsq. BLK: the block's K and 1/det come from
// it.blk (computed by kPrep with the same operations) instead of being recomputed.
template <bool BLK>
__device__ inline void velocityStep(Item& it, float4& A, float4& B, float& dsq, float& lsq) {
    float4 j = it.imp;
    const float4 j0 = j;
    const float4 nr = it.nrm, iv = it.iv, r0 = it.r0, k0 = it.k0;
    const float nx = nr.x, ny = nr.y, tx = -ny, ty = nx;
    const float sF = nr.z, dF = nr.w;
    const float iMA = iv.x, iMB = iv.y, iIA = iv.z, iIB = iv.w;
    auto friction = [&](float rAt, float rBt, float tMass, float& acc, float jn) {
        const float slip = ((B.x - A.x) * tx + (B.y - A.y) * ty) + rBt * B.z - rAt * A.z;
        const float jt = -slip * tMass;
        const float old = acc;
        const float target = old + jt;
        if (fabsf(target) <= jn * sF) acc = target;
        else { const float mx = jn * dF; acc = fminf(fmaxf(target, -mx), mx); }
        const float dj = acc - old;
        const float ix = dj * tx, iy = dj * ty;
        A.x = A.x - ix * iMA; A.y = A.y - iy * iMA; A.z -= dj * rAt * iIA;
        B.x = B.x + ix * iMB; B.y = B.y + iy * iMB; B.z += dj * rBt * iIB;
    };
    auto normal = [&](float rAn, float rBn, float bias, float nMass, float& acc) {
        const float vn = ((B.x - A.x) * nx + (B.y - A.y) * ny) + rBn * B.z - rAn * A.z;
        const float jn = (bias - vn) * nMass;
        const float old = acc;
        acc = fmaxf(0.0f, old + jn);
        const float dj = acc - old;
        const float ix = dj * nx, iy = dj * ny;
        A.x = A.x - ix * iMA; A.y = A.y - iy * iMA; A.z -= dj * rAn * iIA;
        B.x = B.x + ix * iMB; B.y = B.y + iy * iMB; B.z += dj * rBn * iIB;
    };
    if (it.id.z == 1) {
        float jnSingle = 0.0f;
        const float vn = ((B.x - A.x) * nx + (B.y - A.y) * ny) + r0.y * B.z - r0.x * A.z;
        if (!(vn > 0.0f && j.x == 0.0f)) {
            normal(r0.x, r0.y, k0.z, k0.x, j.x);
            jnSingle = j.x;
        }
        friction(r0.z, r0.w, k0.y, j.y, jnSingle);
    } else {
        const float4 r1 = it.r1, k1 = it.k1;
        const float effMass = iMA + iMB;
        const float dvn = (B.x - A.x) * nx + (B.y - A.y) * ny;
        const float vn0 = dvn + r0.y * B.z - r0.x * A.z;
        const float vn1 = dvn + r1.y * B.z - r1.x * A.z;
        const float K00 = BLK ? it.blk.x : effMass + r0.x * r0.x * iIA + r0.y * r0.y * iIB;
        const float K11 = BLK ? it.blk.y : effMass + r1.x * r1.x * iIA + r1.y * r1.y * iIB;
        const float K01 = BLK ? it.blk.z : effMass + r0.x * r1.x * iIA + r0.y * r1.y * iIB;
        const float b0 = k0.z - vn0 + K00 * j.x + K01 * j.z;
        const float b1 = k1.y - vn1 + K01 * j.x + K11 * j.z;
        float x0 = 0.0f, x1 = 0.0f;
        bool ok = false;
        {
            const float EPS = 1e-6f;
            const float invDet = BLK ? it.blk.w : blockInvDet(K00, K11, K01);
            if (invDet == invDet) {
                const float s0 = (K11 * b0 - K01 * b1) * invDet;
                const float s1 = (K00 * b1 - K01 * b0) * invDet;
                if (s0 >= -EPS && s1 >= -EPS) { x0 = fmaxf(s0, 0.0f); x1 = fmaxf(s1, 0.0f); ok = true; }
            }
            if (!ok && K11 > 0.0f) {
                const float s1 = b1 / K11;
                if (s1 >= -EPS && (K01 * s1 - b0) >= -EPS) { x0 = 0.0f; x1 = fmaxf(s1, 0.0f); ok = true; }
            }
            if (!ok && K00 > 0.0f) {
                const float s0 = b0 / K00;
                if (s0 >= -EPS && (K01 * s0 - b1) >= -EPS) { x0 = fmaxf(s0, 0.0f); x1 = 0.0f; ok = true; }
            }
            if (!ok && b0 <= EPS && b1 <= EPS) { x0 = 0.0f; x1 = 0.0f; ok = true; }
        }
        float jnArr0, jnArr1;
        if (!ok) {
            // iterative normal fallback, point 0 then point 1
            normal(r0.x, r0.y, k0.z, k0.x, j.x);
            jnArr0 = j.x;
            normal(r1.x, r1.y, k1.y, k0.w, j.z);
            jnArr1 = j.z;
        } else {
            const float d0 = x0 - j.x, d1 = x1 - j.z;
            j.x = x0; j.z = x1;
            jnArr0 = x0; jnArr1 = x1;
            const float jnSum = d0 + d1;
            const float sA = jnSum * iMA, sB = jnSum * iMB;
            A.x = A.x - nx * sA; A.y = A.y - ny * sA;
            B.x = B.x + nx * sB; B.y = B.y + ny * sB;
            A.z -= (d0 * r0.x + d1 * r1.x) * iIA;
            B.z += (d0 * r0.y + d1 * r1.y) * iIB;
        }
        friction(r0.z, r0.w, k0.y, j.y, jnArr0);
        friction(r1.z, r1.w, k1.x, j.w, jnArr1);
    }
    it.imp = j;
    const float e0 = j.x - j0.x, e1 = j.y - j0.y;
    dsq += e0 * e0 + e1 * e1;
    lsq += j.x * j.x + j.y * j.y;
    if (it.id.z > 1) {
        const float e2 = j.z - j0.z, e3 = j.w - j0.w;
        dsq += e2 * e2 + e3 * e3;
        lsq += j.z * j.z + j.w * j.w;
    }
}

struct PosItem {
    uint2 body;
    uint32_t count;  // contact points; 0 in a padding hole
    uint4 dep;       // predecessor at A, predecessor at B, A dynamic, B dynamic
    float nx, ny, iMA, iMB, penBase;
};

__device__ inline void loadPosItem(const Constraints& cs, uint32_t c, PosItem& pi) {
    const uint4 id = cs.idx[c];
    const float4 nr = cs.nrm[c], iv = cs.inv[c];
    pi.body = make_uint2(id.x, id.y);
    pi.count = id.z;
    pi.dep = cs.dep[c];
    pi.nx = nr.x; pi.ny = nr.y; pi.iMA = iv.x; pi.iMB = iv.y;
    pi.penBase = cs.k1[c].z;
}

// solvePositionConstraints (BodyCollisionSolver.cpp:930-980): linear only; nothing moves when
// the constraint is not penetrating.
__device__ inline void positionStep(const PosItem& pi, float2& A, float2& B) {
    const float pen = pi.penBase - ((B.x - A.x) * pi.nx + (B.y - A.y) * pi.ny);
    if (pen <= 0.0f) return;
    const float total = pen / (pi.iMA + pi.iMB) * 0.5f;
    const float cA = pi.iMA * total, cB = pi.iMB * total;
    A.x -= pi.nx * cA; A.y -= pi.ny * cA;
    B.x += pi.nx * cB; B.y += pi.ny * cB;
}

// Lanes run their constraints in groups of SOLVE_GROUP consecutive lanes (a divisor of the wave
// size and of COLOUR_ALIGN, so a group's constraints of one step share one colour).
constexpr uint32_t SOLVE_GROUP = 4;

// True when the predicate holds on every lane of this lane's group.
__device__ inline bool groupAll(bool pred) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t shift = threadIdx.x & (WAVE - SOLVE_GROUP);
    const unsigned long long bits = SOLVE_GROUP == 64 ? ~0ull : ((1ull << SOLVE_GROUP) - 1);
    return ((mask >> shift) & bits) == bits;
}

// Progress of a group in one flow: after its step at (sweep q, first position b) it publishes
// q * P + min(b + SOLVE_GROUP, P), so every position of the group up to that step reads as done,
// and no position of a later step does. A constraint's predecessor at (sweep q', position p) is
// done once its group's progress reaches q' * P + p + 1.
__device__ inline uint32_t predTarget(uint32_t pred, uint32_t q, uint32_t P, uint32_t T, uint32_t& wave) {
    wave = 0;
    if (pred == NO_PRED) return 0;
    const uint32_t p = pred & ~PRED_PREV_SWEEP;
    uint32_t qp = q;
    if (pred & PRED_PREV_SWEEP) {
        if (q == 0) return 0;
        qp = q - 1;
    }
    wave = (p % T) / SOLVE_GROUP;
    return qp * P + p + 1;
}

__device__ inline bool predecessorsDone(const uint32_t* done, uint32_t predA, uint32_t predB, uint32_t q, uint32_t P,
                                        uint32_t T) {
    uint32_t wa, wb;
    const uint32_t ta = predTarget(predA, q, P, T, wa), tb = predTarget(predB, q, P, T, wb);
    const uint32_t fa = ta ? loadFlag(done + wa) : 0u;
    const uint32_t fb = tb ? loadFlag(done + wb) : 0u;
    return fa >= ta && fb >= tb;
}

__device__ inline void storeCoherentF(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float loadCoherentF(const float* p) {
    return __uint_as_float(loadFlag(reinterpret_cast<const uint32_t*>(p)));
}

// Persistent dataflow over the sweeps. Velocity sweeps read and write only velocities (and the
// constraint's own accumulators), position sweeps only positions, so the two are independent
// flows that run side by side; each alone is the sequential order.
// Work goes by group steps: lane t holds solve positions t, t + T, ... (T = grid threads, a
// multiple of the wave size), and a group of lanes runs a step, all its lanes together, once the
// predecessors of every lane's constraint are done (a done predecessor stays done). Steps are
// taken in (sweep, position) order and each group publishes its progress per flow. Colours start
// on wave boundaries, so the constraints of one step share one sweep and one colour and depend
// only on steps of earlier (sweep, colour): the steps form an acyclic graph, the lowest open step
// is always ready, and with every block resident neither flow can deadlock. Running whole groups
// keeps the SIMDs busy with many lanes at a time; small groups keep a step from waiting on many
// unrelated predecessors. Body state of dynamic bodies goes through coherent loads and stores;
// static bodies are only read.
// With a residual target the velocity sweeps run one at a time behind a grid barrier and every
// block sums the same per-block partials in the same order, so all agree on when to stop.
__global__ __launch_bounds__(256) void kSolveFlow(Counters* __restrict__ ctr, Constraints cs, float4* vel, float4* pos,
                                                  uint32_t* vDone, uint32_t* pDone, float* __restrict__ blockRes,
                                                  uint32_t vIt, uint32_t pIt, int trackResidual, float residualTarget) {
    __shared__ float res[2 * MAX_SWEEPS];
    __shared__ double red[2][4];
    for (uint32_t i = threadIdx.x; i < 2 * MAX_SWEEPS; i += blockDim.x) res[i] = 0.0f;
    __syncthreads();
    const uint32_t P = ctr->colourStart[MAX_COLOURS];   // solve positions including holes
    const uint32_t T = gridDim.x * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t groupBase = tid & ~(SOLVE_GROUP - 1);
    const uint32_t groupId = tid / SOLVE_GROUP;
    const bool leader = (threadIdx.x & (SOLVE_GROUP - 1)) == 0;
    if ((uint64_t)P > (uint64_t)MAX_FLOW_ITEMS * T) {
        if (tid == 0) atomicOr(&ctr->overflow, OVERFLOW_ITEMS);
        return;
    }
    const bool track = trackResidual || residualTarget > 0.0f;
    const uint64_t deadline = wall_clock64() + STALL_TICKS;
    Item it;
    PosItem pi;
    uint32_t loaded = 0xFFFFFFFFu, pLoaded = 0xFFFFFFFFu;
    auto run = [&](uint32_t vBegin, uint32_t vEnd, uint32_t pBegin, uint32_t pEnd) -> bool {
        uint32_t qv = vBegin, kv = 0, qp = pBegin, kp = 0;   // group-uniform cursors
        bool vAct = vBegin < vEnd && groupBase < P, pAct = pBegin < pEnd && groupBase < P;
        bool vReady = false, pReady = false;
        while (true) {
            bool progressed = false;
            if (vAct) {
                const uint32_t c = tid + kv * T;
                bool has = c < P;
                if (has && c != loaded) { loadItem(cs, c, it); loaded = c; }
                has = has && it.id.z != 0;
                if (has && !vReady) vReady = predecessorsDone(vDone, it.dep.x, it.dep.y, qv, P, T);
                if (groupAll(!has || vReady)) {
                    float dsq = 0.0f, lsq = 0.0f;
                    if (has) {
                        float4 A = it.dep.z ? loadCoherent(vel + it.id.x) : vel[it.id.x];
                        float4 B = it.dep.w ? loadCoherent(vel + it.id.y) : vel[it.id.y];
                        velocityStep<false>(it, A, B, dsq, lsq);
                        cs.imp[c] = it.imp;
                        if (it.dep.z) storeCoherent(vel + it.id.x, A);
                        if (it.dep.w) storeCoherent(vel + it.id.y, B);
                    }
                    orderStores();
                    if (track) {
                        for (uint32_t off = SOLVE_GROUP / 2; off > 0; off >>= 1) {
                            dsq += __shfl_xor(dsq, off, SOLVE_GROUP);
                            lsq += __shfl_xor(lsq, off, SOLVE_GROUP);
                        }
                        if (leader) { atomicAdd(&res[qv], dsq); atomicAdd(&res[MAX_SWEEPS + qv], lsq); }
                    }
                    if (leader) storeFlag(vDone + groupId, qv * P + min(groupBase + kv * T + SOLVE_GROUP, P));
                    vReady = false;
                    kv++;
                    if ((uint64_t)groupBase + (uint64_t)kv * T >= P) { kv = 0; qv++; vAct = qv < vEnd; }
                    progressed = true;
                }
            }
            if (pAct) {
                const uint32_t c = tid + kp * T;
                bool has = c < P;
                if (has && c != pLoaded) { loadPosItem(cs, c, pi); pLoaded = c; }
                has = has && pi.count != 0;
                if (has && !pReady) pReady = predecessorsDone(pDone, pi.dep.x, pi.dep.y, qp, P, T);
                if (groupAll(!has || pReady)) {
                    if (has) {
                        float2* pA = reinterpret_cast<float2*>(pos + pi.body.x);
                        float2* pB = reinterpret_cast<float2*>(pos + pi.body.y);
                        float2 A = pi.dep.z ? loadCoherent2(pA) : *pA;
                        float2 B = pi.dep.w ? loadCoherent2(pB) : *pB;
                        positionStep(pi, A, B);
                        if (pi.dep.z) storeCoherent2(pA, A);
                        if (pi.dep.w) storeCoherent2(pB, B);
                    }
                    orderStores();
                    if (leader) storeFlag(pDone + groupId, qp * P + min(groupBase + kp * T + SOLVE_GROUP, P));
                    pReady = false;
                    kp++;
                    if ((uint64_t)groupBase + (uint64_t)kp * T >= P) { kp = 0; qp++; pAct = qp < pEnd; }
                    progressed = true;
                }
            }
            if (__all(!vAct && !pAct)) break;
            if (!__any(progressed)) {
                if (wall_clock64() > deadline) { atomicOr(&ctr->overflow, OVERFLOW_STALL); return false; }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        return true;
    };
    float* myRes = blockRes + (size_t)blockIdx.x * 2 * MAX_SWEEPS;
    uint32_t sweepsDone = vIt;
    bool met = false;
    if (residualTarget > 0.0f) {
        uint32_t gen = 0;
        for (uint32_t s = 0; s < vIt; s++) {
            const bool good = run(s, s + 1, 0, 0);
            if (__syncthreads_or(!good)) break;
            if (threadIdx.x == 0) { storeCoherentF(myRes + s, res[s]); storeCoherentF(myRes + MAX_SWEEPS + s, res[MAX_SWEEPS + s]); }
            if (!gridBarrier(&ctr->barrier[1], gridDim.x, gen, deadline)) {
                if (threadIdx.x == 0) atomicOr(&ctr->overflow, OVERFLOW_STALL);
                break;
            }
            double num = 0.0, den = 0.0;
            for (uint32_t j = threadIdx.x; j < gridDim.x; j += blockDim.x) {
                num += (double)loadCoherentF(blockRes + (size_t)j * 2 * MAX_SWEEPS + s);
                den += (double)loadCoherentF(blockRes + (size_t)j * 2 * MAX_SWEEPS + MAX_SWEEPS + s);
            }
            for (uint32_t off = 32; off > 0; off >>= 1) { num += __shfl_down(num, off, 64); den += __shfl_down(den, off, 64); }
            if ((threadIdx.x & 63u) == 0) { red[0][threadIdx.x >> 6] = num; red[1][threadIdx.x >> 6] = den; }
            __syncthreads();
            num = red[0][0] + red[0][1] + red[0][2] + red[0][3];
            den = red[1][0] + red[1][1] + red[1][2] + red[1][3];
            __syncthreads();
            if (sqrt(num) <= (double)residualTarget * fmax(sqrt(den), 1e-12)) {
                sweepsDone = s + 1;
                met = true;
                break;
            }
        }
        run(0, 0, 0, pIt);
    } else {
        run(0, vIt, 0, pIt);
    }
    __syncthreads();
    if (track)
        for (uint32_t q = threadIdx.x; q < sweepsDone; q += blockDim.x) { myRes[q] = res[q]; myRes[MAX_SWEEPS + q] = res[MAX_SWEEPS + q]; }
    if (tid == 0) {
        ctr->sweeps = sweepsDone;
        ctr->flowMet = met ? 1u : 0u;
        ctr->solveSteps += 1;
        ctr->sweepsTotal += sweepsDone;
    }
}

// ---------------------------------------------------------------- per-scene solve
// Scenes of a batch share no body, so each can be solved by one workgroup on its own, with the
// scene's body records in LDS: the same sequential colour-major order as the dataflow (within a
// colour the constraints share no dynamic body, so the block's lanes run them at once, and a block
// barrier separates colours), hence the same bits, without device-wide flags or barriers. With a
// residual target each scene stops at its own sweep, as a simulation of that scene alone would.

// range[scene * MAX_COLOURS + colour] = (first slot, one past the last slot) of the scene's
// constraints in that colour. The colour scatter keeps manifold order and manifolds come grouped
// by scene, so a range holds few slots of other scenes (the solve skips them by body index).
__global__ __launch_bounds__(256) void kSceneRanges(const Counters* __restrict__ ctr, Constraints cs,
                                                    const uint32_t* __restrict__ bodyScene, uint2* __restrict__ range) {
    __shared__ uint32_t starts[MAX_COLOURS + 1];
    for (uint32_t i = threadIdx.x; i <= MAX_COLOURS; i += blockDim.x) starts[i] = ctr->colourStart[i];
    __syncthreads();
    const uint32_t P = starts[MAX_COLOURS];
    const uint32_t K = ctr->colours;
    auto keyAt = [&](uint32_t c) -> uint32_t {   // scene x colour of the constraint at slot c
        if (c >= P) return NO_KEY;
        const uint4 id = cs.idx[c];
        if (id.z == 0) return NO_KEY;
        uint32_t lo = 0, hi = K;   // starts[lo] <= c < starts[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (starts[mid] <= c) lo = mid; else hi = mid;
        }
        return bodyScene[id.x] * MAX_COLOURS + lo;
    };
    for (uint32_t base = blockIdx.x * blockDim.x; base < P; base += gridDim.x * blockDim.x)
        runRanges(base + threadIdx.x, keyAt, range);
}

constexpr uint32_t SCENE_THREADS = 1024;   // one block per CU

// Sum of (num, den) over the group's lanes: a butterfly within each wave (every lane ends with
// the same bits), then the waves' partials in wave order; every lane returns the total.
__device__ inline double2 sceneSum(SceneLanes& L, double num, double den, double2* part) {
    for (uint32_t off = WAVE / 2; off > 0; off >>= 1) {
        num += __shfl_xor(num, off, WAVE);
        den += __shfl_xor(den, off, WAVE);
    }
    const uint32_t w = L.t / WAVE;
    if ((L.t & (WAVE - 1)) == 0) part[w] = make_double2(num, den);
    L.sync();
    double2 tot = part[0];
    for (uint32_t k = 1; k < L.waves; k++) { tot.x += part[k].x; tot.y += part[k].y; }
    L.sync();   // before the partials are written again
    return tot;
}

// Colour-major rounds of the lanes L over a scene's constraint ranges rng[0..K): lane t takes slot
// base + t of the rounds base = rng[k].x, + T, ... below rng[k].y, and the lanes sync after each
// colour. The next round's record is loaded (load(record, slot)) before the current one is run
// (run(record, slot)), also across the sync that ends a colour, into the other of two records, so
// nothing is copied. This needs that only the lane that runs a slot writes to its constraint, and
// that nothing a record holds changes while the rounds run.
template <class Rec, class Load, class Run>
__device__ inline void colourRounds(SceneLanes& L, const uint2* rng, uint32_t K, Load load, Run run) {
    const uint32_t t = L.t, T = L.T;
    uint32_t nk = 0, nb = 0;   // the next round: colour and base (uniform)
    auto firstRoundFrom = [&](uint32_t k) {
        nk = k;
        while (nk < K && uniform(rng[nk].x) >= uniform(rng[nk].y)) nk++;
        if (nk < K) nb = uniform(rng[nk].x);
    };
    Rec I0, I1;
    firstRoundFrom(0);
    if (nk < K && nb + t < uniform(rng[nk].y)) load(I0, nb + t);
    bool odd = false;   // the current round's record is I1
    for (uint32_t k = 0; k < K; k++) {
        const uint32_t end = uniform(rng[k].y);
        for (uint32_t base = uniform(rng[k].x); base < end; base += T) {
            if (base + T < end) { nk = k; nb = base + T; } else firstRoundFrom(k + 1);
            const bool loadNext = nk < K && nb + t < uniform(rng[nk].y);
            const bool mine = base + t < end;
            if (odd) {
                if (loadNext) load(I0, nb + t);
                if (mine) run(I1, base + t);
            } else {
                if (loadNext) load(I1, nb + t);
                if (mine) run(I0, base + t);
            }
            odd = !odd;
        }
        L.sync();
    }
}

// What the warm start and the position sweeps of the per-scene solve read of a constraint.
struct WarmRec {
    uint4 id;
    float4 nrm, r0, r1, imp;
};
struct PosRec {
    uint4 id;
    float2 n;
    float penBase;
};

// One scene (bodies [b0, b1)) by the lanes L, body records in `body`: warm start, velocity sweeps
// until the residual target or vIt, position sweeps; returns the velocity sweeps run.
__device__ inline uint32_t solveScene(SceneLanes& L, uint32_t s, Counters* __restrict__ ctr, const Constraints& cs,
                                      float4* __restrict__ vel, float4* __restrict__ pos, uint32_t b0, uint32_t b1,
                                      const uint2* __restrict__ range, float2* __restrict__ sceneRes, uint32_t K,
                                      uint32_t vIt, uint32_t pIt, int track, float residualTarget, float4* body,
                                      float2* binv, const float2* __restrict__ inv, uint2* rng, double2* part,
                                      bool& met) {
    const uint32_t t = L.t, T = L.T;
    met = false;
    for (uint32_t k = t; k < K; k += T) rng[k] = range[(size_t)s * MAX_COLOURS + k];
    for (uint32_t b = b0 + t; b < b1; b += T) {
        body[b - b0] = vel[b];
        binv[b - b0] = inv[b];   // inverse mass, inverse inertia: what the constraints' inv plane holds
    }
    L.sync();
    // the constraint's inverse masses and inertias (A, B, A, B), as kPrep writes them
    auto invOf = [&](const uint4& id) {
        const float2 a = binv[id.x - b0], b = binv[id.y - b0];
        return make_float4(a.x, b.x, a.y, b.y);
    };
    // a constraint of this scene (not a hole, nor one of a neighbouring scene's in the range)
    auto ours = [&](const uint4& id) { return id.z != 0 && id.x >= b0 && id.x < b1 && id.y >= b0 && id.y < b1; };
    // applyWarmStarting (BodyCollisionSolver.cpp:416-490) colour by colour: every body adds its
    // constraints' impulses in solve order, with the operations of kWarm
    colourRounds<WarmRec>(
        L, rng, K,
        [&](WarmRec& w, uint32_t c) {
            w.id = cs.idx[c]; w.nrm = cs.nrm[c]; w.r0 = cs.r0[c]; w.r1 = cs.r1[c]; w.imp = cs.imp[c];
        },
        [&](WarmRec& w, uint32_t) {
            if (!ours(w.id)) return;
            const float4 nr = w.nrm, iv = invOf(w.id), r0 = w.r0, r1 = w.r1, j = w.imp;
            const float nx = nr.x, ny = nr.y, tx = -ny, ty = nx;
            const float jn0 = j.x, jt0 = j.y, jn1 = j.z, jt1 = j.w;
            const float sn = jn0 + jn1, st = jt0 + jt1;
            const float sx = sn * nx + st * tx, sy = sn * ny + st * ty;
            if (iv.x != 0.0f) {
                float4 v = body[w.id.x - b0];
                v.x = v.x - sx * iv.x; v.y = v.y - sy * iv.x;
                v.z -= (jn0 * r0.x + jt0 * r0.z + jn1 * r1.x + jt1 * r1.z) * iv.z;
                body[w.id.x - b0] = v;
            }
            if (iv.y != 0.0f) {
                float4 v = body[w.id.y - b0];
                v.x = v.x + sx * iv.y; v.y = v.y + sy * iv.y;
                v.z += (jn0 * r0.y + jt0 * r0.w + jn1 * r1.y + jt1 * r1.w) * iv.w;
                body[w.id.y - b0] = v;
            }
        });
    uint32_t done = vIt;
    for (uint32_t q = 0; q < vIt; q++) {
        float dsq = 0.0f, lsq = 0.0f;
        // only the lane that solves a constraint writes its accumulators (the same lane every sweep)
        colourRounds<Item>(
            L, rng, K, [&](Item& it, uint32_t c) { loadSceneItem(cs, c, it); },
            [&](Item& it, uint32_t c) {
                if (it.id.z == 0 || it.id.x < b0 || it.id.x >= b1) return;   // a hole or another scene
                if (it.id.y < b0 || it.id.y >= b1) {   // both bodies of a contact share a scene
                    atomicOr(&ctr->overflow, OVERFLOW_ITEMS);
                    return;
                }
                float4 A = body[it.id.x - b0], B = body[it.id.y - b0];
                it.iv = invOf(it.id);
                velocityStep<true>(it, A, B, dsq, lsq);
                cs.imp[c] = it.imp;
                if (it.iv.x != 0.0f) body[it.id.x - b0] = A;
                if (it.iv.y != 0.0f) body[it.id.y - b0] = B;
            });
        const double2 r = sceneSum(L, (double)dsq, (double)lsq, part);
        if (t == 0 && track) sceneRes[(size_t)s * MAX_SWEEPS + q] = make_float2((float)r.x, (float)r.y);
        if (residualTarget > 0.0f && sqrt(r.x) <= (double)residualTarget * fmax(sqrt(r.y), 1e-12)) {
            done = q + 1;
            met = true;
            break;
        }
    }
    for (uint32_t b = b0 + t; b < b1; b += T) vel[b] = body[b - b0];
    L.sync();
    if (pIt) {
        float2* bp = reinterpret_cast<float2*>(body);
        for (uint32_t b = b0 + t; b < b1; b += T) {
            const float4 x = pos[b];
            bp[b - b0] = make_float2(x.x, x.y);
        }
        L.sync();
        // solvePositionConstraints (BodyCollisionSolver.cpp:930-980), the operations of positionStep
        for (uint32_t q = 0; q < pIt; q++)
            colourRounds<PosRec>(
                L, rng, K,
                [&](PosRec& pr, uint32_t c) {
                    pr.id = cs.idx[c];
                    const float4 nr = cs.nrm[c];
                    pr.n = make_float2(nr.x, nr.y);
                    pr.penBase = cs.k1[c].z;
                },
                [&](PosRec& pr, uint32_t) {
                    if (!ours(pr.id)) return;
                    PosItem pi;
                    pi.nx = pr.n.x; pi.ny = pr.n.y; pi.penBase = pr.penBase;
                    pi.iMA = binv[pr.id.x - b0].x; pi.iMB = binv[pr.id.y - b0].x;
                    float2 A = bp[pr.id.x - b0], B = bp[pr.id.y - b0];
                    positionStep(pi, A, B);
                    if (pi.iMA != 0.0f) bp[pr.id.x - b0] = A;
                    if (pi.iMB != 0.0f) bp[pr.id.y - b0] = B;
                });
        for (uint32_t b = b0 + t; b < b1; b += T) reinterpret_cast<float2*>(pos + b)[0] = bp[b - b0];
    }
    L.sync();   // before the next scene reuses the LDS
    return done;
}

// Scenes are taken in the step's order (kSceneWorkOrder): the first ctr->sceneWhole by whole
// blocks, the rest (none larger than groupBodies bodies) by the blocks' lane groups (groups per
// block, each with room for groupBodies bodies) from a second cursor. LDS per body: its velocity
// (then position) record and its inverse mass and inertia. A scene's results do not depend on
// which block or group solves it.
__global__ __launch_bounds__(SCENE_THREADS) void kSceneSolve(Counters* __restrict__ ctr, Constraints cs, float4* __restrict__ vel,
                                                             float4* __restrict__ pos, const float2* __restrict__ inv,
                                                             const uint32_t* __restrict__ sceneBodyStart,
                                                             const uint32_t* __restrict__ sceneOrder,
                                                             const uint32_t* __restrict__ sceneDyn, const uint2* __restrict__ range,
                                                             float2* __restrict__ sceneRes, uint32_t* __restrict__ sceneSweeps,
                                                             uint32_t S, uint32_t maxBodies, uint32_t groups, uint32_t groupBodies,
                                                             uint32_t vIt, uint32_t pIt, int track, float residualTarget) {
    extern __shared__ float4 body[];   // the scene's body velocities, then positions (x, y); per group
    __shared__ uint2 rng[SCENE_GROUPS][MAX_COLOURS];
    __shared__ double2 part[SCENE_THREADS / WAVE];
    __shared__ uint32_t taken[SCENE_GROUPS];
    __shared__ uint32_t bar[SCENE_GROUPS];
    __shared__ uint32_t abort;
    const uint32_t K = ctr->colours;
    const uint32_t nWhole = ctr->sceneWhole;
    if (threadIdx.x < SCENE_GROUPS) bar[threadIdx.x] = 0u;
    if (threadIdx.x == 0) abort = 0u;
    SceneLanes L{ threadIdx.x, SCENE_THREADS, 0u, SCENE_THREADS / WAVE, true, nullptr, &abort, 0u, ctr };
    float4* lds = body;
    uint32_t cap = maxBodies;   // bodies of the current LDS region
    bool grouped = false;
    while (true) {   // one call site of solveScene for both phases
        if (L.t == 0) taken[L.g] = grouped ? nWhole + atomicAdd(&ctr->sceneCursor2, 1u) : atomicAdd(&ctr->sceneCursor, 1u);
        L.sync();
        const uint32_t k0 = taken[L.g];
        if (!grouped && k0 >= nWhole) {
            __syncthreads();   // every wave has read taken[0] before the groups reuse it
            if (nWhole >= S) break;
            const uint32_t T = SCENE_THREADS / groups;
            L.g = threadIdx.x / T;
            L.t = threadIdx.x % T;
            L.T = T;
            L.waves = T / WAVE;
            L.whole = groups == 1;
            L.bar = &bar[L.g];
            lds = reinterpret_cast<float4*>(reinterpret_cast<char*>(body) + (size_t)L.g * groupBodies * 24u);
            cap = groupBodies;
            grouped = true;
            continue;
        }
        if (k0 >= S) break;
        const uint32_t s = sceneOrder[k0];
        if (sceneBodyStart[s + 1] - sceneBodyStart[s] > cap) {   // larger than this LDS region: not solved
            if (L.t == 0) {
                sceneSweeps[s] = 0u;
                atomicOr(&ctr->overflow, OVERFLOW_LDS);
            }
            L.sync();   // every wave has read taken[] before the next claim rewrites it
            continue;
        }
        bool met;
        const uint32_t done = solveScene(L, s, ctr, cs, vel, pos, sceneBodyStart[s], sceneBodyStart[s + 1], range, sceneRes, K,
                                         vIt, pIt, track, residualTarget, lds, reinterpret_cast<float2*>(lds + cap), inv,
                                         rng[L.g], part + L.g * L.waves, met);
        if (L.t == 0) {
            sceneSweeps[s] = done;
            atomicMax(&ctr->sweeps, done);
            atomicAdd(&ctr->bodySweeps, (unsigned long long)sceneDyn[s] * done);
            if (met) {
                atomicAdd(&ctr->convergedBodies, (unsigned long long)sceneDyn[s]);
                atomicAdd(&ctr->convergedScenes, 1u);
            }
        }
    }
}

// This step's order of the per-scene solve and how many of its first scenes whole blocks take:
// the scenes a lane group cannot hold (more than groupBodies bodies), then by expected work, most
// first (dynamic bodies x the velocity sweeps the scene ran in the previous step; ties in size
// order); at least minWhole scenes go to whole blocks, so the heaviest get a CU each instead of a
// quarter of one. With more than WORK_SORT_MAX scenes: size order and minWhole whole scenes.
__global__ __launch_bounds__(1024) void kSceneWorkOrder(Counters* __restrict__ ctr, const uint32_t* __restrict__ sceneOrder,
                                                        const uint32_t* __restrict__ sceneBodyStart,
                                                        const uint32_t* __restrict__ sceneSweeps,
                                                        const uint32_t* __restrict__ sceneDyn, uint32_t S, uint32_t minWhole,
                                                        uint32_t groupBodies, uint32_t* __restrict__ out) {
    __shared__ unsigned long long key[WORK_SORT_MAX];
    __shared__ uint32_t must;
    const uint32_t t = threadIdx.x;
    if (S > WORK_SORT_MAX) {
        for (uint32_t i = t; i < S; i += blockDim.x) out[i] = sceneOrder[i];
        if (t == 0) ctr->sceneWhole = minWhole;
        return;
    }
    if (t == 0) must = 0u;
    uint32_t n2 = 1;
    while (n2 < S) n2 <<= 1;
    __syncthreads();
    for (uint32_t i = t; i < n2; i += blockDim.x) {
        unsigned long long k = 0ull;   // padding sorts last
        if (i < S) {
            const uint32_t s = sceneOrder[i];
            const bool big = sceneBodyStart[s + 1] - sceneBodyStart[s] > groupBodies;
            if (big) atomicAdd(&must, 1u);
            const unsigned long long work = min((unsigned long long)max(sceneDyn[s], 1u) * max(sceneSweeps[s], 1u),
                                                0x7FFFFFFFull);
            k = ((unsigned long long)big << 63) | (work << 32) | (0xFFFFFFFFu - i);
        }
        key[i] = k;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= n2; k <<= 1)   // bitonic sort, descending
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = t; i < n2; i += blockDim.x) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const unsigned long long a = key[i], b = key[p];
                    if (((i & k) == 0) ? a < b : a > b) { key[i] = b; key[p] = a; }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = t; i < S; i += blockDim.x) out[i] = sceneOrder[0xFFFFFFFFu - (uint32_t)key[i]];
    if (t == 0) ctr->sceneWhole = max(must, minWhole);
}

// Step totals of the per-scene solve.
__global__ void kSceneTotals(Counters* ctr) {
    ctr->perScene = 1u;
    ctr->solveSteps += 1;
    ctr->sweepsTotal += ctr->sweeps;
    ctr->bodySweepsTotal += ctr->bodySweeps;
    ctr->convergedBodiesTotal += ctr->convergedBodies;
}

// Step totals of the dataflow solve (all dynamic bodies ran the same sweeps and share one
// residual: every scene met the target or none did).
__global__ void kFlowTotals(Counters* ctr, uint32_t nDynBodies, uint32_t S) {
    ctr->bodySweeps = (unsigned long long)nDynBodies * ctr->sweeps;
    ctr->bodySweepsTotal += ctr->bodySweeps;
    if (ctr->flowMet) {
        ctr->convergedBodies = nDynBodies;
        ctr->convergedScenes = S;
        ctr->convergedBodiesTotal += nDynBodies;
    }
}

// Accumulators back to the manifolds (the warm-start source of the next step) and
// computeConstraintErrors (BodyCollisionSolver.cpp:132-183) as per-block partial sums.
__global__ __launch_bounds__(256) void kSolveFinish(const Counters* __restrict__ ctr, Constraints cs,
                                                    const float4* __restrict__ vel, const float4* __restrict__ pos,
                                                    agx_manifold* __restrict__ man, float* __restrict__ errPartial,
                                                    int trackErrors) {
    const uint32_t P = ctr->colourStart[MAX_COLOURS];
    float app = 0.0f, pen = 0.0f;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < P; c += gridDim.x * blockDim.x) {
        const uint4 id = cs.idx[c];
        if (id.z == 0) continue;
        const float4 j = cs.imp[c];
        agx_manifold& m = man[id.w];
        m.jn0 = j.x; m.jt0 = j.y;
        if (id.z > 1) { m.jn1 = j.z; m.jt1 = j.w; }
        if (trackErrors) {
            const float4 nr = cs.nrm[c];
            const float4 A = pos[id.x], B = pos[id.y];
            pen += fmaxf(cs.k1[c].z - ((B.x - A.x) * nr.x + (B.y - A.y) * nr.y), 0.0f);
            const float4 vA = vel[id.x], vB = vel[id.y];
            const float4 r0 = cs.r0[c], r1 = cs.r1[c];
            const float dv = (vB.x - vA.x) * nr.x + (vB.y - vA.y) * nr.y;
            const float a0 = -(dv + r0.y * vB.z - r0.x * vA.z);
            app += a0 * (a0 > 0.0f ? 1.0f : 0.0f);
            if (id.z > 1) { const float a1 = -(dv + r1.y * vB.z - r1.x * vA.z); app += a1 * (a1 > 0.0f ? 1.0f : 0.0f); }
        }
    }
    if (!trackErrors) return;
    __shared__ float sa[4], sp[4];
    for (uint32_t off = 32; off > 0; off >>= 1) { app += __shfl_down(app, off, 64); pen += __shfl_down(pen, off, 64); }
    if ((threadIdx.x & 63u) == 0) { sa[threadIdx.x >> 6] = app; sp[threadIdx.x >> 6] = pen; }
    __syncthreads();
    if (threadIdx.x == 0) {
        errPartial[2 * blockIdx.x] = sa[0] + sa[1] + sa[2] + sa[3];
        errPartial[2 * blockIdx.x + 1] = sp[0] + sp[1] + sp[2] + sp[3];
    }
}

// One block: residual per sweep and error totals from the per-block partials (fixed order).
__device__ inline double blockSum(double v, double* scratch) {
    for (uint32_t off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
    __syncthreads();
    if ((threadIdx.x & (WAVE - 1)) == 0) scratch[threadIdx.x / WAVE] = v;
    __syncthreads();
    double t = 0.0;
    for (uint32_t k = 0; k < blockDim.x / WAVE; k++) t += scratch[k];
    return t;
}

// Residual after each velocity sweep over all scenes still running at it (one block per sweep;
// scenes summed in index order).
__global__ __launch_bounds__(256) void kSceneReduce(Counters* __restrict__ ctr, const float2* __restrict__ sceneRes,
                                                    const uint32_t* __restrict__ sceneSweeps, uint32_t S) {
    __shared__ double scratch[4];
    const uint32_t q = blockIdx.x;
    if (q >= ctr->sweeps) return;
    double num = 0.0, den = 0.0;
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x)
        if (sceneSweeps[s] > q) {
            const float2 v = sceneRes[(size_t)s * MAX_SWEEPS + q];
            num += (double)v.x;
            den += (double)v.y;
        }
    num = blockSum(num, scratch);
    den = blockSum(den, scratch);
    if (threadIdx.x == 0) { ctr->residualNum[q] = num; ctr->residualDen[q] = den; }
}

__global__ __launch_bounds__(256) void kSolveReduce(Counters* __restrict__ ctr, const float* __restrict__ blockRes,
                                                    uint32_t resBlocks, const float* __restrict__ errPartial,
                                                    uint32_t errBlocks, int track, int trackErrors) {
    __shared__ double scratch[4];
    const uint32_t sweeps = min(ctr->sweeps, MAX_SWEEPS);
    if (track) {
        for (uint32_t q = 0; q < sweeps; q++) {
            double num = 0.0, den = 0.0;
            for (uint32_t j = threadIdx.x; j < resBlocks; j += blockDim.x) {
                num += (double)blockRes[(size_t)j * 2 * MAX_SWEEPS + q];
                den += (double)blockRes[(size_t)j * 2 * MAX_SWEEPS + MAX_SWEEPS + q];
            }
            num = blockSum(num, scratch);
            den = blockSum(den, scratch);
            if (threadIdx.x == 0) { ctr->residualNum[q] = num; ctr->residualDen[q] = den; }
        }
    }
    if (trackErrors) {
        double a = 0.0, p = 0.0;
        for (uint32_t j = threadIdx.x; j < errBlocks; j += blockDim.x) { a += errPartial[2 * j]; p += errPartial[2 * j + 1]; }
        a = blockSum(a, scratch);
        p = blockSum(p, scratch);
        if (threadIdx.x == 0) { ctr->velocityError = (float)a; ctr->positionError = (float)p; }
    }
}

// ---------------------------------------------------------------- islands
// Union-find over contact edges between dynamic bodies where every link points from a larger
// body index to a smaller one, so each component's root is its smallest body index. Island ids
// number the roots in ascending order (IslandManager.cpp:36-80 numbers DSU roots in index order;
// the partition is the same).
//
// Parents are followed with plain loads. Every link points from a body to a smaller body of the
// same set, so a stale read can miss a recent link but never invents one: a walk always ends at a
// real ancestor, and a root it takes for current is checked by the compare-and-swap that hooks it,
// which fails and returns the newer parent when that root was already hooked.
//
// Path halving: a walk re-points each body it passes at its grandparent with a plain store. The
// grandparent is an ancestor, so the link still points downwards into the same set, and only
// bodies that already have a parent are re-pointed (roots change only by the compare-and-swap).
// Racing stores to one body each leave some ancestor there, so walks stay short on the long
// parent chains of large islands.
__device__ inline uint32_t ufFind(uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        const uint32_t g = parent[p];
        if (g == p) return p;
        parent[x] = g;
        x = g;
        p = parent[x];
    }
    return x;
}

__global__ __launch_bounds__(256) void kUfInit(uint32_t* parent, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) parent[i] = i;
}

// First links without any walk: each body points at its smallest-index dynamic neighbour below it.
__global__ __launch_bounds__(256) void kUfLinkNeighbours(const agx_manifold* __restrict__ man,
                                                         const Counters* __restrict__ ctr,
                                                         const uint32_t* __restrict__ colBody,
                                                         const float2* __restrict__ inv, uint32_t* parent) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
        if (!isDynamic(inv, a) || !isDynamic(inv, b)) continue;
        atomicMin(parent + max(a, b), min(a, b));
    }
}

// Every dynamic body re-pointed at the root of its tree after the first links (read-only walks,
// then one store each: links still point downwards into the same set), so the unions that follow
// start from trees of depth one.
__global__ __launch_bounds__(256) void kUfCompress(const uint32_t* __restrict__ list, uint32_t n, uint32_t* parent) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        uint32_t x = b, p = parent[x];
        while (p != x) { x = p; p = parent[x]; }
        if (x != b) parent[b] = x;
    }
}

// A merging pass without retries: each contact edge whose two trees differ hooks the larger root
// under the smaller with an atomic min. Two edges that hook one root keep only the smaller target,
// so some unions may be left out here; every link still joins bodies of one set, and the
// compare-and-swap pass that follows unites over every edge. Most of its edges then find one
// root at once, instead of contending for the same few roots of large islands.
__global__ __launch_bounds__(256) void kUfHook(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                               const uint32_t* __restrict__ colBody, const float2* __restrict__ inv,
                                               uint32_t* parent) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
        if (!isDynamic(inv, a) || !isDynamic(inv, b)) continue;
        uint32_t p = parent[a];
        while (p != a) { a = p; p = parent[a]; }
        p = parent[b];
        while (p != b) { b = p; p = parent[b]; }
        if (a != b) atomicMin(parent + max(a, b), min(a, b));
    }
}

__global__ __launch_bounds__(256) void kUfUnite(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                                const uint32_t* __restrict__ colBody, const float2* __restrict__ inv,
                                                uint32_t* parent) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
        if (!isDynamic(inv, a) || !isDynamic(inv, b)) continue;
        a = ufFind(parent, a);
        b = ufFind(parent, b);
        while (a != b) {
            if (a > b) { const uint32_t t = a; a = b; b = t; }
            const uint32_t old = atomicCAS(parent + b, b, a);
            if (old == b) break;
            b = ufFind(parent, old);
        }
    }
}

// Island edges of the joints between two dynamic bodies (IslandManager::step over each constraint
// system's active edges, Simulation.cpp:331-336); joints of sleeping bodies are not active.
__global__ __launch_bounds__(256) void kUfUniteEdges(const uint2* __restrict__ edges, uint32_t n,
                                                     const uint32_t* __restrict__ sleepId, uint32_t* parent) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t a = edges[k].x, b = edges[k].y;
        if (a == 0xFFFFFFFFu) continue;
        if (sleepId && (sleepId[a] != 0u || sleepId[b] != 0u)) continue;
        a = ufFind(parent, a);
        b = ufFind(parent, b);
        while (a != b) {
            if (a > b) { const uint32_t t = a; a = b; b = t; }
            const uint32_t old = atomicCAS(parent + b, b, a);
            if (old == b) break;
            b = ufFind(parent, old);
        }
    }
}

// Read-only root lookup: no path splitting, so concurrent lookups never rewrite parents and
// every body sees the final root (all unions finished in the previous kernel).
__device__ inline uint32_t ufRootOf(const uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) { x = p; p = parent[x]; }
    return x;
}

// root[b] (kept in islandId as scratch) and a root flag per dynamic body for the rank scan.
// Sleeping bodies join no island (IslandManager::reset covers the Dynamic range only); they touch
// no manifold of the step, so each is its own root and drops out of the ranks.
__global__ __launch_bounds__(256) void kUfRoots(const uint32_t* __restrict__ parent, const uint32_t* __restrict__ list,
                                                uint32_t n, uint32_t* __restrict__ rootFlag, uint32_t* __restrict__ rootOf,
                                                const uint32_t* __restrict__ sleepId) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        const uint32_t r = ufRootOf(parent, b);
        rootOf[b] = r;
        rootFlag[k] = (r == b && !(sleepId && sleepId[b] != 0u)) ? 1u : 0u;
    }
}

// Root rank -> island id for every dynamic body (the dynamic list is ascending, so the rank of a
// root in the list is its rank among roots); static bodies keep UINT32_MAX.
__global__ __launch_bounds__(256) void kIslandIds(const uint32_t* __restrict__ list, uint32_t n,
                                                  const uint32_t* __restrict__ rootScan,
                                                  const uint32_t* __restrict__ bodyToDyn, uint32_t* __restrict__ islandId,
                                                  const uint32_t* __restrict__ sleepId) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        islandId[b] = (sleepId && sleepId[b] != 0u) ? 0xFFFFFFFFu : rootScan[bodyToDyn[islandId[b]]];
    }
}

__global__ void kIslandCount(Counters* ctr) { ctr->islands = ctr->scratch[3]; }

// ---------------------------------------------------------------- per-scene islands
// After a per-scene solve, each scene's islands come from one workgroup: union-find over the
// scene's dynamic contacts in LDS (a root is hooked under the smaller root, so every root is its
// component's smallest body), then the roots ranked in body order. A scan over the scenes'
// counts turns the local ranks into the ids the device-wide union-find gives: islands numbered by
// ascending root body.
constexpr uint32_t ISLAND_THREADS = 512;

__device__ inline uint32_t ldsFind(uint32_t* par, uint32_t x) {
    uint32_t p = par[x];
    while (p != x) {
        const uint32_t g = par[p];
        par[x] = g;   // path halving: both values are ancestors of x
        x = p;
        p = g;
    }
    return x;
}

__global__ __launch_bounds__(ISLAND_THREADS) void kSceneIslands(Constraints cs, const float2* __restrict__ inv,
                                                                const uint32_t* __restrict__ sceneBodyStart,
                                                                const uint2* __restrict__ range, const Counters* __restrict__ ctr,
                                                                uint32_t* __restrict__ islandId,
                                                                uint32_t* __restrict__ sceneIslands,
                                                                const uint32_t* __restrict__ sleepId) {
    // a member: a dynamic body that is awake (sleeping bodies touch no manifold and join no island)
    auto member = [&](uint32_t body) { return inv[body].x != 0.0f && !(sleepId && sleepId[body] != 0u); };
    extern __shared__ uint32_t par[];   // the scene's bodies, then the root ranks
    __shared__ uint32_t waveSum[ISLAND_THREADS / WAVE];
    const uint32_t s = blockIdx.x;
    const uint32_t b0 = sceneBodyStart[s], b1 = sceneBodyStart[s + 1], n = b1 - b0;
    const uint32_t t = threadIdx.x;
    const uint32_t K = ctr->colours;
    for (uint32_t i = t; i < n; i += ISLAND_THREADS) par[i] = i;
    __syncthreads();
    for (uint32_t k = 0; k < K; k++) {
        const uint2 r = range[(size_t)s * MAX_COLOURS + k];
        for (uint32_t c = r.x + t; c < r.y; c += ISLAND_THREADS) {
            const uint4 id = cs.idx[c];
            if (id.z == 0 || id.x < b0 || id.x >= b1 || id.y < b0 || id.y >= b1) continue;
            const uint4 dep = cs.dep[c];
            if (!dep.z || !dep.w) continue;   // contacts with a static body join nothing
            uint32_t a = ldsFind(par, id.x - b0), b = ldsFind(par, id.y - b0);
            while (a != b) {
                if (a > b) { const uint32_t x = a; a = b; b = x; }
                const uint32_t old = atomicCAS(par + b, b, a);
                if (old == b) break;
                b = ldsFind(par, old);
            }
        }
    }
    __syncthreads();
    // roots of dynamic bodies, ranked in body order: each thread takes a contiguous run of bodies
    const uint32_t per = (n + ISLAND_THREADS - 1) / ISLAND_THREADS;
    const uint32_t lo = min(n, t * per), hi = min(n, lo + per);
    uint32_t roots = 0;
    for (uint32_t i = lo; i < hi; i++) {
        uint32_t x = i, p = par[x];
        while (p != x) { x = p; p = par[x]; }
        roots += (x == i && member(b0 + i)) ? 1u : 0u;
    }
    uint32_t incl = roots;
    const uint32_t lane = t & (WAVE - 1), wid = t / WAVE;
    for (uint32_t off = 1; off < WAVE; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, WAVE);
        if (lane >= off) incl += v;
    }
    if (lane == WAVE - 1) waveSum[wid] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < ISLAND_THREADS / WAVE; w++) {
        if (w < wid) base += waveSum[w];
        total += waveSum[w];
    }
    // root ranks go to the second half of the LDS array (par stays intact for the lookups)
    uint32_t* rankOf = par + n;
    uint32_t next = base + incl - roots;
    for (uint32_t i = lo; i < hi; i++) {
        uint32_t x = i, p = par[x];
        while (p != x) { x = p; p = par[x]; }
        if (x == i && member(b0 + i)) rankOf[i] = next++;
    }
    __syncthreads();
    for (uint32_t i = t; i < n; i += ISLAND_THREADS) {
        uint32_t x = i, p = par[x];
        while (p != x) { x = p; p = par[x]; }
        islandId[b0 + i] = member(b0 + i) ? rankOf[x] : 0xFFFFFFFFu;
    }
    if (t == 0) sceneIslands[s] = total;
}

// Local island ranks plus the island count of all earlier scenes.
__global__ __launch_bounds__(256) void kSceneIslandOffsets(const uint32_t* __restrict__ bodyScene,
                                                           const uint32_t* __restrict__ offset, uint32_t n,
                                                           uint32_t* __restrict__ islandId) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        const uint32_t id = islandId[b];
        if (id != 0xFFFFFFFFu) islandId[b] = id + offset[bodyScene[b]];
    }
}

}  // namespace

// Launches a dataflow kernel whose blocks wait on one another, so all of them must be resident at
// once. The grid never exceeds the kernel's occupancy (blocks per CU times CUs), and a cooperative
// launch has the runtime check that. The runtime refuses a cooperative launch with
// hipErrorIllegalState while another process on the device holds its cooperative-launch resource;
// the same grid is then launched as an ordinary kernel. Without other work on the device every block
// is resident; if another process's kernels keep some blocks out, every wait ends at its 2 s limit
// and flags OVERFLOW_STALL, so the step reports an error instead of hanging. Inside a stream capture
// (step graphs) the ordinary launch is what gets recorded, so the grid is checked against the
// occupancy here, before either launch: co-residency then rests on that bound alone.
static hipError_t launchResident(const void* kernel, uint32_t blocks, int numCU, void** args, hipStream_t st) {
    int perCu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, 256, 0) != hipSuccess) return hipErrorInvalidValue;
    if (blocks == 0 || (uint64_t)blocks > (uint64_t)std::max(perCu, 0) * (uint64_t)std::max(numCU, 0))
        return hipErrorInvalidConfiguration;
    hipError_t e = hipLaunchCooperativeKernel(kernel, dim3(blocks), dim3(256), args, 0, st);
    if (e != hipErrorIllegalState) return e;
    (void)hipGetLastError();
    return hipLaunchKernel(kernel, dim3(blocks), dim3(256), args, 0, st);
}

void dataflowOccupancy(int& colour, int& solve) {
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void*>(kColourFlow), 256, 0) != hipSuccess) a = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(kSolveFlow), 256, 0) != hipSuccess) b = 1;
    colour = std::max(1, a);
    solve = std::max(1, b);
}

void dataflowBlocksPerCU(int& colour, int& solve) {
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void*>(kColourFlow), 256, 0) != hipSuccess) a = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(kSolveFlow), 256, 0) != hipSuccess) b = 1;
    colour = std::max(1, std::min(a, 4));
    solve = std::max(1, std::min(b, 4));
}

hipError_t launchColourPrep(hipStream_t st, DeviceWorld& w, bool coloursGiven) {
    const uint32_t gm = blocksForManifolds(w);
    hipLaunchKernelGGL(kFlowClear, dim3(blocksFor(w.N)), dim3(256), 0, st, w.N, w.degree, w.incFill, w.colRec);
    hipLaunchKernelGGL(kDegree, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody, w.inv, w.manBodies, w.degree);
    hipError_t e = launchScan(st, w.degree, w.incStart, nullptr, &w.ctr->scratch[4], w.N, w.scanBlock, w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kIncident, dim3(gm), dim3(256), 0, st, w.ctr, w.manBodies, w.inv, w.incStart, w.incFill, w.inc);
    const uint32_t ge = blocksFor((uint32_t)std::min<uint64_t>(2 * w.manCap, 1u << 22));
    hipLaunchKernelGGL(kRankEntries<false>, dim3(ge), dim3(256), 0, st, w.ctr, w.manBodies, w.degree, w.incStart, w.inc,
                       w.posOf, w.rank, w.incSorted);
    void* args[] = { (void*)&w.ctr, (void*)&w.manBodies, (void*)&w.inv, (void*)&w.rank, (void*)&w.colRec,
                     (void*)&w.colourId };
    if (!coloursGiven) {
        e = launchResident(reinterpret_cast<const void*>(kColourFlow), (uint32_t)w.colourBlocks, w.numCU, args, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kColourHist, dim3(COLOUR_BLOCKS), dim3(256), 0, st, w.ctr, w.colourId, w.colourHist);
    hipLaunchKernelGGL(kColourColumnScan, dim3(MAX_COLOURS), dim3(256), 0, st, w.ctr, w.colourHist, COLOUR_BLOCKS);
    hipLaunchKernelGGL(kColourStarts, dim3(1), dim3(64), 0, st, w.ctr);
    hipLaunchKernelGGL(kColourScatter, dim3(COLOUR_BLOCKS), dim3(256), 0, st, w.ctr, w.colourId, w.colourHist, w.order,
                       w.posOf);
    hipLaunchKernelGGL(kMarkHoles, dim3((MAX_COLOURS * COLOUR_ALIGN + 255) / 256), dim3(256), 0, st, w.ctr, w.order, w.cs);
    hipLaunchKernelGGL(kRankEntries<true>, dim3(ge), dim3(256), 0, st, w.ctr, w.manBodies, w.degree, w.incStart, w.inc,
                       w.posOf, w.rank, w.incSorted);
    hipLaunchKernelGGL(kPrep<true>, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.manBodies, w.posOf, w.degree, w.rank,
                       w.incStart, w.incSorted, w.inv, w.pos, w.vel, w.shape, w.materials, w.cs);
    return hipGetLastError();
}

// Colouring and constraint data for the per-scene solve: each scene coloured by one workgroup in
// LDS, then the same stable colour scatter and constraint data (without predecessor slots).
hipError_t launchColourPrepScene(hipStream_t st, DeviceWorld& w) {
    const uint32_t gm = blocksForManifolds(w);
    // the colour ranges are filled by the scene solve that always follows in the same phase
    hipLaunchKernelGGL(kSceneRangesClear, dim3(blocksFor(w.S * (MAX_COLOURS + 1))), dim3(256), 0, st, w.sceneManRange,
                       w.sceneRange, w.S);
    hipLaunchKernelGGL(kSceneManRanges, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colScene, w.sceneManRange,
                       w.manBodies);
    const SceneLaunch& sl = w.sceneLaunch;
    auto capped = [&](uint32_t layout) {   // at most sceneColourCap manifolds in LDS
        return (layout & 0xFFFF0000u) | std::min<uint32_t>(layout & 0xFFFFu, w.sceneColourCap);
    };
    const uint32_t blocks = (uint32_t)std::max(1, std::min<int>(sl.colourBlocks, (int)w.S));
    hipLaunchKernelGGL(kSceneColour, dim3(blocks), dim3(COLOUR_THREADS), w.sceneColourLdsBytes, st, w.man[w.cur], w.ctr,
                       w.colBody, w.inv, w.sceneBodyStart, w.sceneOrder, w.sceneManRange, w.S, sl.colourWhole, sl.colourGroups,
                       capped(sl.colourLayoutWhole), capped(sl.colourLayoutGroup), sl.colourGroupWords, w.manBodies,
                       w.degree, w.rank, w.colourId);
    hipLaunchKernelGGL(kColourHist, dim3(COLOUR_BLOCKS), dim3(256), 0, st, w.ctr, w.colourId, w.colourHist);
    hipLaunchKernelGGL(kColourColumnScan, dim3(MAX_COLOURS), dim3(256), 0, st, w.ctr, w.colourHist, COLOUR_BLOCKS);
    hipLaunchKernelGGL(kColourStarts, dim3(1), dim3(64), 0, st, w.ctr);
    hipLaunchKernelGGL(kColourScatter, dim3(COLOUR_BLOCKS), dim3(256), 0, st, w.ctr, w.colourId, w.colourHist, w.order,
                       w.posOf);
    hipLaunchKernelGGL(kMarkHoles, dim3((MAX_COLOURS * COLOUR_ALIGN + 255) / 256), dim3(256), 0, st, w.ctr, w.order, w.cs);
    hipLaunchKernelGGL(kPrep<false>, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.manBodies, w.posOf, w.degree, w.rank,
                       w.incStart, w.incSorted, w.inv, w.pos, w.vel, w.shape, w.materials, w.cs);
    return hipGetLastError();
}

hipError_t launchWarmStart(hipStream_t st, DeviceWorld& w) {
    const uint32_t gd = blocksFor(std::max(1u, w.nDynBodies));
    hipLaunchKernelGGL(kWarm, dim3(gd), dim3(256), 0, st, w.dynBodies, w.nDynBodies, w.degree, w.incStart, w.incSorted,
                       w.posOf, w.cs, w.vel);
    // cleared for the solve kernel: its grid-barrier counter and the group progress flags
    hipError_t e = launchFillU32(st, &w.ctr->barrier[1], 0u, 1);
    if (e != hipSuccess) return e;
    return launchFillU32(st, w.groupDone, 0u, 2ull * w.solveGroups);
}

hipError_t launchSolve(hipStream_t st, DeviceWorld& w, uint32_t vIt, uint32_t pIt, int trackResidual,
                       float residualTarget) {
    vIt = std::min(vIt, MAX_SWEEPS);
    // group progress markers count sweep * positions: keep them within 32 bits
    if ((uint64_t)(std::max(vIt, pIt) + 1) * w.solveCap >= (1ull << 32)) return hipErrorInvalidValue;
    uint32_t* vDone = w.groupDone;
    uint32_t* pDone = w.groupDone + w.solveGroups;
    void* args[] = { (void*)&w.ctr, (void*)&w.cs, (void*)&w.vel, (void*)&w.pos, (void*)&vDone, (void*)&pDone,
                     (void*)&w.blockRes, (void*)&vIt, (void*)&pIt, (void*)&trackResidual, (void*)&residualTarget };
    hipError_t e = launchResident(reinterpret_cast<const void*>(kSolveFlow), (uint32_t)w.solveBlocks, w.numCU, args, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kFlowTotals, dim3(1), dim3(1), 0, st, w.ctr, w.nDynBodies, w.S);
    return hipGetLastError();
}

// Body records each of `groups` equal LDS regions holds, rounded down to a multiple of four so that
// every region starts 16-byte aligned.
static uint32_t groupBodiesOf(size_t bytes, uint32_t groups) {
    return (uint32_t)(bytes / groups / (sizeof(float4) + sizeof(float2))) & ~3u;
}

uint32_t sceneSolveSetup(const uint32_t* orderedBodies, uint32_t S, int numCU, uint32_t wholeCap, SceneLaunch& out,
                         uint32_t& colourLdsBytes) {
    out = SceneLaunch{};
    colourLdsBytes = 0;
    const uint32_t maxBodies = S ? orderedBodies[0] : 0u;
    hipDeviceProp_t prop{};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    const size_t staticBytes = SCENE_GROUPS * MAX_COLOURS * sizeof(uint2) + (SCENE_THREADS / WAVE) * sizeof(double2) +
                               2 * SCENE_GROUPS * sizeof(uint32_t) + 256;
    const size_t limit = std::max<size_t>(prop.sharedMemPerBlock, (size_t)prop.maxSharedMemoryPerMultiProcessor);
    if (limit <= staticBytes) return 0;
    // all the LDS a block may have (the block holds the CU alone): the largest scene's body records
    // for a whole block, or `groups` smaller scenes side by side
    const size_t bytes = (limit - staticBytes) & ~(size_t)255;
    constexpr size_t BODY_LDS = sizeof(float4) + sizeof(float2);   // velocity or position record, inverse mass/inertia
    if (std::max<size_t>(maxBodies, 1) * BODY_LDS > bytes) return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kSceneSolve), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)bytes) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    // the per-scene colouring, also one block per CU with all the LDS: per scene 24 bytes per body,
    // then as many of its manifolds (6 bytes each) as the rest of the region holds (a scene with
    // more colours them in global scratch)
    if (maxBodies >= 0x8000u) return 0;   // local body indices are packed in 15 bits
    const size_t colourStatic = 4 * SCENE_GROUPS * sizeof(uint32_t) + 256;
    const size_t colourBytes = (limit - colourStatic) & ~(size_t)255;
    auto capOf = [](size_t region, uint32_t bodies) -> size_t {   // manifolds after the body records
        const size_t bodyBytes = (size_t)std::max(bodies, 1u) * 6 * sizeof(uint32_t) + 8;   // + rounding of the colours
        return region > bodyBytes ? std::min<size_t>((region - bodyBytes) / 6, 0xFFFF) : 0;
    };
    if (capOf(colourBytes, maxBodies) < 64) return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kSceneColour), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)colourBytes) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    colourLdsBytes = (uint32_t)colourBytes;
    int perCu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, reinterpret_cast<const void*>(kSceneSolve), SCENE_THREADS,
                                                     bytes) != hipSuccess || perCu < 1) {
        (void)hipGetLastError();
        return 0;
    }
    out.blocks = numCU * perCu;
    // one and a half heavy scenes per block take whole blocks (measured on the 1,024-scene batch:
    // 1 % faster than one per block, 3 % faster than two)
    out.whole = std::min<uint32_t>(S, wholeCap == UINT32_MAX ? (uint32_t)out.blocks * 3 / 2 : wholeCap);
    out.groups = 1;
    if (out.whole < S) {
        if (S <= WORK_SORT_MAX) {
            out.groups = SCENE_GROUPS;   // scenes larger than a group holds go to whole blocks
        } else {                         // size order: the scenes after the whole ones fit
            const uint32_t next = std::max(orderedBodies[out.whole], 1u);
            while (out.groups < SCENE_GROUPS && groupBodiesOf(bytes, 2 * out.groups) >= next) out.groups *= 2;
        }
    }
    out.groupBodies = groupBodiesOf(bytes, out.groups);
    // colouring: the same whole scenes, then as many groups (at most 4) as leave room for three
    // manifolds per body of the largest remaining scene
    int colourPerCu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&colourPerCu, reinterpret_cast<const void*>(kSceneColour),
                                                     COLOUR_THREADS, colourBytes) != hipSuccess || colourPerCu < 1) {
        (void)hipGetLastError();
        return 0;
    }
    out.colourBlocks = numCU * colourPerCu;
    out.colourWhole = std::min<uint32_t>(S, wholeCap == UINT32_MAX ? (uint32_t)out.colourBlocks : wholeCap);
    out.colourLayoutWhole = (std::max(maxBodies, 1u) << 16) | (uint32_t)capOf(colourBytes, maxBodies);
    out.colourGroups = 1;
    out.colourLayoutGroup = out.colourLayoutWhole;
    auto groupWords = [&](uint32_t groups) { return (uint32_t)(colourBytes / groups / sizeof(uint32_t)) & ~3u; };
    if (out.colourWhole < S) {
        const uint32_t next = std::max(orderedBodies[out.colourWhole], 1u);
        out.colourGroups = SCENE_GROUPS;
        while (out.colourGroups > 1 && capOf(4ull * groupWords(out.colourGroups), next) < 3ull * next) out.colourGroups /= 2;
        out.colourLayoutGroup = (next << 16) | (uint32_t)capOf(4ull * groupWords(out.colourGroups), next);
    }
    out.colourGroupWords = groupWords(out.colourGroups);
    // each colouring region starts with 64-bit round-tagged minima: region offsets (group index x
    // colourGroupWords words) must stay 8-byte aligned
    if (out.colourGroupWords % 2u != 0u) return 0;
    return (uint32_t)bytes;
}

hipError_t launchSceneSolve(hipStream_t st, DeviceWorld& w, uint32_t vIt, uint32_t pIt, int trackResidual,
                            float residualTarget) {
    vIt = std::min(vIt, MAX_SWEEPS);
    // sceneRange was emptied by the per-scene colouring of this phase
    hipLaunchKernelGGL(kSceneRanges, dim3(blocksForManifolds(w)), dim3(256), 0, st, w.ctr, w.cs, w.bodyScene, w.sceneRange);
    const int track = (trackResidual || residualTarget > 0.0f) ? 1 : 0;
    const SceneLaunch& sl = w.sceneLaunch;
    hipLaunchKernelGGL(kSceneWorkOrder, dim3(1), dim3(1024), 0, st, w.ctr, w.sceneOrder, w.sceneBodyStart, w.sceneSweeps,
                       w.sceneDyn, w.S, sl.whole, sl.groupBodies, w.sceneWorkOrder);
    // every block takes a whole scene first; with fewer scenes than blocks, no more blocks than scenes
    const uint32_t blocks = (uint32_t)std::max(1, std::min<int>(sl.blocks, (int)std::max(w.S, sl.whole)));
    hipLaunchKernelGGL(kSceneSolve, dim3(blocks), dim3(SCENE_THREADS), w.sceneLdsBytes, st, w.ctr, w.cs, w.vel, w.pos,
                       w.inv, w.sceneBodyStart, w.sceneWorkOrder, w.sceneDyn, w.sceneRange, w.sceneRes, w.sceneSweeps, w.S,
                       std::max(w.sceneMaxBodies, 1u), sl.groups, sl.groupBodies, vIt, pIt, track, residualTarget);
    hipLaunchKernelGGL(kSceneTotals, dim3(1), dim3(1), 0, st, w.ctr);
    return hipGetLastError();
}

hipError_t launchSolveFinish(hipStream_t st, DeviceWorld& w, int trackResidual, float residualTarget, int trackErrors,
                             int perScene) {
    hipLaunchKernelGGL(kSolveFinish, dim3(FINISH_BLOCKS), dim3(256), 0, st, w.ctr, w.cs, w.vel, w.pos, w.man[w.cur],
                       w.errPartial, trackErrors);
    const int track = (trackResidual || residualTarget > 0.0f) ? 1 : 0;
    if (perScene && track)
        hipLaunchKernelGGL(kSceneReduce, dim3(MAX_SWEEPS), dim3(256), 0, st, w.ctr, w.sceneRes, w.sceneSweeps, w.S);
    const int flowTrack = perScene ? 0 : track;
    if (flowTrack || trackErrors)
        hipLaunchKernelGGL(kSolveReduce, dim3(1), dim3(256), 0, st, w.ctr, w.blockRes, (uint32_t)w.solveBlocks, w.errPartial,
                           FINISH_BLOCKS, flowTrack, trackErrors);
    return hipGetLastError();
}

hipError_t launchSceneIslands(hipStream_t st, DeviceWorld& w) {
    const uint32_t bytes = 2 * std::max<uint32_t>(w.sceneMaxBodies, 1) * sizeof(uint32_t);
    hipLaunchKernelGGL(kSceneIslands, dim3(w.S), dim3(ISLAND_THREADS), bytes, st, w.cs, w.inv, w.sceneBodyStart, w.sceneRange,
                       w.ctr, w.islandId, w.sceneIslands, w.sleepId);
    hipError_t e = launchScan(st, w.sceneIslands, w.sceneIslandOffset, nullptr, &w.ctr->islands, w.S, w.scanBlock,
                              w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kSceneIslandOffsets, dim3(blocksFor(w.N)), dim3(256), 0, st, w.bodyScene, w.sceneIslandOffset, w.N,
                       w.islandId);
    return hipGetLastError();
}

hipError_t launchIslands(hipStream_t st, DeviceWorld& w) {
    hipLaunchKernelGGL(kUfInit, dim3(blocksFor(w.N)), dim3(256), 0, st, w.ufParent, w.N);
    hipLaunchKernelGGL(kUfLinkNeighbours, dim3(blocksForManifolds(w)), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody,
                       w.inv, w.ufParent);
    if (w.nDynBodies) {
        hipLaunchKernelGGL(kUfCompress, dim3(blocksFor(w.nDynBodies)), dim3(256), 0, st, w.dynBodies, w.nDynBodies, w.ufParent);
        // one hooking round (a second merges little more than it costs on a 65,540-body pile)
        hipLaunchKernelGGL(kUfHook, dim3(blocksForManifolds(w)), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody, w.inv,
                           w.ufParent);
        hipLaunchKernelGGL(kUfCompress, dim3(blocksFor(w.nDynBodies)), dim3(256), 0, st, w.dynBodies, w.nDynBodies,
                           w.ufParent);
    }
    hipLaunchKernelGGL(kUfUnite, dim3(blocksFor((uint32_t)std::min<uint64_t>(w.manCap, 1u << 22))), dim3(256), 0, st,
                       w.man[w.cur], w.ctr, w.colBody, w.inv, w.ufParent);
    if (w.nJoints)
        hipLaunchKernelGGL(kUfUniteEdges, dim3(blocksFor(w.nJoints)), dim3(256), 0, st, w.jointEdges, w.nJoints, w.sleepId,
                           w.ufParent);
    if (w.nDynBodies == 0) return launchFillU32(st, &w.ctr->islands, 0u, 1);
    hipLaunchKernelGGL(kUfRoots, dim3(blocksFor(w.nDynBodies)), dim3(256), 0, st, w.ufParent, w.dynBodies, w.nDynBodies,
                       w.ufRoot, w.islandId, w.sleepId);
    hipError_t e = launchScan(st, w.ufRoot, w.ufRoot, nullptr, &w.ctr->scratch[3], w.nDynBodies, w.scanBlock, w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kIslandIds, dim3(blocksFor(w.nDynBodies)), dim3(256), 0, st, w.dynBodies, w.nDynBodies, w.ufRoot,
                       w.bodyToDyn, w.islandId, w.sleepId);
    hipLaunchKernelGGL(kIslandCount, dim3(1), dim3(1), 0, st, w.ctr);
    return hipGetLastError();
}

}  // namespace agx
