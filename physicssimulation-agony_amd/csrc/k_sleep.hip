// Sleeping and waking on the device: ObjectManager::updateSleeping / sleepBody / wakeSleepIslands
// (SimulationImpl/ObjectManager.cpp:1314-1491), wakeIslandsOfContacts (ObjectManager.h:139-156) and
// the contact stash of NarrowPhaseCollisionDetector::sleepContactData / wakeContactData
// (Narrow/NarrowPhaseCollisionDetector.cpp:299-414).
//
// Bodies keep their indices: a sleeping body carries its island's sleep id (2 + the island's
// smallest body index, unique among the islands asleep at once) in sleepId, 0 when awake. Sleeping
// bodies are not integrated and pair only with awake dynamic bodies (k_broad.hip); a step's
// manifolds that touch a sleeping body wake its island, and the island's stashed manifolds are
// appended to the step's manifolds before the solve. After the islands, quiet islands fall asleep
// and their manifolds move to the stash. Every count lives on the device, so the kernels run in
// captured step graphs; they return at once when no body woke or fell asleep.
#include "agx_internal.h"

#include <algorithm>

namespace agx {

namespace {

constexpr uint32_t FIRST_SLEEP_ISLAND_ID = 2;   // SoA/BodySoA.h:45

inline uint32_t gridFor(uint64_t n) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192));
}

// ---------------------------------------------------------------- wake, after the narrow phase

// Sleep ids of the bodies in the step's manifolds (only pairs with an awake body exist).
__global__ __launch_bounds__(256) void kWakeMark(const agx_manifold* __restrict__ man, Counters* __restrict__ ctr,
                                                 const uint32_t* __restrict__ colBody, const uint32_t* __restrict__ sleepId,
                                                 uint32_t* __restrict__ woken) {
    const uint32_t M = ctr->manifolds;
    bool any = false;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint32_t sa = sleepId[colBody[man[m].a]], sb = sleepId[colBody[man[m].b]];
        if (sa) { woken[sa - FIRST_SLEEP_ISLAND_ID] = 1u; any = true; }
        if (sb) { woken[sb - FIRST_SLEEP_ISLAND_ID] = 1u; any = true; }
    }
    if (__ballot(any) != 0 && (threadIdx.x & (WAVE - 1)) == 0) ctr->wokeAny = 1u;
}

// wakeOne for every body of a woken island (ObjectManager.cpp:1344-1361): awake, timer 0.
__global__ __launch_bounds__(256) void kWakeBodies(const uint32_t* __restrict__ dynBodies, uint32_t n,
                                                   const Counters* __restrict__ ctr, uint32_t* __restrict__ sleepId,
                                                   float* __restrict__ sleepTimer, const uint32_t* __restrict__ woken) {
    if (ctr->wokeAny == 0u) return;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = dynBodies[k], s = sleepId[b];
        if (s && woken[s - FIRST_SLEEP_ISLAND_ID]) {
            sleepId[b] = 0u;
            sleepTimer[b] = 0.0f;
        }
    }
}

// Which stashed manifolds belong to a woken island (none scanned when nothing woke).
__global__ __launch_bounds__(256) void kStashWakeFlags(Counters* __restrict__ ctr, const uint32_t* __restrict__ stashIsland,
                                                       const uint32_t* __restrict__ woken, uint32_t* __restrict__ flags) {
    const uint32_t n = ctr->wokeAny ? ctr->stashCount : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->stashScanCount = n;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
        flags[s] = woken[stashIsland[s] - FIRST_SLEEP_ISLAND_ID];
}

// A woken island's manifolds go after the step's own, in stash order (wakeContactData: no pair
// of them was generated this frame, as both bodies, or one and a static body, slept through the
// broad phase); the rest of the stash is compacted into the scratch copy.
__global__ __launch_bounds__(256) void kStashRestore(Counters* __restrict__ ctr, const agx_manifold* __restrict__ stash,
                                                     const uint32_t* __restrict__ stashIsland,
                                                     const uint32_t* __restrict__ flags, const uint32_t* __restrict__ scan,
                                                     agx_manifold* __restrict__ man, uint64_t manCap,
                                                     agx_manifold* __restrict__ keepMan, uint32_t* __restrict__ keepIsland,
                                                     float4* __restrict__ histOut) {
    const uint32_t n = ctr->stashScanCount, M = ctr->manifolds;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const uint32_t before = scan[s];
        if (flags[s]) {
            const uint64_t dst = (uint64_t)M + before;
            if (dst < manCap) {
                man[dst] = stash[s];
                if (histOut) histOut[dst] = make_float4(__int_as_float(0x7fc00000), __int_as_float(0x7fc00000),
                                                        __int_as_float(0x7fc00000), __int_as_float(0x7fc00000));
            } else {
                atomicOr(&ctr->overflow, OVERFLOW_PAIRS);
            }
        } else {
            keepMan[s - before] = stash[s];
            keepIsland[s - before] = stashIsland[s];
        }
    }
}

__global__ void kStashRestoreEnd(Counters* ctr, uint64_t manCap) {
    if (threadIdx.x != 0 || ctr->stashScanCount == 0u) return;
    const uint32_t restored = ctr->restoreTotal;
    const uint64_t total = (uint64_t)ctr->manifolds + restored;
    ctr->manifolds = (uint32_t)std::min<uint64_t>(total, manCap);
    ctr->stashCount -= restored;
}

__global__ __launch_bounds__(256) void kStashCopyBack(const Counters* __restrict__ ctr, const agx_manifold* __restrict__ keepMan,
                                                      const uint32_t* __restrict__ keepIsland, agx_manifold* __restrict__ stash,
                                                      uint32_t* __restrict__ stashIsland) {
    const uint32_t n = ctr->stashScanCount ? ctr->stashCount : 0u;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        stash[s] = keepMan[s];
        stashIsland[s] = keepIsland[s];
    }
}

// ---------------------------------------------------------------- sleep, after the islands

__global__ __launch_bounds__(256) void kSleepInit(const Counters* __restrict__ ctr, uint32_t* __restrict__ islandFlag,
                                                  uint32_t* __restrict__ islandMin, uint32_t n) {
    const uint32_t K = min(ctr->islands, n);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x) {
        islandFlag[k] = 1u;
        islandMin[k] = 0xFFFFFFFFu;
    }
}

// Timers of the awake dynamic bodies and the AND of readiness per island (ObjectManager.cpp:1444-1470):
// speed = |v| + |w| x bound radius; quiet when at most the speed threshold; the timer grows by dt
// while quiet and restarts at 0 otherwise; an island is ready when all its bodies are quiet and have
// been for the time threshold. Bodies without an island (woken after the islands) stay awake.
__global__ __launch_bounds__(256) void kSleepTimers(const uint32_t* __restrict__ dynBodies, uint32_t n,
                                                    const float4* __restrict__ vel, const float* __restrict__ boundRadius,
                                                    const uint32_t* __restrict__ islandId, const Counters* __restrict__ ctr,
                                                    const uint32_t* __restrict__ sleepId, float* __restrict__ sleepTimer,
                                                    uint32_t* __restrict__ islandFlag, uint32_t* __restrict__ islandMin,
                                                    float dt, float speedThreshold, float timeThreshold) {
    const uint32_t K = ctr->islands;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = dynBodies[k];
        if (sleepId[b]) continue;
        const float4 v = vel[b];
        const float speed = sqrtf(v.x * v.x + v.y * v.y) + fabsf(v.z) * boundRadius[b];
        const bool quiet = speed <= speedThreshold;
        const float timer = (sleepTimer[b] + dt) * (float)quiet;
        sleepTimer[b] = timer;
        const uint32_t isl = islandId[b];
        if (isl >= K) continue;
        if (!(quiet && timer >= timeThreshold)) islandFlag[isl] = 0u;
        atomicMin(&islandMin[isl], b);
    }
}

// sleepBody for every body of a ready island (ObjectManager.cpp:1314-1342): sleep id, timer 0.
__global__ __launch_bounds__(256) void kSleepMark(const uint32_t* __restrict__ dynBodies, uint32_t n,
                                                  const uint32_t* __restrict__ islandId, Counters* __restrict__ ctr,
                                                  const uint32_t* __restrict__ islandFlag,
                                                  const uint32_t* __restrict__ islandMin, uint32_t* __restrict__ sleepId,
                                                  float* __restrict__ sleepTimer) {
    const uint32_t K = ctr->islands;
    bool any = false;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = dynBodies[k], isl = islandId[b];
        if (sleepId[b] || isl >= K || !islandFlag[isl]) continue;
        sleepId[b] = FIRST_SLEEP_ISLAND_ID + islandMin[isl];
        sleepTimer[b] = 0.0f;
        any = true;
    }
    if (__ballot(any) != 0 && (threadIdx.x & (WAVE - 1)) == 0) ctr->sleptAny = 1u;
}

// sleepContactData: which of the step's manifolds touch a body that fell asleep (none scanned when
// no body did).
__global__ __launch_bounds__(256) void kStashSleepFlags(const agx_manifold* __restrict__ man, Counters* __restrict__ ctr,
                                                        const uint32_t* __restrict__ colBody,
                                                        const uint32_t* __restrict__ sleepId, uint32_t* __restrict__ flags) {
    const uint32_t n = ctr->sleptAny ? ctr->manifolds : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->sleepScanCount = n;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x)
        flags[m] = (sleepId[colBody[man[m].a]] | sleepId[colBody[man[m].b]]) != 0u ? 1u : 0u;
}

// A stable split: the sleepers' manifolds to the end of the stash, filed under A's island if A
// sleeps, else B's; the rest compacted into the scratch copy.
__global__ __launch_bounds__(256) void kStashSplit(const agx_manifold* __restrict__ man, Counters* __restrict__ ctr,
                                                   const uint32_t* __restrict__ colBody, const uint32_t* __restrict__ sleepId,
                                                   const uint32_t* __restrict__ flags, const uint32_t* __restrict__ scan,
                                                   agx_manifold* __restrict__ stash, uint32_t* __restrict__ stashIsland,
                                                   uint64_t stashCap, agx_manifold* __restrict__ keepMan) {
    const uint32_t n = ctr->sleepScanCount, base = ctr->stashCount;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
        const uint32_t before = scan[m];
        const agx_manifold d = man[m];
        if (flags[m]) {
            const uint64_t dst = (uint64_t)base + before;
            if (dst < stashCap) {
                const uint32_t sa = sleepId[colBody[d.a]];
                stash[dst] = d;
                stashIsland[dst] = sa ? sa : sleepId[colBody[d.b]];
            } else {
                atomicOr(&ctr->overflow, OVERFLOW_PAIRS);
            }
        } else {
            keepMan[m - before] = d;
        }
    }
}

__global__ void kStashSplitEnd(Counters* ctr, uint64_t stashCap) {
    if (threadIdx.x != 0 || ctr->sleepScanCount == 0u) return;
    const uint32_t stashed = ctr->stashTotal;
    ctr->stashCount = (uint32_t)std::min<uint64_t>((uint64_t)ctr->stashCount + stashed, stashCap);
    ctr->manifolds -= stashed;
}

__global__ __launch_bounds__(256) void kKeptBack(const Counters* __restrict__ ctr, const agx_manifold* __restrict__ keepMan,
                                                 agx_manifold* __restrict__ man) {
    const uint32_t n = ctr->sleepScanCount ? ctr->manifolds : 0u;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) man[m] = keepMan[m];
}

}  // namespace

hipError_t launchWake(hipStream_t st, DeviceWorld& w) {
    hipError_t e = launchFillU32(st, w.woken, 0u, w.N);
    if (e != hipSuccess) return e;
    const uint32_t gm = gridFor(std::min<uint64_t>(w.manCap, 1u << 22));
    hipLaunchKernelGGL(kWakeMark, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody, w.sleepId, w.woken);
    hipLaunchKernelGGL(kWakeBodies, dim3(gridFor(w.nDynBodies)), dim3(256), 0, st, w.dynBodies, w.nDynBodies, w.ctr,
                       w.sleepId, w.sleepTimer, w.woken);
    const uint32_t gs = gridFor(std::min<uint64_t>(w.manCap, 1u << 22));
    hipLaunchKernelGGL(kStashWakeFlags, dim3(gs), dim3(256), 0, st, w.ctr, w.stashIsland, w.woken, w.flags);
    e = launchScan(st, w.flags, w.flagScan, &w.ctr->stashScanCount, &w.ctr->restoreTotal, w.manCap, w.scanBlock,
                   w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kStashRestore, dim3(gs), dim3(256), 0, st, w.ctr, w.stash, w.stashIsland, w.flags, w.flagScan,
                       w.man[w.cur], w.manCap, w.slots, w.stashIslandAlt, w.histOut);
    hipLaunchKernelGGL(kStashRestoreEnd, dim3(1), dim3(64), 0, st, w.ctr, w.manCap);
    hipLaunchKernelGGL(kStashCopyBack, dim3(gs), dim3(256), 0, st, w.ctr, w.slots, w.stashIslandAlt, w.stash,
                       w.stashIsland);
    return hipGetLastError();
}

hipError_t launchSleep(hipStream_t st, DeviceWorld& w, float dt, float speedThreshold, float timeThreshold) {
    if (w.nDynBodies == 0) return hipSuccess;
    const uint32_t gd = gridFor(w.nDynBodies);
    hipLaunchKernelGGL(kSleepInit, dim3(gd), dim3(256), 0, st, w.ctr, w.islandFlag, w.islandMin, w.nDynBodies);
    hipLaunchKernelGGL(kSleepTimers, dim3(gd), dim3(256), 0, st, w.dynBodies, w.nDynBodies, w.vel, w.boundRadius,
                       w.islandId, w.ctr, w.sleepId, w.sleepTimer, w.islandFlag, w.islandMin, dt, speedThreshold,
                       timeThreshold);
    hipLaunchKernelGGL(kSleepMark, dim3(gd), dim3(256), 0, st, w.dynBodies, w.nDynBodies, w.islandId, w.ctr, w.islandFlag,
                       w.islandMin, w.sleepId, w.sleepTimer);
    const uint32_t gm = gridFor(std::min<uint64_t>(w.manCap, 1u << 22));
    hipLaunchKernelGGL(kStashSleepFlags, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody, w.sleepId, w.flags);
    hipError_t e = launchScan(st, w.flags, w.flagScan, &w.ctr->sleepScanCount, &w.ctr->stashTotal, w.manCap, w.scanBlock,
                              w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kStashSplit, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody, w.sleepId, w.flags,
                       w.flagScan, w.stash, w.stashIsland, w.manCap, w.slots);
    hipLaunchKernelGGL(kStashSplitEnd, dim3(1), dim3(64), 0, st, w.ctr, w.manCap);
    hipLaunchKernelGGL(kKeptBack, dim3(gm), dim3(256), 0, st, w.ctr, w.slots, w.man[w.cur]);
    return hipGetLastError();
}

}  // namespace agx
