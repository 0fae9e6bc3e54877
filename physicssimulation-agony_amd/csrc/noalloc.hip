// agx_broadphase_pairs_noalloc: the standalone broad phase over device buffers with every
// intermediate array (collider lists, per-scene ranges and bounds, both trees with their sort
// scratch, pair staging, counters) carved from one file-scope device array, so a call allocates
// nothing. Limits: AGX_NOALLOC_MAX_COLLIDERS colliders, AGX_NOALLOC_MAX_SCENES scenes,
// AGX_NOALLOC_MAX_PAIRS pairs (include/agx.h).
#include "agx_internal.h"

#include <algorithm>
#include <cstdio>
#include <mutex>

namespace {

constexpr size_t ALIGN = 256;
constexpr size_t alignUp(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// Bytes of each array, in the order they are carved. The same walk sizes the device array at
// compile time (for the largest inputs) and places the arrays at run time.
struct Layout {
    size_t off = 0;
    template <class T>
    constexpr size_t take(size_t count) {
        const size_t at = alignUp(off);
        off = at + (count ? count : 1) * sizeof(T);
        return at;
    }
};

struct TreeOffsets {
    size_t keys, vals, keysAlt, valsAlt, leafBody, nodeAabb, child, range, parent, rec, hist;
};

using agx::sortTileFor;
constexpr uint32_t sortBlocksFor(uint32_t n) {
    return std::max(1u, (n + sortTileFor(n) - 1) / sortTileFor(n));
}

constexpr TreeOffsets carveTree(Layout& L, uint32_t n) {
    TreeOffsets t{};
    const size_t nodes = n ? 2ull * n - 1 : 1, internal = n > 1 ? n - 1 : 1;
    t.keys = L.take<uint32_t>(n);
    t.vals = L.take<uint32_t>(n);
    t.keysAlt = L.take<uint32_t>(n);
    t.valsAlt = L.take<uint32_t>(n);
    t.leafBody = L.take<uint32_t>(n);
    t.nodeAabb = L.take<float4>(nodes);
    t.child = L.take<uint2>(internal);
    t.range = L.take<uint2>(internal);
    t.parent = L.take<uint32_t>(nodes);
    t.rec = L.take<agx::NodeRecord>(internal);
    t.hist = L.take<uint32_t>(4ull * (sortBlocksFor(n) + 1) * 256 + 4);
    return t;
}

constexpr uint32_t SCAN_BLOCK_CAP = 4096;

struct Offsets {
    // set-up (sized by n and S)
    size_t sceneZero, dynList, statList, dynStart, statStart, dynBounds, statBounds, header;
    // trees, pairs and counters (sized by the dynamic and static counts)
    TreeOffsets dyn, stat;
    size_t pairs, waveCount, waveOffset, staged, scanBlock, ctr;
    size_t total;
};

constexpr Offsets carve(uint32_t n, uint32_t nDyn, uint32_t S) {
    Layout L;
    Offsets o{};
    o.sceneZero = L.take<uint32_t>(n);
    o.dynList = L.take<uint32_t>(n);
    o.statList = L.take<uint32_t>(n);
    o.dynStart = L.take<uint32_t>(S + 1);
    o.statStart = L.take<uint32_t>(S + 1);
    o.dynBounds = L.take<int4>(S);
    o.statBounds = L.take<int4>(S);
    o.header = L.take<uint32_t>(4);
    o.dyn = carveTree(L, nDyn);
    o.stat = carveTree(L, n - nDyn);
    const uint32_t waves = (nDyn + 63) / 64;
    o.pairs = L.take<uint2>(AGX_NOALLOC_MAX_PAIRS);
    o.waveCount = L.take<uint32_t>(2ull * std::max(1u, waves));
    o.waveOffset = L.take<uint32_t>(2ull * std::max(1u, waves));
    o.staged = L.take<uint2>(2ull * std::max(1u, waves) * agx::STAGE_PAIRS);
    o.scanBlock = L.take<uint32_t>(SCAN_BLOCK_CAP);
    o.ctr = L.take<agx::Counters>(1);
    o.total = alignUp(L.off);
    return o;
}

// Either tree's arrays grow linearly with its collider count plus per-array rounding, so the
// all-dynamic split bounds every split of the largest input; the slack covers the rounding of
// the second tree's arrays.
constexpr size_t SCRATCH_BYTES =
    carve(AGX_NOALLOC_MAX_COLLIDERS, AGX_NOALLOC_MAX_COLLIDERS, AGX_NOALLOC_MAX_SCENES).total + 64 * ALIGN;

// The device array every call on a device works in (one copy per device, held until the process
// ends). Host code uses its per-device address from hipGetSymbolAddress, never this name.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;                              // calls of this process run one at a time
unsigned char* g_scratch[MAX_DEVICES] = {};   // device address of `scratch` per device

int failWith(int code, const char* fmt, unsigned long long a = 0, unsigned long long b = 0) {
    char buf[256];
    snprintf(buf, sizeof buf, fmt, a, b);
    agx::setLastError(buf);
    return code;
}

#define NA_CHECK(call)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            char m_[256];                                                                      \
            snprintf(m_, sizeof m_, "%s: %s", #call, hipGetErrorString(e_));                   \
            agx::setLastError(m_);                                                             \
            return AGX_ERR_HIP;                                                                \
        }                                                                                      \
    } while (0)

void bindTree(agx::Tree& t, unsigned char* base, const TreeOffsets& o, uint32_t n, uint32_t* list,
              uint32_t* sceneStart, int4* bounds) {
    t.n = n;
    t.sortTile = sortTileFor(n);
    t.sortBlocks = sortBlocksFor(n);
    t.keys = reinterpret_cast<uint32_t*>(base + o.keys);
    t.vals = reinterpret_cast<uint32_t*>(base + o.vals);
    t.keysAlt = reinterpret_cast<uint32_t*>(base + o.keysAlt);
    t.valsAlt = reinterpret_cast<uint32_t*>(base + o.valsAlt);
    t.leafBody = reinterpret_cast<uint32_t*>(base + o.leafBody);
    t.nodeAabb = reinterpret_cast<float4*>(base + o.nodeAabb);
    t.child = reinterpret_cast<uint2*>(base + o.child);
    t.range = reinterpret_cast<uint2*>(base + o.range);
    t.parent = reinterpret_cast<uint32_t*>(base + o.parent);
    t.rec = reinterpret_cast<agx::NodeRecord*>(base + o.rec);
    t.hist = reinterpret_cast<uint32_t*>(base + o.hist);
    t.list = list;
    t.sceneStart = sceneStart;
    t.sceneBounds = reinterpret_cast<float4*>(bounds);
    t.boundsReady = true;
    t.chunks = 0;
}

uint32_t ceilLog2(uint32_t v) {
    uint32_t b = 0;
    while ((1u << b) < v) b++;
    return b;
}

}  // namespace

extern "C" {

int agx_broadphase_pairs_noalloc(const float* aabb, const uint32_t* owner, const uint8_t* is_dynamic,
                                 const uint32_t* scene, uint32_t n, uint32_t num_scenes, uint32_t* pairs_a,
                                 uint32_t* pairs_b, uint64_t cap, uint64_t* count) {
    if (!count) return failWith(AGX_ERR_INVALID_ARGUMENT, "count is null");
    *count = 0;
    if (n > AGX_NOALLOC_MAX_COLLIDERS)
        return failWith(AGX_ERR_CAPACITY, "%llu colliders exceed the allocation-free maximum of %llu", n,
                        AGX_NOALLOC_MAX_COLLIDERS);
    if (num_scenes == 0 || num_scenes > AGX_NOALLOC_MAX_SCENES)
        return failWith(AGX_ERR_INVALID_ARGUMENT, "num_scenes %llu outside [1, %llu]", num_scenes, AGX_NOALLOC_MAX_SCENES);
    if (n == 0) return AGX_OK;
    if (!aabb || !owner || !is_dynamic || (cap && (!pairs_a || !pairs_b)))
        return failWith(AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (reinterpret_cast<uintptr_t>(aabb) % 16 != 0)
        return failWith(AGX_ERR_INVALID_ARGUMENT, "aabb must be 16-byte aligned");
    std::lock_guard<std::mutex> lock(g_mu);
    int dev = 0;
    NA_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return failWith(AGX_ERR_NO_DEVICE, "device %llu out of range", (unsigned)dev);
    if (!g_scratch[dev]) {   // first call on this device: the device's own copy of the array
        void* p = nullptr;
        NA_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    unsigned char* base = g_scratch[dev];
    const hipStream_t st = nullptr;   // the null stream: no stream is created
    const uint32_t S = num_scenes;

    // set-up on the device; the layout of its outputs does not depend on the split
    const Offsets pre = carve(n, 0, S);
    agx::BroadPrep prep{};
    prep.dynList = reinterpret_cast<uint32_t*>(base + pre.dynList);
    prep.statList = reinterpret_cast<uint32_t*>(base + pre.statList);
    prep.dynStart = reinterpret_cast<uint32_t*>(base + pre.dynStart);
    prep.statStart = reinterpret_cast<uint32_t*>(base + pre.statStart);
    prep.dynBounds = reinterpret_cast<int4*>(base + pre.dynBounds);
    prep.statBounds = reinterpret_cast<int4*>(base + pre.statBounds);
    prep.header = reinterpret_cast<uint32_t*>(base + pre.header);
    prep.sceneZero = scene ? nullptr : reinterpret_cast<uint32_t*>(base + pre.sceneZero);
    const float4* box = reinterpret_cast<const float4*>(aabb);
    NA_CHECK(agx::launchBroadPrep(st, box, is_dynamic, scene, n, S, prep));
    uint32_t header[4] = {};
    NA_CHECK(hipMemcpyAsync(header, prep.header, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    NA_CHECK(hipStreamSynchronize(st));
    if (header[2]) return failWith(AGX_ERR_INVALID_ARGUMENT, "a scene index is >= num_scenes (%llu)", S);
    const uint32_t nDyn = header[0], nStat = header[1];
    if (nDyn + nStat != n) return failWith(AGX_ERR_HIP, "set-up counted %llu of %llu colliders", nDyn + nStat, n);

    const Offsets o = carve(n, nDyn, S);
    if (o.total > SCRATCH_BYTES) return failWith(AGX_ERR_CAPACITY, "layout of %llu bytes exceeds %llu", o.total, SCRATCH_BYTES);
    agx::DeviceWorld d{};
    d.S = S;
    d.sceneBits = ceilLog2(S);
    d.mortonBits = std::min(30u, (32u - d.sceneBits) & ~1u);
    d.C = n;
    d.nDyn = nDyn;
    d.nStat = nStat;
    d.aabb = const_cast<float4*>(box);
    d.colBody = const_cast<uint32_t*>(owner);
    d.colScene = scene ? const_cast<uint32_t*>(scene) : prep.sceneZero;
    bindTree(d.dyn, base, o.dyn, nDyn, prep.dynList, prep.dynStart, prep.dynBounds);
    bindTree(d.stat, base, o.stat, nStat, prep.statList, prep.statStart, prep.statBounds);
    d.pairCap = AGX_NOALLOC_MAX_PAIRS;
    d.pairs = reinterpret_cast<uint2*>(base + o.pairs);
    d.queryWaves = (nDyn + 63) / 64;
    d.waveCount = reinterpret_cast<uint32_t*>(base + o.waveCount);
    d.waveOffset = reinterpret_cast<uint32_t*>(base + o.waveOffset);
    d.stagedPairs = reinterpret_cast<uint2*>(base + o.staged);
    d.scanBlock = reinterpret_cast<uint32_t*>(base + o.scanBlock);
    d.scanBlockCap = SCAN_BLOCK_CAP;
    d.ctr = reinterpret_cast<agx::Counters*>(base + o.ctr);
    // the last call's counters are still there: clear them with a kernel
    NA_CHECK(agx::launchFillU32(st, reinterpret_cast<uint32_t*>(d.ctr), 0u, sizeof(agx::Counters) / 4));
    NA_CHECK(agx::launchBuildTree(st, d, d.stat));
    NA_CHECK(agx::launchBuildTree(st, d, d.dyn));
    NA_CHECK(agx::launchPairs(st, d));
    uint32_t found[2] = {};   // pairs, overflow bits
    static_assert(offsetof(agx::Counters, overflow) - offsetof(agx::Counters, pairs) == 7 * sizeof(uint32_t),
                  "counter layout");
    uint32_t counters[8] = {};
    NA_CHECK(hipMemcpyAsync(counters, &d.ctr->pairs, sizeof counters, hipMemcpyDeviceToHost, st));
    NA_CHECK(hipStreamSynchronize(st));
    found[0] = counters[0];
    found[1] = counters[7];
    *count = found[0];
    if (found[1] & 2u) return failWith(AGX_ERR_CAPACITY, "traversal stack overflow");
    if ((uint64_t)found[0] > cap) return failWith(AGX_ERR_CAPACITY, "%llu pairs exceed cap %llu", found[0], cap);
    if (found[0] > AGX_NOALLOC_MAX_PAIRS)
        return failWith(AGX_ERR_CAPACITY, "%llu pairs exceed the allocation-free maximum of %llu", found[0],
                        AGX_NOALLOC_MAX_PAIRS);
    NA_CHECK(agx::launchSplitPairs(st, d.pairs, found[0], pairs_a, pairs_b));
    NA_CHECK(hipStreamSynchronize(st));
    return AGX_OK;
}

}  // extern "C"
