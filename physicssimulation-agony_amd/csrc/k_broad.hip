// Broad phase: per-scene bounds, Morton keys, LDS radix sort, two-level Karras LBVH, refit, pair traversal.
//
// The pair set equals the reference's (Broad/BroadPhaseCollisionDetector.cpp:109-244): node
// AABBs are exact unions, so pruning with the strict overlap test never drops a strictly
// overlapping leaf pair and the set does not depend on tree topology.
#include "agx_internal.h"

#include <cfloat>

namespace agx {

namespace {

// Order-preserving float <-> int mapping for atomic min/max on floats.
__device__ inline int floatOrder(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ inline float orderFloat(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void kBoundsInit(int4* b, uint32_t S) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) b[s] = make_int4(floatOrder(FLT_MAX), floatOrder(FLT_MAX), floatOrder(-FLT_MAX), floatOrder(-FLT_MAX));
}

// Bounds of each scene's AABBs (BvhTree.cpp:148-188), one block per chunk of one scene.
__global__ __launch_bounds__(256) void kBounds(const uint32_t* __restrict__ list, const float4* __restrict__ aabb,
                                               const uint32_t* __restrict__ chunkScene, const uint2* __restrict__ chunkRange,
                                               int4* __restrict__ bounds) {
    const uint2 r = chunkRange[blockIdx.x];
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    for (uint32_t k = r.x + threadIdx.x; k < r.y; k += 256) {
        const float4 a = aabb[list[k]];
        mnx = fminf(mnx, a.x); mny = fminf(mny, a.y); mxx = fmaxf(mxx, a.z); mxy = fmaxf(mxy, a.w);
    }
    for (uint32_t off = 32; off > 0; off >>= 1) {
        mnx = fminf(mnx, __shfl_down(mnx, off, 64)); mny = fminf(mny, __shfl_down(mny, off, 64));
        mxx = fmaxf(mxx, __shfl_down(mxx, off, 64)); mxy = fmaxf(mxy, __shfl_down(mxy, off, 64));
    }
    __shared__ float4 ws[4];
    if ((threadIdx.x & 63u) == 0) ws[threadIdx.x >> 6] = make_float4(mnx, mny, mxx, mxy);
    __syncthreads();
    if (threadIdx.x == 0) {
        float4 t = ws[0];
        for (int k = 1; k < 4; k++) {
            t.x = fminf(t.x, ws[k].x); t.y = fminf(t.y, ws[k].y); t.z = fmaxf(t.z, ws[k].z); t.w = fmaxf(t.w, ws[k].w);
        }
        int4* b = bounds + chunkScene[blockIdx.x];
        atomicMin(&b->x, floatOrder(t.x));
        atomicMin(&b->y, floatOrder(t.y));
        atomicMax(&b->z, floatOrder(t.z));
        atomicMax(&b->w, floatOrder(t.w));
    }
}

__device__ inline uint32_t part1By1(uint32_t x) {   // spreads the low 16 bits to even positions
    x &= 0x0000ffffu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// Morton key of the AABB centre quantised to the scene bounds (BvhTree.cpp:191-280, with
// mortonBits/2 bits per axis), scene id above the Morton bits.
__global__ __launch_bounds__(256) void kMorton(const uint32_t* __restrict__ list, uint32_t n, const float4* __restrict__ aabb,
                                               const uint32_t* __restrict__ colScene, const int4* __restrict__ bounds,
                                               uint32_t mortonBits, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t c = list[k];
    const uint32_t s = colScene[c];
    const int4 bi = bounds[s];
    const float bMinX = orderFloat(bi.x), bMinY = orderFloat(bi.y);
    const float bMaxX = fmaxf(orderFloat(bi.z), bMinX + 1e-5f), bMaxY = fmaxf(orderFloat(bi.w), bMinY + 1e-5f);
    const uint32_t axisBits = mortonBits >> 1;
    const float qMax = (float)((1u << axisBits) - 1u);
    const float scaleX = qMax / (bMaxX - bMinX), scaleY = qMax / (bMaxY - bMinY);
    const float4 a = aabb[c];
    const float tx = fminf(fmaxf((a.x + a.z) * (scaleX * 0.5f) + (-bMinX * scaleX), 0.0f), qMax);
    const float ty = fminf(fmaxf((a.y + a.w) * (scaleY * 0.5f) + (-bMinY * scaleY), 0.0f), qMax);
    const uint32_t code = (part1By1((uint32_t)ty) << 1) | part1By1((uint32_t)tx);
    keys[k] = mortonBits >= 32 ? code : ((s << mortonBits) | code);
    vals[k] = c;
}

// ---------------------------------------------------------------- radix sort (8-bit digits)
// One histogram pass for every digit, one scan of the 4 x 256 digit totals, then one stable
// scatter per digit. A scatter tile finds the keys of its digit in earlier tiles by decoupled
// look-back: it publishes its own count per digit at once, then walks back over earlier tiles'
// published counts until it meets a tile that has published its inclusive prefix. Tiles take
// their index from a ticket counter, so a tile only ever waits for tiles that are running.
// Layout of the sort scratch (cleared once per sort): tile status [passes][tiles][256], digit
// totals [passes][256], tickets [passes].
constexpr uint32_t SORT_PASSES_MAX = 4;
constexpr uint32_t STATUS_AGG = 1u << 30, STATUS_INC = 2u << 30, STATUS_VALUE = (1u << 30) - 1;
constexpr uint64_t SORT_STALL_TICKS = 200000000ull;   // 2 s of the 100 MHz wall clock

__global__ __launch_bounds__(256) void kRadixHistAll(const uint32_t* __restrict__ keys, uint32_t n, uint32_t passes,
                                                     uint32_t tileKeys, uint32_t* __restrict__ digitTotal) {
    __shared__ uint32_t h[SORT_PASSES_MAX][256];
    for (uint32_t p = 0; p < SORT_PASSES_MAX; p++) h[p][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * tileKeys;
    for (uint32_t k = threadIdx.x; k < tileKeys; k += 256) {
        const uint32_t i = base + k;
        if (i >= n) break;
        const uint32_t key = keys[i];
        for (uint32_t p = 0; p < passes; p++) atomicAdd(&h[p][(key >> (8 * p)) & 0xFFu], 1u);
    }
    __syncthreads();
    for (uint32_t p = 0; p < passes; p++)
        if (h[p][threadIdx.x]) atomicAdd(&digitTotal[p * 256 + threadIdx.x], h[p][threadIdx.x]);
}

// One block: exclusive scan of each pass's 256 digit totals, in place.
__global__ __launch_bounds__(256) void kRadixDigitScan(uint32_t* __restrict__ digitTotal, uint32_t passes) {
    __shared__ uint32_t waveSum[4];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    for (uint32_t p = 0; p < passes; p++) {
        const uint32_t v = digitTotal[p * 256 + threadIdx.x];
        uint32_t x = v;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) waveSum[wid] = x;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t k = 0; k < wid; k++) before += waveSum[k];
        digitTotal[p * 256 + threadIdx.x] = before + x - v;
        __syncthreads();
    }
}

// Stable scatter of one digit. A tile of THREADS x ROUNDS keys: each wave ranks ROUNDS rounds of
// 64 consecutive keys with ballot matching (8 ballots per round) and per-wave digit counters in
// LDS. The tile is then put in digit order in LDS, so that it leaves as contiguous runs, one per
// digit, each at its digit's start + the keys of that digit in earlier tiles. Earlier tiles'
// counts come by decoupled look-back (threads 0..255, one digit each): a tile publishes its own
// counts at once and reads back to the nearest tile that has published its inclusive prefix.
template <uint32_t THREADS, uint32_t ROUNDS>
__global__ __launch_bounds__(THREADS) void kRadixScatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                         uint32_t n, uint32_t shift, const uint32_t* __restrict__ digitStart,
                                                         uint32_t* status, uint32_t* ticket, uint32_t* overflow,
                                                         uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut) {
    constexpr uint32_t WAVES = THREADS / 64, TILE = THREADS * ROUNDS;
    static_assert(THREADS >= 256, "one look-back thread per digit");
    static_assert(TILE <= 65536, "rank and digit share a register");
    __shared__ uint32_t cnt[WAVES][256];   // per-wave digit counts, then each wave's offset in its digit's run
    __shared__ uint32_t localStart[256];   // digit runs in the tile's digit order
    __shared__ uint32_t outStart[256];     // digit runs in the output
    __shared__ uint32_t tileKeys[TILE], tileVals[TILE];
    __shared__ uint32_t tileShared;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) tileShared = atomicAdd(ticket, 1u);
    for (uint32_t k = threadIdx.x; k < WAVES * 256; k += THREADS) (&cnt[0][0])[k] = 0;
    __syncthreads();
    const uint32_t tile = tileShared;
    const uint32_t tileFirst = tile * TILE, tileCount = min(TILE, n - tileFirst);
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t rd[ROUNDS], ky[ROUNDS], vl[ROUNDS];   // rd: rank in the wave's digit run | digit << 16
    const uint32_t waveBase = tileFirst + wid * (TILE / WAVES);
    // every load of the tile in flight before the ranking starts (the ranking's LDS updates would
    // otherwise hold each round's load back behind the previous round)
    for (uint32_t r = 0; r < ROUNDS; r++) {
        const uint32_t i = waveBase + r * 64 + lane;
        ky[r] = i < n ? keys[i] : 0u;
        vl[r] = i < n ? vals[i] : 0u;
    }
    for (uint32_t r = 0; r < ROUNDS; r++) {
        const bool valid = waveBase + r * 64 + lane < n;
        const uint32_t d = (ky[r] >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
        for (uint32_t bit = 0; bit < 8; bit++) {
            const bool set = (d >> bit) & 1u;
            const uint64_t m = __ballot(set);
            peers &= set ? m : ~m;
        }
        uint32_t before = 0;
        if (valid) before = cnt[wid][d];
        const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)peers) - 1);
        if (valid && lane == leader) cnt[wid][d] = before + (uint32_t)__popcll(peers);
        rd[r] = (before + (uint32_t)__popcll(peers & ltMask)) | (d << 16);
    }
    __syncthreads();
    uint32_t mine = 0;   // threads 0..255: this tile's keys of digit threadIdx.x
    if (threadIdx.x < 256) {
        const uint32_t d = threadIdx.x;
        for (uint32_t w = 0; w < WAVES; w++) {
            const uint32_t c = cnt[w][d];
            cnt[w][d] = mine;
            mine += c;
        }
        storeFlag(status + tile * 256 + d, (tile == 0 ? STATUS_INC : STATUS_AGG) | mine);
        localStart[d] = mine;
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // exclusive scan of the 256 digit counts, four per lane
        uint32_t v[4], sum = 0;
        for (uint32_t k = 0; k < 4; k++) { v[k] = localStart[4 * lane + k]; sum += v[k]; }
        uint32_t inc = sum;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off, 64);
            if (lane >= off) inc += y;
        }
        uint32_t run = inc - sum;
        for (uint32_t k = 0; k < 4; k++) { localStart[4 * lane + k] = run; run += v[k]; }
    }
    __syncthreads();
    for (uint32_t r = 0; r < ROUNDS; r++) {
        if (waveBase + r * 64 + lane < n) {
            const uint32_t d = rd[r] >> 16;
            const uint32_t at = localStart[d] + cnt[wid][d] + (rd[r] & 0xFFFFu);
            tileKeys[at] = ky[r];
            tileVals[at] = vl[r];
        }
    }
    if (threadIdx.x < 256) {
        const uint32_t d = threadIdx.x;
        uint32_t prefix = 0;
        if (tile != 0) {
            // look back LOOKBACK tiles at a time (all loads in flight together)
            constexpr int32_t LOOKBACK = 16;
            const uint64_t deadline = wall_clock64() + SORT_STALL_TICKS;
            int32_t j = (int32_t)tile - 1;
            while (j >= 0) {
                uint32_t stv[LOOKBACK];
                for (int32_t k = 0; k < LOOKBACK; k++) stv[k] = j - k >= 0 ? loadFlag(status + (j - k) * 256 + d) : 0u;
                int32_t k = 0;
                bool found = false;
                for (; k < LOOKBACK && j - k >= 0; k++) {
                    const uint32_t st = stv[k];
                    if (st & STATUS_INC) { prefix += st & STATUS_VALUE; found = true; break; }
                    if (!(st & STATUS_AGG)) break;   // not published yet: look again from here
                    prefix += st & STATUS_VALUE;
                }
                if (found) break;
                j -= k;
                if (k == 0) {
                    if (wall_clock64() > deadline) { atomicOr(overflow, OVERFLOW_STALL); break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            storeFlag(status + tile * 256 + d, STATUS_INC | (prefix + mine));
        }
        outStart[d] = digitStart[d] + prefix - localStart[d];
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < tileCount; k += THREADS) {
        const uint32_t key = tileKeys[k];
        const uint32_t dst = outStart[(key >> shift) & 0xFFu] + k;
        keysOut[dst] = key;
        valsOut[dst] = tileVals[k];
    }
}

// ---------------------------------------------------------------- two-level LBVH
// The sorted leaves are cut into tiles of BVH_TILE consecutive leaves. Each tile gets its own
// Karras 2012 tree over its keys, and a Karras tree over the tiles' first keys joins the tile
// roots. Every node still covers a contiguous range of sorted leaves, so the traversals and the
// pair set are those of a single tree; what the cut buys is that each tile's nodes live in one
// block's LDS during the refit and only the top tree (one node per tile) is left above them.
// Node ids: top internal nodes [0, T-1) (node 0 is the root when T > 1), then each tile's
// internal nodes in tile order, tile t from (T-1) + t*(BVH_TILE-1); leaves [n-1, 2n-1).
constexpr uint32_t BVH_TILE = 1024;
constexpr uint32_t BVH_TOP_THREADS = 1024;
constexpr uint32_t BVH_TOP_MAX_TILES = 16384;   // LDS arrival counters of the top refit (64 KiB)

__device__ inline uint32_t tileBase(uint32_t tiles, uint32_t t) { return (tiles - 1) + t * (BVH_TILE - 1); }

// Karras split of internal node i over count sorted keys key(0..count-1), ties broken by index:
// its first and last key and the split (the last key of its left half).
template <class Key>
__device__ inline void karrasNode(int i, int count, const Key& key, int& lo, int& hi, int& gamma) {
    auto delta = [&](int a, int b) {
        if (b < 0 || b >= count) return -1;
        const uint32_t ka = key(a), kb = key(b);
        return ka == kb ? 32 + __clz((int)((uint32_t)a ^ (uint32_t)b)) : __clz((int)(ka ^ kb));
    };
    const int d = delta(i, i + 1) > delta(i, i - 1) ? 1 : -1;
    const int dMin = delta(i, i - d);
    int lMax = 2;
    while (delta(i, i + lMax * d) > dMin) lMax <<= 1;
    int l = 0;
    for (int t = lMax >> 1; t >= 1; t >>= 1)
        if (delta(i, i + (l + t) * d) > dMin) l += t;
    const int j = i + l * d;
    const int dNode = delta(i, j);
    int s = 0;
    int t = l;
    while (t > 1) {
        t = (t + 1) >> 1;
        if (delta(i, i + (s + t) * d) > dNode) s += t;
    }
    gamma = i + s * d + min(d, 0);
    lo = min(i, j);
    hi = max(i, j);
}

// One thread per sorted leaf; leaf x of a tile with m leaves is internal node x of the tile's tree
// for x < m - 1. With one tile its root is the tree's root.
__global__ __launch_bounds__(256) void kKarrasTiles(const uint32_t* __restrict__ keys, uint32_t n, uint2* __restrict__ child,
                                                    uint2* __restrict__ range, uint32_t* __restrict__ parent) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t tiles = (n + BVH_TILE - 1) / BVH_TILE;
    const uint32_t t = i / BVH_TILE, t0 = t * BVH_TILE, x = i - t0;
    const uint32_t m = min(BVH_TILE, n - t0);
    if (x + 1 >= m) return;
    const uint32_t* tk = keys + t0;
    int lo, hi, gamma;
    karrasNode((int)x, (int)m, [&](int k) { return tk[k]; }, lo, hi, gamma);
    const uint32_t base = tileBase(tiles, t), leafBase = n - 1;
    const uint32_t left = lo == gamma ? leafBase + t0 + (uint32_t)gamma : base + (uint32_t)gamma;
    const uint32_t right = hi == gamma + 1 ? leafBase + t0 + (uint32_t)(gamma + 1) : base + (uint32_t)(gamma + 1);
    const uint32_t id = base + x;
    child[id] = make_uint2(left, right);
    range[id] = make_uint2(t0 + (uint32_t)lo, t0 + (uint32_t)hi);
    parent[left] = id;
    parent[right] = id;
    if (tiles == 1 && x == 0) parent[id] = 0xFFFFFFFFu;
}

// The root of tile t: its tree's node 0, or its only leaf.
__device__ inline uint32_t tileRoot(uint32_t n, uint32_t tiles, uint32_t t) {
    const uint32_t t0 = t * BVH_TILE;
    return min(BVH_TILE, n - t0) > 1 ? tileBase(tiles, t) : n - 1 + t0;
}

// One thread per top internal node (tiles - 1 of them, tiles >= 2).
__global__ __launch_bounds__(256) void kKarrasTop(const uint32_t* __restrict__ keys, uint32_t n, uint2* __restrict__ child,
                                                  uint2* __restrict__ range, uint32_t* __restrict__ parent) {
    const uint32_t tiles = (n + BVH_TILE - 1) / BVH_TILE;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k + 1 >= tiles) return;
    int lo, hi, gamma;
    karrasNode((int)k, (int)tiles, [&](int t) { return keys[(size_t)t * BVH_TILE]; }, lo, hi, gamma);
    const uint32_t left = lo == gamma ? tileRoot(n, tiles, (uint32_t)gamma) : (uint32_t)gamma;
    const uint32_t right = hi == gamma + 1 ? tileRoot(n, tiles, (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    child[k] = make_uint2(left, right);
    range[k] = make_uint2((uint32_t)lo * BVH_TILE, min(((uint32_t)hi + 1) * BVH_TILE, n) - 1);
    parent[left] = k;
    parent[right] = k;
    if (k == 0) parent[0] = 0xFFFFFFFFu;
}

// Bottom-up refit (BvhTree.cpp:426-530 computes the same unions top-down in reverse order):
// the second walk to reach a node merges its children and writes the node's record (both
// children's boxes, ranges and references).
__device__ inline float4 boxUnion(const float4& a, const float4& b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

// One block per tile: the tile's leaves (box, reference, body) and its internal nodes' parents,
// ranges and children are loaded into LDS with coalesced loads, then every leaf walks up its
// tile's tree with LDS arrival counters, without touching memory except to write the results.
__global__ __launch_bounds__(BVH_TILE) void kRefitTiles(uint32_t n, const uint2* __restrict__ child,
                                                        const uint2* __restrict__ range, uint32_t* __restrict__ parent,
                                                        const uint32_t* __restrict__ vals, const float4* __restrict__ aabb,
                                                        const uint32_t* __restrict__ colBody, uint32_t* __restrict__ leafBody,
                                                        float4* __restrict__ nodeAabb, NodeRecord* __restrict__ rec) {
    __shared__ float4 leafBox[BVH_TILE], innerBox[BVH_TILE];
    __shared__ uint2 innerRange[BVH_TILE], innerChild[BVH_TILE];
    __shared__ uint32_t innerParent[BVH_TILE], leafParent[BVH_TILE], leafRef[BVH_TILE], leafOwner[BVH_TILE];
    __shared__ uint32_t arrived[BVH_TILE];
    const uint32_t tiles = (n + BVH_TILE - 1) / BVH_TILE;
    const uint32_t x = threadIdx.x, t0 = blockIdx.x * BVH_TILE, j = t0 + x;
    const uint32_t m = min(BVH_TILE, n - t0);
    const uint32_t leafBase = n - 1, base = tileBase(tiles, blockIdx.x);
    arrived[x] = 0;
    if (x < m) {
        const uint32_t c = vals[j];
        const uint32_t b = colBody[c];
        const float4 box = aabb[c];
        leafRef[x] = c;
        leafOwner[x] = b;
        leafBox[x] = box;
        leafBody[j] = b;
        nodeAabb[leafBase + j] = box;
        leafParent[x] = n > 1 ? parent[leafBase + j] : 0xFFFFFFFFu;
    }
    if (x + 1 < m) {
        innerParent[x] = parent[base + x];
        innerRange[x] = range[base + x];
        innerChild[x] = child[base + x];
    }
    if (n == 1 && x == 0) parent[0] = 0xFFFFFFFFu;
    __syncthreads();
    if (x >= m) return;
    auto info = [&](uint32_t c) {
        ChildInfo ci;
        if (c >= leafBase) {
            const uint32_t k = c - leafBase - t0;
            ci.box = leafBox[k];
            ci.first = ci.last = c - leafBase;
            ci.ref = leafRef[k];
            ci.body = leafOwner[k];
        } else {
            const uint2 r = innerRange[c - base];
            ci.box = innerBox[c - base];
            ci.first = r.x; ci.last = r.y;
            ci.ref = c;
            ci.body = 0xFFFFFFFFu;
        }
        return ci;
    };
    uint32_t p = leafParent[x];
    while (p - base < m - 1) {   // p is an internal node of this tile (ids below base wrap around)
        __threadfence_block();
        if (atomicAdd(&arrived[p - base], 1u) == 0) break;
        __threadfence_block();
        const uint2 ch = innerChild[p - base];
        NodeRecord r;
        r.c[0] = info(ch.x);
        r.c[1] = info(ch.y);
        const float4 box = boxUnion(r.c[0].box, r.c[1].box);
        innerBox[p - base] = box;
        nodeAabb[p] = box;
        rec[p] = r;
        p = innerParent[p - base];
    }
}

__device__ inline ChildInfo childInfo(uint32_t c, uint32_t leafBase, const float4& box, const uint2* range,
                                      const uint32_t* vals, const uint32_t* colBody) {
    ChildInfo ci;
    ci.box = box;
    if (c >= leafBase) {
        const uint32_t j = c - leafBase;
        ci.first = ci.last = j;
        ci.ref = vals[j];
        ci.body = colBody[ci.ref];
    } else {
        const uint2 r = range[c];
        ci.first = r.x; ci.last = r.y;
        ci.ref = c;
        ci.body = 0xFFFFFFFFu;
    }
    return ci;
}

// One workgroup refits the top tree from the tile roots (tiles >= 2), with arrival counters in
// LDS; the boxes its walks hand over go through device memory, ordered by the workgroup fences
// around each counter bump (all its waves run on one CU).
__global__ __launch_bounds__(BVH_TOP_THREADS) void kRefitTop(uint32_t n, const uint2* __restrict__ child,
                                                             const uint2* __restrict__ range,
                                                             const uint32_t* __restrict__ parent,
                                                             const uint32_t* __restrict__ vals,
                                                             const uint32_t* __restrict__ colBody, float4* nodeAabb,
                                                             NodeRecord* __restrict__ rec) {
    __shared__ uint32_t arrived[BVH_TOP_MAX_TILES];
    const uint32_t tiles = (n + BVH_TILE - 1) / BVH_TILE;
    for (uint32_t k = threadIdx.x; k + 1 < tiles; k += blockDim.x) arrived[k] = 0;
    __syncthreads();
    const uint32_t leafBase = n - 1;
    for (uint32_t t = threadIdx.x; t < tiles; t += blockDim.x) {
        uint32_t p = parent[tileRoot(n, tiles, t)];
        while (p != 0xFFFFFFFFu) {
            const uint2 ch = child[p];
            const uint32_t up = parent[p];
            __threadfence_block();
            if (atomicAdd(&arrived[p], 1u) == 0) break;
            __threadfence_block();
            const float4 a = nodeAabb[ch.x], b = nodeAabb[ch.y];
            NodeRecord r;
            r.c[0] = childInfo(ch.x, leafBase, a, range, vals, colBody);
            r.c[1] = childInfo(ch.y, leafBase, b, range, vals, colBody);
            nodeAabb[p] = boxUnion(a, b);
            rec[p] = r;
            p = up;
        }
    }
}

// ---------------------------------------------------------------- pair traversal
// One query per lane, one wave-uniform stack of internal nodes in LDS (packet traversal):
// queries are consecutive Morton-sorted dynamic leaves, so a wave's paths overlap. A step loads
// one node record (both children) at a wave-uniform address; a child is visited when any
// lane's strict overlap test passes (ballot), and leaf children emit pairs for the passing
// lanes. SELF restricts a query to leaves after it (each pair once) and skips pairs of one body.
// STAGED: pairs go to the wave's staging area (up to STAGE_PAIRS) and the wave's total to
// waveCount; the gather then moves them to their scanned place. Otherwise (a second pass run
// only by waves whose total exceeded the staging area) pairs go straight to the scanned offsets.
// Both passes emit in the same deterministic order.
struct Query { float4 box; uint32_t leaf, col, body, sf, sl; bool valid; };

template <bool SELF>
__device__ inline bool childActive(const Query& q, const float4& box, uint32_t first, uint32_t last) {
    if (!q.valid) return false;
    if (first >= q.sl || last < q.sf) return false;
    if (SELF && last <= q.leaf) return false;
    return q.box.x < box.z && q.box.z > box.x && q.box.y < box.w && q.box.w > box.y;
}

// Query wave w writes its pairs with the query tree itself (part 0) and with the other tree
// (part 1) to region 2w + part, so the pair list runs wave by wave in query (scene-major Morton)
// order: the pairs, and the manifolds compacted from them, are grouped by scene except inside a
// wave whose queries straddle two scenes.
template <bool SELF, bool STAGED>
__global__ __launch_bounds__(256) void kTraverse(Tree qt, Tree tt, const uint32_t* __restrict__ colScene,
                                                 uint32_t* __restrict__ waveCount, const uint32_t* __restrict__ waveOffset,
                                                 uint32_t part, uint2* __restrict__ pairs, uint64_t cap,
                                                 uint2* __restrict__ staged, Counters* ctr,
                                                 const uint32_t* __restrict__ sleepId) {
    __shared__ uint32_t stackMem[4][TRAVERSAL_STACK];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t* stack = stackMem[wid];
    const uint32_t waveGlobal = blockIdx.x * 4 + wid;
    if (waveGlobal * 64 >= qt.n) return;   // whole wave past the last query (wave-uniform exit)
    const uint32_t region = 2 * waveGlobal + part;
    if (!STAGED && waveCount[region] <= STAGE_PAIRS) return;   // already gathered (wave-uniform)
    const uint32_t i = waveGlobal * 64 + lane;
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    Query q;
    q.valid = i < qt.n;
    q.leaf = i;
    q.col = q.valid ? qt.vals[i] : 0u;
    q.box = q.valid ? qt.nodeAabb[qt.n - 1 + i] : make_float4(0, 0, 0, 0);
    q.body = q.valid ? qt.leafBody[i] : 0u;
    const uint32_t scene = q.valid ? colScene[q.col] : 0u;
    q.sf = q.valid ? tt.sceneStart[scene] : 0u;
    q.sl = q.valid ? tt.sceneStart[scene + 1] : 0u;
    // a sleeping query pairs only with awake dynamic bodies: nothing against the static tree, and
    // in the self pass only leaves whose body is awake (the reference keeps sleeping colliders in a
    // third tree that only dynamic leaves query, BroadPhaseCollisionDetector.cpp:109-244)
    const bool qAsleep = sleepId && q.valid && sleepId[q.body] != 0u;
    if (!SELF && qAsleep) q.valid = false;
    uint32_t written = 0;
    const uint64_t outBase = STAGED ? 0ull : (uint64_t)waveOffset[region];
    uint2* stage = staged + (size_t)region * STAGE_PAIRS;
    auto emit = [&](uint32_t other, uint32_t otherBody, bool act) {
        if (SELF) act = act && otherBody != q.body;
        if (SELF && qAsleep && act) act = sleepId[otherBody] == 0u;
        const uint64_t m = __ballot(act);
        if (act) {
            const uint32_t k = written + (uint32_t)__popcll(m & ltMask);
            if (STAGED) {
                if (k < STAGE_PAIRS) stage[k] = make_uint2(q.col, other);
            } else {
                const uint64_t dst = outBase + k;
                if (dst < cap) pairs[dst] = make_uint2(q.col, other);
            }
        }
        written += (uint32_t)__popcll(m);
    };
    if (tt.n > 0 && __ballot(q.valid) != 0) {
        uint32_t sp = 0;
        const bool rootAct = childActive<SELF>(q, tt.nodeAabb[0], 0u, tt.n - 1);
        if (tt.n == 1) emit(tt.vals[0], tt.leafBody[0], rootAct);
        else if (__ballot(rootAct) != 0) stack[sp++] = 0;
        while (sp > 0) {
            const uint32_t node = __builtin_amdgcn_readfirstlane(stack[sp - 1]);
            sp--;
            const NodeRecord r = tt.rec[node];
            const bool al = childActive<SELF>(q, r.c[0].box, r.c[0].first, r.c[0].last);
            const bool ar = childActive<SELF>(q, r.c[1].box, r.c[1].first, r.c[1].last);
            const bool leafL = r.c[0].first == r.c[0].last, leafR = r.c[1].first == r.c[1].last;
            if (leafL) emit(r.c[0].ref, r.c[0].body, al);
            if (leafR) emit(r.c[1].ref, r.c[1].body, ar);
            const bool pushR = !leafR && __ballot(ar) != 0;
            const bool pushL = !leafL && __ballot(al) != 0;
            if (sp + (pushR ? 1 : 0) + (pushL ? 1 : 0) > TRAVERSAL_STACK) {
                if (lane == 0) atomicOr(&ctr->overflow, OVERFLOW_STACK);
                break;
            }
            if (pushR) { if (lane == 0) stack[sp] = r.c[1].ref; sp++; }
            if (pushL) { if (lane == 0) stack[sp] = r.c[0].ref; sp++; }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (STAGED && lane == 0) waveCount[region] = written;
}

// One wave per staging area: copies a wave's staged pairs to their scanned place.
__global__ __launch_bounds__(256) void kGatherPairs(const uint32_t* __restrict__ waveCount,
                                                    const uint32_t* __restrict__ waveOffset, uint32_t regions,
                                                    const uint2* __restrict__ staged, uint2* __restrict__ pairs,
                                                    uint64_t cap, Counters* ctr, const uint32_t* __restrict__ total) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // the scanned pair total; no kernel of this phase reads it
        ctr->pairs = *total;
        if ((uint64_t)*total > cap) atomicOr(&ctr->overflow, 1u);
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t region = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (region >= regions) return;
    const uint32_t n = waveCount[region];
    if (n > STAGE_PAIRS) return;   // re-traversed straight into place
    const uint64_t base = waveOffset[region];
    const uint2* src = staged + (size_t)region * STAGE_PAIRS;
    for (uint32_t k = lane; k < n; k += 64)
        if (base + k < cap) pairs[base + k] = src[k];
}

inline uint32_t blocksFor(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

// Exclusive scan of one value per thread over a block of BROAD_PREP_THREADS; *total gets the sum.
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t* waveSums, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t k = 0; k < BROAD_PREP_THREADS / 64; k++) {
            const uint32_t s = waveSums[k];
            waveSums[k] = run;
            run += s;
        }
        *total = run;
    }
    __syncthreads();
    const uint32_t r = waveSums[wave] + inc - v;
    __syncthreads();
    return r;
}

// One block: the dynamic and static collider lists (ascending index), each tree's per-scene
// sorted ranges (from per-scene counts) and per-scene bounds (BvhTree.cpp:148-188), from device
// inputs. Bounds are exact min/max, so they equal those of the chunked bounds pass. Dynamic LDS:
// 8 S ints of bounds and 2 S counts.
__global__ __launch_bounds__(BROAD_PREP_THREADS) void kBroadPrep(const float4* __restrict__ aabb,
                                                                 const uint8_t* __restrict__ isDyn,
                                                                 const uint32_t* __restrict__ scene, uint32_t n,
                                                                 uint32_t S, BroadPrep o) {
    extern __shared__ int prepLds[];
    int* bnd = prepLds;                                   // [tree][scene][4]: tree 0 dynamic, 1 static
    uint32_t* cnt = reinterpret_cast<uint32_t*>(prepLds + 8 * S);   // [tree][scene]
    __shared__ uint32_t waveSums[BROAD_PREP_THREADS / 64];
    __shared__ uint32_t nDynTotal, sumA, sumB, bad;
    const uint32_t t = threadIdx.x;
    for (uint32_t k = t; k < 2 * S; k += BROAD_PREP_THREADS) {
        int* b = bnd + 4 * k;
        b[0] = b[1] = floatOrder(FLT_MAX);
        b[2] = b[3] = floatOrder(-FLT_MAX);
        cnt[k] = 0;
    }
    if (t == 0) bad = 0;
    __syncthreads();
    const uint32_t per = (n + BROAD_PREP_THREADS - 1) / BROAD_PREP_THREADS;
    const uint32_t lo = min(n, t * per), hi = min(n, lo + per);
    uint32_t dynHere = 0;
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t d = isDyn[i] ? 1u : 0u;
        dynHere += d;
        const uint32_t s = scene ? scene[i] : 0u;
        if (o.sceneZero) o.sceneZero[i] = 0u;
        if (s >= S) { atomicOr(&bad, 1u); continue; }
        const uint32_t k = (d ? 0u : S) + s;
        const float4 a = aabb[i];
        atomicMin(&bnd[4 * k + 0], floatOrder(a.x));
        atomicMin(&bnd[4 * k + 1], floatOrder(a.y));
        atomicMax(&bnd[4 * k + 2], floatOrder(a.z));
        atomicMax(&bnd[4 * k + 3], floatOrder(a.w));
        atomicAdd(&cnt[k], 1u);
    }
    // stable partition: this thread's colliders go after those of lower threads
    uint32_t dOff = blockExclusiveScan(dynHere, waveSums, &nDynTotal);
    uint32_t sOff = lo - dOff;
    for (uint32_t i = lo; i < hi; i++) {
        if (isDyn[i]) o.dynList[dOff++] = i;
        else o.statList[sOff++] = i;
    }
    // per-scene sorted ranges: exclusive scans of the counts, S <= BROAD_PREP_THREADS
    const uint32_t cd = t < S ? cnt[t] : 0u, cs = t < S ? cnt[S + t] : 0u;
    const uint32_t startD = blockExclusiveScan(cd, waveSums, &sumA);
    const uint32_t startS = blockExclusiveScan(cs, waveSums, &sumB);
    if (t < S) {
        o.dynStart[t] = startD;
        o.statStart[t] = startS;
        o.dynBounds[t] = make_int4(bnd[4 * t], bnd[4 * t + 1], bnd[4 * t + 2], bnd[4 * t + 3]);
        const int* b = bnd + 4 * (S + t);
        o.statBounds[t] = make_int4(b[0], b[1], b[2], b[3]);
    }
    if (t == 0) {
        o.dynStart[S] = sumA;
        o.statStart[S] = sumB;
        o.header[0] = nDynTotal;
        o.header[1] = n - nDynTotal;
        o.header[2] = bad;
    }
}

// Pairs into the caller's two index arrays.
__global__ __launch_bounds__(256) void kSplitPairs(const uint2* __restrict__ pairs, uint32_t n, uint32_t* __restrict__ a,
                                                   uint32_t* __restrict__ b) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint2 p = pairs[i];
        a[i] = p.x;
        b[i] = p.y;
    }
}

}  // namespace

hipError_t launchBroadPrep(hipStream_t st, const float4* aabb, const uint8_t* isDynamic, const uint32_t* scene,
                           uint32_t n, uint32_t S, const BroadPrep& out) {
    if (S == 0 || S > BROAD_PREP_THREADS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kBroadPrep, dim3(1), dim3(BROAD_PREP_THREADS), 40 * S, st, aabb, isDynamic, scene, n, S, out);
    return hipGetLastError();
}

hipError_t launchSplitPairs(hipStream_t st, const uint2* pairs, uint32_t n, uint32_t* a, uint32_t* b) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kSplitPairs, dim3(std::min(blocksFor(n, 256), 4096u)), dim3(256), 0, st, pairs, n, a, b);
    return hipGetLastError();
}

hipError_t launchBuildTree(hipStream_t st, DeviceWorld& w, Tree& t) {
    if (t.n == 0) return hipSuccess;
    if (!t.boundsReady) {
        hipLaunchKernelGGL(kBoundsInit, dim3(blocksFor(w.S, 256)), dim3(256), 0, st, reinterpret_cast<int4*>(t.sceneBounds),
                           w.S);
        hipLaunchKernelGGL(kBounds, dim3(t.chunks), dim3(256), 0, st, t.list, w.aabb, t.chunkScene, t.chunkRange,
                           reinterpret_cast<int4*>(t.sceneBounds));
    }
    hipLaunchKernelGGL(kMorton, dim3(blocksFor(t.n, 256)), dim3(256), 0, st, t.list, t.n, w.aabb, w.colScene,
                       reinterpret_cast<const int4*>(t.sceneBounds), w.mortonBits, t.keys, t.vals);
    const uint32_t keyBits = w.mortonBits + w.sceneBits;
    const uint32_t passes = (keyBits + 7) / 8;
    // sort scratch: tile status, digit totals, tickets (see kRadixScatter)
    uint32_t* status = t.hist;
    uint32_t* digitTotal = status + (size_t)SORT_PASSES_MAX * t.sortBlocks * 256;
    uint32_t* tickets = digitTotal + SORT_PASSES_MAX * 256;
    hipError_t e = launchFillU32(st, t.hist, 0u, (size_t)SORT_PASSES_MAX * (t.sortBlocks + 1) * 256 + SORT_PASSES_MAX);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kRadixHistAll, dim3(blocksFor(t.n, SORT_TILE_LARGE)), dim3(256), 0, st, t.keys, t.n, passes,
                       SORT_TILE_LARGE, digitTotal);
    hipLaunchKernelGGL(kRadixDigitScan, dim3(1), dim3(256), 0, st, digitTotal, passes);
    for (uint32_t p = 0; p < passes; p++) {
#define AGX_SCATTER(THREADS, ROUNDS)                                                                                   \
    hipLaunchKernelGGL((kRadixScatter<THREADS, ROUNDS>), dim3(t.sortBlocks), dim3(THREADS), 0, st, t.keys, t.vals, t.n, \
                       p * 8, digitTotal + p * 256, status + (size_t)p * t.sortBlocks * 256, tickets + p,              \
                       &w.ctr->overflow, t.keysAlt, t.valsAlt)
        static_assert(SORT_TILE_SMALL == 256 * 4 && SORT_TILE_LARGE == 256 * 16, "tile shapes");
        if (t.sortTile == SORT_TILE_LARGE) AGX_SCATTER(256, 16);
        else AGX_SCATTER(256, 4);
#undef AGX_SCATTER
        std::swap(t.keys, t.keysAlt);
        std::swap(t.vals, t.valsAlt);
    }
    const uint32_t tiles = (t.n + BVH_TILE - 1) / BVH_TILE;
    if (tiles > BVH_TOP_MAX_TILES) return hipErrorInvalidValue;   // more leaves than the top refit holds
    if (t.n > 1)
        hipLaunchKernelGGL(kKarrasTiles, dim3(blocksFor(t.n, 256)), dim3(256), 0, st, t.keys, t.n, t.child, t.range,
                           t.parent);
    if (tiles > 1)
        hipLaunchKernelGGL(kKarrasTop, dim3(blocksFor(tiles - 1, 256)), dim3(256), 0, st, t.keys, t.n, t.child, t.range,
                           t.parent);
    hipLaunchKernelGGL(kRefitTiles, dim3(tiles), dim3(BVH_TILE), 0, st, t.n, t.child, t.range, t.parent, t.vals, w.aabb,
                       w.colBody, t.leafBody, t.nodeAabb, t.rec);
    if (tiles > 1)
        hipLaunchKernelGGL(kRefitTop, dim3(1), dim3(BVH_TOP_THREADS), 0, st, t.n, t.child, t.range, t.parent, t.vals,
                           w.colBody, t.nodeAabb, t.rec);
    return hipGetLastError();
}

hipError_t launchPairs(hipStream_t st, DeviceWorld& w) {
    if (w.dyn.n == 0) {
        return launchFillU32(st, &w.ctr->pairs, 0u, 1);
    }
    const uint32_t waves = w.queryWaves;          // ceil(nDyn / 64)
    const uint32_t blocks = blocksFor(waves, 4);
    hipLaunchKernelGGL((kTraverse<true, true>), dim3(blocks), dim3(256), 0, st, w.dyn, w.dyn, w.colScene, w.waveCount,
                       w.waveOffset, 0u, w.pairs, w.pairCap, w.stagedPairs, w.ctr, w.sleepId);
    hipLaunchKernelGGL((kTraverse<false, true>), dim3(blocks), dim3(256), 0, st, w.dyn, w.stat, w.colScene, w.waveCount,
                       w.waveOffset, 1u, w.pairs, w.pairCap, w.stagedPairs, w.ctr, w.sleepId);
    uint32_t* total = &w.ctr->scratch[0];
    hipError_t e = launchScan(st, w.waveCount, w.waveOffset, nullptr, total, 2ull * waves, w.scanBlock, w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kGatherPairs, dim3(blocksFor(2 * waves, 4)), dim3(256), 0, st, w.waveCount, w.waveOffset, 2 * waves,
                       w.stagedPairs, w.pairs, w.pairCap, w.ctr, total);
    hipLaunchKernelGGL((kTraverse<true, false>), dim3(blocks), dim3(256), 0, st, w.dyn, w.dyn, w.colScene, w.waveCount,
                       w.waveOffset, 0u, w.pairs, w.pairCap, w.stagedPairs, w.ctr, w.sleepId);
    hipLaunchKernelGGL((kTraverse<false, false>), dim3(blocks), dim3(256), 0, st, w.dyn, w.stat, w.colScene, w.waveCount,
                       w.waveOffset, 1u, w.pairs, w.pairCap, w.stagedPairs, w.ctr, w.sleepId);
    return hipGetLastError();
}

}  // namespace agx
