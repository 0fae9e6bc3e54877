// Broad phase: per-scene bounds, Morton keys, LDS radix sort, Karras LBVH, refit, pair traversal.
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

// Stable scatter of one digit. Each wave ranks SORT_TILE / 256 rounds of 64 consecutive keys with ballot
// matching (8 ballots per round) and per-wave digit counters in LDS; waves are combined in order,
// earlier tiles by look-back, earlier digits by the scanned totals.
template <uint32_t SORT_TILE>
__global__ __launch_bounds__(256) void kRadixScatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                     uint32_t n, uint32_t shift, const uint32_t* __restrict__ digitStart,
                                                     uint32_t* status, uint32_t* ticket, uint32_t* overflow,
                                                     uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut) {
    constexpr uint32_t ROUNDS = SORT_TILE / 256;   // keys per lane
    __shared__ uint32_t cnt[4][256];
    __shared__ uint32_t base[4][256];
    __shared__ uint32_t tileShared;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) tileShared = atomicAdd(ticket, 1u);
    for (uint32_t w = 0; w < 4; w++) cnt[w][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t tile = tileShared;
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t rk[ROUNDS], dg[ROUNDS], ky[ROUNDS], vl[ROUNDS];
    const uint32_t waveBase = tile * SORT_TILE + wid * (SORT_TILE / 4);
    for (uint32_t r = 0; r < ROUNDS; r++) {
        const uint32_t i = waveBase + r * 64 + lane;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t d = (key >> shift) & 0xFFu;
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
        rk[r] = before + (uint32_t)__popcll(peers & ltMask);
        dg[r] = d;
        ky[r] = key;
        vl[r] = valid ? vals[i] : 0u;
    }
    __syncthreads();
    {
        const uint32_t d = threadIdx.x;
        const uint32_t mine = cnt[0][d] + cnt[1][d] + cnt[2][d] + cnt[3][d];
        uint32_t* own = status + tile * 256 + d;
        uint32_t prefix = 0;
        if (tile == 0) {
            storeFlag(own, STATUS_INC | mine);
        } else {
            storeFlag(own, STATUS_AGG | mine);
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
            storeFlag(own, STATUS_INC | (prefix + mine));
        }
        uint32_t run = digitStart[d] + prefix;
        for (uint32_t w = 0; w < 4; w++) { base[w][d] = run; run += cnt[w][d]; }
    }
    __syncthreads();
    for (uint32_t r = 0; r < ROUNDS; r++) {
        const uint32_t i = waveBase + r * 64 + lane;
        if (i < n) {
            const uint32_t dst = base[wid][dg[r]] + rk[r];
            keysOut[dst] = ky[r];
            valsOut[dst] = vl[r];
        }
    }
}

// ---------------------------------------------------------------- Karras LBVH
__device__ inline int prefixDelta(const uint32_t* __restrict__ keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    return a == b ? 32 + __clz((int)((uint32_t)i ^ (uint32_t)j)) : __clz((int)(a ^ b));
}

// Karras 2012: one thread per internal node; ties broken by leaf index.
__global__ __launch_bounds__(256) void kKarras(const uint32_t* __restrict__ keys, uint32_t n, uint2* __restrict__ child,
                                               uint2* __restrict__ range, uint32_t* __restrict__ parent,
                                               uint32_t* __restrict__ flags) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int nn = (int)n;
    if (i >= nn - 1) return;
    const int dR = prefixDelta(keys, nn, i, i + 1), dL = prefixDelta(keys, nn, i, i - 1);
    const int d = dR > dL ? 1 : -1;
    const int dMin = prefixDelta(keys, nn, i, i - d);
    int lMax = 2;
    while (prefixDelta(keys, nn, i, i + lMax * d) > dMin) lMax <<= 1;
    int l = 0;
    for (int t = lMax >> 1; t >= 1; t >>= 1)
        if (prefixDelta(keys, nn, i, i + (l + t) * d) > dMin) l += t;
    const int j = i + l * d;
    const int dNode = prefixDelta(keys, nn, i, j);
    int s = 0;
    int t = l;
    while (t > 1) {
        t = (t + 1) >> 1;
        if (prefixDelta(keys, nn, i, i + (s + t) * d) > dNode) s += t;
    }
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const uint32_t leafBase = n - 1;
    const uint32_t left = (lo == gamma) ? leafBase + (uint32_t)gamma : (uint32_t)gamma;
    const uint32_t right = (hi == gamma + 1) ? leafBase + (uint32_t)(gamma + 1) : (uint32_t)(gamma + 1);
    child[i] = make_uint2(left, right);
    range[i] = make_uint2((uint32_t)lo, (uint32_t)hi);
    parent[left] = (uint32_t)i;
    parent[right] = (uint32_t)i;
    flags[i] = 0;
    if (i == 0) parent[0] = 0xFFFFFFFFu;
}

__global__ __launch_bounds__(256) void kLeaves(const uint32_t* __restrict__ vals, uint32_t n, const float4* __restrict__ aabb,
                                               const uint32_t* __restrict__ colBody, float4* __restrict__ nodeAabb,
                                               uint32_t* __restrict__ leafBody, uint32_t* __restrict__ parent) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t c = vals[j];
    nodeAabb[n - 1 + j] = aabb[c];
    leafBody[j] = colBody[c];
    if (n == 1) parent[0] = 0xFFFFFFFFu;
}

// Bottom-up refit (BvhTree.cpp:426-530 computes the same unions top-down in reverse order):
// the second thread to reach a node merges its children and writes the node's record (both
// children's boxes, ranges and references). Node boxes written here use agent-coherent stores,
// completed before the arrival counter is bumped.
__device__ inline ChildInfo childInfo(uint32_t c, uint32_t leafBase, const float4& box, const uint2* range,
                                      const uint32_t* vals, const uint32_t* leafBody) {
    ChildInfo ci;
    ci.box = box;
    if (c >= leafBase) {
        const uint32_t j = c - leafBase;
        ci.first = ci.last = j;
        ci.ref = vals[j];
        ci.body = leafBody[j];
    } else {
        const uint2 r = range[c];
        ci.first = r.x; ci.last = r.y;
        ci.ref = c;
        ci.body = 0xFFFFFFFFu;
    }
    return ci;
}

__global__ __launch_bounds__(256) void kRefit(uint32_t n, const uint2* __restrict__ child, const uint2* __restrict__ range,
                                              const uint32_t* __restrict__ parent, const uint32_t* __restrict__ vals,
                                              const uint32_t* __restrict__ leafBody, uint32_t* __restrict__ flags,
                                              float4* __restrict__ nodeAabb, NodeRecord* __restrict__ rec) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || n < 2) return;
    const uint32_t leafBase = n - 1;
    uint32_t node = leafBase + j;
    while (true) {
        const uint32_t p = parent[node];
        if (p == 0xFFFFFFFFu) break;
        orderStores();
        const uint32_t old = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) break;
        const uint2 ch = child[p];
        const float4 a = loadCoherent(nodeAabb + ch.x);
        const float4 b = loadCoherent(nodeAabb + ch.y);
        storeCoherent(nodeAabb + p, make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w)));
        NodeRecord r;
        r.c[0] = childInfo(ch.x, leafBase, a, range, vals, leafBody);
        r.c[1] = childInfo(ch.y, leafBase, b, range, vals, leafBody);
        rec[p] = r;
        node = p;
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
                                                 uint2* __restrict__ staged, Counters* ctr) {
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
    uint32_t written = 0;
    const uint64_t outBase = STAGED ? 0ull : (uint64_t)waveOffset[region];
    uint2* stage = staged + (size_t)region * STAGE_PAIRS;
    auto emit = [&](uint32_t other, uint32_t otherBody, bool act) {
        if (SELF) act = act && otherBody != q.body;
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
    hipLaunchKernelGGL(kRadixHistAll, dim3(t.sortBlocks), dim3(256), 0, st, t.keys, t.n, passes, t.sortTile, digitTotal);
    hipLaunchKernelGGL(kRadixDigitScan, dim3(1), dim3(256), 0, st, digitTotal, passes);
    for (uint32_t p = 0; p < passes; p++) {
        if (t.sortTile == SORT_TILE_LARGE)
            hipLaunchKernelGGL(kRadixScatter<SORT_TILE_LARGE>, dim3(t.sortBlocks), dim3(256), 0, st, t.keys, t.vals, t.n, p * 8,
                               digitTotal + p * 256, status + (size_t)p * t.sortBlocks * 256, tickets + p, &w.ctr->overflow,
                               t.keysAlt, t.valsAlt);
        else
            hipLaunchKernelGGL(kRadixScatter<SORT_TILE_SMALL>, dim3(t.sortBlocks), dim3(256), 0, st, t.keys, t.vals, t.n, p * 8,
                               digitTotal + p * 256, status + (size_t)p * t.sortBlocks * 256, tickets + p, &w.ctr->overflow,
                               t.keysAlt, t.valsAlt);
        std::swap(t.keys, t.keysAlt);
        std::swap(t.vals, t.valsAlt);
    }
    if (t.n > 1)
        hipLaunchKernelGGL(kKarras, dim3(blocksFor(t.n - 1, 256)), dim3(256), 0, st, t.keys, t.n, t.child, t.range,
                           t.parent, t.flags);
    hipLaunchKernelGGL(kLeaves, dim3(blocksFor(t.n, 256)), dim3(256), 0, st, t.vals, t.n, w.aabb, w.colBody, t.nodeAabb,
                       t.leafBody, t.parent);
    if (t.n > 1)
        hipLaunchKernelGGL(kRefit, dim3(blocksFor(t.n, 256)), dim3(256), 0, st, t.n, t.child, t.range, t.parent, t.vals,
                           t.leafBody, t.flags, t.nodeAabb, t.rec);
    return hipGetLastError();
}

hipError_t launchPairs(hipStream_t st, DeviceWorld& w) {
    if (w.dyn.n == 0) {
        return launchFillU32(st, &w.ctr->pairs, 0u, 1);
    }
    const uint32_t waves = w.queryWaves;          // ceil(nDyn / 64)
    const uint32_t blocks = blocksFor(waves, 4);
    hipLaunchKernelGGL((kTraverse<true, true>), dim3(blocks), dim3(256), 0, st, w.dyn, w.dyn, w.colScene, w.waveCount,
                       w.waveOffset, 0u, w.pairs, w.pairCap, w.stagedPairs, w.ctr);
    hipLaunchKernelGGL((kTraverse<false, true>), dim3(blocks), dim3(256), 0, st, w.dyn, w.stat, w.colScene, w.waveCount,
                       w.waveOffset, 1u, w.pairs, w.pairCap, w.stagedPairs, w.ctr);
    uint32_t* total = &w.ctr->scratch[0];
    hipError_t e = launchScan(st, w.waveCount, w.waveOffset, nullptr, total, 2ull * waves, w.scanBlock, w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kGatherPairs, dim3(blocksFor(2 * waves, 4)), dim3(256), 0, st, w.waveCount, w.waveOffset, 2 * waves,
                       w.stagedPairs, w.pairs, w.pairCap, w.ctr, total);
    hipLaunchKernelGGL((kTraverse<true, false>), dim3(blocks), dim3(256), 0, st, w.dyn, w.dyn, w.colScene, w.waveCount,
                       w.waveOffset, 0u, w.pairs, w.pairCap, w.stagedPairs, w.ctr);
    hipLaunchKernelGGL((kTraverse<false, false>), dim3(blocks), dim3(256), 0, st, w.dyn, w.stat, w.colScene, w.waveCount,
                       w.waveOffset, 1u, w.pairs, w.pairCap, w.stagedPairs, w.ctr);
    return hipGetLastError();
}

}  // namespace agx
