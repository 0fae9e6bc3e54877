// Oracle broad phase (test infrastructure only).
// broadPhaseBvh restates Broad/BvhTree.cpp + BvhPairTraversal.h + LeafPairTesting.h so the
// CPU baseline has the reference's structure and pair order; broadPhaseBrute is an exact
// sweep-and-prune over the same predicate, used to pin the pair set.
#include "oracle.h"

#include <algorithm>
#include <bit>
#include <cfloat>
#include <cmath>

namespace orc {

namespace {

constexpr uint32_t LEAF = 16;            // BvhNode.h:13 (KD_LEAF_SIZE)
constexpr uint32_t INVALID = 0xFFFFFFFFu;

struct Node {
    float minX, maxX, minY, maxY;
    uint32_t left = INVALID;
    uint32_t start, end;
    uint32_t leafIndex = 0;
    bool overlaps(const Node& o) const {    // BvhNode.h:33-38
        return minX < o.maxX && maxX > o.minX && minY < o.maxY && maxY > o.minY;
    }
};

struct Tree {
    std::vector<Node> nodes;
    std::vector<uint32_t> indices;                       // sorted collider indices
    std::vector<float> lminX, lmaxX, lminY, lmaxY;       // per-leaf 16-wide SoA
    bool empty() const { return nodes.empty(); }
};

uint32_t part1By1(uint32_t x) {                          // BvhTree.cpp:27-35
    x &= 0x0000ffffu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// BvhTree.cpp:76-104, 148-530.
void buildTree(Tree& t, const World& w, uint32_t begin, uint32_t end) {
    t.nodes.clear();
    const uint32_t count = end - begin;
    if (count == 0) return;
    // gatherCentroids (:148-188).
    std::vector<float> cx(count), cy(count);
    float bMinX = FLT_MAX, bMinY = FLT_MAX, bMaxX = -FLT_MAX, bMaxY = -FLT_MAX;
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t c = begin + i;
        cx[i] = w.minX[c] + w.maxX[c];
        cy[i] = w.minY[c] + w.maxY[c];
        bMinX = std::fmin(bMinX, w.minX[c]); bMinY = std::fmin(bMinY, w.minY[c]);
        bMaxX = std::fmax(bMaxX, w.maxX[c]); bMaxY = std::fmax(bMaxY, w.maxY[c]);
    }
    // transformCentroids (:191-250).
    const float kEps = 1e-5f;
    const float maxX = std::fmax(bMaxX, bMinX + kEps), maxY = std::fmax(bMaxY, bMinY + kEps);
    const float clampMax = (float)0xFFFFu;
    const float scaleX = clampMax / (maxX - bMinX), scaleY = clampMax / (maxY - bMinY);
    const float nsx = -bMinX * scaleX, nsy = -bMinY * scaleY;
    const float hsx = scaleX * 0.5f, hsy = scaleY * 0.5f;
    // computeMortonCodes (:252-280) + sortByMortonCodes (:282-354): stable 4 x 8-bit LSD.
    std::vector<std::pair<uint32_t, uint32_t>> a(count), b(count);
    for (uint32_t i = 0; i < count; i++) {
        const float tx = std::clamp(cx[i] * hsx + nsx, 0.0f, clampMax);
        const float ty = std::clamp(cy[i] * hsy + nsy, 0.0f, clampMax);
        a[i] = { (part1By1((uint32_t)ty) << 1) | part1By1((uint32_t)tx), begin + i };
    }
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        uint32_t counts[256] = {};
        for (const auto& e : a) counts[(e.first >> shift) & 0xFF]++;
        uint32_t sum = 0;
        for (uint32_t& c : counts) { const uint32_t v = c; c = sum; sum += v; }
        for (const auto& e : a) b[counts[(e.first >> shift) & 0xFF]++] = e;
        a.swap(b);
    }
    t.indices.resize(count);
    std::vector<uint32_t> codes(count);
    for (uint32_t i = 0; i < count; i++) { t.indices[i] = a[i].second; codes[i] = a[i].first; }
    // buildNodes (:358-424).
    uint32_t leafCount = 0;
    std::vector<uint32_t> stack;
    Node root; root.start = 0; root.end = count;
    t.nodes.push_back(root);
    stack.push_back(0);
    while (!stack.empty()) {
        const uint32_t ni = stack.back(); stack.pop_back();
        const uint32_t ns = t.nodes[ni].start, ne = t.nodes[ni].end, size = ne - ns;
        if (size <= LEAF) { t.nodes[ni].leafIndex = leafCount++; continue; }
        const uint32_t first = codes[ns], last = codes[ne - 1];
        uint32_t mid;
        if (first == last) {
            mid = ns + (size >> 1);
        } else {
            const uint32_t splitBit = std::bit_floor(first ^ last);
            uint32_t lo = ns, hi = ne - 1;
            while (lo < hi) {
                const uint32_t m = (lo + hi) >> 1;
                if ((codes[m] & splitBit) == 0u) lo = m + 1; else hi = m;
            }
            mid = std::clamp(lo, ns + 1u, ne - 1u);
        }
        const uint32_t li = (uint32_t)t.nodes.size();
        t.nodes[ni].left = li;
        stack.push_back(li + 1);
        stack.push_back(li);
        Node l; l.start = ns; l.end = mid;
        Node r; r.start = mid; r.end = ne;
        t.nodes.push_back(l);
        t.nodes.push_back(r);
    }
    // fitNodes (:426-530).
    t.lminX.assign((size_t)leafCount * LEAF, FLT_MAX); t.lmaxX.assign((size_t)leafCount * LEAF, -FLT_MAX);
    t.lminY.assign((size_t)leafCount * LEAF, FLT_MAX); t.lmaxY.assign((size_t)leafCount * LEAF, -FLT_MAX);
    for (size_t idx = t.nodes.size(); idx-- > 0;) {
        Node& n = t.nodes[idx];
        if (n.left != INVALID) {
            const Node& l = t.nodes[n.left];
            const Node& r = t.nodes[n.left + 1];
            n.minX = std::fmin(l.minX, r.minX); n.maxX = std::fmax(l.maxX, r.maxX);
            n.minY = std::fmin(l.minY, r.minY); n.maxY = std::fmax(l.maxY, r.maxY);
            continue;
        }
        n.minX = FLT_MAX; n.maxX = -FLT_MAX; n.minY = FLT_MAX; n.maxY = -FLT_MAX;
        for (uint32_t k = 0; k < n.end - n.start; k++) {
            const uint32_t c = t.indices[n.start + k];
            const size_t s = (size_t)n.leafIndex * LEAF + k;
            t.lminX[s] = w.minX[c]; t.lmaxX[s] = w.maxX[c]; t.lminY[s] = w.minY[c]; t.lmaxY[s] = w.maxY[c];
            n.minX = std::fmin(n.minX, w.minX[c]); n.maxX = std::fmax(n.maxX, w.maxX[c]);
            n.minY = std::fmin(n.minY, w.minY[c]); n.maxY = std::fmax(n.maxY, w.maxY[c]);
        }
    }
}

struct NodePair { uint32_t a, b; };

// BvhPairTraversal.h:16-76.
void traverseRoot(const Tree& t, std::vector<NodePair>& nodePairs, std::vector<NodePair>& leafPairs,
                  std::vector<uint32_t>& leaves) {
    if (t.nodes[0].left == INVALID) { leaves.push_back(0); return; }
    std::vector<uint32_t> stack{ 0 };
    while (!stack.empty()) {
        const uint32_t ni = stack.back(); stack.pop_back();
        const uint32_t L = t.nodes[ni].left, R = L + 1;
        const bool lLeaf = t.nodes[L].left == INVALID, rLeaf = t.nodes[R].left == INVALID;
        if (rLeaf) leaves.push_back(R); else stack.push_back(R);
        if (lLeaf) leaves.push_back(L); else stack.push_back(L);
        if (t.nodes[L].overlaps(t.nodes[R])) {
            if (lLeaf && rLeaf) leafPairs.push_back({ L, R });
            else nodePairs.push_back({ L, R });
        }
    }
}

// BvhPairTraversal.h:128-212.
void traverseStack(const Tree& ta, const Tree& tb, std::vector<NodePair>& stack, std::vector<NodePair>& out) {
    auto emit = [&](NodePair l, NodePair r, bool oL, bool oR, bool llL, bool llR) {
        if (oR) {
            if (llR) {
                out.push_back(r);
                if (oL) { if (llL) out.push_back(l); else stack.push_back(l); }
            } else {
                if (oL) stack.push_back(l);
                stack.push_back(r);
            }
        } else if (oL) {
            if (llL) out.push_back(l); else stack.push_back(l);
        }
    };
    while (!stack.empty()) {
        const NodePair p = stack.back(); stack.pop_back();
        const Node& A = ta.nodes[p.a];
        const Node& B = tb.nodes[p.b];
        const bool aLeaf = A.left == INVALID, bLeaf = B.left == INVALID;
        if (aLeaf && bLeaf) { out.push_back(p); continue; }
        const float areaA = (A.maxX - A.minX) * (A.maxY - A.minY);
        const float areaB = (B.maxX - B.minX) * (B.maxY - B.minY);
        const bool splitB = aLeaf || (!bLeaf && (areaB > areaA));
        if (splitB) {
            const uint32_t l = B.left, r = l + 1;
            emit({ p.a, l }, { p.a, r }, A.overlaps(tb.nodes[l]), A.overlaps(tb.nodes[r]),
                 aLeaf && tb.nodes[l].left == INVALID, aLeaf && tb.nodes[r].left == INVALID);
        } else {
            const uint32_t l = A.left, r = l + 1;
            emit({ l, p.b }, { r, p.b }, B.overlaps(ta.nodes[l]), B.overlaps(ta.nodes[r]),
                 bLeaf && ta.nodes[l].left == INVALID, bLeaf && ta.nodes[r].left == INVALID);
        }
    }
}

// LeafPairTesting.h:71-203 (the mask build order reproduces the SIMD pass pair order).
void leafSelf(const Tree& t, uint32_t leaf, const World& w, std::vector<uint32_t>& pa, std::vector<uint32_t>& pb) {
    const Node& n = t.nodes[leaf];
    const size_t s = (size_t)n.leafIndex * LEAF;
    const uint32_t count = n.end - n.start;
    for (uint32_t i = 0; i < count; i++) {
        for (uint32_t j = i + 1; j < LEAF; j++) {
            if (!(t.lminX[s + i] < t.lmaxX[s + j] && t.lmaxX[s + i] > t.lminX[s + j] &&
                  t.lminY[s + i] < t.lmaxY[s + j] && t.lmaxY[s + i] > t.lminY[s + j])) continue;
            const uint32_t A = t.indices[n.start + i], B = t.indices[n.start + j];
            if (w.colBody[A] == w.colBody[B]) continue;
            pa.push_back(A); pb.push_back(B);
        }
    }
}

void leafCross(const Tree& ta, uint32_t la, const Tree& tb, uint32_t lb, bool sameBodyCheck, const World& w,
               std::vector<uint32_t>& pa, std::vector<uint32_t>& pb) {
    const Node& A = ta.nodes[la];
    const Node& B = tb.nodes[lb];
    const size_t sa = (size_t)A.leafIndex * LEAF, sb = (size_t)B.leafIndex * LEAF;
    for (uint32_t i = 0; i < A.end - A.start; i++) {
        for (uint32_t j = 0; j < LEAF; j++) {
            if (!(ta.lminX[sa + i] < tb.lmaxX[sb + j] && ta.lmaxX[sa + i] > tb.lminX[sb + j] &&
                  ta.lminY[sa + i] < tb.lmaxY[sb + j] && ta.lmaxY[sa + i] > tb.lminY[sb + j])) continue;
            const uint32_t ca = ta.indices[A.start + i], cb = tb.indices[B.start + j];
            if (sameBodyCheck && w.colBody[ca] == w.colBody[cb]) continue;
            pa.push_back(ca); pb.push_back(cb);
        }
    }
}

}  // namespace

// BroadPhaseCollisionDetector.cpp:109-244 (single-threaded path, as the step forces).
void broadPhaseBvh(World& w) {
    w.pairA.clear(); w.pairB.clear();
    Tree dyn, stat;
    buildTree(dyn, w, w.staticColliders, w.colliderCount());
    buildTree(stat, w, 0, w.staticColliders);
    if (dyn.empty()) return;
    std::vector<NodePair> nodePairs, leafPairs;
    std::vector<uint32_t> leaves;
    traverseRoot(dyn, nodePairs, leafPairs, leaves);
    traverseStack(dyn, dyn, nodePairs, leafPairs);
    std::vector<NodePair> crossLeafPairs;
    if (!stat.empty() && dyn.nodes[0].overlaps(stat.nodes[0])) {
        std::vector<NodePair> crossNodes{ { 0, 0 } };
        traverseStack(dyn, stat, crossNodes, crossLeafPairs);
    }
    for (uint32_t leaf : leaves) leafSelf(dyn, leaf, w, w.pairA, w.pairB);
    for (const NodePair& p : leafPairs) leafCross(dyn, p.a, dyn, p.b, true, w, w.pairA, w.pairB);
    for (const NodePair& p : crossLeafPairs) leafCross(dyn, p.a, stat, p.b, false, w, w.pairA, w.pairB);
    // Sleeping colliders form the reference's third tree, queried by dynamic leaves only
    // (BroadPhaseCollisionDetector.cpp:109-244): no pair of two sleeping bodies or of a sleeping
    // and a static body.
    if (!w.sleepId.empty()) {
        size_t k = 0;
        for (size_t i = 0; i < w.pairA.size(); i++) {
            const uint32_t a = w.colBody[w.pairA[i]], b = w.colBody[w.pairB[i]];
            const bool sa = asleep(w, a), sb = asleep(w, b);
            if ((sa && (sb || w.invMass[b] == 0.0f)) || (sb && w.invMass[a] == 0.0f)) continue;
            w.pairA[k] = w.pairA[i];
            w.pairB[k] = w.pairB[i];
            k++;
        }
        w.pairA.resize(k);
        w.pairB.resize(k);
    }
}

// Exact pair set by sweep-and-prune on minX: {strictly overlapping AABBs, different owner
// bodies, at least one dynamic}; self pairs as (lower, higher) index, cross as (dyn, static).
void broadPhaseBrute(const World& w, std::vector<uint32_t>& pa, std::vector<uint32_t>& pb) {
    pa.clear(); pb.clear();
    const uint32_t n = w.colliderCount();
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&w](uint32_t x, uint32_t y) {
        return w.minX[x] < w.minX[y] || (w.minX[x] == w.minX[y] && x < y);
    });
    for (uint32_t ii = 0; ii < n; ii++) {
        const uint32_t i = order[ii];
        for (uint32_t jj = ii + 1; jj < n; jj++) {
            const uint32_t j = order[jj];
            if (!(w.minX[j] < w.maxX[i])) break;     // sorted: no later j can overlap i in x
            if (!(w.minX[i] < w.maxX[j] && w.maxX[i] > w.minX[j] && w.minY[i] < w.maxY[j] && w.maxY[i] > w.minY[j])) continue;
            const bool di = i >= w.staticColliders, dj = j >= w.staticColliders;
            if (!di && !dj) continue;
            if (w.colBody[i] == w.colBody[j]) continue;
            const bool si = asleep(w, w.colBody[i]), sj = asleep(w, w.colBody[j]);
            if ((si && (sj || !dj)) || (sj && !di)) continue;   // awake dynamic x sleeping only
            if (di && dj) { pa.push_back(std::min(i, j)); pb.push_back(std::max(i, j)); }
            else if (di) { pa.push_back(i); pb.push_back(j); }
            else { pa.push_back(j); pb.push_back(i); }
        }
    }
}

}  // namespace orc
