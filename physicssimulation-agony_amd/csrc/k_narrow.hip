// Narrow phase: one lane per broad-phase pair, canonical A/B, SAT + clipping manifolds
// (Narrow/CollisionMethods/*.cpp), deterministic compaction and the warm-start cache
// (Narrow/ContactCache.h, NarrowPhaseCollisionDetector.cpp:272-297, 692-767).
#include "agx_internal.h"

#include <cfloat>

namespace agx {

namespace {

struct V { float x, y; };
__device__ inline V operator+(V a, V b) { return { a.x + b.x, a.y + b.y }; }
__device__ inline V operator-(V a, V b) { return { a.x - b.x, a.y - b.y }; }
__device__ inline V operator-(V a) { return { -a.x, -a.y }; }
__device__ inline V operator*(V a, float s) { return { a.x * s, a.y * s }; }
__device__ inline V operator*(float s, V a) { return { s * a.x, s * a.y }; }
__device__ inline float dot(V a, V b) { return a.x * b.x + a.y * b.y; }
__device__ inline V ld(const float2* p, uint32_t i) { const float2 v = p[i]; return { v.x, v.y }; }

constexpr float ZERO_DIV = 1e-4f;          // Narrow/Constants.h:7-9
constexpr float ZERO_DIV_SQ = 1e-8f;
constexpr float SAT_EPS = 1e-4f;
constexpr uint32_t REF_TAG = 0x80000000u;  // Narrow/Constants.h:23

struct Shapes {
    const float4* pose;
    const uint4* shape;
    const float2* dims;
    const float2* verts;
    const float2* normals;
};

__device__ inline uint32_t nextIdx(uint32_t i, uint32_t n) { return i + 1 == n ? 0u : i + 1; }

__device__ inline V flipSign(V v, float sign) {      // HelperMethods.h:23-31
    const uint32_t sb = __float_as_uint(sign) & 0x80000000u;
    return { __uint_as_float(__float_as_uint(v.x) ^ sb), __uint_as_float(__float_as_uint(v.y) ^ sb) };
}

__device__ inline bool clipSegment(V& p1, V& p2, V pp, V pn) {   // HelperMethods.h:33-52
    const float d1 = dot(p1 - pp, pn);
    const float d2 = dot(p2 - pp, pn);
    if (d1 >= 0 && d2 >= 0) return false;
    if (d1 < 0 && d2 < 0) return true;
    const float t = d1 / (d1 - d2);
    const V dir = p2 - p1;
    const V inter = p1 + dir * t;
    if (d1 >= 0) p2 = inter; else p1 = inter;
    return false;
}

__device__ inline void emit(agx_manifold& m, uint32_t a, uint32_t b, V n, float depth, uint32_t count, V p0, V p1,
                            uint32_t id0, uint32_t id1) {
    m.a = a; m.b = b; m.normal_x = n.x; m.normal_y = n.y; m.depth = depth; m.count = count;
    m.p0_x = p0.x; m.p0_y = p0.y;
    m.p1_x = count > 1 ? p1.x : 0.0f; m.p1_y = count > 1 ? p1.y : 0.0f;
    m.id0 = id0; m.id1 = count > 1 ? id1 : 0u;
    m.jn0 = 0.0f; m.jt0 = 0.0f; m.jn1 = 0.0f; m.jt1 = 0.0f;
}

// CircleCircle.cpp:130-176.
__device__ bool circleCircle(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V pA{ pa.x, pa.y }, pB{ pb.x, pb.y };
    const float rA = s.dims[a].x, rB = s.dims[b].x;
    const V delta = pB - pA;
    const float rsum = rA + rB;
    const float sq = dot(delta, delta);
    if (sq >= rsum * rsum) return false;
    const float dist = sqrtf(sq);
    const float depth = rsum - dist;
    V n;
    if (dist < ZERO_DIV) n = { 1.0f, 0.0f };
    else { const float inv = 1.0f / dist; n = delta * inv; }
    emit(out, a, b, n, depth, 1, pA + n * rA, { 0, 0 }, 0, 0);
    return true;
}

// CircleBox.cpp:9-120.
__device__ bool circleBox(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V pA{ pa.x, pa.y }, pB{ pb.x, pb.y };
    const float rA = s.dims[a].x, hw = s.dims[b].x, hh = s.dims[b].y;
    const float cosB = pb.z, sinB = pb.w;
    const V dg = pA - pB;
    const V cl{ dg.x * cosB + dg.y * sinB, dg.x * -sinB + dg.y * cosB };
    const V closest{ fminf(fmaxf(cl.x, -hw), hw), fminf(fmaxf(cl.y, -hh), hh) };
    const V dl = closest - cl;
    const float sq = dot(dl, dl);
    if (sq >= rA * rA) return false;
    if (sq >= ZERO_DIV_SQ) {
        const float dist = sqrtf(sq);
        const float inv = 1.0f / dist;
        const V wd{ cosB * dl.x - sinB * dl.y, sinB * dl.x + cosB * dl.y };
        const V n = wd * inv;
        uint32_t id = (uint32_t)(cl.x > hw);
        id |= (uint32_t)(cl.x < -hw) << 1;
        id |= (uint32_t)(cl.y > hh) << 2;
        id |= (uint32_t)(cl.y < -hh) << 3;
        emit(out, a, b, n, rA - dist, 1, pA + n * rA, { 0, 0 }, id, 0);
        return true;
    }
    const float dx = hw - fabsf(cl.x);
    const float dy = hh - fabsf(cl.y);
    const bool useY = dy <= dx;
    const float minPen = useY ? dy : dx;
    const float sign = copysignf(1.0f, useY ? cl.y : cl.x);
    V nl{ 0.0f, 0.0f };
    if (useY) nl.y = sign; else nl.x = sign;
    const uint32_t id = 1u << ((uint32_t)(sign <= 0) + ((uint32_t)useY << 1));
    const V n = -V{ cosB * nl.x - sinB * nl.y, sinB * nl.x + cosB * nl.y };
    emit(out, a, b, n, rA + minPen, 1, pA + n * rA, { 0, 0 }, id, 0);
    return true;
}

// CirclePolygon.cpp:10-103 with HelperMethods.h:73-106.
__device__ bool circlePolygon(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V pA{ pa.x, pa.y }, pB{ pb.x, pb.y };
    const float cosB = pb.z, sinB = pb.w;
    const float rA = s.dims[a].x;
    const uint4 sh = s.shape[b];
    const float2* v = s.verts + sh.y;
    const float2* nn = s.normals + sh.y;
    const uint32_t count = sh.z;
    const V dg = pA - pB;
    const V cl{ dg.x * cosB + dg.y * sinB, dg.x * -sinB + dg.y * cosB };
    float maxSep = -FLT_MAX;
    uint32_t edge = 0;
    for (uint32_t i = 0; i < count; i++) {
        const float sep = dot(ld(nn, i), cl - ld(v, i));
        if (sep > rA) return false;
        if (sep > maxSep) { maxSep = sep; edge = i; }
    }
    V nl;
    float sep;
    if (maxSep >= ZERO_DIV) {
        const V va = ld(v, edge), vb = ld(v, nextIdx(edge, count));
        const V ab = vb - va;
        const float denom = dot(ab, ab);
        V cp;
        if (denom < ZERO_DIV_SQ) cp = va;
        else { const float t = dot(cl - va, ab) / denom; cp = va + ab * fminf(fmaxf(t, 0.0f), 1.0f); }
        const V d = cl - cp;
        const float dsq = dot(d, d);
        if (dsq > rA * rA) return false;
        if (dsq < ZERO_DIV_SQ) { sep = 0.0f; nl = ld(nn, edge); }
        else { sep = sqrtf(dsq); nl = d * (1.0f / sep); }
    } else {
        nl = ld(nn, edge);
        sep = maxSep;
    }
    const V n = -V{ cosB * nl.x - sinB * nl.y, sinB * nl.x + cosB * nl.y };
    emit(out, a, b, n, rA - sep, 1, pA + n * rA, { 0, 0 }, edge, 0);
    return true;
}

// BoxBox.cpp:13-227.
__device__ bool boxBox(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V pA{ pa.x, pa.y }, pB{ pb.x, pb.y };
    const float cosA = pa.z, sinA = pa.w, cosB = pb.z, sinB = pb.w;
    const float hwA = s.dims[a].x, hhA = s.dims[a].y, hwB = s.dims[b].x, hhB = s.dims[b].y;
    const V rightA{ cosA, sinA }, upA{ -sinA, cosA }, rightB{ cosB, sinB }, upB{ -sinB, cosB };
    const float relCos = cosA * cosB + sinA * sinB;
    const float relSin = cosA * sinB - sinA * cosB;
    const float arc = fabsf(relCos), ars = fabsf(relSin);
    const V cd = pB - pA;
    // Axes [A right, A up, B right, B up]: projected radii sum and centre distance.
    const float ov0r = hwB * arc + hhB * ars + hwA, ov1r = hwB * ars + hhB * arc + hhA;
    const float ov2r = hwA * arc + hhA * ars + hwB, ov3r = hwA * ars + hhA * arc + hhB;
    const float ax[4] = { cosA, -sinA, cosB, -sinB };
    const float ay[4] = { sinA, cosA, sinB, cosB };
    float cdo[4], ov[4];
    const float rr[4] = { ov0r, ov1r, ov2r, ov3r };
#pragma unroll
    for (int k = 0; k < 4; k++) { cdo[k] = cd.x * ax[k] + cd.y * ay[k]; ov[k] = rr[k] - fabsf(cdo[k]); }
    if (ov[0] < 0.0f || ov[1] < 0.0f || ov[2] < 0.0f || ov[3] < 0.0f) return false;
    const float depth = fminf(fminf(ov[0], ov[1]), fminf(ov[2], ov[3]));
    const uint32_t n0 = ov[0] != depth;
    const uint32_t n01 = n0 & (uint32_t)(ov[1] != depth);
    const uint32_t n012 = n01 & (uint32_t)(ov[2] != depth);
    const uint32_t best = n0 + n01 + n012;
    float bax = ax[0], bay = ay[0], bcd = cdo[0];
#pragma unroll
    for (uint32_t k = 1; k < 4; k++) if (best == k) { bax = ax[k]; bay = ay[k]; bcd = cdo[k]; }
    const V normal = flipSign({ bax, bay }, bcd);
    const bool refIsA = best < 2;
    const V pRef = refIsA ? pA : pB, pInc = refIsA ? pB : pA;
    const V rRef = refIsA ? rightA : rightB, uRef = refIsA ? upA : upB;
    const V rInc = refIsA ? rightB : rightA, uInc = refIsA ? upB : upA;
    const float hwRef = refIsA ? hwA : hwB, hhRef = refIsA ? hhA : hhB;
    const float hwInc = refIsA ? hwB : hwA, hhInc = refIsA ? hhB : hhA;
    const V refNormal = refIsA ? normal : -normal;
    const float sign = copysignf(1.0f, refIsA ? bcd : -bcd);
    V refStart, refEnd, refCenter, side;
    {
        const bool useX = best == 0 || best == 2;
        if (useX) { refCenter = pRef + rRef * (sign * hwRef); side = uRef; }
        else { refCenter = pRef + uRef * (sign * hhRef); side = rRef; }
        const V off = side * (useX ? hhRef : hwRef);
        refStart = refCenter + off;
        refEnd = refCenter - off;
    }
    V incStart, incEnd;
    {
        const bool axisIsUp = (best & 1u) != 0;
        const float sideTerm = (best == 0 || best == 3) ? -relSin : relSin;
        const float dotX = sign * (axisIsUp ? sideTerm : relCos);
        const float dotY = sign * (axisIsUp ? relCos : sideTerm);
        const bool useX = fabsf(dotX) > fabsf(dotY);
        const float incSign = -copysignf(1.0f, useX ? dotX : dotY);
        V fc, off;
        if (useX) { fc = pInc + rInc * (incSign * hwInc); off = uInc * hhInc; }
        else { fc = pInc + uInc * (incSign * hhInc); off = rInc * hwInc; }
        incStart = fc + off;
        incEnd = fc - off;
    }
    V c0 = incStart, c1 = incEnd;
    if (clipSegment(c0, c1, refStart, -side)) return false;
    if (clipSegment(c0, c1, refEnd, side)) return false;
    const float refPlane = dot(refCenter, refNormal);
    const bool k0 = dot(c0, refNormal) <= refPlane + SAT_EPS;
    const bool k1 = dot(c1, refNormal) <= refPlane + SAT_EPS;
    const uint32_t idBase = best << 1;
    if (!k0 && !k1) return false;
    if (k0 && k1) emit(out, a, b, normal, depth, 2, c0, c1, idBase, idBase | 1u);
    else if (k0) emit(out, a, b, normal, depth, 1, c0, { 0, 0 }, idBase, 0);
    else emit(out, a, b, normal, depth, 1, c1, { 0, 0 }, idBase | 1u, 0);
    return true;
}

// PolygonPolygon.cpp:18-235. B's vertices and normals in A's frame are recomputed on use
// (the same expression each time, so values are identical to a stored copy).
__device__ bool polygonPolygon(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V pA{ pa.x, pa.y }, pB{ pb.x, pb.y };
    const float cosA = pa.z, sinA = pa.w, cosB = pb.z, sinB = pb.w;
    const uint4 sa = s.shape[a], sb = s.shape[b];
    const uint32_t countA = sa.z, countB = sb.z;
    const float2* vA = s.verts + sa.y;
    const float2* nA = s.normals + sa.y;
    const float2* vB = s.verts + sb.y;
    const float2* nB = s.normals + sb.y;
    const V rightA{ cosA, sinA }, upA{ -sinA, cosA };
    const V cd = pB - pA;
    const V dl{ dot(cd, rightA), dot(cd, upA) };
    const float relCos = cosA * cosB + sinA * sinB;
    const float relSin = cosA * sinB - sinA * cosB;
    auto vBA = [&](uint32_t i) -> V {
        const V v = ld(vB, i);
        return { dl.x + relCos * v.x - relSin * v.y, dl.y + relSin * v.x + relCos * v.y };
    };
    auto nBA = [&](uint32_t i) -> V {
        const V nb = ld(nB, i);
        return { relCos * nb.x - relSin * nb.y, relSin * nb.x + relCos * nb.y };
    };
    V nl{ 0.0f, 0.0f };
    float depth = FLT_MAX;
    bool refIsA = true;
    uint32_t bestIdx = 0;
    float bestFacing = FLT_MAX;
    for (uint32_t i = 0; i < countA; i++) {
        const V axis = ld(nA, i);
        float minB = FLT_MAX;
        for (uint32_t j = 0; j < countB; j++) minB = fminf(minB, dot(vBA(j), axis));
        const float overlap = dot(ld(vA, i), axis) - minB;
        if (overlap < 0.0f) return false;
        const float facing = -dot(axis, dl);
        if (overlap > depth) continue;
        if (overlap < depth || facing < bestFacing) { depth = overlap; nl = axis; refIsA = true; bestIdx = i; bestFacing = facing; }
    }
    for (uint32_t i = 0; i < countB; i++) {
        const V axis = nBA(i);
        float minA = FLT_MAX;
        for (uint32_t j = 0; j < countA; j++) minA = fminf(minA, dot(ld(vA, j), axis));
        const float overlap = dot(vBA(i), axis) - minA;
        if (overlap < 0.0f) return false;
        const float facing = dot(axis, dl);
        if (overlap > depth) continue;
        if (overlap < depth || facing < bestFacing) { depth = overlap; nl = -axis; refIsA = false; bestIdx = i; bestFacing = facing; }
    }
    const uint32_t refCount = refIsA ? countA : countB;
    const uint32_t incCount = refIsA ? countB : countA;
    const V refN = refIsA ? nl : -nl;
    const uint32_t i0 = bestIdx, i1 = nextIdx(i0, refCount);
    const V refStart = refIsA ? ld(vA, i0) : vBA(i0);
    const V refEnd = refIsA ? ld(vA, i1) : vBA(i1);
    const V refCenter = (refStart + refEnd) * 0.5f;
    const V edge = refEnd - refStart;
    const float elen2 = dot(edge, edge);
    const V side = edge * (1.0f / sqrtf(elen2));
    uint32_t inc = 0;
    float minDot = FLT_MAX;
    for (uint32_t i = 0; i < incCount; i++) {
        const V nLocal = refIsA ? nBA(i) : ld(nA, i);
        const float d = dot(refN, nLocal);
        if (d < minDot) { minDot = d; inc = i; }
    }
    const uint32_t inc2 = nextIdx(inc, incCount);
    V c0 = refIsA ? vBA(inc) : ld(vA, inc);
    V c1 = refIsA ? vBA(inc2) : ld(vA, inc2);
    if (clipSegment(c0, c1, refStart, side)) return false;
    if (clipSegment(c0, c1, refEnd, -side)) return false;
    const float refPlane = dot(refCenter, refN);
    const uint32_t tag = refIsA ? REF_TAG : 0u;
    const bool k0 = dot(c0, refN) <= refPlane + SAT_EPS;
    const bool k1 = dot(c1, refN) <= refPlane + SAT_EPS;
    if (!k0 && !k1) return false;
    const V nWorld = rightA * nl.x + upA * nl.y;
    const V w0 = pA + rightA * c0.x + upA * c0.y;
    const V w1 = pA + rightA * c1.x + upA * c1.y;
    if (k0 && k1) emit(out, a, b, nWorld, depth, 2, w0, w1, tag | inc, tag | inc2);
    else if (k0) emit(out, a, b, nWorld, depth, 1, w0, { 0, 0 }, tag | inc, 0);
    else emit(out, a, b, nWorld, depth, 1, w1, { 0, 0 }, tag | inc2, 0);
    return true;
}

// BoxPolygon.cpp:14-319 (polygon B's frame).
__device__ bool boxPolygon(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V pB{ pb.x, pb.y };
    const float cosA = pa.z, sinA = pa.w, cosB = pb.z, sinB = pb.w;
    const float hwA = s.dims[a].x, hhA = s.dims[a].y;
    const uint4 sb = s.shape[b];
    const float2* vB = s.verts + sb.y;
    const float2* nB = s.normals + sb.y;
    const uint32_t count = sb.z;
    const V rightB{ cosB, sinB }, upB{ -sinB, cosB };
    const float relCos = cosA * cosB + sinA * sinB;
    const float relSin = cosA * sinB - sinA * cosB;
    const V rA{ relCos, -relSin }, uA{ relSin, relCos };
    const V dBA = V{ pa.x, pa.y } - pB;
    const V bc{ dot(dBA, rightB), dot(dBA, upB) };
    V nl;
    float depth;
    uint32_t bestType;
    uint32_t bestIdx = 0;
    float bestFacing;
    float pcx2, pcy2;
    {
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (uint32_t i = 0; i < count; i++) { const float p = dot(ld(vB, i), rA); mn = fminf(mn, p); mx = fmaxf(mx, p); }
        const float bp = dot(bc, rA);
        mn -= bp; mx -= bp;
        const float overlap = hwA + fminf(-mn, mx);
        if (overlap < 0.0f) return false;
        pcx2 = mn + mx;
        depth = overlap;
        nl = flipSign(rA, pcx2);
        bestType = 0;
        bestFacing = -fabsf(pcx2);
    }
    {
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (uint32_t i = 0; i < count; i++) { const float p = dot(ld(vB, i), uA); mn = fminf(mn, p); mx = fmaxf(mx, p); }
        const float bp = dot(bc, uA);
        mn -= bp; mx -= bp;
        const float overlap = hhA + fminf(-mn, mx);
        if (overlap < 0.0f) return false;
        pcy2 = mn + mx;
        const float facing = -fabsf(pcy2);
        if (overlap < depth || (overlap == depth && facing < bestFacing)) {
            depth = overlap; nl = flipSign(uA, pcy2); bestType = 1; bestFacing = facing;
        }
    }
    const V pcRel = pcx2 * rA + pcy2 * uA;
    for (uint32_t i = 0; i < count; i++) {
        const V axis = ld(nB, i);
        const float boxRadius = hwA * fabsf(dot(axis, rA)) + hhA * fabsf(dot(axis, uA));
        const float bcp = dot(bc, axis);
        const float overlap = dot(axis, ld(vB, i)) - (bcp - boxRadius);
        if (overlap < 0.0f) return false;
        const float facing = dot(axis, pcRel);
        if (overlap < depth || (overlap == depth && facing < bestFacing)) {
            depth = overlap; nl = -axis; bestType = 2; bestIdx = i; bestFacing = facing;
        }
    }
    const bool refIsBox = bestType != 2;
    V refN, refCenter, side, refStart, refEnd, incStart, incEnd;
    uint32_t inc1, inc2;
    if (refIsBox) {
        refN = nl;
        const bool useX = bestType == 0;
        const float sign = useX ? dot(refN, rA) : dot(refN, uA);
        if (useX) { refCenter = bc + rA * (sign * hwA); side = uA; }
        else { refCenter = bc + uA * (sign * hhA); side = rA; }
        const V off = side * (useX ? hhA : hwA);
        refStart = refCenter + off;
        refEnd = refCenter - off;
        inc1 = 0;
        float minDot = FLT_MAX;
        for (uint32_t i = 0; i < count; i++) {
            const float d = dot(refN, ld(nB, i));
            if (d < minDot) { minDot = d; inc1 = i; }
        }
        inc2 = nextIdx(inc1, count);
        incStart = ld(vB, inc1);
        incEnd = ld(vB, inc2);
    } else {
        refN = -nl;
        const uint32_t i0 = bestIdx, i1 = nextIdx(i0, count);
        refStart = ld(vB, i0);
        refEnd = ld(vB, i1);
        refCenter = (refStart + refEnd) * 0.5f;
        const V edge = refEnd - refStart;
        const float elen2 = dot(edge, edge);
        side = edge * (-1.0f / sqrtf(elen2));
        const float dotX = dot(refN, rA);
        const float dotY = dot(refN, uA);
        const bool useX = fabsf(dotX) > fabsf(dotY);
        const float sign = -copysignf(1.0f, useX ? dotX : dotY);
        // Narrow/Constants.h:34-37 BOX_INCIDENT_EDGE_IDS[useX][sign > 0].
        if (useX) { inc1 = sign > 0 ? 0u : 3u; inc2 = sign > 0 ? 1u : 2u; }
        else { inc1 = sign > 0 ? 0u : 1u; inc2 = sign > 0 ? 3u : 2u; }
        V fc, off;
        if (useX) { fc = bc + rA * (sign * hwA); off = uA * hhA; }
        else { fc = bc + uA * (sign * hhA); off = rA * hwA; }
        incStart = fc + off;
        incEnd = fc - off;
    }
    V c0 = incStart, c1 = incEnd;
    if (clipSegment(c0, c1, refStart, -side)) return false;
    if (clipSegment(c0, c1, refEnd, side)) return false;
    const uint32_t tag = refIsBox ? REF_TAG : 0u;
    const float refPlane = dot(refCenter, refN);
    const bool k0 = dot(c0, refN) <= refPlane + SAT_EPS;
    const bool k1 = dot(c1, refN) <= refPlane + SAT_EPS;
    if (!k0 && !k1) return false;
    const V nWorld = rightB * nl.x + upB * nl.y;
    const V w0 = pB + rightB * c0.x + upB * c0.y;
    const V w1 = pB + rightB * c1.x + upB * c1.y;
    if (k0 && k1) emit(out, a, b, nWorld, depth, 2, w0, w1, tag | inc1, tag | inc2);
    else if (k0) emit(out, a, b, nWorld, depth, 1, w0, { 0, 0 }, tag | inc1, 0);
    else emit(out, a, b, nWorld, depth, 1, w1, { 0, 0 }, tag | inc2, 0);
    return true;
}

// ---------------------------------------------------------------- capsules
// A capsule is its core segment (half length hl along its local +x) grown by its radius. The
// manifolds keep the reference's operation order, contact ids and tie-breaks, for the same bits.
constexpr uint32_t CAPSULE_VERTEX_TAG = 0x40000000u;   // Narrow/Constants.h:40-41
constexpr float CAPSULE_PARALLEL_SIN2 = 1e-3f;

// Point on the core segment from s0 along unit dir (length len) nearest to w.
__device__ inline V coreNearest(V s0, V dir, float len, V w) {
    return s0 + dir * fminf(fmaxf(dot(w - s0, dir), 0.0f), len);
}

// CircleCapsule.cpp:92-147 (its batched path): circle A, capsule B.
__device__ bool circleCapsule(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V centre{ pa.x, pa.y };
    const float rA = s.dims[a].x;
    const float2 cap = s.dims[b];
    const float c = pb.z, sn = pb.w;
    const V off{ pa.x - pb.x, pa.y - pb.y };
    // nearest core point, along the axis (clamped with selects, as the batched path does)
    const float along = off.x * c + off.y * sn, lo = 0.0f - cap.x;
    float u = along < lo ? lo : along;
    u = cap.x < u ? cap.x : u;
    const V gap{ c * u - off.x, sn * u - off.y };
    const float reach = rA + cap.y;
    const float gg = gap.x * gap.x + gap.y * gap.y;
    if (!(gg < reach * reach)) return false;
    const float dist = sqrtf(gg);
    const float inv = 1.0f / dist;
    V n;
    if (dist < ZERO_DIV) {   // centre on the core: out of the capsule's side, away from the centre's side
        const bool below = off.y * c - off.x * sn < 0.0f;
        n = below ? V{ -sn, c } : V{ sn, -c };
    } else {
        n = gap * inv;
    }
    emit(out, a, b, n, reach - dist, 1, centre + n * rA, { 0, 0 }, 0, 0);
    return true;
}

// CapsuleCapsule.cpp:25-157: the nearest points of the two cores (from their centres); side by
// side (near-parallel, overlapping along A's axis) one contact at each end of the overlap.
__device__ bool capsuleCapsule(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const V oA{ pa.x, pa.y }, oB{ pb.x, pb.y };
    const float2 ca = s.dims[a], cb = s.dims[b];
    const float hA = ca.x, hB = cb.x, rA = ca.y, rB = cb.y;
    const V between = oB - oA;
    const float rsum = rA + rB;
    const float bound = hA + hB + rsum;
    if (dot(between, between) >= bound * bound) return false;
    const V axA{ pa.z, pa.w }, axB{ pb.z, pb.w };
    const float cosAB = dot(axA, axB);
    const float alongA = dot(between, axA);
    const float negAlongA = -alongA;
    const float negAlongB = -dot(axB, between);
    const float sin2 = 1.0f - cosAB * cosAB;
    const bool parallel = sin2 < CAPSULE_PARALLEL_SIN2;
    const float reachB = fabsf(cosAB) * hB;
    const float lo = fmaxf(-hA, alongA - reachB), hi = fminf(hA, alongA + reachB);
    const bool overlap = parallel && (hi - lo >= ZERO_DIV);
    auto clampTo = [](float v, float h) { return fminf(fmaxf(v, -h), h); };
    float sa, sb;
    if (overlap) {
        sa = 0.5f * (lo + hi);
        sb = clampTo(cosAB * sa + negAlongB, hB);
    } else {
        sa = parallel ? 0.0f : clampTo((cosAB * negAlongB - negAlongA) / sin2, hA);
        sb = clampTo(cosAB * sa + negAlongB, hB);
        sa = clampTo(cosAB * sb - negAlongA, hA);
    }
    const V nearA = oA + axA * sa, nearB = oB + axB * sb;
    const V d = nearB - nearA;
    const float dd = dot(d, d);
    if (dd >= rsum * rsum) return false;
    V n;
    float depth;
    if (dd >= ZERO_DIV_SQ) {
        const float dist = sqrtf(dd);
        n = d * (1.0f / dist);
        depth = rsum - dist;
    } else {   // the cores cross: out along the side axis that needs the shorter push
        const V sideA{ -axA.y, axA.x }, sideB{ -axB.y, axB.x };
        const float offA = dot(between, sideA), offB = dot(between, sideB);
        const float pushA = fabsf(dot(axB, sideA)) * hB - fabsf(offA);
        const float pushB = fabsf(dot(axA, sideB)) * hA - fabsf(offB);
        if (pushA <= pushB) { n = flipSign(sideA, offA); depth = pushA; }
        else { n = flipSign(sideB, offB); depth = pushB; }
        depth += rsum;
    }
    V pt[2] = { { 0, 0 }, { 0, 0 } };
    uint32_t id[2] = { 0, 0 };
    uint32_t cnt = 0;
    if (overlap) {
        const V sideB{ -axB.y, axB.x };
        const float ends[2] = { lo, hi };
        float nearest = FLT_MAX;
        for (uint32_t k = 0; k < 2; k++) {
            const V p = oA + axA * ends[k];
            const float g = fabsf(dot(p - oB, sideB));
            if (g >= rsum) continue;
            pt[cnt] = p + n * rA;
            id[cnt] = k;
            cnt++;
            nearest = fminf(nearest, g);
        }
        if (cnt) depth = rsum - nearest;
    }
    if (cnt == 0) {
        pt[0] = nearA + n * rA;
        id[0] = 2;
        cnt = 1;
    }
    emit(out, a, b, n, depth, cnt, pt[0], pt[1], id[0], id[1]);
    return true;
}

// Contacts of a core segment (c0, c1 in the reference face's frame, the face's outward normal
// along n) with a face: a point of each end within reach, on the capsule surface.
struct FaceContacts { V p[2]; uint32_t id[2]; uint32_t n; };
__device__ inline FaceContacts faceContacts(V c0, V c1, float e0, float e1, float radius, V n, uint32_t idBase) {
    FaceContacts f{ { { 0, 0 }, { 0, 0 } }, { 0, 0 }, 0 };
    const V span = c1 - c0;
    const bool second = dot(span, span) >= ZERO_DIV_SQ;
    const float reach = radius + SAT_EPS;
    if (e0 <= reach) { f.p[0] = c0 - n * radius; f.id[0] = idBase; f.n = 1; }
    if (second && e1 <= reach) { f.p[f.n] = c1 - n * radius; f.id[f.n] = idBase | 1u; f.n++; }
    return f;
}

// BoxCapsule.cpp:26-242: box A, capsule B, in A's frame folded into the (+, +) quadrant; the
// reference face is turned to x = ext.x.
__device__ bool boxCapsule(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const float2 half = s.dims[a], cap = s.dims[b];
    const V oA{ pa.x, pa.y }, xA{ pa.z, pa.w }, axB{ pb.z, pb.w };
    const float hl = cap.x, radius = cap.y;
    const V yA{ -xA.y, xA.x };
    const V between = V{ pb.x, pb.y } - oA;
    const float lx = dot(between, xA), ly = dot(between, yA);
    const float fx = copysignf(1.0f, lx), fy = copysignf(1.0f, ly);
    const V mid0{ fabsf(lx), fabsf(ly) };
    const V ax0{ (xA.x * axB.x + xA.y * axB.y) * fx, (xA.x * axB.y - xA.y * axB.x) * fy };
    const float aax = fabsf(ax0.x), aay = fabsf(ax0.y);
    const float sepX = mid0.x - hl * aax - half.x;
    const float sepY = mid0.y - hl * aay - half.y;
    const float faceSep = fmaxf(sepX, sepY);
    if (faceSep >= radius) return false;
    const V side{ -ax0.y, ax0.x };
    const float sideOff = dot(side, mid0);
    const float sideSep = fabsf(sideOff) - (half.x * aay + half.y * aax);
    if (sideSep >= radius) return false;
    const bool sideWins = sideSep > faceSep + SAT_EPS;
    const bool turn = !sideWins && sepY > sepX;
    const V ext = turn ? V{ half.y, half.x } : V{ half.x, half.y };
    const V dir = turn ? V{ ax0.y, ax0.x } : ax0;
    const V mid = turn ? V{ mid0.y, mid0.x } : mid0;
    const V bx = xA * fx, by = yA * fy;
    const V wx = turn ? by : bx, wy = turn ? bx : by;
    const V hc = dir * hl;
    const V s0 = mid - hc, s1 = mid + hc;
    const float len = hl * 2.0f;
    auto coreDist2 = [&](V w) { const V d = coreNearest(s0, dir, len, w) - w; return dot(d, d); };

    V n{ 1, 0 };
    float depth = 0.0f;
    FaceContacts fc{ { { 0, 0 }, { 0, 0 } }, { 0, 0 }, 0 };
    bool cornerX = true, cornerY = true;   // the corner a vertex contact takes (folded, turned frame)
    V cornerNormal{ 1, 0 };
    float cornerSep = faceSep;
    bool face = false;
    if (sideWins) {
        cornerNormal = sideOff >= 0.0f ? side : -side;
        cornerSep = sideSep;
        cornerX = cornerNormal.x >= 0.0f;
        cornerY = cornerNormal.y >= 0.0f;
    } else {
        V c0 = s0, c1 = s1;
        float e0 = s0.x - ext.x, e1 = s1.x - ext.x;
        bool corner = false;
        if (faceSep > 0.0f) {   // cores apart: only the part of the core beside the face sees it
            const float yLo = fminf(s0.y, s1.y), yHi = fmaxf(s0.y, s1.y);
            if (yLo > ext.y) {
                corner = true;
            } else {
                const bool under = yLo < -ext.y, over = yHi > ext.y;
                if (under || over) {
                    const float invDy = 1.0f / (s1.y - s0.y);
                    const float tA = (-ext.y - s0.y) * invDy, tB = (ext.y - s0.y) * invDy;
                    const float t0 = fmaxf(0.0f, fminf(tA, tB)), t1 = fminf(1.0f, fmaxf(tA, tB));
                    const V sp = s1 - s0;
                    c0 = s0 + sp * t0;
                    c1 = s0 + sp * t1;
                    e0 = c0.x - ext.x;
                    e1 = c1.x - ext.x;
                    const float eMin = fminf(e0, e1);
                    const float faceDist2 = eMin * eMin;
                    if (under && coreDist2(V{ ext.x, -ext.y }) + SAT_EPS < faceDist2) { corner = true; cornerY = false; }
                    else if (over && coreDist2(ext) + SAT_EPS < faceDist2) corner = true;
                }
            }
        }
        if (!corner) {
            const uint32_t faceId = turn ? (fy < 0.0f ? 0u : 2u) : (fx < 0.0f ? 3u : 1u);
            fc = faceContacts(V{ c0.x, c0.y }, V{ c1.x, c1.y }, e0, e1, radius, V{ 1, 0 }, REF_TAG | (faceId << 1));
            if (fc.n == 0) return false;
            depth = radius - fminf(e0, e1);
            face = true;
        }
    }
    if (!face) {   // one contact: a box corner against the core
        const V w{ cornerX ? ext.x : -ext.x, cornerY ? ext.y : -ext.y };
        const V q = coreNearest(s0, dir, len, w);
        n = cornerNormal;
        depth = radius - cornerSep;
        if (cornerSep > 0.0f) {   // cores apart: the corner's distance decides
            const V d = q - w;
            const float dd = dot(d, d);
            if (dd >= radius * radius) return false;
            if (dd >= ZERO_DIV_SQ) {
                const float dist = sqrtf(dd);
                n = d * (1.0f / dist);
                depth = radius - dist;
            }
        }
        // the corner in the box's own frame (turn undone, then the fold); order (-,-) (+,-) (+,+) (-,+)
        const bool px = (turn ? cornerY : cornerX) != (fx < 0.0f);
        const bool py = (turn ? cornerX : cornerY) != (fy < 0.0f);
        fc.n = 1;
        fc.p[0] = q - n * radius;
        fc.id[0] = CAPSULE_VERTEX_TAG | ((py ? 2u : 0u) | (px != py ? 1u : 0u));
    }
    const V nW = wx * n.x + wy * n.y;
    const V p0 = oA + wx * fc.p[0].x + wy * fc.p[0].y;
    const V p1 = fc.n > 1 ? oA + wx * fc.p[1].x + wy * fc.p[1].y : V{ 0, 0 };
    emit(out, a, b, nW, depth, fc.n, p0, p1, fc.id[0], fc.id[1]);
    return true;
}

// PolygonCapsule.cpp:27-240: polygon A, capsule B, in A's frame.
__device__ bool polygonCapsule(const Shapes& s, uint32_t a, uint32_t b, agx_manifold& out) {
    const float4 pa = s.pose[a], pb = s.pose[b];
    const uint4 sh = s.shape[a];
    const float2* vs = s.verts + sh.y;
    const float2* ns = s.normals + sh.y;
    const uint32_t count = sh.z;
    const float2 cap = s.dims[b];
    const V oA{ pa.x, pa.y }, xA{ pa.z, pa.w }, axB{ pb.z, pb.w };
    const float hl = cap.x, radius = cap.y;
    const V yA{ -xA.y, xA.x };
    const V between = V{ pb.x, pb.y } - oA;
    const V mid{ dot(between, xA), dot(between, yA) };
    const V dir{ xA.x * axB.x + xA.y * axB.y, xA.x * axB.y - xA.y * axB.x };
    const V hc = dir * hl;
    const V s0 = mid - hc, s1 = mid + hc;
    const float len = hl * 2.0f;
    const V side{ -dir.y, dir.x };
    float faceSep = -FLT_MAX, sideLo = FLT_MAX, sideHi = -FLT_MAX;
    uint32_t face = 0, loVertex = 0, hiVertex = 0;
    for (uint32_t i = 0; i < count; i++) {
        const V v = ld(vs, i), nv = ld(ns, i);
        const float sep = fminf(dot(nv, s0 - v), dot(nv, s1 - v));
        if (sep >= radius) return false;
        if (sep > faceSep) { faceSep = sep; face = i; }
        const float sd = dot(v, side);
        if (sd < sideLo) { sideLo = sd; loVertex = i; }
        if (sd > sideHi) { sideHi = sd; hiVertex = i; }
    }
    const float sideOff = dot(side, mid);
    const float abovePlus = sideOff - sideHi, belowMinus = sideLo - sideOff;
    const float sideSep = fmaxf(abovePlus, belowMinus);
    if (sideSep >= radius) return false;
    auto coreDist2 = [&](V w) { const V d = coreNearest(s0, dir, len, w) - w; return dot(d, d); };

    V n{ 0, 0 };
    float depth = 0.0f;
    FaceContacts fc{ { { 0, 0 }, { 0, 0 } }, { 0, 0 }, 0 };
    uint32_t vertex = 0;
    V vertexNormal{ 0, 0 };
    float vertexSep = 0.0f;
    bool faceHit = false;
    if (sideSep > faceSep + SAT_EPS) {   // a vertex pokes the capsule's side
        const bool plus = abovePlus >= belowMinus;
        vertex = plus ? hiVertex : loVertex;
        vertexNormal = plus ? side : -side;
        vertexSep = sideSep;
    } else {
        const uint32_t i0 = face, i1 = nextIdx(i0, count);
        const V va = ld(vs, i0), vb = ld(vs, i1), nf = ld(ns, i0);
        const V tg{ -nf.y, nf.x };
        vertexNormal = nf;
        vertexSep = faceSep;
        V c0 = s0, c1 = s1;
        float e0 = dot(s0 - va, nf), e1 = dot(s1 - va, nf);
        bool atVertex = false;
        if (faceSep > 0.0f) {   // cores apart: an overhanging part of the core may be nearer an end vertex
            const float edgeLen = dot(vb - va, tg);
            const float t0 = dot(s0 - va, tg), t1 = dot(s1 - va, tg);
            const bool beforeStart = fminf(t0, t1) < 0.0f, pastEnd = fmaxf(t0, t1) > edgeLen;
            if (beforeStart || pastEnd) {
                if (clipSegment(c0, c1, va, tg)) { vertex = i0; atVertex = true; }
                else if (clipSegment(c0, c1, vb, -tg)) { vertex = i1; atVertex = true; }
                else {
                    e0 = dot(c0 - va, nf);
                    e1 = dot(c1 - va, nf);
                    const float eMin = fminf(e0, e1);
                    const float faceDist2 = eMin * eMin;
                    if (beforeStart && coreDist2(va) + SAT_EPS < faceDist2) { vertex = i0; atVertex = true; }
                    else if (pastEnd && coreDist2(vb) + SAT_EPS < faceDist2) { vertex = i1; atVertex = true; }
                }
            }
        }
        if (!atVertex) {
            fc = faceContacts(c0, c1, e0, e1, radius, nf, REF_TAG | (i0 << 1));
            if (fc.n == 0) return false;
            n = nf;
            depth = radius - fminf(e0, e1);
            faceHit = true;
        }
    }
    if (!faceHit) {   // one contact: polygon vertex against the core
        const V w = ld(vs, vertex);
        const V q = coreNearest(s0, dir, len, w);
        n = vertexNormal;
        depth = radius - vertexSep;
        if (vertexSep > 0.0f) {
            const V d = q - w;
            const float dd = dot(d, d);
            if (dd >= radius * radius) return false;
            if (dd >= ZERO_DIV_SQ) {
                const float dist = sqrtf(dd);
                n = d * (1.0f / dist);
                depth = radius - dist;
            }
        }
        fc.n = 1;
        fc.p[0] = q - n * radius;
        fc.id[0] = CAPSULE_VERTEX_TAG | vertex;
    }
    const V nW = xA * n.x + yA * n.y;
    const V p0 = oA + xA * fc.p[0].x + yA * fc.p[0].y;
    const V p1 = fc.n > 1 ? oA + xA * fc.p[1].x + yA * fc.p[1].y : V{ 0, 0 };
    emit(out, a, b, nW, depth, fc.n, p0, p1, fc.id[0], fc.id[1]);
    return true;
}

// NarrowPhaseCollisionDetector.cpp:653-667 canonical order, then SymmetricMatrix dispatch.
// A block takes up to NARROW_CHUNK pairs at a time, bins them in LDS by shape combination and then
// runs the bins in order, so the lanes of a wave mostly take the same collision method (a wave runs
// every method any of its lanes needs). Chunks shrink when there are too few pairs to keep about
// NARROW_BUSY_BLOCKS blocks working. Results go to the pair's own slot, so the binning order does
// not show in them.
constexpr uint32_t NARROW_CHUNK = 2048;
constexpr uint32_t NARROW_PER_LANE = NARROW_CHUNK / 256;
constexpr uint32_t NARROW_BUSY_BLOCKS = 1024;
constexpr uint32_t METHODS = 10;   // shape pairs (A type <= B type) of the four shapes
constexpr uint32_t NO_METHOD = 15u;
__global__ __launch_bounds__(256) void kNarrow(const uint2* __restrict__ pairs, Counters* __restrict__ ctr, uint64_t cap,
                                               Shapes s, agx_manifold* __restrict__ slots, uint32_t* __restrict__ flags) {
    __shared__ uint2 binAB[NARROW_CHUNK];       // canonical (A, B) of the chunk's pairs, by method
    __shared__ uint32_t binP[NARROW_CHUNK];     // pair index | method << 28
    __shared__ uint32_t count[16], start[16];
    const uint32_t n = (uint32_t)min((uint64_t)ctr->pairs, cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->scratch[1] = n;   // the pairs the flag scan covers
    const uint32_t t = threadIdx.x, lane = t & (WAVE - 1);
    const uint64_t below = lane ? (~0ull >> (WAVE - lane)) : 0ull;
    const uint32_t busy = min(gridDim.x, NARROW_BUSY_BLOCKS);
    const uint32_t perLane = max(1u, min(NARROW_PER_LANE, (n + busy * 256 - 1) / (busy * 256)));
    const uint32_t chunk = 256 * perLane;
    for (uint32_t base = blockIdx.x * chunk; base < n; base += gridDim.x * chunk) {   // block-uniform
        if (t < 16) count[t] = 0u;
        __syncthreads();
        uint32_t method[NARROW_PER_LANE], rank[NARROW_PER_LANE];
        uint2 ab[NARROW_PER_LANE];
#pragma unroll
        for (uint32_t r = 0; r < NARROW_PER_LANE; r++) {
            const uint32_t p = base + r * 256 + t;
            method[r] = NO_METHOD;
            ab[r] = make_uint2(0u, 0u);
            if (r < perLane && p < n) {
                uint32_t a = pairs[p].x, b = pairs[p].y;
                uint32_t ta = s.shape[a].x, tb = s.shape[b].x;
                if (ta > tb) { const uint32_t x = a; a = b; b = x; const uint32_t u = ta; ta = tb; tb = u; }
                else if (ta == tb && b > a) { const uint32_t x = a; a = b; b = x; }
                method[r] = tb * (tb + 1) / 2 + ta;   // 0..METHODS-1
                ab[r] = make_uint2(a, b);
            }
            // rank within the method's bin: one LDS atomic per method present in the wave
            rank[r] = 0u;
            for (uint32_t c = 0; c < METHODS; c++) {
                const uint64_t m = __ballot(method[r] == c);
                if (!m) continue;
                const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)m) - 1);
                uint32_t first = 0;
                if (lane == leader) first = atomicAdd(&count[c], (uint32_t)__popcll(m));
                first = __shfl(first, leader, WAVE);
                if (method[r] == c) rank[r] = first + (uint32_t)__popcll(m & below);
            }
        }
        __syncthreads();
        if (t == 0) {
            uint32_t run = 0;
            for (uint32_t c = 0; c < METHODS; c++) { start[c] = run; run += count[c]; }
            start[METHODS] = run;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < NARROW_PER_LANE; r++) {
            if (method[r] == NO_METHOD) continue;
            const uint32_t e = start[method[r]] + rank[r];
            binAB[e] = ab[r];
            binP[e] = (base + r * 256 + t) | (method[r] << 28);
        }
        __syncthreads();
        const uint32_t total = start[METHODS];
        for (uint32_t e = t; e < total; e += 256) {
            const uint2 pab = binAB[e];
            const uint32_t pm = binP[e];
            const uint32_t p = pm & ((1u << 28) - 1u), c = pm >> 28;
            agx_manifold m;
            bool hit;
            switch (c) {
            case 0: hit = circleCircle(s, pab.x, pab.y, m); break;
            case 1: hit = circleBox(s, pab.x, pab.y, m); break;
            case 2: hit = boxBox(s, pab.x, pab.y, m); break;
            case 3: hit = circlePolygon(s, pab.x, pab.y, m); break;
            case 4: hit = boxPolygon(s, pab.x, pab.y, m); break;
            case 5: hit = polygonPolygon(s, pab.x, pab.y, m); break;
            case 6: hit = circleCapsule(s, pab.x, pab.y, m); break;
            case 7: hit = boxCapsule(s, pab.x, pab.y, m); break;
            case 8: hit = polygonCapsule(s, pab.x, pab.y, m); break;
            default: hit = capsuleCapsule(s, pab.x, pab.y, m); break;
            }
            if (hit) slots[p] = m;
            flags[p] = hit ? 1u : 0u;
        }
        __syncthreads();   // before the next round reuses the bins
    }
}

__device__ inline unsigned long long pairKey(uint32_t sa, uint32_t sb) {   // NarrowPhaseCollisionDetector.h:215-218
    const unsigned long long lo = min(sa, sb), hi = max(sa, sb);
    return (lo << 32) | hi;
}
__device__ inline uint32_t hashSlot(unsigned long long key, uint32_t capLog2) {   // ContactCache.h:27-30
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - capLog2));
}
constexpr unsigned long long EMPTY_KEY = ~0ull;

// Compaction in pair order and warm start (finalizeResults, NarrowPhaseCollisionDetector.cpp:692-767):
// impulses are copied from the previous manifold of the same collider pair when A/B order and
// the contact id match.
__global__ __launch_bounds__(256) void kCompact(const agx_manifold* __restrict__ slots, const uint32_t* __restrict__ flags,
                                                const uint32_t* __restrict__ scan, Counters* __restrict__ ctr,
                                                agx_manifold* __restrict__ cur, const agx_manifold* __restrict__ prev,
                                                const unsigned long long* __restrict__ hashKeys,
                                                const uint32_t* __restrict__ hashVals, const uint32_t* __restrict__ colSlot,
                                                const float4* __restrict__ histIn, float4* __restrict__ histOut) {
    const uint32_t n = ctr->scratch[1];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->manifolds = ctr->scratch[2];   // the flag scan's total
    const uint32_t prevCount = ctr->prevManifolds;
    const uint32_t cap = ctr->hashCapacity;
    const uint32_t capLog2 = cap ? (uint32_t)__builtin_ctz(cap) : 0u;
    uint32_t contacts = 0;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        if (!flags[p]) continue;
        agx_manifold m = slots[p];
        contacts += m.count;
        const float none = __builtin_nanf("");
        float4 hist = make_float4(none, none, none, none);   // accumulators of two steps ago, per point
        if (prevCount > 0 && cap > 0) {
            const unsigned long long key = pairKey(colSlot[m.a], colSlot[m.b]);
            uint32_t h = hashSlot(key, capLog2);
            uint32_t found = 0xFFFFFFFFu;
            for (uint32_t probe = 0; probe < cap; probe++) {
                const unsigned long long k = hashKeys[h];
                if (k == key) { found = hashVals[h]; break; }
                if (k == EMPTY_KEY) break;
                h = (h + 1) & (cap - 1);
            }
            if (found != 0xFFFFFFFFu) {
                const agx_manifold& q = prev[found];
                if (q.a == m.a && q.b == m.b) {
                    const float4 qh = histIn ? histIn[found] : hist;
                    if (q.count > 0 && m.id0 == q.id0) { m.jn0 = q.jn0; m.jt0 = q.jt0; hist.x = qh.x; hist.y = qh.y; }
                    else if (q.count > 1 && m.id0 == q.id1) { m.jn0 = q.jn1; m.jt0 = q.jt1; hist.x = qh.z; hist.y = qh.w; }
                    if (m.count > 1) {
                        if (q.count > 0 && m.id1 == q.id0) { m.jn1 = q.jn0; m.jt1 = q.jt0; hist.z = qh.x; hist.w = qh.y; }
                        else if (q.count > 1 && m.id1 == q.id1) { m.jn1 = q.jn1; m.jt1 = q.jt1; hist.z = qh.z; hist.w = qh.w; }
                    }
                }
            }
        }
        cur[scan[p]] = m;
        if (histOut) histOut[scan[p]] = hist;
    }
    // one update of the shared counter per block (per wave, the updates queue on its address)
    __shared__ uint32_t waveContacts[256 / 64];
    for (uint32_t off = 32; off > 0; off >>= 1) contacts += __shfl_down(contacts, off, 64);
    if ((threadIdx.x & 63u) == 0) waveContacts[threadIdx.x / 64] = contacts;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = waveContacts[0] + waveContacts[1] + waveContacts[2] + waveContacts[3];
        if (total) atomicAdd(&ctr->contacts, total);
    }
}


// rebuildContactCache (NarrowPhaseCollisionDetector.cpp:272-297): table sized to the manifold
// count, only manifolds with a nonzero accumulator are keyed.
// Power of two, at least twice the key count, capped by the allocation.
__device__ __forceinline__ uint32_t hashCapacityFor(uint32_t m, uint32_t capMax) {
    uint32_t cap = 16;
    while (cap < 2u * m && cap < capMax) cap <<= 1;
    return cap;
}

// Sizes the table for this step's manifolds (the warm-start source of the next step) and clears it.
__global__ __launch_bounds__(256) void kHashClear(unsigned long long* keys, Counters* ctr, uint32_t capMax) {
    const uint32_t m = ctr->manifolds;
    const uint32_t cap = hashCapacityFor(m, capMax);
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // no thread of this kernel reads them
        ctr->hashCapacity = cap;
        ctr->prevManifolds = m;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) keys[i] = EMPTY_KEY;
}

__global__ __launch_bounds__(256) void kHashInsert(const agx_manifold* __restrict__ cur, Counters* ctr,
                                                   unsigned long long* keys, uint32_t* vals, const uint32_t* __restrict__ colSlot) {
    const uint32_t n = ctr->prevManifolds;
    const uint32_t cap = ctr->hashCapacity;
    const uint32_t capLog2 = (uint32_t)__builtin_ctz(cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const agx_manifold m = cur[i];
        bool any = m.jn0 != 0.0f || m.jt0 != 0.0f;
        if (m.count > 1) any = any || m.jn1 != 0.0f || m.jt1 != 0.0f;
        if (!any) continue;
        const unsigned long long key = pairKey(colSlot[m.a], colSlot[m.b]);
        uint32_t h = hashSlot(key, capLog2);
        for (uint32_t probe = 0; probe < cap; probe++) {
            const unsigned long long old = atomicCAS(keys + h, EMPTY_KEY, key);
            if (old == EMPTY_KEY || old == key) { vals[h] = i; break; }
            h = (h + 1) & (cap - 1);
        }
    }
}

inline uint32_t streamBlocks(const DeviceWorld& w, uint64_t cap) {
    const uint64_t want = (cap + 255) / 256;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)w.numCU * 16));
}

}  // namespace

hipError_t launchNarrow(hipStream_t st, DeviceWorld& w) {
    const Shapes s{ w.pose, w.shape, w.dims, w.verts, w.normals };
    const uint32_t blocks = streamBlocks(w, w.pairCap);
    hipLaunchKernelGGL(kNarrow, dim3(blocks), dim3(256), 0, st, w.pairs, w.ctr, w.pairCap, s, w.slots, w.flags);
    hipError_t e = launchScan(st, w.flags, w.flagScan, &w.ctr->scratch[1], &w.ctr->scratch[2], w.pairCap, w.scanBlock,
                              w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kCompact, dim3(blocks), dim3(256), 0, st, w.slots, w.flags, w.flagScan, w.ctr, w.man[w.cur],
                       w.man[1 - w.cur], w.hashKeys, w.hashVals, w.colSlot, w.histIn, w.histOut);
    return hipGetLastError();
}

hipError_t launchHashBuild(hipStream_t st, DeviceWorld& w) {
    const uint32_t blocks = streamBlocks(w, w.hashCapMax);
    hipLaunchKernelGGL(kHashClear, dim3(blocks), dim3(256), 0, st, w.hashKeys, w.ctr, (uint32_t)w.hashCapMax);
    hipLaunchKernelGGL(kHashInsert, dim3(streamBlocks(w, w.manCap)), dim3(256), 0, st, w.man[w.cur], w.ctr, w.hashKeys,
                       w.hashVals, w.colSlot);
    return hipGetLastError();
}

}  // namespace agx
