// Oracle narrow phase (test infrastructure only): Narrow/NarrowPhaseCollisionDetector.cpp
// and the ten CollisionMethods/*.cpp, restated in the reference's operation order.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace orc {

namespace {

constexpr float ZERO_DIV = 1e-4f;           // Narrow/Constants.h:7-9
constexpr float ZERO_DIV_SQ = 1e-8f;
constexpr float SAT_EPS = 1e-4f;
constexpr uint32_t REF_TAG = 0x80000000u;   // Narrow/Constants.h:23
constexpr uint32_t BOX_INCIDENT[2][2][2] = { { { 1, 2 }, { 0, 3 } }, { { 3, 2 }, { 0, 1 } } };  // :34-37

inline uint32_t nextIdx(uint32_t i, uint32_t n) { return i + 1 == n ? 0u : i + 1; }

inline void flipSignIfNegative(V2& v, float sign) {   // HelperMethods.h:23-31
    uint32_t sb, x, y;
    std::memcpy(&sb, &sign, 4); std::memcpy(&x, &v.x, 4); std::memcpy(&y, &v.y, 4);
    sb &= 0x80000000u; x ^= sb; y ^= sb;
    std::memcpy(&v.x, &x, 4); std::memcpy(&v.y, &y, 4);
}

inline bool clipSegment(V2& p1, V2& p2, V2 planePoint, V2 planeNormal) {   // HelperMethods.h:33-52
    const float d1 = dot(p1 - planePoint, planeNormal);
    const float d2 = dot(p2 - planePoint, planeNormal);
    if (d1 >= 0 && d2 >= 0) return false;
    if (d1 < 0 && d2 < 0) return true;
    const float t = d1 / (d1 - d2);
    const V2 dir = p2 - p1;
    const V2 inter = p1 + dir * t;
    if (d1 >= 0) p2 = inter; else p1 = inter;
    return false;
}

inline void projectOnAxis(const V2* v, uint32_t n, V2 axis, float& mn, float& mx) {   // :54-71
    mn = FLT_MAX; mx = -FLT_MAX;
    for (uint32_t i = 0; i < n; i++) { const float p = dot(v[i], axis); mn = std::fmin(mn, p); mx = std::fmax(mx, p); }
}

inline V2 closestPointOnSegment(V2 p, V2 a, V2 b) {   // :73-80
    const V2 ab = b - a;
    const float denom = dot(ab, ab);
    if (denom < ZERO_DIV_SQ) return a;
    const float t = dot(p - a, ab) / denom;
    return a + ab * std::fmin(std::fmax(t, 0.0f), 1.0f);
}

inline bool findSupportEdge(const V2* v, const V2* n, uint32_t count, V2 point, float maxDist,
                            uint32_t& edgeOut, float& sepOut) {   // :82-106
    float maxSep = -FLT_MAX;
    uint32_t edge = 0;
    for (uint32_t i = 0; i < count; i++) {
        const float sep = dot(n[i], point - v[i]);
        if (sep > maxDist) return false;
        if (sep > maxSep) { maxSep = sep; edge = i; }
    }
    edgeOut = edge; sepOut = maxSep;
    return true;
}

Manifold makeManifold(uint32_t a, uint32_t b, V2 n, float depth, uint32_t count, V2 p0, V2 p1, uint32_t id0, uint32_t id1) {
    Manifold m{};
    m.a = a; m.b = b; m.nx = n.x; m.ny = n.y; m.depth = depth; m.count = count;
    m.p0x = p0.x; m.p0y = p0.y;
    m.p1x = count > 1 ? p1.x : 0.0f; m.p1y = count > 1 ? p1.y : 0.0f;
    m.id0 = id0; m.id1 = count > 1 ? id1 : 0u;
    return m;
}

struct Pose { V2 p; float c, s; };
inline Pose pose(const World& w, uint32_t i) { return { { w.wX[i], w.wY[i] }, w.wCos[i], w.wSin[i] }; }

// CollisionMethods/CircleCircle.cpp:130-176.
bool circleCircle(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 pA{ w.wX[a], w.wY[a] }, pB{ w.wX[b], w.wY[b] };
    const float rA = w.p0[a], rB = w.p0[b];
    const V2 delta = pB - pA;
    const float rsum = rA + rB;
    const float sq = dot(delta, delta);
    if (sq >= rsum * rsum) return false;
    const float dist = std::sqrt(sq);
    const float depth = rsum - dist;
    V2 n;
    if (dist < ZERO_DIV) n = { 1.0f, 0.0f };
    else { const float inv = 1.0f / dist; n = delta * inv; }
    out = makeManifold(a, b, n, depth, 1, pA + n * rA, { 0, 0 }, 0, 0);
    return true;
}

// CollisionMethods/CircleBox.cpp:9-120.
bool circleBox(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 pA{ w.wX[a], w.wY[a] }, pB{ w.wX[b], w.wY[b] };
    const float rA = w.p0[a], hw = w.p0[b], hh = w.p1[b];
    const float cosB = w.wCos[b], sinB = w.wSin[b];
    const V2 dg = pA - pB;
    const V2 cl{ dg.x * cosB + dg.y * sinB, dg.x * -sinB + dg.y * cosB };
    const V2 closest{ std::fmin(std::fmax(cl.x, -hw), hw), std::fmin(std::fmax(cl.y, -hh), hh) };
    const V2 dl = closest - cl;
    const float sq = dot(dl, dl);
    if (sq >= rA * rA) return false;
    if (sq >= ZERO_DIV_SQ) {
        const float dist = std::sqrt(sq);
        const float inv = 1.0f / dist;
        const V2 wd{ cosB * dl.x - sinB * dl.y, sinB * dl.x + cosB * dl.y };
        const V2 n = wd * inv;
        uint32_t id = 0;
        id |= (uint32_t)(cl.x > hw);
        id |= (uint32_t)(cl.x < -hw) << 1;
        id |= (uint32_t)(cl.y > hh) << 2;
        id |= (uint32_t)(cl.y < -hh) << 3;
        out = makeManifold(a, b, n, rA - dist, 1, pA + n * rA, { 0, 0 }, id, 0);
        return true;
    }
    const float dx = hw - std::fabs(cl.x);
    const float dy = hh - std::fabs(cl.y);
    const bool useY = dy <= dx;
    const float minPen = useY ? dy : dx;
    const float sign = std::copysign(1.0f, useY ? cl.y : cl.x);
    V2 nl{ 0.0f, 0.0f };
    if (useY) nl.y = sign; else nl.x = sign;
    const uint32_t id = 1u << ((uint32_t)(sign <= 0) + ((uint32_t)useY << 1));
    const V2 n = -V2{ cosB * nl.x - sinB * nl.y, sinB * nl.x + cosB * nl.y };
    out = makeManifold(a, b, n, rA + minPen, 1, pA + n * rA, { 0, 0 }, id, 0);
    return true;
}

// CollisionMethods/CirclePolygon.cpp:10-103.
bool circlePolygon(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 pA{ w.wX[a], w.wY[a] }, pB{ w.wX[b], w.wY[b] };
    const float cosB = w.wCos[b], sinB = w.wSin[b];
    const float rA = w.p0[a];
    const V2* v = &w.verts[w.polyStart[b]];
    const V2* nn = &w.normals[w.polyStart[b]];
    const uint32_t count = w.polyCount[b];
    const V2 dg = pA - pB;
    const V2 cl{ dg.x * cosB + dg.y * sinB, dg.x * -sinB + dg.y * cosB };
    uint32_t edge = 0;
    float maxSep = 0;
    if (!findSupportEdge(v, nn, count, cl, rA, edge, maxSep)) return false;
    V2 nl;
    float sep;
    if (maxSep >= ZERO_DIV) {
        const V2 cp = closestPointOnSegment(cl, v[edge], v[nextIdx(edge, count)]);
        const V2 d = cl - cp;
        const float dsq = dot(d, d);
        if (dsq > rA * rA) return false;
        if (dsq < ZERO_DIV_SQ) { sep = 0.0f; nl = nn[edge]; }
        else { sep = std::sqrt(dsq); nl = d * (1.0f / sep); }
    } else {
        nl = nn[edge];
        sep = maxSep;
    }
    const V2 n = -V2{ cosB * nl.x - sinB * nl.y, sinB * nl.x + cosB * nl.y };
    out = makeManifold(a, b, n, rA - sep, 1, pA + n * rA, { 0, 0 }, edge, 0);
    return true;
}

// CollisionMethods/BoxBox.cpp:13-227.
bool boxBox(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 pA{ w.wX[a], w.wY[a] }, pB{ w.wX[b], w.wY[b] };
    const float cosA = w.wCos[a], sinA = w.wSin[a], cosB = w.wCos[b], sinB = w.wSin[b];
    const float hwA = w.p0[a], hhA = w.p1[a], hwB = w.p0[b], hhB = w.p1[b];
    const V2 rightA{ cosA, sinA }, upA{ -sinA, cosA }, rightB{ cosB, sinB }, upB{ -sinB, cosB };
    const float relCos = cosA * cosB + sinA * sinB;
    const float relSin = cosA * sinB - sinA * cosB;
    const float arc = std::fabs(relCos), ars = std::fabs(relSin);
    const V2 cd = pB - pA;
    const float hwv[4] = { hwB, hwB, hwA, hwA };
    const float hhv[4] = { hhB, hhB, hhA, hhA };
    const float hwh[4] = { hwA, hhA, hwB, hhB };
    const float cs1[4] = { arc, ars, arc, ars };
    const float cs2[4] = { ars, arc, ars, arc };
    const float ax[4] = { cosA, -sinA, cosB, -sinB };
    const float ay[4] = { sinA, cosA, sinB, cosB };
    float ov[4], cdo[4];
    for (int k = 0; k < 4; k++) {
        const float rs = hwv[k] * cs1[k] + hhv[k] * cs2[k] + hwh[k];
        cdo[k] = cd.x * ax[k] + cd.y * ay[k];
        ov[k] = rs - std::fabs(cdo[k]);
    }
    for (int k = 0; k < 4; k++) if (ov[k] < 0.0f) return false;
    const float depth = std::fmin(std::fmin(ov[0], ov[1]), std::fmin(ov[2], ov[3]));
    const uint32_t n0 = ov[0] != depth;
    const uint32_t n01 = n0 & (uint32_t)(ov[1] != depth);
    const uint32_t n012 = n01 & (uint32_t)(ov[2] != depth);
    const uint32_t best = n0 + n01 + n012;
    V2 normal{ ax[best], ay[best] };
    flipSignIfNegative(normal, cdo[best]);
    const bool refIsA = best < 2;
    V2 pRef, pInc, rRef, uRef, rInc, uInc, refNormal;
    float hwRef, hhRef, hwInc, hhInc;
    if (refIsA) { pRef = pA; hwRef = hwA; hhRef = hhA; rRef = rightA; uRef = upA; pInc = pB; hwInc = hwB; hhInc = hhB; rInc = rightB; uInc = upB; refNormal = normal; }
    else { pRef = pB; hwRef = hwB; hhRef = hhB; rRef = rightB; uRef = upB; pInc = pA; hwInc = hwA; hhInc = hhA; rInc = rightA; uInc = upA; refNormal = -normal; }
    const float sign = std::copysign(1.0f, refIsA ? cdo[best] : -cdo[best]);
    V2 refStart, refEnd, refCenter, side;
    {
        const bool useX = best == 0 || best == 2;
        if (useX) { refCenter = pRef + rRef * (sign * hwRef); side = uRef; }
        else { refCenter = pRef + uRef * (sign * hhRef); side = rRef; }
        const V2 off = side * (useX ? hhRef : hwRef);
        refStart = refCenter + off;
        refEnd = refCenter - off;
    }
    V2 incStart, incEnd;
    {
        const bool axisIsUp = (best & 1u) != 0;
        const float sideTerm = (best == 0 || best == 3) ? -relSin : relSin;
        const float dotX = sign * (axisIsUp ? sideTerm : relCos);
        const float dotY = sign * (axisIsUp ? relCos : sideTerm);
        const bool useX = std::fabs(dotX) > std::fabs(dotY);
        const float incSign = -std::copysign(1.0f, useX ? dotX : dotY);
        V2 fc, off;
        if (useX) { fc = pInc + rInc * (incSign * hwInc); off = uInc * hhInc; }
        else { fc = pInc + uInc * (incSign * hhInc); off = rInc * hwInc; }
        incStart = fc + off;
        incEnd = fc - off;
    }
    V2 cl[2] = { incStart, incEnd };
    if (clipSegment(cl[0], cl[1], refStart, -side)) return false;
    if (clipSegment(cl[0], cl[1], refEnd, side)) return false;
    V2 pts[2] = { { 0, 0 }, { 0, 0 } };
    uint32_t ids[2] = { 0, 0 };
    uint32_t count = 0;
    const float refPlane = dot(refCenter, refNormal);
    const uint32_t idBase = best << 1;
    for (uint32_t i = 0; i < 2; i++) {
        if (dot(cl[i], refNormal) <= refPlane + SAT_EPS) { pts[count] = cl[i]; ids[count] = idBase | i; count++; }
    }
    if (count == 0) return false;
    out = makeManifold(a, b, normal, depth, count, pts[0], pts[1], ids[0], ids[1]);
    return true;
}

// CollisionMethods/PolygonPolygon.cpp:18-235.
bool polygonPolygon(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 pA{ w.wX[a], w.wY[a] }, pB{ w.wX[b], w.wY[b] };
    const float cosA = w.wCos[a], sinA = w.wSin[a], cosB = w.wCos[b], sinB = w.wSin[b];
    const uint32_t countA = w.polyCount[a], countB = w.polyCount[b];
    const V2* vA = &w.verts[w.polyStart[a]];
    const V2* nA = &w.normals[w.polyStart[a]];
    const V2* vB = &w.verts[w.polyStart[b]];
    const V2* nB = &w.normals[w.polyStart[b]];
    const V2 rightA{ cosA, sinA }, upA{ -sinA, cosA };
    const V2 cd = pB - pA;
    const V2 dl{ dot(cd, rightA), dot(cd, upA) };
    const float relCos = cosA * cosB + sinA * sinB;
    const float relSin = cosA * sinB - sinA * cosB;
    V2 vBA[16], nBA[16];
    for (uint32_t i = 0; i < countB; i++) {
        const V2 v = vB[i];
        vBA[i] = { dl.x + relCos * v.x - relSin * v.y, dl.y + relSin * v.x + relCos * v.y };
    }
    V2 nl{ 0, 0 };
    float depth = FLT_MAX;
    bool refIsA = true;
    uint32_t bestIdx = 0;
    float bestFacing = FLT_MAX;
    for (uint32_t i = 0; i < countA; i++) {
        const V2 axis = nA[i];
        float minB = FLT_MAX;
        for (uint32_t j = 0; j < countB; j++) minB = std::fmin(minB, dot(vBA[j], axis));
        const float overlap = dot(vA[i], axis) - minB;
        if (overlap < 0.0f) return false;
        const float facing = -dot(axis, dl);
        if (overlap > depth) continue;
        if (overlap < depth || facing < bestFacing) { depth = overlap; nl = axis; refIsA = true; bestIdx = i; bestFacing = facing; }
    }
    for (uint32_t i = 0; i < countB; i++) {
        const V2 nb = nB[i];
        const V2 axis{ relCos * nb.x - relSin * nb.y, relSin * nb.x + relCos * nb.y };
        nBA[i] = axis;
        float minA = FLT_MAX;
        for (uint32_t j = 0; j < countA; j++) minA = std::fmin(minA, dot(vA[j], axis));
        const float overlap = dot(vBA[i], axis) - minA;
        if (overlap < 0.0f) return false;
        const float facing = dot(axis, dl);
        if (overlap > depth) continue;
        if (overlap < depth || facing < bestFacing) { depth = overlap; nl = -axis; refIsA = false; bestIdx = i; bestFacing = facing; }
    }
    const V2* refV = refIsA ? vA : vBA;
    const uint32_t refCount = refIsA ? countA : countB;
    const V2* incV = refIsA ? vBA : vA;
    const uint32_t incCount = refIsA ? countB : countA;
    const V2 refN = refIsA ? nl : -nl;
    const uint32_t i0 = bestIdx, i1 = nextIdx(i0, refCount);
    const V2 refStart = refV[i0], refEnd = refV[i1];
    const V2 refCenter = (refStart + refEnd) * 0.5f;
    const V2 edge = refEnd - refStart;
    const float elen2 = dot(edge, edge);
    const V2 side = edge * (1.0f / std::sqrt(elen2));
    uint32_t inc = 0;
    float minDot = FLT_MAX;
    for (uint32_t i = 0; i < incCount; i++) {
        const V2 nLocal = refIsA ? nBA[i] : nA[i];
        const float d = dot(refN, nLocal);
        if (d < minDot) { minDot = d; inc = i; }
    }
    V2 cl[2] = { incV[inc], incV[nextIdx(inc, incCount)] };
    if (clipSegment(cl[0], cl[1], refStart, side)) return false;
    if (clipSegment(cl[0], cl[1], refEnd, -side)) return false;
    const float refPlane = dot(refCenter, refN);
    const uint32_t tag = refIsA ? REF_TAG : 0u;
    const uint32_t id1 = tag | inc, id2 = tag | nextIdx(inc, incCount);
    V2 pts[2] = { { 0, 0 }, { 0, 0 } };
    uint32_t ids[2] = { 0, 0 };
    uint32_t count = 0;
    const V2 nWorld = rightA * nl.x + upA * nl.y;
    for (uint32_t i = 0; i < 2; i++) {
        if (dot(cl[i], refN) <= refPlane + SAT_EPS) {
            pts[count] = pA + rightA * cl[i].x + upA * cl[i].y;
            ids[count] = i == 0 ? id1 : id2;
            count++;
        }
    }
    if (count == 0) return false;
    out = makeManifold(a, b, nWorld, depth, count, pts[0], pts[1], ids[0], ids[1]);
    return true;
}

// CollisionMethods/BoxPolygon.cpp:14-319 (works in polygon B's frame).
bool boxPolygon(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 pA{ w.wX[a], w.wY[a] }, pB{ w.wX[b], w.wY[b] };
    const float cosA = w.wCos[a], sinA = w.wSin[a], cosB = w.wCos[b], sinB = w.wSin[b];
    const float hwA = w.p0[a], hhA = w.p1[a];
    const V2* vB = &w.verts[w.polyStart[b]];
    const V2* nB = &w.normals[w.polyStart[b]];
    const uint32_t count = w.polyCount[b];
    const V2 rightB{ cosB, sinB }, upB{ -sinB, cosB };
    const float relCos = cosA * cosB + sinA * sinB;
    const float relSin = cosA * sinB - sinA * cosB;
    const V2 rA{ relCos, -relSin }, uA{ relSin, relCos };
    const V2 dBA = pA - pB;
    const V2 bc{ dot(dBA, rightB), dot(dBA, upB) };
    V2 nl;
    float depth;
    int bestType;   // 0 box right, 1 box up, 2 polygon edge
    uint32_t bestIdx = 0;
    float bestFacing;
    float pMinX, pMaxX, pMinY, pMaxY;
    float pcx2, pcy2;
    {
        projectOnAxis(vB, count, rA, pMinX, pMaxX);
        const float bp = dot(bc, rA);
        pMinX -= bp; pMaxX -= bp;
        const float overlap = hwA + std::fmin(-pMinX, pMaxX);
        if (overlap < 0.0f) return false;
        pcx2 = pMinX + pMaxX;
        depth = overlap;
        nl = rA;
        flipSignIfNegative(nl, pcx2);
        bestType = 0;
        bestFacing = -std::fabs(pcx2);
    }
    {
        projectOnAxis(vB, count, uA, pMinY, pMaxY);
        const float bp = dot(bc, uA);
        pMinY -= bp; pMaxY -= bp;
        const float overlap = hhA + std::fmin(-pMinY, pMaxY);
        if (overlap < 0.0f) return false;
        pcy2 = pMinY + pMaxY;
        const float facing = -std::fabs(pcy2);
        if (overlap < depth || (overlap == depth && facing < bestFacing)) {
            depth = overlap; nl = uA; flipSignIfNegative(nl, pcy2); bestType = 1; bestFacing = facing;
        }
    }
    const V2 pcRel = pcx2 * rA + pcy2 * uA;
    for (uint32_t i = 0; i < count; i++) {
        const V2 axis = nB[i];
        const float boxRadius = hwA * std::fabs(dot(axis, rA)) + hhA * std::fabs(dot(axis, uA));
        const float bcp = dot(bc, axis);
        const float overlap = dot(axis, vB[i]) - (bcp - boxRadius);
        if (overlap < 0.0f) return false;
        const float facing = dot(axis, pcRel);
        if (overlap < depth || (overlap == depth && facing < bestFacing)) {
            depth = overlap; nl = -axis; bestType = 2; bestIdx = i; bestFacing = facing;
        }
    }
    const bool refIsBox = bestType != 2;
    V2 refN, refCenter, side, refStart, refEnd, incStart, incEnd;
    uint32_t inc1, inc2;
    if (refIsBox) {
        refN = nl;
        const bool useX = bestType == 0;
        const float sign = useX ? dot(refN, rA) : dot(refN, uA);
        if (useX) { refCenter = bc + rA * (sign * hwA); side = uA; }
        else { refCenter = bc + uA * (sign * hhA); side = rA; }
        const V2 off = side * (useX ? hhA : hwA);
        refStart = refCenter + off;
        refEnd = refCenter - off;
        inc1 = 0;
        float minDot = FLT_MAX;
        for (uint32_t i = 0; i < count; i++) {
            const float d = dot(refN, nB[i]);
            if (d < minDot) { minDot = d; inc1 = i; }
        }
        inc2 = nextIdx(inc1, count);
        incStart = vB[inc1];
        incEnd = vB[inc2];
    } else {
        refN = -nl;
        const uint32_t i0 = bestIdx, i1 = nextIdx(i0, count);
        refStart = vB[i0];
        refEnd = vB[i1];
        refCenter = (refStart + refEnd) * 0.5f;
        const V2 edge = refEnd - refStart;
        const float elen2 = dot(edge, edge);
        side = edge * (-1.0f / std::sqrt(elen2));
        const float dotX = dot(refN, rA);
        const float dotY = dot(refN, uA);
        const bool useX = std::fabs(dotX) > std::fabs(dotY);
        const float sign = -std::copysign(1.0f, useX ? dotX : dotY);
        inc1 = BOX_INCIDENT[useX][sign > 0][0];
        inc2 = BOX_INCIDENT[useX][sign > 0][1];
        V2 fc, off;
        if (useX) { fc = bc + rA * (sign * hwA); off = uA * hhA; }
        else { fc = bc + uA * (sign * hhA); off = rA * hwA; }
        incStart = fc + off;
        incEnd = fc - off;
    }
    V2 cl[2] = { incStart, incEnd };
    if (clipSegment(cl[0], cl[1], refStart, -side)) return false;
    if (clipSegment(cl[0], cl[1], refEnd, side)) return false;
    const uint32_t tag = refIsBox ? REF_TAG : 0u;
    const uint32_t id1 = tag | inc1, id2 = tag | inc2;
    V2 pts[2] = { { 0, 0 }, { 0, 0 } };
    uint32_t ids[2] = { 0, 0 };
    uint32_t cnt = 0;
    const float refPlane = dot(refCenter, refN);
    const V2 nWorld = rightB * nl.x + upB * nl.y;
    for (uint32_t i = 0; i < 2; i++) {
        if (dot(cl[i], refN) <= refPlane + SAT_EPS) {
            pts[cnt] = pB + rightB * cl[i].x + upB * cl[i].y;
            ids[cnt] = i == 0 ? id1 : id2;
            cnt++;
        }
    }
    if (cnt == 0) return false;
    out = makeManifold(a, b, nWorld, depth, cnt, pts[0], pts[1], ids[0], ids[1]);
    return true;
}

// ---------------------------------------------------------------- capsules
constexpr uint32_t CAPSULE_VERTEX_CONTACT_TAG = 0x40000000u;   // Narrow/Constants.h:40-41
constexpr float CAPSULE_PARALLEL_SIN_SQUARED = 1e-3f;
inline float clampValue(float v, float lo, float hi) { return std::fmin(std::fmax(v, lo), hi); }   // HelperMethods.h:111-114

// CollisionMethods/CircleCapsule.cpp:58-147, the batched (8-lane) path for every pair. The
// reference's scalar tail (the last pairs of a bucket that do not fill a batch, :150-203) tests
// the side of a centre lying on the core (distance < 1e-4) with a different sign expression; that
// branch is restated as the batched path has it.
bool circleCapsule(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const float positionAX = w.wX[a], positionAY = w.wY[a];
    const float positionBX = w.wX[b], positionBY = w.wY[b];
    const float cosB = w.wCos[b], sinB = w.wSin[b];
    const float radiusA = w.p0[a], halfLengthB = w.p0[b], radiusB = w.p1[b];
    const float dX = positionAX - positionBX;
    const float dY = positionAY - positionBY;
    const float negHalfLength = 0.0f - halfLengthB;
    const float alongAxis = dX * cosB + dY * sinB;
    const float clampedLow = alongAxis < negHalfLength ? negHalfLength : alongAxis;
    const float t = halfLengthB < clampedLow ? halfLengthB : clampedLow;
    const float deltaX = cosB * t - dX;
    const float deltaY = sinB * t - dY;
    const float radiusSum = radiusA + radiusB;
    const float squaredDistance = deltaX * deltaX + deltaY * deltaY;
    if (!(squaredDistance < radiusSum * radiusSum)) return false;
    const float distance = std::sqrt(squaredDistance);
    const float invDistance = 1.0f / distance;
    const bool isCloseToZero = distance < ZERO_DIV;
    const float sideOffset = dY * cosB - dX * sinB;
    const bool flip = sideOffset < 0.0f;
    const float fallbackX = flip ? -sinB : sinB;
    const float fallbackY = flip ? cosB : -cosB;
    const V2 normal{ isCloseToZero ? fallbackX : deltaX * invDistance, isCloseToZero ? fallbackY : deltaY * invDistance };
    const float depth = radiusSum - distance;
    const V2 positionA{ positionAX, positionAY };
    out = makeManifold(a, b, normal, depth, 1, positionA + normal * radiusA, { 0, 0 }, 0, 0);
    return true;
}

// CollisionMethods/CapsuleCapsule.cpp:25-157.
bool capsuleCapsule(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 positionA{ w.wX[a], w.wY[a] }, positionB{ w.wX[b], w.wY[b] };
    const float halfLengthA = w.p0[a], halfLengthB = w.p0[b], radiusA = w.p1[a], radiusB = w.p1[b];
    const V2 centerDelta = positionB - positionA;
    const float radiusSum = radiusA + radiusB;
    const float reach = halfLengthA + halfLengthB + radiusSum;
    if (dot(centerDelta, centerDelta) >= reach * reach) return false;
    const V2 axisA{ w.wCos[a], w.wSin[a] }, axisB{ w.wCos[b], w.wSin[b] };
    const float bb = dot(axisA, axisB);
    const float centerAlongA = dot(centerDelta, axisA);
    const float c = -centerAlongA;
    const float f = -dot(axisB, centerDelta);
    const float sinSquared = 1.0f - bb * bb;
    const bool parallel = sinSquared < CAPSULE_PARALLEL_SIN_SQUARED;
    const float extentB = std::fabs(bb) * halfLengthB;
    const float overlapLo = std::fmax(-halfLengthA, centerAlongA - extentB);
    const float overlapHi = std::fmin(halfLengthA, centerAlongA + extentB);
    const bool hasOverlap = parallel && (overlapHi - overlapLo >= ZERO_DIV);
    float s, t;
    if (hasOverlap) {
        s = 0.5f * (overlapLo + overlapHi);
        t = clampValue(bb * s + f, -halfLengthB, halfLengthB);
    } else {
        s = parallel ? 0.0f : clampValue((bb * f - c) / sinSquared, -halfLengthA, halfLengthA);
        t = bb * s + f;
        t = clampValue(t, -halfLengthB, halfLengthB);
        s = clampValue(bb * t - c, -halfLengthA, halfLengthA);
    }
    const V2 pointA = positionA + axisA * s;
    const V2 pointB = positionB + axisB * t;
    const V2 delta = pointB - pointA;
    const float squaredDistance = dot(delta, delta);
    if (squaredDistance >= radiusSum * radiusSum) return false;
    V2 normal;
    float depth;
    if (squaredDistance >= ZERO_DIV_SQ) {
        const float distance = std::sqrt(squaredDistance);
        normal = delta * (1.0f / distance);
        depth = radiusSum - distance;
    } else {
        const V2 sideA{ -axisA.y, axisA.x }, sideB{ -axisB.y, axisB.x };
        const float offsetA = dot(centerDelta, sideA), offsetB = dot(centerDelta, sideB);
        const float depthA = std::fabs(dot(axisB, sideA)) * halfLengthB - std::fabs(offsetA);
        const float depthB = std::fabs(dot(axisA, sideB)) * halfLengthA - std::fabs(offsetB);
        if (depthA <= depthB) { normal = sideA; flipSignIfNegative(normal, offsetA); depth = depthA; }
        else { normal = sideB; flipSignIfNegative(normal, offsetB); depth = depthB; }
        depth += radiusSum;
    }
    V2 contactPoints[2] = { { 0, 0 }, { 0, 0 } };
    uint32_t contactIds[2] = { 0, 0 };
    uint32_t contactCount = 0;
    if (hasOverlap) {
        const V2 sideB{ -axisB.y, axisB.x };
        const float ends[2] = { overlapLo, overlapHi };
        float minGap = FLT_MAX;
        for (uint32_t k = 0; k < 2; k++) {
            const V2 p = positionA + axisA * ends[k];
            const float gap = std::fabs(dot(p - positionB, sideB));
            if (gap >= radiusSum) continue;
            contactPoints[contactCount] = p + normal * radiusA;
            contactIds[contactCount] = k;
            contactCount++;
            minGap = std::fmin(minGap, gap);
        }
        if (contactCount != 0) depth = radiusSum - minGap;
    }
    if (contactCount == 0) {
        contactPoints[0] = pointA + normal * radiusA;
        contactIds[0] = 2;
        contactCount = 1;
    }
    out = makeManifold(a, b, normal, depth, contactCount, contactPoints[0], contactPoints[1], contactIds[0], contactIds[1]);
    return true;
}

// CollisionMethods/BoxCapsule.cpp:26-242.
bool boxCapsule(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2 halfExtents{ w.p0[a], w.p1[a] };
    const V2 positionA{ w.wX[a], w.wY[a] }, rightA{ w.wCos[a], w.wSin[a] };
    const V2 positionB{ w.wX[b], w.wY[b] }, axisB{ w.wCos[b], w.wSin[b] };
    const float halfLengthB = w.p0[b], radius = w.p1[b];
    const V2 upA{ -rightA.y, rightA.x };
    const V2 centerDelta = positionB - positionA;
    const float localX = dot(centerDelta, rightA);
    const float localY = dot(centerDelta, upA);
    const float sx = std::copysign(1.0f, localX);
    const float sy = std::copysign(1.0f, localY);
    const V2 center{ std::fabs(localX), std::fabs(localY) };
    const V2 axis{ (rightA.x * axisB.x + rightA.y * axisB.y) * sx, (rightA.x * axisB.y - rightA.y * axisB.x) * sy };
    const float absAxisX = std::fabs(axis.x), absAxisY = std::fabs(axis.y);
    const float sepX = center.x - halfLengthB * absAxisX - halfExtents.x;
    const float sepY = center.y - halfLengthB * absAxisY - halfExtents.y;
    const float faceSeparation = std::fmax(sepX, sepY);
    if (faceSeparation >= radius) return false;
    const V2 sideAxis{ -axis.y, axis.x };
    const float sideOffset = dot(sideAxis, center);
    const float sideSeparation = std::fabs(sideOffset) - (halfExtents.x * absAxisY + halfExtents.y * absAxisX);
    if (sideSeparation >= radius) return false;
    const bool sideWins = sideSeparation > faceSeparation + SAT_EPS;
    const bool swap = !sideWins && sepY > sepX;
    const V2 ext = swap ? V2{ halfExtents.y, halfExtents.x } : halfExtents;
    const V2 dir = swap ? V2{ axis.y, axis.x } : axis;
    const V2 mid = swap ? V2{ center.y, center.x } : center;
    const V2 basisX = rightA * sx, basisY = upA * sy;
    const V2 worldX = swap ? basisY : basisX, worldY = swap ? basisX : basisY;
    const V2 halfCore = dir * halfLengthB;
    const V2 s0 = mid - halfCore, s1 = mid + halfCore;
    const float length = halfLengthB * 2.0f;
    auto coreDistSquared = [&](const V2& p) {
        const V2 d = s0 + dir * std::fmin(std::fmax(dot(p - s0, dir), 0.0f), length) - p;
        return dot(d, d);
    };
    V2 normal{ 1, 0 };
    float depth = 0.0f;
    uint32_t contactCount = 0;
    V2 point0{ 0, 0 }, point1{ 0, 0 };
    uint32_t id0 = 0, id1 = 0;
    bool faceContact = false, rejected = false, cornerPosX = true, cornerPosY = true;
    V2 vertexNormal{ 1, 0 };
    float vertexSeparation = faceSeparation;
    if (sideWins) {
        vertexNormal = sideOffset >= 0 ? sideAxis : -sideAxis;
        vertexSeparation = sideSeparation;
        cornerPosX = vertexNormal.x >= 0;
        cornerPosY = vertexNormal.y >= 0;
    } else do {
        V2 c0 = s0, c1 = s1;
        float e0 = s0.x - ext.x, e1 = s1.x - ext.x;
        if (faceSeparation > 0) {
            const float minY = std::fmin(s0.y, s1.y), maxY = std::fmax(s0.y, s1.y);
            if (minY > ext.y) break;
            const bool beyondLow = minY < -ext.y, beyondHigh = maxY > ext.y;
            if (beyondLow || beyondHigh) {
                const float invDy = 1.0f / (s1.y - s0.y);
                const float tA = (-ext.y - s0.y) * invDy, tB = (ext.y - s0.y) * invDy;
                const float tLow = std::fmax(0.0f, std::fmin(tA, tB)), tHigh = std::fmin(1.0f, std::fmax(tA, tB));
                const V2 span = s1 - s0;
                c0 = s0 + span * tLow;
                c1 = s0 + span * tHigh;
                e0 = c0.x - ext.x;
                e1 = c1.x - ext.x;
                const float minE = std::fmin(e0, e1);
                const float faceDistSquared = minE * minE;
                if (beyondLow && coreDistSquared(V2{ ext.x, -ext.y }) + SAT_EPS < faceDistSquared) { cornerPosY = false; break; }
                if (beyondHigh && coreDistSquared(ext) + SAT_EPS < faceDistSquared) break;
            }
        }
        const V2 clipped = c1 - c0;
        const bool hasSecond = dot(clipped, clipped) >= ZERO_DIV_SQ;
        const float contactLimit = radius + SAT_EPS;
        const uint32_t faceId = swap ? (sy < 0.0f ? 0u : 2u) : (sx < 0.0f ? 3u : 1u);
        const uint32_t idBase = REF_TAG | (faceId << 1);
        if (e0 <= contactLimit) { point0 = V2{ c0.x - radius, c0.y }; id0 = idBase; contactCount = 1; }
        if (hasSecond && e1 <= contactLimit) {
            const V2 p{ c1.x - radius, c1.y };
            if (contactCount == 0) { point0 = p; id0 = idBase | 1u; }
            else { point1 = p; id1 = idBase | 1u; }
            contactCount++;
        }
        if (contactCount == 0) { rejected = true; break; }
        depth = radius - std::fmin(e0, e1);
        faceContact = true;
    } while (false);
    if (rejected) return false;
    if (!faceContact) {
        const V2 wc{ cornerPosX ? ext.x : -ext.x, cornerPosY ? ext.y : -ext.y };
        const V2 q = s0 + dir * std::fmin(std::fmax(dot(wc - s0, dir), 0.0f), length);
        normal = vertexNormal;
        depth = radius - vertexSeparation;
        if (vertexSeparation > 0.0f) {
            const V2 d = q - wc;
            const float distSquared = dot(d, d);
            if (distSquared >= radius * radius) return false;
            if (distSquared >= ZERO_DIV_SQ) {
                const float distance = std::sqrt(distSquared);
                normal = d * (1.0f / distance);
                depth = radius - distance;
            }
        }
        const bool foldedPosX = swap ? cornerPosY : cornerPosX;
        const bool foldedPosY = swap ? cornerPosX : cornerPosY;
        const bool realPosX = foldedPosX != (sx < 0.0f);
        const bool realPosY = foldedPosY != (sy < 0.0f);
        const uint32_t vertexId = (realPosY ? 2u : 0u) | ((realPosX != realPosY) ? 1u : 0u);
        contactCount = 1;
        point0 = q - normal * radius;
        id0 = CAPSULE_VERTEX_CONTACT_TAG | vertexId;
    }
    normal = worldX * normal.x + worldY * normal.y;
    point0 = positionA + worldX * point0.x + worldY * point0.y;
    if (contactCount > 1) point1 = positionA + worldX * point1.x + worldY * point1.y;
    out = makeManifold(a, b, normal, depth, contactCount, point0, point1, id0, id1);
    return true;
}

// CollisionMethods/PolygonCapsule.cpp:27-240.
bool polygonCapsule(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    const V2* vertsPtr = &w.verts[w.polyStart[a]];
    const V2* normalsPtr = &w.normals[w.polyStart[a]];
    const uint32_t vertexCount = w.polyCount[a];
    const V2 positionA{ w.wX[a], w.wY[a] }, rightA{ w.wCos[a], w.wSin[a] };
    const V2 positionB{ w.wX[b], w.wY[b] }, axisB{ w.wCos[b], w.wSin[b] };
    const float halfLengthB = w.p0[b], radius = w.p1[b];
    const V2 upA{ -rightA.y, rightA.x };
    const V2 centerDelta = positionB - positionA;
    const V2 centerLocal{ dot(centerDelta, rightA), dot(centerDelta, upA) };
    const V2 axis{ rightA.x * axisB.x + rightA.y * axisB.y, rightA.x * axisB.y - rightA.y * axisB.x };
    const V2 halfCore = axis * halfLengthB;
    const V2 s0 = centerLocal - halfCore, s1 = centerLocal + halfCore;
    const float length = halfLengthB * 2.0f;
    const V2 sideAxis{ -axis.y, axis.x };
    float faceSeparation = -FLT_MAX;
    uint32_t faceIndex = 0;
    float minSide = FLT_MAX, maxSide = -FLT_MAX;
    uint32_t minSideVertex = 0, maxSideVertex = 0;
    for (uint32_t i = 0; i < vertexCount; i++) {
        const V2 v = vertsPtr[i], n = normalsPtr[i];
        const float separation = std::fmin(dot(n, s0 - v), dot(n, s1 - v));
        if (separation >= radius) return false;
        if (separation > faceSeparation) { faceSeparation = separation; faceIndex = i; }
        const float side = dot(v, sideAxis);
        if (side < minSide) { minSide = side; minSideVertex = i; }
        if (side > maxSide) { maxSide = side; maxSideVertex = i; }
    }
    const float sideOffset = dot(sideAxis, centerLocal);
    const float sidePlus = sideOffset - maxSide;
    const float sideMinus = minSide - sideOffset;
    const float sideSeparation = std::fmax(sidePlus, sideMinus);
    if (sideSeparation >= radius) return false;
    auto coreDistSquared = [&](const V2& p) {
        const V2 d = s0 + axis * std::fmin(std::fmax(dot(p - s0, axis), 0.0f), length) - p;
        return dot(d, d);
    };
    V2 normal{ 0, 0 };
    float depth = 0.0f;
    uint32_t contactCount = 0;
    V2 point0{ 0, 0 }, point1{ 0, 0 };
    uint32_t id0 = 0, id1 = 0;
    bool faceContact = false;
    uint32_t vertexId = 0;
    V2 vertexNormal{ 0, 0 };
    float vertexSeparation = 0.0f;
    do {
        if (sideSeparation > faceSeparation + SAT_EPS) {
            const bool onPlusSide = sidePlus >= sideMinus;
            vertexId = onPlusSide ? maxSideVertex : minSideVertex;
            vertexNormal = onPlusSide ? sideAxis : -sideAxis;
            vertexSeparation = sideSeparation;
            break;
        }
        const uint32_t i0 = faceIndex, i1 = nextIdx(i0, vertexCount);
        const V2 va = vertsPtr[i0], vb = vertsPtr[i1], n = normalsPtr[i0];
        const V2 tangent{ -n.y, n.x };
        vertexNormal = n;
        vertexSeparation = faceSeparation;
        V2 c0 = s0, c1 = s1;
        float e0 = dot(s0 - va, n), e1 = dot(s1 - va, n);
        if (faceSeparation > 0.0f) {
            const float edgeLength = dot(vb - va, tangent);
            const float t0 = dot(s0 - va, tangent), t1 = dot(s1 - va, tangent);
            const bool beyondStart = std::fmin(t0, t1) < 0.0f;
            const bool beyondEnd = std::fmax(t0, t1) > edgeLength;
            if (beyondStart || beyondEnd) {
                if (clipSegment(c0, c1, va, tangent)) { vertexId = i0; break; }
                if (clipSegment(c0, c1, vb, -tangent)) { vertexId = i1; break; }
                e0 = dot(c0 - va, n);
                e1 = dot(c1 - va, n);
                const float minE = std::fmin(e0, e1);
                const float faceDistSquared = minE * minE;
                if (beyondStart && coreDistSquared(va) + SAT_EPS < faceDistSquared) { vertexId = i0; break; }
                if (beyondEnd && coreDistSquared(vb) + SAT_EPS < faceDistSquared) { vertexId = i1; break; }
            }
        }
        const V2 span = c1 - c0;
        const bool hasSecond = dot(span, span) >= ZERO_DIV_SQ;
        const float contactLimit = radius + SAT_EPS;
        const uint32_t idBase = REF_TAG | (i0 << 1);
        if (e0 <= contactLimit) { point0 = c0 - n * radius; id0 = idBase; contactCount = 1; }
        if (hasSecond && e1 <= contactLimit) {
            const V2 p = c1 - n * radius;
            if (contactCount == 0) { point0 = p; id0 = idBase | 1u; }
            else { point1 = p; id1 = idBase | 1u; }
            contactCount++;
        }
        if (contactCount == 0) return false;
        normal = n;
        depth = radius - std::fmin(e0, e1);
        faceContact = true;
    } while (false);
    if (!faceContact) {
        const V2 wv = vertsPtr[vertexId];
        const V2 q = s0 + axis * std::fmin(std::fmax(dot(wv - s0, axis), 0.0f), length);
        normal = vertexNormal;
        depth = radius - vertexSeparation;
        if (vertexSeparation > 0.0f) {
            const V2 d = q - wv;
            const float distSquared = dot(d, d);
            if (distSquared >= radius * radius) return false;
            if (distSquared >= ZERO_DIV_SQ) {
                const float distance = std::sqrt(distSquared);
                normal = d * (1.0f / distance);
                depth = radius - distance;
            }
        }
        contactCount = 1;
        point0 = q - normal * radius;
        id0 = CAPSULE_VERTEX_CONTACT_TAG | vertexId;
    }
    normal = rightA * normal.x + upA * normal.y;
    point0 = positionA + rightA * point0.x + upA * point0.y;
    if (contactCount > 1) point1 = positionA + rightA * point1.x + upA * point1.y;
    out = makeManifold(a, b, normal, depth, contactCount, point0, point1, id0, id1);
    return true;
}

// SymmetricMatrix.h:14-18 storage index of (typeA <= typeB): CC, CB, BB, CP, BP, PP, CK, BK, PK, KK
// (K: capsule).
inline uint32_t typeSlot(uint32_t lo, uint32_t hi) { return hi * (hi + 1) / 2 + lo; }

bool collide(const World& w, uint32_t a, uint32_t b, Manifold& out) {
    switch (typeSlot(w.colType[a], w.colType[b])) {
    case 0: return circleCircle(w, a, b, out);
    case 1: return circleBox(w, a, b, out);
    case 2: return boxBox(w, a, b, out);
    case 3: return circlePolygon(w, a, b, out);
    case 4: return boxPolygon(w, a, b, out);
    case 5: return polygonPolygon(w, a, b, out);
    case 6: return circleCapsule(w, a, b, out);
    case 7: return boxCapsule(w, a, b, out);
    case 8: return polygonCapsule(w, a, b, out);
    default: return capsuleCapsule(w, a, b, out);
    }
}

}  // namespace

// NarrowPhaseCollisionDetector.cpp:501-516 + :638-690: 1,024-pair chunks, canonical A/B,
// type buckets dispatched in symmetric-matrix order.
void narrowPhase(World& w, const uint32_t* pa, const uint32_t* pb, size_t n, std::vector<Manifold>& out) {
    out.clear();
    std::vector<std::pair<uint32_t, uint32_t>> buckets[10];
    for (size_t start = 0; start < n; start += 1024) {
        const size_t end = std::min(start + 1024, n);
        for (auto& bk : buckets) bk.clear();
        for (size_t i = start; i < end; i++) {
            uint32_t a = pa[i], b = pb[i];
            uint32_t ta = w.colType[a], tb = w.colType[b];
            if (ta > tb) { std::swap(a, b); std::swap(ta, tb); }
            else if (ta == tb && b > a) std::swap(a, b);
            buckets[typeSlot(ta, tb)].push_back({ a, b });
        }
        for (auto& bk : buckets) {
            for (const auto& p : bk) {
                Manifold m;
                if (collide(w, p.first, p.second, m)) out.push_back(m);
            }
        }
    }
}

// ContactCache.h:9-77 + rebuildContactCache (:272-297) + finalizeResults (:692-767).
void warmStartFromCache(const World& w, std::vector<Manifold>& cur) {
    std::unordered_map<uint64_t, uint32_t> cache;
    auto key = [&w](uint32_t a, uint32_t b) {
        const uint64_t sa = w.colSlot[a], sb = w.colSlot[b];
        return (std::min(sa, sb) << 32) | std::max(sa, sb);
    };
    for (uint32_t i = 0; i < w.prevManifolds.size(); i++) {
        const Manifold& p = w.prevManifolds[i];
        const float imp[4] = { p.jn0, p.jt0, p.jn1, p.jt1 };
        bool any = false;
        for (uint32_t k = 0; k < 2 * p.count && k < 4; k++) any |= imp[k] != 0.0f;
        if (any) cache[key(p.a, p.b)] = i;
    }
    for (Manifold& c : cur) {
        c.jn0 = c.jt0 = c.jn1 = c.jt1 = 0.0f;
        if (cache.empty()) continue;
        auto it = cache.find(key(c.a, c.b));
        if (it == cache.end()) continue;
        const Manifold& p = w.prevManifolds[it->second];
        if (p.a != c.a || p.b != c.b) continue;
        const uint32_t ids[2] = { c.id0, c.id1 };
        for (uint32_t i = 0; i < c.count; i++) {
            float jn = 0, jt = 0;
            bool hit = false;
            if (p.count > 0 && ids[i] == p.id0) { jn = p.jn0; jt = p.jt0; hit = true; }
            else if (p.count > 1 && ids[i] == p.id1) { jn = p.jn1; jt = p.jt1; hit = true; }
            if (!hit) continue;
            if (i == 0) { c.jn0 = jn; c.jt0 = jt; } else { c.jn1 = jn; c.jt1 = jt; }
        }
    }
}

}  // namespace orc
