// Host-side scene building (ObjectManager creation semantics), see builder.h.
#include "builder.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace agx {

namespace {
constexpr float kTwoPi = (float)(3.14159265358979323846 * 2.0);
inline Vec sub(Vec a, Vec b) { return { a.x - b.x, a.y - b.y }; }
inline float dot2(Vec a, Vec b) { return a.x * b.x + a.y * b.y; }

// PhysicsGeometry.cpp:28-84 (area moments; centroid and inertia about it).
void polygonMoments(float mass, const std::vector<Vec>& v, float& inertia, Vec& centroid) {
    const size_t n = v.size();
    float area = 0, cx = 0, cy = 0, xx = 0, yy = 0;
    for (size_t i = 0; i < n; i++) {
        const Vec p0 = v[i];
        const Vec p1 = v[i + 1 == n ? 0 : i + 1];
        const float cross = p0.x * p1.y - p1.x * p0.y;
        area += cross;
        cx += (p0.x + p1.x) * cross;
        cy += (p0.y + p1.y) * cross;
        xx += (p0.y * p0.y + p0.y * p1.y + p1.y * p1.y) * cross;
        yy += (p0.x * p0.x + p0.x * p1.x + p1.x * p1.x) * cross;
    }
    area *= 0.5f;
    if (area < FLT_EPSILON) { inertia = 0.0f; centroid = { 0.0f, 0.0f }; return; }
    const float inv6A = 1.0f / (6.0f * area);
    centroid = { cx * inv6A, cy * inv6A };
    inertia = (mass / area) * ((xx + yy) / 12.0f) - mass * dot2(centroid, centroid);
}

// PhysicsGeometry.cpp:173-242: drop near-duplicate neighbours (1e-5), strip straight vertices,
// orient counter-clockwise, reject degenerate outlines. Only convex outlines of at most 16
// vertices are accepted here (convex input is a single piece, PolygonDecomposition.cpp:714-717).
bool sanitizeOutline(const Vec* in, uint32_t count, std::vector<Vec>& out) {
    const float epsSq = 1e-5f * 1e-5f;
    auto close = [epsSq](Vec a, Vec b) { const Vec d = sub(a, b); return dot2(d, d) <= epsSq; };
    out.clear();
    for (uint32_t i = 0; i < count; i++)
        if (out.empty() || !close(out.back(), in[i])) out.push_back(in[i]);
    while (out.size() > 1 && close(out.back(), out.front())) out.pop_back();
    for (bool again = true; again && out.size() >= 3;) {
        again = false;
        const size_t n = out.size();
        std::vector<Vec> keep;
        keep.reserve(n);
        for (size_t i = 0; i < n; i++) {
            const Vec e1 = sub(out[i], out[(i + n - 1) % n]);
            const Vec e2 = sub(out[(i + 1) % n], out[i]);
            const bool straight = dot2(e1, e2) > 0.0f &&
                std::fabs(e1.x * e2.y - e1.y * e2.x) <= 1e-6f * std::sqrt(dot2(e1, e1)) * std::sqrt(dot2(e2, e2));
            if (straight) again = true; else keep.push_back(out[i]);
        }
        out.swap(keep);
    }
    if (out.size() < 3) return false;
    float twiceArea = 0.0f;
    for (size_t i = 0, n = out.size(); i < n; i++) {
        const Vec p0 = out[i], p1 = out[(i + 1) % n];
        twiceArea += p0.x * p1.y - p1.x * p0.y;
    }
    if (twiceArea < 0.0f) std::reverse(out.begin(), out.end());
    if (std::fabs(twiceArea) * 0.5f <= FLT_EPSILON) return false;
    const size_t n = out.size();
    if (n > 16) return false;
    for (size_t i = 0; i < n; i++) {
        const Vec e1 = sub(out[(i + 1) % n], out[i]);
        const Vec e2 = sub(out[(i + 2) % n], out[(i + 1) % n]);
        if (e1.x * e2.y - e1.y * e2.x < 0.0f) return false;
    }
    return true;
}
}  // namespace

// ObjectManager::createConstraint (ObjectManager.cpp:430-475) for host-side scenes; the prismatic
// joint's reference angle as computeReferenceAngle (PrismaticJointConstraintSystem.h:118-125) and the
// axis as sanitizeAxis (ConstraintSystemBase.h:146-158).
bool SceneData::addJoint(JointSpec spec, uint32_t a, uint32_t b) {
    const bool single = spec.type == 4;
    if (spec.type > 4 || a >= bodyCount() || (!single && b >= bodyCount())) return false;
    spec.slotA = bodySlot[a];
    spec.slotB = single ? UINT32_MAX : bodySlot[b];
    const float lenSq = spec.axis.x * spec.axis.x + spec.axis.y * spec.axis.y;
    if (lenSq > 1e-12f) {
        const float len = std::sqrt(lenSq);
        spec.axis = { spec.axis.x / len, spec.axis.y / len };
    } else {
        spec.axis = { 1.0f, 0.0f };
    }
    spec.referenceAngle = 0.0f;
    if (spec.type == 2) {
        const float pi = (float)3.14159265358979323846, twoPi = 2.0f * pi;
        float r = rot[b] - rot[a];
        r -= twoPi * std::floor((r + pi) / twoPi);
        spec.referenceAngle = r;
    }
    joints.push_back(spec);
    return true;
}

void SceneData::swapBodies(uint32_t a, uint32_t b) {
    if (a == b) return;
    for (auto* v : { &comX, &comY, &velX, &velY, &rot, &angVel, &mass, &invMass, &inertia, &invInertia, &cosv, &sinv,
                     &localComX, &localComY, &boundRadius })
        std::swap((*v)[a], (*v)[b]);
    std::swap(bodySlot[a], bodySlot[b]);
    std::swap(bodyColliders[a], bodyColliders[b]);
    for (uint32_t c : bodyColliders[a]) colBody[c] = a;
    for (uint32_t c : bodyColliders[b]) colBody[c] = b;
}

void SceneData::swapColliders(uint32_t a, uint32_t b) {
    if (a == b) return;
    const uint32_t oa = colBody[a], ob = colBody[b];
    auto relink = [this, a, b](uint32_t owner) {
        for (uint32_t& c : bodyColliders[owner]) c = (c == a) ? b : (c == b ? a : c);
    };
    relink(oa);
    if (ob != oa) relink(ob);
    for (auto* v : { &localOffX, &localOffY, &comOffX, &comOffY, &localRot, &wCos, &wSin, &p0, &p1, &minExtent, &maxExtent })
        std::swap((*v)[a], (*v)[b]);
    for (auto* v : { &colBody, &colMaterial, &colSlot, &polyStart, &polyCount }) std::swap((*v)[a], (*v)[b]);
    std::swap(colType[a], colType[b]);
}

// ObjectManager.cpp:1269-1286 + moveBodyToGroup: a new body is appended; a static one then takes
// the place of the first dynamic body, which moves to the end.
uint32_t SceneData::appendBody(Vec pos, Vec vel, float r, float av, float m, float im, float in, float iin, Vec lc) {
    const float c = std::cos(r), s = std::sin(r);
    const uint32_t idx = bodyCount();
    comX.push_back(pos.x + (c * lc.x - s * lc.y));
    comY.push_back(pos.y + (s * lc.x + c * lc.y));
    velX.push_back(vel.x); velY.push_back(vel.y); rot.push_back(r); angVel.push_back(av);
    mass.push_back(m); invMass.push_back(im); inertia.push_back(in); invInertia.push_back(iin);
    cosv.push_back(std::cos(r)); sinv.push_back(std::sin(r));
    localComX.push_back(lc.x); localComY.push_back(lc.y); boundRadius.push_back(0.0f);
    bodySlot.push_back(nextBodySlot++);
    bodyColliders.emplace_back();
    if (im != 0.0f) return idx;
    const uint32_t dst = staticBodies++;
    swapBodies(idx, dst);
    return dst;
}

uint32_t SceneData::appendCollider(uint32_t body, Vec off, float lr, uint32_t material, uint8_t type, float a0, float a1,
                                   uint32_t ps, uint32_t pc) {
    const uint32_t idx = colliderCount();
    float wrapped = std::fmod(lr, kTwoPi);
    if (wrapped < 0.0f) wrapped += kTwoPi;
    colBody.push_back(body);
    colMaterial.push_back(material < materialCount ? material : 0u);
    colSlot.push_back(nextColSlot++);
    polyStart.push_back(ps); polyCount.push_back(pc);
    colType.push_back(type);
    localOffX.push_back(off.x); localOffY.push_back(off.y);
    comOffX.push_back(off.x - localComX[body]); comOffY.push_back(off.y - localComY[body]);
    localRot.push_back(wrapped);
    wCos.push_back(std::cos(wrapped)); wSin.push_back(std::sin(wrapped));
    p0.push_back(a0); p1.push_back(a1);
    minExtent.push_back(0.0f); maxExtent.push_back(0.0f);
    bodyColliders[body].push_back(idx);
    return idx;
}

// ObjectManager.cpp:1103-1108: bound radius, then into the owner's group.
uint32_t SceneData::placeCollider(uint32_t c, float minExt, float maxExt) {
    minExtent[c] = minExt;
    maxExtent[c] = maxExt;
    const uint32_t body = colBody[c];
    const float x = comOffX[c], y = comOffY[c];
    boundRadius[body] = std::fmax(boundRadius[body], std::sqrt(x * x + y * y) + maxExt);
    if (body >= staticBodies) return c;
    const uint32_t dst = staticColliders++;
    swapColliders(c, dst);
    return dst;
}

uint32_t SceneData::createBody(float px, float py, float r, float vx, float vy, float av, float massIn, float inertiaIn,
                               float comXIn, float comYIn) {
    const float m = std::fmax(0.0f, massIn), in = std::fmax(0.0f, inertiaIn);
    return appendBody({ px, py }, { vx, vy }, r, av, m, m == 0.0f ? 0.0f : 1.0f / m, in, in == 0.0f ? 0.0f : 1.0f / in,
                      { comXIn, comYIn });
}

uint32_t SceneData::createCollider(uint32_t body, uint8_t type, float offX, float offY, float localRotIn,
                                   uint32_t material, float a0, float a1) {
    if (body >= bodyCount()) return UINT32_MAX;
    const Vec off{ offX, offY };
    if (type == 0) {   // circle: radius
        const float radius = std::fmax(0.0f, a0);
        return placeCollider(appendCollider(body, off, localRotIn, material, 0, radius, 0.0f, 0, 0), radius, radius);
    }
    if (type == 1) {   // box: width, height
        const float hw = std::fmax(0.0f, a0) * 0.5f, hh = std::fmax(0.0f, a1) * 0.5f;
        return placeCollider(appendCollider(body, off, localRotIn, material, 1, hw, hh, 0, 0), std::fmin(hw, hh),
                             std::fmax(hw, hh));
    }
    if (type == 3) {   // capsule: full core length, radius (axis local +x)
        const float half = std::fmax(0.0f, a0) * 0.5f, rad = std::fmax(0.0f, a1);
        return placeCollider(appendCollider(body, off, localRotIn, material, 3, half, rad, 0, 0), rad, half + rad);
    }
    return UINT32_MAX;
}

bool SceneData::destroyBody(uint32_t body) {
    if (body >= bodyCount()) return false;
    const uint32_t slot = bodySlot[body];
    // joints on the body go with it
    std::vector<JointSpec> kept;
    for (const JointSpec& j : joints)
        if (j.slotA != slot && !(j.type != 4 && j.slotB == slot)) kept.push_back(j);
    joints.swap(kept);
    // its colliders: up through the static group (if there), then to the end, then removed
    while (!bodyColliders[body].empty()) {
        uint32_t c = bodyColliders[body].back();
        if (c < staticColliders) {
            swapColliders(c, staticColliders - 1);
            c = --staticColliders;
        }
        swapColliders(c, colliderCount() - 1);
        bodyColliders[body].pop_back();
        for (auto* v : { &localOffX, &localOffY, &comOffX, &comOffY, &localRot, &wCos, &wSin, &p0, &p1, &minExtent, &maxExtent })
            v->pop_back();
        for (auto* v : { &colBody, &colMaterial, &colSlot, &polyStart, &polyCount }) v->pop_back();
        colType.pop_back();
    }
    // the body: the same way
    if (body < staticBodies) {
        swapBodies(body, staticBodies - 1);
        body = --staticBodies;
    }
    swapBodies(body, bodyCount() - 1);
    for (auto* v : { &comX, &comY, &velX, &velY, &rot, &angVel, &mass, &invMass, &inertia, &invInertia, &cosv, &sinv,
                     &localComX, &localComY, &boundRadius })
        v->pop_back();
    bodySlot.pop_back();
    bodyColliders.pop_back();
    return true;
}

uint32_t SceneData::createCircle(float px, float py, float r, float vx, float vy, float av, float massIn,
                                 uint32_t material, float radiusIn) {
    const float m = std::fmax(0.0f, massIn);
    const float radius = std::fmax(0.0f, radiusIn);
    const float inertiaV = (0.5f * (radius * radius) + 0.0f) * m;
    const uint32_t b = appendBody({ px, py }, { vx, vy }, r, av, m, m == 0.0f ? 0.0f : 1.0f / m, inertiaV,
                                  inertiaV == 0.0f ? 0.0f : 1.0f / inertiaV, { 0.0f, 0.0f });
    placeCollider(appendCollider(b, { 0.0f, 0.0f }, 0.0f, material, 0, radius, 0.0f, 0, 0), radius, radius);
    return b;
}

uint32_t SceneData::createBox(float px, float py, float r, float vx, float vy, float av, float massIn, uint32_t material,
                              float widthIn, float heightIn) {
    const float m = std::fmax(0.0f, massIn);
    const float w = std::fmax(0.0f, widthIn), h = std::fmax(0.0f, heightIn);
    const float inertiaV = m * ((float)(1.0 / 12.0) * (w * w + h * h) + 0.0f);
    const uint32_t b = appendBody({ px, py }, { vx, vy }, r, av, m, m == 0.0f ? 0.0f : 1.0f / m, inertiaV,
                                  inertiaV == 0.0f ? 0.0f : 1.0f / inertiaV, { 0.0f, 0.0f });
    const float hw = w * 0.5f, hh = h * 0.5f;
    placeCollider(appendCollider(b, { 0.0f, 0.0f }, 0.0f, material, 1, hw, hh, 0, 0), std::fmin(hw, hh), std::fmax(hw, hh));
    return b;
}

uint32_t SceneData::createPolygon(float px, float py, float r, float vx, float vy, float av, float massIn,
                                  uint32_t material, const Vec* in, uint32_t count) {
    if (in == nullptr || count < 3) return UINT32_MAX;
    const float m = std::fmax(0.0f, massIn);
    std::vector<Vec> outline;
    if (!sanitizeOutline(in, count, outline)) return UINT32_MAX;
    float inertiaTrue;
    Vec com;
    polygonMoments(m, outline, inertiaTrue, com);
    const Vec needed = sub(com, com);
    for (Vec& v : outline) v = sub(v, com);
    const float inertiaV = inertiaTrue + m * dot2(needed, needed);
    const float c = std::cos(r), s = std::sin(r);
    const Vec origin{ px + (c * com.x - s * com.y), py + (s * com.x + c * com.y) };
    const uint32_t b = appendBody(origin, { vx, vy }, r, av, m, m == 0.0f ? 0.0f : 1.0f / m, inertiaV,
                                  inertiaV == 0.0f ? 0.0f : 1.0f / inertiaV, needed);
    // Collider piece recentred on its own centroid; the shift becomes the local offset.
    float unused;
    Vec cen;
    polygonMoments(1.0f, outline, unused, cen);
    for (Vec& v : outline) v = sub(v, cen);
    const float cl = std::cos(0.0f), sl = std::sin(0.0f);
    const Vec off{ 0.0f + (cl * cen.x - sl * cen.y), 0.0f + (sl * cen.x + cl * cen.y) };
    const uint32_t start = (uint32_t)verts.size(), n = (uint32_t)outline.size();
    float inRadius = FLT_MAX, outSq = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        const Vec e = sub(outline[i + 1 == n ? 0 : i + 1], outline[i]);
        Vec nn{ e.y, -e.x };
        const float lsq = nn.x * nn.x + nn.y * nn.y;
        if (lsq >= 1e-8f) { const float k = 1.0f / std::sqrt(lsq); nn = { nn.x * k, nn.y * k }; }
        else nn = { 0.0f, 0.0f };
        verts.push_back(outline[i]);
        normals.push_back(nn);
        inRadius = std::fmin(inRadius, dot2(nn, outline[i]));
        outSq = std::fmax(outSq, dot2(outline[i], outline[i]));
    }
    placeCollider(appendCollider(b, off, 0.0f, material, 2, 0.0f, 0.0f, start, n), std::fmax(inRadius, 0.0f), std::sqrt(outSq));
    return b;
}

// ObjectManager.cpp:351-389 + :197-230; the mass splits between the core's rectangle and the two
// end caps by area (PhysicsGeometry.cpp:87-117): the rectangle about its centre, each cap about its
// own centroid shifted to the capsule centre.
uint32_t SceneData::createCapsule(float px, float py, float r, float vx, float vy, float av, float massIn,
                                  uint32_t material, float lengthIn, float radiusIn) {
    const float m = std::fmax(0.0f, massIn);
    const float len = std::fmax(0.0f, lengthIn), rad = std::fmax(0.0f, radiusIn);
    const float pi = (float)3.14159265358979323846;
    const float coreArea = len * 2.0f * rad, capArea = pi * rad * rad;
    const float area = coreArea + capArea;
    const float coreMass = m * coreArea / area, capMass = m * capArea / area;
    const float width = 2.0f * rad;
    const float coreInertia = coreMass * (len * len + width * width) / 12.0f;
    const float capCentroid = 4.0f * rad / (3.0f * pi);
    const float capInertia = capMass * (0.5f * rad * rad + 0.25f * len * len + len * capCentroid);
    const float inertiaV = coreInertia + capInertia + m * 0.0f;
    const uint32_t b = appendBody({ px, py }, { vx, vy }, r, av, m, m == 0.0f ? 0.0f : 1.0f / m, inertiaV,
                                  inertiaV == 0.0f ? 0.0f : 1.0f / inertiaV, { 0.0f, 0.0f });
    const float half = len * 0.5f;
    placeCollider(appendCollider(b, { 0.0f, 0.0f }, 0.0f, material, 3, half, rad, 0, 0), rad, half + rad);
    return b;
}

}  // namespace agx
