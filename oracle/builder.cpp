// Oracle builder: restates the reference's object creation (test infrastructure only).
// Dense order follows ObjectManager's group moves so collider indices, and with them
// the canonical A/B choice of the narrow phase, match the reference's.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace orc {

namespace {

constexpr float TWO_PI = (float)(3.14159265358979323846 * 2.0);   // Constants.h:15

enum Group : uint32_t { G_STATIC = 0, G_SLEEPING = 1, G_DYNAMIC = 2, G_UNUSED = 3 };

// GroupRanges.h:24-40: starts[g] .. starts[g + 1] is group g, Unused is starts[3] .. count.
struct Groups {
    uint32_t starts[4] = { 0, 0, 0, 0 };
    uint32_t groupOf(uint32_t index) const {
        for (uint32_t g = 0; g < 3; g++)
            if (index >= starts[g] && index < starts[g + 1]) return g;
        return G_UNUSED;
    }
};

// Per-world builder state kept beside the World (group boundaries of bodies and colliders).
struct BuildState { Groups bodies, colliders; };

// Called right after one element was appended: that element is the only one in Unused.
BuildState& buildState(World& w) {
    // The builder keeps group boundaries in the world's counters: staticBodies and
    // staticColliders are the Static group ends; Sleeping is always empty.
    static thread_local BuildState s;
    s.bodies.starts[0] = 0;
    s.bodies.starts[1] = w.staticBodies;
    s.bodies.starts[2] = w.staticBodies;
    s.bodies.starts[3] = w.bodyCount() - 1;
    s.colliders.starts[0] = 0;
    s.colliders.starts[1] = w.staticColliders;
    s.colliders.starts[2] = w.staticColliders;
    s.colliders.starts[3] = w.colliderCount() - 1;
    return s;
}

// ObjectManager.cpp:977-1003 (moveInGroups). Only downward moves occur at creation.
template <class SwapFn>
uint32_t moveInGroups(Groups& g, uint32_t index, uint32_t target, SwapFn&& swapFn) {
    uint32_t cur = g.groupOf(index);
    while (cur != target) {
        if (cur < target) {
            const uint32_t last = g.starts[cur + 1] - 1;
            swapFn(index, last);
            index = last;
            g.starts[cur + 1]--;
            cur++;
        } else {
            const uint32_t first = g.starts[cur];
            swapFn(index, first);
            index = first;
            g.starts[cur]++;
            cur--;
        }
    }
    return index;
}

void swapBodies(World& w, uint32_t a, uint32_t b) {
    if (a == b) return;
    auto sw = [a, b](auto& v) { std::swap(v[a], v[b]); };
    sw(w.comX); sw(w.comY); sw(w.velX); sw(w.velY); sw(w.rot); sw(w.angVel);
    sw(w.mass); sw(w.invMass); sw(w.inertia); sw(w.invInertia); sw(w.cosv); sw(w.sinv);
    sw(w.localComX); sw(w.localComY); sw(w.boundRadius); sw(w.bodySlot); sw(w.islandId);
    sw(w.bodyColliders);
    for (uint32_t c : w.bodyColliders[a]) w.colBody[c] = a;
    for (uint32_t c : w.bodyColliders[b]) w.colBody[c] = b;
}

void swapColliders(World& w, uint32_t a, uint32_t b) {
    if (a == b) return;
    auto sw = [a, b](auto& v) { std::swap(v[a], v[b]); };
    // ObjectManager::swapColliders: fix the owners' collider lists first.
    auto fix = [&w, a, b](uint32_t owner) {
        for (uint32_t& c : w.bodyColliders[owner]) { if (c == a) c = b; else if (c == b) c = a; }
    };
    const uint32_t oa = w.colBody[a], ob = w.colBody[b];
    fix(oa);
    if (ob != oa) fix(ob);
    sw(w.colBody); sw(w.colMaterial); sw(w.colSlot); sw(w.colType);
    sw(w.localOffX); sw(w.localOffY); sw(w.comOffX); sw(w.comOffY); sw(w.localRot);
    sw(w.wX); sw(w.wY); sw(w.wRot); sw(w.wCos); sw(w.wSin);
    sw(w.minX); sw(w.minY); sw(w.maxX); sw(w.maxY);
    sw(w.p0); sw(w.p1); sw(w.polyStart); sw(w.polyCount); sw(w.minExtent); sw(w.maxExtent);
}

// ObjectManager.cpp:1269-1286 (appendBody) + BodySoA.cpp:6-40 (append).
uint32_t appendBody(World& w, V2 pos, V2 vel, float rot, float av, float mass, float invMass,
                    float inertia, float invInertia, V2 localCom) {
    const float cosR = std::cos(rot);
    const float sinR = std::sin(rot);
    const V2 com = pos + V2{ cosR * localCom.x - sinR * localCom.y, sinR * localCom.x + cosR * localCom.y };
    const uint32_t index = w.bodyCount();
    w.comX.push_back(com.x); w.comY.push_back(com.y);
    w.localComX.push_back(localCom.x); w.localComY.push_back(localCom.y);
    w.velX.push_back(vel.x); w.velY.push_back(vel.y);
    w.rot.push_back(rot); w.angVel.push_back(av);
    w.mass.push_back(mass); w.invMass.push_back(invMass);
    w.inertia.push_back(inertia); w.invInertia.push_back(invInertia);
    w.cosv.push_back(std::cos(rot)); w.sinv.push_back(std::sin(rot));
    w.boundRadius.push_back(0.0f);
    w.bodySlot.push_back(index);                     // handles are never recycled here
    w.islandId.push_back(UINT32_MAX);
    w.bodyColliders.emplace_back();
    BuildState& s = buildState(w);
    const uint32_t target = invMass == 0.0f ? G_STATIC : G_DYNAMIC;
    const uint32_t placed = moveInGroups(s.bodies, index, target, [&w](uint32_t a, uint32_t b) { swapBodies(w, a, b); });
    w.staticBodies = s.bodies.starts[1];
    return placed;
}

// SoA/ColliderSoA.cpp:7-45 (append) + ObjectManager.cpp:522-533 (createColliderInternal).
uint32_t appendCollider(World& w, uint32_t body, V2 localOffset, float localRot, uint32_t material,
                        ShapeType type, float p0, float p1, uint32_t ps, uint32_t pc) {
    const uint32_t index = w.colliderCount();
    w.colBody.push_back(body);
    w.localOffX.push_back(localOffset.x); w.localOffY.push_back(localOffset.y);
    float wrapped = std::fmod(localRot, TWO_PI);
    if (wrapped < 0.0f) wrapped += TWO_PI;
    w.localRot.push_back(wrapped);
    w.wX.push_back(0.0f); w.wY.push_back(0.0f);
    w.wRot.push_back(wrapped);
    w.wCos.push_back(std::cos(wrapped)); w.wSin.push_back(std::sin(wrapped));
    w.colMaterial.push_back(material);
    w.colType.push_back(type);
    w.minX.push_back(0.0f); w.minY.push_back(0.0f); w.maxX.push_back(0.0f); w.maxY.push_back(0.0f);
    w.p0.push_back(p0); w.p1.push_back(p1); w.polyStart.push_back(ps); w.polyCount.push_back(pc);
    w.minExtent.push_back(0.0f); w.maxExtent.push_back(0.0f);
    w.colSlot.push_back(index);
    // refreshComOffset (ObjectManager.cpp:17-22).
    w.comOffX.push_back(localOffset.x - w.localComX[body]);
    w.comOffY.push_back(localOffset.y - w.localComY[body]);
    w.bodyColliders[body].push_back(index);
    w.spatialDirty = true;
    return index;
}

// ObjectManager.cpp:1103-1108 (placeNewCollider) + colliderReach (:30-35).
uint32_t placeNewCollider(World& w, uint32_t c, float minExt, float maxExt) {
    w.minExtent[c] = minExt;
    w.maxExtent[c] = maxExt;
    const uint32_t body = w.colBody[c];
    const float x = w.comOffX[c], y = w.comOffY[c];
    w.boundRadius[body] = std::fmax(w.boundRadius[body], std::sqrt(x * x + y * y) + w.maxExtent[c]);
    BuildState& s = buildState(w);
    const uint32_t target = body < w.staticBodies ? G_STATIC : G_DYNAMIC;
    const uint32_t placed = moveInGroups(s.colliders, c, target, [&w](uint32_t a, uint32_t b) { swapColliders(w, a, b); });
    w.staticColliders = s.colliders.starts[1];
    return placed;
}

uint32_t validMaterial(const World& w, uint32_t m) { return m < w.materials.size() ? m : 0; }

// PhysicsGeometry.cpp:28-84.
void polygonInertia(float mass, const std::vector<V2>& v, float& inertia, V2& centroid) {
    const size_t n = v.size();
    float signedArea = 0, cx = 0, cy = 0, xx = 0, yy = 0;
    for (size_t i = 0; i < n; i++) {
        const V2 p0 = v[i];
        const V2 p1 = v[(i + 1 == n) ? 0 : i + 1];
        const float cross = p0.x * p1.y - p1.x * p0.y;
        signedArea += cross;
        cx += (p0.x + p1.x) * cross;
        cy += (p0.y + p1.y) * cross;
        xx += (p0.y * p0.y + p0.y * p1.y + p1.y * p1.y) * cross;
        yy += (p0.x * p0.x + p0.x * p1.x + p1.x * p1.x) * cross;
    }
    signedArea *= 0.5f;
    if (signedArea < FLT_EPSILON) { inertia = 0.0f; centroid = { 0.0f, 0.0f }; return; }
    const float invSixArea = 1.0f / (6.0f * signedArea);
    centroid = { cx * invSixArea, cy * invSixArea };
    inertia = (mass / signedArea) * ((xx + yy) / 12.0f) - mass * dot(centroid, centroid);
}

// PhysicsGeometry.cpp:173-210 (preparePolygonOutline) + :212-242 (filterDuplicateVertices).
// Convex outlines only: concave decomposition is not on the collide-and-solve path.
bool prepareOutline(const V2* in, uint32_t count, std::vector<V2>& out) {
    const float epsSq = 1e-5f * 1e-5f;
    auto isClose = [epsSq](V2 a, V2 b) { const V2 d = a - b; return dot(d, d) <= epsSq; };
    out.clear();
    for (uint32_t i = 0; i < count; i++)
        if (out.empty() || !isClose(out.back(), in[i])) out.push_back(in[i]);
    while (out.size() > 1 && isClose(out.back(), out.front())) out.pop_back();

    bool removed = true;
    while (removed && out.size() >= 3) {
        removed = false;
        const size_t n = out.size();
        std::vector<V2> kept;
        kept.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            const V2 e1 = out[i] - out[(i + n - 1) % n];
            const V2 e2 = out[(i + 1) % n] - out[i];
            const bool straight = dot(e1, e2) > 0.0f &&
                std::fabs(e1.x * e2.y - e1.y * e2.x) <= 1e-6f * std::sqrt(dot(e1, e1)) * std::sqrt(dot(e2, e2));
            if (straight) removed = true;
            else kept.push_back(out[i]);
        }
        out.swap(kept);
    }
    if (out.size() < 3) return false;
    float area2 = 0.0f;
    for (size_t i = 0, n = out.size(); i < n; ++i) {
        const V2 p0 = out[i];
        const V2 p1 = out[(i + 1) % n];
        area2 += p0.x * p1.y - p1.x * p0.y;
    }
    if (area2 < 0.0f) std::reverse(out.begin(), out.end());
    if (std::fabs(area2) * 0.5f <= FLT_EPSILON) return false;
    // Convexity (PolygonDecomposition.cpp:714-717 returns convex input as one piece).
    const size_t n = out.size();
    for (size_t i = 0; i < n; i++) {
        const V2 e1 = out[(i + 1) % n] - out[i];
        const V2 e2 = out[(i + 2) % n] - out[(i + 1) % n];
        if (e1.x * e2.y - e1.y * e2.x < 0.0f) return false;
    }
    return n <= 16;   // Limits.h:6 (MAX_POLYGON_VERTEX_COUNT)
}

}  // namespace

uint32_t addMaterial(World& w, const Material& m) {
    Material c = m;   // SimulationImpl/MaterialManager.h:31 clamps
    c.elasticity = std::fmin(std::fmax(c.elasticity, 0.0f), 1.0f);
    c.staticFriction = std::fmax(c.staticFriction, 0.0f);
    c.dynamicFriction = std::fmax(c.dynamicFriction, 0.0f);
    if (w.materials.empty()) w.materials.push_back(Material{});
    w.materials.push_back(c);
    return (uint32_t)w.materials.size() - 1;
}

// ObjectManager.cpp:238-262 + :89-112.
uint32_t createCircle(World& w, float px, float py, float rot, float vx, float vy, float av,
                      float massIn, uint32_t material, float radiusIn) {
    if (w.materials.empty()) w.materials.push_back(Material{});
    const float mass = std::fmax(0.0f, massIn);
    const float radius = std::fmax(0.0f, radiusIn);
    const V2 com{ 0.0f, 0.0f };
    const float inertia = (0.5f * (radius * radius) + dot(com, com)) * mass;
    const float invMass = mass == 0.0f ? 0.0f : 1.0f / mass;
    const float invInertia = inertia == 0.0f ? 0.0f : 1.0f / inertia;
    const uint32_t body = appendBody(w, { px, py }, { vx, vy }, rot, av, mass, invMass, inertia, invInertia, com);
    const uint32_t c = appendCollider(w, body, { 0.0f, 0.0f }, 0.0f, validMaterial(w, material), SHAPE_CIRCLE, radius, 0.0f, 0, 0);
    placeNewCollider(w, c, radius, radius);
    return body;
}

// ObjectManager.cpp:264-290 + :114-139.
uint32_t createBox(World& w, float px, float py, float rot, float vx, float vy, float av,
                   float massIn, uint32_t material, float widthIn, float heightIn) {
    if (w.materials.empty()) w.materials.push_back(Material{});
    const float mass = std::fmax(0.0f, massIn);
    const float width = std::fmax(0.0f, widthIn);
    const float height = std::fmax(0.0f, heightIn);
    const V2 com{ 0.0f, 0.0f };
    const float div = (float)(1.0 / 12.0);
    const float inertia = mass * (div * (width * width + height * height) + dot(com, com));
    const float invMass = mass == 0.0f ? 0.0f : 1.0f / mass;
    const float invInertia = inertia == 0.0f ? 0.0f : 1.0f / inertia;
    const uint32_t body = appendBody(w, { px, py }, { vx, vy }, rot, av, mass, invMass, inertia, invInertia, com);
    const float hw = width * 0.5f, hh = height * 0.5f;
    const uint32_t c = appendCollider(w, body, { 0.0f, 0.0f }, 0.0f, validMaterial(w, material), SHAPE_BOX, hw, hh, 0, 0);
    placeNewCollider(w, c, std::fmin(hw, hh), std::fmax(hw, hh));
    return body;
}

// ObjectManager.cpp:292-349 + PhysicsGeometry.cpp:259-298 + ObjectManager.cpp:535-562.
uint32_t createPolygon(World& w, float px, float py, float rot, float vx, float vy, float av,
                       float massIn, uint32_t material, const V2* localVerts, uint32_t count) {
    if (w.materials.empty()) w.materials.push_back(Material{});
    if (localVerts == nullptr || count < 3) return UINT32_MAX;
    const float mass = std::fmax(0.0f, massIn);
    std::vector<V2> outline;
    if (!prepareOutline(localVerts, count, outline)) return UINT32_MAX;

    float inertiaAboutTrueCom;
    V2 trueCom;
    polygonInertia(mass, outline, inertiaAboutTrueCom, trueCom);
    const V2 neededCom = trueCom - trueCom;
    for (V2& v : outline) v = v - trueCom;
    const float inertia = inertiaAboutTrueCom + mass * dot(neededCom, neededCom);

    const float cosR = std::cos(rot);
    const float sinR = std::sin(rot);
    const V2 origin = V2{ px, py } + V2{ cosR * trueCom.x - sinR * trueCom.y, sinR * trueCom.x + cosR * trueCom.y };
    const float invMass = mass == 0.0f ? 0.0f : 1.0f / mass;
    const float invInertia = inertia == 0.0f ? 0.0f : 1.0f / inertia;
    const uint32_t body = appendBody(w, origin, { vx, vy }, rot, av, mass, invMass, inertia, invInertia, neededCom);

    // createConvexPolygonColliderInternal: recentre the piece on its centroid.
    float unusedInertia;
    V2 centroid;
    polygonInertia(1.0f, outline, unusedInertia, centroid);
    for (V2& v : outline) v = v - centroid;
    const float cosL = std::cos(0.0f), sinL = std::sin(0.0f);
    const V2 localOffset = V2{ 0.0f, 0.0f } + V2{ cosL * centroid.x - sinL * centroid.y, sinL * centroid.x + cosL * centroid.y };

    // PolygonGeometry (SoA/Shapes/PolygonSoA.cpp:14-30): outward unit edge normals.
    const uint32_t start = (uint32_t)w.verts.size();
    const uint32_t n = (uint32_t)outline.size();
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t next = (i + 1 == n) ? 0 : i + 1;
        const V2 edge = outline[next] - outline[i];
        V2 normal{ edge.y, -edge.x };
        const float lenSquared = normal.x * normal.x + normal.y * normal.y;
        if (lenSquared >= 1e-8f) normal = normal * (1.0f / std::sqrt(lenSquared));
        else normal = { 0.0f, 0.0f };
        w.verts.push_back(outline[i]);
        w.normals.push_back(normal);
    }
    // PhysicsGeometry.cpp:146-171 (in/out radius).
    float inRadius = FLT_MAX, maxRadiusSq = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        inRadius = std::fmin(inRadius, dot(w.normals[start + i], w.verts[start + i]));
        maxRadiusSq = std::fmax(maxRadiusSq, dot(w.verts[start + i], w.verts[start + i]));
    }
    const uint32_t c = appendCollider(w, body, localOffset, 0.0f, validMaterial(w, material), SHAPE_POLYGON, 0.0f, 0.0f, start, n);
    placeNewCollider(w, c, std::fmax(inRadius, 0.0f), std::sqrt(maxRadiusSq));
    return body;
}

// PhysicsGeometry.cpp:87-117, with the centre of mass at the capsule centre.
float capsuleInertia(float mass, float length, float radius) {
    const float PI = (float)3.14159265358979323846;   // Constants.h:8
    const float rectArea = length * 2.0f * radius;
    const float circleArea = PI * radius * radius;
    const float totalArea = rectArea + circleArea;
    const float rectMass = mass * rectArea / totalArea;
    const float capMass = mass * circleArea / totalArea;
    const float diameter = 2.0f * radius;
    const float rectInertia = rectMass * (length * length + diameter * diameter) / 12.0f;
    const float d = 4.0f * radius / (3.0f * PI);
    const float capInertia = capMass * (0.5f * radius * radius + 0.25f * length * length + length * d);
    const V2 com{ 0.0f, 0.0f };
    const float deltaSquared = dot(com, com);
    return rectInertia + capInertia + mass * deltaSquared;
}

// ObjectManager.cpp:351-389 + :197-230 (createCapsuleCollider with zero offset and rotation).
uint32_t createCapsule(World& w, float px, float py, float rot, float vx, float vy, float av,
                       float massIn, uint32_t material, float lengthIn, float radiusIn) {
    if (w.materials.empty()) w.materials.push_back(Material{});
    const float length = std::fmax(0.0f, lengthIn);
    const float mass = std::fmax(0.0f, massIn);
    const float radius = std::fmax(0.0f, radiusIn);
    const V2 com{ 0.0f, 0.0f };
    const float inertia = capsuleInertia(mass, length, radius);
    const float invMass = mass == 0.0f ? 0.0f : 1.0f / mass;
    const float invInertia = inertia == 0.0f ? 0.0f : 1.0f / inertia;
    const uint32_t body = appendBody(w, { px, py }, { vx, vy }, rot, av, mass, invMass, inertia, invInertia, com);
    const uint32_t c = appendCollider(w, body, { 0.0f, 0.0f }, 0.0f, validMaterial(w, material), SHAPE_CAPSULE,
                                      length * 0.5f, radius, 0, 0);
    placeNewCollider(w, c, radius, length * 0.5f + radius);
    return body;
}

}  // namespace orc
