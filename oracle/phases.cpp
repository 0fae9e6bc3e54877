// Oracle step phases (test infrastructure only): integration, poses/AABBs, CCD,
// islands and the step driver. Narrow phase, broad phase and solver live in
// narrow.cpp, broad.cpp and solver.cpp.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace orc {

namespace {
constexpr double FULL_PI = 3.14159265358979323846;           // Constants.h:8-14
constexpr float HALF_PI = (float)(FULL_PI * 0.5);
constexpr float INV_HALF_PI = (float)(1.0 / (FULL_PI * 0.5));
constexpr float TWO_PI = (float)(FULL_PI * 2.0);
constexpr float INV_TWO_PI = (float)(1.0 / (FULL_PI * 2.0));
// FastCosSin.cpp:10-14 (the reference's own polynomial data).
constexpr float C0 = (float)9.9999999999953215e-01;
constexpr float C1 = (float)-1.6666666534403832e-01;
constexpr float C2 = (float)8.3332302391869520e-03;
constexpr float C3 = (float)-1.9818200679665358e-04;
constexpr float C4 = (float)2.6182346015248291e-06;
}  // namespace

// FastCosSin.cpp:44-75 (8-lane path): truncated quadrant, mirrored angle, degree-9 odd
// polynomial, sin clamped to 1, cos = sqrt(1 - sin^2); signs from quadrant bit 1.
void fastCosSin(float x, float& c, float& s) {
    const int q = (int)(x * INV_HALF_PI);
    const float u = x - (float)q * HALF_PI;
    const float mirror = (float)(q & 1);
    const float a = mirror * (HALF_PI - 2.0f * u) + u;
    const float a2 = a * a;
    float sm = C4;
    sm = sm * a2 + C3;
    sm = sm * a2 + C2;
    sm = sm * a2 + C1;
    sm = sm * a2 + C0;
    sm = sm * a;
    sm = std::fmin(sm, 1.0f);
    const float cm = std::sqrt(1.0f - sm * sm);
    s = (q & 2) ? -sm : sm;
    c = ((q + 1) & 2) ? -cm : cm;
}

// LeftOutFunctions.cpp:30-43 (8-lane path, fused negMulAdd).
float wrapRotation(float r) {
    const float wrapCount = std::trunc(r * INV_TWO_PI);
    r = std::fma(-wrapCount, TWO_PI, r);
    if (r < 0.0f) r += TWO_PI;
    return r;
}

// CCD/ContinuousCollisionDetector.cpp:276-299.
void snapshotPoses(World& w) {
    const uint32_t n = w.bodyCount(), c = w.colliderCount();
    w.prevComX.resize(n); w.prevComY.resize(n);
    w.prevWX.resize(c); w.prevWY.resize(c);
    for (uint32_t i = w.staticBodies; i < n; i++) { w.prevComX[i] = w.comX[i]; w.prevComY[i] = w.comY[i]; }
    for (uint32_t i = w.staticColliders; i < c; i++) { w.prevWX[i] = w.wX[i]; w.prevWY[i] = w.wY[i]; }
}

// Integrator.cpp:99-209 (velocity Verlet; angular damping is a no-op at its default 1.0,
// Integrator.cpp:214-217, and the mouse holder is not on this path), joints between the halves.
void integrate(World& w, float dt) {
    const float halfDT = dt * 0.5f;
    const float gx = w.settings.gravityX * halfDT, gy = w.settings.gravityY * halfDT;
    for (uint32_t i = w.staticBodies; i < w.bodyCount(); i++) {
        if (asleep(w, i)) continue;   // the Dynamic range only (sleeping bodies are not in it)
        const float movable = w.invMass[i] != 0.0f ? 1.0f : 0.0f;
        w.velX[i] += gx * movable;
        w.velY[i] += gy * movable;
    }
    solveJoints(w, dt);   // Simulation.cpp:384-405: constraints between the two halves
    for (uint32_t i = w.staticBodies; i < w.bodyCount(); i++) {
        if (asleep(w, i)) continue;
        w.comX[i] = w.velX[i] * dt + w.comX[i];
        w.comY[i] = w.velY[i] * dt + w.comY[i];
        w.rot[i] = w.angVel[i] * dt + w.rot[i];
        const float movable = w.invMass[i] != 0.0f ? 1.0f : 0.0f;
        w.velX[i] += gx * movable;
        w.velY[i] += gy * movable;
    }
}

// Simulation.cpp:291-292.
void wrapAndCosSin(World& w) {
    for (uint32_t i = w.staticBodies; i < w.bodyCount(); i++) {
        w.rot[i] = wrapRotation(w.rot[i]);
        fastCosSin(w.rot[i], w.cosv[i], w.sinv[i]);
    }
}

// Simulation.cpp:673-736.
void colliderPoses(World& w, bool all) {
    const uint32_t begin = all ? 0 : w.staticColliders;
    for (uint32_t i = begin; i < w.colliderCount(); i++) {
        const uint32_t b = w.colBody[i];
        const float offX = w.comOffX[i], offY = w.comOffY[i];
        const float bc = w.cosv[b], bs = w.sinv[b];
        w.wX[i] = w.comX[b] + (offX * bc - offY * bs);
        w.wY[i] = w.comY[b] + (offX * bs + offY * bc);
        w.wRot[i] = wrapRotation(w.rot[b] + w.localRot[i]);
        fastCosSin(w.wRot[i], w.wCos[i], w.wSin[i]);
    }
}

// ColliderAABBBuilder.cpp:8-175.
void colliderAABBs(World& w, bool all) {
    const uint32_t begin = all ? 0 : w.staticColliders;
    for (uint32_t i = begin; i < w.colliderCount(); i++) {
        const float x = w.wX[i], y = w.wY[i];
        if (w.colType[i] == SHAPE_CIRCLE) {
            const float r = w.p0[i];
            w.minX[i] = x - r; w.minY[i] = y - r; w.maxX[i] = x + r; w.maxY[i] = y + r;
        } else if (w.colType[i] == SHAPE_CAPSULE) {   // :138-175
            const float ex = std::fabs(w.wCos[i]) * w.p0[i] + w.p1[i];
            const float ey = std::fabs(w.wSin[i]) * w.p0[i] + w.p1[i];
            w.minX[i] = x - ex; w.minY[i] = y - ey; w.maxX[i] = x + ex; w.maxY[i] = y + ey;
        } else if (w.colType[i] == SHAPE_BOX) {
            const float ac = std::fabs(w.wCos[i]), as = std::fabs(w.wSin[i]);
            const float ex = ac * w.p0[i] + as * w.p1[i];
            const float ey = as * w.p0[i] + ac * w.p1[i];
            w.minX[i] = x - ex; w.minY[i] = y - ey; w.maxX[i] = x + ex; w.maxY[i] = y + ey;
        } else {
            const float c = w.wCos[i], s = w.wSin[i];
            float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
            for (uint32_t v = 0; v < w.polyCount[i]; v++) {
                const V2 p = w.verts[w.polyStart[i] + v];
                const float wx = c * p.x - s * p.y;
                const float wy = s * p.x + c * p.y;
                mnx = std::fmin(mnx, wx); mxx = std::fmax(mxx, wx);
                mny = std::fmin(mny, wy); mxy = std::fmax(mxy, wy);
            }
            w.minX[i] = x + mnx; w.minY[i] = y + mny; w.maxX[i] = x + mxx; w.maxY[i] = y + mxy;
        }
    }
}

// ---------------------------------------------------------------- CCD
namespace {
constexpr float NO_HIT = 2.0f;

// CCD/ContinuousCollisionDetector.cpp:48-61.
float castCircle(float mx, float my, float dx, float dy, float dd, float radius, float best) {
    const float c = mx * mx + my * my - radius * radius;
    const float b = mx * dx + my * dy;
    if (c <= 0.0f || b >= 0.0f) return NO_HIT;
    if (c >= -2.0f * b * best) return NO_HIT;
    const float disc = b * b - dd * c;
    if (disc < 0.0f) return NO_HIT;
    return c / (std::sqrt(disc) - b);
}

// :63-106.
float castBox(V2 q0, V2 dl, float dd, float hw, float hh, float radius, float best) {
    const float sx = std::copysign(1.0f, q0.x), sy = std::copysign(1.0f, q0.y);
    const float qx = q0.x * sx, qy = q0.y * sy;
    const float vx = dl.x * sx, vy = dl.y * sy;
    const float w = hw + radius, h = hh + radius;
    const float gapX = qx - w, gapY = qy - h;
    if (gapX <= 0.0f && gapY <= 0.0f) return NO_HIT;
    float tx = 0.0f;
    if (gapX > 0.0f) {
        if (vx >= 0.0f || gapX >= best * -vx) return NO_HIT;
        tx = gapX / -vx;
    }
    float ty = 0.0f;
    if (gapY > 0.0f) {
        if (vy >= 0.0f || gapY >= best * -vy) return NO_HIT;
        ty = gapY / -vy;
    }
    const float t = std::fmax(tx, ty);
    const float px = qx + vx * t, py = qy + vy * t;
    if (tx >= ty ? std::fabs(py) > h : std::fabs(px) > w) return NO_HIT;
    if (std::fabs(px) > hw && std::fabs(py) > hh)
        return castCircle(qx - std::copysign(hw, px), qy - std::copysign(hh, py), vx, vy, dd, radius, best);
    return t;
}

// :128-220.
float castPolygon(V2 q0, V2 dl, float dd, const V2* vertices, const V2* normals, uint32_t count,
                  float radius, float best) {
    float tEnter = 0.0f, tExit = best;
    uint32_t enterEdge = 0;
    float gapArray[16];
    for (uint32_t i = 0; i < count; i++) {
        const V2 n = normals[i];
        const float speed = dot(n, dl);
        const float gap = dot(n, q0 - vertices[i]) - radius;
        gapArray[i] = gap;
        if (speed < 0.0f) {
            if (gap > 0.0f) {
                const float approach = -speed;
                if (gap >= tExit * approach) return NO_HIT;
                const float t = gap / approach;
                if (t > tEnter) { tEnter = t; enterEdge = i; }
            }
        } else if (gap > 0.0f) {
            return NO_HIT;
        } else if (speed > 0.0f) {
            tExit = std::fmin(tExit, -gap / speed);
        }
    }
    if (tEnter >= tExit) return NO_HIT;
    uint32_t edge = enterEdge;
    const bool startedInside = tEnter == 0.0f;
    if (startedInside) {
        float maxGap = -FLT_MAX;
        for (uint32_t i = 0; i < count; i++)
            if (gapArray[i] > maxGap) { maxGap = gapArray[i]; edge = i; }
    }
    const V2 p = q0 + dl * tEnter;
    const uint32_t next = edge + 1 == count ? 0 : edge + 1;
    const V2 a = vertices[edge], b = vertices[next];
    const V2 ab = b - a;
    const float u = dot(p - a, ab);
    if (u >= 0.0f && u <= dot(ab, ab)) return startedInside ? NO_HIT : tEnter;
    const V2 corner = u < 0.0f ? a : b;
    return castCircle(q0.x - corner.x, q0.y - corner.y, dl.x, dl.y, dd, radius, tExit);
}

// :222-240.
float castCapsule(V2 q0, V2 dl, float dd, float halfLength, float capsuleRadius, float radius, float best) {
    const float inflated = capsuleRadius + radius;
    const float nvy = -std::copysign(1.0f, q0.y) * dl.y;
    const float gap = std::fabs(q0.y) - inflated;
    if (nvy > 0.0f && gap >= 0.0f && gap < best * nvy) {
        const float time = gap / nvy;
        if (std::fabs(q0.x + dl.x * time) <= halfLength) return time;
    }
    const float t = std::fmin(castCircle(q0.x + halfLength, q0.y, dl.x, dl.y, dd, inflated, best), best);
    return std::fmin(castCircle(q0.x - halfLength, q0.y, dl.x, dl.y, dd, inflated, t), t);
}

// :243-273.
float castCore(const World& w, uint32_t col, V2 start, V2 d, float dd, float coreRadius, float best) {
    const float cs = w.wCos[col], sn = w.wSin[col];
    const float rx = start.x - w.wX[col], ry = start.y - w.wY[col];
    const V2 q0{ cs * rx + sn * ry, cs * ry - sn * rx };
    const V2 dl{ cs * d.x + sn * d.y, cs * d.y - sn * d.x };
    switch (w.colType[col]) {
    case SHAPE_CIRCLE: return castCircle(q0.x, q0.y, dl.x, dl.y, dd, w.p0[col] + coreRadius, best);
    case SHAPE_BOX: return castBox(q0, dl, dd, w.p0[col], w.p1[col], coreRadius, best);
    case SHAPE_CAPSULE: return castCapsule(q0, dl, dd, w.p0[col], w.p1[col], coreRadius, best);
    default: return castPolygon(q0, dl, dd, &w.verts[w.polyStart[col]], &w.normals[w.polyStart[col]],
                                w.polyCount[col], coreRadius, best);
    }
}
}  // namespace

// CCD/ContinuousCollisionDetector.cpp:307-531. The static set is scanned directly: the
// reference's static BVH only prunes, and every static collider whose leaf AABB strictly
// overlaps the swept AABB is cast, so the earliest time of impact is the same.
bool ccdResolve(World& w, std::vector<float>* hitTimeOut) {
    const uint32_t n = w.bodyCount();
    std::vector<float> hit(n, 1.0f);
    std::vector<uint32_t> hitBodies;
    if (w.staticColliders == 0 || w.staticColliders == w.colliderCount()) {
        if (hitTimeOut) *hitTimeOut = hit;
        return false;
    }
    // Shape ranges are processed circles, boxes, polygons (:495-511); the result per body is a min.
    for (uint32_t col = w.staticColliders; col < w.colliderCount(); col++) {
        const uint32_t body = w.colBody[col];
        const float bdx = w.comX[body] - w.prevComX[body];
        const float bdy = w.comY[body] - w.prevComY[body];
        const float bd2 = bdx * bdx + bdy * bdy;
        if (bd2 <= 1e-8f) continue;
        const float minExtent = w.minExtent[col];
        if (minExtent <= 0.0f) continue;
        const float threshold = w.settings.ccdFastFactor * minExtent;
        const float thresholdSq = threshold * threshold;
        if (bd2 <= thresholdSq) continue;
        const V2 start{ w.prevWX[col], w.prevWY[col] };
        const V2 end{ w.wX[col], w.wY[col] };
        const V2 d = end - start;
        const float dd = dot(d, d);
        if (dd <= thresholdSq) continue;
        const float coreRadius = w.settings.ccdCoreFraction * minExtent;
        const float sMinX = std::fmin(start.x, end.x) - coreRadius;
        const float sMaxX = std::fmax(start.x, end.x) + coreRadius;
        const float sMinY = std::fmin(start.y, end.y) - coreRadius;
        const float sMaxY = std::fmax(start.y, end.y) + coreRadius;
        float best = 1.0f;
        for (uint32_t s = 0; s < w.staticColliders; s++) {
            if (!(sMinX < w.maxX[s] && sMaxX > w.minX[s] && sMinY < w.maxY[s] && sMaxY > w.minY[s])) continue;
            best = std::fmin(best, castCore(w, s, start, d, dd, coreRadius, best));
        }
        if (best < 1.0f) {
            if (hit[body] >= 1.0f) hitBodies.push_back(body);
            hit[body] = std::fmin(hit[body], best);
        }
    }
    if (hitTimeOut) *hitTimeOut = hit;
    if (hitBodies.empty()) return false;
    for (uint32_t body : hitBodies) {
        const float back = hit[body] - 1.0f;
        const float shiftX = (w.comX[body] - w.prevComX[body]) * back;
        const float shiftY = (w.comY[body] - w.prevComY[body]) * back;
        w.comX[body] += shiftX;
        w.comY[body] += shiftY;
        for (uint32_t c : w.bodyColliders[body]) { w.wX[c] += shiftX; w.wY[c] += shiftY; }
    }
    return true;
}

// ---------------------------------------------------------------- islands
// IslandManager.cpp:7-116: DSU over Dynamic bodies, union by size, path halving,
// ids numbered by ascending root index.
void islands(World& w, const std::vector<Manifold>& m) {
    const uint32_t dynBegin = w.staticBodies, dynEnd = w.bodyCount();
    const uint32_t count = dynEnd - dynBegin;
    std::vector<int32_t> parent(count, -1);
    auto find = [&parent](uint32_t x) {
        while (parent[x] >= 0) {
            const uint32_t p = (uint32_t)parent[x];
            const int32_t gp = parent[p];
            if (gp < 0) return p;
            parent[x] = gp;
            x = (uint32_t)gp;
        }
        return x;
    };
    for (uint32_t i = 0; i < w.staticBodies; i++) w.islandId[i] = UINT32_MAX;
    std::vector<uint8_t> member(count, 1);   // IslandManager::reset covers the Dynamic range only
    for (uint32_t i = 0; i < count; i++) member[i] = asleep(w, i + dynBegin) ? 0 : 1;
    for (const Manifold& mf : m) {
        const uint32_t a = w.colBody[mf.a] - dynBegin;
        const uint32_t b = w.colBody[mf.b] - dynBegin;
        if (a >= count || b >= count || !member[a] || !member[b]) continue;
        uint32_t ra = find(a), rb = find(b);
        if (ra == rb) continue;
        if (parent[ra] > parent[rb]) std::swap(ra, rb);
        parent[ra] += parent[rb];
        parent[rb] = (int32_t)ra;
    }
    // the joint systems' active edges (Simulation.cpp:331-336): joints between two dynamic bodies
    for (const Joint& j : w.joints) {
        if (j.d.type == JOINT_MOTOR) continue;
        const uint32_t a = j.d.bodyA - dynBegin, b = j.d.bodyB - dynBegin;
        if (a >= count || b >= count || !member[a] || !member[b]) continue;
        uint32_t ra = find(a), rb = find(b);
        if (ra == rb) continue;
        if (parent[ra] > parent[rb]) std::swap(ra, rb);
        parent[ra] += parent[rb];
        parent[rb] = (int32_t)ra;
    }
    std::vector<uint32_t> rootToId(count);
    w.islandStart.assign(count + 1, 0);
    uint32_t k = 0;
    for (uint32_t i = 0; i < count; i++)
        if (parent[i] < 0 && member[i]) { w.islandStart[k] = (uint32_t)(-parent[i]); rootToId[i] = k; k++; }
    uint32_t sum = 0;
    for (uint32_t j = 0; j < k; j++) { const uint32_t s = w.islandStart[j]; w.islandStart[j] = sum; sum += s; }
    w.islandStart[k] = sum;
    w.islandMembers.assign(sum, 0);
    for (uint32_t i = 0; i < count; i++) {
        if (!member[i]) { w.islandId[i + dynBegin] = UINT32_MAX; continue; }
        const uint32_t id = rootToId[find(i)];
        w.islandId[i + dynBegin] = id;
        w.islandMembers[w.islandStart[id]++] = i + dynBegin;
    }
    for (uint32_t j = k; j > 0; --j) w.islandStart[j] = w.islandStart[j - 1];
    w.islandStart[0] = 0;
    w.islandStart.resize(k + 1);
    w.islandCount = k;
}

// ---------------------------------------------------------------- deferred modifications
namespace {
float wrapAngle(float angle) {   // ObjectManager.cpp:614-619
    float wrapped = std::fmod(angle, TWO_PI);
    if (wrapped < 0.0f) wrapped += TWO_PI;
    return wrapped;
}
}  // namespace

// ObjectManager::flushCommands / applyCommand (ObjectManager.cpp:638-865), body commands only; a
// position or rotation edit makes the next step rebuild every collider's pose (spatialDirty).
void flushCommands(World& w) {
    std::vector<uint32_t> commanded, editedStatic;
    for (const World::BodyCommand& c : w.commands) {
        const uint32_t b = c.body;
        commanded.push_back(b);   // every command wakes its body's island (ObjectManager.cpp:684-693)
        if ((c.kind == 0 || c.kind == 1) && w.invMass[b] == 0.0f) editedStatic.push_back(b);   // markStaticBodyEdited
        switch (c.kind) {
        case 0: {
            const float r = w.rot[b];
            const float cosR = std::cos(r), sinR = std::sin(r);
            const float lx = w.localComX[b], ly = w.localComY[b];
            w.comX[b] = c.a + cosR * lx - sinR * ly;
            w.comY[b] = c.b + sinR * lx + cosR * ly;
            w.spatialDirty = true;
            break;
        }
        case 1: {
            const float r = wrapAngle(c.a);
            w.rot[b] = r;
            w.cosv[b] = std::cos(r);
            w.sinv[b] = std::sin(r);
            w.spatialDirty = true;
            break;
        }
        case 2: w.velX[b] = c.a; w.velY[b] = c.b; break;
        case 3: w.angVel[b] = c.a; break;
        case 4: {   // applyBodyMass (:867-905), dynamic bodies only
            const float newMass = std::fmax(0.0f, c.a), oldMass = w.mass[b];
            if (!(newMass > 0.0f) || !(oldMass > 0.0f)) break;
            const float inertia = w.inertia[b] * (newMass / oldMass);
            w.mass[b] = newMass;
            w.invMass[b] = 1.0f / newMass;
            w.inertia[b] = inertia;
            w.invInertia[b] = inertia > 0.0f ? 1.0f / inertia : 0.0f;
            break;
        }
        case 5: {
            if (w.invMass[b] == 0.0f) break;
            const float inertia = std::fmax(0.0f, c.a);
            w.inertia[b] = inertia;
            w.invInertia[b] = inertia > 0.0f ? 1.0f / inertia : 0.0f;
            break;
        }
        default: break;
        }
    }
    w.commands.clear();
    if (!commanded.empty()) wakeForCommands(w, commanded, editedStatic);
}

// ---------------------------------------------------------------- step
// Simulation.cpp:640-655: poses (and AABBs) of dynamic colliders after the last step.
void postUpdate(World& w) {
    colliderPoses(w, false);
    colliderAABBs(w, false);
    w.postUpdateDone = true;
}

// Simulation.cpp:255-351 (joints and the mouse holder are not on this path).
void step(World& w, float dt, Pool* pool) {
    if (w.bodyCount() == 0) { w.prevManifolds.clear(); w.manifolds.clear(); return; }
    if (w.settings.ccd) snapshotPoses(w);
    integrate(w, dt);
    wrapAndCosSin(w);
    const bool rebuild = w.spatialDirty;
    colliderPoses(w, rebuild);
    colliderAABBs(w, rebuild);
    w.spatialDirty = false;
    if (w.settings.ccd && w.staticColliders > 0) {
        if (ccdResolve(w, nullptr)) colliderAABBs(w, false);
    }
    w.manifolds.clear();
    if (w.colliderCount() > 1) {
        broadPhaseBvh(w);
        if (w.pairA.empty()) {
            w.prevManifolds.clear();
        } else {
            if (pool) narrowPhaseThreaded(w, w.pairA.data(), w.pairB.data(), w.pairA.size(), w.manifolds, *pool);
            else narrowPhase(w, w.pairA.data(), w.pairB.data(), w.pairA.size(), w.manifolds);
            warmStartFromCache(w, w.manifolds);
            wakeContacts(w, w.manifolds);   // Simulation.cpp:428-434, before the solve
            if (!w.manifolds.empty()) {
                collisionEvents(w, w.manifolds, nullptr);
                if (pool) solveThreaded(w, w.manifolds, *pool, w.settings.velocityIterations, w.settings.positionIterations);
                else solve(w, w.manifolds, nullptr, w.settings.velocityIterations, w.settings.positionIterations, nullptr);
            }
            w.prevManifolds = w.manifolds;
        }
    } else {
        w.prevManifolds.clear();
    }
    islands(w, w.manifolds);
    updateSleeping(w, dt, w.manifolds);   // Simulation.cpp:348-350: sleep, then stash the sleepers' manifolds
    w.prevManifolds = w.manifolds;
    w.postUpdateDone = false;
    bool anyCallback = false;
    for (uint32_t cb : w.colCallback) anyCallback = anyCallback || cb != 0;
    if (w.settings.collisionEvents && anyCallback) w.eventStep++;
}

}  // namespace orc
