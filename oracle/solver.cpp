// Oracle contact solver (test infrastructure only): Solving/BodyCollisionSolver.cpp,
// single-threaded path, generalised to an explicit constraint order (the threaded path
// solves in wave order; any order in which no two simultaneous constraints share a dynamic
// body is equivalent to a sequential order).
#include "oracle.h"
#include "solver_internal.h"

#include <cmath>
#include <vector>

namespace orc {

namespace {

constexpr float SLOP = 0.001f;            // BodyCollisionSolver.h:71-72
constexpr float BAUMGARTE = 0.5f;
constexpr float MAX_CONDITION = 1000.0f;  // BodyCollisionSolver.h:65

inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }

}  // namespace

void Solver::prepare(const uint32_t* order, size_t lo, size_t hi) {
    for (size_t c = lo; c < hi; c++) {
        const uint32_t mi = order ? order[c] : (uint32_t)c;
        const Manifold& d = ms[mi];
        Constraint& k = cs[c];
        const uint32_t bA = w.colBody[d.a], bB = w.colBody[d.b];
        const V2 comA{ w.comX[bA], w.comY[bA] }, comB{ w.comX[bB], w.comY[bB] };
        k.bodyA = bA; k.bodyB = bB; k.count = d.count; k.manifold = mi;
        k.nx = d.nx; k.ny = d.ny;
        k.imp[0] = d.jn0; k.imp[1] = d.jt0; k.imp[2] = d.jn1; k.imp[3] = d.jt1;
        k.invMassA = w.invMass[bA]; k.invMassB = w.invMass[bB];
        k.invInertiaA = w.invInertia[bA]; k.invInertiaB = w.invInertia[bB];
        const V2 n{ d.nx, d.ny };
        k.penetrationBase = d.depth + dot(comB - comA, n) - SLOP;
        const Material& mA = w.materials[w.colMaterial[d.a]];
        const Material& mB = w.materials[w.colMaterial[d.b]];
        const float e = (mA.elasticity + mB.elasticity) * 0.5f;
        k.staticFriction = std::sqrt(std::fmax(0.0f, mA.staticFriction * mB.staticFriction));
        k.dynamicFriction = std::sqrt(std::fmax(0.0f, mA.dynamicFriction * mB.dynamicFriction));
        const float effMass = k.invMassA + k.invMassB;
        const V2 t{ -n.y, n.x };
        const V2 vA{ w.velX[bA], w.velY[bA] }, vB{ w.velX[bB], w.velY[bB] };
        const float wA = w.angVel[bA], wB = w.angVel[bB];
        k.p[1] = PointData{ 0, 0, 0, 0, 0, 0, 0 };
        for (uint32_t i = 0; i < d.count; i++) {
            const V2 cp = i == 0 ? V2{ d.p0x, d.p0y } : V2{ d.p1x, d.p1y };
            PointData& p = k.p[i];
            const V2 rA = cp - comA, rB = cp - comB;
            const V2 rAp{ -rA.y, rA.x }, rBp{ -rB.y, rB.x };
            p.rAn = dot(rAp, n); p.rBn = dot(rBp, n); p.rAt = dot(rAp, t); p.rBt = dot(rBp, t);
            const float nd = effMass + p.rAn * p.rAn * k.invInertiaA + p.rBn * p.rBn * k.invInertiaB;
            p.normalMass = nd > 0.0f ? 1.0f / nd : 0.0f;
            const float td = effMass + p.rAt * p.rAt * k.invInertiaA + p.rBt * p.rBt * k.invInertiaB;
            p.tangentMass = td > 0.0f ? 1.0f / td : 0.0f;
            const float negVn = -(dot(vB - vA, n) + p.rBn * wB - p.rAn * wA);
            p.bias = (negVn > 0.0f) ? e * negVn : 0.0f;
        }
    }
}

void Solver::warmStart(size_t lo, size_t hi) {
    for (size_t c = lo; c < hi; c++) {
        const Constraint& k = cs[c];
        V2 vA, vB; float wA, wB;
        load(k.bodyA, vA, wA); load(k.bodyB, vB, wB);
        const V2 n{ k.nx, k.ny }, t{ -n.y, n.x };
        float jn[2] = { 0, 0 }, jt[2] = { 0, 0 };
        for (uint32_t i = 0; i < k.count; i++) { jn[i] = k.imp[2 * i]; jt[i] = k.imp[2 * i + 1]; }
        const PointData& p0 = k.p[0];
        const PointData& p1 = k.p[1];
        const V2 sum = (jn[0] + jn[1]) * n + (jt[0] + jt[1]) * t;
        vA = vA - sum * k.invMassA;
        vB = vB + sum * k.invMassB;
        wA -= (jn[0] * p0.rAn + jt[0] * p0.rAt + jn[1] * p1.rAn + jt[1] * p1.rAt) * k.invInertiaA;
        wB += (jn[0] * p0.rBn + jt[0] * p0.rBt + jn[1] * p1.rBn + jt[1] * p1.rBt) * k.invInertiaB;
        store(k.bodyA, vA, wA, k.invMassA);
        store(k.bodyB, vB, wB, k.invMassB);
    }
}

// Block normal, iterative friction (.h:67-68).
void Solver::velocity(size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; c++) {
            Constraint& k = cs[c];
            V2 vA, vB; float wA, wB;
            load(k.bodyA, vA, wA); load(k.bodyB, vB, wB);
            const V2 n{ k.nx, k.ny }, t{ -n.y, n.x };
            const float iMA = k.invMassA, iMB = k.invMassB, iIA = k.invInertiaA, iIB = k.invInertiaB;
            auto friction = [&](uint32_t i, float jn) {
                const PointData& p = k.p[i];
                const float slip = dot(vB - vA, t) + p.rBt * wB - p.rAt * wA;
                const float jt = -slip * p.tangentMass;
                float& acc = k.imp[2 * i + 1];
                const float old = acc;
                const float target = old + jt;
                if (std::fabs(target) <= jn * k.staticFriction) acc = target;
                else { const float mx = jn * k.dynamicFriction; acc = clampf(target, -mx, mx); }
                const float dj = acc - old;
                const V2 imp = dj * t;
                vA = vA - imp * iMA; wA -= dj * p.rAt * iIA;
                vB = vB + imp * iMB; wB += dj * p.rBt * iIB;
            };
            if (k.count == 1) {
                const PointData& p = k.p[0];
                float jnSingle = 0.0f;
                const float vn = dot(vB - vA, n) + p.rBn * wB - p.rAn * wA;
                float& acc = k.imp[0];
                if (!(vn > 0.0f && acc == 0.0f)) {
                    const float jn = (p.bias - vn) * p.normalMass;
                    const float old = acc;
                    acc = std::fmax(0.0f, old + jn);
                    const float dj = acc - old;
                    jnSingle = acc;
                    const V2 imp = dj * n;
                    vA = vA - imp * iMA; wA -= dj * p.rAn * iIA;
                    vB = vB + imp * iMB; wB += dj * p.rBn * iIB;
                }
                friction(0, jnSingle);
            } else {
                const float effMass = iMA + iMB;
                float jnArr[2] = { 0, 0 };
                auto normalIterative = [&]() {
                    for (uint32_t i = 0; i < 2; i++) {
                        const PointData& p = k.p[i];
                        const float vn = dot(vB - vA, n) + p.rBn * wB - p.rAn * wA;
                        const float jn = (p.bias - vn) * p.normalMass;
                        float& acc = k.imp[2 * i];
                        const float old = acc;
                        acc = std::fmax(0.0f, old + jn);
                        const float dj = acc - old;
                        jnArr[i] = acc;
                        const V2 imp = dj * n;
                        vA = vA - imp * iMA; wA -= dj * p.rAn * iIA;
                        vB = vB + imp * iMB; wB += dj * p.rBn * iIB;
                    }
                };
                const PointData& c0 = k.p[0];
                const PointData& c1 = k.p[1];
                const float dvn = dot(vB - vA, n);
                const float vn0 = dvn + c0.rBn * wB - c0.rAn * wA;
                const float vn1 = dvn + c1.rBn * wB - c1.rAn * wA;
                float& a0 = k.imp[0];
                float& a1 = k.imp[2];
                const float K00 = effMass + c0.rAn * c0.rAn * iIA + c0.rBn * c0.rBn * iIB;
                const float K11 = effMass + c1.rAn * c1.rAn * iIA + c1.rBn * c1.rBn * iIB;
                const float K01 = effMass + c0.rAn * c1.rAn * iIA + c0.rBn * c1.rBn * iIB;
                const float b0 = c0.bias - vn0 + K00 * a0 + K01 * a1;
                const float b1 = c1.bias - vn1 + K01 * a0 + K11 * a1;
                float x0 = 0, x1 = 0;
                bool ok = false;
                {
                    const float EPS = 1e-6f;
                    const float det = K00 * K11 - K01 * K01;
                    const float maxK = std::fmax(K00, K11);
                    if (maxK * maxK < MAX_CONDITION * det) {
                        const float invDet = 1.0f / det;
                        const float s0 = (K11 * b0 - K01 * b1) * invDet;
                        const float s1 = (K00 * b1 - K01 * b0) * invDet;
                        if (s0 >= -EPS && s1 >= -EPS) { x0 = std::fmax(s0, 0.0f); x1 = std::fmax(s1, 0.0f); ok = true; }
                    }
                    if (!ok && K11 > 0) {
                        const float s1 = b1 / K11;
                        if (s1 >= -EPS && (K01 * s1 - b0) >= -EPS) { x0 = 0; x1 = std::fmax(s1, 0.0f); ok = true; }
                    }
                    if (!ok && K00 > 0) {
                        const float s0 = b0 / K00;
                        if (s0 >= -EPS && (K01 * s0 - b1) >= -EPS) { x0 = std::fmax(s0, 0.0f); x1 = 0; ok = true; }
                    }
                    if (!ok && b0 <= EPS && b1 <= EPS) { x0 = 0; x1 = 0; ok = true; }
                }
                if (!ok) {
                    normalIterative();
                } else {
                    const float d0 = x0 - a0, d1 = x1 - a1;
                    a0 = x0; a1 = x1;
                    jnArr[0] = x0; jnArr[1] = x1;
                    const float jnSum = d0 + d1;
                    vA = vA - n * (jnSum * iMA);
                    vB = vB + n * (jnSum * iMB);
                    wA -= (d0 * c0.rAn + d1 * c1.rAn) * iIA;
                    wB += (d0 * c0.rBn + d1 * c1.rBn) * iIB;
                }
                friction(0, jnArr[0]);
                friction(1, jnArr[1]);
            }
            store(k.bodyA, vA, wA, iMA);
            store(k.bodyB, vB, wB, iMB);
        }
}

void Solver::position(size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; c++) {
            const Constraint& k = cs[c];
            const uint32_t bA = k.bodyA, bB = k.bodyB;
            const V2 pA{ w.comX[bA], w.comY[bA] }, pB{ w.comX[bB], w.comY[bB] };
            const V2 n{ k.nx, k.ny };
            const float pen = k.penetrationBase - dot(pB - pA, n);
            if (pen <= 0.0f) continue;
            const float iMA = w.invMass[bA], iMB = w.invMass[bB];
            const float total = pen / (iMA + iMB) * BAUMGARTE;
            const float cA = iMA * total, cB = iMB * total;
            const V2 vcA = n * cA, vcB = n * cB;
            if (iMA > 0) { w.comX[bA] -= vcA.x; w.comY[bA] -= vcA.y; }
            if (iMB > 0) { w.comX[bB] += vcB.x; w.comY[bB] += vcB.y; }
        }
}

void Solver::finish() {
    const size_t M = cs.size();
    for (size_t c = 0; c < M; c++) {
        Manifold& d = ms[cs[c].manifold];
        d.jn0 = cs[c].imp[0]; d.jt0 = cs[c].imp[1]; d.jn1 = cs[c].imp[2]; d.jt1 = cs[c].imp[3];
    }
    // computeConstraintErrors (:132-183).
    float tApp = 0.0f, tPen = 0.0f;
    for (size_t c = 0; c < M; c++) {
        const Constraint& k = cs[c];
        const V2 pA{ w.comX[k.bodyA], w.comY[k.bodyA] }, pB{ w.comX[k.bodyB], w.comY[k.bodyB] };
        const V2 n{ k.nx, k.ny };
        tPen += std::fmax(k.penetrationBase - dot(pB - pA, n), 0.0f);
        const V2 vA{ w.velX[k.bodyA], w.velY[k.bodyA] }, vB{ w.velX[k.bodyB], w.velY[k.bodyB] };
        const float wA = w.angVel[k.bodyA], wB = w.angVel[k.bodyB];
        for (uint32_t i = 0; i < k.count; i++) {
            const float app = -(dot(vB - vA, n) + k.p[i].rBn * wB - k.p[i].rAn * wA);
            tApp += app * (app > 0.0f ? 1.0f : 0.0f);
        }
    }
    w.velocityError = tApp;
    w.positionError = tPen;
}

void collisionEvents(World& w, const std::vector<Manifold>& ms, const uint32_t* order) {
    bool any = false;
    for (uint32_t cb : w.colCallback) any = any || cb != 0;
    if (!w.settings.collisionEvents || !any) return;
    auto callbackOf = [&](uint32_t c) { return c < w.colCallback.size() ? w.colCallback[c] : 0u; };
    auto thresholdOf = [&](uint32_t c) { return c < w.colThreshold.size() ? w.colThreshold[c] : 3.40282347e38f; };
    for (size_t c = 0; c < ms.size(); c++) {
        const Manifold& d = ms[order ? order[c] : c];
        const uint32_t bA = w.colBody[d.a], bB = w.colBody[d.b];
        const V2 comA{ w.comX[bA], w.comY[bA] }, comB{ w.comX[bB], w.comY[bB] };
        const float iMA = w.invMass[bA], iMB = w.invMass[bB], iIA = w.invInertia[bA], iIB = w.invInertia[bB];
        const V2 n{ d.nx, d.ny };
        const V2 vA{ w.velX[bA], w.velY[bA] }, vB{ w.velX[bB], w.velY[bB] };
        const float wA = w.angVel[bA], wB = w.angVel[bB];
        const float effMass = iMA + iMB;
        float impact = 0.0f;
        for (uint32_t i = 0; i < d.count; i++) {   // :333-356
            const V2 cp = i == 0 ? V2{ d.p0x, d.p0y } : V2{ d.p1x, d.p1y };
            const V2 rA = cp - comA, rB = cp - comB;
            const float rAn = dot(V2{ -rA.y, rA.x }, n), rBn = dot(V2{ -rB.y, rB.x }, n);
            const float nd = effMass + rAn * rAn * iIA + rBn * rBn * iIB;
            const float normalMass = nd > 0.0f ? 1.0f / nd : 0.0f;
            const float negVn = -(dot(vB - vA, n) + rBn * wB - rAn * wA);
            impact = std::fmax(impact, negVn * normalMass);
        }
        if (impact <= 0.0f) continue;   // :359
        const uint32_t cbA = callbackOf(d.a), cbB = callbackOf(d.b);
        const bool fireA = cbA != 0 && impact >= thresholdOf(d.a);
        const bool fireB = cbB != 0 && impact >= thresholdOf(d.b);
        if (!fireA && !fireB) continue;
        const V2 point = d.count > 1 ? (V2{ d.p0x, d.p0y } + V2{ d.p1x, d.p1y }) * 0.5f : V2{ d.p0x, d.p0y };
        if (fireA) w.events.push_back(Event{ d.a, d.b, bA, bB, point.x, point.y, n.x, n.y, impact, cbA, w.eventStep, 0 });
        if (fireB) w.events.push_back(Event{ d.b, d.a, bB, bA, point.x, point.y, -n.x, -n.y, impact, cbB, w.eventStep, 0 });
    }
}

void solve(World& w, std::vector<Manifold>& ms, const uint32_t* order, uint32_t vIt, uint32_t pIt, float* residualOut) {
    Solver sv(w, ms);
    const size_t M = ms.size();
    sv.prepare(order, 0, M);
    sv.warmStart(0, M);
    std::vector<float> prev;
    for (uint32_t it = 0; it < vIt; it++) {
        if (residualOut) {
            prev.resize(4 * M);
            for (size_t c = 0; c < M; c++) for (int q = 0; q < 4; q++) prev[4 * c + q] = sv.cs[c].imp[q];
        }
        sv.velocity(0, M);
        if (residualOut) {
            double dsq = 0.0, lsq = 0.0;
            for (size_t c = 0; c < M; c++)
                for (uint32_t q = 0; q < 2 * sv.cs[c].count; q++) {
                    const double d = (double)sv.cs[c].imp[q] - (double)prev[4 * c + q];
                    dsq += d * d;
                    lsq += (double)sv.cs[c].imp[q] * (double)sv.cs[c].imp[q];
                }
            residualOut[it] = (float)(std::sqrt(dsq) / std::fmax(std::sqrt(lsq), 1e-12));
        }
    }
    for (uint32_t it = 0; it < pIt; it++) sv.position(0, M);
    sv.finish();
}

}  // namespace orc
