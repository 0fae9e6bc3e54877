// Oracle joints (test infrastructure only; see oracle.h): the constraint data and velocity passes of
// Solving/SpringSolver.cpp, RevoluteJointSolver.cpp, PrismaticJointSolver.cpp, SliderJointSolver.cpp
// and MotorSolver.cpp, scheduled as ConstraintScheduler::solve (ConstraintScheduler.cpp:9-84).
#include "oracle.h"

#include <algorithm>
#include <cmath>

namespace orc {

namespace {

constexpr float PJ_PI = 3.14159265358979323846f;   // PrismaticJointSolver.cpp:13-14
constexpr float PJ_TWO_PI = 2.0f * PJ_PI;
constexpr float LIMIT_SLOP = 0.005f;                // PrismaticJointSolver.cpp:54
constexpr float LIMIT_BAUMGARTE = 0.2f;

V2 rotate(V2 v, float c, float s) { return { v.x * c - v.y * s, v.x * s + v.y * c }; }
float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
float wrapToPi(float a) { return a - PJ_TWO_PI * std::floor((a + PJ_PI) / PJ_TWO_PI); }
float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }

// Per-joint constraint data of one step.
struct Data {
    bool active = false;
    V2 rA{}, rB{}, dir{}, axis{}, perp{};
    float invEff = 0, bias = 0;                          // spring / slider perpendicular row (invK, bias)
    float invK11 = 0, invK12 = 0, invK22 = 0;            // revolute, prismatic block
    V2 biasV{};                                          // revolute
    float s1 = 0, s2 = 0, a1 = 0, a2 = 0;
    bool blockValid = false;
    float bias1 = 0, bias2 = 0;
    float axisInvEff = 0, axisBias = 0, limitDirection = 0, accAxis = 0;
    float invI = 0, inertia = 0, target = 0, maxImpulse = 0, impulse = 0;   // motor
};

void softness(float dt, float k, float c, float& gamma, float& beta) {
    gamma = 0.0f;
    beta = 0.0f;
    const float denom = dt * (c + dt * k);
    if (denom > 1e-12f) {
        gamma = 1.0f / denom;
        beta = dt * k * gamma;
    }
}

Data prepare(World& w, uint32_t j, float dt) {
    const Joint& J = w.joints[j];
    const JointDesc& p = J.d;
    Data D;
    if (p.type == JOINT_MOTOR) {   // MotorSolver.cpp:8-51
        if (asleep(w, p.bodyA)) return D;
        const float invI = w.invInertia[p.bodyA];
        if (invI == 0.0f) { w.motorAcc[j] = 0.0f; return D; }
        const float maxImpulse = p.maxTorque * dt;
        const float impulse = clampf(w.motorAcc[j], -maxImpulse, maxImpulse);
        D.invI = invI;
        D.inertia = 1.0f / invI;
        D.target = p.targetSpeed;
        D.maxImpulse = maxImpulse;
        D.impulse = impulse;
        w.motorAcc[j] = impulse;
        D.active = true;
        return D;
    }
    const uint32_t a = p.bodyA, b = p.bodyB;
    const float iMA = w.invMass[a], iMB = w.invMass[b], iIA = w.invInertia[a], iIB = w.invInertia[b];
    if (iMA == 0.0f && iMB == 0.0f) return D;
    if (asleep(w, a) || asleep(w, b)) return D;
    const V2 cA{ w.comX[a], w.comY[a] }, cB{ w.comX[b], w.comY[b] };
    const V2 rA = rotate({ p.anchorAX, p.anchorAY }, w.cosv[a], w.sinv[a]);
    const V2 rB = rotate({ p.anchorBX, p.anchorBY }, w.cosv[b], w.sinv[b]);
    const V2 d = (cB + rB) - (cA + rA);
    float gamma, beta;
    softness(dt, p.stiffness, p.damping, gamma, beta);
    if (p.type == JOINT_SPRING) {   // SpringSolver.cpp:8-112
        const float lenSq = dot(d, d);
        if (lenSq < 1e-12f) return D;
        const float len = std::sqrt(lenSq);
        const V2 dir{ d.x / len, d.y / len };
        const float C = len - p.restLength;
        const float raCn = rA.x * dir.y - rA.y * dir.x;
        const float rbCn = rB.x * dir.y - rB.y * dir.x;
        const float eff = iMA + iMB + iIA * raCn * raCn + iIB * rbCn * rbCn;
        if (eff <= 0.0f) return D;
        D.dir = dir; D.rA = rA; D.rB = rB;
        D.invEff = 1.0f / (eff + gamma);
        D.bias = beta * C;
        D.active = true;
        return D;
    }
    if (p.type == JOINT_REVOLUTE) {   // RevoluteJointSolver.cpp:8-108
        const float k11 = iMA + iMB + iIA * rA.y * rA.y + iIB * rB.y * rB.y;
        const float k22 = iMA + iMB + iIA * rA.x * rA.x + iIB * rB.x * rB.x;
        const float k12 = -iIA * rA.x * rA.y - iIB * rB.x * rB.y;
        const float k11S = k11 + gamma, k22S = k22 + gamma;
        const float det = k11S * k22S - k12 * k12;
        if (det <= 1e-12f) return D;
        const float invDet = 1.0f / det;
        D.rA = rA; D.rB = rB;
        D.invK11 = k22S * invDet;
        D.invK22 = k11S * invDet;
        D.invK12 = -k12 * invDet;
        D.biasV = beta * d;
        D.active = true;
        return D;
    }
    // prismatic (PrismaticJointSolver.cpp:24-186) and slider (SliderJointSolver.cpp:9-151)
    const V2 axis = rotate({ p.axisX, p.axisY }, w.cosv[a], w.sinv[a]);
    const V2 perp{ -axis.y, axis.x };
    D.axis = axis; D.perp = perp;
    D.s1 = cross(d + rA, perp);
    D.s2 = cross(rB, perp);
    if (p.type == JOINT_PRISMATIC) {
        const float k11 = iMA + iMB + iIA * D.s1 * D.s1 + iIB * D.s2 * D.s2;
        const float k12 = iIA * D.s1 + iIB * D.s2;
        float k22 = iIA + iIB;
        if (k22 == 0.0f) k22 = 1.0f;
        const float k11S = k11 + gamma, k22S = k22 + gamma;
        const float det = k11S * k22S - k12 * k12;
        if (det > 1e-12f) {
            const float invDet = 1.0f / det;
            D.invK11 = k22S * invDet;
            D.invK22 = k11S * invDet;
            D.invK12 = -k12 * invDet;
            D.blockValid = true;
            const float C1 = dot(perp, d);
            const float rel = wrapToPi(w.rot[b] - w.rot[a]);
            const float C2 = wrapToPi(rel - J.referenceAngle);
            D.bias1 = beta * C1;
            D.bias2 = beta * C2;
        }
    } else {
        const float k = iMA + iMB + iIA * D.s1 * D.s1 + iIB * D.s2 * D.s2;
        const float kSoft = k + gamma;
        if (kSoft > 1e-12f) {
            D.invEff = 1.0f / kSoft;
            D.bias = beta * dot(perp, d);
        }
    }
    if (p.enableLimit) {
        const float translation = dot(axis, d);
        float direction = 0.0f, C = 0.0f;
        if (translation <= p.lowerLimit + LIMIT_SLOP) { direction = 1.0f; C = translation - p.lowerLimit; }
        else if (translation >= p.upperLimit - LIMIT_SLOP) { direction = -1.0f; C = translation - p.upperLimit; }
        if (direction != 0.0f) {
            const float a1 = cross(d + rA, axis), a2 = cross(rB, axis);
            const float kAxis = iMA + iMB + iIA * a1 * a1 + iIB * a2 * a2;
            if (kAxis > 1e-12f) {
                const float invDt = dt > 1e-12f ? 1.0f / dt : 0.0f;
                D.a1 = a1; D.a2 = a2;
                D.axisInvEff = 1.0f / kAxis;
                D.limitDirection = direction;
                D.axisBias = LIMIT_BAUMGARTE * invDt * C;
            }
        }
    }
    D.active = D.blockValid || D.invEff > 0.0f || D.limitDirection != 0.0f;
    return D;
}

void solve(World& w, uint32_t j, Data& D) {
    if (!D.active) return;
    const JointDesc& p = w.joints[j].d;
    if (p.type == JOINT_MOTOR) {   // MotorSolver.cpp:68-96
        const uint32_t b = p.bodyA;
        const float wv = w.angVel[b];
        const float lambda = (D.target - wv) * D.inertia;
        const float oldI = D.impulse;
        const float newI = clampf(oldI + lambda, -D.maxImpulse, D.maxImpulse);
        D.impulse = newI;
        w.motorAcc[j] = newI;
        w.angVel[b] = wv + (newI - oldI) * D.invI;
        return;
    }
    const uint32_t a = p.bodyA, b = p.bodyB;
    const float iMA = w.invMass[a], iMB = w.invMass[b], iIA = w.invInertia[a], iIB = w.invInertia[b];
    V2 vA{ w.velX[a], w.velY[a] }, vB{ w.velX[b], w.velY[b] };
    float wA = w.angVel[a], wB = w.angVel[b];
    if (p.type == JOINT_SPRING || p.type == JOINT_REVOLUTE) {
        const V2 angA = V2{ -D.rA.y, D.rA.x } * wA;
        const V2 angB = V2{ -D.rB.y, D.rB.x } * wB;
        const V2 rel = (vB + angB) - (vA + angA);
        V2 imp;
        if (p.type == JOINT_SPRING) {
            const float cdot = dot(rel, D.dir);
            const float mag = (cdot + D.bias) * D.invEff;
            imp = D.dir * mag;
        } else {
            const V2 rhs = rel + D.biasV;
            imp = { D.invK11 * rhs.x + D.invK12 * rhs.y, D.invK12 * rhs.x + D.invK22 * rhs.y };
        }
        if (iMA > 0.0f) {
            vA = vA + imp * iMA;
            wA += (D.rA.x * imp.y - D.rA.y * imp.x) * iIA;
            w.velX[a] = vA.x; w.velY[a] = vA.y; w.angVel[a] = wA;
        }
        if (iMB > 0.0f) {
            vB = vB - imp * iMB;
            wB -= (D.rB.x * imp.y - D.rB.y * imp.x) * iIB;
            w.velX[b] = vB.x; w.velY[b] = vB.y; w.angVel[b] = wB;
        }
        return;
    }
    if (p.type == JOINT_PRISMATIC && D.blockValid) {
        const float cdot1 = dot(D.perp, vB - vA) + D.s2 * wB - D.s1 * wA;
        const float cdot2 = wB - wA;
        const float rhs1 = cdot1 + D.bias1, rhs2 = cdot2 + D.bias2;
        const float l1 = -(D.invK11 * rhs1 + D.invK12 * rhs2);
        const float l2 = -(D.invK12 * rhs1 + D.invK22 * rhs2);
        const V2 P = D.perp * l1;
        const float LA = l1 * D.s1 + l2, LB = l1 * D.s2 + l2;
        vA = vA - P * iMA; wA -= LA * iIA;
        vB = vB + P * iMB; wB += LB * iIB;
    } else if (p.type == JOINT_SLIDER && D.invEff > 0.0f) {
        const float cdot = dot(D.perp, vB - vA) + D.s2 * wB - D.s1 * wA;
        const float l = -D.invEff * (cdot + D.bias);
        const V2 P = D.perp * l;
        const float LA = l * D.s1, LB = l * D.s2;
        vA = vA - P * iMA; wA -= LA * iIA;
        vB = vB + P * iMB; wB += LB * iIB;
    }
    if (D.limitDirection != 0.0f) {
        const float cdot = dot(D.axis, vB - vA) + D.a2 * wB - D.a1 * wA;
        const float raw = -(cdot + D.axisBias) * D.axisInvEff;
        const float oldI = D.accAxis;
        const float newI = D.limitDirection > 0.0f ? std::max(oldI + raw, 0.0f) : std::min(oldI + raw, 0.0f);
        const float applied = newI - oldI;
        D.accAxis = newI;
        const V2 P = D.axis * applied;
        const float LA = applied * D.a1, LB = applied * D.a2;
        vA = vA - P * iMA; wA -= LA * iIA;
        vB = vB + P * iMB; wB += LB * iIB;
    }
    if (iMA > 0.0f) { w.velX[a] = vA.x; w.velY[a] = vA.y; w.angVel[a] = wA; }
    if (iMB > 0.0f) { w.velX[b] = vB.x; w.velY[b] = vB.y; w.angVel[b] = wB; }
}

}  // namespace

bool addJoint(World& w, const JointDesc& d) {
    const bool single = d.type == JOINT_MOTOR;
    if (d.type >= JOINT_TYPES || d.bodyA >= w.bodyCount() || (!single && d.bodyB >= w.bodyCount())) return false;
    Joint j;
    j.d = d;
    if (single) j.d.bodyB = d.bodyA;
    V2 axis{ d.axisX, d.axisY };   // sanitizeAxis (ConstraintSystemBase.h:146-158)
    const float lenSq = dot(axis, axis);
    if (lenSq > 1e-12f) {
        const float len = std::sqrt(lenSq);
        axis = { axis.x / len, axis.y / len };
    } else {
        axis = { 1.0f, 0.0f };
    }
    j.d.axisX = axis.x;
    j.d.axisY = axis.y;
    if (d.type == JOINT_PRISMATIC) {   // computeReferenceAngle (PrismaticJointConstraintSystem.h:118-125)
        float r = w.rot[d.bodyB] - w.rot[d.bodyA];
        r -= PJ_TWO_PI * std::floor((r + PJ_PI) / PJ_TWO_PI);
        j.referenceAngle = r;
    }
    w.joints.push_back(j);
    w.motorAcc.push_back(0.0f);
    return true;
}

void solveJoints(World& w, float dt) {
    const uint32_t n = (uint32_t)w.joints.size();
    if (n == 0) return;
    // the order of the velocity passes: the given colours, or each type's joints in index order
    std::vector<uint32_t> order = w.jointOrder, colourStart = w.jointColourStart, typeColour = w.jointTypeColour;
    if (order.empty()) {
        colourStart.clear();
        typeColour.assign(JOINT_TYPES + 1, 0);
        for (uint32_t t = 0; t < JOINT_TYPES; t++) {
            typeColour[t] = (uint32_t)colourStart.size();
            const size_t before = order.size();
            for (uint32_t j = 0; j < n; j++)
                if (w.joints[j].d.type == t) order.push_back(j);
            if (order.size() > before) colourStart.push_back((uint32_t)before);
        }
        typeColour[JOINT_TYPES] = (uint32_t)colourStart.size();
        colourStart.push_back((uint32_t)order.size());
    }
    std::vector<Data> data(n);
    bool awake[JOINT_TYPES] = {};
    for (uint32_t j = 0; j < n; j++) {   // the Compute tick
        data[j] = prepare(w, j, dt);
        const JointDesc& p = w.joints[j].d;
        if (!asleep(w, p.bodyA) && !asleep(w, p.bodyB)) awake[p.type] = true;
    }
    uint32_t maxIt = 0;
    for (uint32_t t = 0; t < JOINT_TYPES; t++)
        if (awake[t]) maxIt = std::max(maxIt, w.settings.jointIterations[t]);
    auto colourRange = [&](uint32_t c) { return std::make_pair(colourStart[c], colourStart[c + 1]); };
    for (uint32_t c = typeColour[JOINT_MOTOR]; c < typeColour[JOINT_MOTOR + 1]; c++) {   // WarmStart ticks
        const auto r = colourRange(c);
        for (uint32_t i = r.first; i < r.second; i++) {
            const uint32_t j = order[i];
            if (data[j].active) w.angVel[w.joints[j].d.bodyA] += data[j].impulse * data[j].invI;
        }
    }
    for (uint32_t k = 0; k < maxIt; k++)   // Solve ticks
        for (uint32_t t = 0; t < JOINT_TYPES; t++) {
            const uint64_t it = awake[t] ? w.settings.jointIterations[t] : 0;
            if (it == 0 || (uint64_t(k + 1) * it) / maxIt == (uint64_t(k) * it) / maxIt) continue;
            for (uint32_t c = typeColour[t]; c < typeColour[t + 1]; c++) {
                const auto r = colourRange(c);
                for (uint32_t i = r.first; i < r.second; i++) solve(w, order[i], data[order[i]]);
            }
        }
}

}  // namespace orc
