// Joints: the five constraint solvers of Solving/ (SpringSolver.cpp, RevoluteJointSolver.cpp,
// PrismaticJointSolver.cpp, SliderJointSolver.cpp, MotorSolver.cpp) under the ConstraintScheduler's
// tick order (ConstraintScheduler.cpp:9-84), between the two halves of the velocity-Verlet step.
//
// One workgroup runs the whole schedule: constraint data for every joint, the motors' warm start,
// then the velocity passes of each type colour by colour (a colour's joints share no dynamic body, so
// its lanes solve them at once), with a workgroup barrier between colours. Body velocities stay in
// device memory; all waves of the workgroup share one CU and its L1, and the barrier orders their
// stores and loads. The per-joint arithmetic follows the reference's expressions in order, so the
// oracle replaying the same order gives the same bits.
#include "agx_internal.h"

#include <algorithm>

namespace agx {

namespace {

constexpr uint32_t JOINT_THREADS = 1024;
constexpr float LIMIT_SLOP = 0.005f;        // PrismaticJointSolver.cpp:54, SliderJointSolver.cpp:45
constexpr float LIMIT_BAUMGARTE = 0.2f;
constexpr float PJ_PI = 3.14159265358979323846f;
constexpr float PJ_TWO_PI = 2.0f * PJ_PI;

__device__ inline float2 rotate(float2 v, float c, float s) { return make_float2(v.x * c - v.y * s, v.x * s + v.y * c); }
__device__ inline float cross2(float2 a, float2 b) { return a.x * b.y - a.y * b.x; }
__device__ inline float wrapToPi(float a) { return a - PJ_TWO_PI * floorf((a + PJ_PI) / PJ_TWO_PI); }
__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// Soft-constraint coefficients from stiffness k and damping c (SpringSolver.cpp:93-103).
__device__ inline void softness(float dt, float k, float c, float& gamma, float& beta) {
    gamma = 0.0f;
    beta = 0.0f;
    const float denom = dt * (c + dt * k);
    if (denom > 1e-12f) {
        gamma = 1.0f / denom;
        beta = dt * k * gamma;
    }
}

// Per-step data slots (JointData::v, 20 floats). Zero in slot 19 marks a joint with nothing to solve.
enum : uint32_t { D_ACTIVE = 19 };

struct Bodies {
    const float4* pos;
    const float2* rcs;
    const float2* inv;
    float4* vel;
    const uint32_t* sleepId;
};

__device__ inline bool asleep(const Bodies& B, uint32_t b) { return B.sleepId && B.sleepId[b] != 0u; }

__device__ void prepJoint(const JointRec& J, JointData& D, float* motorAcc, uint32_t j, const Bodies& B, float dt) {
    for (int k = 0; k < 20; k++) D.v[k] = 0.0f;
    if (J.type == AGX_JOINT_MOTOR) {   // MotorSolver.cpp:8-51
        if (asleep(B, J.a)) return;   // in the sleeping group: not prepared, its impulse kept
        const float invI = B.inv[J.a].y;
        if (invI == 0.0f) { motorAcc[j] = 0.0f; return; }
        const float maxImpulse = J.motor.y * dt;
        const float impulse = clampf(motorAcc[j], -maxImpulse, maxImpulse);
        D.v[0] = invI;
        D.v[1] = 1.0f / invI;
        D.v[2] = J.motor.x;
        D.v[3] = maxImpulse;
        D.v[4] = impulse;
        motorAcc[j] = impulse;
        D.v[D_ACTIVE] = 1.0f;
        return;
    }
    const float2 iA = B.inv[J.a], iB = B.inv[J.b];
    if (iA.x == 0.0f && iB.x == 0.0f) return;
    if (asleep(B, J.a) || asleep(B, J.b)) return;   // constraints of sleeping bodies are not active
    const float4 pA = B.pos[J.a], pB = B.pos[J.b];
    const float2 csA = B.rcs[J.a], csB = B.rcs[J.b];
    const float2 rA = rotate(make_float2(J.anchors.x, J.anchors.y), csA.x, csA.y);
    const float2 rB = rotate(make_float2(J.anchors.z, J.anchors.w), csB.x, csB.y);
    const float2 wA = make_float2(pA.x + rA.x, pA.y + rA.y);
    const float2 wB = make_float2(pB.x + rB.x, pB.y + rB.y);
    const float2 d = make_float2(wB.x - wA.x, wB.y - wA.y);
    const float iMA = iA.x, iMB = iB.x, iIA = iA.y, iIB = iB.y;
    float gamma, beta;
    softness(dt, J.params.x, J.params.y, gamma, beta);
    if (J.type == AGX_JOINT_SPRING) {   // SpringSolver.cpp:8-112
        const float lenSq = d.x * d.x + d.y * d.y;
        if (lenSq < 1e-12f) return;
        const float len = sqrtf(lenSq);
        const float2 dir = make_float2(d.x / len, d.y / len);
        const float C = len - J.axisRest.z;
        const float raCn = rA.x * dir.y - rA.y * dir.x;
        const float rbCn = rB.x * dir.y - rB.y * dir.x;
        const float eff = iMA + iMB + iIA * raCn * raCn + iIB * rbCn * rbCn;
        if (eff <= 0.0f) return;
        D.v[0] = dir.x; D.v[1] = dir.y;
        D.v[2] = rA.x; D.v[3] = rA.y; D.v[4] = rB.x; D.v[5] = rB.y;
        D.v[6] = 1.0f / (eff + gamma);
        D.v[7] = beta * C;
        D.v[D_ACTIVE] = 1.0f;
        return;
    }
    if (J.type == AGX_JOINT_REVOLUTE) {   // RevoluteJointSolver.cpp:8-108
        const float k11 = iMA + iMB + iIA * rA.y * rA.y + iIB * rB.y * rB.y;
        const float k22 = iMA + iMB + iIA * rA.x * rA.x + iIB * rB.x * rB.x;
        const float k12 = -iIA * rA.x * rA.y - iIB * rB.x * rB.y;
        const float k11S = k11 + gamma, k22S = k22 + gamma;
        const float det = k11S * k22S - k12 * k12;
        if (det <= 1e-12f) return;
        const float invDet = 1.0f / det;
        D.v[0] = rA.x; D.v[1] = rA.y; D.v[2] = rB.x; D.v[3] = rB.y;
        D.v[4] = k22S * invDet;   // invK11
        D.v[5] = -k12 * invDet;   // invK12
        D.v[6] = k11S * invDet;   // invK22
        D.v[7] = beta * d.x; D.v[8] = beta * d.y;
        D.v[D_ACTIVE] = 1.0f;
        return;
    }
    // prismatic (PrismaticJointSolver.cpp:24-186) and slider (SliderJointSolver.cpp:9-151)
    const float2 axis = rotate(make_float2(J.axisRest.x, J.axisRest.y), csA.x, csA.y);
    const float2 perp = make_float2(-axis.y, axis.x);
    const float2 dA = make_float2(d.x + rA.x, d.y + rA.y);
    const float s1 = cross2(dA, perp), s2 = cross2(rB, perp);
    D.v[0] = axis.x; D.v[1] = axis.y; D.v[2] = perp.x; D.v[3] = perp.y; D.v[4] = s1; D.v[5] = s2;
    bool any = false;
    if (J.type == AGX_JOINT_PRISMATIC) {
        const float k11 = iMA + iMB + iIA * s1 * s1 + iIB * s2 * s2;
        const float k12 = iIA * s1 + iIB * s2;
        float k22 = iIA + iIB;
        if (k22 == 0.0f) k22 = 1.0f;
        const float k11S = k11 + gamma, k22S = k22 + gamma;
        const float det = k11S * k22S - k12 * k12;
        if (det > 1e-12f) {
            const float invDet = 1.0f / det;
            D.v[8] = k22S * invDet;   // invK11
            D.v[9] = -k12 * invDet;   // invK12
            D.v[10] = k11S * invDet;  // invK22
            D.v[11] = 1.0f;           // block valid
            const float C1 = perp.x * d.x + perp.y * d.y;
            const float rel = wrapToPi(pB.z - pA.z);
            const float C2 = wrapToPi(rel - J.axisRest.w);
            D.v[12] = beta * C1;
            D.v[13] = beta * C2;
            any = true;
        }
    } else {
        const float k = iMA + iMB + iIA * s1 * s1 + iIB * s2 * s2;
        const float kSoft = k + gamma;
        if (kSoft > 1e-12f) {
            D.v[8] = 1.0f / kSoft;   // invK
            D.v[12] = beta * (perp.x * d.x + perp.y * d.y);
            any = true;
        }
    }
    if (J.enableLimit) {
        const float translation = axis.x * d.x + axis.y * d.y;
        const float lower = J.params.z, upper = J.params.w;
        float direction = 0.0f, C = 0.0f;
        if (translation <= lower + LIMIT_SLOP) { direction = 1.0f; C = translation - lower; }
        else if (translation >= upper - LIMIT_SLOP) { direction = -1.0f; C = translation - upper; }
        if (direction != 0.0f) {
            const float a1 = cross2(dA, axis), a2 = cross2(rB, axis);
            const float kAxis = iMA + iMB + iIA * a1 * a1 + iIB * a2 * a2;
            if (kAxis > 1e-12f) {
                const float invDt = dt > 1e-12f ? 1.0f / dt : 0.0f;
                D.v[6] = a1; D.v[7] = a2;
                D.v[14] = 1.0f / kAxis;
                D.v[15] = LIMIT_BAUMGARTE * invDt * C;
                D.v[16] = direction;
                any = true;
            }
        }
    }
    if (any) D.v[D_ACTIVE] = 1.0f;
}

__device__ void warmMotor(const JointRec& J, const JointData& D, const Bodies& B) {   // MotorSolver.cpp:53-66
    if (D.v[D_ACTIVE] == 0.0f) return;
    float4 v = B.vel[J.a];
    v.z += D.v[4] * D.v[0];
    B.vel[J.a] = v;
}

__device__ void solveJoint(const JointRec& J, JointData& D, float* motorAcc, uint32_t j, const Bodies& B) {
    if (D.v[D_ACTIVE] == 0.0f) return;
    if (J.type == AGX_JOINT_MOTOR) {   // MotorSolver.cpp:68-96
        float4 v = B.vel[J.a];
        const float w = v.z;
        const float lambda = (D.v[2] - w) * D.v[1];
        const float oldImpulse = D.v[4];
        const float newImpulse = clampf(oldImpulse + lambda, -D.v[3], D.v[3]);
        D.v[4] = newImpulse;
        motorAcc[j] = newImpulse;
        v.z = w + (newImpulse - oldImpulse) * D.v[0];
        B.vel[J.a] = v;
        return;
    }
    const float2 iA = B.inv[J.a], iB = B.inv[J.b];
    const float iMA = iA.x, iMB = iB.x, iIA = iA.y, iIB = iB.y;
    float4 vA = B.vel[J.a], vB = B.vel[J.b];
    if (J.type == AGX_JOINT_SPRING || J.type == AGX_JOINT_REVOLUTE) {
        const bool spring = J.type == AGX_JOINT_SPRING;
        const float2 rA = spring ? make_float2(D.v[2], D.v[3]) : make_float2(D.v[0], D.v[1]);
        const float2 rB = spring ? make_float2(D.v[4], D.v[5]) : make_float2(D.v[2], D.v[3]);
        const float2 angA = make_float2(-rA.y * vA.z, rA.x * vA.z);
        const float2 angB = make_float2(-rB.y * vB.z, rB.x * vB.z);
        const float2 rel = make_float2((vB.x + angB.x) - (vA.x + angA.x), (vB.y + angB.y) - (vA.y + angA.y));
        float2 imp;
        if (spring) {   // SpringSolver.cpp:146-151
            const float cdot = rel.x * D.v[0] + rel.y * D.v[1];
            const float mag = (cdot + D.v[7]) * D.v[6];
            imp = make_float2(D.v[0] * mag, D.v[1] * mag);
        } else {        // RevoluteJointSolver.cpp:146-154
            const float rx = rel.x + D.v[7], ry = rel.y + D.v[8];
            imp = make_float2(D.v[4] * rx + D.v[5] * ry, D.v[5] * rx + D.v[6] * ry);
        }
        if (iMA > 0.0f) {
            vA.x = vA.x + imp.x * iMA;
            vA.y = vA.y + imp.y * iMA;
            const float tA = rA.x * imp.y - rA.y * imp.x;
            vA.z += tA * iIA;
            B.vel[J.a] = vA;
        }
        if (iMB > 0.0f) {
            vB.x = vB.x - imp.x * iMB;
            vB.y = vB.y - imp.y * iMB;
            const float tB = rB.x * imp.y - rB.y * imp.x;
            vB.z -= tB * iIB;
            B.vel[J.b] = vB;
        }
        return;
    }
    // prismatic / slider
    const float2 axis = make_float2(D.v[0], D.v[1]), perp = make_float2(D.v[2], D.v[3]);
    const float s1 = D.v[4], s2 = D.v[5];
    if (J.type == AGX_JOINT_PRISMATIC && D.v[11] != 0.0f) {   // PrismaticJointSolver.cpp:222-244
        const float cdot1 = (perp.x * (vB.x - vA.x) + perp.y * (vB.y - vA.y)) + s2 * vB.z - s1 * vA.z;
        const float cdot2 = vB.z - vA.z;
        const float rhs1 = cdot1 + D.v[12], rhs2 = cdot2 + D.v[13];
        const float l1 = -(D.v[8] * rhs1 + D.v[9] * rhs2);
        const float l2 = -(D.v[9] * rhs1 + D.v[10] * rhs2);
        const float px = perp.x * l1, py = perp.y * l1;
        const float LA = l1 * s1 + l2, LB = l1 * s2 + l2;
        vA.x = vA.x - px * iMA; vA.y = vA.y - py * iMA; vA.z -= LA * iIA;
        vB.x = vB.x + px * iMB; vB.y = vB.y + py * iMB; vB.z += LB * iIB;
    } else if (J.type == AGX_JOINT_SLIDER && D.v[8] > 0.0f) {   // SliderJointSolver.cpp:180-196
        const float cdot = (perp.x * (vB.x - vA.x) + perp.y * (vB.y - vA.y)) + s2 * vB.z - s1 * vA.z;
        const float l = -D.v[8] * (cdot + D.v[12]);
        const float px = perp.x * l, py = perp.y * l;
        const float LA = l * s1, LB = l * s2;
        vA.x = vA.x - px * iMA; vA.y = vA.y - py * iMA; vA.z -= LA * iIA;
        vB.x = vB.x + px * iMB; vB.y = vB.y + py * iMB; vB.z += LB * iIB;
    }
    if (D.v[16] != 0.0f) {   // the axis limit: a one-sided stop, accumulated and clamped
        const float a1 = D.v[6], a2 = D.v[7];
        const float cdot = (axis.x * (vB.x - vA.x) + axis.y * (vB.y - vA.y)) + a2 * vB.z - a1 * vA.z;
        const float raw = -(cdot + D.v[15]) * D.v[14];
        const float oldI = D.v[17];
        const float sum = oldI + raw;
        const float newI = D.v[16] > 0.0f ? (sum < 0.0f ? 0.0f : sum) : (0.0f < sum ? 0.0f : sum);
        const float applied = newI - oldI;
        D.v[17] = newI;
        const float px = axis.x * applied, py = axis.y * applied;
        const float LA = applied * a1, LB = applied * a2;
        vA.x = vA.x - px * iMA; vA.y = vA.y - py * iMA; vA.z -= LA * iIA;
        vB.x = vB.x + px * iMB; vB.y = vB.y + py * iMB; vB.z += LB * iIB;
    }
    if (iMA > 0.0f) B.vel[J.a] = vA;
    if (iMB > 0.0f) B.vel[J.b] = vB;
}

// Every velocity pass of type t whose turn comes at tick k: iterations spread evenly over the
// largest count of the active types (ConstraintScheduler.cpp:42-54).
__device__ inline bool fires(uint32_t k, uint32_t n, uint32_t maxIt) {
    return (uint64_t)(k + 1) * n / maxIt != (uint64_t)k * n / maxIt;
}

__global__ __launch_bounds__(JOINT_THREADS) void kJoints(const JointRec* __restrict__ joints, JointData* __restrict__ data,
                                                         float* __restrict__ motorAcc, uint32_t nJoints,
                                                         const uint32_t* __restrict__ order, const uint2* __restrict__ colours,
                                                         JointSchedule sched, Bodies B, float dt) {
    const uint32_t t = threadIdx.x;
    // A type takes part when some joint of it is awake (ConstraintScheduler.cpp:20-31: a system whose
    // constraints all sleep has no waves); the passes spread over the largest count of those types.
    __shared__ uint32_t awake[AGX_JOINT_TYPES];
    if (t < AGX_JOINT_TYPES) awake[t] = 0u;
    __syncthreads();
    for (uint32_t j = t; j < nJoints; j += JOINT_THREADS) {
        const JointRec J = joints[j];
        prepJoint(J, data[j], motorAcc, j, B, dt);
        const bool sleeping = asleep(B, J.a) || (J.type != AGX_JOINT_MOTOR && asleep(B, J.b));
        if (!sleeping) awake[J.type] = 1u;
    }
    __syncthreads();
    uint32_t maxIt = 0;
    for (uint32_t type = 0; type < AGX_JOINT_TYPES; type++)
        if (awake[type]) maxIt = max(maxIt, sched.iterations[type]);
    // the motors' warm start, colour by colour
    for (uint32_t c = sched.typeColour[AGX_JOINT_MOTOR]; c < sched.typeColour[AGX_JOINT_MOTOR + 1]; c++) {
        const uint2 r = colours[c];
        for (uint32_t i = r.x + t; i < r.y; i += JOINT_THREADS) warmMotor(joints[order[i]], data[order[i]], B);
        __syncthreads();
    }
    for (uint32_t k = 0; k < maxIt; k++) {
        for (uint32_t type = 0; type < AGX_JOINT_TYPES; type++) {
            const uint32_t n = awake[type] ? sched.iterations[type] : 0u;
            if (n == 0 || !fires(k, n, maxIt)) continue;
            for (uint32_t c = sched.typeColour[type]; c < sched.typeColour[type + 1]; c++) {
                const uint2 r = colours[c];
                for (uint32_t i = r.x + t; i < r.y; i += JOINT_THREADS) {
                    const uint32_t j = order[i];
                    solveJoint(joints[j], data[j], motorAcc, j, B);
                }
                __syncthreads();
            }
        }
    }
}

}  // namespace

hipError_t launchJoints(hipStream_t st, DeviceWorld& w, float dt) {
    if (w.nJoints == 0) return hipSuccess;
    Bodies B{ w.pos, w.rcs, w.inv, w.vel, w.sleepId };
    hipLaunchKernelGGL(kJoints, dim3(1), dim3(JOINT_THREADS), 0, st, w.joints, w.jointData, w.motorAcc, w.nJoints, w.jointOrder,
                       w.jointColours, w.jointSchedule, B, dt);
    return hipGetLastError();
}

}  // namespace agx
