// Collision events: the event recording of BodyCollisionSolver::computeConstraintData
// (Solving/BodyCollisionSolver.cpp:300-388), on the step's pre-warm-start velocities.
#include "agx_internal.h"

#include <algorithm>

namespace agx {

namespace {

// Per manifold: the impact impulse (the most, over its contact points, of approach speed x normal
// effective mass; BodyCollisionSolver.cpp:333-356) and how many sides fire. The normal mass and the
// approach speed are computed with kPrep's operations, so they carry the same bits as the solver's.
__global__ __launch_bounds__(256) void kEventCount(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                                   const uint32_t* __restrict__ colBody, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const float2* __restrict__ inv,
                                                   const uint32_t* __restrict__ colCallback,
                                                   const float* __restrict__ colThreshold, uint32_t* __restrict__ evCount,
                                                   uint32_t* __restrict__ evSides, float* __restrict__ evImpulse) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const agx_manifold d = man[m];
        const uint32_t cbA = colCallback[d.a], cbB = colCallback[d.b];
        uint32_t fires = 0;
        float impact = 0.0f;
        if (cbA != 0u || cbB != 0u) {
            const uint32_t bA = colBody[d.a], bB = colBody[d.b];
            const float4 pA = pos[bA], pB = pos[bB], vA = vel[bA], vB = vel[bB];
            const float2 iA = inv[bA], iB = inv[bB];
            const float nx = d.normal_x, ny = d.normal_y;
            const float effMass = iA.x + iB.x;
            for (uint32_t i = 0; i < d.count; i++) {
                const float cx = i == 0 ? d.p0_x : d.p1_x, cy = i == 0 ? d.p0_y : d.p1_y;
                const float rAx = cx - pA.x, rAy = cy - pA.y, rBx = cx - pB.x, rBy = cy - pB.y;
                const float rAn = -rAy * nx + rAx * ny, rBn = -rBy * nx + rBx * ny;
                const float nd = effMass + rAn * rAn * iA.y + rBn * rBn * iB.y;
                const float normalMass = nd > 0.0f ? 1.0f / nd : 0.0f;
                const float negVn = -(((vB.x - vA.x) * nx + (vB.y - vA.y) * ny) + rBn * vB.z - rAn * vA.z);
                impact = fmaxf(impact, negVn * normalMass);
            }
            if (impact > 0.0f) {
                fires += (cbA != 0u && impact >= colThreshold[d.a]) ? 1u : 0u;
                fires += (cbB != 0u && impact >= colThreshold[d.b]) ? 2u : 0u;
            }
        }
        evCount[m] = fires;     // bit 0: side A fires, bit 1: side B fires
        evSides[m] = __popc(fires);
        evImpulse[m] = impact;
    }
}

__global__ __launch_bounds__(256) void kEventWrite(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                                   const uint32_t* __restrict__ colBody,
                                                   const uint32_t* __restrict__ colCallback,
                                                   const uint32_t* __restrict__ evCount,
                                                   const uint32_t* __restrict__ evOffset,
                                                   const float* __restrict__ evImpulse,
                                                   agx_collision_event* __restrict__ out, uint32_t cap) {
    const uint32_t M = ctr->manifolds;
    const uint32_t base = ctr->eventCount, step = ctr->eventStep;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint32_t fires = evCount[m];
        if (fires == 0u) continue;
        const agx_manifold d = man[m];
        agx_collision_event e;
        e.point_x = d.count > 1 ? (d.p0_x + d.p1_x) * 0.5f : d.p0_x;
        e.point_y = d.count > 1 ? (d.p0_y + d.p1_y) * 0.5f : d.p0_y;
        e.impulse = evImpulse[m];
        e.step = step;
        e.reserved = 0u;
        uint32_t at = base + evOffset[m];
        if (fires & 1u) {
            e.self = d.a; e.other = d.b;
            e.self_body = colBody[d.a]; e.other_body = colBody[d.b];
            e.normal_x = d.normal_x; e.normal_y = d.normal_y;
            e.callback_id = colCallback[d.a];
            if (at < cap) out[at] = e;
            at++;
        }
        if (fires & 2u) {
            e.self = d.b; e.other = d.a;
            e.self_body = colBody[d.b]; e.other_body = colBody[d.a];
            e.normal_x = -d.normal_x; e.normal_y = -d.normal_y;
            e.callback_id = colCallback[d.b];
            if (at < cap) out[at] = e;
        }
    }
}

// The call's running event count and step index advance once per step.
__global__ void kEventAdvance(Counters* ctr) {
    if (threadIdx.x == 0) {
        ctr->eventCount += ctr->eventStepTotal;
        ctr->eventStep += 1u;
    }
}

}  // namespace

hipError_t launchEvents(hipStream_t st, DeviceWorld& w) {
    const uint32_t gm = std::max(1u, std::min((uint32_t)((std::min<uint64_t>(w.manCap, 1u << 22) + 255u) / 256u), 8192u));
    hipLaunchKernelGGL(kEventCount, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody, w.pos, w.vel, w.inv,
                       w.colCallback, w.colThreshold, w.evCount, w.evSides, w.evImpulse);
    hipError_t e = launchScan(st, w.evSides, w.evOffset, &w.ctr->manifolds, &w.ctr->eventStepTotal, w.manCap, w.scanBlock,
                              w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kEventWrite, dim3(gm), dim3(256), 0, st, w.man[w.cur], w.ctr, w.colBody, w.colCallback, w.evCount,
                       w.evOffset, w.evImpulse, w.events, w.eventCap);
    hipLaunchKernelGGL(kEventAdvance, dim3(1), dim3(64), 0, st, w.ctr);
    return hipGetLastError();
}

}  // namespace agx
