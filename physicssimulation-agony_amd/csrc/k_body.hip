// Per-body and per-collider streaming kernels: integration, poses/AABBs, CCD.
#include "agx_internal.h"

#include <cfloat>

namespace agx {

namespace {

// Integration/Integrator.cpp:99-209 + Simulation.cpp:291-292, fused: CCD snapshot of the
// pre-integration pose, velocity-Verlet half kick / drift / half kick, wrap, cos/sin.
__global__ __launch_bounds__(256) void kIntegrate(float4* __restrict__ pos, float4* __restrict__ vel,
                                                  float2* __restrict__ rcs, const float2* __restrict__ inv,
                                                  float4* __restrict__ prev, const uint32_t* __restrict__ list,
                                                  uint32_t n, float dt, float gxh, float gyh, int snapshot) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        float4 p = pos[b];
        float4 v = vel[b];
        const float2 rc = rcs[b];
        if (snapshot) prev[b] = make_float4(p.x, p.y, rc.x, rc.y);
        const float movable = inv[b].x != 0.0f ? 1.0f : 0.0f;
        v.x += gxh * movable;
        v.y += gyh * movable;
        p.x = v.x * dt + p.x;
        p.y = v.y * dt + p.y;
        p.z = v.z * dt + p.z;
        v.x += gxh * movable;
        v.y += gyh * movable;
        p.z = dm::wrapRotation(p.z);
        float c, s;
        dm::fastCosSin(p.z, c, s);
        pos[b] = p;
        vel[b] = v;
        rcs[b] = make_float2(c, s);
    }
}

// ColliderAABBBuilder.cpp:8-136.
__device__ inline float4 colliderAabb(uint4 sh, float2 dims, float4 ps, const float2* __restrict__ verts) {
    const float x = ps.x, y = ps.y;
    if (sh.x == AGX_SHAPE_CIRCLE) {
        const float r = dims.x;
        return make_float4(x - r, y - r, x + r, y + r);
    }
    if (sh.x == AGX_SHAPE_BOX) {
        const float ac = fabsf(ps.z), as = fabsf(ps.w);
        const float ex = ac * dims.x + as * dims.y;
        const float ey = as * dims.x + ac * dims.y;
        return make_float4(x - ex, y - ey, x + ex, y + ey);
    }
    const float c = ps.z, s = ps.w;
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    for (uint32_t v = 0; v < sh.z; v++) {
        const float2 p = verts[sh.y + v];
        const float wx = c * p.x - s * p.y;
        const float wy = s * p.x + c * p.y;
        mnx = fminf(mnx, wx); mxx = fmaxf(mxx, wx);
        mny = fminf(mny, wy); mxy = fmaxf(mxy, wy);
    }
    return make_float4(x + mnx, y + mny, x + mxx, y + mxy);
}

// Simulation.cpp:673-736 + ColliderAABBBuilder.cpp. prevW receives the collider origin the CCD
// sweep starts from: the post-update pose of the snapshot (first step of an update call) or
// the pose left by the previous step.
__global__ __launch_bounds__(256) void kPose(const float4* __restrict__ pos, const float2* __restrict__ rcs,
                                             const float4* __restrict__ prev, const uint32_t* __restrict__ colBody,
                                             const float2* __restrict__ comOff, const float* __restrict__ localRot,
                                             const uint4* __restrict__ shape, const float2* __restrict__ dims,
                                             const float2* __restrict__ verts, float4* __restrict__ pose,
                                             float* __restrict__ wRot, float2* __restrict__ prevW,
                                             float4* __restrict__ aabb, const uint32_t* __restrict__ list,
                                             uint32_t n, int prevFromSnapshot, int writePrev) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t c = list ? list[k] : k;
        const uint32_t b = colBody[c];
        const float4 p = pos[b];
        const float2 rc = rcs[b];
        const float2 off = comOff[c];
        if (writePrev) {
            if (prevFromSnapshot) {
                const float4 q = prev[b];
                prevW[c] = make_float2(q.x + (off.x * q.z - off.y * q.w), q.y + (off.x * q.w + off.y * q.z));
            } else {
                const float4 old = pose[c];
                prevW[c] = make_float2(old.x, old.y);
            }
        }
        const float wx = p.x + (off.x * rc.x - off.y * rc.y);
        const float wy = p.y + (off.x * rc.y + off.y * rc.x);
        const float wr = dm::wrapRotation(p.z + localRot[c]);
        float cs, sn;
        dm::fastCosSin(wr, cs, sn);
        const float4 ps = make_float4(wx, wy, cs, sn);
        pose[c] = ps;
        wRot[c] = wr;
        aabb[c] = colliderAabb(shape[c], dims[c], ps, verts);
    }
}

// ---------------------------------------------------------------- CCD
// CCD/ContinuousCollisionDetector.cpp:48-273 (casts of the core circle against inflated shapes).
constexpr float NO_HIT = 2.0f;

__device__ inline float castCircle(float mx, float my, float dx, float dy, float dd, float radius, float best) {
    const float c = mx * mx + my * my - radius * radius;
    const float b = mx * dx + my * dy;
    if (c <= 0.0f || b >= 0.0f) return NO_HIT;
    if (c >= -2.0f * b * best) return NO_HIT;
    const float disc = b * b - dd * c;
    if (disc < 0.0f) return NO_HIT;
    return c / (sqrtf(disc) - b);
}

__device__ inline float castBox(float q0x, float q0y, float dlx, float dly, float dd, float hw, float hh,
                                float radius, float best) {
    const float sx = copysignf(1.0f, q0x), sy = copysignf(1.0f, q0y);
    const float qx = q0x * sx, qy = q0y * sy;
    const float vx = dlx * sx, vy = dly * sy;
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
    const float t = fmaxf(tx, ty);
    const float px = qx + vx * t, py = qy + vy * t;
    if (tx >= ty ? fabsf(py) > h : fabsf(px) > w) return NO_HIT;
    if (fabsf(px) > hw && fabsf(py) > hh)
        return castCircle(qx - copysignf(hw, px), qy - copysignf(hh, py), vx, vy, dd, radius, best);
    return t;
}

__device__ inline float castPolygon(float q0x, float q0y, float dlx, float dly, float dd, const float2* __restrict__ v,
                                    const float2* __restrict__ nrm, uint32_t count, float radius, float best) {
    float tEnter = 0.0f, tExit = best;
    uint32_t enterEdge = 0;
    for (uint32_t i = 0; i < count; i++) {
        const float2 n = nrm[i];
        const float speed = n.x * dlx + n.y * dly;
        const float gap = (n.x * (q0x - v[i].x) + n.y * (q0y - v[i].y)) - radius;
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
            tExit = fminf(tExit, -gap / speed);
        }
    }
    if (tEnter >= tExit) return NO_HIT;
    uint32_t edge = enterEdge;
    const bool startedInside = tEnter == 0.0f;
    if (startedInside) {
        float maxGap = -FLT_MAX;
        for (uint32_t i = 0; i < count; i++) {
            const float gap = (nrm[i].x * (q0x - v[i].x) + nrm[i].y * (q0y - v[i].y)) - radius;
            if (gap > maxGap) { maxGap = gap; edge = i; }
        }
    }
    const float px = q0x + dlx * tEnter, py = q0y + dly * tEnter;
    const uint32_t next = edge + 1 == count ? 0 : edge + 1;
    const float2 a = v[edge], b = v[next];
    const float abx = b.x - a.x, aby = b.y - a.y;
    const float u = (px - a.x) * abx + (py - a.y) * aby;
    if (u >= 0.0f && u <= abx * abx + aby * aby) return startedInside ? NO_HIT : tEnter;
    const float2 corner = u < 0.0f ? a : b;
    return castCircle(q0x - corner.x, q0y - corner.y, dlx, dly, dd, radius, tExit);
}

__device__ inline float castCore(uint4 sh, float2 dims, float4 ps, const float2* __restrict__ verts,
                                 const float2* __restrict__ normals, float sx, float sy, float dx, float dy,
                                 float dd, float core, float best) {
    const float cs = ps.z, sn = ps.w;
    const float rx = sx - ps.x, ry = sy - ps.y;
    const float q0x = cs * rx + sn * ry, q0y = cs * ry - sn * rx;
    const float dlx = cs * dx + sn * dy, dly = cs * dy - sn * dx;
    if (sh.x == AGX_SHAPE_CIRCLE) return castCircle(q0x, q0y, dlx, dly, dd, dims.x + core, best);
    if (sh.x == AGX_SHAPE_BOX) return castBox(q0x, q0y, dlx, dly, dd, dims.x, dims.y, core, best);
    return castPolygon(q0x, q0y, dlx, dly, dd, verts + sh.y, normals + sh.y, sh.z, core, best);
}

// CCD/ContinuousCollisionDetector.cpp:307-440. Queries run in the previous step's Morton order
// of dynamic colliders (spatially coherent waves); each lane walks the static LBVH with its
// own stack, visiting nodes whose AABB touches the swept core AABB and casting against
// static colliders whose AABB strictly overlaps it. Earliest hit per body by atomic min.
__global__ __launch_bounds__(256) void kCcd(const uint32_t* __restrict__ order, uint32_t n,
                                            const uint32_t* __restrict__ colBody, const uint32_t* __restrict__ colScene,
                                            const float4* __restrict__ pos, const float4* __restrict__ prev,
                                            const float2* __restrict__ prevW, const float4* __restrict__ pose,
                                            const float2* __restrict__ extent, const uint4* __restrict__ shape,
                                            const float2* __restrict__ dims, const float2* __restrict__ verts,
                                            const float2* __restrict__ normals, Tree st, float fastFactor,
                                            float coreFraction, float* __restrict__ hitTime, Counters* ctr) {
    __shared__ uint32_t stackMem[256 * 48];
    uint32_t* stack = stackMem + threadIdx.x * 48;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || st.n == 0) return;
    const uint32_t c = order[k];
    const uint32_t body = colBody[c];
    const float4 p = pos[body];
    const float4 q = prev[body];
    const float bdx = p.x - q.x, bdy = p.y - q.y;
    const float bd2 = bdx * bdx + bdy * bdy;
    if (bd2 <= 1e-8f) return;
    const float minExtent = extent[c].x;
    if (minExtent <= 0.0f) return;
    const float threshold = fastFactor * minExtent;
    const float thresholdSq = threshold * threshold;
    if (bd2 <= thresholdSq) return;
    const float2 s = prevW[c];
    const float4 e = pose[c];
    const float dx = e.x - s.x, dy = e.y - s.y;
    const float dd = dx * dx + dy * dy;
    if (dd <= thresholdSq) return;
    const float core = coreFraction * minExtent;
    const float sMinX = fminf(s.x, e.x) - core, sMaxX = fmaxf(s.x, e.x) + core;
    const float sMinY = fminf(s.y, e.y) - core, sMaxY = fmaxf(s.y, e.y) + core;
    const uint32_t scene = colScene[c];
    const uint32_t sf = st.sceneStart[scene], sl = st.sceneStart[scene + 1];
    if (sf == sl) return;
    const uint32_t leafBase = st.n - 1;
    float best = 1.0f;
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const uint32_t node = stack[--sp];
        const float4 nb = st.nodeAabb[node];
        if (node >= leafBase) {
            const uint32_t j = node - leafBase;
            if (j < sf || j >= sl) continue;
            if (!(sMinX < nb.z && sMaxX > nb.x && sMinY < nb.w && sMaxY > nb.y)) continue;
            const uint32_t sc = st.vals[j];
            best = fminf(best, castCore(shape[sc], dims[sc], pose[sc], verts, normals, s.x, s.y, dx, dy, dd, core, best));
            continue;
        }
        const uint2 rg = st.range[node];
        if (rg.y < sf || rg.x >= sl) continue;
        if (nb.z < sMinX || nb.x > sMaxX || nb.w < sMinY || nb.y > sMaxY) continue;
        if (sp + 2 > 48) { atomicOr(&ctr->overflow, 2u); break; }
        const uint2 ch = st.child[node];
        stack[sp++] = ch.y;
        stack[sp++] = ch.x;
    }
    bool first = false;   // this cast gave the body its first hit of the step
    if (best < 1.0f) {
        const uint32_t old = atomicMin(reinterpret_cast<uint32_t*>(hitTime) + body, __float_as_uint(best));
        first = __uint_as_float(old) >= 1.0f;
    }
    // one update of the hit counter per wave (updates of one address from every lane queue there)
    const uint64_t hits = __ballot(first);
    if (first && (threadIdx.x & (WAVE - 1)) == (uint32_t)(__ffsll((unsigned long long)hits) - 1))
        atomicAdd(&ctr->ccdHits, (uint32_t)__popcll(hits));
}

// CCD/ContinuousCollisionDetector.cpp:513-529: move hit colliders back along their body's
// displacement and refresh their AABBs (bodies are moved by kCcdShiftBodies afterwards).
__global__ __launch_bounds__(256) void kCcdShiftColliders(const uint32_t* __restrict__ list, uint32_t n,
                                                          const uint32_t* __restrict__ colBody,
                                                          const float4* __restrict__ pos, const float4* __restrict__ prev,
                                                          const float* __restrict__ hitTime, const uint4* __restrict__ shape,
                                                          const float2* __restrict__ dims, const float2* __restrict__ verts,
                                                          float4* __restrict__ pose, float4* __restrict__ aabb) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t c = list[k];
        const uint32_t b = colBody[c];
        const float t = hitTime[b];
        if (!(t < 1.0f)) continue;
        const float back = t - 1.0f;
        const float4 p = pos[b];
        const float4 q = prev[b];
        const float shX = (p.x - q.x) * back, shY = (p.y - q.y) * back;
        float4 ps = pose[c];
        ps.x += shX;
        ps.y += shY;
        pose[c] = ps;
        aabb[c] = colliderAabb(shape[c], dims[c], ps, verts);
    }
}

__global__ __launch_bounds__(256) void kCcdShiftBodies(const uint32_t* __restrict__ list, uint32_t n,
                                                       float4* __restrict__ pos, const float4* __restrict__ prev,
                                                       float* __restrict__ hitTime) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        const float t = hitTime[b];
        if (!(t < 1.0f)) continue;
        const float back = t - 1.0f;
        float4 p = pos[b];
        const float4 q = prev[b];
        const float shX = (p.x - q.x) * back, shY = (p.y - q.y) * back;
        p.x += shX;
        p.y += shY;
        pos[b] = p;
        hitTime[b] = 1.0f;
    }
}

// Simulation::refreshSpatialData (Simulation.cpp:561-579): cos/sin of every body's rotation.
__global__ __launch_bounds__(256) void kRefreshCosSin(const float4* __restrict__ pos, float2* __restrict__ rcs, uint32_t n) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        float c, s;
        dm::fastCosSin(pos[b].z, c, s);
        rcs[b] = make_float2(c, s);
    }
}

inline uint32_t gridFor(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 8192u)); }

}  // namespace

hipError_t launchIntegrate(hipStream_t st, DeviceWorld& w, float dt, float gx, float gy, int snapshot) {
    if (w.nDynBodies == 0) return hipSuccess;
    const float halfDT = dt * 0.5f;
    hipLaunchKernelGGL(kIntegrate, dim3(gridFor(w.nDynBodies)), dim3(256), 0, st, w.pos, w.vel, w.rcs, w.inv, w.prev,
                       w.dynBodies, w.nDynBodies, dt, gx * halfDT, gy * halfDT, snapshot);
    return hipGetLastError();
}

hipError_t launchRefreshCosSin(hipStream_t st, DeviceWorld& w) {
    if (w.N == 0) return hipSuccess;
    hipLaunchKernelGGL(kRefreshCosSin, dim3(gridFor(w.N)), dim3(256), 0, st, w.pos, w.rcs, w.N);
    return hipGetLastError();
}

hipError_t launchPose(hipStream_t st, DeviceWorld& w, int all, int prevFromSnapshot, int writePrev) {
    const uint32_t n = all ? w.C : w.nDyn;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kPose, dim3(gridFor(n)), dim3(256), 0, st, w.pos, w.rcs, w.prev, w.colBody, w.comOff, w.localRot,
                       w.shape, w.dims, w.verts, w.pose, w.wRot, w.prevW, w.aabb, all ? nullptr : w.dyn.list, n,
                       prevFromSnapshot, writePrev);
    return hipGetLastError();
}

hipError_t launchCcd(hipStream_t st, DeviceWorld& w, float fastFactor, float coreFraction) {
    if (w.nStat == 0 || w.nDyn == 0) return hipSuccess;
    // dyn.vals holds the last Morton order of the dynamic colliders (built at creation and every step).
    hipLaunchKernelGGL(kCcd, dim3((w.nDyn + 255) / 256), dim3(256), 0, st, w.dyn.vals, w.nDyn, w.colBody, w.colScene,
                       w.pos, w.prev, w.prevW, w.pose, w.extent, w.shape, w.dims, w.verts, w.normals, w.stat,
                       fastFactor, coreFraction, w.hitTime, w.ctr);
    hipLaunchKernelGGL(kCcdShiftColliders, dim3(gridFor(w.nDyn)), dim3(256), 0, st, w.dyn.list, w.nDyn, w.colBody,
                       w.pos, w.prev, w.hitTime, w.shape, w.dims, w.verts, w.pose, w.aabb);
    hipLaunchKernelGGL(kCcdShiftBodies, dim3(gridFor(w.nDynBodies)), dim3(256), 0, st, w.dynBodies, w.nDynBodies,
                       w.pos, w.prev, w.hitTime);
    return hipGetLastError();
}

}  // namespace agx
