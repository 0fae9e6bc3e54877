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
                                                  uint32_t n, float dt, float gxh, float gyh, int snapshot,
                                                  const uint32_t* __restrict__ sleepId, int part) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        float4 p = pos[b];
        float4 v = vel[b];
        const float2 rc = rcs[b];
        if (snapshot && part != 2) prev[b] = make_float4(p.x, p.y, rc.x, rc.y);
        if (sleepId && sleepId[b] != 0u) continue;   // sleeping bodies are not in the Dynamic range
        const float movable = inv[b].x != 0.0f ? 1.0f : 0.0f;
        if (part != 2) {   // integratePhase1_VelocityVerlet (Integrator.cpp:99-139)
            v.x += gxh * movable;
            v.y += gyh * movable;
        }
        if (part == 1) {   // joints run before the drift
            vel[b] = v;
            continue;
        }
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

// ColliderAABBBuilder.cpp:8-175.
__device__ inline float4 colliderAabb(uint4 sh, float2 dims, float4 ps, const float2* __restrict__ verts) {
    const float x = ps.x, y = ps.y;
    if (sh.x == AGX_SHAPE_CIRCLE) {
        const float r = dims.x;
        return make_float4(x - r, y - r, x + r, y + r);
    }
    if (sh.x == AGX_SHAPE_CAPSULE) {   // the core's half length along each axis, grown by the radius
        const float ex = fabsf(ps.z) * dims.x + dims.y;
        const float ey = fabsf(ps.w) * dims.x + dims.y;
        return make_float4(x - ex, y - ey, x + ex, y + ey);
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
// The core circle (radius rho) of a fast dynamic collider sweeps the segment from its previous
// to its current origin; it hits a static collider when it comes within rho of its outline, i.e.
// when its centre enters the outline grown by rho (the Minkowski sum with a disc). The time of
// impact is found in the static collider's frame. For a polygon, by feature candidates: the
// boundary of a convex outline grown by rho consists of each edge's line pushed out by rho (between
// the edge's end normals) and of discs of radius rho around the vertices, so for a start outside the
// grown outline the first entry is the earliest of those crossings. A box is a rounded rectangle,
// entered through its bounding rectangle. Times are fractions of the sweep in [0, 1); NO_HIT marks
// none before the caller's limit. Follows the behaviour of CCD/ContinuousCollisionDetector.cpp:48-273
// (starts inside the grown outline never hit; for boxes, starts inside its bounding rectangle).
constexpr float NO_HIT = 2.0f;

// First time t in [0, limit) at which |o + t m| = rho for a point offset o from a disc centre moving
// by m per sweep (mm = |m|^2); NO_HIT when it starts inside or on the disc, moves away, misses or
// arrives at or after limit. The first root of mm t^2 + 2 k t + g is taken as g / (sqrt(k^2 - mm g) - k),
// which has no cancellation for k < 0.
__host__ __device__ inline float sweepDisc(float ox, float oy, float mx, float my, float mm, float rho, float limit) {
    const float g = ox * ox + oy * oy - rho * rho;
    const float k = ox * mx + oy * my;
    if (!(g > 0.0f) || !(k < 0.0f)) return NO_HIT;
    const float disc = k * k - mm * g;
    if (disc < 0.0f) return NO_HIT;
    const float t = g / (sqrtf(disc) - k);
    return t < limit ? t : NO_HIT;
}

// Convex outline given by corner(i) and outward unit edge normal faceNormal(i) of edge
// corner(i) -> corner(i+1), in the collider's frame; q, m: the core's start and displacement in
// that frame.
template <class Corner, class FaceNormal>
__host__ __device__ inline float sweepOutline(float qx, float qy, float mx, float my, float mm, float rho, float limit,
                                     uint32_t count, Corner corner, FaceNormal faceNormal) {
    float hit = NO_HIT;
    bool beyondGrown = false;   // the start lies outside some edge's pushed-out line
    bool outsideHull = false;   // the start lies outside some edge's own line
    bool nearFeature = false;   // the start is within rho of an edge span or a vertex
    for (uint32_t i = 0; i < count; i++) {
        const float2 a = corner(i);
        const float2 nrm = faceNormal(i);
        const float ax = qx - a.x, ay = qy - a.y;
        const float height = nrm.x * ax + nrm.y * ay;          // start's distance beyond edge i's line
        const float rate = nrm.x * mx + nrm.y * my;            // its change over the whole sweep
        if (height > rho) {
            beyondGrown = true;
            if (!(rate < 0.0f)) return NO_HIT;                 // never crosses this pushed-out line: separated
            const float t = (height - rho) / -rate;
            if (t < fminf(hit, limit)) {
                const float2 b = corner(i + 1 == count ? 0 : i + 1);
                const float ex = b.x - a.x, ey = b.y - a.y;
                const float along = (ax + mx * t) * ex + (ay + my * t) * ey;
                if (along >= 0.0f && along <= ex * ex + ey * ey) hit = t;
            }
        } else if (height > 0.0f) {
            outsideHull = true;
            const float2 b = corner(i + 1 == count ? 0 : i + 1);
            const float ex = b.x - a.x, ey = b.y - a.y;
            const float along = ax * ex + ay * ey;
            if (along >= 0.0f && along <= ex * ex + ey * ey) nearFeature = true;
        }
        if (ax * ax + ay * ay <= rho * rho) nearFeature = true;
        hit = fminf(hit, sweepDisc(ax, ay, mx, my, mm, rho, fminf(hit, limit)));
    }
    const bool startsInside = !beyondGrown && (!outsideHull || nearFeature);
    return startsInside ? NO_HIT : hit;
}

// Box of half extents (hw, hh) grown by rho: a rounded rectangle inside the rectangle of half
// extents (W, H) = (hw + rho, hh + rho). A start anywhere within that rectangle, the corner squares
// outside the rounding included, counts as inside and never hits (CCD/ContinuousCollisionDetector.cpp:
// 63-106 behaves so). Otherwise the start enters the rectangle when it has closed the last of its
// per-axis excesses (the slab entry); an entry point beyond both the box's own extents lies in a corner
// square, where only the disc of that corner can be met before the sweep leaves the square again.
__host__ __device__ inline float sweepRoundedBox(float qx, float qy, float lx, float ly, float mm, float hw, float hh,
                                                 float rho, float limit) {
    const float W = hw + rho, H = hh + rho;
    const float ax = fabsf(qx), ay = fabsf(qy);
    const float excessX = ax - W, excessY = ay - H;       // start's distance beyond the rectangle's faces
    if (!(excessX > 0.0f) && !(excessY > 0.0f)) return NO_HIT;
    // the sweep mirrored into the start's quadrant, where the box lies towards -x and -y
    const float sx = copysignf(1.0f, qx), sy = copysignf(1.0f, qy);
    const float ux = lx * sx, uy = ly * sy;
    const float closeX = -ux, closeY = -uy;               // speeds towards the box along each axis
    float enterX = 0.0f, enterY = 0.0f;
    if (excessX > 0.0f) {
        if (!(closeX > 0.0f)) return NO_HIT;
        enterX = excessX / closeX;
        if (!(enterX < limit)) return NO_HIT;
    }
    if (excessY > 0.0f) {
        if (!(closeY > 0.0f)) return NO_HIT;
        enterY = excessY / closeY;
        if (!(enterY < limit)) return NO_HIT;
    }
    const bool throughX = enterX >= enterY;               // the x faces are crossed last
    const float t = throughX ? enterX : enterY;
    const float ex = ax + ux * t, ey = ay + uy * t;       // entry point (mirrored)
    if (throughX ? fabsf(ey) > H : fabsf(ex) > W) return NO_HIT;   // passes beside the rectangle
    if (fabsf(ex) > hw && fabsf(ey) > hh)
        return sweepDisc(ax - copysignf(hw, ex), ay - copysignf(hh, ey), ux, uy, mm, rho, limit);
    return t;
}

// Capsule of core half length hl along x grown by rho (R = capsule radius + rho): a stadium, the
// union of the slab |y| <= R over |x| <= hl and the discs of radius R at both core ends. The sweep
// enters through the slab face on the start's side when it reaches that face within the core's span
// (a start on or beyond the face, approaching it); otherwise the earlier of the end discs decides
// (CCD/ContinuousCollisionDetector.cpp:222-240 behaves so: a start inside the slab is tested
// against the end discs only).
__host__ __device__ inline float sweepCapsule(float qx, float qy, float lx, float ly, float mm, float hl, float R,
                                              float limit) {
    const float excess = fabsf(qy) - R;                    // start's distance beyond the face on its side
    const float closing = -copysignf(1.0f, qy) * ly;       // speed towards that face
    if (closing > 0.0f && excess >= 0.0f) {
        const float t = excess / closing;
        if (t < limit && fabsf(qx + lx * t) <= hl) return t;
    }
    const float first = sweepDisc(qx + hl, qy, lx, ly, mm, R, limit);
    return fminf(first, sweepDisc(qx - hl, qy, lx, ly, mm, R, fminf(first, limit)));
}

// Earliest time of impact (< limit) of the core with centre s, displacement m and radius rho
// against static collider pose ps (x, y, cos, sin), or NO_HIT. Worked in the collider's frame.
__host__ __device__ inline float sweepCore(uint4 sh, float2 dims, float4 ps, const float2* __restrict__ verts,
                                  const float2* __restrict__ normals, float sx, float sy, float mx, float my,
                                  float mm, float rho, float limit) {
    const float ox = sx - ps.x, oy = sy - ps.y;
    // into the collider's frame: rotate by the inverse of its rotation
    const float c = ps.z, sn = ps.w;
    const float qx = c * ox + sn * oy, qy = c * oy - sn * ox;
    const float lx = c * mx + sn * my, ly = c * my - sn * mx;
    if (sh.x == AGX_SHAPE_CIRCLE) return sweepDisc(qx, qy, lx, ly, mm, dims.x + rho, limit);
    if (sh.x == AGX_SHAPE_BOX) return sweepRoundedBox(qx, qy, lx, ly, mm, dims.x, dims.y, rho, limit);
    if (sh.x == AGX_SHAPE_CAPSULE) return sweepCapsule(qx, qy, lx, ly, mm, dims.x, dims.y + rho, limit);
    const float2* v = verts + sh.y;
    const float2* nr = normals + sh.y;
    return sweepOutline(qx, qy, lx, ly, mm, rho, limit, sh.z, [=](uint32_t i) { return v[i]; },
                        [=](uint32_t i) { return nr[i]; });
}

// The fast-mover test of CCD/ContinuousCollisionDetector.cpp:343-369: the body's origin and the
// collider's origin both moved by more than fastFactor x the collider's smallest extent.
__device__ inline bool ccdFast(uint32_t c, const uint32_t* __restrict__ colBody, const float4* __restrict__ pos,
                               const float4* __restrict__ prev, const float2* __restrict__ prevW,
                               const float4* __restrict__ pose, const float2* __restrict__ extent, float fastFactor) {
    const uint32_t body = colBody[c];
    const float4 p = pos[body];
    const float4 q = prev[body];
    const float bdx = p.x - q.x, bdy = p.y - q.y;
    const float bd2 = bdx * bdx + bdy * bdy;
    if (bd2 <= 1e-8f) return false;
    const float minExtent = extent[c].x;
    if (minExtent <= 0.0f) return false;
    const float threshold = fastFactor * minExtent;
    const float thresholdSq = threshold * threshold;
    if (bd2 <= thresholdSq) return false;
    const float2 s = prevW[c];
    const float4 e = pose[c];
    const float dx = e.x - s.x, dy = e.y - s.y;
    return dx * dx + dy * dy > thresholdSq;
}

// The fast dynamic colliders, in the previous step's Morton order of dynamic colliders within each
// block of 1,024 (spatially coherent), appended block by block to `fast`; ctr->ccdCandidates
// counts them. One counter update per block: updates of one address from every wave queue there.
constexpr uint32_t CCD_SELECT_THREADS = 1024;
__global__ __launch_bounds__(CCD_SELECT_THREADS) void kCcdSelect(
    const uint32_t* __restrict__ order, uint32_t n, const uint32_t* __restrict__ colBody,
    const float4* __restrict__ pos, const float4* __restrict__ prev, const float2* __restrict__ prevW,
    const float4* __restrict__ pose, const float2* __restrict__ extent, float fastFactor,
    uint32_t* __restrict__ fast, Counters* ctr) {
    constexpr uint32_t WAVES = CCD_SELECT_THREADS / WAVE;
    __shared__ uint32_t waveCount[WAVES];
    __shared__ uint32_t blockBase;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = k < n ? order[k] : 0u;
    const bool isFast = k < n && ccdFast(c, colBody, pos, prev, prevW, pose, extent, fastFactor);
    const uint64_t mask = __ballot(isFast);
    const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) waveCount[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t i = 0; i < WAVES; i++) {
            const uint32_t t = waveCount[i];
            waveCount[i] = total;
            total += t;
        }
        blockBase = total ? atomicAdd(&ctr->ccdCandidates, total) : 0u;
    }
    __syncthreads();
    if (isFast) fast[blockBase + waveCount[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = c;
}

// CCD/ContinuousCollisionDetector.cpp:307-440 for the fast colliders only: each lane walks the
// static LBVH with its own stack, visiting nodes whose AABB touches the swept core AABB and
// casting against static colliders whose AABB strictly overlaps it. Earliest hit per body by
// atomic min (independent of the order the casts run in). One wave per block, so that the few
// thousand packed waves spread over every CU.
constexpr uint32_t CCD_THREADS = 64;
__global__ __launch_bounds__(CCD_THREADS) void kCcd(const uint32_t* __restrict__ fast, const uint32_t* __restrict__ colBody,
                                            const uint32_t* __restrict__ colScene,
                                            const float2* __restrict__ prevW, const float4* __restrict__ pose,
                                            const float2* __restrict__ extent, const uint4* __restrict__ shape,
                                            const float2* __restrict__ dims, const float2* __restrict__ verts,
                                            const float2* __restrict__ normals, Tree st, float coreFraction,
                                            float* __restrict__ hitTime, Counters* ctr) {
    __shared__ uint32_t stackMem[CCD_THREADS * 48];
    uint32_t* stack = stackMem + threadIdx.x * 48;
    const uint32_t n = ctr->ccdCandidates;
    if (st.n == 0) return;
    const uint32_t leafBase = st.n - 1;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t k = base + threadIdx.x;
        float best = 1.0f;
        uint32_t body = 0;
        if (k < n) {
            const uint32_t c = fast[k];
            body = colBody[c];
            const float minExtent = extent[c].x;
            const float2 s = prevW[c];
            const float4 e = pose[c];
            const float dx = e.x - s.x, dy = e.y - s.y;
            const float dd = dx * dx + dy * dy;
            const float core = coreFraction * minExtent;
            const float sMinX = fminf(s.x, e.x) - core, sMaxX = fmaxf(s.x, e.x) + core;
            const float sMinY = fminf(s.y, e.y) - core, sMaxY = fmaxf(s.y, e.y) + core;
            const uint32_t scene = colScene[c];
            const uint32_t sf = st.sceneStart[scene], sl = st.sceneStart[scene + 1];
            uint32_t sp = 0;
            if (sf != sl) stack[sp++] = 0;
            while (sp) {
                const uint32_t node = stack[--sp];
                const float4 nb = st.nodeAabb[node];
                if (node >= leafBase) {
                    const uint32_t j = node - leafBase;
                    if (j < sf || j >= sl) continue;
                    if (!(sMinX < nb.z && sMaxX > nb.x && sMinY < nb.w && sMaxY > nb.y)) continue;
                    const uint32_t sc = st.vals[j];
                    best = fminf(best, sweepCore(shape[sc], dims[sc], pose[sc], verts, normals, s.x, s.y, dx, dy, dd, core,
                                                 best));
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

hipError_t launchIntegrate(hipStream_t st, DeviceWorld& w, float dt, float gx, float gy, int snapshot, int part) {
    if (w.nDynBodies == 0) return hipSuccess;
    const float halfDT = dt * 0.5f;
    hipLaunchKernelGGL(kIntegrate, dim3(gridFor(w.nDynBodies)), dim3(256), 0, st, w.pos, w.vel, w.rcs, w.inv, w.prev,
                       w.dynBodies, w.nDynBodies, dt, gx * halfDT, gy * halfDT, snapshot, w.sleepId, part);
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
    hipLaunchKernelGGL(kCcdSelect, dim3((w.nDyn + CCD_SELECT_THREADS - 1) / CCD_SELECT_THREADS), dim3(CCD_SELECT_THREADS),
                       0, st, w.dyn.vals, w.nDyn, w.colBody, w.pos, w.prev, w.prevW, w.pose, w.extent, fastFactor,
                       w.ccdFast, w.ctr);
    // the candidate count is known only on the device: enough one-wave blocks for 64 Ki
    // candidates at once, more in grid-stride turns
    hipLaunchKernelGGL(kCcd, dim3(std::min((w.nDyn + CCD_THREADS - 1) / CCD_THREADS, 1024u)), dim3(CCD_THREADS), 0, st,
                       w.ccdFast, w.colBody, w.colScene,
                       w.prevW, w.pose, w.extent, w.shape, w.dims, w.verts, w.normals, w.stat, coreFraction, w.hitTime,
                       w.ctr);
    hipLaunchKernelGGL(kCcdShiftColliders, dim3(gridFor(w.nDyn)), dim3(256), 0, st, w.dyn.list, w.nDyn, w.colBody,
                       w.pos, w.prev, w.hitTime, w.shape, w.dims, w.verts, w.pose, w.aabb);
    hipLaunchKernelGGL(kCcdShiftBodies, dim3(gridFor(w.nDynBodies)), dim3(256), 0, st, w.dynBodies, w.nDynBodies,
                       w.pos, w.prev, w.hitTime);
    return hipGetLastError();
}

}  // namespace agx
