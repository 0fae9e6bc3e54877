// Contact solver (Solving/BodyCollisionSolver.cpp) and islands (IslandManager.cpp).
//
// Colouring replaces SolvingPlanner's waves (Solving/SolvingPlanner.cpp:141-310): constraints
// of one colour share no dynamic body (static bodies are never written), so a colour is solved
// in parallel with the same result as any sequential order of it; colours run in sequence.
#include "agx_internal.h"

namespace agx {

namespace {

constexpr uint16_t UNCOLOURED = 0xFFFFu;
constexpr uint32_t FIRST_FIT = 128;   // colours chosen first-fit from two 64-bit masks per body

__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

__device__ inline bool isDynamic(const float2* inv, uint32_t b) { return inv[b].x != 0.0f; }

// ---------------------------------------------------------------- colouring + constraint prep
// Round r: every uncoloured manifold claims its dynamic bodies with a priority stamped with r
// (atomic max), so claims of earlier rounds always lose. A manifold holding all its claims is
// alone at its bodies this round and takes the lowest colour free at both (first fit).
__global__ __launch_bounds__(256) void kColourPrep(const agx_manifold* __restrict__ man, Counters* __restrict__ ctr,
                                                   const uint32_t* __restrict__ colBody, const float2* __restrict__ inv,
                                                   const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                   const uint4* __restrict__ shape, const float4* __restrict__ materials,
                                                   uint16_t* __restrict__ colourId, unsigned long long* __restrict__ claim,
                                                   unsigned long long* __restrict__ used, uint32_t* __restrict__ order,
                                                   Constraints cs) {
    uint32_t gen = 0;
    uint32_t* bar = &ctr->barrier[0];
    const uint32_t M = ctr->manifolds;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t m = tid; m < M; m += stride) {
        colourId[m] = UNCOLOURED;
        const uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
        if (isDynamic(inv, a)) { claim[a] = 0; used[2 * a] = 0; used[2 * a + 1] = 0; }
        if (isDynamic(inv, b)) { claim[b] = 0; used[2 * b] = 0; used[2 * b + 1] = 0; }
    }
    gridBarrier(bar, gridDim.x, gen);
    uint32_t rounds = 0;
    for (uint32_t r = 0; r < MAX_COLOUR_ROUNDS; r++) {
        rounds = r + 1;
        for (uint32_t m = tid; m < M; m += stride) {
            if (colourId[m] != UNCOLOURED) continue;
            const unsigned long long prio = ((unsigned long long)(r + 1) << 48) |
                                            ((unsigned long long)(mix32(m) & 0xFFFFu) << 32) | m;
            const uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
            if (isDynamic(inv, a)) atomicMax(claim + a, prio);
            if (isDynamic(inv, b)) atomicMax(claim + b, prio);
        }
        gridBarrier(bar, gridDim.x, gen);
        uint32_t left = 0;
        for (uint32_t m = tid; m < M; m += stride) {
            if (colourId[m] != UNCOLOURED) continue;
            const unsigned long long prio = ((unsigned long long)(r + 1) << 48) |
                                            ((unsigned long long)(mix32(m) & 0xFFFFu) << 32) | m;
            const uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
            const bool da = isDynamic(inv, a), db = isDynamic(inv, b);
            const bool win = (!da || claim[a] == prio) && (!db || claim[b] == prio);
            if (!win) { left++; continue; }
            const unsigned long long u0 = (da ? used[2 * a] : 0ull) | (db ? used[2 * b] : 0ull);
            const unsigned long long u1 = (da ? used[2 * a + 1] : 0ull) | (db ? used[2 * b + 1] : 0ull);
            uint32_t c;
            if (~u0) c = (uint32_t)__ffsll((unsigned long long)~u0) - 1;
            else if (~u1) c = 64 + (uint32_t)__ffsll((unsigned long long)~u1) - 1;
            else c = FIRST_FIT + r;   // all first-fit colours taken: this round's winners are independent
            colourId[m] = (uint16_t)c;
            if (c < 64) {
                const unsigned long long bit = 1ull << c;
                if (da) used[2 * a] |= bit;
                if (db) used[2 * b] |= bit;
            } else if (c < FIRST_FIT) {
                const unsigned long long bit = 1ull << (c - 64);
                if (da) used[2 * a + 1] |= bit;
                if (db) used[2 * b + 1] |= bit;
            }
        }
        for (uint32_t off = 32; off > 0; off >>= 1) left += __shfl_down(left, off, 64);
        if ((threadIdx.x & 63u) == 0 && left) atomicAdd(&ctr->remaining[r], left);
        gridBarrier(bar, gridDim.x, gen);
        if (__hip_atomic_load(&ctr->remaining[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        if (r + 1 == MAX_COLOUR_ROUNDS && tid == 0) atomicOr(&ctr->overflow, 4u);
    }
    // Colour histogram, scan, scatter into colour order with constraint prep.
    for (uint32_t m = tid; m < M; m += stride) {
        const uint16_t c = colourId[m];
        if (c == UNCOLOURED) continue;
        atomicAdd(&ctr->colourCount[c], 1u);
        atomicMax(&ctr->colours, (uint32_t)c + 1u);
    }
    gridBarrier(bar, gridDim.x, gen);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t K = ctr->colours;
        uint32_t run = 0;
        for (uint32_t k = 0; k < K; k++) { ctr->colourStart[k] = run; run += ctr->colourCount[k]; }
        ctr->colourStart[K] = run;
        ctr->colourRounds = rounds;
    }
    gridBarrier(bar, gridDim.x, gen);
    // computeConstraintData (BodyCollisionSolver.cpp:229-390).
    for (uint32_t m = tid; m < M; m += stride) {
        const uint16_t col = colourId[m];
        if (col == UNCOLOURED) continue;
        const uint32_t slot = ctr->colourStart[col] + atomicAdd(&ctr->colourFill[col], 1u);
        order[slot] = m;
        const agx_manifold d = man[m];
        const uint32_t bA = colBody[d.a], bB = colBody[d.b];
        const float4 pA = pos[bA], pB = pos[bB];
        const float2 iA = inv[bA], iB = inv[bB];
        const float nx = d.normal_x, ny = d.normal_y;
        const float penBase = d.depth + ((pB.x - pA.x) * nx + (pB.y - pA.y) * ny) - 0.001f;
        const float4 mA = materials[shape[d.a].w], mB = materials[shape[d.b].w];
        const float e = (mA.x + mB.x) * 0.5f;
        const float sf = sqrtf(fmaxf(0.0f, mA.y * mB.y));
        const float df = sqrtf(fmaxf(0.0f, mA.z * mB.z));
        const float effMass = iA.x + iB.x;
        const float tx = -ny, ty = nx;
        const float4 vA = vel[bA], vB = vel[bB];
        float pr[2][7] = { { 0, 0, 0, 0, 0, 0, 0 }, { 0, 0, 0, 0, 0, 0, 0 } };
        for (uint32_t i = 0; i < d.count; i++) {
            const float cx = i == 0 ? d.p0_x : d.p1_x, cy = i == 0 ? d.p0_y : d.p1_y;
            const float rAx = cx - pA.x, rAy = cy - pA.y, rBx = cx - pB.x, rBy = cy - pB.y;
            const float rApx = -rAy, rApy = rAx, rBpx = -rBy, rBpy = rBx;
            const float rAn = rApx * nx + rApy * ny, rBn = rBpx * nx + rBpy * ny;
            const float rAt = rApx * tx + rApy * ty, rBt = rBpx * tx + rBpy * ty;
            const float nd = effMass + rAn * rAn * iA.y + rBn * rBn * iB.y;
            const float td = effMass + rAt * rAt * iA.y + rBt * rBt * iB.y;
            const float negVn = -(((vB.x - vA.x) * nx + (vB.y - vA.y) * ny) + rBn * vB.z - rAn * vA.z);
            pr[i][0] = rAn; pr[i][1] = rBn; pr[i][2] = rAt; pr[i][3] = rBt;
            pr[i][4] = nd > 0.0f ? 1.0f / nd : 0.0f;
            pr[i][5] = td > 0.0f ? 1.0f / td : 0.0f;
            pr[i][6] = negVn > 0.0f ? e * negVn : 0.0f;
        }
        cs.idx[slot] = make_uint4(bA, bB, d.count, m);
        cs.nrm[slot] = make_float4(nx, ny, sf, df);
        cs.inv[slot] = make_float4(iA.x, iB.x, iA.y, iB.y);
        cs.r0[slot] = make_float4(pr[0][0], pr[0][1], pr[0][2], pr[0][3]);
        cs.r1[slot] = make_float4(pr[1][0], pr[1][1], pr[1][2], pr[1][3]);
        cs.k0[slot] = make_float4(pr[0][4], pr[0][5], pr[0][6], pr[1][4]);
        cs.k1[slot] = make_float4(pr[1][5], pr[1][6], penBase, 0.0f);
        cs.imp[slot] = make_float4(d.jn0, d.jt0, d.count > 1 ? d.jn1 : 0.0f, d.count > 1 ? d.jt1 : 0.0f);
    }
}

// ---------------------------------------------------------------- sweeps
struct Vel { float vx, vy, w; };

// applyWarmStarting (BodyCollisionSolver.cpp:416-490).
__device__ inline void warmOne(const Constraints& cs, uint32_t c, float4* __restrict__ vel) {
    const uint4 id = cs.idx[c];
    const float4 nr = cs.nrm[c], iv = cs.inv[c], r0 = cs.r0[c], r1 = cs.r1[c], j = cs.imp[c];
    float4 a = vel[id.x], b = vel[id.y];
    const float nx = nr.x, ny = nr.y, tx = -ny, ty = nx;
    const float jn0 = j.x, jt0 = j.y, jn1 = j.z, jt1 = j.w;
    const float sn = jn0 + jn1, st = jt0 + jt1;
    const float sx = sn * nx + st * tx, sy = sn * ny + st * ty;
    a.x = a.x - sx * iv.x; a.y = a.y - sy * iv.x;
    b.x = b.x + sx * iv.y; b.y = b.y + sy * iv.y;
    a.z -= (jn0 * r0.x + jt0 * r0.z + jn1 * r1.x + jt1 * r1.z) * iv.z;
    b.z += (jn0 * r0.y + jt0 * r0.w + jn1 * r1.y + jt1 * r1.w) * iv.w;
    if (iv.x > 0.0f) vel[id.x] = a;
    if (iv.y > 0.0f) vel[id.y] = b;
}

// solveVelocityConstraints (BodyCollisionSolver.cpp:492-928): one point = clamped normal step
// (skipped while separating with zero impulse) + iterative friction; two points = 2x2 block LCP
// (4 cases, iterative fallback) + iterative friction per point. Returns sum of squared
// accumulator changes and of squared accumulators through dsq/lsq.
__device__ inline void velocityOne(const Constraints& cs, uint32_t c, float4* __restrict__ vel, float& dsq, float& lsq) {
    const uint4 id = cs.idx[c];
    const float4 nr = cs.nrm[c], iv = cs.inv[c], r0 = cs.r0[c], k0 = cs.k0[c];
    float4 j = cs.imp[c];
    const float4 j0 = j;
    float4 A = vel[id.x], B = vel[id.y];
    const float nx = nr.x, ny = nr.y, tx = -ny, ty = nx;
    const float sF = nr.z, dF = nr.w;
    const float iMA = iv.x, iMB = iv.y, iIA = iv.z, iIB = iv.w;
    auto friction = [&](float rAt, float rBt, float tMass, float& acc, float jn) {
        const float slip = ((B.x - A.x) * tx + (B.y - A.y) * ty) + rBt * B.z - rAt * A.z;
        const float jt = -slip * tMass;
        const float old = acc;
        const float target = old + jt;
        if (fabsf(target) <= jn * sF) acc = target;
        else { const float mx = jn * dF; acc = fminf(fmaxf(target, -mx), mx); }
        const float dj = acc - old;
        const float ix = dj * tx, iy = dj * ty;
        A.x = A.x - ix * iMA; A.y = A.y - iy * iMA; A.z -= dj * rAt * iIA;
        B.x = B.x + ix * iMB; B.y = B.y + iy * iMB; B.z += dj * rBt * iIB;
    };
    if (id.z == 1) {
        const float rAn = r0.x, rBn = r0.y;
        float jnSingle = 0.0f;
        const float vn = ((B.x - A.x) * nx + (B.y - A.y) * ny) + rBn * B.z - rAn * A.z;
        if (!(vn > 0.0f && j.x == 0.0f)) {
            const float jn = (k0.z - vn) * k0.x;
            const float old = j.x;
            j.x = fmaxf(0.0f, old + jn);
            const float dj = j.x - old;
            jnSingle = j.x;
            const float ix = dj * nx, iy = dj * ny;
            A.x = A.x - ix * iMA; A.y = A.y - iy * iMA; A.z -= dj * rAn * iIA;
            B.x = B.x + ix * iMB; B.y = B.y + iy * iMB; B.z += dj * rBn * iIB;
        }
        friction(r0.z, r0.w, k0.y, j.y, jnSingle);
    } else {
        const float4 r1 = cs.r1[c], k1 = cs.k1[c];
        const float effMass = iMA + iMB;
        float jnArr0, jnArr1;
        const float dvn = (B.x - A.x) * nx + (B.y - A.y) * ny;
        const float vn0 = dvn + r0.y * B.z - r0.x * A.z;
        const float vn1 = dvn + r1.y * B.z - r1.x * A.z;
        const float K00 = effMass + r0.x * r0.x * iIA + r0.y * r0.y * iIB;
        const float K11 = effMass + r1.x * r1.x * iIA + r1.y * r1.y * iIB;
        const float K01 = effMass + r0.x * r1.x * iIA + r0.y * r1.y * iIB;
        const float b0 = k0.z - vn0 + K00 * j.x + K01 * j.z;
        const float b1 = k1.y - vn1 + K01 * j.x + K11 * j.z;
        float x0 = 0.0f, x1 = 0.0f;
        bool ok = false;
        {
            const float EPS = 1e-6f;
            const float det = K00 * K11 - K01 * K01;
            const float maxK = fmaxf(K00, K11);
            if (maxK * maxK < 1000.0f * det) {
                const float invDet = 1.0f / det;
                const float s0 = (K11 * b0 - K01 * b1) * invDet;
                const float s1 = (K00 * b1 - K01 * b0) * invDet;
                if (s0 >= -EPS && s1 >= -EPS) { x0 = fmaxf(s0, 0.0f); x1 = fmaxf(s1, 0.0f); ok = true; }
            }
            if (!ok && K11 > 0.0f) {
                const float s1 = b1 / K11;
                if (s1 >= -EPS && (K01 * s1 - b0) >= -EPS) { x0 = 0.0f; x1 = fmaxf(s1, 0.0f); ok = true; }
            }
            if (!ok && K00 > 0.0f) {
                const float s0 = b0 / K00;
                if (s0 >= -EPS && (K01 * s0 - b1) >= -EPS) { x0 = fmaxf(s0, 0.0f); x1 = 0.0f; ok = true; }
            }
            if (!ok && b0 <= EPS && b1 <= EPS) { x0 = 0.0f; x1 = 0.0f; ok = true; }
        }
        if (!ok) {
            // Iterative normal fallback, point 0 then point 1.
            {
                const float vn = ((B.x - A.x) * nx + (B.y - A.y) * ny) + r0.y * B.z - r0.x * A.z;
                const float jn = (k0.z - vn) * k0.x;
                const float old = j.x;
                j.x = fmaxf(0.0f, old + jn);
                const float dj = j.x - old;
                jnArr0 = j.x;
                const float ix = dj * nx, iy = dj * ny;
                A.x = A.x - ix * iMA; A.y = A.y - iy * iMA; A.z -= dj * r0.x * iIA;
                B.x = B.x + ix * iMB; B.y = B.y + iy * iMB; B.z += dj * r0.y * iIB;
            }
            {
                const float vn = ((B.x - A.x) * nx + (B.y - A.y) * ny) + r1.y * B.z - r1.x * A.z;
                const float jn = (k1.y - vn) * k0.w;
                const float old = j.z;
                j.z = fmaxf(0.0f, old + jn);
                const float dj = j.z - old;
                jnArr1 = j.z;
                const float ix = dj * nx, iy = dj * ny;
                A.x = A.x - ix * iMA; A.y = A.y - iy * iMA; A.z -= dj * r1.x * iIA;
                B.x = B.x + ix * iMB; B.y = B.y + iy * iMB; B.z += dj * r1.y * iIB;
            }
        } else {
            const float d0 = x0 - j.x, d1 = x1 - j.z;
            j.x = x0; j.z = x1;
            jnArr0 = x0; jnArr1 = x1;
            const float jnSum = d0 + d1;
            const float sA = jnSum * iMA, sB = jnSum * iMB;
            A.x = A.x - nx * sA; A.y = A.y - ny * sA;
            B.x = B.x + nx * sB; B.y = B.y + ny * sB;
            A.z -= (d0 * r0.x + d1 * r1.x) * iIA;
            B.z += (d0 * r0.y + d1 * r1.y) * iIB;
        }
        friction(r0.z, r0.w, k0.y, j.y, jnArr0);
        friction(r1.z, r1.w, k1.x, j.w, jnArr1);
    }
    cs.imp[c] = j;
    if (iMA > 0.0f) vel[id.x] = A;
    if (iMB > 0.0f) vel[id.y] = B;
    const float e0 = j.x - j0.x, e1 = j.y - j0.y;
    dsq += e0 * e0 + e1 * e1;
    lsq += j.x * j.x + j.y * j.y;
    if (id.z > 1) {
        const float e2 = j.z - j0.z, e3 = j.w - j0.w;
        dsq += e2 * e2 + e3 * e3;
        lsq += j.z * j.z + j.w * j.w;
    }
}

// solvePositionConstraints (BodyCollisionSolver.cpp:930-980): linear only.
__device__ inline void positionOne(const Constraints& cs, uint32_t c, float4* __restrict__ pos,
                                   const float2* __restrict__ inv) {
    const uint4 id = cs.idx[c];
    const float4 nr = cs.nrm[c];
    const float penBase = cs.k1[c].z;
    float4 A = pos[id.x], B = pos[id.y];
    const float pen = penBase - ((B.x - A.x) * nr.x + (B.y - A.y) * nr.y);
    if (pen <= 0.0f) return;
    const float iMA = inv[id.x].x, iMB = inv[id.y].x;
    const float total = pen / (iMA + iMB) * 0.5f;
    const float cA = iMA * total, cB = iMB * total;
    if (iMA > 0.0f) { A.x -= nr.x * cA; A.y -= nr.y * cA; pos[id.x] = A; }
    if (iMB > 0.0f) { B.x += nr.x * cB; B.y += nr.y * cB; pos[id.y] = B; }
}

__device__ inline void blockAddDouble(double* dst, float v) {
    for (uint32_t off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63u) == 0 && v != 0.0f) atomicAdd(dst, (double)v);
}

// Persistent solve: warm start, velocity sweeps, position sweeps; colours in sequence with a
// grid barrier after each; finally scatter the accumulators back to the manifolds.
__global__ __launch_bounds__(256) void kSolve(Counters* __restrict__ ctr, Constraints cs, float4* __restrict__ vel,
                                              float4* __restrict__ pos, const float2* __restrict__ inv,
                                              const uint32_t* __restrict__ order, agx_manifold* __restrict__ man,
                                              uint32_t vIt, uint32_t pIt, int trackResidual, float residualTarget,
                                              int trackErrors) {
    uint32_t gen = 0;
    uint32_t* bar = &ctr->barrier[1];
    const uint32_t K = ctr->colours;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = 0; k < K; k++) {
        const uint32_t lo = ctr->colourStart[k], hi = ctr->colourStart[k + 1];
        for (uint32_t c = lo + tid; c < hi; c += stride) warmOne(cs, c, vel);
        gridBarrier(bar, gridDim.x, gen);
    }
    uint32_t sweeps = 0;
    for (uint32_t it = 0; it < vIt && it < MAX_SWEEPS; it++) {
        float dsq = 0.0f, lsq = 0.0f;
        for (uint32_t k = 0; k < K; k++) {
            const uint32_t lo = ctr->colourStart[k], hi = ctr->colourStart[k + 1];
            for (uint32_t c = lo + tid; c < hi; c += stride) velocityOne(cs, c, vel, dsq, lsq);
            if (k + 1 < K) gridBarrier(bar, gridDim.x, gen);
        }
        sweeps = it + 1;
        if (trackResidual || residualTarget > 0.0f) {
            blockAddDouble(&ctr->residualNum[it], dsq);
            blockAddDouble(&ctr->residualDen[it], lsq);
        }
        gridBarrier(bar, gridDim.x, gen);
        if (residualTarget > 0.0f) {
            const double num = __hip_atomic_load(&ctr->residualNum[it], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const double den = __hip_atomic_load(&ctr->residualDen[it], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (sqrt(num) <= (double)residualTarget * fmax(sqrt(den), 1e-12)) break;
        }
    }
    if (tid == 0) ctr->sweeps = sweeps;
    for (uint32_t it = 0; it < pIt; it++) {
        for (uint32_t k = 0; k < K; k++) {
            const uint32_t lo = ctr->colourStart[k], hi = ctr->colourStart[k + 1];
            for (uint32_t c = lo + tid; c < hi; c += stride) positionOne(cs, c, pos, inv);
            gridBarrier(bar, gridDim.x, gen);
        }
    }
    const uint32_t total = K ? ctr->colourStart[K] : 0u;
    float app = 0.0f, pen = 0.0f;
    for (uint32_t c = tid; c < total; c += stride) {
        const float4 j = cs.imp[c];
        agx_manifold& m = man[order[c]];
        m.jn0 = j.x; m.jt0 = j.y;
        if (m.count > 1) { m.jn1 = j.z; m.jt1 = j.w; }
        if (trackErrors) {
            // computeConstraintErrors (BodyCollisionSolver.cpp:132-183).
            const uint4 id = cs.idx[c];
            const float4 nr = cs.nrm[c];
            const float4 A = pos[id.x], B = pos[id.y];
            pen += fmaxf(cs.k1[c].z - ((B.x - A.x) * nr.x + (B.y - A.y) * nr.y), 0.0f);
            const float4 vA = vel[id.x], vB = vel[id.y];
            const float4 r0 = cs.r0[c], r1 = cs.r1[c];
            const float dv = (vB.x - vA.x) * nr.x + (vB.y - vA.y) * nr.y;
            const float a0 = -(dv + r0.y * vB.z - r0.x * vA.z);
            app += a0 * (a0 > 0.0f ? 1.0f : 0.0f);
            if (id.z > 1) { const float a1 = -(dv + r1.y * vB.z - r1.x * vA.z); app += a1 * (a1 > 0.0f ? 1.0f : 0.0f); }
        }
    }
    if (trackErrors) {
        for (uint32_t off = 32; off > 0; off >>= 1) { app += __shfl_down(app, off, 64); pen += __shfl_down(pen, off, 64); }
        if ((threadIdx.x & 63u) == 0) { atomicAdd(&ctr->velocityError, app); atomicAdd(&ctr->positionError, pen); }
    }
}

// ---------------------------------------------------------------- islands
// Lock-free union-find over contact edges between dynamic bodies: the larger root is hooked
// under the smaller by compare-and-swap, so each component's root is its smallest body index.
// Island ids number the roots in ascending order (IslandManager.cpp:36-80 numbers DSU roots
// in index order; the partition is the same).
__device__ inline uint32_t ufFind(uint32_t* parent, uint32_t x) {
    uint32_t p = __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        const uint32_t gp = __hip_atomic_load(parent + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = gp;
    }
    return x;
}

__global__ __launch_bounds__(256) void kUfInit(uint32_t* parent, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) parent[i] = i;
}

__global__ __launch_bounds__(256) void kUfUnite(const agx_manifold* __restrict__ man, const Counters* __restrict__ ctr,
                                                const uint32_t* __restrict__ colBody, const float2* __restrict__ inv,
                                                uint32_t* parent) {
    const uint32_t M = ctr->manifolds;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        uint32_t a = colBody[man[m].a], b = colBody[man[m].b];
        if (!isDynamic(inv, a) || !isDynamic(inv, b)) continue;
        while (true) {
            a = ufFind(parent, a);
            b = ufFind(parent, b);
            if (a == b) break;
            if (a > b) { const uint32_t t = a; a = b; b = t; }
            const uint32_t old = atomicCAS(parent + b, b, a);
            if (old == b) break;
            b = old;
        }
    }
}

// Read-only root lookup: no path splitting, so concurrent lookups never rewrite parents and
// every body sees the final root (all unions finished in the previous kernel).
__device__ inline uint32_t ufRootOf(const uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) { x = p; p = parent[x]; }
    return x;
}

// root[b] (kept in islandId as scratch) and a root flag per dynamic body for the rank scan.
__global__ __launch_bounds__(256) void kUfRoots(const uint32_t* __restrict__ parent, const uint32_t* __restrict__ list,
                                                uint32_t n, uint32_t* __restrict__ rootFlag, uint32_t* __restrict__ rootOf) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        const uint32_t r = ufRootOf(parent, b);
        rootOf[b] = r;
        rootFlag[k] = (r == b) ? 1u : 0u;
    }
}

// Root rank -> island id for every dynamic body (the dynamic list is ascending, so the rank of a
// root in the list is its rank among roots); static bodies keep UINT32_MAX.
__global__ __launch_bounds__(256) void kIslandIds(const uint32_t* __restrict__ list, uint32_t n,
                                                  const uint32_t* __restrict__ rootScan,
                                                  const uint32_t* __restrict__ bodyToDyn, uint32_t* __restrict__ islandId) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t b = list[k];
        islandId[b] = rootScan[bodyToDyn[islandId[b]]];
    }
}

__global__ void kIslandCount(Counters* ctr) { ctr->islands = ctr->scratch[3]; }

inline uint32_t blocksFor(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 8192u)); }

}  // namespace

int cooperativeBlocksPerCU() {
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void*>(kColourPrep), 256, 0) != hipSuccess) a = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(kSolve), 256, 0) != hipSuccess) b = 1;
    return std::max(1, std::min(a, b));
}

hipError_t launchColourPrep(hipStream_t st, DeviceWorld& w) {
    void* args[] = { (void*)&w.man[w.cur], (void*)&w.ctr, (void*)&w.colBody, (void*)&w.inv, (void*)&w.pos, (void*)&w.vel,
                     (void*)&w.shape, (void*)&w.materials, (void*)&w.colourId, (void*)&w.claim, (void*)&w.used,
                     (void*)&w.order, (void*)&w.cs };
    hipError_t e = launchFillU32(st, &w.ctr->barrier[0], 0u, 1);
    if (e != hipSuccess) return e;
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kColourPrep), dim3(w.coopBlocks), dim3(256), args, 0, st);
}

hipError_t launchSolve(hipStream_t st, DeviceWorld& w, uint32_t vIt, uint32_t pIt, int trackResidual,
                       float residualTarget, int trackErrors) {
    void* args[] = { (void*)&w.ctr, (void*)&w.cs, (void*)&w.vel, (void*)&w.pos, (void*)&w.inv, (void*)&w.order,
                     (void*)&w.man[w.cur], (void*)&vIt, (void*)&pIt, (void*)&trackResidual, (void*)&residualTarget,
                     (void*)&trackErrors };
    hipError_t e = launchFillU32(st, &w.ctr->barrier[1], 0u, 1);
    if (e != hipSuccess) return e;
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kSolve), dim3(w.coopBlocks), dim3(256), args, 0, st);
}

hipError_t launchIslands(hipStream_t st, DeviceWorld& w) {
    hipLaunchKernelGGL(kUfInit, dim3(blocksFor(w.N)), dim3(256), 0, st, w.ufParent, w.N);
    hipLaunchKernelGGL(kUfUnite, dim3(blocksFor((uint32_t)std::min<uint64_t>(w.manCap, 1u << 22))), dim3(256), 0, st,
                       w.man[w.cur], w.ctr, w.colBody, w.inv, w.ufParent);
    if (w.nDynBodies == 0) return launchFillU32(st, &w.ctr->islands, 0u, 1);
    hipLaunchKernelGGL(kUfRoots, dim3(blocksFor(w.nDynBodies)), dim3(256), 0, st, w.ufParent, w.dynBodies, w.nDynBodies,
                       w.ufRoot, w.islandId);
    hipError_t e = launchScan(st, w.ufRoot, w.ufRoot, nullptr, &w.ctr->scratch[3], w.nDynBodies, w.scanBlock, w.scanBlockCap);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kIslandIds, dim3(blocksFor(w.nDynBodies)), dim3(256), 0, st, w.dynBodies, w.nDynBodies, w.ufRoot,
                       w.bodyToDyn, w.islandId);
    hipLaunchKernelGGL(kIslandCount, dim3(1), dim3(1), 0, st, w.ctr);
    return hipGetLastError();
}

}  // namespace agx
