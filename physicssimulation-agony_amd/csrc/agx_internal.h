// Internal declarations of the MI355X collide-and-solve library (device layout, kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#include "../../include/agx.h"

namespace agx {

constexpr uint32_t WAVE = 64;
constexpr uint32_t FIRST_FIT_COLOURS = 128;  // chosen first-fit from two 64-bit masks per body
constexpr uint32_t MAX_COLOURS = FIRST_FIT_COLOURS + 1;   // + one shared overflow colour
static_assert(MAX_COLOURS == AGX_MAX_COLOURS, "colour count of the C ABI");
constexpr uint32_t COLOUR_BLOCKS = 512;     // blocks of the colour histogram / scatter kernels
constexpr uint32_t MAX_FLOW_ITEMS = 64;     // constraints per lane in the dataflow kernels
constexpr uint32_t COLOUR_ALIGN = 64;       // solve positions of each colour start on a wave boundary
constexpr uint32_t MAX_SWEEPS = AGX_MAX_SWEEPS;
constexpr uint32_t TRAVERSAL_STACK = 64;
constexpr uint32_t SORT_TILE_SMALL = 1024;  // keys per radix-sort block (256 threads x 4) ...
constexpr uint32_t SORT_TILE_LARGE = 4096;  // ... and from SORT_LARGE_KEYS keys (256 x 16): fewer
constexpr uint32_t SORT_LARGE_KEYS = 1u << 18;   // tiles shorten the look-back, more keep CUs busy
constexpr uint32_t sortTileFor(uint32_t n) { return n >= SORT_LARGE_KEYS ? SORT_TILE_LARGE : SORT_TILE_SMALL; }
constexpr uint32_t BOUNDS_CHUNK = 4096;     // colliders per bounds block (one scene per chunk)

// Device counters of one step (zeroed at step start, copied to agx_stats at FINISH).
// Residual gate of the learned warm start: exponential averages of what a step cost with the
// cached warm start (0) and with the network's prediction (1). The cheaper method is used, and the
// other one runs once as a probe every `period` steps, and its cost replaces that method's stale
// average; a probe that loses doubles the period (up to GATE_PROBE_MAX), one that wins resets it.
// A step whose prediction cost more than the cached average is followed by a cached step at once.
constexpr uint32_t GATE_PROBE_PERIOD = 16;
constexpr uint32_t GATE_PROBE_MAX = 1024;
struct LearnGate {
    float cost[2];
    uint32_t use;            // method of the next step (1: learned)
    uint32_t sinceProbe;
    uint32_t steps[2];       // steps measured per method
    uint32_t period;         // steps between probes (0: GATE_PROBE_PERIOD)
    uint32_t probing;        // the next step is a probe
};

__host__ __device__ inline void gateUpdate(LearnGate& g, uint32_t used, float cost) {
    const uint32_t k = used ? 1u : 0u;
    // a probe replaces its method's average: the old one is stale by a whole period
    g.cost[k] = (g.steps[k] == 0 || g.probing) ? cost : 0.75f * g.cost[k] + 0.25f * cost;
    g.steps[k]++;
    if (g.period == 0) g.period = GATE_PROBE_PERIOD;
    if (g.steps[1] == 0) { g.use = 1; return; }
    if (g.steps[0] == 0) { g.use = 0; return; }
    const uint32_t best = g.cost[1] < g.cost[0] ? 1u : 0u;
    if (g.probing) {
        g.probing = 0;
        g.period = best == k ? GATE_PROBE_PERIOD : (g.period * 2 < GATE_PROBE_MAX ? g.period * 2 : GATE_PROBE_MAX);
    }
    uint32_t next = best;
    if (++g.sinceProbe >= g.period) {
        g.sinceProbe = 0;
        next = 1u - best;
        g.probing = 1;
    }
    if (k == 1 && cost > g.cost[0]) {   // this prediction did worse than the cached average
        next = 0;
        g.probing = 0;
    }
    g.use = next;
}

struct Counters {
    // persistent across steps (not cleared at step start)
    uint32_t prevManifolds;    // manifolds of the previous step (warm-start source)
    uint32_t hashCapacity;     // power of two >= 2 * prevManifolds (>= 16)
    LearnGate gate;
    unsigned long long solveSteps;   // steps solved since the world was created
    unsigned long long sweepsTotal;  // velocity sweeps run over those steps (per step: the most any scene ran)
    unsigned long long bodySweepsTotal;   // sum over those steps of dynamic bodies x the sweeps of their scene
    unsigned long long convergedBodiesTotal;   // sum over those steps of convergedBodies
    uint32_t eventCount;       // collision events recorded since the update/step call began (may exceed the capacity)
    uint32_t eventStep;        // physics steps with event recording since the call began
    uint32_t stashCount;       // manifolds of sleeping islands kept in the stash
    // per step
    uint32_t pairs;            // total pairs found (may exceed capacity -> overflow)
    uint32_t manifolds;
    uint32_t contacts;
    uint32_t colours;
    uint32_t islands;
    uint32_t ccdHits;
    uint32_t sweeps;
    uint32_t overflow;         // OVERFLOW_* bits
    uint32_t scratch[8];       // scan totals
    uint32_t colourStart[MAX_COLOURS + 1];   // solve position of each colour's first slot (wave-aligned)
    uint32_t colourCount[MAX_COLOURS];       // constraints of each colour (the rest of its slots are holes)
    double residualNum[MAX_SWEEPS];
    double residualDen[MAX_SWEEPS];
    float velocityError, positionError;
    uint32_t barrier[4];       // grid-barrier arrival counters (one per cooperative kernel)
    uint32_t learnRows;        // feature rows of this step (2 per manifold; 0 when nothing is learned)
    uint32_t learnPredictRows; // rows the network runs on (learnRows when the prediction is used, else 0)
    uint32_t learnUsed;        // this step's warm start came from the network
    uint32_t sampleBase;       // first training row this step claimed
    uint32_t learnCompactRows; // new-contacts mode: rows written (one per contact point without a cached impulse)
    uint32_t sceneCursor;      // per-scene solve: scenes taken by whole blocks
    uint32_t sceneCursor2;     // per-scene solve: scenes taken by lane groups
    uint32_t sceneWhole;       // per-scene solve: scenes of this step's order that whole blocks take
    uint32_t colourCursor;     // per-scene colouring: scenes taken by whole blocks
    uint32_t colourCursor2;    // per-scene colouring: scenes taken by lane groups
    uint32_t perScene;         // this step was solved scene by scene
    unsigned long long bodySweeps;   // this step: dynamic bodies x the sweeps of their scene
    unsigned long long convergedBodies;   // this step: dynamic bodies of scenes that met the residual target
    uint32_t convergedScenes;  // this step: scenes that met the residual target
    uint32_t ccdCandidates;    // this step: fast dynamic colliders the CCD casts
    uint32_t flowMet;          // this step: the dataflow solve met the residual target
    uint32_t eventStepTotal;   // this step: collision events recorded
    uint32_t wokeAny;          // this step: a manifold touched a sleeping body
    uint32_t sleptAny;         // this step: a body fell asleep
    uint32_t stashScanCount;   // this step: stash entries scanned by the wake (0 when nothing woke)
    uint32_t restoreTotal;     // this step: stashed manifolds restored
    uint32_t sleepScanCount;   // this step: manifolds scanned by the sleep (0 when nothing slept)
    uint32_t stashTotal;       // this step: manifolds moved to the stash
};

constexpr uint32_t OVERFLOW_PAIRS = 1u;       // pair buffer too small (error)
constexpr uint32_t OVERFLOW_STACK = 2u;       // traversal stack too small (error)
constexpr uint32_t OVERFLOW_COLOURING = 4u;   // a constraint found no free first-fit colour (still exact)
constexpr uint32_t OVERFLOW_HASH = 8u;        // warm-start hash full (error)
constexpr uint32_t OVERFLOW_STALL = 16u;      // a dataflow kernel hit its time limit (error)
constexpr uint32_t OVERFLOW_ITEMS = 32u;      // more constraints than the dataflow kernels hold (error)
constexpr uint32_t OVERFLOW_LDS = 64u;        // a scene larger than its per-scene LDS region was not solved (error)
static_assert(OVERFLOW_PAIRS == AGX_OVERFLOW_PAIRS && OVERFLOW_STACK == AGX_OVERFLOW_STACK &&
                  OVERFLOW_COLOURING == AGX_OVERFLOW_COLOURING && OVERFLOW_HASH == AGX_OVERFLOW_HASH &&
                  OVERFLOW_STALL == AGX_OVERFLOW_STALL && OVERFLOW_ITEMS == AGX_OVERFLOW_ITEMS &&
                  OVERFLOW_LDS == AGX_OVERFLOW_LDS,
              "overflow bits match the public header");

// Both children of an internal node in one 64-byte record, so a traversal step is one load.
// A leaf child has first == last (its sorted position), ref = its collider and body = the
// collider's body; an internal child has first < last and ref = its node index.
struct ChildInfo {
    float4 box;
    uint32_t first, last, ref, body;
};
struct NodeRecord {
    ChildInfo c[2];
};
static_assert(sizeof(NodeRecord) == 64, "node record is one 64-byte line");

constexpr uint32_t STAGE_PAIRS = 1024;   // pairs a traversal wave stages before the gather

// One BVH over a fixed collider list (dynamic or static), two-level Karras layout (tile trees
// joined by a tree over the tiles): internal nodes [0, n-1), leaves [n-1, 2n-1); node 0 is the
// root (a leaf when n == 1).
struct Tree {
    uint32_t n = 0;
    uint32_t sortTile = 0;         // keys per radix-sort block (sortTileFor(n))
    uint32_t sortBlocks = 0;
    uint32_t* keys = nullptr;      // sorted (scene << mortonBits | morton)
    uint32_t* vals = nullptr;      // sorted collider indices
    uint32_t* keysAlt = nullptr;
    uint32_t* valsAlt = nullptr;
    uint32_t* hist = nullptr;      // radix sort scratch: 4 x (sortBlocks + 1) x 256 + 4 words
    float4* nodeAabb = nullptr;    // 2n-1
    uint2* child = nullptr;        // n-1: unified child indices
    uint2* range = nullptr;        // n-1: first, last sorted leaf
    uint32_t* parent = nullptr;    // 2n-1
    NodeRecord* rec = nullptr;     // n-1 internal-node records (written by the refit)
    uint32_t* leafBody = nullptr;  // n
    uint32_t* sceneStart = nullptr;  // S+1: sorted leaf range of each scene
    uint32_t* list = nullptr;      // n: colliders of this tree (ascending index)
    uint32_t* chunkScene = nullptr;  // bounds chunks
    uint2* chunkRange = nullptr;
    uint32_t chunks = 0;
    float4* sceneBounds = nullptr;   // S: (minX, minY, maxX, maxY)
    bool boundsReady = false;        // sceneBounds already computed (no chunked bounds pass)
};

// Device-side set-up of a broad phase over device inputs (one block of BROAD_PREP_THREADS):
// dynamic and static collider lists in ascending index order, each tree's per-scene sorted ranges
// (S + 1 entries) and per-scene bounds (order-preserving int encoding, as the tree build keeps
// them); header[0..2] = dynamic count, static count, 1 when a scene index was >= S. sceneZero,
// when not null, receives n zeros (the scene of every collider when the caller gives none).
constexpr uint32_t BROAD_PREP_THREADS = 1024;
struct BroadPrep {
    uint32_t* dynList;
    uint32_t* statList;
    uint32_t* dynStart;
    uint32_t* statStart;
    int4* dynBounds;
    int4* statBounds;
    uint32_t* header;
    uint32_t* sceneZero;
};

// Constraint planes, one float4 per constraint per plane, in solve order (colour-major).
struct Constraints {
    uint4* idx;      // bodyA, bodyB, count, manifold
    uint4* dep;      // predecessor slot at A and at B (dataflow only), A dynamic, B dynamic
    float4* nrm;     // nx, ny, staticFriction, dynamicFriction
    float4* inv;     // invMassA, invMassB, invInertiaA, invInertiaB
    float4* r0;      // rAn0, rBn0, rAt0, rBt0
    float4* r1;      // rAn1, rBn1, rAt1, rBt1
    float4* k0;      // normalMass0, tangentMass0, bias0, normalMass1
    float4* k1;      // tangentMass1, bias1, penetrationBase, unused
    float4* imp;     // jn0, jt0, jn1, jt1
    float4* blk;     // two points: K00, K11, K01 of the 2x2 normal block and 1/det (NaN when the
                     // block solve is skipped as ill-conditioned); zero for one point
};

// Launch shape of the per-scene solve: `blocks` resident blocks (one per CU) first take whole
// scenes (at least `whole` of them), then split into `groups` lane groups of `groupBodies` LDS body
// records.
// The per-scene colouring likewise: colourWhole whole scenes (largest first), then colourGroups
// groups of colourGroupWords LDS words; layouts (bodies << 16) | manifolds kept in LDS.
struct SceneLaunch {
    int blocks = 0;
    uint32_t whole = 0, groups = 1, groupBodies = 0;
    int colourBlocks = 0;
    uint32_t colourWhole = 0, colourGroups = 1, colourLayoutWhole = 0, colourLayoutGroup = 0, colourGroupWords = 0;
};

// A joint as built (include/agx.h agx_joint_desc): bodies by world body index (motor: a only).
struct JointRec {
    uint32_t type, a, b, enableLimit;
    float4 anchors;    // local anchor A (x, y), local anchor B (x, y)
    float4 axisRest;   // local axis of A (x, y; unit), spring rest length, prismatic reference angle
    float4 params;     // stiffness, damping, lower limit, upper limit
    float2 motor;      // target speed, maximum torque
    float2 pad;
};
// A joint's constraint data of the step (the solver's ConstraintData), 20 floats.
struct JointData { float v[20]; };
// The step's joint schedule: velocity passes per type and each type's colours
// [typeColour[t], typeColour[t + 1]) in the colour ranges.
struct JointSchedule {
    uint32_t iterations[AGX_JOINT_TYPES];
    uint32_t typeColour[AGX_JOINT_TYPES + 1];
};

struct DeviceWorld {
    uint32_t S = 0, N = 0, C = 0, nDyn = 0, nStat = 0, nVerts = 0, nMat = 0;
    uint32_t sceneBits = 0, mortonBits = 30;
    // bodies
    float4* pos = nullptr;       // comX, comY, rotation, 0
    float4* vel = nullptr;       // vx, vy, angular velocity, 0
    float2* rcs = nullptr;       // cos, sin
    float2* inv = nullptr;       // invMass, invInertia (static iff invMass == 0)
    float4* prev = nullptr;      // CCD snapshot: comX, comY, cos, sin before integration
    uint32_t* bodyScene = nullptr;
    uint32_t* dynBodies = nullptr;   // dynamic body list (ascending)
    uint32_t* bodyToDyn = nullptr;   // body -> position in dynBodies (UINT32_MAX for static)
    uint32_t nDynBodies = 0;
    uint32_t* islandId = nullptr;
    uint32_t* ufParent = nullptr;
    uint32_t* ufRoot = nullptr;      // scan scratch (N)
    float* hitTime = nullptr;        // CCD, per body, 1 between steps
    uint32_t* ccdFast = nullptr;     // CCD: the step's fast dynamic colliders (nDyn entries)
    // colliders
    uint32_t* colBody = nullptr;
    uint32_t* colScene = nullptr;
    uint32_t* colSlot = nullptr;
    float2* comOff = nullptr;
    float* localRot = nullptr;
    uint4* shape = nullptr;          // type, polyStart, polyCount, material
    float2* dims = nullptr;          // p0, p1
    float2* extent = nullptr;        // minExtent, maxExtent
    float4* pose = nullptr;          // wX, wY, wCos, wSin
    float* wRot = nullptr;
    float2* prevW = nullptr;
    float4* aabb = nullptr;
    float2* verts = nullptr;
    float2* normals = nullptr;
    float4* materials = nullptr;     // elasticity, staticFriction, dynamicFriction, 0
    // broad phase
    Tree dyn, stat;
    uint64_t pairCap = 0;
    uint2* pairs = nullptr;
    uint32_t* waveCount = nullptr;   // per query wave (self + cross passes)
    uint2* stagedPairs = nullptr;    // per query wave: STAGE_PAIRS pairs
    uint32_t* waveOffset = nullptr;
    uint32_t queryWaves = 0;
    // narrow phase
    uint64_t manCap = 0;
    agx_manifold* slots = nullptr;   // per pair slot
    uint32_t* flags = nullptr;       // per pair: 1 if a manifold was written
    uint32_t* flagScan = nullptr;
    agx_manifold* man[2] = { nullptr, nullptr };   // current / previous (ping-pong)
    int cur = 0;
    unsigned long long* hashKeys = nullptr;
    uint32_t* hashVals = nullptr;
    uint64_t hashCapMax = 0;
    // solver: per-body incidence lists (CSR over bodies) drive a dataflow colouring and solve
    uint2* manBodies = nullptr;      // per manifold: body of A, body of B
    uint32_t* degree = nullptr;      // per body: manifolds touching it (dynamic bodies only)
    uint32_t* incStart = nullptr;    // per body: start of its list in inc (scan of degree)
    uint32_t* incFill = nullptr;     // per body: fill cursor
    uint32_t* inc = nullptr;         // 2 * manCap entries 2m + side, grouped by body
    uint32_t* incSorted = nullptr;   // the same lists sorted by solve position
    uint32_t* rank = nullptr;        // per entry 2m + side: rank in its body's list
    // Per-body records of the dataflow kernels: 64-bit words (sequence << 32 | 32-bit payload).
    // A word's sequence counts the operations finished on the body, so one atomic 64-bit load
    // tells a waiting constraint both that its turn has come and the current value.
    unsigned long long* colRec = nullptr;   // 4 words: first-fit colours taken (4 x 32 bits)
    uint32_t* groupDone = nullptr;   // 2 x solveGroups: per lane-group progress of the velocity and position flows
    uint32_t solveGroups = 0;
    uint16_t* colourId = nullptr;    // per manifold
    uint32_t* colourHist = nullptr;  // COLOUR_BLOCKS x MAX_COLOURS block histograms, then block bases
    uint32_t* posOf = nullptr;       // manifold -> solve position
    uint32_t* order = nullptr;       // solve position -> manifold (UINT32_MAX in padding holes)
    uint64_t solveCap = 0;           // solve positions: manCap + COLOUR_ALIGN * MAX_COLOURS
    float* blockRes = nullptr;       // per solve block: residual partial sums (2 x MAX_SWEEPS)
    // per-scene solve (scenes whose bodies fit one block's LDS)
    uint32_t* sceneBodyStart = nullptr;   // S + 1: first body of each scene (scenes are concatenated)
    uint32_t* sceneOrder = nullptr;       // scenes by decreasing body count (the blocks' work order)
    uint32_t* sceneDyn = nullptr;         // dynamic bodies per scene
    uint2* sceneRange = nullptr;          // S x MAX_COLOURS: first and one past the last slot of the scene in a colour
    float2* sceneRes = nullptr;           // S x MAX_SWEEPS: the scene's squared accumulator change and accumulators
    uint32_t* sceneSweeps = nullptr;      // velocity sweeps each scene ran this step
    uint2* sceneManRange = nullptr;       // S: first manifold and one past the last of the scene
    uint32_t* sceneIslands = nullptr;     // islands per scene (per-scene islands)
    uint32_t* sceneIslandOffset = nullptr;   // islands of all earlier scenes
    uint32_t sceneMaxBodies = 0;          // bodies of the largest scene
    uint32_t sceneLdsBytes = 0;           // dynamic LDS of the per-scene solve (0: some scene does not fit)
    uint32_t sceneColourLdsBytes = 0;     // dynamic LDS of the per-scene colouring
    uint32_t sceneColourCap = 0xFFFFu;    // most manifolds of a scene the colouring keeps in LDS
    SceneLaunch sceneLaunch;              // launch shape of the per-scene solve
    uint32_t* sceneWorkOrder = nullptr;   // this step's scene order of the per-scene solve
    float* errPartial = nullptr;     // per finish block: velocity and position error partial sums
    Constraints cs{};
    // scan scratch
    uint32_t* scanBlock = nullptr;
    uint32_t scanBlockCap = 0;
    Counters* ctr = nullptr;
    // launch shape
    int numCU = 0;
    int colourBlocks = 0;            // co-resident blocks of the dataflow colouring kernel
    int solveBlocks = 0;             // co-resident blocks of the dataflow solve kernel
    // learned warm start: per contact point row 2m + p
    uint16_t* learnX = nullptr;      // features (bf16, 32 per row)
    float* learnY = nullptr;         // network outputs (2 per row)
    uint32_t* learnMap = nullptr;    // new-contacts mode: row -> 2 * manifold + point
    // accumulator history: warmHist[c][m] holds the cached accumulators (jn0, jt0, jn1, jt1) manifold
    // m of the step with cur == c started from; the narrow phase carries the matched previous
    // manifold's entry into histOf (NaN for a point without one), i.e. the accumulators of two steps ago
    float4* warmHist[2] = { nullptr, nullptr };
    float4* histOf = nullptr;
    const float4* histIn = nullptr;  // this step: warmHist of the previous step (null: not written)
    float4* histOut = nullptr;       // this step: histOf (null: no learned warm start)
    // collision events (allocated with the first collider callback)
    uint32_t* colCallback = nullptr; // per collider: callback id (0: none)
    float* colThreshold = nullptr;   // per collider: least impact impulse that fires
    uint32_t* evCount = nullptr;     // per manifold: bit 0 side A fires, bit 1 side B fires
    uint32_t* evSides = nullptr;     // per manifold: sides that fire (0..2)
    uint32_t* evOffset = nullptr;    // per manifold: exclusive scan of evSides
    float* evImpulse = nullptr;      // per manifold: impact impulse
    agx_collision_event* events = nullptr;
    uint32_t eventCap = 0;
    // sleeping (allocated when a step first runs with a finite sleep time threshold; null: off)
    uint32_t* sleepId = nullptr;     // per body: sleep island id (>= 2) or 0 when awake
    float* sleepTimer = nullptr;     // per body: time quiet
    float* boundRadius = nullptr;    // per body: reach of its colliders from the centre of mass
    uint32_t* woken = nullptr;       // per sleep id (id - 2 < N): woken this step
    uint32_t* islandFlag = nullptr;  // per island: all bodies ready to sleep
    uint32_t* islandMin = nullptr;   // per island: smallest body
    agx_manifold* stash = nullptr;   // manifolds of sleeping islands (manCap)
    uint32_t* stashIsland = nullptr; // their sleep ids
    uint32_t* stashIslandAlt = nullptr;   // compaction scratch
    // joints (k_joints.hip)
    uint32_t nJoints = 0;
    JointRec* joints = nullptr;
    JointData* jointData = nullptr;
    float* motorAcc = nullptr;       // per joint: a motor's accumulated impulse (warm start)
    uint32_t* jointOrder = nullptr;  // joint indices type by type, colour by colour
    uint2* jointColours = nullptr;   // colour ranges in jointOrder
    JointSchedule jointSchedule{};
    uint2* jointEdges = nullptr;     // per joint: its two bodies when both are dynamic, else (UINT32_MAX, ...)
};

// ---------------------------------------------------------------- device math

namespace dm {
constexpr double FULL_PI = 3.14159265358979323846;
constexpr float HALF_PI = (float)(FULL_PI * 0.5);
constexpr float INV_HALF_PI = (float)(1.0 / (FULL_PI * 0.5));
constexpr float TWO_PI = (float)(FULL_PI * 2.0);
constexpr float INV_TWO_PI = (float)(1.0 / (FULL_PI * 2.0));
// The reference's odd polynomial for sin on [0, pi/2] (FastCosSin.cpp:10-14).
constexpr float C0 = (float)9.9999999999953215e-01;
constexpr float C1 = (float)-1.6666666534403832e-01;
constexpr float C2 = (float)8.3332302391869520e-03;
constexpr float C3 = (float)-1.9818200679665358e-04;
constexpr float C4 = (float)2.6182346015248291e-06;

// FastCosSin.cpp:44-75: quadrant by truncation, mirrored angle, polynomial, clamp, sqrt.
__host__ __device__ inline void fastCosSin(float x, float& c, float& s) {
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
    sm = fminf(sm, 1.0f);
    const float cm = sqrtf(1.0f - sm * sm);
    s = (q & 2) ? -sm : sm;
    c = ((q + 1) & 2) ? -cm : cm;
}

// LeftOutFunctions.cpp:30-43: wrap to [0, 2pi) with a fused multiply-subtract.
__host__ __device__ inline float wrapRotation(float r) {
    const float wrapCount = truncf(r * INV_TWO_PI);
    r = fmaf(-wrapCount, TWO_PI, r);
    if (r < 0.0f) r += TWO_PI;
    return r;
}
}  // namespace dm

// ---------------------------------------------------------------- coherent access
// Data shared between waves of one running kernel on different XCDs (each XCD has its own L2)
// goes through agent-scope relaxed atomics, which the hardware keeps coherent without L2
// write-backs or invalidations. A writer publishes with orderStores() (all of its earlier
// memory operations complete) before the flag store; a reader issues its data loads only
// after observing the flag.
__device__ inline uint32_t loadFlag(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void storeFlag(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long loadCoherent64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void storeCoherent64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float2 loadCoherent2(const float2* p) {
    const unsigned long long a = loadCoherent64(reinterpret_cast<const unsigned long long*>(p));
    return make_float2(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)));
}
__device__ inline void storeCoherent2(float2* p, float2 v) {
    storeCoherent64(reinterpret_cast<unsigned long long*>(p),
                    ((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x));
}
__device__ inline float4 loadCoherent(const float4* p) {
    const float2* q = reinterpret_cast<const float2*>(p);
    const float2 a = loadCoherent2(q), b = loadCoherent2(q + 1);
    return make_float4(a.x, a.y, b.x, b.y);
}
__device__ inline void storeCoherent(float4* p, float4 v) {
    float2* q = reinterpret_cast<float2*>(p);
    storeCoherent2(q, make_float2(v.x, v.y));
    storeCoherent2(q + 1, make_float2(v.z, v.w));
}
__device__ inline void orderStores() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ---------------------------------------------------------------- grid barrier
// Cooperative kernels only (all blocks co-resident). Every wave waits for its own memory
// operations, the block synchronises, thread 0 arrives on the counter and polls it. Only data
// written with coherent stores is visible to other blocks afterwards. Gives up (returns false)
// after the deadline (wall clock, 100 MHz), so a broken invariant cannot hang the device.
__device__ inline bool gridBarrier(uint32_t* counter, uint32_t nblocks, uint32_t& generation, uint64_t deadline) {
    orderStores();
    __syncthreads();
    __shared__ uint32_t ok;
    if (threadIdx.x == 0) {
        ok = 1u;
        generation++;
        const uint32_t target = generation * nblocks;
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (loadFlag(counter) < target) {
            if (wall_clock64() > deadline) { ok = 0u; break; }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    return ok != 0u;
}

// ---------------------------------------------------------------- host-side launchers
// Each returns hipSuccess or the first launch error. All run on 'st'.
hipError_t launchFillU32(hipStream_t st, uint32_t* p, uint32_t v, size_t n);
hipError_t launchFillBytes(hipStream_t st, void* p, size_t bytes);   // zero fill
// Exclusive scan of in[0..n) into out (in place allowed); total into *totalPtr (may be null).
// n = *countPtr when countPtr is not null (device-side count, bounded by countHost), else countHost.
hipError_t launchScan(hipStream_t st, const uint32_t* in, uint32_t* out, const uint32_t* countPtr,
                      uint32_t* totalPtr, uint64_t countHost, uint32_t* blockScratch, uint32_t blockCap);

hipError_t launchRefreshCosSin(hipStream_t st, DeviceWorld& w);
// part 0: the whole velocity-Verlet step; 1: the CCD snapshot and the first half kick only; 2: the
// drift, the second half kick, wrap and cos/sin (joints are solved between parts 1 and 2).
hipError_t launchIntegrate(hipStream_t st, DeviceWorld& w, float dt, float gx, float gy, int snapshot, int part = 0);
hipError_t launchPose(hipStream_t st, DeviceWorld& w, int all, int prevFromSnapshot, int writePrev);
hipError_t launchCcd(hipStream_t st, DeviceWorld& w, float fastFactor, float coreFraction);

hipError_t launchBuildTree(hipStream_t st, DeviceWorld& w, Tree& t);
hipError_t launchPairs(hipStream_t st, DeviceWorld& w);
// S <= BROAD_PREP_THREADS scenes; 40 S bytes of LDS.
hipError_t launchBroadPrep(hipStream_t st, const float4* aabb, const uint8_t* isDynamic, const uint32_t* scene,
                           uint32_t n, uint32_t S, const BroadPrep& out);
hipError_t launchSplitPairs(hipStream_t st, const uint2* pairs, uint32_t n, uint32_t* a, uint32_t* b);

hipError_t launchNarrow(hipStream_t st, DeviceWorld& w);
hipError_t launchHashBuild(hipStream_t st, DeviceWorld& w);

// Colouring, solve order and constraint prep; launchWarmStart then applies the warm-start
// impulses (and clears the solve kernel's flags).
// coloursGiven: colourId already holds every manifold's colour (the first-fit colouring is skipped)
hipError_t launchColourPrep(hipStream_t st, DeviceWorld& w, bool coloursGiven = false);
hipError_t launchWarmStart(hipStream_t st, DeviceWorld& w);
// launchSolve: the velocity and position sweeps (one kernel); launchSolveFinish: accumulators
// back to the manifolds, residual and error totals.
// Per-scene solve: its dynamic LDS bytes (0 when the largest scene does not fit) and launch shape,
// for scenes whose body counts, largest first, are orderedBodies; wholeCap: how many scenes whole
// blocks take (UINT32_MAX: one and a half per resident block).
uint32_t sceneSolveSetup(const uint32_t* orderedBodies, uint32_t S, int numCU, uint32_t wholeCap, SceneLaunch& out,
                         uint32_t& colourLdsBytes);
hipError_t launchSceneSolve(hipStream_t st, DeviceWorld& w, uint32_t vIt, uint32_t pIt, int trackResidual,
                            float residualTarget);
// Colouring (one workgroup per scene, in LDS) and constraint data without predecessor slots.
hipError_t launchColourPrepScene(hipStream_t st, DeviceWorld& w);
// Islands from the slot ranges of the last per-scene solve (one workgroup per scene).
hipError_t launchSceneIslands(hipStream_t st, DeviceWorld& w);
hipError_t launchSolve(hipStream_t st, DeviceWorld& w, uint32_t vIt, uint32_t pIt, int trackResidual,
                       float residualTarget);
hipError_t launchSolveFinish(hipStream_t st, DeviceWorld& w, int trackResidual, float residualTarget, int trackErrors,
                             int perScene);
hipError_t launchIslands(hipStream_t st, DeviceWorld& w);
// Sleeping (k_sleep.hip): the wake after the narrow phase, the sleep after the islands.
hipError_t launchWake(hipStream_t st, DeviceWorld& w);
hipError_t launchSleep(hipStream_t st, DeviceWorld& w, float dt, float speedThreshold, float timeThreshold);
// Joints between the two velocity-Verlet halves (k_joints.hip).
hipError_t launchJoints(hipStream_t st, DeviceWorld& w, float dt);
// Collision events of the step's manifolds on its pre-warm-start velocities (k_events.hip).
hipError_t launchEvents(hipStream_t st, DeviceWorld& w);
constexpr uint32_t SOLVE_FINISH_BLOCKS = 256;
// Co-resident 256-thread blocks per CU of the dataflow colouring and solve kernels: the default
// (at most 4) and what the occupancy calculator allows.
void dataflowBlocksPerCU(int& colour, int& solve);
void dataflowOccupancy(int& colour, int& solve);
// Message of the last failure outside a world (agx_last_error).
void setLastError(const char* msg);
// The one stream every world and network of the current device runs on. The dataflow kernels
// (colouring, sweeps, radix scatter) wait on other blocks of their own launch, which needs those
// blocks to get CUs; two such kernels of different worlds side by side could each hold what the
// other waits for. One stream per device keeps all launches in issue order.
hipError_t deviceStream(hipStream_t* out);

}  // namespace agx
