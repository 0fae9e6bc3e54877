// C ABI (include/agx.h): scene building, device world lifetime, step orchestration, read-outs.
#include "agx_internal.h"
#include "builder.h"
#include "mlp_internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>
#include <unordered_map>

struct agx_builder {
    agx::Builder b;
};

namespace {
enum Segment : uint32_t {
    SEG_INTEGRATE, SEG_POSE, SEG_CCD, SEG_TREE, SEG_PAIRS, SEG_NARROW, SEG_COLOUR, SEG_SOLVE, SEG_SOLVE_FINISH,
    SEG_ISLANDS, SEG_FINISH, SEG_POST_UPDATE, SEG_LEARN, SEG_COUNT
};
const char* const kSegmentNames[SEG_COUNT] = { "integrate", "pose", "ccd", "tree_build", "pair_traversal", "narrow",
                                               "colour_prep", "solve", "solve_finish", "islands", "finish",
                                               "post_update", "learned" };
struct EventPair { hipEvent_t a, b; uint32_t seg; };
}  // namespace

struct agx_world {
    agx::DeviceWorld d;
    agx_settings settings{};
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<void*> allocations;
    std::vector<float> mass, inertia, localComX, localComY;   // per body, constant
    std::vector<uint32_t> colType, colMaterial, colSlot, colScene, polyStart, polyCount, colBodyHost;
    std::vector<float> boundRadius;  // per body: the reach of its colliders (ObjectManager.cpp:30-43)
    std::vector<agx::JointRec> jointRecsHost;   // joints as built, their solve order and colour ranges
    std::vector<uint32_t> jointOrderHost;
    std::vector<uint2> jointColoursHost;
    std::vector<float> colP0, colP1, colMinExt, colMaxExt;
    uint32_t staticBodies = 0, staticColliders = 0;
    float accumulator = 0.0f;
    uint32_t lastStepCount = 1;
    bool postUpdatePending = true;   // next step's CCD sweep starts from the post-update pose
    bool inStep = false;             // phases of a step have run but FINISH has not
    bool solvedPerScene = false;     // this step's SOLVE phase ran the per-scene solve
    int nextPhase = 0;
    bool profiling = false;
    std::vector<EventPair> pending;  // recorded, not yet read
    std::vector<hipEvent_t> eventPool;
    double segMs[SEG_COUNT] = {};
    uint32_t segLaunches[SEG_COUNT] = {};
    agx_mlp* mlp = nullptr;          // learned warm start (not owned)
    uint32_t learnMode = 0;          // AGX_LEARN_* flags
    bool histReady = false;          // the previous step recorded its manifolds' starting accumulators
    bool learnedNarrow = false;      // this step's narrow phase ran with a network set
    int colourPerCuMax = 1, solvePerCuMax = 1;   // launch-shape limits (occupancy, allocations)
    std::vector<uint32_t> sceneSizes;   // body counts of the scenes in size order (per-scene launch shape)
    // deferred body modifications (agx_world_queue_body), applied in order at the next flush
    struct BodyCommand { uint32_t kind, body; float a, b; };
    std::vector<BodyCommand> commands;
    // collision events (agx_world_set_collider_callback / agx_world_read_events)
    std::vector<uint32_t> colCallbackHost;
    uint32_t callbackColliders = 0;  // colliders with a callback id
    bool recordEvents = false;       // this call's steps record events
    bool staticDirty = false;        // a static body moved: static poses, AABBs and tree are rebuilt
    bool poseEdited = false;         // a body moved or turned since the last pose pass
    // the scenes as built, kept for creation after the upload (agx_world_add_bodies rebuilds from it)
    agx_builder* built = nullptr;
    // step graphs (agx_world_step_graph): one per ping-pong parity of the manifold buffers
    hipStream_t captureStream = nullptr;
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<unsigned char> key;   // what the captured launches depend on
        bool histReadyAfter = false;
    } graphs[2];
};

namespace {

thread_local std::string g_lastError;

int fail(agx_world* w, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (w) w->err = buf;
    g_lastError = buf;
    return code;
}

}  // namespace

namespace agx {
void setLastError(const char* msg) { g_lastError = msg ? msg : ""; }

hipError_t deviceStream(hipStream_t* out) {
    static std::mutex mu;
    static hipStream_t streams[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!streams[dev]) {
        e = hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
        if (e != hipSuccess) return e;
    }
    *out = streams[dev];
    return hipSuccess;
}
}  // namespace agx

namespace {

#define HIP_CHECK(w, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return fail((w), AGX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));  \
    } while (0)

template <class T>
hipError_t devAlloc(agx_world* w, T** p, size_t count) {
    *p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
    if (e == hipSuccess) w->allocations.push_back(*p);
    return e;
}

template <class T>
hipError_t upload(agx_world* w, T** p, const std::vector<T>& host) {
    hipError_t e = devAlloc(w, p, host.size());
    if (e != hipSuccess || host.empty()) return e;
    return hipMemcpy(*p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
}

uint32_t ceilLog2(uint32_t v) {
    uint32_t b = 0;
    while ((1u << b) < v) b++;
    return b;
}

// Allocates a tree over 'list' (ascending collider indices) grouped by scene.
int allocTree(agx_world* w, agx::Tree& t, const std::vector<uint32_t>& list, const std::vector<uint32_t>& colScene,
              uint32_t S) {
    t.n = (uint32_t)list.size();
    t.sortTile = agx::sortTileFor(t.n);
    t.sortBlocks = std::max(1u, (t.n + t.sortTile - 1) / t.sortTile);
    std::vector<uint32_t> sceneStart(S + 1, 0);
    for (uint32_t c : list) sceneStart[colScene[c] + 1]++;
    for (uint32_t s = 0; s < S; s++) sceneStart[s + 1] += sceneStart[s];
    std::vector<uint32_t> chunkScene;
    std::vector<uint2> chunkRange;
    for (uint32_t s = 0; s < S; s++)
        for (uint32_t lo = sceneStart[s]; lo < sceneStart[s + 1]; lo += agx::BOUNDS_CHUNK) {
            chunkScene.push_back(s);
            chunkRange.push_back(make_uint2(lo, std::min(sceneStart[s + 1], lo + agx::BOUNDS_CHUNK)));
        }
    t.chunks = (uint32_t)chunkScene.size();
    const size_t n = t.n, nodes = n ? 2 * n - 1 : 1, internal = n > 1 ? n - 1 : 1;
    HIP_CHECK(w, upload(w, &t.list, list));
    HIP_CHECK(w, upload(w, &t.sceneStart, sceneStart));
    HIP_CHECK(w, upload(w, &t.chunkScene, chunkScene));
    HIP_CHECK(w, upload(w, &t.chunkRange, chunkRange));
    HIP_CHECK(w, devAlloc(w, &t.keys, n));
    HIP_CHECK(w, devAlloc(w, &t.vals, n));
    HIP_CHECK(w, devAlloc(w, &t.keysAlt, n));
    HIP_CHECK(w, devAlloc(w, &t.valsAlt, n));
    HIP_CHECK(w, devAlloc(w, &t.hist, 4ull * (t.sortBlocks + 1) * 256 + 4));   // radix sort scratch
    HIP_CHECK(w, devAlloc(w, &t.nodeAabb, nodes));
    HIP_CHECK(w, devAlloc(w, &t.child, internal));
    HIP_CHECK(w, devAlloc(w, &t.range, internal));
    HIP_CHECK(w, devAlloc(w, &t.parent, nodes));
    HIP_CHECK(w, devAlloc(w, &t.rec, internal));
    HIP_CHECK(w, devAlloc(w, &t.leafBody, n));
    HIP_CHECK(w, devAlloc(w, &t.sceneBounds, S));
    if (n) HIP_CHECK(w, hipMemcpy(t.vals, list.data(), n * 4, hipMemcpyHostToDevice));   // initial CCD order
    return AGX_OK;
}

int allocPairBuffers(agx_world* w, uint64_t pairCap) {
    agx::DeviceWorld& d = w->d;
    // the narrow phase packs a pair index into 29 bits
    if (pairCap >= (1ull << 28)) return fail(w, AGX_ERR_CAPACITY, "pair capacity above 2^28 pairs");
    d.pairCap = pairCap;
    d.queryWaves = (d.nDyn + 63) / 64;
    HIP_CHECK(w, devAlloc(w, &d.pairs, pairCap));
    HIP_CHECK(w, devAlloc(w, &d.waveCount, 2ull * std::max(1u, d.queryWaves)));
    HIP_CHECK(w, devAlloc(w, &d.waveOffset, 2ull * std::max(1u, d.queryWaves)));
    HIP_CHECK(w, devAlloc(w, &d.stagedPairs, 2ull * std::max(1u, d.queryWaves) * agx::STAGE_PAIRS));
    d.scanBlockCap = 4096;
    HIP_CHECK(w, devAlloc(w, &d.scanBlock, d.scanBlockCap));
    HIP_CHECK(w, devAlloc(w, &d.ctr, 1));
    HIP_CHECK(w, hipMemset(d.ctr, 0, sizeof(agx::Counters)));
    return AGX_OK;
}

hipEvent_t takeEvent(agx_world* w) {
    if (!w->eventPool.empty()) { hipEvent_t e = w->eventPool.back(); w->eventPool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one segment's launches with events when profiling is on.
struct SegmentTimer {
    agx_world* w;
    EventPair p{ nullptr, nullptr, 0 };
    SegmentTimer(agx_world* world, uint32_t seg) : w(world) {
        if (!w->profiling) return;
        p.seg = seg;
        p.a = takeEvent(w);
        p.b = takeEvent(w);
        if (p.a) (void)hipEventRecord(p.a, w->stream);
    }
    ~SegmentTimer() {
        if (!w->profiling || !p.a || !p.b) return;
        (void)hipEventRecord(p.b, w->stream);
        w->pending.push_back(p);
    }
};

int beginStep(agx_world* w) {
    agx::DeviceWorld& d = w->d;
    const size_t off = offsetof(agx::Counters, pairs);
    HIP_CHECK(w, agx::launchFillBytes(w->stream, reinterpret_cast<char*>(d.ctr) + off, sizeof(agx::Counters) - off));
    return AGX_OK;
}

float wrapAngle(float angle) {   // ObjectManager.cpp:614-619
    const float twoPi = (float)(3.14159265358979323846 * 2.0);
    float wrapped = std::fmod(angle, twoPi);
    if (wrapped < 0.0f) wrapped += twoPi;
    return wrapped;
}

// ObjectManager::flushCommands / applyCommand (SimulationImpl/ObjectManager.cpp:638-865) for the
// body commands, in queue order, on host copies of the body arrays. Collider poses are not
// refreshed here: as in the reference, the next step's CCD snapshot still holds the pose before
// the edit, and the step's pose pass computes the new one (static bodies: a full rebuild).
// The wake paths of a flush (ObjectManager::flushCommands, ObjectManager.cpp:638-700, and
// Simulation::wakeIslandsOfEditedStatics, Simulation.cpp:463-486), on host copies: every commanded
// body wakes its island; stashed manifolds of an edited static body's colliders are dropped and
// their islands woken; awake partners of those colliders in the warm-start source restart their
// timers. Woken islands' stashed manifolds are appended, in stash order, to the warm-start source
// (wakeSleepingContactData), whose hash is then rebuilt.
int wakeForCommands(agx_world* w, const std::vector<uint32_t>& commanded, const std::vector<uint32_t>& editedStatic) {
    agx::DeviceWorld& d = w->d;
    if (!d.sleepId) return AGX_OK;
    const uint32_t n = d.N;
    std::vector<uint32_t> ids(n);
    std::vector<float> timer(n);
    HIP_CHECK(w, hipMemcpy(ids.data(), d.sleepId, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_CHECK(w, hipMemcpy(timer.data(), d.sleepTimer, n * sizeof(float), hipMemcpyDeviceToHost));
    agx::Counters c;
    HIP_CHECK(w, hipMemcpy(&c, d.ctr, sizeof c, hipMemcpyDeviceToHost));
    std::vector<agx_manifold> stash(c.stashCount), prev(c.prevManifolds);
    std::vector<uint32_t> stashIsland(c.stashCount);
    if (c.stashCount) {
        HIP_CHECK(w, hipMemcpy(stash.data(), d.stash, stash.size() * sizeof(agx_manifold), hipMemcpyDeviceToHost));
        HIP_CHECK(w, hipMemcpy(stashIsland.data(), d.stashIsland, stash.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    if (c.prevManifolds)
        HIP_CHECK(w, hipMemcpy(prev.data(), d.man[1 - d.cur], prev.size() * sizeof(agx_manifold), hipMemcpyDeviceToHost));
    std::vector<uint8_t> woken(n, 0);
    bool any = false;
    for (uint32_t b : commanded)
        if (ids[b]) { woken[ids[b] - 2] = 1; any = true; }
    std::vector<uint8_t> edited(d.C, 0);
    for (uint32_t b : editedStatic)
        for (uint32_t col = 0; col < d.C; col++)
            if (w->colBodyHost[col] == b) edited[col] = 1;
    std::vector<uint32_t> partners;
    if (!editedStatic.empty()) {
        size_t k = 0;
        for (size_t i = 0; i < stash.size(); i++) {
            if (edited[stash[i].a] || edited[stash[i].b]) { woken[stashIsland[i] - 2] = 1; any = true; continue; }
            stash[k] = stash[i];
            stashIsland[k] = stashIsland[i];
            k++;
        }
        stash.resize(k);
        stashIsland.resize(k);
        for (const agx_manifold& m : prev)
            if (edited[m.a] != edited[m.b]) partners.push_back(w->colBodyHost[edited[m.a] ? m.b : m.a]);
    }
    if (!any && partners.empty() && editedStatic.empty()) return AGX_OK;
    const size_t prevBefore = prev.size();
    if (any) {
        for (uint32_t b = 0; b < n; b++)
            if (ids[b] && woken[ids[b] - 2]) { ids[b] = 0; timer[b] = 0.0f; }
        size_t k = 0;
        for (size_t i = 0; i < stash.size(); i++) {
            if (woken[stashIsland[i] - 2]) { prev.push_back(stash[i]); continue; }
            stash[k] = stash[i];
            stashIsland[k] = stashIsland[i];
            k++;
        }
        stash.resize(k);
        stashIsland.resize(k);
    }
    for (uint32_t b : partners)
        if (w->mass[b] > 0.0f && ids[b] == 0) timer[b] = 0.0f;
    if (prev.size() > d.manCap) return fail(w, AGX_ERR_CAPACITY, "restored manifolds exceed the manifold capacity");
    HIP_CHECK(w, hipMemcpy(d.sleepId, ids.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_CHECK(w, hipMemcpy(d.sleepTimer, timer.data(), n * sizeof(float), hipMemcpyHostToDevice));
    if (!stash.empty()) {
        HIP_CHECK(w, hipMemcpy(d.stash, stash.data(), stash.size() * sizeof(agx_manifold), hipMemcpyHostToDevice));
        HIP_CHECK(w, hipMemcpy(d.stashIsland, stashIsland.data(), stash.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    const uint32_t stashCount = (uint32_t)stash.size();
    HIP_CHECK(w, hipMemcpy(&d.ctr->stashCount, &stashCount, sizeof stashCount, hipMemcpyHostToDevice));
    if (prev.size() != prevBefore) {   // the warm-start source grew: rebuild its hash as a step's FINISH does
        HIP_CHECK(w, hipMemcpy(d.man[1 - d.cur], prev.data(), prev.size() * sizeof(agx_manifold), hipMemcpyHostToDevice));
        const uint32_t m = (uint32_t)prev.size();
        HIP_CHECK(w, hipMemcpy(&d.ctr->manifolds, &m, sizeof m, hipMemcpyHostToDevice));
        d.cur ^= 1;
        HIP_CHECK(w, agx::launchHashBuild(w->stream, d));
        d.cur ^= 1;
        HIP_CHECK(w, hipStreamSynchronize(w->stream));
        w->histReady = false;
    }
    return AGX_OK;
}

int flushCommands(agx_world* w) {
    if (w->commands.empty()) return AGX_OK;
    std::vector<uint32_t> commanded, editedStatic;
    agx::DeviceWorld& d = w->d;
    const uint32_t n = d.N;
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    std::vector<float4> pos(n), vel(n);
    std::vector<float2> rcs(n), inv(n);
    HIP_CHECK(w, hipMemcpy(pos.data(), d.pos, n * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_CHECK(w, hipMemcpy(vel.data(), d.vel, n * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_CHECK(w, hipMemcpy(rcs.data(), d.rcs, n * sizeof(float2), hipMemcpyDeviceToHost));
    HIP_CHECK(w, hipMemcpy(inv.data(), d.inv, n * sizeof(float2), hipMemcpyDeviceToHost));
    for (const agx_world::BodyCommand& c : w->commands) {
        const uint32_t b = c.body;
        const bool isStatic = inv[b].x == 0.0f;
        commanded.push_back(b);   // every command wakes its body's island (ObjectManager.cpp:684-693)
        if (isStatic && (c.kind == AGX_SET_POSITION || c.kind == AGX_SET_ROTATION)) editedStatic.push_back(b);
        switch (c.kind) {
        case AGX_SET_POSITION: {
            const float r = pos[b].z;
            const float cosR = std::cos(r), sinR = std::sin(r);
            const float lx = w->localComX[b], ly = w->localComY[b];
            pos[b].x = c.a + cosR * lx - sinR * ly;
            pos[b].y = c.b + sinR * lx + cosR * ly;
            if (isStatic) w->staticDirty = true;
            w->poseEdited = true;
            break;
        }
        case AGX_SET_ROTATION: {
            const float r = wrapAngle(c.a);
            pos[b].z = r;
            rcs[b] = make_float2(std::cos(r), std::sin(r));
            if (isStatic) w->staticDirty = true;
            w->poseEdited = true;
            break;
        }
        case AGX_SET_VELOCITY:
            vel[b].x = c.a;
            vel[b].y = c.b;
            break;
        case AGX_SET_ANGULAR_VELOCITY:
            vel[b].z = c.a;
            break;
        case AGX_SET_MASS: {   // applyBodyMass (:867-905) within the dynamic group
            const float newMass = std::fmax(0.0f, c.a), oldMass = w->mass[b];
            float inertia = w->inertia[b];
            if (oldMass > 0.0f) inertia *= newMass / oldMass;
            w->mass[b] = newMass;
            w->inertia[b] = inertia;
            inv[b] = make_float2(1.0f / newMass, inertia > 0.0f ? 1.0f / inertia : 0.0f);
            break;
        }
        case AGX_SET_INERTIA: {
            if (isStatic) break;
            const float inertia = std::fmax(0.0f, c.a);
            w->inertia[b] = inertia;
            inv[b].y = inertia > 0.0f ? 1.0f / inertia : 0.0f;
            break;
        }
        }
    }
    w->commands.clear();
    HIP_CHECK(w, hipMemcpy(d.pos, pos.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIP_CHECK(w, hipMemcpy(d.vel, vel.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIP_CHECK(w, hipMemcpy(d.rcs, rcs.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    HIP_CHECK(w, hipMemcpy(d.inv, inv.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    return wakeForCommands(w, commanded, editedStatic);
}

int runPhase(agx_world* w, uint32_t phase, float dt) {
    agx::DeviceWorld& d = w->d;
    const agx_settings& s = w->settings;
    hipStream_t st = w->stream;
    const bool ccd = s.ccd_enabled != 0;
    switch (phase) {
    case AGX_PHASE_INTEGRATE: {
        if (int r = beginStep(w)) return r;
        w->inStep = true;
        w->solvedPerScene = false;
        SegmentTimer t(w, SEG_INTEGRATE);
        if (d.nJoints == 0) {
            HIP_CHECK(w, agx::launchIntegrate(st, d, dt, s.gravity_x, s.gravity_y, ccd ? 1 : 0));
        } else {   // Simulation::integrateAndApplyConstraints (Simulation.cpp:384-405)
            for (uint32_t k = 0; k < AGX_JOINT_TYPES; k++) d.jointSchedule.iterations[k] = s.joint_iterations[k];
            HIP_CHECK(w, agx::launchIntegrate(st, d, dt, s.gravity_x, s.gravity_y, ccd ? 1 : 0, 1));
            HIP_CHECK(w, agx::launchJoints(st, d, dt));
            HIP_CHECK(w, agx::launchIntegrate(st, d, dt, s.gravity_x, s.gravity_y, ccd ? 1 : 0, 2));
        }
        break;
    }
    case AGX_PHASE_POSE: {
        SegmentTimer t(w, SEG_POSE);
        // the CCD sweep starts from the colliders' last computed poses (snapshotPoses,
        // ContinuousCollisionDetector.cpp:276-299); after a post-update pose pass they follow from the
        // step's body snapshot, unless a flushed command moved a body after that pass
        HIP_CHECK(w, agx::launchPose(st, d, 0, (w->postUpdatePending && !w->poseEdited) ? 1 : 0, ccd ? 1 : 0));
        w->poseEdited = false;
        if (w->staticDirty) {   // rebuildSpatialData after an edit of a static body (Simulation.cpp:295)
            HIP_CHECK(w, agx::launchPose(st, d, 1, 0, 0));
            HIP_CHECK(w, agx::launchBuildTree(st, d, d.stat));
            w->staticDirty = false;
        }
        break;
    }
    case AGX_PHASE_CCD:
        if (ccd && d.nStat > 0) {
            SegmentTimer t(w, SEG_CCD);
            HIP_CHECK(w, agx::launchCcd(st, d, s.ccd_fast_factor, s.ccd_core_fraction));
        }
        break;
    case AGX_PHASE_BROAD: {
        {
            SegmentTimer t(w, SEG_TREE);
            HIP_CHECK(w, agx::launchBuildTree(st, d, d.dyn));
        }
        SegmentTimer t(w, SEG_PAIRS);
        HIP_CHECK(w, agx::launchPairs(st, d));
        break;
    }
    case AGX_PHASE_NARROW: {
        const bool learned = w->mlp && !s.collide_only;
        // accumulator history (features 8, 17, 18) only for the network at every point: a point
        // without a cached impulse has practically never had one two steps ago
        const bool hist = learned && !(w->learnMode & AGX_LEARN_NEW_CONTACTS);
        d.histOut = hist ? d.histOf : nullptr;
        d.histIn = (hist && w->histReady) ? d.warmHist[1 - d.cur] : nullptr;
        w->learnedNarrow = learned;
        SegmentTimer t(w, SEG_NARROW);
        HIP_CHECK(w, agx::launchNarrow(st, d));
        // ObjectManager::wakeIslandsOfContacts + wakeSleepingContactData (Simulation.cpp:428-434)
        if (d.sleepId && !s.collide_only) HIP_CHECK(w, agx::launchWake(st, d));
        break;
    }
    case AGX_PHASE_SOLVE:
        if (!s.collide_only) {
            const bool predict = w->mlp && (w->learnMode & AGX_LEARN_PREDICT);
            const bool collect = w->mlp && (w->learnMode & AGX_LEARN_COLLECT);
            const bool gated = predict && !(w->learnMode & AGX_LEARN_UNGATED);
            // the gate judges each step by its residual, so a gated world always tracks it
            const int trackResidual = (s.track_residual || gated) ? 1 : 0;
            const bool perScene = d.sceneLdsBytes > 0 && s.solve_per_scene != 0;
            w->solvedPerScene = perScene;
            {
                SegmentTimer t(w, SEG_COLOUR);
                if (perScene) HIP_CHECK(w, agx::launchColourPrepScene(st, d));
                else HIP_CHECK(w, agx::launchColourPrep(st, d));
            }
            if (w->recordEvents) {   // on the pre-warm-start velocities, as computeConstraintData
                SegmentTimer t(w, SEG_COLOUR);
                HIP_CHECK(w, agx::launchEvents(st, d));
            }
            if (predict || collect) {
                SegmentTimer t(w, SEG_LEARN);
                if (!w->learnedNarrow) return fail(w, AGX_ERR_STATE, "learned mode was enabled after this step's narrow phase");
                const uint32_t newOnly = (w->learnMode & AGX_LEARN_NEW_CONTACTS) ? 1u : 0u;
                HIP_CHECK(w, agx::launchLearnBeforeWarm(st, d, agx::mlpDevice(w->mlp), predict ? 1u : 0u, gated ? 1u : 0u,
                                                        collect ? 1u : 0u, newOnly, s.gravity_x, s.gravity_y, dt));
            }
            w->histReady = (predict || collect) && !(w->learnMode & AGX_LEARN_NEW_CONTACTS);
            if (!perScene) {   // the per-scene solve applies the warm start itself
                SegmentTimer t(w, SEG_COLOUR);
                HIP_CHECK(w, agx::launchWarmStart(st, d));
            }
            {
                SegmentTimer t(w, SEG_SOLVE);
                if (perScene)
                    HIP_CHECK(w, agx::launchSceneSolve(st, d, s.velocity_iterations, s.position_iterations, trackResidual,
                                                       s.residual_target));
                else
                    HIP_CHECK(w, agx::launchSolve(st, d, s.velocity_iterations, s.position_iterations, trackResidual,
                                                  s.residual_target));
            }
            if (collect) {
                SegmentTimer t(w, SEG_LEARN);
                HIP_CHECK(w, agx::launchLearnCollect(st, d, *agx::mlpSamples(w->mlp),
                                                     (w->learnMode & AGX_LEARN_NEW_CONTACTS) ? 1u : 0u, s.gravity_x,
                                                     s.gravity_y, dt));
            }
            SegmentTimer t(w, SEG_SOLVE_FINISH);
            HIP_CHECK(w, agx::launchSolveFinish(st, d, trackResidual, s.residual_target, s.track_errors ? 1 : 0,
                                                perScene ? 1 : 0));
            if (gated) HIP_CHECK(w, agx::launchGateUpdate(st, d, s.residual_target > 0.0f ? 1u : 0u));
        }
        break;
    case AGX_PHASE_ISLANDS: {
        SegmentTimer t(w, SEG_ISLANDS);
        if (w->solvedPerScene && !s.collide_only && d.nJoints == 0) HIP_CHECK(w, agx::launchSceneIslands(st, d));
        else HIP_CHECK(w, agx::launchIslands(st, d));
        // updateSleeping + sleepContactData (Simulation.cpp:348-350)
        if (d.sleepId && !s.collide_only)
            HIP_CHECK(w, agx::launchSleep(st, d, dt, s.sleep_speed_threshold, s.sleep_time_threshold));
        break;
    }
    case AGX_PHASE_FINISH: {
        SegmentTimer t(w, SEG_FINISH);
        HIP_CHECK(w, agx::launchHashBuild(st, d));
    }
        d.cur ^= 1;
        w->postUpdatePending = false;
        w->inStep = false;
        break;
    default:
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "unknown phase %u", phase);
    }
    w->nextPhase = (int)((phase + 1) % AGX_PHASE_COUNT);
    return AGX_OK;
}

// Sleeping state, from the first call with a finite sleep time threshold on (every body awake, the
// stash empty); it stays once allocated, so islands asleep when sleeping is switched off still wake.
int allocSleep(agx_world* w) {
    agx::DeviceWorld& d = w->d;
    const uint64_t n = d.N, nd = std::max(1u, d.nDynBodies);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    HIP_CHECK(w, devAlloc(w, &d.sleepTimer, n));
    HIP_CHECK(w, upload(w, &d.boundRadius, w->boundRadius));
    HIP_CHECK(w, devAlloc(w, &d.woken, n));
    HIP_CHECK(w, devAlloc(w, &d.islandFlag, nd));
    HIP_CHECK(w, devAlloc(w, &d.islandMin, nd));
    HIP_CHECK(w, devAlloc(w, &d.stash, d.manCap));
    HIP_CHECK(w, devAlloc(w, &d.stashIsland, d.manCap));
    HIP_CHECK(w, devAlloc(w, &d.stashIslandAlt, d.manCap));
    uint32_t* ids = nullptr;
    HIP_CHECK(w, devAlloc(w, &ids, n));
    HIP_CHECK(w, agx::launchFillU32(w->stream, ids, 0u, n));
    HIP_CHECK(w, agx::launchFillU32(w->stream, reinterpret_cast<uint32_t*>(d.sleepTimer), 0u, n));
    HIP_CHECK(w, agx::launchFillBytes(w->stream, &d.ctr->stashCount, sizeof(uint32_t)));
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    d.sleepId = ids;
    return AGX_OK;
}

// Start of an update/step call (or of a step driven phase by phase): whether its steps record
// collision events, with the event list and step index cleared on the device (a fill kernel, so
// nothing here is a memset) and the event buffer sized to the settings.
int beginCall(agx_world* w) {
    agx::DeviceWorld& d = w->d;
    if (!d.sleepId && std::isfinite(w->settings.sleep_time_threshold))
        if (int r = allocSleep(w)) return r;
    w->recordEvents = w->settings.collision_events != 0 && w->callbackColliders > 0;
    if (!w->recordEvents) return AGX_OK;
    const uint32_t cap = std::max(1u, w->settings.event_capacity);
    if (!d.events || d.eventCap != cap) {
        if (d.events) {
            HIP_CHECK(w, hipStreamSynchronize(w->stream));
            w->allocations.erase(std::remove(w->allocations.begin(), w->allocations.end(), (void*)d.events),
                                 w->allocations.end());
            HIP_CHECK(w, hipFree(d.events));
            d.events = nullptr;
        }
        HIP_CHECK(w, devAlloc(w, &d.events, cap));
        d.eventCap = cap;
    }
    HIP_CHECK(w, agx::launchFillBytes(w->stream, &d.ctr->eventCount, 2 * sizeof(uint32_t)));
    return AGX_OK;
}

int postUpdate(agx_world* w) {
    SegmentTimer t(w, SEG_POST_UPDATE);
    HIP_CHECK(w, agx::launchPose(w->stream, w->d, 0, 0, 0));
    w->postUpdatePending = true;
    return AGX_OK;
}

// Everything a captured step's launches depend on besides device memory: the step size, the
// settings, the learned mode and network, whether a history was recorded, and the parity.
std::vector<unsigned char> graphKey(const agx_world* w, float dt) {
    std::vector<unsigned char> k;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        k.insert(k.end(), c, c + n);
    };
    put(&dt, sizeof dt);
    put(&w->settings, sizeof w->settings);
    put(&w->mlp, sizeof w->mlp);
    put(&w->learnMode, sizeof w->learnMode);
    const unsigned char flags[4] = { (unsigned char)w->histReady, (unsigned char)w->postUpdatePending,
                                     (unsigned char)w->d.cur, (unsigned char)w->recordEvents };
    put(flags, sizeof flags);
    const void* bufs[5] = { w->d.learnX, w->d.warmHist[0], w->d.histOf, w->d.events, w->d.sleepId };
    put(bufs, sizeof bufs);
    const int shape[2] = { w->d.colourBlocks, w->d.solveBlocks };
    put(shape, sizeof shape);
    put(&w->d.sceneLaunch, sizeof w->d.sceneLaunch);
    return k;
}

void dropGraph(agx_world::StepGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.key.clear();
}

// Records one full step (all phases) into a graph on the world's private capture stream, so that
// launches other worlds queue on the shared stream meanwhile are not captured. The host-side
// state advances as if the step had run.
int captureStep(agx_world* w, float dt, agx_world::StepGraph& g) {
    dropGraph(g);
    std::vector<unsigned char> key = graphKey(w, dt);
    if (!w->captureStream)
        HIP_CHECK(w, hipStreamCreateWithFlags(&w->captureStream, hipStreamNonBlocking));
    hipStream_t shared = w->stream;
    w->stream = w->captureStream;
    hipError_t e = hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal);
    int rc = AGX_OK;
    if (e == hipSuccess)
        for (uint32_t p = 0; p < AGX_PHASE_COUNT && rc == AGX_OK; p++) rc = runPhase(w, p, dt);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(w->stream, &graph) : e;
    w->stream = shared;
    if (rc != AGX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess || e2 != hipSuccess) return fail(w, AGX_ERR_HIP, "step capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) { g.exec = nullptr; return fail(w, AGX_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
    g.key = std::move(key);
    g.histReadyAfter = w->histReady;
    return AGX_OK;
}

}  // namespace

extern "C" {

void agx_settings_default(agx_settings* s) {
    if (!s) return;
    std::memset(s, 0, sizeof *s);
    s->gravity_x = 0.0f;
    s->gravity_y = -9.81f;
    s->velocity_iterations = 6;
    s->position_iterations = 3;
    s->ccd_enabled = 1;
    s->ccd_fast_factor = 1.0f;
    s->ccd_core_fraction = 0.5f;
    s->update_interval = 1.0f / 300.0f;
    s->max_delta_per_update = 1.0f / 20.0f;
    s->time_scale = 1.0f;
    s->solve_per_scene = 1;
    s->collision_events = 1;
    s->event_capacity = 65536;
    s->sleep_speed_threshold = 0.1f;
    s->sleep_time_threshold = INFINITY;
    for (uint32_t t = 0; t < AGX_JOINT_TYPES; t++) s->joint_iterations[t] = 6;
}

const char* agx_last_error(void) { return g_lastError.c_str(); }

int agx_device_count(int* out) {
    if (!out) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "out is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *out = e == hipSuccess ? n : 0;
    return AGX_OK;
}

int agx_builder_create(uint32_t num_scenes, agx_builder** out) {
    if (!out || num_scenes == 0) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "need out pointer and >= 1 scene");
    agx_builder* b = new (std::nothrow) agx_builder();
    if (!b) return fail(nullptr, AGX_ERR_CAPACITY, "out of host memory");
    b->b.scenes.resize(num_scenes);
    *out = b;
    return AGX_OK;
}

int agx_builder_destroy(agx_builder* b) {
    delete b;
    return AGX_OK;
}

int agx_builder_add_material(agx_builder* b, float e, float sf, float df, uint32_t* out_index) {
    if (!b) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "builder is null");
    // SimulationImpl/MaterialManager.h:31: elasticity clamped to [0, 1], frictions to >= 0.
    b->b.materials.push_back({ std::fmin(std::fmax(e, 0.0f), 1.0f), std::fmax(sf, 0.0f), std::fmax(df, 0.0f) });
    if (out_index) *out_index = (uint32_t)b->b.materials.size() - 1;
    return AGX_OK;
}

int agx_builder_add_bodies(agx_builder* b, uint32_t scene, uint32_t n, const uint8_t* type, const float* px,
                           const float* py, const float* rot, const float* vx, const float* vy, const float* av,
                           const float* mass, const uint32_t* material, const float* p0, const float* p1,
                           const uint32_t* vstart, const uint32_t* vcount, const float* verts, uint32_t* out_index) {
    if (!b || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    if (n && (!type || !px || !py || !rot || !vx || !vy || !av || !mass || !material || !p0 || !p1))
        return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "null body array");
    agx::SceneData& sd = b->b.scenes[scene];
    sd.materialCount = (uint32_t)b->b.materials.size();
    int rejected = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t r = UINT32_MAX;
        if (type[i] == AGX_SHAPE_CIRCLE) r = sd.createCircle(px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i], p0[i]);
        else if (type[i] == AGX_SHAPE_BOX) r = sd.createBox(px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i], p0[i], p1[i]);
        else if (type[i] == AGX_SHAPE_CAPSULE)
            r = sd.createCapsule(px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i], p0[i], p1[i]);
        else if (type[i] == AGX_SHAPE_POLYGON && vstart && vcount && verts)
            r = sd.createPolygon(px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i],
                                 reinterpret_cast<const agx::Vec*>(verts) + vstart[i], vcount[i]);
        if (r == UINT32_MAX) rejected++;
        if (out_index) out_index[i] = r;
    }
    if (rejected) fail(nullptr, 0, "%d bodies rejected (invalid or non-convex polygon outline)", rejected);
    return rejected;
}

int agx_builder_remove_body(agx_builder* b, uint32_t scene, uint32_t body) {
    if (!b || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    if (!b->b.scenes[scene].destroyBody(body)) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "no body %u in scene %u", body, scene);
    return AGX_OK;
}

int agx_builder_create_body(agx_builder* b, uint32_t scene, float px, float py, float rot, float vx, float vy, float av,
                            float mass, float inertia, float comX, float comY, uint32_t* out_index) {
    if (!b || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    const uint32_t r = b->b.scenes[scene].createBody(px, py, rot, vx, vy, av, mass, inertia, comX, comY);
    if (out_index) *out_index = r;
    return AGX_OK;
}

int agx_builder_add_collider(agx_builder* b, uint32_t scene, uint32_t body, uint8_t type, float offX, float offY,
                             float localRot, uint32_t material, float p0, float p1, uint32_t* out_index) {
    if (!b || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    agx::SceneData& sd = b->b.scenes[scene];
    sd.materialCount = (uint32_t)b->b.materials.size();
    if (body >= sd.bodyCount()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "no body %u in scene %u", body, scene);
    if (type != AGX_SHAPE_CIRCLE && type != AGX_SHAPE_BOX && type != AGX_SHAPE_CAPSULE)
        return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "collider shape %u: circle, box or capsule", (unsigned)type);
    const uint32_t c = sd.createCollider(body, type, offX, offY, localRot, material, p0, p1);
    if (out_index) *out_index = c;
    return AGX_OK;
}

int agx_builder_add_joints(agx_builder* b, uint32_t scene, uint32_t n, const agx_joint_desc* joints, uint32_t* out_index) {
    if (!b || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    if (n && !joints) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "null joint array");
    agx::SceneData& sd = b->b.scenes[scene];
    const size_t before = sd.joints.size();
    for (uint32_t i = 0; i < n; i++) {
        const agx_joint_desc& j = joints[i];
        agx::JointSpec spec{};
        spec.type = j.type;
        spec.enableLimit = j.enable_limit ? 1u : 0u;
        spec.anchorA = { j.anchor_a_x, j.anchor_a_y };
        spec.anchorB = { j.anchor_b_x, j.anchor_b_y };
        spec.axis = { j.axis_x, j.axis_y };
        spec.restLength = j.rest_length;
        spec.stiffness = j.stiffness;
        spec.damping = j.damping;
        spec.lowerLimit = j.lower_limit;
        spec.upperLimit = j.upper_limit;
        spec.targetSpeed = j.target_speed;
        spec.maxTorque = j.max_torque;
        if (!sd.addJoint(spec, j.body_a, j.body_b)) {
            sd.joints.resize(before);
            return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "joint %u: unknown type %u or missing body", i, j.type);
        }
        if (out_index) out_index[i] = (uint32_t)sd.joints.size() - 1;
    }
    return AGX_OK;
}

int agx_builder_scene_counts(agx_builder* b, uint32_t scene, agx_counts* out) {
    if (!b || !out || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    const agx::SceneData& sd = b->b.scenes[scene];
    out->scenes = (uint32_t)b->b.scenes.size(); out->bodies = sd.bodyCount(); out->static_bodies = sd.staticBodies;
    out->colliders = sd.colliderCount(); out->static_colliders = sd.staticColliders;
    out->vertices = (uint32_t)sd.verts.size(); out->materials = (uint32_t)b->b.materials.size(); out->reserved = 0;
    return AGX_OK;
}

int agx_builder_read_body_f32(agx_builder* b, uint32_t scene, uint32_t field, float* out, uint32_t n) {
    if (!b || !out || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    const agx::SceneData& sd = b->b.scenes[scene];
    if (n != sd.bodyCount()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "expected %u bodies", sd.bodyCount());
    const std::vector<float>* v = nullptr;
    switch (field) {
    case AGX_BODY_COM_X: v = &sd.comX; break;       case AGX_BODY_COM_Y: v = &sd.comY; break;
    case AGX_BODY_ROTATION: v = &sd.rot; break;     case AGX_BODY_VEL_X: v = &sd.velX; break;
    case AGX_BODY_VEL_Y: v = &sd.velY; break;       case AGX_BODY_ANG_VEL: v = &sd.angVel; break;
    case AGX_BODY_COS: v = &sd.cosv; break;         case AGX_BODY_SIN: v = &sd.sinv; break;
    case AGX_BODY_INV_MASS: v = &sd.invMass; break; case AGX_BODY_INV_INERTIA: v = &sd.invInertia; break;
    case AGX_BODY_MASS: v = &sd.mass; break;        case AGX_BODY_INERTIA: v = &sd.inertia; break;
    case AGX_BODY_LOCAL_COM_X: v = &sd.localComX; break; case AGX_BODY_LOCAL_COM_Y: v = &sd.localComY; break;
    default: return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "unknown body field %u", field);
    }
    std::memcpy(out, v->data(), n * 4u);
    return AGX_OK;
}

int agx_builder_read_collider(agx_builder* b, uint32_t scene, uint32_t field, uint32_t as_u32, void* out, uint32_t n) {
    if (!b || !out || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    const agx::SceneData& sd = b->b.scenes[scene];
    if (n != sd.colliderCount()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "expected %u colliders", sd.colliderCount());
    if (as_u32) {
        uint32_t* o = static_cast<uint32_t*>(out);
        switch (field) {
        case AGX_COL_BODY: std::memcpy(o, sd.colBody.data(), n * 4u); return AGX_OK;
        case AGX_COL_TYPE: for (uint32_t i = 0; i < n; i++) o[i] = sd.colType[i]; return AGX_OK;
        case AGX_COL_MATERIAL: std::memcpy(o, sd.colMaterial.data(), n * 4u); return AGX_OK;
        case AGX_COL_SLOT: std::memcpy(o, sd.colSlot.data(), n * 4u); return AGX_OK;
        case AGX_COL_POLY_START: std::memcpy(o, sd.polyStart.data(), n * 4u); return AGX_OK;
        case AGX_COL_POLY_COUNT: std::memcpy(o, sd.polyCount.data(), n * 4u); return AGX_OK;
        default: return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "unknown collider field %u", field);
        }
    }
    const std::vector<float>* v = nullptr;
    switch (field) {
    case AGX_COL_COM_OFF_X: v = &sd.comOffX; break;  case AGX_COL_COM_OFF_Y: v = &sd.comOffY; break;
    case AGX_COL_LOCAL_ROT: v = &sd.localRot; break; case AGX_COL_P0: v = &sd.p0; break;
    case AGX_COL_P1: v = &sd.p1; break;              case AGX_COL_MIN_EXTENT: v = &sd.minExtent; break;
    case AGX_COL_MAX_EXTENT: v = &sd.maxExtent; break;
    default: return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "unknown collider field %u", field);
    }
    std::memcpy(out, v->data(), n * 4u);
    return AGX_OK;
}

int agx_builder_read_vertices(agx_builder* b, uint32_t scene, float* verts, float* normals, uint32_t n) {
    if (!b || scene >= b->b.scenes.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "bad builder or scene");
    const agx::SceneData& sd = b->b.scenes[scene];
    if (n != sd.verts.size()) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "expected %zu vertices", sd.verts.size());
    if (verts) std::memcpy(verts, sd.verts.data(), n * 8u);
    if (normals) std::memcpy(normals, sd.normals.data(), n * 8u);
    return AGX_OK;
}

int agx_world_create(agx_builder* bld, const agx_settings* settings, agx_world** out) {
    if (!bld || !out) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "builder and out are required");
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) return fail(nullptr, AGX_ERR_NO_DEVICE, "no HIP device");
    agx_world* w = new (std::nothrow) agx_world();
    if (!w) return fail(nullptr, AGX_ERR_CAPACITY, "out of host memory");
    if (settings) w->settings = *settings; else agx_settings_default(&w->settings);
    agx::DeviceWorld& d = w->d;
    const auto& scenes = bld->b.scenes;
    d.S = (uint32_t)scenes.size();
    d.sceneBits = ceilLog2(d.S);
    if (d.sceneBits > 16) { delete w; return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "too many scenes"); }
    d.mortonBits = std::min(30u, (32u - d.sceneBits) & ~1u);

    // Concatenate scenes.
    std::vector<agx::JointRec> jointRecs;
    std::vector<float4> pos, vel, prev;
    std::vector<float2> rcs, inv, comOff, dims, extent, verts, normals;
    std::vector<uint32_t> bodyScene, colBody, colScene, colSlot, dynBodies, bodyToDyn, dynList, statList;
    std::vector<uint32_t> sceneBodyStart, sceneDyn;
    std::vector<float> localRot;
    std::vector<uint4> shape;
    uint32_t slotOff = 0;   // collider creation slots are unique over the world (the warm-start cache key)
    for (uint32_t s = 0; s < d.S; s++) {
        const agx::SceneData& sd = scenes[s];
        const uint32_t bOff = (uint32_t)pos.size(), cOff = (uint32_t)colBody.size(), vOff = (uint32_t)verts.size();
        sceneBodyStart.push_back(bOff);
        sceneDyn.push_back(sd.bodyCount() - sd.staticBodies);
        for (uint32_t i = 0; i < sd.bodyCount(); i++) {
            pos.push_back(make_float4(sd.comX[i], sd.comY[i], sd.rot[i], 0.0f));
            vel.push_back(make_float4(sd.velX[i], sd.velY[i], sd.angVel[i], 0.0f));
            prev.push_back(make_float4(sd.comX[i], sd.comY[i], sd.cosv[i], sd.sinv[i]));
            rcs.push_back(make_float2(sd.cosv[i], sd.sinv[i]));
            inv.push_back(make_float2(sd.invMass[i], sd.invInertia[i]));
            bodyScene.push_back(s);
            w->mass.push_back(sd.mass[i]); w->inertia.push_back(sd.inertia[i]);
            w->localComX.push_back(sd.localComX[i]); w->localComY.push_back(sd.localComY[i]);
            if (sd.invMass[i] != 0.0f) { bodyToDyn.push_back((uint32_t)dynBodies.size()); dynBodies.push_back(bOff + i); }
            else bodyToDyn.push_back(UINT32_MAX);
        }
        for (uint32_t c = 0; c < sd.colliderCount(); c++) {
            const uint32_t gb = bOff + sd.colBody[c];
            colBody.push_back(gb);
            colScene.push_back(s);
            colSlot.push_back(slotOff + sd.colSlot[c]);
            comOff.push_back(make_float2(sd.comOffX[c], sd.comOffY[c]));
            localRot.push_back(sd.localRot[c]);
            const uint32_t ps = sd.colType[c] == AGX_SHAPE_POLYGON ? vOff + sd.polyStart[c] : 0u;
            shape.push_back(make_uint4(sd.colType[c], ps, sd.polyCount[c], sd.colMaterial[c]));
            dims.push_back(make_float2(sd.p0[c], sd.p1[c]));
            extent.push_back(make_float2(sd.minExtent[c], sd.maxExtent[c]));
            (sd.invMass[sd.colBody[c]] != 0.0f ? dynList : statList).push_back(cOff + c);
            w->colType.push_back(sd.colType[c]); w->colMaterial.push_back(sd.colMaterial[c]);
            w->colSlot.push_back(slotOff + sd.colSlot[c]); w->colScene.push_back(s);
            w->polyStart.push_back(ps); w->polyCount.push_back(sd.polyCount[c]);
            w->colP0.push_back(sd.p0[c]); w->colP1.push_back(sd.p1[c]);
            w->colMinExt.push_back(sd.minExtent[c]); w->colMaxExt.push_back(sd.maxExtent[c]);
            w->colBodyHost.push_back(gb);
            if (w->boundRadius.size() < pos.size()) w->boundRadius.resize(pos.size(), 0.0f);
            w->boundRadius[gb] = std::fmax(w->boundRadius[gb], std::sqrt(sd.comOffX[c] * sd.comOffX[c] +
                                                                          sd.comOffY[c] * sd.comOffY[c]) + sd.maxExtent[c]);
        }
        for (size_t v = 0; v < sd.verts.size(); v++) {
            verts.push_back(make_float2(sd.verts[v].x, sd.verts[v].y));
            normals.push_back(make_float2(sd.normals[v].x, sd.normals[v].y));
        }
        w->staticBodies += sd.staticBodies;
        w->staticColliders += sd.staticColliders;
        slotOff += sd.nextColSlot;
        if (!sd.joints.empty()) {   // creation slots to this world's body indices
            std::vector<uint32_t> denseOf(sd.nextBodySlot, UINT32_MAX);
            for (uint32_t i = 0; i < sd.bodyCount(); i++)
                if (sd.bodySlot[i] < denseOf.size()) denseOf[sd.bodySlot[i]] = i;
            for (const agx::JointSpec& j : sd.joints) {
                agx::JointRec r{};
                r.type = j.type;
                r.a = bOff + denseOf[j.slotA];
                r.b = j.type == AGX_JOINT_MOTOR ? r.a : bOff + denseOf[j.slotB];
                r.enableLimit = j.enableLimit;
                r.anchors = make_float4(j.anchorA.x, j.anchorA.y, j.anchorB.x, j.anchorB.y);
                r.axisRest = make_float4(j.axis.x, j.axis.y, j.restLength, j.referenceAngle);
                r.params = make_float4(j.stiffness, j.damping, j.lowerLimit, j.upperLimit);
                r.motor = make_float2(j.targetSpeed, j.maxTorque);
                jointRecs.push_back(r);
            }
        }
    }
    d.N = (uint32_t)pos.size();
    w->boundRadius.resize(d.N, 0.0f);
    sceneBodyStart.push_back(d.N);
    d.C = (uint32_t)colBody.size();
    d.nDyn = (uint32_t)dynList.size();
    d.nStat = (uint32_t)statList.size();
    d.nDynBodies = (uint32_t)dynBodies.size();
    d.nVerts = (uint32_t)verts.size();
    std::vector<float4> mats;
    for (const agx::Material& m : bld->b.materials) mats.push_back(make_float4(m.elasticity, m.staticFriction, m.dynamicFriction, 0.0f));
    d.nMat = (uint32_t)mats.size();

    hipDeviceProp_t prop;
    int dev = 0;
    int rc = AGX_OK;
    auto bail = [&](int code) { agx_world_destroy(w); return code; };
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return bail(fail(nullptr, AGX_ERR_HIP, "hipGetDeviceProperties failed"));
    d.numCU = prop.multiProcessorCount;
    {
        int perCuColour = 1, perCuSolve = 1, occColour = 1, occSolve = 1;
        agx::dataflowBlocksPerCU(perCuColour, perCuSolve);
        agx::dataflowOccupancy(occColour, occSolve);
        d.colourBlocks = d.numCU * perCuColour;
        d.solveBlocks = d.numCU * perCuSolve;
        w->colourPerCuMax = occColour;
        w->solvePerCuMax = perCuSolve;   // the sweep kernel's per-block buffers are sized for this many
    }
    if (agx::deviceStream(&w->stream) != hipSuccess) return bail(fail(nullptr, AGX_ERR_HIP, "hipStreamCreate failed"));
    // per-scene solve: scenes largest first, LDS for the largest scene's bodies
    std::vector<uint32_t> sceneOrder(d.S);
    uint32_t maxSceneBodies = 0;
    for (uint32_t s = 0; s < d.S; s++) {
        sceneOrder[s] = s;
        maxSceneBodies = std::max(maxSceneBodies, sceneBodyStart[s + 1] - sceneBodyStart[s]);
    }
    std::stable_sort(sceneOrder.begin(), sceneOrder.end(), [&](uint32_t a, uint32_t b) {
        return sceneBodyStart[a + 1] - sceneBodyStart[a] > sceneBodyStart[b + 1] - sceneBodyStart[b];
    });
    // AGX_SCENE_WHOLE_CAP sets how many of the largest scenes whole blocks solve (default: one per
    // resident block; the rest go to lane groups); tests use it to reach the groups with few scenes
    uint32_t wholeCap = UINT32_MAX;
    if (const char* cap = std::getenv("AGX_SCENE_WHOLE_CAP")) wholeCap = (uint32_t)std::strtoul(cap, nullptr, 10);
    std::vector<uint32_t> orderedBodies(d.S);
    for (uint32_t k = 0; k < d.S; k++) orderedBodies[k] = sceneBodyStart[sceneOrder[k] + 1] - sceneBodyStart[sceneOrder[k]];
    d.sceneLdsBytes = agx::sceneSolveSetup(orderedBodies.data(), d.S, d.numCU, wholeCap, d.sceneLaunch, d.sceneColourLdsBytes);
    w->sceneSizes = orderedBodies;
    d.sceneMaxBodies = maxSceneBodies;
    // AGX_COLOUR_LDS_CAP lowers how many of a scene's manifolds the per-scene colouring keeps in
    // LDS (larger scenes colour from global scratch instead); tests use it to reach that path
    if (const char* cap = std::getenv("AGX_COLOUR_LDS_CAP")) d.sceneColourCap = (uint32_t)std::strtoul(cap, nullptr, 10);

#define UP(p, v) do { if (upload(w, &(p), (v)) != hipSuccess) return bail(fail(nullptr, AGX_ERR_HIP, "upload " #p)); } while (0)
#define AL(p, n) do { if (devAlloc(w, &(p), (n)) != hipSuccess) return bail(fail(nullptr, AGX_ERR_HIP, "alloc " #p)); } while (0)
    UP(d.pos, pos); UP(d.vel, vel); UP(d.prev, prev); UP(d.rcs, rcs); UP(d.inv, inv);
    UP(d.bodyScene, bodyScene); UP(d.dynBodies, dynBodies); UP(d.bodyToDyn, bodyToDyn);
    UP(d.colBody, colBody); UP(d.colScene, colScene); UP(d.colSlot, colSlot); UP(d.comOff, comOff);
    UP(d.localRot, localRot); UP(d.shape, shape); UP(d.dims, dims); UP(d.extent, extent);
    UP(d.verts, verts); UP(d.normals, normals); UP(d.materials, mats);
    UP(d.sceneBodyStart, sceneBodyStart); UP(d.sceneOrder, sceneOrder); UP(d.sceneDyn, sceneDyn);
    AL(d.sceneRange, (uint64_t)d.S * agx::MAX_COLOURS); AL(d.sceneRes, (uint64_t)d.S * agx::MAX_SWEEPS);
    AL(d.sceneSweeps, d.S); AL(d.sceneWorkOrder, d.S); AL(d.sceneIslands, d.S); AL(d.sceneIslandOffset, d.S); AL(d.sceneManRange, d.S);
    AL(d.islandId, d.N); AL(d.ufParent, d.N); AL(d.ufRoot, std::max(1u, d.nDynBodies)); AL(d.hitTime, d.N); AL(d.ccdFast, std::max(1u, d.nDyn));
    AL(d.pose, d.C); AL(d.wRot, d.C); AL(d.prevW, d.C); AL(d.aabb, d.C);
    if ((rc = allocTree(w, d.dyn, dynList, colScene, d.S))) return bail(rc);
    if ((rc = allocTree(w, d.stat, statList, colScene, d.S))) return bail(rc);
    const uint64_t pairCap = std::max<uint64_t>(1u << 16, 16ull * d.nDyn + 4ull * d.nStat);
    if ((rc = allocPairBuffers(w, pairCap))) return bail(rc);
    d.manCap = pairCap;
    AL(d.slots, pairCap); AL(d.flags, pairCap); AL(d.flagScan, pairCap);
    AL(d.man[0], pairCap); AL(d.man[1], pairCap);
    d.hashCapMax = 16;
    while (d.hashCapMax < 2 * d.manCap) d.hashCapMax <<= 1;
    AL(d.hashKeys, d.hashCapMax); AL(d.hashVals, d.hashCapMax);
    AL(d.manBodies, pairCap); AL(d.degree, d.N); AL(d.incStart, d.N); AL(d.incFill, d.N);
    AL(d.inc, 2 * pairCap); AL(d.incSorted, 2 * pairCap); AL(d.rank, 2 * pairCap);
    AL(d.colRec, 4ull * d.N);
    d.solveGroups = (uint32_t)d.solveBlocks * 256u;   // enough for any group size down to one lane
    AL(d.groupDone, 2ull * d.solveGroups);
    AL(d.colourId, pairCap); AL(d.colourHist, (uint64_t)agx::COLOUR_BLOCKS * agx::MAX_COLOURS);
    const uint64_t solveCap = pairCap + (uint64_t)agx::COLOUR_ALIGN * agx::MAX_COLOURS;
    d.solveCap = solveCap;
    AL(d.posOf, pairCap); AL(d.order, solveCap);
    AL(d.blockRes, (uint64_t)d.solveBlocks * 2 * agx::MAX_SWEEPS); AL(d.errPartial, 2 * agx::SOLVE_FINISH_BLOCKS);
    AL(d.cs.idx, solveCap); AL(d.cs.dep, solveCap); AL(d.cs.nrm, solveCap); AL(d.cs.inv, solveCap); AL(d.cs.r0, solveCap);
    AL(d.cs.r1, solveCap); AL(d.cs.k0, solveCap); AL(d.cs.k1, solveCap); AL(d.cs.imp, solveCap);
    AL(d.cs.blk, solveCap);
    AL(d.colCallback, d.C); AL(d.colThreshold, d.C);
    if (!jointRecs.empty()) {
        // Colours per type: each joint takes the smallest colour that no earlier joint of its type
        // on one of its dynamic bodies has (static bodies are never written, so they never conflict).
        std::vector<uint32_t> order;
        std::vector<uint2> colours, edges(jointRecs.size());
        for (uint32_t t = 0; t < AGX_JOINT_TYPES; t++) {
            d.jointSchedule.typeColour[t] = (uint32_t)colours.size();
            std::vector<std::vector<uint32_t>> byColour;
            std::map<uint64_t, bool> used;
            for (uint32_t j = 0; j < jointRecs.size(); j++) {
                const agx::JointRec& r = jointRecs[j];
                if (r.type != t) continue;
                const bool dynA = inv[r.a].x != 0.0f, dynB = r.type != AGX_JOINT_MOTOR && inv[r.b].x != 0.0f;
                uint32_t c = 0;
                while ((dynA && used.count(((uint64_t)r.a << 32) | c)) || (dynB && used.count(((uint64_t)r.b << 32) | c))) c++;
                if (dynA) used[((uint64_t)r.a << 32) | c] = true;
                if (dynB) used[((uint64_t)r.b << 32) | c] = true;
                if (c >= byColour.size()) byColour.resize(c + 1);
                byColour[c].push_back(j);
            }
            for (const auto& list : byColour) {
                colours.push_back(make_uint2((uint32_t)order.size(), (uint32_t)(order.size() + list.size())));
                order.insert(order.end(), list.begin(), list.end());
            }
        }
        d.jointSchedule.typeColour[AGX_JOINT_TYPES] = (uint32_t)colours.size();
        for (uint32_t j = 0; j < jointRecs.size(); j++) {
            const agx::JointRec& r = jointRecs[j];
            const bool both = r.type != AGX_JOINT_MOTOR && inv[r.a].x != 0.0f && inv[r.b].x != 0.0f;
            edges[j] = both ? make_uint2(r.a, r.b) : make_uint2(UINT32_MAX, UINT32_MAX);
        }
        d.nJoints = (uint32_t)jointRecs.size();
        UP(d.joints, jointRecs); UP(d.jointOrder, order); UP(d.jointColours, colours); UP(d.jointEdges, edges);
        AL(d.jointData, d.nJoints); AL(d.motorAcc, d.nJoints);
        w->jointRecsHost = jointRecs;
        w->jointOrderHost = order;
        w->jointColoursHost = colours;
    }
    AL(d.evCount, pairCap); AL(d.evSides, pairCap); AL(d.evOffset, pairCap); AL(d.evImpulse, pairCap);
    w->colCallbackHost.assign(d.C, 0u);
#undef UP
#undef AL
    // Initial state: island ids, CCD hit times, empty contact cache, refreshSpatialData, trees.
    hipStream_t st = w->stream;
    bool ok = (d.nJoints == 0 || agx::launchFillU32(st, reinterpret_cast<uint32_t*>(d.motorAcc), 0u, d.nJoints) == hipSuccess) &&
              agx::launchFillU32(st, d.islandId, UINT32_MAX, d.N) == hipSuccess &&
              agx::launchFillU32(st, d.colCallback, 0u, d.C) == hipSuccess &&
              agx::launchFillU32(st, reinterpret_cast<uint32_t*>(d.colThreshold), 0x7F7FFFFFu, d.C) == hipSuccess &&
              agx::launchFillU32(st, d.sceneSweeps, 0u, d.S) == hipSuccess &&
              agx::launchFillU32(st, reinterpret_cast<uint32_t*>(d.hitTime), 0x3F800000u, d.N) == hipSuccess &&
              agx::launchFillU32(st, reinterpret_cast<uint32_t*>(d.hashKeys), 0xFFFFFFFFu, 2 * d.hashCapMax) == hipSuccess &&
              agx::launchRefreshCosSin(st, d) == hipSuccess &&
              agx::launchPose(st, d, 1, 0, 0) == hipSuccess &&
              agx::launchBuildTree(st, d, d.stat) == hipSuccess &&
              agx::launchBuildTree(st, d, d.dyn) == hipSuccess &&
              hipStreamSynchronize(st) == hipSuccess;
    if (!ok) return bail(fail(nullptr, AGX_ERR_HIP, "world initialisation failed: %s", hipGetErrorString(hipGetLastError())));
    w->built = new (std::nothrow) agx_builder(*bld);
    if (!w->built) return bail(fail(nullptr, AGX_ERR_CAPACITY, "out of host memory"));
    *out = w;
    return AGX_OK;
}

int agx_world_destroy(agx_world* w) {
    if (!w) return AGX_OK;
    if (w->stream) (void)hipStreamSynchronize(w->stream);   // the device's shared stream stays
    for (agx_world::StepGraph& g : w->graphs) dropGraph(g);
    if (w->captureStream) (void)hipStreamDestroy(w->captureStream);
    for (const EventPair& p : w->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (hipEvent_t e : w->eventPool) (void)hipEventDestroy(e);
    for (void* p : w->allocations) (void)hipFree(p);
    delete w->built;
    delete w;
    return AGX_OK;
}

const char* agx_world_last_error(agx_world* w) { return w ? w->err.c_str() : g_lastError.c_str(); }

int agx_world_set_settings(agx_world* w, const agx_settings* s) {
    if (!w || !s) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (s->velocity_iterations > agx::MAX_SWEEPS) return fail(w, AGX_ERR_INVALID_ARGUMENT, "at most %u velocity iterations", agx::MAX_SWEEPS);
    w->settings = *s;
    return AGX_OK;
}

int agx_world_get_settings(agx_world* w, agx_settings* s) {
    if (!w || !s) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    *s = w->settings;
    return AGX_OK;
}

int agx_world_counts(agx_world* w, agx_counts* out) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    out->scenes = w->d.S; out->bodies = w->d.N; out->static_bodies = w->staticBodies;
    out->colliders = w->d.C; out->static_colliders = w->staticColliders;
    out->vertices = w->d.nVerts; out->materials = w->d.nMat; out->reserved = 0;
    return AGX_OK;
}

int agx_world_run_phase(agx_world* w, uint32_t phase, float dt) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (phase >= AGX_PHASE_COUNT) return fail(w, AGX_ERR_INVALID_ARGUMENT, "unknown phase %u", phase);
    if ((int)phase != w->nextPhase) return fail(w, AGX_ERR_STATE, "phase %u out of order (expected %d)", phase, w->nextPhase);
    if (phase == AGX_PHASE_INTEGRATE) {
        if (int r = flushCommands(w)) return r;
        if (int r = beginCall(w)) return r;
    }
    return runPhase(w, phase, dt);
}

int agx_world_post_update(agx_world* w) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    return postUpdate(w);
}

int agx_world_step(agx_world* w, float dt, uint32_t n_steps) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    if (!(dt > 0.0f)) return AGX_OK;
    if (int r = flushCommands(w)) return r;
    if (int r = beginCall(w)) return r;
    for (uint32_t i = 0; i < n_steps; i++)
        for (uint32_t p = 0; p < AGX_PHASE_COUNT; p++)
            if (int r = runPhase(w, p, dt)) return r;
    return n_steps ? postUpdate(w) : AGX_OK;
}

// ---- learned warm start
int agx_world_set_learned(agx_world* w, agx_mlp* mlp, uint32_t mode) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (mode & ~(AGX_LEARN_PREDICT | AGX_LEARN_COLLECT | AGX_LEARN_UNGATED | AGX_LEARN_NEW_CONTACTS))
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "unknown learned mode bits 0x%x", mode);
    if ((mode & (AGX_LEARN_PREDICT | AGX_LEARN_COLLECT)) && !mlp)
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "learned mode needs a network");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    agx::DeviceWorld& d = w->d;
    if ((mode & (AGX_LEARN_PREDICT | AGX_LEARN_COLLECT)) && !d.learnX) {
        const size_t man = std::max<uint64_t>(d.manCap, 1), rows = 2 * man;
        if (devAlloc(w, &d.learnX, rows * agx::MLP_INP) != hipSuccess || devAlloc(w, &d.learnY, rows * agx::MLP_OUT) != hipSuccess ||
            devAlloc(w, &d.learnMap, rows) != hipSuccess ||
            devAlloc(w, &d.warmHist[0], man) != hipSuccess || devAlloc(w, &d.warmHist[1], man) != hipSuccess ||
            devAlloc(w, &d.histOf, man) != hipSuccess)
            return fail(w, AGX_ERR_HIP, "cannot allocate the learned warm-start buffers (%zu rows)", rows);
    }
    // the accumulator history stays valid across mode changes: every step solved with a network
    // set records it, and a step without one marks it missing
    w->mlp = (mode & (AGX_LEARN_PREDICT | AGX_LEARN_COLLECT)) ? mlp : nullptr;
    w->learnMode = w->mlp ? mode : 0u;
    return agx_world_reset_gate(w);
}

int agx_world_reset_gate(agx_world* w) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    agx::LearnGate g{};
    g.use = 1;   // the first step tries the network
    HIP_CHECK(w, hipMemcpyAsync(&w->d.ctr->gate, &g, sizeof g, hipMemcpyHostToDevice, w->stream));
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    return AGX_OK;
}

int agx_world_read_gate(agx_world* w, agx_gate* out) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    agx::LearnGate g{};
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    HIP_CHECK(w, hipMemcpy(&g, &w->d.ctr->gate, sizeof g, hipMemcpyDeviceToHost));
    out->cost_cached = g.cost[0]; out->cost_learned = g.cost[1];
    out->use_learned = g.use; out->since_probe = g.sinceProbe;
    out->steps_cached = g.steps[0]; out->steps_learned = g.steps[1];
    out->period = g.period; out->probing = g.probing;
    return AGX_OK;
}

int agx_gate_update(agx_gate* out, uint32_t used_learned, float cost) {
    if (!out) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "gate is null");
    agx::LearnGate g{};
    g.cost[0] = out->cost_cached; g.cost[1] = out->cost_learned;
    g.use = out->use_learned; g.sinceProbe = out->since_probe;
    g.steps[0] = out->steps_cached; g.steps[1] = out->steps_learned;
    g.period = out->period; g.probing = out->probing;
    agx::gateUpdate(g, used_learned, cost);
    out->cost_cached = g.cost[0]; out->cost_learned = g.cost[1];
    out->use_learned = g.use; out->since_probe = g.sinceProbe;
    out->steps_cached = g.steps[0]; out->steps_learned = g.steps[1];
    out->period = g.period; out->probing = g.probing;
    return AGX_OK;
}

}  // extern "C"

// ---- launch-shape tuning
namespace {

std::mutex g_tuneMutex;
std::map<std::string, agx_tuning> g_tuneCache;
bool g_tuneFileLoaded = false;

// The GPU, the world's size class (bodies to the power of two below, scenes) and the solver
// settings a tuning was measured for.
std::string tuneKey(agx_world* w) {
    hipDeviceProp_t prop{};
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipGetDeviceProperties(&prop, dev);
    auto bucket = [](uint64_t v) { uint32_t b = 0; while (v > 1) { v >>= 1; b++; } return b; };
    char buf[512];
    snprintf(buf, sizeof buf, "%s|%s|cu%d|bodies2^%u|scenes%u|iv%u|ip%u|target%d|learned%u|perscene%d", prop.name,
             prop.gcnArchName, prop.multiProcessorCount, bucket(w->d.N), w->d.S, w->settings.velocity_iterations,
             w->settings.position_iterations, w->settings.residual_target > 0.0f ? 1 : 0, w->learnMode,
             w->d.sceneLdsBytes && w->settings.solve_per_scene ? 1 : 0);
    for (char* c = buf; *c; c++) if (*c == ' ' || *c == '\n') *c = '_';
    return buf;
}

// AGX_TUNE_CACHE names a text file of "key solve colour ms" lines shared by processes.
void loadTuneFile() {
    if (g_tuneFileLoaded) return;
    g_tuneFileLoaded = true;
    const char* path = std::getenv("AGX_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[768], key[512];
    while (std::fgets(line, sizeof line, f)) {
        agx_tuning t{};
        if (std::sscanf(line, "%511s %u %u %f %u", key, &t.solve_blocks_per_cu, &t.colour_blocks_per_cu, &t.ms_per_step,
                        &t.scene_whole) >= 4)
            g_tuneCache[key] = t;
    }
    std::fclose(f);
}

void storeTune(const std::string& key, const agx_tuning& t) {
    g_tuneCache[key] = t;
    const char* path = std::getenv("AGX_TUNE_CACHE");
    if (!path || !*path) return;
    if (FILE* f = std::fopen(path, "a")) {
        std::fprintf(f, "%s %u %u %.6f %u\n", key.c_str(), t.solve_blocks_per_cu, t.colour_blocks_per_cu, t.ms_per_step,
                     t.scene_whole);
        std::fclose(f);
    }
}

}  // namespace

extern "C" {

int agx_world_get_tuning(agx_world* w, agx_tuning* out) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof *out);
    out->solve_blocks_per_cu = (uint32_t)(w->d.solveBlocks / w->d.numCU);
    out->colour_blocks_per_cu = (uint32_t)(w->d.colourBlocks / w->d.numCU);
    out->max_solve_blocks_per_cu = (uint32_t)w->solvePerCuMax;
    out->max_colour_blocks_per_cu = (uint32_t)w->colourPerCuMax;
    out->scene_whole = w->d.sceneLdsBytes ? w->d.sceneLaunch.whole : 0u;
    return AGX_OK;
}

int agx_world_set_tuning(agx_world* w, const agx_tuning* in) {
    if (!w || !in) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    if (in->solve_blocks_per_cu < 1 || (int)in->solve_blocks_per_cu > w->solvePerCuMax || in->colour_blocks_per_cu < 1 ||
        (int)in->colour_blocks_per_cu > w->colourPerCuMax)
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "blocks per CU must be 1..%d (sweeps) and 1..%d (colouring)",
                    w->solvePerCuMax, w->colourPerCuMax);
    if (in->scene_whole) {
        if (!w->d.sceneLdsBytes) return fail(w, AGX_ERR_INVALID_ARGUMENT, "the world does not solve scene by scene");
        agx::SceneLaunch launch;
        uint32_t colourBytes = 0;
        const uint32_t bytes = agx::sceneSolveSetup(w->sceneSizes.data(), w->d.S, w->d.numCU,
                                                    std::min(in->scene_whole, w->d.S), launch, colourBytes);
        if (bytes != w->d.sceneLdsBytes || colourBytes != w->d.sceneColourLdsBytes)
            return fail(w, AGX_ERR_HIP, "per-scene launch shape not available");
        w->d.sceneLaunch = launch;
    }
    w->d.solveBlocks = w->d.numCU * (int)in->solve_blocks_per_cu;
    w->d.colourBlocks = w->d.numCU * (int)in->colour_blocks_per_cu;
    return AGX_OK;
}

// Times `steps_per_trial` graph-replayed steps for every launch shape in two interleaved rounds
// (so that a scene that settles meanwhile favours no candidate) and keeps the fastest. A shape
// measured before for the same GPU, size class and solver settings is reused without stepping.
int agx_world_autotune(agx_world* w, float dt, uint32_t steps_per_trial, agx_tuning* out) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    if (!(dt > 0.0f) || steps_per_trial == 0) return fail(w, AGX_ERR_INVALID_ARGUMENT, "need dt > 0 and steps");
    const std::string key = tuneKey(w);
    {
        std::lock_guard<std::mutex> lock(g_tuneMutex);
        loadTuneFile();
        auto it = g_tuneCache.find(key);
        if (it != g_tuneCache.end()) {
            agx_tuning t = it->second;
            if ((int)t.solve_blocks_per_cu <= w->solvePerCuMax && (int)t.colour_blocks_per_cu <= w->colourPerCuMax) {
                if (int r = agx_world_set_tuning(w, &t)) return r;
                if (out) { *out = t; out->from_cache = 1; out->max_solve_blocks_per_cu = (uint32_t)w->solvePerCuMax;
                           out->max_colour_blocks_per_cu = (uint32_t)w->colourPerCuMax; }
                return AGX_OK;
            }
        }
    }
    std::vector<agx_tuning> cand;
    if (w->d.sceneLdsBytes && w->settings.solve_per_scene) {
        // the steps solve scene by scene: the dataflow kernels never run, the whole-block share does
        agx_tuning base{};
        if (int r = agx_world_get_tuning(w, &base)) return r;
        const uint32_t blocks = (uint32_t)std::max(w->d.sceneLaunch.blocks, 1);
        for (uint32_t halves : { 1u, 2u, 3u, 4u }) {
            agx_tuning t = base;
            t.scene_whole = std::max(1u, std::min(w->d.S, blocks * halves / 2));
            bool seen = false;
            for (const agx_tuning& c : cand) seen |= c.scene_whole == t.scene_whole;
            if (!seen) cand.push_back(t);
        }
    } else {
        for (uint32_t sb = 1; sb <= (uint32_t)w->solvePerCuMax; sb *= 2)
            for (uint32_t cb = 1; cb <= (uint32_t)w->colourPerCuMax; cb *= 2) {
                agx_tuning t{};
                t.solve_blocks_per_cu = sb;
                t.colour_blocks_per_cu = cb;
                cand.push_back(t);
            }
    }
    // With a residual target the sweep count varies from step to step, so a trial's time is
    // divided by the sweeps it ran (velocity and position) to compare launch shapes fairly.
    const bool perSweep = w->settings.residual_target > 0.0f && !w->settings.collide_only;
    std::vector<double> ms(cand.size(), 0.0), work(cand.size(), 0.0);
    hipEvent_t a = nullptr, b = nullptr;
    HIP_CHECK(w, hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return fail(w, AGX_ERR_HIP, "event"); }
    int rc = AGX_OK;
    for (int round = 0; round < 2 && rc == AGX_OK; round++)
        for (size_t c = 0; c < cand.size() && rc == AGX_OK; c++) {
            if ((rc = agx_world_set_tuning(w, &cand[c]))) break;
            agx_stats s0{};
            if ((rc = agx_world_read_stats(w, &s0))) break;
            if (hipEventRecord(a, w->stream) != hipSuccess) { rc = fail(w, AGX_ERR_HIP, "event"); break; }
            if ((rc = agx_world_step_graph(w, dt, steps_per_trial))) break;
            float e = 0.0f;
            if (hipEventRecord(b, w->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(&e, a, b) != hipSuccess) { rc = fail(w, AGX_ERR_HIP, "event timing"); break; }
            agx_stats st{};
            if ((rc = agx_world_read_stats(w, &st))) break;
            ms[c] += st.overflow ? 1e30 : (double)e;   // a shape that overflowed a capacity is never chosen
            work[c] += perSweep ? (double)(st.total_velocity_sweeps - s0.total_velocity_sweeps) +
                                      (double)steps_per_trial * w->settings.position_iterations
                                : (double)steps_per_trial;
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != AGX_OK) return rc;
    size_t best = 0;
    auto cost = [&](size_t c) { return ms[c] / std::max(work[c], 1.0); };
    for (size_t c = 1; c < cand.size(); c++) if (cost(c) < cost(best)) best = c;
    agx_tuning t = cand[best];
    t.ms_per_step = (float)(ms[best] / (2.0 * steps_per_trial));
    if (int r = agx_world_set_tuning(w, &t)) return r;
    {
        std::lock_guard<std::mutex> lock(g_tuneMutex);
        storeTune(key, t);
    }
    if (out) { *out = t; out->from_cache = 0; out->max_solve_blocks_per_cu = (uint32_t)w->solvePerCuMax;
               out->max_colour_blocks_per_cu = (uint32_t)w->colourPerCuMax; }
    return AGX_OK;
}

// Steps as agx_world_step does, with every step after the first replayed from a captured graph
// (one per parity of the ping-pong manifold buffers, re-captured when what it depends on changed).
// The first step runs as launches because it starts from the post-update pose. Profiling needs the
// per-segment events of plain launches, so a profiled world steps without graphs.
int agx_world_step_graph(agx_world* w, float dt, uint32_t n_steps) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    if (!(dt > 0.0f) || n_steps == 0) return AGX_OK;
    if (int r = flushCommands(w)) return r;
    if (w->profiling) return agx_world_step(w, dt, n_steps);
    if (int r = beginCall(w)) return r;
    for (uint32_t p = 0; p < AGX_PHASE_COUNT; p++)
        if (int r = runPhase(w, p, dt)) return r;
    for (uint32_t i = 1; i < n_steps; i++) {
        agx_world::StepGraph& g = w->graphs[w->d.cur];
        if (!g.exec || g.key != graphKey(w, dt)) {
            if (int r = captureStep(w, dt, g)) return r;
            // captured, not run: the host state already moved one step ahead; run the graph now
            HIP_CHECK(w, hipGraphLaunch(g.exec, w->stream));
            continue;
        }
        HIP_CHECK(w, hipGraphLaunch(g.exec, w->stream));
        // the host-side effects of one step (runPhase INTEGRATE..FINISH)
        w->d.cur ^= 1;
        w->postUpdatePending = false;
        w->histReady = g.histReadyAfter;
        w->nextPhase = 0;
    }
    return postUpdate(w);
}

// Simulation::update + FixedClock::advance (SimulationImpl/FixedClock.h:35-63).
int agx_world_update(agx_world* w, float dt, uint32_t* out_steps) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (out_steps) *out_steps = 0;
    if (dt <= 0.0f) return AGX_OK;
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    if (int r = flushCommands(w)) return r;   // Simulation::update flushes before anything reads the data
    const agx_settings& s = w->settings;
    w->accumulator += std::fmin(dt, s.max_delta_per_update);
    const uint32_t steps = (uint32_t)std::floor(w->accumulator / s.update_interval);
    w->accumulator -= (float)steps * s.update_interval;
    if (steps == 0) return AGX_OK;
    w->lastStepCount = steps;
    if (out_steps) *out_steps = steps;
    return agx_world_step(w, s.update_interval * s.time_scale, steps);
}

int agx_world_queue_body(agx_world* w, uint32_t kind, uint32_t body, float a, float b) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (body >= w->d.N) return fail(w, AGX_ERR_INVALID_ARGUMENT, "body %u out of range (%u bodies)", body, w->d.N);
    if (kind > AGX_SET_INERTIA) return fail(w, AGX_ERR_INVALID_ARGUMENT, "unknown body command %u", kind);
    if (kind == AGX_SET_MASS) {
        const bool dynamic = w->mass[body] > 0.0f;   // static bodies have mass 0
        if (!(a > 0.0f) || !dynamic)
            return fail(w, AGX_ERR_INVALID_ARGUMENT, "mass changes keep a body dynamic (mass > 0 on a dynamic body)");
    }
    w->commands.push_back({ kind, body, a, b });
    return AGX_OK;
}

int agx_world_narrowphase(agx_world* w, const uint32_t* pairs_a, const uint32_t* pairs_b, uint64_t n, agx_manifold* out,
                          uint64_t cap, uint64_t* count) {
    if (!w || !count || (n && (!pairs_a || !pairs_b))) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    agx::DeviceWorld& d = w->d;
    *count = 0;
    if (n > d.pairCap) return fail(w, AGX_ERR_CAPACITY, "%llu pairs exceed the pair capacity %llu", (unsigned long long)n,
                                   (unsigned long long)d.pairCap);
    std::vector<uint2> p(n);
    for (uint64_t i = 0; i < n; i++) {
        if (pairs_a[i] >= d.C || pairs_b[i] >= d.C || pairs_a[i] == pairs_b[i])
            return fail(w, AGX_ERR_INVALID_ARGUMENT, "pair %llu (%u, %u) is not two colliders of the world",
                        (unsigned long long)i, pairs_a[i], pairs_b[i]);
        p[i] = make_uint2(pairs_a[i], pairs_b[i]);
    }
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    if (n) HIP_CHECK(w, hipMemcpy(d.pairs, p.data(), n * sizeof(uint2), hipMemcpyHostToDevice));
    const uint32_t n32 = (uint32_t)n;
    HIP_CHECK(w, agx::launchFillU32(w->stream, &d.ctr->pairs, n32, 1));
    d.histOut = nullptr;
    d.histIn = nullptr;
    HIP_CHECK(w, agx::launchNarrow(w->stream, d));
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    uint32_t m = 0, overflow = 0;
    HIP_CHECK(w, hipMemcpy(&m, &d.ctr->manifolds, sizeof m, hipMemcpyDeviceToHost));
    HIP_CHECK(w, hipMemcpy(&overflow, &d.ctr->overflow, sizeof overflow, hipMemcpyDeviceToHost));
    *count = m;
    if (overflow & AGX_OVERFLOW_HASH) return fail(w, AGX_ERR_CAPACITY, "warm-start lookup overflowed");
    if (m > cap) return fail(w, AGX_ERR_CAPACITY, "%u manifolds exceed cap %llu", m, (unsigned long long)cap);
    if (m && out) HIP_CHECK(w, hipMemcpy(out, d.man[d.cur], m * sizeof(agx_manifold), hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_solve_contacts(agx_world* w, const agx_manifold* manifolds, uint64_t n, const uint32_t* order,
                             agx_manifold* out) {
    if (!w || (n && (!manifolds || !out))) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    agx::DeviceWorld& d = w->d;
    if (n > d.manCap) return fail(w, AGX_ERR_CAPACITY, "%llu manifolds exceed the capacity %llu", (unsigned long long)n,
                                  (unsigned long long)d.manCap);
    for (uint64_t i = 0; i < n; i++) {
        const agx_manifold& m = manifolds[i];
        if (m.a >= d.C || m.b >= d.C || m.a == m.b || m.count < 1 || m.count > 2)
            return fail(w, AGX_ERR_INVALID_ARGUMENT, "manifold %llu: colliders (%u, %u), %u points", (unsigned long long)i,
                        m.a, m.b, m.count);
    }
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    std::vector<uint16_t> level;
    if (order && n) {   // levels of the sequential order: one more than the last level at either dynamic body
        std::vector<float2> inv(d.N);
        HIP_CHECK(w, hipMemcpy(inv.data(), d.inv, d.N * sizeof(float2), hipMemcpyDeviceToHost));
        std::vector<uint32_t> next(d.N, 0);   // level after the last manifold so far at the body
        std::vector<unsigned char> seen(n, 0);
        level.assign(n, 0);
        for (uint64_t k = 0; k < n; k++) {
            const uint32_t m = order[k];
            if (m >= n || seen[m]) return fail(w, AGX_ERR_INVALID_ARGUMENT, "order is not a permutation of 0..%llu",
                                               (unsigned long long)n - 1);
            seen[m] = 1;
            const uint32_t a = w->colBodyHost[manifolds[m].a], b = w->colBodyHost[manifolds[m].b];
            const bool da = inv[a].x != 0.0f, db = inv[b].x != 0.0f;
            const uint32_t lv = std::max(da ? next[a] : 0u, db ? next[b] : 0u);
            if (lv >= agx::MAX_COLOURS)
                return fail(w, AGX_ERR_CAPACITY, "the order needs more than %u levels", agx::MAX_COLOURS);
            level[m] = (uint16_t)lv;
            if (da) next[a] = lv + 1;
            if (db) next[b] = lv + 1;
        }
    }
    if (int rc = beginStep(w)) return rc;
    const uint32_t m32 = (uint32_t)n;
    if (n) HIP_CHECK(w, hipMemcpyAsync(d.man[d.cur], manifolds, n * sizeof(agx_manifold), hipMemcpyHostToDevice, w->stream));
    HIP_CHECK(w, hipMemcpyAsync(&d.ctr->manifolds, &m32, sizeof m32, hipMemcpyHostToDevice, w->stream));
    if (!level.empty())
        HIP_CHECK(w, hipMemcpyAsync(d.colourId, level.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice, w->stream));
    const agx_settings& s = w->settings;
    const int track = s.track_residual ? 1 : 0;
    HIP_CHECK(w, agx::launchColourPrep(w->stream, d, !level.empty()));
    HIP_CHECK(w, agx::launchWarmStart(w->stream, d));
    HIP_CHECK(w, agx::launchSolve(w->stream, d, s.velocity_iterations, s.position_iterations, track, s.residual_target));
    HIP_CHECK(w, agx::launchSolveFinish(w->stream, d, track, s.residual_target, s.track_errors ? 1 : 0, 0));
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    uint32_t overflow = 0;
    HIP_CHECK(w, hipMemcpy(&overflow, &d.ctr->overflow, sizeof overflow, hipMemcpyDeviceToHost));
    if (overflow & (AGX_OVERFLOW_STALL | AGX_OVERFLOW_ITEMS))
        return fail(w, AGX_ERR_CAPACITY, "the solve reported overflow bits 0x%x", overflow);
    if (n) HIP_CHECK(w, hipMemcpy(out, d.man[d.cur], n * sizeof(agx_manifold), hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_set_collider_callback(agx_world* w, uint32_t collider, uint32_t callback_id, float threshold) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    agx::DeviceWorld& d = w->d;
    if (collider >= d.C) return fail(w, AGX_ERR_INVALID_ARGUMENT, "collider %u out of range (%u colliders)", collider, d.C);
    if (callback_id > AGX_MAX_CALLBACK_ID) return fail(w, AGX_ERR_INVALID_ARGUMENT, "callback id %u too large", callback_id);
    if (callback_id == 0) threshold = FLT_MAX;   // NO_COLLISION_EVENT_THRESHOLD (Types/CollisionEvent.h:15)
    else if (!(threshold >= 0.0f)) return fail(w, AGX_ERR_INVALID_ARGUMENT, "event threshold must be >= 0");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    const uint32_t old = w->colCallbackHost[collider];
    w->callbackColliders += (callback_id != 0) - (old != 0);
    w->colCallbackHost[collider] = callback_id;
    HIP_CHECK(w, hipMemcpyAsync(d.colCallback + collider, &callback_id, sizeof callback_id, hipMemcpyHostToDevice, w->stream));
    HIP_CHECK(w, hipMemcpyAsync(d.colThreshold + collider, &threshold, sizeof threshold, hipMemcpyHostToDevice, w->stream));
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    return AGX_OK;
}

int agx_world_read_joint_order(agx_world* w, uint32_t* order, uint32_t cap, uint32_t* colour_start, uint32_t colour_cap,
                               uint32_t* type_colour, uint32_t* joints, uint32_t* colours) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    const agx::DeviceWorld& d = w->d;
    const uint32_t k = (uint32_t)w->jointColoursHost.size();
    if (joints) *joints = d.nJoints;
    if (colours) *colours = k;
    if (order) {
        if (cap < d.nJoints) return fail(w, AGX_ERR_CAPACITY, "%u joints", d.nJoints);
        std::copy(w->jointOrderHost.begin(), w->jointOrderHost.end(), order);
    }
    if (colour_start) {
        if (colour_cap < k + 1) return fail(w, AGX_ERR_CAPACITY, "%u joint colours", k);
        for (uint32_t c = 0; c < k; c++) colour_start[c] = w->jointColoursHost[c].x;
        colour_start[k] = d.nJoints;
    }
    if (type_colour)
        for (uint32_t t = 0; t <= AGX_JOINT_TYPES; t++) type_colour[t] = d.nJoints ? d.jointSchedule.typeColour[t] : 0u;
    return AGX_OK;
}

int agx_world_read_joints(agx_world* w, agx_joint_desc* out, float* motor_impulse, uint32_t n) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    const agx::DeviceWorld& d = w->d;
    if (n != d.nJoints) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u joints", d.nJoints);
    if (out)
        for (uint32_t j = 0; j < n; j++) {
            const agx::JointRec& r = w->jointRecsHost[j];
            agx_joint_desc& o = out[j];
            o.type = r.type; o.body_a = r.a; o.body_b = r.type == AGX_JOINT_MOTOR ? UINT32_MAX : r.b;
            o.enable_limit = r.enableLimit;
            o.anchor_a_x = r.anchors.x; o.anchor_a_y = r.anchors.y; o.anchor_b_x = r.anchors.z; o.anchor_b_y = r.anchors.w;
            o.axis_x = r.axisRest.x; o.axis_y = r.axisRest.y; o.rest_length = r.axisRest.z;
            o.stiffness = r.params.x; o.damping = r.params.y; o.lower_limit = r.params.z; o.upper_limit = r.params.w;
            o.target_speed = r.motor.x; o.max_torque = r.motor.y;
        }
    if (motor_impulse && n) {
        HIP_CHECK(w, hipStreamSynchronize(w->stream));
        HIP_CHECK(w, hipMemcpy(motor_impulse, d.motorAcc, n * sizeof(float), hipMemcpyDeviceToHost));
    }
    return AGX_OK;
}

int agx_world_read_sleep(agx_world* w, uint32_t* sleep_island, float* timer, uint32_t n) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    agx::DeviceWorld& d = w->d;
    if (n != d.N) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u bodies", d.N);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    if (!d.sleepId) {
        if (sleep_island) std::fill(sleep_island, sleep_island + n, 0u);
        if (timer) std::fill(timer, timer + n, 0.0f);
        return AGX_OK;
    }
    if (sleep_island) HIP_CHECK(w, hipMemcpy(sleep_island, d.sleepId, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (timer) HIP_CHECK(w, hipMemcpy(timer, d.sleepTimer, n * sizeof(float), hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_read_stash(agx_world* w, agx_manifold* out, uint32_t* island, uint64_t cap, uint64_t* count) {
    if (!w || !count) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    agx::DeviceWorld& d = w->d;
    *count = 0;
    if (!d.sleepId) return AGX_OK;
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    uint32_t n = 0;
    HIP_CHECK(w, hipMemcpy(&n, &d.ctr->stashCount, sizeof n, hipMemcpyDeviceToHost));
    *count = n;
    if (!out && !island) return AGX_OK;   // a size query
    if (n > cap) return fail(w, AGX_ERR_CAPACITY, "%u stashed manifolds, room for %llu", n, (unsigned long long)cap);
    if (n && out) HIP_CHECK(w, hipMemcpy(out, d.stash, n * sizeof(agx_manifold), hipMemcpyDeviceToHost));
    if (n && island) HIP_CHECK(w, hipMemcpy(island, d.stashIsland, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_read_events(agx_world* w, agx_collision_event* out, uint64_t cap, uint64_t* count, uint64_t* dropped) {
    if (!w || !count) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    agx::DeviceWorld& d = w->d;
    *count = 0;
    if (dropped) *dropped = 0;
    if (!w->recordEvents || !d.events) return AGX_OK;
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    uint32_t recorded = 0;
    HIP_CHECK(w, hipMemcpy(&recorded, &d.ctr->eventCount, sizeof recorded, hipMemcpyDeviceToHost));
    const uint64_t kept = std::min<uint64_t>(recorded, d.eventCap);
    *count = kept;
    if (dropped) *dropped = recorded - kept;
    if (kept > cap) return fail(w, AGX_ERR_CAPACITY, "%llu events, room for %llu", (unsigned long long)kept,
                                (unsigned long long)cap);
    if (kept && !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null output");
    if (kept) HIP_CHECK(w, hipMemcpy(out, d.events, kept * sizeof(agx_collision_event), hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_flush(agx_world* w) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    return flushCommands(w);
}

int agx_world_sync(agx_world* w) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    return AGX_OK;
}

int agx_world_read_body_f32(agx_world* w, uint32_t field, float* out, uint32_t n) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    agx::DeviceWorld& d = w->d;
    if (n != d.N) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u bodies, got %u", d.N, n);
    if (field >= AGX_BODY_FIELD_COUNT) return fail(w, AGX_ERR_INVALID_ARGUMENT, "unknown body field %u", field);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    if (field == AGX_BODY_MASS || field == AGX_BODY_INERTIA || field == AGX_BODY_LOCAL_COM_X || field == AGX_BODY_LOCAL_COM_Y) {
        const std::vector<float>& v = field == AGX_BODY_MASS ? w->mass : field == AGX_BODY_INERTIA ? w->inertia
                                    : field == AGX_BODY_LOCAL_COM_X ? w->localComX : w->localComY;
        std::memcpy(out, v.data(), n * 4u);
        return AGX_OK;
    }
    if (field <= AGX_BODY_ANG_VEL) {
        std::vector<float4> h(n);
        HIP_CHECK(w, hipMemcpy(h.data(), field <= AGX_BODY_ROTATION ? d.pos : d.vel, n * sizeof(float4), hipMemcpyDeviceToHost));
        const uint32_t lane = field <= AGX_BODY_ROTATION ? field : field - AGX_BODY_VEL_X;
        for (uint32_t i = 0; i < n; i++) out[i] = lane == 0 ? h[i].x : lane == 1 ? h[i].y : h[i].z;
        return AGX_OK;
    }
    std::vector<float2> h(n);
    HIP_CHECK(w, hipMemcpy(h.data(), field <= AGX_BODY_SIN ? d.rcs : d.inv, n * sizeof(float2), hipMemcpyDeviceToHost));
    const bool first = field == AGX_BODY_COS || field == AGX_BODY_INV_MASS;
    for (uint32_t i = 0; i < n; i++) out[i] = first ? h[i].x : h[i].y;
    return AGX_OK;
}

int agx_world_write_body_f32(agx_world* w, uint32_t field, const float* in, uint32_t n) {
    if (!w || !in) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    agx::DeviceWorld& d = w->d;
    if (n != d.N) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u bodies, got %u", d.N, n);
    if (field > AGX_BODY_SIN) return fail(w, AGX_ERR_INVALID_ARGUMENT, "body field %u is not writable", field);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    if (field <= AGX_BODY_ANG_VEL) {
        float4* p = field <= AGX_BODY_ROTATION ? d.pos : d.vel;
        std::vector<float4> h(n);
        HIP_CHECK(w, hipMemcpy(h.data(), p, n * sizeof(float4), hipMemcpyDeviceToHost));
        const uint32_t lane = field <= AGX_BODY_ROTATION ? field : field - AGX_BODY_VEL_X;
        for (uint32_t i = 0; i < n; i++) (lane == 0 ? h[i].x : lane == 1 ? h[i].y : h[i].z) = in[i];
        HIP_CHECK(w, hipMemcpy(p, h.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        return AGX_OK;
    }
    std::vector<float2> h(n);
    HIP_CHECK(w, hipMemcpy(h.data(), d.rcs, n * sizeof(float2), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) (field == AGX_BODY_COS ? h[i].x : h[i].y) = in[i];
    HIP_CHECK(w, hipMemcpy(d.rcs, h.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    return AGX_OK;
}

int agx_world_read_collider_f32(agx_world* w, uint32_t field, float* out, uint32_t n) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    agx::DeviceWorld& d = w->d;
    if (n != d.C) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u colliders, got %u", d.C, n);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    switch (field) {
    case AGX_COL_WORLD_X: case AGX_COL_WORLD_Y: case AGX_COL_COS: case AGX_COL_SIN: {
        std::vector<float4> h(n);
        HIP_CHECK(w, hipMemcpy(h.data(), d.pose, n * sizeof(float4), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++)
            out[i] = field == AGX_COL_WORLD_X ? h[i].x : field == AGX_COL_WORLD_Y ? h[i].y : field == AGX_COL_COS ? h[i].z : h[i].w;
        return AGX_OK;
    }
    case AGX_COL_AABB_MIN_X: case AGX_COL_AABB_MIN_Y: case AGX_COL_AABB_MAX_X: case AGX_COL_AABB_MAX_Y: {
        std::vector<float4> h(n);
        HIP_CHECK(w, hipMemcpy(h.data(), d.aabb, n * sizeof(float4), hipMemcpyDeviceToHost));
        const uint32_t k = field - AGX_COL_AABB_MIN_X;
        for (uint32_t i = 0; i < n; i++) out[i] = k == 0 ? h[i].x : k == 1 ? h[i].y : k == 2 ? h[i].z : h[i].w;
        return AGX_OK;
    }
    case AGX_COL_WORLD_ROT:
        HIP_CHECK(w, hipMemcpy(out, d.wRot, n * 4u, hipMemcpyDeviceToHost));
        return AGX_OK;
    case AGX_COL_PREV_X: case AGX_COL_PREV_Y: case AGX_COL_COM_OFF_X: case AGX_COL_COM_OFF_Y: {
        std::vector<float2> h(n);
        const bool pv = field <= AGX_COL_PREV_Y;
        HIP_CHECK(w, hipMemcpy(h.data(), pv ? d.prevW : d.comOff, n * sizeof(float2), hipMemcpyDeviceToHost));
        const bool x = field == AGX_COL_PREV_X || field == AGX_COL_COM_OFF_X;
        for (uint32_t i = 0; i < n; i++) out[i] = x ? h[i].x : h[i].y;
        return AGX_OK;
    }
    case AGX_COL_LOCAL_ROT:
        HIP_CHECK(w, hipMemcpy(out, d.localRot, n * 4u, hipMemcpyDeviceToHost));
        return AGX_OK;
    case AGX_COL_P0: std::memcpy(out, w->colP0.data(), n * 4u); return AGX_OK;
    case AGX_COL_P1: std::memcpy(out, w->colP1.data(), n * 4u); return AGX_OK;
    case AGX_COL_MIN_EXTENT: std::memcpy(out, w->colMinExt.data(), n * 4u); return AGX_OK;
    case AGX_COL_MAX_EXTENT: std::memcpy(out, w->colMaxExt.data(), n * 4u); return AGX_OK;
    default:
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "unknown collider field %u", field);
    }
}

int agx_world_read_collider_u32(agx_world* w, uint32_t field, uint32_t* out, uint32_t n) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (n != w->d.C) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u colliders, got %u", w->d.C, n);
    switch (field) {
    case AGX_COL_BODY: HIP_CHECK(w, hipMemcpy(out, w->d.colBody, n * 4u, hipMemcpyDeviceToHost)); return AGX_OK;
    case AGX_COL_TYPE: std::memcpy(out, w->colType.data(), n * 4u); return AGX_OK;
    case AGX_COL_MATERIAL: std::memcpy(out, w->colMaterial.data(), n * 4u); return AGX_OK;
    case AGX_COL_SLOT: std::memcpy(out, w->colSlot.data(), n * 4u); return AGX_OK;
    case AGX_COL_POLY_START: std::memcpy(out, w->polyStart.data(), n * 4u); return AGX_OK;
    case AGX_COL_POLY_COUNT: std::memcpy(out, w->polyCount.data(), n * 4u); return AGX_OK;
    case AGX_COL_SCENE: std::memcpy(out, w->colScene.data(), n * 4u); return AGX_OK;
    default: return fail(w, AGX_ERR_INVALID_ARGUMENT, "unknown collider field %u", field);
    }
}

int agx_world_read_pairs(agx_world* w, uint32_t* a, uint32_t* b, uint64_t cap, uint64_t* count) {
    if (!w || !count) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    agx::Counters c;
    HIP_CHECK(w, hipMemcpy(&c, w->d.ctr, sizeof c, hipMemcpyDeviceToHost));
    const uint64_t n = std::min<uint64_t>(c.pairs, w->d.pairCap);
    *count = c.pairs;
    if (n > cap) return fail(w, AGX_ERR_CAPACITY, "%llu pairs exceed cap %llu", (unsigned long long)n, (unsigned long long)cap);
    if (n && (!a || !b)) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null pair arrays");
    std::vector<uint2> h(n);
    if (n) HIP_CHECK(w, hipMemcpy(h.data(), w->d.pairs, n * sizeof(uint2), hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; i++) { a[i] = h[i].x; b[i] = h[i].y; }
    if (c.overflow & 1u) return fail(w, AGX_ERR_CAPACITY, "pair buffer overflow");
    return AGX_OK;
}

int agx_world_read_broadphase_order(agx_world* w, uint32_t which, uint32_t* keys, uint32_t* colliders, uint32_t n,
                                    uint32_t* morton_bits) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (which > 1) return fail(w, AGX_ERR_INVALID_ARGUMENT, "which must be 0 (dynamic) or 1 (static)");
    const agx::Tree& t = which ? w->d.stat : w->d.dyn;
    if (n != t.n) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u entries", t.n);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    if (n && keys) HIP_CHECK(w, hipMemcpy(keys, t.keys, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (n && colliders) HIP_CHECK(w, hipMemcpy(colliders, t.vals, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (morton_bits) *morton_bits = w->d.mortonBits;
    return AGX_OK;
}

int agx_world_read_manifolds(agx_world* w, uint32_t which, agx_manifold* out, uint64_t cap, uint64_t* count) {
    if (!w || !count) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    agx::Counters c;
    HIP_CHECK(w, hipMemcpy(&c, w->d.ctr, sizeof c, hipMemcpyDeviceToHost));
    const agx_manifold* src;
    uint64_t n;
    if (w->inStep && which == 0) { src = w->d.man[w->d.cur]; n = c.manifolds; }
    else { src = w->d.man[1 - w->d.cur]; n = c.prevManifolds; }
    *count = n;
    if (n > cap) return fail(w, AGX_ERR_CAPACITY, "%llu manifolds exceed cap", (unsigned long long)n);
    if (n && !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null output");
    if (n) HIP_CHECK(w, hipMemcpy(out, src, n * sizeof(agx_manifold), hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_write_manifolds(agx_world* w, uint32_t which, const agx_manifold* in, uint64_t count) {
    if (!w || (count && !in)) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (which != 1 || w->inStep) return fail(w, AGX_ERR_STATE, "only the warm-start source can be written, between steps");
    if (count > w->d.manCap) return fail(w, AGX_ERR_CAPACITY, "too many manifolds");
    agx::DeviceWorld& d = w->d;
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    if (count) HIP_CHECK(w, hipMemcpy(d.man[1 - d.cur], in, count * sizeof(agx_manifold), hipMemcpyHostToDevice));
    const uint32_t m = (uint32_t)count;
    HIP_CHECK(w, hipMemcpy(&d.ctr->manifolds, &m, 4, hipMemcpyHostToDevice));
    d.cur ^= 1;
    HIP_CHECK(w, agx::launchHashBuild(w->stream, d));
    d.cur ^= 1;
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    w->histReady = false;   // the new source has no recorded starting accumulators
    return AGX_OK;
}

int agx_world_read_solve_order(agx_world* w, uint32_t* order, uint64_t cap, uint32_t* colour_start, uint32_t colour_cap,
                               uint32_t* colours) {
    if (!w || !colours) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    agx::Counters* c = new agx::Counters;
    if (hipMemcpy(c, w->d.ctr, sizeof *c, hipMemcpyDeviceToHost) != hipSuccess) { delete c; return fail(w, AGX_ERR_HIP, "read counters"); }
    const uint32_t K = c->colours;
    *colours = K;
    if (K + 1 > colour_cap) { delete c; return fail(w, AGX_ERR_CAPACITY, "%u colours exceed cap", K); }
    // Device positions are wave-aligned per colour; the exported order drops the padding holes.
    std::vector<uint32_t> start(K + 1), count(K);
    uint32_t total = 0;
    for (uint32_t k = 0; k < K; k++) { start[k] = total; count[k] = c->colourCount[k]; total += count[k]; }
    start[K] = total;
    const uint32_t padded = K ? c->colourStart[K] : 0u;
    std::vector<uint32_t> deviceStart(c->colourStart, c->colourStart + K + 1);
    delete c;
    if (colour_start) std::memcpy(colour_start, start.data(), (K + 1) * 4u);
    if (total > cap) return fail(w, AGX_ERR_CAPACITY, "order needs %u entries", total);
    if (total && order) {
        std::vector<uint32_t> dev(padded);
        HIP_CHECK(w, hipMemcpy(dev.data(), w->d.order, padded * 4u, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < K; k++) std::memcpy(order + start[k], dev.data() + deviceStart[k], count[k] * 4u);
    }
    return AGX_OK;
}

int agx_world_read_islands(agx_world* w, uint32_t* island_id, uint32_t n, uint32_t* island_count) {
    if (!w || !island_id || !island_count) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (n != w->d.N) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u bodies", w->d.N);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    HIP_CHECK(w, hipMemcpy(island_id, w->d.islandId, n * 4u, hipMemcpyDeviceToHost));
    HIP_CHECK(w, hipMemcpy(island_count, &w->d.ctr->islands, 4, hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_read_stats(agx_world* w, agx_stats* out) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    agx::Counters* c = new agx::Counters;
    if (hipMemcpy(c, w->d.ctr, sizeof *c, hipMemcpyDeviceToHost) != hipSuccess) { delete c; return fail(w, AGX_ERR_HIP, "read counters"); }
    std::memset(out, 0, sizeof *out);
    out->pairs = c->pairs;
    out->manifolds = w->inStep ? c->manifolds : c->prevManifolds;
    out->contacts = c->contacts;
    out->colours = c->colours;
    out->islands = c->islands;
    out->ccd_hits = c->ccdHits;
    out->velocity_sweeps = c->sweeps;
    out->overflow = c->overflow;
    out->learned_warm_start = c->learnUsed;
    out->solved_steps = c->solveSteps;
    out->total_velocity_sweeps = c->sweepsTotal;
    out->total_body_sweeps = c->bodySweepsTotal;
    out->mean_velocity_sweeps = (float)((double)c->bodySweeps / (double)std::max(w->d.nDynBodies, 1u));
    out->solved_per_scene = c->perScene;
    out->converged_scenes = c->convergedScenes;
    out->converged_bodies = c->convergedBodies;
    out->total_converged_bodies = c->convergedBodiesTotal;
    out->velocity_error = c->velocityError;
    out->position_error = c->positionError;
    for (uint32_t k = 0; k < agx::MAX_SWEEPS; k++)
        out->residual[k] = (float)(std::sqrt(c->residualNum[k]) / std::fmax(std::sqrt(c->residualDen[k]), 1e-12));
    delete c;
    return AGX_OK;
}

int agx_world_read_scene_residual(agx_world* w, float* out, uint32_t n) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (n != w->d.S) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u scenes, got %u", w->d.S, n);
    if (!w->settings.track_residual) return fail(w, AGX_ERR_STATE, "residuals are tracked only with track_residual");
    if (!w->d.sceneRes || !w->d.sceneSweeps) return fail(w, AGX_ERR_STATE, "the world does not solve scene by scene");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    std::vector<uint32_t> sweeps(n);
    std::vector<float2> res((size_t)n * agx::MAX_SWEEPS);
    HIP_CHECK(w, hipMemcpy(sweeps.data(), w->d.sceneSweeps, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_CHECK(w, hipMemcpy(res.data(), w->d.sceneRes, res.size() * sizeof(float2), hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < n; s++) {
        const uint32_t q = std::min(sweeps[s], agx::MAX_SWEEPS);
        if (q == 0) { out[s] = 0.0f; continue; }
        const float2 r = res[(size_t)s * agx::MAX_SWEEPS + q - 1];
        out[s] = (float)(std::sqrt((double)r.x) / std::fmax(std::sqrt((double)r.y), 1e-12));
    }
    return AGX_OK;
}

int agx_world_read_scene_sweeps(agx_world* w, uint32_t* out, uint32_t n) {
    if (!w || !out) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (n != w->d.S) return fail(w, AGX_ERR_INVALID_ARGUMENT, "expected %u scenes, got %u", w->d.S, n);
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    HIP_CHECK(w, hipMemcpy(out, w->d.sceneSweeps, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return AGX_OK;
}

int agx_world_set_profiling(agx_world* w, uint32_t enabled) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    w->profiling = enabled != 0;
    return AGX_OK;
}

int agx_world_read_profile(agx_world* w, double* ms_total, uint32_t* launches, uint32_t cap, uint32_t* segments) {
    if (!w || !segments) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    for (const EventPair& p : w->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) { w->segMs[p.seg] += ms; w->segLaunches[p.seg]++; }
        w->eventPool.push_back(p.a);
        w->eventPool.push_back(p.b);
    }
    w->pending.clear();
    *segments = SEG_COUNT;
    for (uint32_t i = 0; i < SEG_COUNT && i < cap; i++) {
        if (ms_total) ms_total[i] = w->segMs[i];
        if (launches) launches[i] = w->segLaunches[i];
        w->segMs[i] = 0.0;
        w->segLaunches[i] = 0;
    }
    return AGX_OK;
}

const char* agx_profile_segment_name(uint32_t segment) { return segment < SEG_COUNT ? kSegmentNames[segment] : ""; }

int agx_broadphase_pairs(const float* aabb, const uint32_t* owner, const uint8_t* is_dynamic, const uint32_t* scene,
                         uint32_t n, uint32_t* pairs_a, uint32_t* pairs_b, uint64_t cap, uint64_t* count) {
    if (!count || (n && (!aabb || !owner || !is_dynamic))) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "null argument");
    *count = 0;
    if (n == 0) return AGX_OK;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) return fail(nullptr, AGX_ERR_NO_DEVICE, "no HIP device");
    agx_world* w = new (std::nothrow) agx_world();
    if (!w) return fail(nullptr, AGX_ERR_CAPACITY, "out of host memory");
    auto bail = [&](int code) { const std::string e = w->err; agx_world_destroy(w); g_lastError = e; return code; };
    agx::DeviceWorld& d = w->d;
    uint32_t S = 1;
    std::vector<uint32_t> sc(n, 0);
    if (scene) for (uint32_t i = 0; i < n; i++) { sc[i] = scene[i]; S = std::max(S, scene[i] + 1); }
    d.S = S;
    d.sceneBits = ceilLog2(S);
    if (d.sceneBits > 16) return bail(fail(w, AGX_ERR_INVALID_ARGUMENT, "too many scenes"));
    d.mortonBits = std::min(30u, (32u - d.sceneBits) & ~1u);
    d.C = n;
    std::vector<uint32_t> dynList, statList, body(owner, owner + n);
    for (uint32_t i = 0; i < n; i++) (is_dynamic[i] ? dynList : statList).push_back(i);
    // scene by scene (ascending index within a scene), so every bounds chunk holds one scene
    auto byScene = [&](uint32_t x, uint32_t y) { return sc[x] < sc[y]; };
    std::stable_sort(dynList.begin(), dynList.end(), byScene);
    std::stable_sort(statList.begin(), statList.end(), byScene);
    d.nDyn = (uint32_t)dynList.size();
    d.nStat = (uint32_t)statList.size();
    std::vector<float4> boxes(n);
    for (uint32_t i = 0; i < n; i++) boxes[i] = make_float4(aabb[4 * i], aabb[4 * i + 1], aabb[4 * i + 2], aabb[4 * i + 3]);
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return bail(fail(w, AGX_ERR_HIP, "hipGetDeviceProperties failed"));
    d.numCU = prop.multiProcessorCount;
    if (agx::deviceStream(&w->stream) != hipSuccess) return bail(fail(w, AGX_ERR_HIP, "stream"));
    if (upload(w, &d.aabb, boxes) != hipSuccess || upload(w, &d.colBody, body) != hipSuccess ||
        upload(w, &d.colScene, sc) != hipSuccess)
        return bail(fail(w, AGX_ERR_HIP, "upload failed"));
    int rc;
    if ((rc = allocTree(w, d.dyn, dynList, sc, S))) return bail(rc);
    if ((rc = allocTree(w, d.stat, statList, sc, S))) return bail(rc);
    if ((rc = allocPairBuffers(w, std::max<uint64_t>(cap, 1)))) return bail(rc);
    hipStream_t st = w->stream;
    if (agx::launchBuildTree(st, d, d.stat) != hipSuccess || agx::launchBuildTree(st, d, d.dyn) != hipSuccess ||
        agx::launchPairs(st, d) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return bail(fail(w, AGX_ERR_HIP, "broad phase failed: %s", hipGetErrorString(hipGetLastError())));
    agx::Counters c;
    if (hipMemcpy(&c, d.ctr, sizeof c, hipMemcpyDeviceToHost) != hipSuccess) return bail(fail(w, AGX_ERR_HIP, "read counters"));
    *count = c.pairs;
    if (c.overflow & 2u) return bail(fail(w, AGX_ERR_CAPACITY, "traversal stack overflow"));
    if ((uint64_t)c.pairs > cap) return bail(fail(w, AGX_ERR_CAPACITY, "%u pairs exceed cap", c.pairs));
    std::vector<uint2> h(c.pairs);
    if (c.pairs && hipMemcpy(h.data(), d.pairs, c.pairs * sizeof(uint2), hipMemcpyDeviceToHost) != hipSuccess)
        return bail(fail(w, AGX_ERR_HIP, "read pairs"));
    for (uint32_t i = 0; i < c.pairs; i++) { pairs_a[i] = h[i].x; pairs_b[i] = h[i].y; }
    agx_world_destroy(w);
    return AGX_OK;
}

}  // extern "C"

// ---- world snapshots (agx_world_save / agx_world_load)
namespace {

constexpr char SNAP_MAGIC[8] = { 'A', 'G', 'X', 'S', 'N', 'A', 'P', '1' };
constexpr uint32_t SNAP_VERSION = 1;
struct SnapHeader {
    char magic[8];
    uint32_t version, scenes, bodies, colliders, vertices, materials;
    uint64_t fingerprint;
    uint64_t payload;
    uint32_t manifolds, reserved[3];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");

enum SnapTag : uint32_t {
    TAG_SETTINGS = 1, TAG_HOST, TAG_POS, TAG_VEL, TAG_RCS, TAG_INV, TAG_MASS, TAG_INERTIA, TAG_POSE, TAG_WROT,
    TAG_PREVW, TAG_AABB, TAG_ISLAND, TAG_HIT, TAG_COUNTERS, TAG_MANIFOLDS, TAG_SCENE_SWEEPS, TAG_CALLBACK,
    TAG_THRESHOLD, TAG_COMMANDS, TAG_SLEEP_ID, TAG_SLEEP_TIMER, TAG_STASH, TAG_STASH_ISLAND, TAG_MOTOR, TAG_END
};
struct SnapHost {
    float accumulator;
    uint32_t lastStepCount, postUpdatePending, poseEdited;
};

// FNV-1a over the world's fixed structure: what a snapshot's arrays are indexed by.
uint64_t fingerprint(const agx_world* w) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    };
    const agx::DeviceWorld& d = w->d;
    const uint32_t counts[7] = { d.S, d.N, d.C, d.nVerts, d.nMat, w->staticBodies, d.nJoints };
    mix(counts, sizeof counts);
    auto vec = [&](const auto& v) { if (!v.empty()) mix(v.data(), v.size() * sizeof(v[0])); };
    vec(w->colType); vec(w->colMaterial); vec(w->colScene); vec(w->polyStart); vec(w->polyCount);
    vec(w->colP0); vec(w->colP1); vec(w->localComX); vec(w->localComY); vec(w->jointRecsHost);
    return h;
}

struct SnapWriter {
    unsigned char* out;   // null: count only
    uint64_t cap, at = 0;
    bool overflow = false;
    void raw(const void* p, uint64_t n) {
        if (out && at + n <= cap) std::memcpy(out + at, p, n);
        else if (out) overflow = true;
        at += n;
    }
    void section(uint32_t tag, const void* p, uint64_t n) {
        const uint32_t head[2] = { tag, 0u };
        raw(head, sizeof head);
        raw(&n, sizeof n);
        raw(p, n);
        static const unsigned char zeros[16] = {};
        raw(zeros, (16 - (n & 15)) & 15);
    }
};

// Copies of the device arrays a snapshot holds (host-side staging).
struct SnapArrays {
    std::vector<float4> pos, vel, pose, aabb;
    std::vector<float2> rcs, inv, prevW;
    std::vector<float> wRot, hit, threshold;
    std::vector<uint32_t> island, sceneSweeps, callback;
    std::vector<agx_manifold> man, stash;
    std::vector<uint32_t> sleepId, stashIsland;
    std::vector<float> sleepTimer, motorAcc;
    agx::Counters ctr;
};

template <class T>
hipError_t down(std::vector<T>& v, const T* p, size_t n) {
    v.resize(n);
    return n ? hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost) : hipSuccess;
}
template <class T>
hipError_t up(T* p, const std::vector<T>& v) {
    return v.empty() ? hipSuccess : hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// The state a snapshot holds, copied from the device.
int gatherState(agx_world* w, SnapArrays& a) {
    agx::DeviceWorld& d = w->d;
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    HIP_CHECK(w, hipMemcpy(&a.ctr, d.ctr, sizeof a.ctr, hipMemcpyDeviceToHost));
    HIP_CHECK(w, down(a.pos, d.pos, d.N)); HIP_CHECK(w, down(a.vel, d.vel, d.N)); HIP_CHECK(w, down(a.rcs, d.rcs, d.N));
    HIP_CHECK(w, down(a.inv, d.inv, d.N)); HIP_CHECK(w, down(a.island, d.islandId, d.N));
    HIP_CHECK(w, down(a.hit, d.hitTime, d.N));
    HIP_CHECK(w, down(a.pose, d.pose, d.C)); HIP_CHECK(w, down(a.wRot, d.wRot, d.C));
    HIP_CHECK(w, down(a.prevW, d.prevW, d.C)); HIP_CHECK(w, down(a.aabb, d.aabb, d.C));
    HIP_CHECK(w, down(a.callback, d.colCallback, d.C)); HIP_CHECK(w, down(a.threshold, d.colThreshold, d.C));
    HIP_CHECK(w, down(a.sceneSweeps, d.sceneSweeps, d.S));
    HIP_CHECK(w, down(a.man, d.man[1 - d.cur], a.ctr.prevManifolds));
    HIP_CHECK(w, down(a.motorAcc, d.motorAcc, d.nJoints));
    if (d.sleepId) {
        HIP_CHECK(w, down(a.sleepId, d.sleepId, d.N)); HIP_CHECK(w, down(a.sleepTimer, d.sleepTimer, d.N));
        HIP_CHECK(w, down(a.stash, d.stash, a.ctr.stashCount)); HIP_CHECK(w, down(a.stashIsland, d.stashIsland, a.ctr.stashCount));
    }
    return AGX_OK;
}

int writeSnapshot(agx_world* w, SnapWriter& sw) {
    agx::DeviceWorld& d = w->d;
    SnapArrays a;
    if (int rc = gatherState(w, a)) return rc;
    SnapHeader h{};
    std::memcpy(h.magic, SNAP_MAGIC, sizeof h.magic);
    h.version = SNAP_VERSION;
    h.scenes = d.S; h.bodies = d.N; h.colliders = d.C; h.vertices = d.nVerts; h.materials = d.nMat;
    h.fingerprint = fingerprint(w);
    h.manifolds = a.ctr.prevManifolds;
    const uint64_t headerAt = sw.at;
    sw.raw(&h, sizeof h);
    const uint64_t payloadStart = sw.at;
    const SnapHost host{ w->accumulator, w->lastStepCount, w->postUpdatePending ? 1u : 0u, w->poseEdited ? 1u : 0u };
    sw.section(TAG_SETTINGS, &w->settings, sizeof w->settings);
    sw.section(TAG_HOST, &host, sizeof host);
    sw.section(TAG_POS, a.pos.data(), a.pos.size() * sizeof(float4));
    sw.section(TAG_VEL, a.vel.data(), a.vel.size() * sizeof(float4));
    sw.section(TAG_RCS, a.rcs.data(), a.rcs.size() * sizeof(float2));
    sw.section(TAG_INV, a.inv.data(), a.inv.size() * sizeof(float2));
    sw.section(TAG_MASS, w->mass.data(), w->mass.size() * sizeof(float));
    sw.section(TAG_INERTIA, w->inertia.data(), w->inertia.size() * sizeof(float));
    sw.section(TAG_POSE, a.pose.data(), a.pose.size() * sizeof(float4));
    sw.section(TAG_WROT, a.wRot.data(), a.wRot.size() * sizeof(float));
    sw.section(TAG_PREVW, a.prevW.data(), a.prevW.size() * sizeof(float2));
    sw.section(TAG_AABB, a.aabb.data(), a.aabb.size() * sizeof(float4));
    sw.section(TAG_ISLAND, a.island.data(), a.island.size() * sizeof(uint32_t));
    sw.section(TAG_HIT, a.hit.data(), a.hit.size() * sizeof(float));
    sw.section(TAG_COUNTERS, &a.ctr, sizeof a.ctr);
    sw.section(TAG_MANIFOLDS, a.man.data(), a.man.size() * sizeof(agx_manifold));
    sw.section(TAG_SCENE_SWEEPS, a.sceneSweeps.data(), a.sceneSweeps.size() * sizeof(uint32_t));
    sw.section(TAG_CALLBACK, a.callback.data(), a.callback.size() * sizeof(uint32_t));
    sw.section(TAG_THRESHOLD, a.threshold.data(), a.threshold.size() * sizeof(float));
    sw.section(TAG_COMMANDS, w->commands.data(), w->commands.size() * sizeof(agx_world::BodyCommand));
    sw.section(TAG_SLEEP_ID, a.sleepId.data(), a.sleepId.size() * sizeof(uint32_t));
    sw.section(TAG_SLEEP_TIMER, a.sleepTimer.data(), a.sleepTimer.size() * sizeof(float));
    sw.section(TAG_STASH, a.stash.data(), a.stash.size() * sizeof(agx_manifold));
    sw.section(TAG_STASH_ISLAND, a.stashIsland.data(), a.stashIsland.size() * sizeof(uint32_t));
    sw.section(TAG_MOTOR, a.motorAcc.data(), a.motorAcc.size() * sizeof(float));
    sw.section(TAG_END, nullptr, 0);
    h.payload = sw.at - payloadStart;
    if (sw.out && !sw.overflow) std::memcpy(sw.out + headerAt, &h, sizeof h);
    return AGX_OK;
}

struct SnapReader {
    const unsigned char* in;
    uint64_t size, at = 0;
    // Next section, which must carry `tag` and (when expect != ~0) exactly `expect` bytes.
    const unsigned char* section(uint32_t tag, uint64_t expect, uint64_t* got) {
        if (at + 16 > size) return nullptr;
        uint32_t head[2];
        uint64_t n;
        std::memcpy(head, in + at, sizeof head);
        std::memcpy(&n, in + at + 8, sizeof n);
        if (head[0] != tag || (expect != ~0ull && n != expect) || n > size - at - 16) return nullptr;
        const unsigned char* p = in + at + 16;
        at += 16 + n + ((16 - (n & 15)) & 15);
        if (got) *got = n;
        return p;
    }
};

}  // namespace

extern "C" {

int agx_world_snapshot_size(agx_world* w, uint64_t* bytes) {
    if (!w || !bytes) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    SnapWriter sw{ nullptr, 0 };
    if (int r = writeSnapshot(w, sw)) return r;
    *bytes = sw.at;
    return AGX_OK;
}

int agx_world_save(agx_world* w, void* buffer, uint64_t cap, uint64_t* written) {
    if (!w || !buffer || !written) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    SnapWriter sw{ static_cast<unsigned char*>(buffer), cap };
    if (int r = writeSnapshot(w, sw)) return r;
    *written = sw.at;
    if (sw.overflow) return fail(w, AGX_ERR_CAPACITY, "snapshot needs %llu bytes", (unsigned long long)sw.at);
    return AGX_OK;
}

int agx_world_load(agx_world* w, const void* buffer, uint64_t bytes) {
    if (!w || !buffer) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    agx::DeviceWorld& d = w->d;
    const unsigned char* in = static_cast<const unsigned char*>(buffer);
    SnapHeader h;
    if (bytes < sizeof h) return fail(w, AGX_ERR_INVALID_ARGUMENT, "not a snapshot (too short)");
    std::memcpy(&h, in, sizeof h);
    if (std::memcmp(h.magic, SNAP_MAGIC, sizeof h.magic) != 0 || h.version != SNAP_VERSION)
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "not a snapshot of this version");
    if (h.scenes != d.S || h.bodies != d.N || h.colliders != d.C || h.vertices != d.nVerts || h.materials != d.nMat ||
        h.fingerprint != fingerprint(w))
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "snapshot of a different world");
    if (h.payload > bytes - sizeof h || h.manifolds > d.manCap)
        return fail(w, AGX_ERR_INVALID_ARGUMENT, "truncated or corrupt snapshot");
    SnapReader r{ in + sizeof h, h.payload };
    const uint64_t N = d.N, C = d.C;
    SnapArrays a;
    agx_settings settings;
    SnapHost host;
    uint64_t commandBytes = 0;
    auto take = [&](uint32_t tag, void* dst, uint64_t n) {
        const unsigned char* p = r.section(tag, n, nullptr);
        if (!p) return false;
        if (n) std::memcpy(dst, p, n);
        return true;
    };
    auto takeVec = [&](uint32_t tag, auto& v, uint64_t count) {
        v.resize(count);
        return take(tag, v.data(), count * sizeof(v[0]));
    };
    std::vector<float> mass, inertia;
    bool ok = take(TAG_SETTINGS, &settings, sizeof settings) && take(TAG_HOST, &host, sizeof host) &&
              takeVec(TAG_POS, a.pos, N) && takeVec(TAG_VEL, a.vel, N) && takeVec(TAG_RCS, a.rcs, N) &&
              takeVec(TAG_INV, a.inv, N) && takeVec(TAG_MASS, mass, N) && takeVec(TAG_INERTIA, inertia, N) &&
              takeVec(TAG_POSE, a.pose, C) && takeVec(TAG_WROT, a.wRot, C) && takeVec(TAG_PREVW, a.prevW, C) &&
              takeVec(TAG_AABB, a.aabb, C) && takeVec(TAG_ISLAND, a.island, N) && takeVec(TAG_HIT, a.hit, N) &&
              take(TAG_COUNTERS, &a.ctr, sizeof a.ctr) && takeVec(TAG_MANIFOLDS, a.man, h.manifolds) &&
              takeVec(TAG_SCENE_SWEEPS, a.sceneSweeps, d.S) && takeVec(TAG_CALLBACK, a.callback, C) &&
              takeVec(TAG_THRESHOLD, a.threshold, C);
    const unsigned char* cmd = ok ? r.section(TAG_COMMANDS, ~0ull, &commandBytes) : nullptr;
    ok = ok && cmd && commandBytes % sizeof(agx_world::BodyCommand) == 0;
    // sleeping state: both arrays or neither (a world that never slept), then the stash
    uint64_t sleepBytes = 0, timerBytes = 0;
    const unsigned char* sid = ok ? r.section(TAG_SLEEP_ID, ~0ull, &sleepBytes) : nullptr;
    const unsigned char* stm = sid ? r.section(TAG_SLEEP_TIMER, ~0ull, &timerBytes) : nullptr;
    const bool slept = sleepBytes != 0;
    ok = ok && sid && stm && (slept ? sleepBytes == N * 4 && timerBytes == N * 4 : timerBytes == 0) &&
         takeVec(TAG_STASH, a.stash, slept ? a.ctr.stashCount : 0) &&
         takeVec(TAG_STASH_ISLAND, a.stashIsland, slept ? a.ctr.stashCount : 0) &&
         takeVec(TAG_MOTOR, a.motorAcc, d.nJoints) && r.section(TAG_END, 0, nullptr);
    if (ok && slept) {
        a.sleepId.assign(reinterpret_cast<const uint32_t*>(sid), reinterpret_cast<const uint32_t*>(sid) + N);
        a.sleepTimer.assign(reinterpret_cast<const float*>(stm), reinterpret_cast<const float*>(stm) + N);
        for (uint32_t id : a.sleepId) ok = ok && (id == 0 || (id >= 2 && id - 2 < N));
        for (uint32_t id : a.stashIsland) ok = ok && id >= 2 && id - 2 < N;
        ok = ok && a.ctr.stashCount <= d.manCap;
    }
    if (!ok || a.ctr.prevManifolds != h.manifolds) return fail(w, AGX_ERR_INVALID_ARGUMENT, "truncated or corrupt snapshot");
    if (slept && !d.sleepId)
        if (int rc = allocSleep(w)) return rc;
    for (uint32_t cb : a.callback)
        if (cb > AGX_MAX_CALLBACK_ID) return fail(w, AGX_ERR_INVALID_ARGUMENT, "corrupt snapshot (callback id)");
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    HIP_CHECK(w, up(d.pos, a.pos)); HIP_CHECK(w, up(d.vel, a.vel)); HIP_CHECK(w, up(d.rcs, a.rcs));
    HIP_CHECK(w, up(d.inv, a.inv)); HIP_CHECK(w, up(d.islandId, a.island)); HIP_CHECK(w, up(d.hitTime, a.hit));
    HIP_CHECK(w, up(d.pose, a.pose)); HIP_CHECK(w, up(d.wRot, a.wRot)); HIP_CHECK(w, up(d.prevW, a.prevW));
    HIP_CHECK(w, up(d.aabb, a.aabb)); HIP_CHECK(w, up(d.colCallback, a.callback));
    HIP_CHECK(w, up(d.colThreshold, a.threshold)); HIP_CHECK(w, up(d.sceneSweeps, a.sceneSweeps));
    HIP_CHECK(w, up(d.motorAcc, a.motorAcc));
    if (d.sleepId) {   // a snapshot of a world that never slept: every body awake, the stash empty
        if (!slept) { a.sleepId.assign(N, 0u); a.sleepTimer.assign(N, 0.0f); a.ctr.stashCount = 0; }
        HIP_CHECK(w, up(d.sleepId, a.sleepId)); HIP_CHECK(w, up(d.sleepTimer, a.sleepTimer));
        HIP_CHECK(w, up(d.stash, a.stash)); HIP_CHECK(w, up(d.stashIsland, a.stashIsland));
    }
    // the warm-start source and its hash: built the way a step's FINISH builds them, then every
    // counter as saved
    HIP_CHECK(w, up(d.man[1 - d.cur], a.man));
    agx::Counters c = a.ctr;
    c.manifolds = h.manifolds;
    HIP_CHECK(w, hipMemcpy(d.ctr, &c, sizeof c, hipMemcpyHostToDevice));
    d.cur ^= 1;
    HIP_CHECK(w, agx::launchHashBuild(w->stream, d));
    d.cur ^= 1;
    HIP_CHECK(w, hipStreamSynchronize(w->stream));
    HIP_CHECK(w, hipMemcpy(d.ctr, &a.ctr, sizeof a.ctr, hipMemcpyHostToDevice));
    w->settings = settings;
    w->mass = mass;
    w->inertia = inertia;
    w->accumulator = host.accumulator;
    w->lastStepCount = host.lastStepCount;
    w->postUpdatePending = host.postUpdatePending != 0;
    w->poseEdited = host.poseEdited != 0;
    w->commands.assign(reinterpret_cast<const agx_world::BodyCommand*>(cmd),
                       reinterpret_cast<const agx_world::BodyCommand*>(cmd + commandBytes));
    w->colCallbackHost = a.callback;
    w->callbackColliders = 0;
    for (uint32_t cb : a.callback) w->callbackColliders += cb != 0 ? 1u : 0u;
    w->staticDirty = d.nStat > 0;   // static colliders' poses and tree follow the loaded bodies at the next pose pass
    w->histReady = false;
    w->nextPhase = 0;
    return AGX_OK;
}

}  // extern "C"

namespace {

// The world's copy of its builder brought to the current body state (positions, velocities,
// rotations, masses as changed by commands), and every body and collider keyed by (scene, slot).
int syncBuilder(agx_world* w, SnapArrays& a, std::vector<uint64_t>& oldBody, std::vector<uint64_t>& oldCol) {
    agx::DeviceWorld& d = w->d;
    if (int rc = gatherState(w, a)) return rc;
    agx::Builder& B = w->built->b;
    oldBody.assign(d.N, 0);
    oldCol.assign(d.C, 0);
    uint32_t g = 0, c = 0;
    for (uint32_t s = 0; s < d.S; s++) {
        agx::SceneData& sd = B.scenes[s];
        for (uint32_t i = 0; i < sd.bodyCount(); i++, g++) {
            oldBody[g] = ((uint64_t)s << 32) | sd.bodySlot[i];
            sd.comX[i] = a.pos[g].x; sd.comY[i] = a.pos[g].y; sd.rot[i] = a.pos[g].z;
            sd.velX[i] = a.vel[g].x; sd.velY[i] = a.vel[g].y; sd.angVel[i] = a.vel[g].z;
            sd.cosv[i] = a.rcs[g].x; sd.sinv[i] = a.rcs[g].y;
            sd.invMass[i] = a.inv[g].x; sd.invInertia[i] = a.inv[g].y;
            sd.mass[i] = w->mass[g]; sd.inertia[i] = w->inertia[g];
        }
        for (uint32_t k = 0; k < sd.colliderCount(); k++, c++) oldCol[c] = ((uint64_t)s << 32) | sd.colSlot[k];
    }
    return AGX_OK;
}

// Rebuilds w from its (changed) builder and moves everything the old world carried to the new
// dense indices by creation slot; bodies and colliders no longer in the builder are dropped with
// the manifolds, stashed contacts and commands that name them, and a sleeping island that lost a
// body wakes (its stashed contacts return to the warm-start source), as removing a body from a
// sleeping island does (ObjectManager.cpp:394-395). sceneStart receives each scene's first body.
int rebuildWorld(agx_world* w, const SnapArrays& a, const std::vector<uint64_t>& oldBody,
                 const std::vector<uint64_t>& oldCol, std::vector<uint32_t>& sceneStart) {
    agx::DeviceWorld& d = w->d;
    agx::Builder& B = w->built->b;
    agx_world* nw = nullptr;
    if (int rc = agx_world_create(w->built, &w->settings, &nw)) return fail(w, rc, "rebuild failed: %s", g_lastError.c_str());
    agx::DeviceWorld& e = nw->d;
    std::unordered_map<uint64_t, uint32_t> bodyAt, colAt;
    sceneStart.assign(e.S + 1, 0);
    {
        uint32_t g = 0, c = 0;
        for (uint32_t s = 0; s < e.S; s++) {
            const agx::SceneData& sd = B.scenes[s];
            sceneStart[s] = g;
            for (uint32_t i = 0; i < sd.bodyCount(); i++, g++) bodyAt[((uint64_t)s << 32) | sd.bodySlot[i]] = g;
            for (uint32_t k = 0; k < sd.colliderCount(); k++, c++) colAt[((uint64_t)s << 32) | sd.colSlot[k]] = c;
        }
        sceneStart[e.S] = g;
    }
    constexpr uint32_t GONE = UINT32_MAX;
    std::vector<uint32_t> nb(d.N), nc(d.C);
    for (uint32_t g = 0; g < d.N; g++) { auto it = bodyAt.find(oldBody[g]); nb[g] = it == bodyAt.end() ? GONE : it->second; }
    for (uint32_t c = 0; c < d.C; c++) { auto it = colAt.find(oldCol[c]); nc[c] = it == colAt.end() ? GONE : it->second; }
    auto kept = [&](const agx_manifold& m) { return nc[m.a] != GONE && nc[m.b] != GONE; };
    SnapArrays b;
    if (int rc = gatherState(nw, b)) { agx_world_destroy(nw); return fail(w, rc, "rebuild: reading the new world failed"); }
    for (uint32_t g = 0; g < d.N; g++) {
        const uint32_t k = nb[g];
        if (k == GONE) continue;
        b.pos[k] = a.pos[g]; b.vel[k] = a.vel[g]; b.rcs[k] = a.rcs[g]; b.inv[k] = a.inv[g];
        b.island[k] = a.island[g]; b.hit[k] = a.hit[g];
    }
    for (uint32_t c = 0; c < d.C; c++) {
        const uint32_t k = nc[c];
        if (k == GONE) continue;
        b.pose[k] = a.pose[c]; b.wRot[k] = a.wRot[c]; b.prevW[k] = a.prevW[c]; b.aabb[k] = a.aabb[c];
        b.callback[k] = a.callback[c]; b.threshold[k] = a.threshold[c];
    }
    b.man.clear();   // the warm-start source follows its colliders
    for (agx_manifold m : a.man)
        if (kept(m)) { m.a = nc[m.a]; m.b = nc[m.b]; b.man.push_back(m); }
    b.sceneSweeps = a.sceneSweeps;
    b.motorAcc.assign(e.nJoints, 0.0f);
    if (e.nJoints == d.nJoints) b.motorAcc = a.motorAcc;   // joints removed with a body restart their motors
    const bool slept = !a.sleepId.empty();
    uint32_t stashCount = 0;
    if (slept) {   // an island keeps sleeping (id 2 + its smallest new body index) unless it lost a body
        if (int rc = allocSleep(nw)) { agx_world_destroy(nw); return fail(w, rc, "rebuild: sleeping state"); }
        std::unordered_map<uint32_t, uint32_t> idOf;
        for (uint32_t g = 0; g < d.N; g++)
            if (a.sleepId[g]) {
                uint32_t& id = idOf.emplace(a.sleepId[g], UINT32_MAX).first->second;
                if (nb[g] == GONE) id = 0u;   // woken
                else if (id != 0u) id = std::min(id, 2 + nb[g]);
            }
        b.sleepId.assign(e.N, 0u);
        b.sleepTimer.assign(e.N, 0.0f);
        for (uint32_t g = 0; g < d.N; g++) {
            if (nb[g] == GONE) continue;
            const uint32_t id = a.sleepId[g] ? idOf.at(a.sleepId[g]) : 0u;
            b.sleepId[nb[g]] = id;
            b.sleepTimer[nb[g]] = id ? a.sleepTimer[g] : (a.sleepId[g] ? 0.0f : a.sleepTimer[g]);
        }
        b.stash.clear();
        b.stashIsland.clear();
        for (size_t i = 0; i < a.stash.size(); i++) {
            agx_manifold m = a.stash[i];
            if (!kept(m)) continue;
            m.a = nc[m.a]; m.b = nc[m.b];
            const uint32_t id = idOf.at(a.stashIsland[i]);
            if (id) { b.stash.push_back(m); b.stashIsland.push_back(id); }
            else b.man.push_back(m);   // a woken island's contacts warm-start again
        }
        stashCount = (uint32_t)b.stash.size();
    }
    int rc = AGX_OK;
    auto upAll = [&]() -> int {
        HIP_CHECK(nw, hipStreamSynchronize(nw->stream));
        HIP_CHECK(nw, up(e.pos, b.pos)); HIP_CHECK(nw, up(e.vel, b.vel)); HIP_CHECK(nw, up(e.rcs, b.rcs));
        HIP_CHECK(nw, up(e.inv, b.inv)); HIP_CHECK(nw, up(e.islandId, b.island)); HIP_CHECK(nw, up(e.hitTime, b.hit));
        HIP_CHECK(nw, up(e.pose, b.pose)); HIP_CHECK(nw, up(e.wRot, b.wRot)); HIP_CHECK(nw, up(e.prevW, b.prevW));
        HIP_CHECK(nw, up(e.aabb, b.aabb)); HIP_CHECK(nw, up(e.colCallback, b.callback));
        HIP_CHECK(nw, up(e.colThreshold, b.threshold)); HIP_CHECK(nw, up(e.sceneSweeps, b.sceneSweeps));
        HIP_CHECK(nw, up(e.motorAcc, b.motorAcc));
        if (slept) {
            HIP_CHECK(nw, up(e.sleepId, b.sleepId)); HIP_CHECK(nw, up(e.sleepTimer, b.sleepTimer));
            HIP_CHECK(nw, up(e.stash, b.stash)); HIP_CHECK(nw, up(e.stashIsland, b.stashIsland));
        }
        if (b.man.size() > e.manCap) return fail(nw, AGX_ERR_CAPACITY, "rebuild: too many warm-start manifolds");
        // the warm-start source and its hash, built the way a step's FINISH builds them; then the
        // counters that last across steps
        HIP_CHECK(nw, up(e.man[1 - e.cur], b.man));
        agx::Counters c = b.ctr;
        c.manifolds = (uint32_t)b.man.size();
        c.gate = a.ctr.gate;
        c.solveSteps = a.ctr.solveSteps; c.sweepsTotal = a.ctr.sweepsTotal;
        c.bodySweepsTotal = a.ctr.bodySweepsTotal; c.convergedBodiesTotal = a.ctr.convergedBodiesTotal;
        c.stashCount = stashCount;
        HIP_CHECK(nw, hipMemcpy(e.ctr, &c, sizeof c, hipMemcpyHostToDevice));
        e.cur ^= 1;
        HIP_CHECK(nw, agx::launchHashBuild(nw->stream, e));
        e.cur ^= 1;
        HIP_CHECK(nw, hipStreamSynchronize(nw->stream));
        return AGX_OK;
    };
    if ((rc = upAll())) { std::string m = nw->err; agx_world_destroy(nw); return fail(w, rc, "rebuild: %s", m.c_str()); }
    nw->accumulator = w->accumulator;
    nw->lastStepCount = w->lastStepCount;
    nw->postUpdatePending = w->postUpdatePending;
    nw->poseEdited = w->poseEdited;
    nw->staticDirty = true;   // static poses, AABBs and tree follow the bodies at the next pose pass
    nw->profiling = w->profiling;
    if (w->mlp) {   // the learned warm start's buffers, then the gate as it was
        if ((rc = agx_world_set_learned(nw, w->mlp, w->learnMode)) ||
            hipMemcpy(&nw->d.ctr->gate, &a.ctr.gate, sizeof a.ctr.gate, hipMemcpyHostToDevice) != hipSuccess) {
            agx_world_destroy(nw);
            return fail(w, rc ? rc : AGX_ERR_HIP, "rebuild: learned warm start");
        }
    }
    for (agx_world::BodyCommand c : w->commands)
        if (nb[c.body] != GONE) { c.body = nb[c.body]; nw->commands.push_back(c); }
    nw->colCallbackHost = b.callback;
    nw->callbackColliders = 0;
    for (uint32_t cb : b.callback) nw->callbackColliders += cb != 0 ? 1u : 0u;
    if (w->d.solveBlocks / w->d.numCU <= nw->solvePerCuMax && w->d.colourBlocks / w->d.numCU <= nw->colourPerCuMax) {
        e.solveBlocks = w->d.solveBlocks;
        e.colourBlocks = w->d.colourBlocks;
    }
    std::swap(*w, *nw);
    agx_world_destroy(nw);
    return AGX_OK;
}

}  // namespace

extern "C" {

// ObjectAPI::create* after the upload: the bodies are created in the world's copy of its builder
// (the reference's group moves included) and the world is rebuilt from it.
int agx_world_add_bodies(agx_world* w, uint32_t scene, uint32_t n, const uint8_t* type, const float* px, const float* py,
                         const float* rot, const float* vx, const float* vy, const float* av, const float* mass,
                         const uint32_t* material, const float* p0, const float* p1, const uint32_t* vstart,
                         const uint32_t* vcount, const float* verts, uint32_t* out_index) {
    if (!w) return fail(nullptr, AGX_ERR_INVALID_ARGUMENT, "world is null");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    if (!w->built || scene >= w->d.S) return fail(w, AGX_ERR_INVALID_ARGUMENT, "scene %u out of range", scene);
    SnapArrays a;
    std::vector<uint64_t> oldBody, oldCol;
    if (int rc = syncBuilder(w, a, oldBody, oldCol)) return rc;
    agx::Builder& B = w->built->b;
    const agx::SceneData backup = B.scenes[scene];
    std::vector<uint32_t> local(n, UINT32_MAX);
    const int rejected = agx_builder_add_bodies(w->built, scene, n, type, px, py, rot, vx, vy, av, mass, material, p0, p1,
                                                vstart, vcount, verts, local.data());
    if (rejected < 0) { B.scenes[scene] = backup; return fail(w, rejected, "%s", g_lastError.c_str()); }
    const std::string rejectMsg = rejected ? g_lastError : std::string();
    std::vector<uint32_t> sceneStart;
    if (int rc = rebuildWorld(w, a, oldBody, oldCol, sceneStart)) { B.scenes[scene] = backup; return rc; }
    if (out_index)
        for (uint32_t i = 0; i < n; i++) out_index[i] = local[i] == UINT32_MAX ? UINT32_MAX : sceneStart[scene] + local[i];
    if (rejected) fail(w, 0, "%s", rejectMsg.c_str());
    return rejected;
}

// ObjectAPI::destroyBody after the upload (ObjectManager::destroyBodyImmediate,
// ObjectManager.cpp:390-418): the bodies leave the world's copy of its builder with their colliders
// and joints, and the world is rebuilt from it.
int agx_world_remove_bodies(agx_world* w, uint32_t n, const uint32_t* bodies) {
    if (!w || (n && !bodies)) return fail(w, AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (w->inStep) return fail(w, AGX_ERR_STATE, "a step is in progress");
    if (!w->built) return fail(w, AGX_ERR_STATE, "the world keeps no scene copy");
    agx::DeviceWorld& d = w->d;
    for (uint32_t i = 0; i < n; i++)
        if (bodies[i] >= d.N) return fail(w, AGX_ERR_INVALID_ARGUMENT, "body %u out of range (%u bodies)", bodies[i], d.N);
    SnapArrays a;
    std::vector<uint64_t> oldBody, oldCol;
    if (int rc = syncBuilder(w, a, oldBody, oldCol)) return rc;
    agx::Builder& B = w->built->b;
    const std::vector<agx::SceneData> backup = B.scenes;
    for (uint32_t i = 0; i < n; i++) {   // by slot: each removal moves others
        const uint32_t s = (uint32_t)(oldBody[bodies[i]] >> 32), slot = (uint32_t)oldBody[bodies[i]];
        agx::SceneData& sd = B.scenes[s];
        uint32_t at = UINT32_MAX;
        for (uint32_t k = 0; k < sd.bodyCount(); k++)
            if (sd.bodySlot[k] == slot) { at = k; break; }
        if (at != UINT32_MAX) sd.destroyBody(at);   // a body named twice is removed once
    }
    std::vector<uint32_t> sceneStart;
    if (int rc = rebuildWorld(w, a, oldBody, oldCol, sceneStart)) { B.scenes = backup; return rc; }
    return AGX_OK;
}

}  // extern "C"
