// C API of the oracle for ctypes (test infrastructure only; see oracle.h).
#include "oracle.h"
#include "solver_internal.h"

#include <cstring>
#include <string>

using namespace orc;

namespace {

std::vector<float>* f32Field(World& w, const char* name) {
    struct E { const char* n; std::vector<float> World::*p; };
    static const E table[] = {
        { "comX", &World::comX }, { "comY", &World::comY }, { "velX", &World::velX }, { "velY", &World::velY },
        { "rot", &World::rot }, { "angVel", &World::angVel }, { "mass", &World::mass }, { "invMass", &World::invMass },
        { "inertia", &World::inertia }, { "invInertia", &World::invInertia }, { "cos", &World::cosv }, { "sin", &World::sinv },
        { "localComX", &World::localComX }, { "localComY", &World::localComY }, { "boundRadius", &World::boundRadius },
        { "localOffX", &World::localOffX }, { "localOffY", &World::localOffY }, { "comOffX", &World::comOffX },
        { "comOffY", &World::comOffY }, { "localRot", &World::localRot }, { "wX", &World::wX }, { "wY", &World::wY },
        { "wRot", &World::wRot }, { "wCos", &World::wCos }, { "wSin", &World::wSin }, { "minX", &World::minX },
        { "minY", &World::minY }, { "maxX", &World::maxX }, { "maxY", &World::maxY }, { "p0", &World::p0 }, { "p1", &World::p1 },
        { "minExtent", &World::minExtent }, { "maxExtent", &World::maxExtent }, { "prevComX", &World::prevComX },
        { "prevComY", &World::prevComY }, { "prevWX", &World::prevWX }, { "prevWY", &World::prevWY },
    };
    for (const E& e : table) if (std::strcmp(e.n, name) == 0) return &(w.*(e.p));
    return nullptr;
}

std::vector<uint32_t>* u32Field(World& w, const char* name) {
    struct E { const char* n; std::vector<uint32_t> World::*p; };
    static const E table[] = {
        { "bodySlot", &World::bodySlot }, { "islandId", &World::islandId }, { "colBody", &World::colBody },
        { "colMaterial", &World::colMaterial }, { "colSlot", &World::colSlot }, { "polyStart", &World::polyStart },
        { "polyCount", &World::polyCount },
    };
    for (const E& e : table) if (std::strcmp(e.n, name) == 0) return &(w.*(e.p));
    return nullptr;
}

World& W(void* h) { return *static_cast<World*>(h); }

}  // namespace

extern "C" {

void* orc_new(void) { World* w = new World(); w->materials.push_back(Material{}); return w; }
void orc_free(void* h) { delete static_cast<World*>(h); }

void orc_set_settings(void* h, float gx, float gy, uint32_t vIt, uint32_t pIt, int ccd, float fast, float core) {
    Settings& s = W(h).settings;
    s.gravityX = gx; s.gravityY = gy; s.velocityIterations = vIt; s.positionIterations = pIt;
    s.ccd = ccd != 0; s.ccdFastFactor = fast; s.ccdCoreFraction = core;
}

uint32_t orc_add_material(void* h, float e, float us, float ud) { return addMaterial(W(h), Material{ e, us, ud }); }

// Vectorised creation in call order; out_index receives each body's dense index at its
// creation (later static creations may move earlier bodies; read colBody/bodySlot after).
int orc_add_bodies(void* h, uint32_t n, const uint8_t* type, const float* px, const float* py, const float* rot,
                   const float* vx, const float* vy, const float* av, const float* mass, const uint32_t* material,
                   const float* p0, const float* p1, const uint32_t* vstart, const uint32_t* vcount,
                   const float* vxy, uint32_t* out_index) {
    World& w = W(h);
    int bad = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t r;
        if (type[i] == SHAPE_CIRCLE) r = createCircle(w, px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i], p0[i]);
        else if (type[i] == SHAPE_BOX) r = createBox(w, px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i], p0[i], p1[i]);
        else if (type[i] == SHAPE_CAPSULE) r = createCapsule(w, px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i], p0[i], p1[i]);
        else r = createPolygon(w, px[i], py[i], rot[i], vx[i], vy[i], av[i], mass[i], material[i],
                               reinterpret_cast<const V2*>(vxy) + vstart[i], vcount[i]);
        if (out_index) out_index[i] = r;
        if (r == UINT32_MAX) bad++;
    }
    return bad;
}

// Simulation::refreshSpatialData (Simulation.cpp:561-579): cos/sin of every body, poses and AABBs of every collider.
void orc_refresh(void* h) {
    World& w = W(h);
    for (uint32_t i = 0; i < w.bodyCount(); i++) fastCosSin(w.rot[i], w.cosv[i], w.sinv[i]);
    colliderPoses(w, true);
    colliderAABBs(w, true);
    w.spatialDirty = false;
    w.postUpdateDone = true;
}

void orc_counts(void* h, uint32_t* out) {
    World& w = W(h);
    out[0] = w.bodyCount(); out[1] = w.staticBodies; out[2] = w.colliderCount(); out[3] = w.staticColliders;
    out[4] = (uint32_t)w.verts.size(); out[5] = (uint32_t)w.materials.size();
    out[6] = (uint32_t)w.manifolds.size(); out[7] = (uint32_t)w.pairA.size(); out[8] = w.islandCount;
    out[9] = (uint32_t)w.prevManifolds.size();
}

int orc_get_f32(void* h, const char* name, float* out, uint32_t n) {
    std::vector<float>* v = f32Field(W(h), name);
    if (!v || v->size() != n) return -1;
    std::memcpy(out, v->data(), n * 4u);
    return 0;
}
int orc_set_f32(void* h, const char* name, const float* in, uint32_t n) {
    std::vector<float>* v = f32Field(W(h), name);
    if (!v) return -1;
    v->assign(in, in + n);
    return 0;
}
int orc_get_u32(void* h, const char* name, uint32_t* out, uint32_t n) {
    std::vector<uint32_t>* v = u32Field(W(h), name);
    if (!v || v->size() != n) return -1;
    std::memcpy(out, v->data(), n * 4u);
    return 0;
}
int orc_get_types(void* h, uint8_t* out, uint32_t n) {
    World& w = W(h);
    if (w.colType.size() != n) return -1;
    std::memcpy(out, w.colType.data(), n);
    return 0;
}
int orc_get_verts(void* h, float* vxy, float* nxy, uint32_t n) {
    World& w = W(h);
    if (w.verts.size() != n) return -1;
    std::memcpy(vxy, w.verts.data(), n * 8u);
    std::memcpy(nxy, w.normals.data(), n * 8u);
    return 0;
}

void orc_fast_cossin(const float* x, float* c, float* s, uint32_t n) { for (uint32_t i = 0; i < n; i++) fastCosSin(x[i], c[i], s[i]); }
void orc_wrap(const float* x, float* y, uint32_t n) { for (uint32_t i = 0; i < n; i++) y[i] = wrapRotation(x[i]); }

void orc_snapshot(void* h) { snapshotPoses(W(h)); }
void orc_integrate(void* h, float dt) { integrate(W(h), dt); }
void orc_wrap_cossin(void* h) { wrapAndCosSin(W(h)); }
void orc_poses(void* h, int all) { colliderPoses(W(h), all != 0); }
void orc_aabbs(void* h, int all) { colliderAABBs(W(h), all != 0); }
int orc_ccd(void* h, float* hitTime) {
    std::vector<float> t;
    const bool any = ccdResolve(W(h), &t);
    if (hitTime) std::memcpy(hitTime, t.data(), t.size() * 4u);
    return any ? 1 : 0;
}

int64_t orc_broad_bvh(void* h, uint32_t* a, uint32_t* b, uint64_t cap) {
    World& w = W(h);
    broadPhaseBvh(w);
    const uint64_t n = w.pairA.size();
    if (n > cap) return -(int64_t)n;
    std::memcpy(a, w.pairA.data(), n * 4u);
    std::memcpy(b, w.pairB.data(), n * 4u);
    return (int64_t)n;
}
int64_t orc_broad_brute(void* h, uint32_t* a, uint32_t* b, uint64_t cap) {
    std::vector<uint32_t> pa, pb;
    broadPhaseBrute(W(h), pa, pb);
    if (pa.size() > cap) return -(int64_t)pa.size();
    std::memcpy(a, pa.data(), pa.size() * 4u);
    std::memcpy(b, pb.data(), pb.size() * 4u);
    return (int64_t)pa.size();
}

int64_t orc_narrow(void* h, const uint32_t* pa, const uint32_t* pb, uint64_t n, Manifold* out, uint64_t cap) {
    std::vector<Manifold> m;
    narrowPhase(W(h), pa, pb, n, m);
    if (m.size() > cap) return -(int64_t)m.size();
    std::memcpy(out, m.data(), m.size() * sizeof(Manifold));
    return (int64_t)m.size();
}

void orc_set_prev_manifolds(void* h, const Manifold* m, uint64_t n) { W(h).prevManifolds.assign(m, m + n); }
void orc_warm_start(void* h, Manifold* m, uint64_t n) {
    std::vector<Manifold> v(m, m + n);
    warmStartFromCache(W(h), v);
    std::memcpy(m, v.data(), n * sizeof(Manifold));
}

void orc_solve(void* h, Manifold* m, uint64_t n, const uint32_t* order, uint32_t vIt, uint32_t pIt, float* residual) {
    std::vector<Manifold> v(m, m + n);
    solve(W(h), v, order, vIt, pIt, residual);
    std::memcpy(m, v.data(), n * sizeof(Manifold));
}

uint32_t orc_islands(void* h, const Manifold* m, uint64_t n) {
    World& w = W(h);
    islands(w, std::vector<Manifold>(m, m + n));
    return w.islandCount;
}

void orc_step(void* h, float dt) { step(W(h), dt); }
void orc_post_update(void* h) { postUpdate(W(h)); }
void orc_update(void* h, float dt, uint32_t steps) {
    flushCommands(W(h));   // Simulation::update flushes first (Simulation.cpp:38-39)
    W(h).events.clear();
    W(h).eventStep = 0;
    for (uint32_t i = 0; i < steps; i++) step(W(h), dt);
    postUpdate(W(h));
}
int64_t orc_get_manifolds(void* h, Manifold* out, uint64_t cap) {
    World& w = W(h);
    if (w.manifolds.size() > cap) return -(int64_t)w.manifolds.size();
    std::memcpy(out, w.manifolds.data(), w.manifolds.size() * sizeof(Manifold));
    return (int64_t)w.manifolds.size();
}
int64_t orc_get_prev_manifolds(void* h, Manifold* out, uint64_t cap) {
    World& w = W(h);
    if (w.prevManifolds.size() > cap) return -(int64_t)w.prevManifolds.size();
    if (!w.prevManifolds.empty()) std::memcpy(out, w.prevManifolds.data(), w.prevManifolds.size() * sizeof(Manifold));
    return (int64_t)w.prevManifolds.size();
}
void orc_errors(void* h, float* out) { out[0] = W(h).velocityError; out[1] = W(h).positionError; }

// Simulation::setColliderCollisionCallback (Simulation.cpp:152-172): id 0 clears.
void orc_set_callback(void* h, uint32_t collider, uint32_t id, float threshold) {
    World& w = W(h);
    if (w.colCallback.size() < w.colliderCount()) {
        w.colCallback.resize(w.colliderCount(), 0u);
        w.colThreshold.resize(w.colliderCount(), 3.40282347e38f);
    }
    w.colCallback[collider] = id;
    w.colThreshold[collider] = id ? threshold : 3.40282347e38f;
}
void orc_set_events_enabled(void* h, int on) { W(h).settings.collisionEvents = on != 0; }
// Events of manifolds m on the current state (index order), then every event kept since the last update call.
int64_t orc_collision_events(void* h, const Manifold* m, uint64_t n, Event* out, uint64_t cap) {
    World& w = W(h);
    std::vector<Event> keep;
    keep.swap(w.events);
    collisionEvents(w, std::vector<Manifold>(m, m + n), nullptr);
    std::vector<Event> got;
    got.swap(w.events);
    w.events.swap(keep);
    if (got.size() > cap) return -(int64_t)got.size();
    if (!got.empty()) std::memcpy(out, got.data(), got.size() * sizeof(Event));
    return (int64_t)got.size();
}
int64_t orc_get_events(void* h, Event* out, uint64_t cap) {
    World& w = W(h);
    if (w.events.size() > cap) return -(int64_t)w.events.size();
    if (!w.events.empty()) std::memcpy(out, w.events.data(), w.events.size() * sizeof(Event));
    return (int64_t)w.events.size();
}

// Sleeping: thresholds (SimulationSettings::sleepSpeedThreshold / sleepTimeThreshold), per-body
// state, the stash, and the wake and sleep passes of one step on given manifolds.
void orc_set_sleep(void* h, float speed, float time) {
    W(h).settings.sleepSpeedThreshold = speed;
    W(h).settings.sleepTimeThreshold = time;
}
void orc_get_sleep(void* h, uint32_t* ids, float* timers, uint32_t n) {
    World& w = W(h);
    for (uint32_t b = 0; b < n; b++) {
        ids[b] = b < w.sleepId.size() ? w.sleepId[b] : 0u;
        timers[b] = b < w.sleepTimer.size() ? w.sleepTimer[b] : 0.0f;
    }
}
void orc_set_sleep_state(void* h, const uint32_t* ids, const float* timers, uint32_t n) {
    W(h).sleepId.assign(ids, ids + n);
    W(h).sleepTimer.assign(timers, timers + n);
}
int64_t orc_get_stash(void* h, Manifold* out, uint32_t* island, uint64_t cap) {
    World& w = W(h);
    if (w.stash.size() > cap) return -(int64_t)w.stash.size();
    if (!w.stash.empty()) {
        std::memcpy(out, w.stash.data(), w.stash.size() * sizeof(Manifold));
        std::memcpy(island, w.stashIsland.data(), w.stash.size() * sizeof(uint32_t));
    }
    return (int64_t)w.stash.size();
}
void orc_set_stash(void* h, const Manifold* m, const uint32_t* island, uint64_t n) {
    W(h).stash.assign(m, m + n);
    W(h).stashIsland.assign(island, island + n);
}
// wakeContacts on m (n manifolds): out receives m plus the restored manifolds; returns the count.
int64_t orc_wake_contacts(void* h, const Manifold* m, uint64_t n, Manifold* out, uint64_t cap) {
    std::vector<Manifold> v(m, m + n);
    wakeContacts(W(h), v);
    if (v.size() > cap) return -(int64_t)v.size();
    std::memcpy(out, v.data(), v.size() * sizeof(Manifold));
    return (int64_t)v.size();
}
// updateSleeping after islands on m: out receives the manifolds kept awake; returns their count.
int64_t orc_update_sleeping(void* h, float dt, const Manifold* m, uint64_t n, Manifold* out) {
    std::vector<Manifold> v(m, m + n);
    updateSleeping(W(h), dt, v);
    std::memcpy(out, v.data(), v.size() * sizeof(Manifold));
    return (int64_t)v.size();
}

// Joints: creation (-1 if one names a missing body), per-type iterations, the pass order to replay
// (order: n joint indices; colourStart: k + 1 boundaries; typeColour: 6 entries; n = 0 restores each
// type's index order), and the motors' accumulated impulses.
int orc_add_joints(void* h, uint32_t n, const JointDesc* d) {
    for (uint32_t i = 0; i < n; i++)
        if (!addJoint(W(h), d[i])) return -1;
    return 0;
}
void orc_set_joint_iterations(void* h, const uint32_t* it) {
    for (int t = 0; t < 5; t++) W(h).settings.jointIterations[t] = it[t];
}
void orc_set_joint_order(void* h, const uint32_t* order, uint32_t n, const uint32_t* colourStart, uint32_t k,
                         const uint32_t* typeColour) {
    World& w = W(h);
    w.jointOrder.assign(order, order + n);
    w.jointColourStart.assign(colourStart, colourStart + (n ? k + 1 : 0));
    w.jointTypeColour.assign(typeColour, typeColour + (n ? JOINT_TYPES + 1 : 0));
}
int orc_get_motor_impulses(void* h, float* out, uint32_t n) {
    World& w = W(h);
    if (w.motorAcc.size() != n) return -1;
    if (n) std::memcpy(out, w.motorAcc.data(), n * sizeof(float));
    return 0;
}
int orc_set_motor_impulses(void* h, const float* in, uint32_t n) {
    World& w = W(h);
    if (w.motorAcc.size() != n) return -1;
    if (n) std::memcpy(w.motorAcc.data(), in, n * sizeof(float));
    return 0;
}

void orc_queue_body(void* h, uint32_t kind, uint32_t body, float a, float b) { W(h).commands.push_back({ kind, body, a, b }); }
void orc_flush(void* h) { flushCommands(W(h)); }

// The threaded step (oracle/threaded.cpp): a pool of `participants` (pinned threads + the caller).
void* orc_pool_new(uint32_t participants) { return poolCreate(participants); }
void orc_pool_free(void* p) { poolDestroy(static_cast<Pool*>(p)); }
uint32_t orc_pool_size(void* p) { return poolSize(static_cast<Pool*>(p)); }
void orc_update_pool(void* h, void* pool, float dt, uint32_t steps) {
    flushCommands(W(h));
    W(h).events.clear();
    W(h).eventStep = 0;
    for (uint32_t i = 0; i < steps; i++) step(W(h), dt, static_cast<Pool*>(pool));
    postUpdate(W(h));
}
// The solver's wave plan of manifolds m for `workers` workers: flat order (n entries), then the
// passes' starts and sizes (wave-major) into passStart/passSize up to cap; returns the pass count.
int64_t orc_plan_waves(void* h, const Manifold* m, uint64_t n, uint32_t workers, uint32_t* flat, uint32_t* passStart,
                       uint32_t* passSize, uint64_t cap) {
    const WavePlan plan = planWaves(W(h), std::vector<Manifold>(m, m + n), workers);
    if (plan.passStart.size() > cap) return -(int64_t)plan.passStart.size();
    std::memcpy(flat, plan.flat.data(), plan.flat.size() * sizeof(uint32_t));
    std::memcpy(passStart, plan.passStart.data(), plan.passStart.size() * sizeof(uint32_t));
    std::memcpy(passSize, plan.passSize.data(), plan.passSize.size() * sizeof(uint32_t));
    return (int64_t)plan.passStart.size();
}
// BodyCollisionSolver::solveCollisions with its threaded path on the pool (m solved in place).
void orc_solve_threaded(void* h, void* pool, Manifold* m, uint64_t n, uint32_t vIt, uint32_t pIt) {
    std::vector<Manifold> ms(m, m + n);
    solveThreaded(W(h), ms, *static_cast<Pool*>(pool), vIt, pIt);
    std::memcpy(m, ms.data(), n * sizeof(Manifold));
}
// NarrowPhaseCollisionDetector::findCollisions with its threaded path on the pool.
int64_t orc_narrow_threaded(void* h, void* pool, const uint32_t* pa, const uint32_t* pb, uint64_t n, Manifold* out,
                            uint64_t cap) {
    std::vector<Manifold> ms;
    narrowPhaseThreaded(W(h), pa, pb, n, ms, *static_cast<Pool*>(pool));
    if (ms.size() > cap) return -(int64_t)ms.size();
    std::memcpy(out, ms.data(), ms.size() * sizeof(Manifold));
    return (int64_t)ms.size();
}

}  // extern "C"
