// Oracle: CPU restatement of PS_AGONY2D's per-step collide-and-solve path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
// path: it may be loaded by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, and by nothing else. The product never links,
// loads or calls it.
//
// Every function follows the reference file:line named in its comment. The
// reference source (a C++20 engine) is not buildable in this image (it
// includes <tracy/Tracy.hpp> unconditionally and needs the C++17 library
// constant std::hardware_destructive_interference_size, absent from
// libstdc++ 11), so parity is pinned by closed forms, brute force and
// invariants in tests/test_oracle_*.py instead of by a compiled reference.
//
// Build with -ffp-contract=off: arithmetic is written in the reference's
// operation order, and the HIP path uses the same order, so phases with the
// same inputs agree bit for bit.
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>

namespace orc {

struct V2 { float x, y; };
inline V2 operator+(V2 a, V2 b) { return { a.x + b.x, a.y + b.y }; }
inline V2 operator-(V2 a, V2 b) { return { a.x - b.x, a.y - b.y }; }
inline V2 operator-(V2 a) { return { -a.x, -a.y }; }
inline V2 operator*(V2 a, float s) { return { a.x * s, a.y * s }; }
inline V2 operator*(float s, V2 a) { return { s * a.x, s * a.y }; }
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

enum ShapeType : uint8_t { SHAPE_CIRCLE = 0, SHAPE_BOX = 1, SHAPE_POLYGON = 2, SHAPE_CAPSULE = 3 };   // Types/Types.h:35-42

struct Material { float elasticity = 1.0f, staticFriction = 0.0f, dynamicFriction = 0.0f; };

// One manifold; field order matches the device record (BodyCollisionData,
// Narrow/NarrowPhaseCollisionDetector.h:26-61, with handles replaced by the
// dense collider indices of the canonical A/B).
struct Manifold {
    uint32_t a, b;
    float nx, ny;
    float depth;
    uint32_t count;
    float p0x, p0y, p1x, p1y;
    uint32_t id0, id1;
    float jn0, jt0, jn1, jt1;
};
static_assert(sizeof(Manifold) == 64, "manifold record is 64 bytes");

// Collision event (Types/CollisionEvent.h:18-27) with handles replaced by dense indices; the
// field order matches the device record agx_collision_event.
struct Event {
    uint32_t self, other, selfBody, otherBody;
    float px, py, nx, ny, impulse;
    uint32_t callback, step, reserved;
};
static_assert(sizeof(Event) == 48, "event record is 48 bytes");

// A joint (ConstraintSystems/*ConstraintSystem.h CreateParams); field order matches agx_joint_desc.
enum JointType : uint32_t { JOINT_SPRING = 0, JOINT_REVOLUTE, JOINT_PRISMATIC, JOINT_SLIDER, JOINT_MOTOR, JOINT_TYPES };
struct JointDesc {
    uint32_t type, bodyA, bodyB, enableLimit;
    float anchorAX, anchorAY, anchorBX, anchorBY, axisX, axisY, restLength, stiffness, damping;
    float lowerLimit, upperLimit, targetSpeed, maxTorque;
};
struct Joint {
    JointDesc d;
    float referenceAngle = 0.0f;   // prismatic: relative rotation at creation
};

struct Settings {
    float gravityX = 0.0f, gravityY = -9.81f;
    uint32_t velocityIterations = 6;   // Simulation.h:56
    uint32_t positionIterations = 3;   // Simulation.h:57
    bool ccd = true;                   // Simulation.h:66
    bool collisionEvents = true;       // Simulation.h:60
    uint32_t jointIterations[5] = { 6, 6, 6, 6, 6 };   // ConstraintSystemBase.h:34
    float sleepSpeedThreshold = 0.1f;  // Simulation.h:71
    float sleepTimeThreshold = 3.40282347e38f;   // Simulation.h:72 (0.5 s there); the default here keeps sleeping off
    float ccdFastFactor = 1.0f;        // CCD/ContinuousCollisionDetector.h:31
    float ccdCoreFraction = 0.5f;      // CCD/ContinuousCollisionDetector.h:32
};

struct World {
    Settings settings;
    std::vector<Material> materials;   // index 0 = default material

    // Bodies, dense order: [Static | Dynamic] (GroupRanges.h:13; sleeping is off).
    uint32_t staticBodies = 0;
    std::vector<float> comX, comY, velX, velY, rot, angVel;
    std::vector<float> mass, invMass, inertia, invInertia, cosv, sinv;
    std::vector<float> localComX, localComY, boundRadius;
    std::vector<uint32_t> bodySlot, islandId;
    std::vector<std::vector<uint32_t>> bodyColliders;

    // Colliders, dense order follows the owner's group.
    uint32_t staticColliders = 0;
    std::vector<uint32_t> colBody, colMaterial, colSlot;
    std::vector<uint8_t> colType;
    std::vector<float> localOffX, localOffY, comOffX, comOffY, localRot;
    std::vector<float> wX, wY, wRot, wCos, wSin;
    std::vector<float> minX, minY, maxX, maxY;
    std::vector<float> p0, p1;                 // circle: radius; box: half width, half height
    std::vector<uint32_t> polyStart, polyCount;
    std::vector<float> minExtent, maxExtent;
    std::vector<V2> verts, normals;            // polygon pool (local frame, CCW)

    // Step state.
    std::vector<float> prevComX, prevComY, prevWX, prevWY;
    bool spatialDirty = true;
    bool postUpdateDone = true;
    std::vector<Manifold> prevManifolds;       // previous step's manifolds (warm start cache)
    std::vector<Manifold> manifolds;           // this step's manifolds
    std::vector<uint32_t> pairA, pairB;        // this step's broad-phase pairs
    std::vector<uint32_t> islandStart, islandMembers;
    uint32_t islandCount = 0;
    // BodyCollisionSolver::computeConstraintErrors (BodyCollisionSolver.cpp:132-183), after the solve.
    float velocityError = 0.0f, positionError = 0.0f;
    // Deferred body modifications (ObjectAPI.h:76-81), applied by flushCommands.
    struct BodyCommand { uint32_t kind, body; float a, b; };
    std::vector<BodyCommand> commands;
    // Collision callbacks per collider (ColliderSoA callbackId / eventThreshold; absent entries are
    // "none") and the events recorded by the steps of the current update call.
    std::vector<uint32_t> colCallback;
    std::vector<float> colThreshold;
    std::vector<Event> events;
    uint32_t eventStep = 0;
    // Sleeping (ObjectManager.cpp:1314-1491) with stable indices: a body sleeps in place, with
    // sleepId = its sleep island's id (>= 2, FIRST_SLEEP_ISLAND_ID, BodySoA.h:45: here 2 + the
    // island's smallest body index), 0 when awake; the manifolds of a sleeping island are kept in
    // the stash with their island id (NarrowPhaseCollisionDetector.cpp:299-346) until it wakes.
    // Joints (oracle/joints.cpp) and the order their velocity passes run in: jointOrder type by type,
    // colour by colour, jointColourStart the colour boundaries, jointTypeColour[t] the first colour of
    // type t (empty: each type's joints in index order, the reference's single-threaded order).
    std::vector<Joint> joints;
    std::vector<float> motorAcc;
    std::vector<uint32_t> jointOrder, jointColourStart, jointTypeColour;
    std::vector<uint32_t> sleepId;
    std::vector<float> sleepTimer;
    std::vector<Manifold> stash;
    std::vector<uint32_t> stashIsland;

    uint32_t bodyCount() const { return (uint32_t)comX.size(); }
    uint32_t colliderCount() const { return (uint32_t)colBody.size(); }
};

// ---- builder (SimulationImpl/ObjectManager.cpp:71-389, PhysicsGeometry.cpp) ----
uint32_t addMaterial(World& w, const Material& m);
// Returns the dense body index, or UINT32_MAX on invalid input (reference: nullopt).
uint32_t createCircle(World& w, float px, float py, float rot, float vx, float vy, float av,
                      float mass, uint32_t material, float radius);
uint32_t createBox(World& w, float px, float py, float rot, float vx, float vy, float av,
                   float mass, uint32_t material, float width, float height);
uint32_t createPolygon(World& w, float px, float py, float rot, float vx, float vy, float av,
                       float mass, uint32_t material, const V2* localVerts, uint32_t count);
// Capsule of full core length `length` along its local +x and radius `radius` (p0 = half length,
// p1 = radius of its collider).
uint32_t createCapsule(World& w, float px, float py, float rot, float vx, float vy, float av,
                       float mass, uint32_t material, float length, float radius);
float capsuleInertia(float mass, float length, float radius);   // PhysicsGeometry.cpp:87-117 (centre of mass 0)

// ---- phases ----
void fastCosSin(float x, float& c, float& s);                 // FastCosSin.cpp:26-109
float wrapRotation(float r);                                  // LeftOutFunctions.cpp:13-102
void snapshotPoses(World& w);                                 // CCD/ContinuousCollisionDetector.cpp:276-299
void integrate(World& w, float dt);                           // Integration/Integrator.cpp:99-209
void wrapAndCosSin(World& w);                                 // Simulation.cpp:291-292
void colliderPoses(World& w, bool all);                       // Simulation.cpp:673-736
void colliderAABBs(World& w, bool all);                       // ColliderAABBBuilder.cpp:8-136
bool ccdResolve(World& w, std::vector<float>* hitTimeOut);    // CCD/ContinuousCollisionDetector.cpp:442-531
void broadPhaseBvh(World& w);                                 // Broad/*: BVH build + pair traversal
void broadPhaseBrute(const World& w, std::vector<uint32_t>& a, std::vector<uint32_t>& b);
void narrowPhase(World& w, const uint32_t* pa, const uint32_t* pb, size_t n,
                 std::vector<Manifold>& out);                 // Narrow/NarrowPhaseCollisionDetector.cpp:638-767
void warmStartFromCache(const World& w, std::vector<Manifold>& cur);  // :272-297 + :692-767
// Sequential impulses in the given order (order[c] = manifold index of constraint c).
// residualOut (optional, size velocityIterations) receives r_k after velocity sweep k.
void solve(World& w, std::vector<Manifold>& m, const uint32_t* order,
           uint32_t vIt, uint32_t pIt, float* residualOut);   // Solving/BodyCollisionSolver.cpp:38-980
void islands(World& w, const std::vector<Manifold>& m);       // IslandManager.cpp:7-116
// The event recording of computeConstraintData (Solving/BodyCollisionSolver.cpp:300-388) over the
// manifolds in order[c] (null: index order) on the current, pre-warm-start velocities; appends to
// w.events with step index w.eventStep. No-op unless events are on and some collider has a callback.
void collisionEvents(World& w, const std::vector<Manifold>& m, const uint32_t* order);
// ---- joints (oracle/joints.cpp) ----
// Appends joints (body indices are current dense indices; false if one does not exist).
bool addJoint(World& w, const JointDesc& d);
// ConstraintScheduler::solve (ConstraintScheduler.cpp:9-84) over the joints, on the velocities
// between the velocity-Verlet halves.
void solveJoints(World& w, float dt);

// ---- sleeping (oracle/sleep.cpp) ----
bool asleep(const World& w, uint32_t body);
// ObjectManager::wakeIslandsOfContacts (ObjectManager.h:139-156) on this step's manifolds, then
// wakeContactData (NarrowPhaseCollisionDetector.cpp:348-414): the woken islands' stashed manifolds
// are appended to `cur` in stash order.
void wakeContacts(World& w, std::vector<Manifold>& cur);
// ObjectManager::updateSleeping (ObjectManager.cpp:1438-1491) after the islands, then
// sleepContactData (NarrowPhaseCollisionDetector.cpp:299-346) on `cur`.
void updateSleeping(World& w, float dt, std::vector<Manifold>& cur);
// Wake paths of ObjectManager::flushCommands (ObjectManager.cpp:638-700) and
// Simulation::wakeIslandsOfEditedStatics (Simulation.cpp:463-486): commanded bodies wake their
// islands; stashed manifolds of edited static bodies' colliders are dropped and their islands woken;
// awake partners of those colliders restart their timers. Restored manifolds go to prevManifolds.
void wakeForCommands(World& w, const std::vector<uint32_t>& commandedBodies, const std::vector<uint32_t>& editedStatic);

// ---- threaded step (the CPU baseline's structure; oracle/threaded.cpp) ----
struct Pool;                                                  // pinned std::threads + the caller
Pool* poolCreate(uint32_t participants);
void poolDestroy(Pool* p);
uint32_t poolSize(const Pool* p);
void narrowPhaseThreaded(World& w, const uint32_t* pa, const uint32_t* pb, size_t n, std::vector<Manifold>& out,
                         Pool& pool);
void solveThreaded(World& w, std::vector<Manifold>& m, Pool& pool, uint32_t vIt, uint32_t pIt);

// Simulation.cpp:255-351; with a pool, the narrow phase and the solver take its threaded paths.
void step(World& w, float dt, Pool* pool = nullptr);
void postUpdate(World& w);                                    // Simulation.cpp:640-655
// Kinds: 0 position (body origin), 1 rotation, 2 velocity, 3 angular velocity, 4 mass, 5 inertia.
void flushCommands(World& w);                                 // SimulationImpl/ObjectManager.cpp:638-865

}  // namespace orc
