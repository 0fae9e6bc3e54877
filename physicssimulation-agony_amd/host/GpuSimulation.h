// GpuSimulation.h - C++ mirror of PS_AGONY2D::Simulation and its ObjectAPI over the agx C ABI.
//
// Header-only; links against libagx.so. Names, parameter structs and error behaviour follow the
// reference so that code written against PS_AGONY2D::Simulation ports by changing the namespace:
//   Simulation::update / getSimulationSettings / createMaterial / objects()   (Simulation.h:148-203)
//   ObjectAPI::createCircle / createBox / createPolygon / createCapsule (+ *H), getBodyData (ObjectAPI.h:34-109)
//   ObjectAPI::setBodyPosition / Rotation / Velocity / AngularVelocity / Mass / Inertia (deferred,
//   ObjectAPI.h:76-81), getBodies / getColliders / getPolygons viewers as host copies (:96-101)
//   BodyCreateParams, CircleCreateParams, BoxCreateParams, PolygonCreateParams (ObjectCreateParams.h:14-95)
//   Material                                                                  (Types/Material.h:6-11)
//   ConstraintAPI::createSpring / createRevoluteJoint / createPrismaticJoint / createSliderJoint /
//   createMotor with the systems' CreateParams (ConstraintAPI.h:31-35, ConstraintSystems/*.h:14-26)
//   CollisionEvent, collision callbacks: createCollisionCallback / destroyCollisionCallback,
//   setColliderCollisionCallback (Types/CollisionEvent.h:18-30, CollisionCallbackManager.h:14-118,
//   Simulation.cpp:152-180)
// Bad user input returns std::nullopt / false and writes a message to std::cerr, like the
// reference's ObjectManager (ObjectManager.cpp:92-97).
//
// Differences a caller must know: the scene is uploaded to the GPU at the first update() (or an
// explicit upload()). Bodies created afterwards rebuild the world around them (a cost like the
// upload); materials and constraints are created before the upload (afterwards they return
// std::nullopt). Read-outs copy device state to host vectors (the reference returns non-owning views).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/agx.h"

namespace PS_AGONY2D_GPU {

using Real = float;
using ObjectDenseIndex = uint32_t;
using MaterialIndex = uint32_t;
using Handle = uint32_t;   // the dense index at creation; bodies are never destroyed here

struct Vec2 {
    Real x = 0, y = 0;
};

// Types/Material.h:6-11 (default elasticity 1).
struct Material {
    Real elasticity = 1.0f;
    Real staticFriction = 0.0f;
    Real dynamicFriction = 0.0f;
};

// ObjectCreateParams.h:14-24. Mass 0 makes a static body; inertia and centre of mass come from the
// shape (a caller-supplied inertia is not used on this path).
struct BodyCreateParams {
    Vec2 position{};
    Vec2 velocity{};
    Real rotation = 0;
    Real angularVelocity = 0;
    Real mass = 0;
    Real inertia = 0;
    MaterialIndex materialIndex = 0;
};
// ObjectCreateParams.h: a body on its own (createBody: mass, inertia and local centre of mass as
// given) and colliders attached to it at a local offset and rotation (create*Collider).
struct BodyOnlyCreateParams {
    Vec2 position{};
    Vec2 velocity{};
    Real rotation = 0;
    Real angularVelocity = 0;
    Real mass = 0;
    Real inertia = 0;
    Vec2 centerOfMass{};
};
struct ColliderCreateParams { ObjectDenseIndex body = 0; Vec2 localOffset{}; Real localRotation = 0; MaterialIndex materialIndex = 0; };
struct CircleColliderCreateParams { ColliderCreateParams base; Real radius = 0; };
struct BoxColliderCreateParams { ColliderCreateParams base; Vec2 size{}; };
struct CapsuleColliderCreateParams { ColliderCreateParams base; Real length = 0; Real radius = 0; };
struct CircleCreateParams { BodyCreateParams base; Real radius = 0; };
struct BoxCreateParams { BodyCreateParams base; Vec2 size{}; };
struct PolygonCreateParams { BodyCreateParams base; std::vector<Vec2> localVertices; };
// ObjectCreateParams.h:97-104: core length along the local +x axis and radius; when both world
// end points are given they set the length, position and rotation (ObjectManager.cpp:351-363).
struct CapsuleCreateParams {
    BodyCreateParams base;
    Real length = 0;
    Real radius = 0;
    std::optional<Vec2> start{}, end{};
};

// BodySoA fields of one body (SoA/BodySoA.h:14-36).
struct BodyData {
    Vec2 centerOfMass{}, velocity{}, localCenterOfMass{};
    Real rotation = 0, angularVelocity = 0, mass = 0, inertia = 0, invMass = 0, invInertia = 0;
};

// Host copies of the body arrays, indexed by dense body index.
struct BodyArrays {
    std::vector<float> comX, comY, rotation, velX, velY, angularVelocity, invMass;
};

// Host copies of the collider arrays (SoA/ColliderSoA.h:10-32), indexed by dense collider index.
struct ColliderArrays {
    std::vector<float> worldX, worldY, worldRotation, aabbMinX, aabbMinY, aabbMaxX, aabbMaxY;
    std::vector<uint32_t> body, shapeType, material;
};

// Polygon outlines (SoA/Shapes/PolygonSoA.h:14-52): the local vertices and outward edge normals of
// collider c are vertices[start[c] .. start[c] + count[c]) (count 0 for circles and boxes).
struct PolygonArrays {
    std::vector<Vec2> vertices, normals;
    std::vector<uint32_t> start, count;
};

class Simulation;

// Types/CollisionEvent.h:11-30, with dense indices for handles. Handlers run after update()/step()
// returns, for every physics step of that call in step order (the reference runs them at the end of
// each physics step; commands they queue take effect at the next update either way).
using CollisionCallbackIndex = uint16_t;
struct CollisionEvent {
    ObjectDenseIndex self, other;
    ObjectDenseIndex selfBody, otherBody;
    Vec2 point;
    Vec2 normal;        // from self towards other
    Real impulse;
    CollisionCallbackIndex callbackId;
    uint32_t step;      // physics step of the call
};
using CollisionCallback = void (*)(const CollisionEvent& event, Simulation& simulation, void* userData);

// ConstraintSystems' CreateParams (SimulationImpl/ConstraintSystems/*ConstraintSystem.h:14-26), with
// body indices in place of handles.
struct SpringCreateParams { Handle bodyHandleA, bodyHandleB; Vec2 localAnchorA, localAnchorB; Real restLength, stiffness, damping; };
struct RevoluteJointCreateParams { Handle bodyHandleA, bodyHandleB; Vec2 localAnchorA, localAnchorB; Real stiffness, damping; };
struct PrismaticJointCreateParams {
    Handle bodyHandleA, bodyHandleB;
    Vec2 localAnchorA, localAnchorB, localAxisA;
    Real stiffness, damping;
    Real lowerLimit = 0, upperLimit = 0;
    bool enableLimit = false;
};
using SliderJointCreateParams = PrismaticJointCreateParams;
struct MotorCreateParams { Handle bodyHandle; Real targetSpeed, maxTorque; };

// ConstraintAPI creation (SimulationAPIs/ConstraintAPI.h:31-35): before the upload, like bodies;
// std::nullopt and a message for a missing body. Returns the joint's index among the scene's joints.
class ConstraintAPI {
public:
    explicit ConstraintAPI(Simulation& s) : sim(s) {}
    ConstraintAPI(const ConstraintAPI&) = delete;
    ConstraintAPI& operator=(const ConstraintAPI&) = delete;
    std::optional<uint32_t> createSpring(const SpringCreateParams& p) {
        agx_joint_desc d = pair(AGX_JOINT_SPRING, p.bodyHandleA, p.bodyHandleB, p.localAnchorA, p.localAnchorB);
        d.rest_length = p.restLength; d.stiffness = p.stiffness; d.damping = p.damping;
        return add(d);
    }
    std::optional<uint32_t> createRevoluteJoint(const RevoluteJointCreateParams& p) {
        agx_joint_desc d = pair(AGX_JOINT_REVOLUTE, p.bodyHandleA, p.bodyHandleB, p.localAnchorA, p.localAnchorB);
        d.stiffness = p.stiffness; d.damping = p.damping;
        return add(d);
    }
    std::optional<uint32_t> createPrismaticJoint(const PrismaticJointCreateParams& p) { return axisJoint(AGX_JOINT_PRISMATIC, p); }
    std::optional<uint32_t> createSliderJoint(const SliderJointCreateParams& p) { return axisJoint(AGX_JOINT_SLIDER, p); }
    std::optional<uint32_t> createMotor(const MotorCreateParams& p) {
        agx_joint_desc d{};
        d.type = AGX_JOINT_MOTOR; d.body_a = p.bodyHandle; d.target_speed = p.targetSpeed; d.max_torque = p.maxTorque;
        return add(d);
    }

private:
    static agx_joint_desc pair(uint32_t type, Handle a, Handle b, Vec2 anchorA, Vec2 anchorB) {
        agx_joint_desc d{};
        d.type = type; d.body_a = a; d.body_b = b;
        d.anchor_a_x = anchorA.x; d.anchor_a_y = anchorA.y; d.anchor_b_x = anchorB.x; d.anchor_b_y = anchorB.y;
        d.axis_x = 1.0f;
        return d;
    }
    std::optional<uint32_t> axisJoint(uint32_t type, const PrismaticJointCreateParams& p) {
        agx_joint_desc d = pair(type, p.bodyHandleA, p.bodyHandleB, p.localAnchorA, p.localAnchorB);
        d.axis_x = p.localAxisA.x; d.axis_y = p.localAxisA.y;
        d.stiffness = p.stiffness; d.damping = p.damping;
        d.lower_limit = p.lowerLimit; d.upper_limit = p.upperLimit; d.enable_limit = p.enableLimit ? 1u : 0u;
        return add(d);
    }
    std::optional<uint32_t> add(const agx_joint_desc& d);
    Simulation& sim;
};

class ObjectAPI {
public:
    explicit ObjectAPI(Simulation& s) : sim(s) {}
    ObjectAPI(const ObjectAPI&) = delete;
    ObjectAPI& operator=(const ObjectAPI&) = delete;

    std::optional<ObjectDenseIndex> createCircle(const CircleCreateParams& p);
    std::optional<ObjectDenseIndex> createBox(const BoxCreateParams& p);
    std::optional<ObjectDenseIndex> createPolygon(const PolygonCreateParams& p);
    std::optional<ObjectDenseIndex> createCapsule(const CapsuleCreateParams& p);
    // ObjectAPI::createBody / createCircleCollider / createBoxCollider / createCapsuleCollider
    // (ObjectAPI.h:28-32): before the upload; colliders of one body never collide with each other
    std::optional<ObjectDenseIndex> createBody(const BodyOnlyCreateParams& p);
    // ObjectAPI::destroyBody (ObjectAPI.h:70): with its colliders and joints, before or after the
    // upload (afterwards the world is rebuilt); false for a body that does not exist
    bool destroyBody(ObjectDenseIndex body);
    std::optional<ObjectDenseIndex> createCircleCollider(const CircleColliderCreateParams& p) {
        return collider(p.base, AGX_SHAPE_CIRCLE, p.radius, 0.0f);
    }
    std::optional<ObjectDenseIndex> createBoxCollider(const BoxColliderCreateParams& p) {
        return collider(p.base, AGX_SHAPE_BOX, p.size.x, p.size.y);
    }
    std::optional<ObjectDenseIndex> createCapsuleCollider(const CapsuleColliderCreateParams& p) {
        return collider(p.base, AGX_SHAPE_CAPSULE, p.length, p.radius);
    }
    std::optional<Handle> createCircleH(const CircleCreateParams& p) { return createCircle(p); }
    std::optional<Handle> createBoxH(const BoxCreateParams& p) { return createBox(p); }
    std::optional<Handle> createPolygonH(const PolygonCreateParams& p) { return createPolygon(p); }
    std::optional<Handle> createCapsuleH(const CapsuleCreateParams& p) { return createCapsule(p); }

    std::optional<BodyData> getBodyData(ObjectDenseIndex index) const;
    BodyArrays getBodies() const;
    ColliderArrays getColliders() const;
    PolygonArrays getPolygons() const;

    // Deferred modifications, applied at the next update() (ObjectManager::flushCommands). A
    // mass change must keep the body dynamic (mass > 0); false and a message otherwise.
    bool setBodyPosition(Handle body, Vec2 position) { return queue(AGX_SET_POSITION, body, position.x, position.y); }
    bool setBodyRotation(Handle body, Real rotation) { return queue(AGX_SET_ROTATION, body, rotation, 0); }
    bool setBodyVelocity(Handle body, Vec2 velocity) { return queue(AGX_SET_VELOCITY, body, velocity.x, velocity.y); }
    bool setBodyAngularVelocity(Handle body, Real w) { return queue(AGX_SET_ANGULAR_VELOCITY, body, w, 0); }
    bool setBodyMass(Handle body, Real mass) { return queue(AGX_SET_MASS, body, mass, 0); }
    bool setBodyInertia(Handle body, Real inertia) { return queue(AGX_SET_INERTIA, body, inertia, 0); }

private:
    friend class Simulation;
    struct Command { uint32_t kind; Handle body; Real a, b; };
    bool queue(uint32_t kind, Handle body, Real a, Real b);
    std::vector<Command> pending;   // commands given before the scene is on the GPU
    std::optional<ObjectDenseIndex> collider(const ColliderCreateParams& c, uint8_t type, float p0, float p1);
    std::optional<ObjectDenseIndex> add(uint8_t type, const BodyCreateParams& b, float p0, float p1,
                                        const std::vector<Vec2>* verts);
    Simulation& sim;
};

class Simulation {
public:
    // Simulation.h:50-76 defaults (6 velocity / 3 position iterations, CCD on, 1/300 s interval).
    using SimulationSettings = agx_settings;

    explicit Simulation(uint32_t scenes = 1) : api(*this), constraintApi(*this) {
        agx_settings_default(&settings);
        if (agx_builder_create(scenes, &builder) != AGX_OK) fail("agx_builder_create", agx_last_error());
    }
    ~Simulation() {
        if (world) agx_world_destroy(world);
        if (builder) agx_builder_destroy(builder);
    }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    // Settings take effect at the next update().
    SimulationSettings& getSimulationSettings() noexcept { return settings; }

    std::optional<MaterialIndex> createMaterial(const Material& m) {
        if (world) { fail("createMaterial", "the scene is already on the GPU"); return std::nullopt; }
        uint32_t idx = 0;
        if (agx_builder_add_material(builder, m.elasticity, m.staticFriction, m.dynamicFriction, &idx) != AGX_OK) {
            fail("createMaterial", agx_last_error());
            return std::nullopt;
        }
        return idx;
    }

    ObjectAPI& objects() noexcept { return api; }
    ConstraintAPI& constraints() noexcept { return constraintApi; }

    // CollisionCallbackManager::create / destroy (CollisionCallbackManager.h:38-80): slot 0 is the
    // permanent "none", destroyed slots are reused last-in first-out.
    std::optional<CollisionCallbackIndex> createCollisionCallback(CollisionCallback function, void* userData) {
        if (!function) return std::nullopt;
        CollisionCallbackIndex index;
        if (!freeCallbacks.empty()) {
            index = freeCallbacks.back();
            freeCallbacks.pop_back();
            callbacks[index] = { function, userData, true };
        } else {
            if (callbacks.size() >= AGX_MAX_CALLBACK_ID) return std::nullopt;
            index = (CollisionCallbackIndex)callbacks.size();
            callbacks.push_back({ function, userData, true });
        }
        return index;
    }
    bool destroyCollisionCallback(CollisionCallbackIndex index) {
        if (index == 0 || !callbackAlive(index)) return false;
        callbacks[index] = {};
        freeCallbacks.push_back(index);
        return true;
    }
    bool callbackAlive(CollisionCallbackIndex index) const noexcept {
        return index < callbacks.size() && callbacks[index].alive;
    }
    // Simulation::setColliderCollisionCallback (Simulation.cpp:152-172): index 0 clears; otherwise
    // the callback must be alive and the threshold >= 0.
    bool setColliderCollisionCallback(ObjectDenseIndex collider, CollisionCallbackIndex index, Real impulseThreshold) {
        if (collider >= counts().colliders) return false;
        if (index != 0 && (!callbackAlive(index) || !(impulseThreshold >= Real(0)))) return false;
        if (!world) { pendingCallbacks.push_back({ collider, index, impulseThreshold }); return true; }
        if (agx_world_set_collider_callback(world, collider, index, impulseThreshold) != AGX_OK) {
            fail("setColliderCollisionCallback", agx_world_last_error(world));
            return false;
        }
        return true;
    }

    // Uploads the scene (Simulation::refreshSpatialData runs here). Called by the first update().
    bool upload() {
        if (world) return true;
        if (agx_world_create(builder, &settings, &world) != AGX_OK) {
            fail("upload", agx_last_error());
            world = nullptr;
            return false;
        }
        for (const ObjectAPI::Command& c : api.pending)   // commands given before the upload, in order
            if (agx_world_queue_body(world, c.kind, c.body, c.a, c.b) != AGX_OK) fail("set", agx_world_last_error(world));
        api.pending.clear();
        for (const PendingCallback& c : pendingCallbacks)
            if (agx_world_set_collider_callback(world, c.collider, c.index, c.threshold) != AGX_OK)
                fail("setColliderCollisionCallback", agx_world_last_error(world));
        pendingCallbacks.clear();
        return true;
    }

    // Simulation::dispatchCollisionEvents (Simulation.cpp:174-180) for the last update()/step():
    // each event goes to its callback if that is still alive; a handler may create or destroy
    // callbacks (the entry is copied before the call).
    void dispatchCollisionEvents() {
        if (!world) return;
        uint64_t count = 0, dropped = 0;
        if (agx_world_read_events(world, nullptr, 0, &count, &dropped) != AGX_OK && count == 0) return;
        std::vector<agx_collision_event> ev(count);
        if (count && agx_world_read_events(world, ev.data(), count, &count, &dropped) != AGX_OK) return;
        if (dropped) fail("dispatchCollisionEvents", "events beyond settings.event_capacity were dropped");
        for (const agx_collision_event& e : ev) {
            const CollisionCallbackIndex id = (CollisionCallbackIndex)e.callback_id;
            if (id == 0 || !callbackAlive(id)) continue;
            const CallbackEntry entry = callbacks[id];
            const CollisionEvent out{ e.self, e.other, e.self_body, e.other_body, { e.point_x, e.point_y },
                                      { e.normal_x, e.normal_y }, e.impulse, id, e.step };
            entry.function(out, *this, entry.userData);
        }
    }

    // Simulation::update (Simulation.cpp:32-74): fixed-interval sub-steps, then the pose refresh.
    uint32_t update(Real dt) {
        if (!upload()) return 0;
        agx_world_set_settings(world, &settings);
        uint32_t steps = 0;
        if (agx_world_update(world, dt, &steps) != AGX_OK) fail("update", agx_world_last_error(world));
        if (steps) dispatchCollisionEvents();
        return steps;
    }

    // Exactly n physics steps of dt (a clock that yields n steps), then the pose refresh.
    bool step(Real dt, uint32_t n) {
        if (!upload()) return false;
        agx_world_set_settings(world, &settings);
        if (agx_world_step(world, dt, n) != AGX_OK || agx_world_sync(world) != AGX_OK) {
            fail("step", agx_world_last_error(world));
            return false;
        }
        if (n) dispatchCollisionEvents();
        return true;
    }

    // Simulation::getBodyCollisionData (Simulation.h:186): the last step's solved manifolds.
    std::vector<agx_manifold> getBodyCollisionData() const {
        std::vector<agx_manifold> out;
        if (!world) return out;
        uint64_t count = 0;
        agx_world_read_manifolds(world, 0, nullptr, 0, &count);
        out.resize(count);
        if (count && agx_world_read_manifolds(world, 0, out.data(), count, &count) != AGX_OK) out.clear();
        return out;
    }

    // Simulation::islands (Simulation.h:203): island id per body (UINT32_MAX for static bodies).
    std::pair<std::vector<uint32_t>, uint32_t> islands() const {
        std::vector<uint32_t> ids;
        uint32_t n = 0;
        if (!world) return { ids, 0 };
        ids.resize(counts().bodies);
        agx_world_read_islands(world, ids.data(), (uint32_t)ids.size(), &n);
        return { ids, n };
    }

    agx_stats getStats() const {
        agx_stats st{};
        if (world) agx_world_read_stats(world, &st);
        return st;
    }

    agx_counts counts() const {
        agx_counts c{};
        if (world) agx_world_counts(world, &c);
        else agx_builder_scene_counts(builder, 0, &c);
        return c;
    }

    agx_world* nativeWorld() noexcept { return world; }
    agx_builder* nativeBuilder() noexcept { return builder; }

    static void fail(const char* what, const char* why) {
        std::cerr << "PS_AGONY2D_GPU::" << what << ": " << (why ? why : "error") << '\n';
    }

private:
    friend class ObjectAPI;
    struct CallbackEntry { CollisionCallback function = nullptr; void* userData = nullptr; bool alive = false; };
    struct PendingCallback { ObjectDenseIndex collider; CollisionCallbackIndex index; Real threshold; };
    std::vector<CallbackEntry> callbacks{ CallbackEntry{ nullptr, nullptr, false } };   // slot 0: none
    std::vector<CollisionCallbackIndex> freeCallbacks;
    std::vector<PendingCallback> pendingCallbacks;
    agx_settings settings{};
    agx_builder* builder = nullptr;
    agx_world* world = nullptr;
    ObjectAPI api;
    ConstraintAPI constraintApi;
    friend class ConstraintAPI;
};

inline std::optional<uint32_t> ConstraintAPI::add(const agx_joint_desc& d) {
    if (sim.world) { Simulation::fail("createConstraint", "the scene is already on the GPU"); return std::nullopt; }
    uint32_t index = 0;
    if (agx_builder_add_joints(sim.builder, 0, 1, &d, &index) != AGX_OK) {
        Simulation::fail("createConstraint", agx_last_error());
        return std::nullopt;
    }
    return index;
}

inline std::optional<ObjectDenseIndex> ObjectAPI::add(uint8_t type, const BodyCreateParams& b, float p0, float p1,
                                                      const std::vector<Vec2>* verts) {
    const uint32_t vcount = verts ? (uint32_t)verts->size() : 0u, vstart = 0;
    std::vector<float> flat;
    if (verts) for (const Vec2& v : *verts) { flat.push_back(v.x); flat.push_back(v.y); }
    uint32_t out = UINT32_MAX;
    // before the upload into the builder; afterwards into the world, which is rebuilt around the
    // new body with everything it carried (agx_world_add_bodies)
    const int rc = sim.world
        ? agx_world_add_bodies(sim.world, 0, 1, &type, &b.position.x, &b.position.y, &b.rotation, &b.velocity.x,
                               &b.velocity.y, &b.angularVelocity, &b.mass, &b.materialIndex, &p0, &p1, &vstart, &vcount,
                               flat.empty() ? nullptr : flat.data(), &out)
        : agx_builder_add_bodies(sim.builder, 0, 1, &type, &b.position.x, &b.position.y, &b.rotation, &b.velocity.x,
                                 &b.velocity.y, &b.angularVelocity, &b.mass, &b.materialIndex, &p0, &p1, &vstart, &vcount,
                                 flat.empty() ? nullptr : flat.data(), &out);
    if (rc < 0 || out == UINT32_MAX) {
        Simulation::fail("create", rc < 0 ? (sim.world ? agx_world_last_error(sim.world) : agx_last_error())
                                          : "invalid or non-convex polygon");
        return std::nullopt;
    }
    return out;
}

inline bool ObjectAPI::destroyBody(ObjectDenseIndex body) {
    const int rc = sim.world ? agx_world_remove_bodies(sim.world, 1, &body) : agx_builder_remove_body(sim.builder, 0, body);
    if (rc != AGX_OK) Simulation::fail("destroyBody", sim.world ? agx_world_last_error(sim.world) : agx_last_error());
    return rc == AGX_OK;
}

inline std::optional<ObjectDenseIndex> ObjectAPI::createBody(const BodyOnlyCreateParams& p) {
    if (sim.world) { Simulation::fail("createBody", "the scene is already on the GPU"); return std::nullopt; }
    uint32_t out = UINT32_MAX;
    if (agx_builder_create_body(sim.builder, 0, p.position.x, p.position.y, p.rotation, p.velocity.x, p.velocity.y,
                                p.angularVelocity, p.mass, p.inertia, p.centerOfMass.x, p.centerOfMass.y, &out) != AGX_OK)
        return std::nullopt;
    return out;
}

inline std::optional<ObjectDenseIndex> ObjectAPI::collider(const ColliderCreateParams& c, uint8_t type, float p0, float p1) {
    if (sim.world) { Simulation::fail("createCollider", "the scene is already on the GPU"); return std::nullopt; }
    uint32_t out = UINT32_MAX;
    if (agx_builder_add_collider(sim.builder, 0, c.body, type, c.localOffset.x, c.localOffset.y, c.localRotation,
                                 c.materialIndex, p0, p1, &out) != AGX_OK) {
        Simulation::fail("createCollider", agx_last_error());
        return std::nullopt;
    }
    return out;
}

inline std::optional<ObjectDenseIndex> ObjectAPI::createCircle(const CircleCreateParams& p) {
    return add(AGX_SHAPE_CIRCLE, p.base, p.radius, 0.0f, nullptr);
}
inline std::optional<ObjectDenseIndex> ObjectAPI::createBox(const BoxCreateParams& p) {
    return add(AGX_SHAPE_BOX, p.base, p.size.x, p.size.y, nullptr);
}
inline std::optional<ObjectDenseIndex> ObjectAPI::createCapsule(const CapsuleCreateParams& p) {
    if (p.start && p.end) {
        const Vec2 d{ p.end->x - p.start->x, p.end->y - p.start->y };
        CapsuleCreateParams q = p;
        q.length = std::sqrt(d.x * d.x + d.y * d.y);
        q.base.position = { (p.start->x + p.end->x) * 0.5f, (p.start->y + p.end->y) * 0.5f };
        q.base.rotation = std::atan2(d.y, d.x);
        q.start.reset();
        q.end.reset();
        return createCapsule(q);
    }
    return add(AGX_SHAPE_CAPSULE, p.base, p.length, p.radius, nullptr);
}

inline std::optional<ObjectDenseIndex> ObjectAPI::createPolygon(const PolygonCreateParams& p) {
    return add(AGX_SHAPE_POLYGON, p.base, 0.0f, 0.0f, &p.localVertices);
}

inline BodyArrays ObjectAPI::getBodies() const {
    BodyArrays a;
    const uint32_t n = sim.counts().bodies;
    auto read = [&](uint32_t field, std::vector<float>& v) {
        v.resize(n);
        if (sim.world) agx_world_read_body_f32(sim.world, field, v.data(), n);
        else agx_builder_read_body_f32(sim.builder, 0, field, v.data(), n);
    };
    read(AGX_BODY_COM_X, a.comX); read(AGX_BODY_COM_Y, a.comY); read(AGX_BODY_ROTATION, a.rotation);
    read(AGX_BODY_VEL_X, a.velX); read(AGX_BODY_VEL_Y, a.velY); read(AGX_BODY_ANG_VEL, a.angularVelocity);
    read(AGX_BODY_INV_MASS, a.invMass);
    return a;
}

inline bool ObjectAPI::queue(uint32_t kind, Handle body, Real a, Real b) {
    if (!sim.world) {
        const agx_counts c = sim.counts();
        if (body >= c.bodies) { Simulation::fail("set", "no such body"); return false; }
        if (kind == AGX_SET_MASS && !(a > 0)) { Simulation::fail("set", "mass changes keep a body dynamic"); return false; }
        pending.push_back({ kind, body, a, b });
        return true;
    }
    if (agx_world_queue_body(sim.world, kind, body, a, b) != AGX_OK) {
        Simulation::fail("set", agx_world_last_error(sim.world));
        return false;
    }
    return true;
}

inline ColliderArrays ObjectAPI::getColliders() const {
    ColliderArrays a;
    const uint32_t n = sim.counts().colliders;
    auto readF = [&](uint32_t field, std::vector<float>& v) {
        v.resize(n);
        if (sim.world) agx_world_read_collider_f32(sim.world, field, v.data(), n);
        else agx_builder_read_collider(sim.builder, 0, field, 0, v.data(), n);
    };
    auto readU = [&](uint32_t field, std::vector<uint32_t>& v) {
        v.resize(n);
        if (sim.world) agx_world_read_collider_u32(sim.world, field, v.data(), n);
        else agx_builder_read_collider(sim.builder, 0, field, 1, v.data(), n);
    };
    readF(AGX_COL_WORLD_X, a.worldX); readF(AGX_COL_WORLD_Y, a.worldY); readF(AGX_COL_WORLD_ROT, a.worldRotation);
    readF(AGX_COL_AABB_MIN_X, a.aabbMinX); readF(AGX_COL_AABB_MIN_Y, a.aabbMinY);
    readF(AGX_COL_AABB_MAX_X, a.aabbMaxX); readF(AGX_COL_AABB_MAX_Y, a.aabbMaxY);
    readU(AGX_COL_BODY, a.body); readU(AGX_COL_TYPE, a.shapeType); readU(AGX_COL_MATERIAL, a.material);
    return a;
}

inline PolygonArrays ObjectAPI::getPolygons() const {
    PolygonArrays p;
    const agx_counts c = sim.counts();
    std::vector<float> v(2 * (size_t)c.vertices), nrm(2 * (size_t)c.vertices);
    if (c.vertices) agx_builder_read_vertices(sim.builder, 0, v.data(), nrm.data(), c.vertices);
    for (uint32_t i = 0; i < c.vertices; i++) {
        p.vertices.push_back({ v[2 * i], v[2 * i + 1] });
        p.normals.push_back({ nrm[2 * i], nrm[2 * i + 1] });
    }
    p.start.resize(c.colliders);
    p.count.resize(c.colliders);
    agx_builder_read_collider(sim.builder, 0, AGX_COL_POLY_START, 1, p.start.data(), c.colliders);
    agx_builder_read_collider(sim.builder, 0, AGX_COL_POLY_COUNT, 1, p.count.data(), c.colliders);
    return p;
}

inline std::optional<BodyData> ObjectAPI::getBodyData(ObjectDenseIndex index) const {
    const uint32_t n = sim.counts().bodies;
    if (index >= n) return std::nullopt;
    auto field = [&](uint32_t f) {
        std::vector<float> v(n);
        if (sim.world) agx_world_read_body_f32(sim.world, f, v.data(), n);
        else agx_builder_read_body_f32(sim.builder, 0, f, v.data(), n);
        return v[index];
    };
    BodyData d;
    d.centerOfMass = { field(AGX_BODY_COM_X), field(AGX_BODY_COM_Y) };
    d.velocity = { field(AGX_BODY_VEL_X), field(AGX_BODY_VEL_Y) };
    d.localCenterOfMass = { field(AGX_BODY_LOCAL_COM_X), field(AGX_BODY_LOCAL_COM_Y) };
    d.rotation = field(AGX_BODY_ROTATION);
    d.angularVelocity = field(AGX_BODY_ANG_VEL);
    d.mass = field(AGX_BODY_MASS);
    d.inertia = field(AGX_BODY_INERTIA);
    d.invMass = field(AGX_BODY_INV_MASS);
    d.invInertia = field(AGX_BODY_INV_INERTIA);
    return d;
}

}  // namespace PS_AGONY2D_GPU
