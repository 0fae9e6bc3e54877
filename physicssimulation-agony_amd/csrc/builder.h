// Host-side scene building for the device world: ObjectManager creation semantics
// (SimulationImpl/ObjectManager.cpp:71-389, PhysicsGeometry.cpp:14-298).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace agx {

struct Vec { float x, y; };

// One scene in dense order: [Static | Dynamic] bodies, colliders following their owner's group.
// A joint as created (ConstraintSystems/*ConstraintSystem.h CreateParams); bodies by creation slot,
// which stays with the body when later creations move it in the dense arrays.
struct JointSpec {
    uint32_t type, slotA, slotB, enableLimit;
    Vec anchorA, anchorB, axis;
    float restLength, stiffness, damping, lowerLimit, upperLimit, targetSpeed, maxTorque;
    float referenceAngle;   // prismatic: relative rotation of B to A at creation, in (-pi, pi]
};

struct SceneData {
    std::vector<float> comX, comY, velX, velY, rot, angVel, mass, invMass, inertia, invInertia, cosv, sinv;
    std::vector<float> localComX, localComY, boundRadius;
    std::vector<uint32_t> bodySlot;
    std::vector<std::vector<uint32_t>> bodyColliders;
    uint32_t staticBodies = 0;
    std::vector<uint32_t> colBody, colMaterial, colSlot, polyStart, polyCount;
    std::vector<uint8_t> colType;
    std::vector<float> localOffX, localOffY, comOffX, comOffY, localRot, wCos, wSin, p0, p1, minExtent, maxExtent;
    uint32_t staticColliders = 0;
    std::vector<Vec> verts, normals;

    uint32_t bodyCount() const { return (uint32_t)comX.size(); }
    uint32_t colliderCount() const { return (uint32_t)colBody.size(); }

    // Returns the new body's dense index or UINT32_MAX (invalid polygon outline).
    uint32_t createCircle(float px, float py, float rot, float vx, float vy, float av, float mass,
                          uint32_t material, float radius);
    uint32_t createBox(float px, float py, float rot, float vx, float vy, float av, float mass,
                       uint32_t material, float width, float height);
    uint32_t createPolygon(float px, float py, float rot, float vx, float vy, float av, float mass,
                           uint32_t material, const Vec* verts, uint32_t count);
    uint32_t createCapsule(float px, float py, float rot, float vx, float vy, float av, float mass,
                           uint32_t material, float length, float radius);
    // ObjectManager::createBody (ObjectManager.cpp:71-88): a body without colliders, its mass,
    // inertia and local centre of mass as given (negative values become 0).
    uint32_t createBody(float px, float py, float rot, float vx, float vy, float av, float mass, float inertia,
                        float comX, float comY);
    // ObjectManager::create{Circle,Box,Capsule}Collider (ObjectManager.cpp:90-181): a collider on dense
    // body `body` at local offset and rotation, into its owner's group; the body's mass properties
    // stay. Returns the collider's dense index or UINT32_MAX (no such body, unknown shape).
    uint32_t createCollider(uint32_t body, uint8_t type, float offX, float offY, float localRot, uint32_t material,
                            float p0, float p1);
    // ObjectManager::destroyBodyImmediate (ObjectManager.cpp:390-418): the body's colliders, then the
    // body, each moved to the end of its arrays through its group (swapping with the last member of
    // each group on the way) and removed; joints on the body are removed. False when there is no
    // such body.
    bool destroyBody(uint32_t body);
    uint32_t materialCount = 1;   // set by the builder owner so out-of-range materials map to 0
    uint32_t nextBodySlot = 0, nextColSlot = 0;   // creation slots of the next body and collider (never reused)
    std::vector<JointSpec> joints;
    // Appends a joint between dense bodies a and b (b ignored for a motor, type 4); false when a
    // body does not exist. The axis is normalised, (1, 0) when shorter than 1e-6.
    bool addJoint(JointSpec spec, uint32_t a, uint32_t b);

private:
    uint32_t appendBody(Vec pos, Vec vel, float rot, float av, float mass, float invMass, float inertia,
                        float invInertia, Vec localCom);
    uint32_t appendCollider(uint32_t body, Vec off, float localRot, uint32_t material, uint8_t type, float p0,
                            float p1, uint32_t ps, uint32_t pc);
    uint32_t placeCollider(uint32_t c, float minExt, float maxExt);   // returns the collider's index after the move
    void swapBodies(uint32_t a, uint32_t b);
    void swapColliders(uint32_t a, uint32_t b);
};

struct Material { float elasticity, staticFriction, dynamicFriction; };

struct Builder {
    std::vector<SceneData> scenes;
    std::vector<Material> materials{ Material{ 1.0f, 0.0f, 0.0f } };   // Types/Material.h:8-10 default
};

}  // namespace agx
