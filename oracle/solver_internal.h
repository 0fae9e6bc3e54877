// Oracle solver internals shared by the sequential and the threaded solve (test infrastructure
// only).
#pragma once
#include "oracle.h"

#include <vector>

namespace orc {

struct PointData { float rAn, rBn, rAt, rBt, normalMass, tangentMass, bias; };
struct Constraint {                // VelocityHotData + VelocityConstraintData + FrictionData + PositionHotData
    uint32_t bodyA, bodyB, count, manifold;
    float nx, ny;
    float imp[4];                  // jn0, jt0, jn1, jt1
    float invMassA, invMassB, invInertiaA, invInertiaB;
    PointData p[2];
    float staticFriction, dynamicFriction;
    float penetrationBase;
};

// The solver's phases over ranges of constraints (solve positions [lo, hi)), so a sequential
// solve runs each over all positions and the threaded one over each worker's pass of a wave.
struct Solver {
    World& w;
    std::vector<Manifold>& ms;
    std::vector<Constraint> cs;

    Solver(World& world, std::vector<Manifold>& manifolds) : w(world), ms(manifolds), cs(manifolds.size()) {}

    void load(uint32_t b, V2& v, float& av) const { v = { w.velX[b], w.velY[b] }; av = w.angVel[b]; }
    void store(uint32_t b, V2 v, float av, float invMass) {
        if (invMass > 0.0f) { w.velX[b] = v.x; w.velY[b] = v.y; w.angVel[b] = av; }
    }

    // Solving/BodyCollisionSolver.cpp: computeConstraintData (:229-390), order[c] = manifold at
    // solve position c (null: identity); applyWarmStarting (:416-490); solveVelocityConstraints
    // (:492-928); solvePositionConstraints (:930-980); scatterImpulses (:392-414) and
    // computeConstraintErrors (:132-183).
    void prepare(const uint32_t* order, size_t lo, size_t hi);
    void warmStart(size_t lo, size_t hi);
    void velocity(size_t lo, size_t hi);
    void position(size_t lo, size_t hi);
    void finish();
};


// A wave plan (SolvingPlan, Solving/SolvingPlanner.h:12-32): flat[c] = manifold at solve
// position c; pass (wave, worker) = positions [passStart[p], passStart[p] + passSize[p]) with
// p = wave * workers + worker.
struct WavePlan {
    uint32_t workers = 0;
    std::vector<uint32_t> flat, passStart, passSize;
};

WavePlan planWaves(const World& w, const std::vector<Manifold>& ms, uint32_t workers);

}  // namespace orc
