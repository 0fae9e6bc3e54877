/* agx.h - C ABI of the MI355X-native PS_AGONY2D collide-and-solve path.
 *
 * The reference has no FFI: its boundary is the C++ class PS_AGONY2D::Simulation
 * (PhysicsSimulation-Agony/PS_AGONY2D/Simulation.h:36-255) and its ObjectAPI facade
 * (SimulationAPIs/ObjectAPI.h:14-127). This header is the thin C layer a host binding
 * calls instead; the C++ mirror of Simulation/ObjectAPI lives in
 * physicssimulation-agony_amd/host/GpuSimulation.h and calls only these functions.
 *
 * Conventions: every function returns 0 (AGX_OK) or a negative AGX_ERR_* code and never
 * throws. Callers own host buffers; a world owns its device buffers. All worlds and networks
 * on one device share one HIP stream, so their device work runs in issue order (the solver's
 * dataflow kernels must not run beside another world's). One world must be externally
 * synchronised; distinct worlds may be used from distinct threads, and waiting on one world
 * also waits for the others' queued work. agx_world_last_error() describes the last failure on a world,
 * agx_last_error() the last failure of a call without a world.
 */
#ifndef AGX_H
#define AGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AGX_OK 0
#define AGX_ERR_INVALID_ARGUMENT (-1)
#define AGX_ERR_CAPACITY (-2)
#define AGX_ERR_HIP (-3)
#define AGX_ERR_STATE (-4)
#define AGX_ERR_NO_DEVICE (-5)

/* Shape types: Types/Types.h:35-42 (ColliderShapeType). */
#define AGX_SHAPE_CIRCLE 0
#define AGX_SHAPE_BOX 1
#define AGX_SHAPE_POLYGON 2
#define AGX_SHAPE_CAPSULE 3

typedef struct agx_builder agx_builder;
typedef struct agx_world agx_world;
typedef struct agx_mlp agx_mlp;

/* Simulation::SimulationSettings (Simulation.h:50-76), FixedClock::Settings
 * (SimulationImpl/FixedClock.h:12-17), Integrator::Settings (Integration/Integrator.h:14-19)
 * and ContinuousCollisionDetector::Settings (CCD/ContinuousCollisionDetector.h:29-33). */
typedef struct agx_settings {
    float gravity_x, gravity_y;            /* default (0, -9.81) */
    uint32_t velocity_iterations;          /* default 6 */
    uint32_t position_iterations;          /* default 3 */
    uint32_t ccd_enabled;                  /* default 1 */
    float ccd_fast_factor;                 /* default 1.0 */
    float ccd_core_fraction;               /* default 0.5 */
    float update_interval;                 /* default 1/300 s */
    float max_delta_per_update;            /* default 1/20 s */
    float time_scale;                      /* default 1 */
    uint32_t track_errors;                 /* computeConstraintErrors each step (Simulation.h:75) */
    uint32_t track_residual;               /* relative impulse change after every velocity sweep */
    float residual_target;                 /* > 0: stop velocity sweeps once residual <= target */
    uint32_t collide_only;                 /* 1: integrate + collide, skip the solver (broad/narrow timing) */
    uint32_t solve_per_scene;              /* 1 (default): each scene solved by one workgroup with its bodies
                                              in LDS, stopping at its own residual, whenever every scene fits
                                              (about 6,500 bodies); 0: one dataflow over all scenes (one
                                              residual for the batch), also used for larger scenes. Fixed
                                              sweep counts give the same bits either way. */
    uint32_t collision_events;             /* default 1: record collision events (collisionEventsEnabled,
                                              Simulation.h:60) while some collider has a callback id */
    uint32_t event_capacity;               /* events kept per update/step call, default 65,536; more are
                                              counted and dropped */
    float sleep_speed_threshold;           /* default 0.1 (sleepSpeedThreshold, Simulation.h:70-72) */
    float sleep_time_threshold;            /* default +infinity here: sleeping off, as in every benchmark
                                              scene; the reference's default is 0.5 s (Simulation.h:73) */
    uint32_t joint_iterations[5];          /* velocity iterations per joint type (AGX_JOINT_*), default 6 each
                                              (IConstraintSystem::iterations, ConstraintSystemBase.h:34) */
} agx_settings;

/* One manifold record (64 bytes): NarrowPhaseCollisionDetector.h:26-61 with the collider
 * handles replaced by dense collider indices of the canonical A/B; impulses are the
 * persistent accumulators (jn, jt) of the two contact points. */
typedef struct agx_manifold {
    uint32_t a, b;
    float normal_x, normal_y;
    float depth;
    uint32_t count;
    float p0_x, p0_y, p1_x, p1_y;
    uint32_t id0, id1;
    float jn0, jt0, jn1, jt1;
} agx_manifold;

/* Bits of agx_stats.overflow. */
#define AGX_OVERFLOW_PAIRS 1u      /* pair buffer too small */
#define AGX_OVERFLOW_STACK 2u      /* BVH traversal stack too small */
#define AGX_OVERFLOW_COLOURING 4u  /* a constraint found no free first-fit colour (the solve is still exact) */
#define AGX_OVERFLOW_HASH 8u       /* warm-start hash table full */
#define AGX_OVERFLOW_STALL 16u     /* a dataflow kernel's wait hit its 2 s limit */
#define AGX_OVERFLOW_ITEMS 32u     /* more constraints than the dataflow kernels hold */
#define AGX_OVERFLOW_LDS 64u       /* a scene larger than its per-scene LDS region was not solved */

/* Most velocity sweeps of one step (velocity_iterations at most this). */
#define AGX_MAX_SWEEPS 256

/* Per-step counters of the last step. */
typedef struct agx_stats {
    uint64_t pairs;                /* broad-phase pairs */
    uint64_t manifolds;            /* narrow-phase manifolds */
    uint64_t contacts;             /* contact points */
    uint32_t colours;              /* solver colours */
    uint32_t islands;              /* islands of dynamic bodies */
    uint32_t ccd_hits;             /* bodies moved back by CCD */
    uint32_t velocity_sweeps;      /* velocity sweeps actually run (adaptive mode may stop early) */
    uint32_t overflow;             /* AGX_OVERFLOW_* bits; nonzero except COLOURING: results are not valid */
    uint32_t learned_warm_start;   /* 1: the step's warm start came from the network */
    float velocity_error;          /* computeConstraintErrors().x */
    float position_error;          /* computeConstraintErrors().y */
    float residual[AGX_MAX_SWEEPS]; /* relative impulse change after velocity sweep k */
    uint64_t solved_steps;         /* steps solved since the world was created */
    uint64_t total_velocity_sweeps; /* velocity sweeps over those steps (per step: the most any scene ran) */
    uint64_t total_body_sweeps;    /* over those steps: dynamic bodies x the velocity sweeps of their scene */
    float mean_velocity_sweeps;    /* last step: velocity sweeps per dynamic body (scenes may stop apart) */
    uint32_t solved_per_scene;     /* 1: the last step was solved scene by scene */
    uint32_t converged_scenes;     /* last step: scenes whose velocity sweeps met residual_target */
    uint64_t converged_bodies;     /* last step: dynamic bodies of those scenes */
    uint64_t total_converged_bodies; /* over the solved steps: the sum of converged_bodies */
} agx_stats;

/* World sizes. */
typedef struct agx_counts {
    uint32_t scenes, bodies, static_bodies, colliders, static_colliders, vertices, materials, reserved;
} agx_counts;

/* Body fields readable/writable through agx_world_{read,write}_body_f32
 * (SoA/BodySoA.h:14-36). */
enum {
    AGX_BODY_COM_X = 0, AGX_BODY_COM_Y, AGX_BODY_ROTATION, AGX_BODY_VEL_X, AGX_BODY_VEL_Y,
    AGX_BODY_ANG_VEL, AGX_BODY_COS, AGX_BODY_SIN, AGX_BODY_INV_MASS, AGX_BODY_INV_INERTIA,
    AGX_BODY_MASS, AGX_BODY_INERTIA, AGX_BODY_LOCAL_COM_X, AGX_BODY_LOCAL_COM_Y,
    AGX_BODY_FIELD_COUNT
};
/* Collider fields readable through agx_world_read_collider_f32 (SoA/ColliderSoA.h:10-32). */
enum {
    AGX_COL_WORLD_X = 0, AGX_COL_WORLD_Y, AGX_COL_WORLD_ROT, AGX_COL_COS, AGX_COL_SIN,
    AGX_COL_AABB_MIN_X, AGX_COL_AABB_MIN_Y, AGX_COL_AABB_MAX_X, AGX_COL_AABB_MAX_Y,
    AGX_COL_PREV_X, AGX_COL_PREV_Y, AGX_COL_COM_OFF_X, AGX_COL_COM_OFF_Y, AGX_COL_LOCAL_ROT,
    AGX_COL_P0, AGX_COL_P1, AGX_COL_MIN_EXTENT, AGX_COL_MAX_EXTENT,
    AGX_COL_FIELD_COUNT
};
/* Collider integer fields readable through agx_world_read_collider_u32. */
enum { AGX_COL_BODY = 0, AGX_COL_TYPE, AGX_COL_MATERIAL, AGX_COL_SLOT, AGX_COL_POLY_START, AGX_COL_POLY_COUNT, AGX_COL_SCENE };

/* Step phases, in step order (Simulation.cpp:255-351), for agx_world_run_phase. */
enum {
    AGX_PHASE_INTEGRATE = 0,   /* CCD snapshot, velocity-Verlet, wrap, cos/sin */
    AGX_PHASE_POSE,            /* collider world poses + AABBs */
    AGX_PHASE_CCD,             /* swept core circles vs static colliders */
    AGX_PHASE_BROAD,           /* Morton, radix sort, LBVH build/refit, pair traversal */
    AGX_PHASE_NARROW,          /* SAT manifolds, compaction, warm-start lookup */
    AGX_PHASE_SOLVE,           /* colouring, constraint prep, warm start, velocity and position sweeps, scatter */
    AGX_PHASE_ISLANDS,         /* union-find islands */
    AGX_PHASE_FINISH,          /* contact cache for the next step, step counters */
    AGX_PHASE_COUNT
};

void agx_settings_default(agx_settings* s);
const char* agx_last_error(void);
int agx_device_count(int* out);

/* ---- scene building: ObjectAPI::createCircle/createBox/createPolygon (ObjectAPI.h:34-37),
 *      Simulation::createMaterial (Simulation.h:155). Creation semantics follow
 *      SimulationImpl/ObjectManager.cpp:71-389 (mass from the caller, inertia and centre of
 *      mass from the shape, group-ordered dense indices). ---- */
int agx_builder_create(uint32_t num_scenes, agx_builder** out);
int agx_builder_destroy(agx_builder* b);
int agx_builder_add_material(agx_builder* b, float elasticity, float static_friction,
                             float dynamic_friction, uint32_t* out_index);
/* Adds n bodies with one collider each to scene 'scene', in call order. shape_type[i] picks
 * the shape; p0/p1 are radius (circle), width/height (box) or full core length/radius (capsule,
 * CapsuleCreateParams, ObjectCreateParams.h:97-104: its axis is local +x); polygons take vert_count[i]
 * local (x, y) vertices starting at vertex vert_start[i] of 'verts'. out_dense_index (may be
 * NULL) receives each body's dense index at its creation; a rejected body (invalid polygon)
 * gets UINT32_MAX and the call returns the number of rejected bodies. */
int agx_builder_add_bodies(agx_builder* b, uint32_t scene, uint32_t n, const uint8_t* shape_type,
                           const float* pos_x, const float* pos_y, const float* rotation,
                           const float* vel_x, const float* vel_y, const float* ang_vel,
                           const float* mass, const uint32_t* material, const float* p0,
                           const float* p1, const uint32_t* vert_start, const uint32_t* vert_count,
                           const float* verts, uint32_t* out_dense_index);
/* ObjectAPI::createBody + createCircleCollider / createBoxCollider / createCapsuleCollider
 * (ObjectAPI.h:28-32, ObjectManager.cpp:71-181): bodies with several colliders. A body is created
 * without colliders, with its mass, inertia and local centre of mass as given (com_x/com_y, in the
 * body frame; its position is the body origin); colliders are then attached at a local offset and
 * rotation, into their owner's group, and leave the body's mass properties as they are. Colliders
 * of one body never collide with each other. shape_type: AGX_SHAPE_CIRCLE (p0 radius),
 * AGX_SHAPE_BOX (p0 width, p1 height) or AGX_SHAPE_CAPSULE (p0 core length, p1 radius); polygon
 * outlines are created with their body (agx_builder_add_bodies). */
int agx_builder_create_body(agx_builder* b, uint32_t scene, float pos_x, float pos_y, float rotation, float vel_x,
                            float vel_y, float ang_vel, float mass, float inertia, float com_x, float com_y,
                            uint32_t* out_dense_index);
int agx_builder_add_collider(agx_builder* b, uint32_t scene, uint32_t body, uint8_t shape_type, float offset_x,
                             float offset_y, float local_rotation, uint32_t material, float p0, float p1,
                             uint32_t* out_collider_index);
/* Host-side views of a built scene before upload (ObjectAPI viewers, ObjectAPI.h:96-101):
 * dense-order body and collider fields of one scene, and its polygon vertex/normal pool. */
int agx_builder_scene_counts(agx_builder* b, uint32_t scene, agx_counts* out);
int agx_builder_read_body_f32(agx_builder* b, uint32_t scene, uint32_t field, float* out, uint32_t n);
int agx_builder_read_collider(agx_builder* b, uint32_t scene, uint32_t field, uint32_t as_u32, void* out, uint32_t n);
int agx_builder_read_vertices(agx_builder* b, uint32_t scene, float* verts, float* normals, uint32_t n);
/* ObjectAPI::createCircle/createBox/createPolygon/createCapsule after the upload (ObjectAPI.h:34-38,
 * creation semantics as agx_builder_add_bodies, ObjectManager.cpp:71-389): the bodies are created
 * in the world's copy of its scenes, with the reference's group moves, and the world is rebuilt
 * from it with everything it carried (body state, collider poses and CCD origins, the warm-start
 * manifolds and their cache, sleeping islands and their stashed contacts, collider callbacks,
 * queued body commands, joint state, tuning, the learned warm start) moved to the new dense
 * indices by creation slot. Existing bodies keep their state bit for bit. out_dense_index (may be
 * NULL) receives each new body's world index at its creation (UINT32_MAX when rejected); the call
 * returns the number of rejected bodies. Dense indices of existing bodies can change (a static body
 * joins the static group, later scenes shift), as they do in the reference. Between steps only;
 * the rebuild costs about what agx_world_create costs. */
int agx_world_add_bodies(agx_world* w, uint32_t scene, uint32_t n, const uint8_t* shape_type,
                         const float* pos_x, const float* pos_y, const float* rotation,
                         const float* vel_x, const float* vel_y, const float* ang_vel,
                         const float* mass, const uint32_t* material, const float* p0,
                         const float* p1, const uint32_t* vert_start, const uint32_t* vert_count,
                         const float* verts, uint32_t* out_dense_index);
/* ObjectAPI::destroyBody (ObjectAPI.h:70, ObjectManager::destroyBodyImmediate, ObjectManager.cpp:390-418):
 * the bodies (world dense indices, named before any of them is removed) leave the world with their
 * colliders and joints; each is moved to the end of its arrays through its group, swapping with the
 * last member of each group on the way, as the reference does. The world is rebuilt: the warm-start
 * manifolds, stashed contacts and queued commands that name a removed body or collider are dropped,
 * a sleeping island that lost a body wakes, everything else moves to the new dense indices. Between
 * steps only; the rebuild costs about what agx_world_create costs. */
int agx_world_remove_bodies(agx_world* w, uint32_t n, const uint32_t* bodies);
/* The same removal before the upload: body is a dense index of the scene. */
int agx_builder_remove_body(agx_builder* b, uint32_t scene, uint32_t body);

/* ---- joints: ConstraintAPI::createSpring / createRevoluteJoint / createPrismaticJoint /
 * createSliderJoint / createMotor (SimulationAPIs/ConstraintAPI.h:31-35) with the CreateParams of
 * SimulationImpl/ConstraintSystems/(Spring, RevoluteJoint, ...)ConstraintSystem.h:14-26. Types follow ConstraintType
 * (Types/Types.h:50-58), which is also the solving order. body_a / body_b are dense body indices of
 * the scene at the time of the call (the joint follows its bodies if later creations move them);
 * a motor uses body_a only. The prismatic joint's reference angle is the bodies' relative rotation
 * at the call; axes are normalised ((1, 0) when zero). Each step the joints are solved between the
 * two velocity-Verlet halves (Simulation.cpp:384-405, ConstraintScheduler.cpp:9-84): constraint data
 * of every type, the motors' warm start, then joint_iterations[t] velocity passes of type t spread
 * over the largest count, each pass colour by colour (joints of one colour share no dynamic body;
 * the order is read back with agx_world_read_joint_order). */
#define AGX_JOINT_SPRING 0
#define AGX_JOINT_REVOLUTE 1
#define AGX_JOINT_PRISMATIC 2
#define AGX_JOINT_SLIDER 3
#define AGX_JOINT_MOTOR 4
#define AGX_JOINT_TYPES 5
typedef struct agx_joint_desc {
    uint32_t type;
    uint32_t body_a, body_b;
    uint32_t enable_limit;                    /* prismatic, slider */
    float anchor_a_x, anchor_a_y;             /* body-local anchors (not motor) */
    float anchor_b_x, anchor_b_y;
    float axis_x, axis_y;                     /* prismatic, slider: axis in body A's frame */
    float rest_length;                        /* spring */
    float stiffness, damping;                 /* spring, revolute, prismatic, slider */
    float lower_limit, upper_limit;           /* prismatic, slider */
    float target_speed, max_torque;           /* motor */
} agx_joint_desc;
/* Adds n joints to a scene in call order; out_index (may be NULL) receives each one's index among
 * the scene's joints. If a joint names a body that does not exist, or an unknown type, the call adds
 * nothing and returns AGX_ERR_INVALID_ARGUMENT (the reference returns std::nullopt). */
int agx_builder_add_joints(agx_builder* b, uint32_t scene, uint32_t n, const agx_joint_desc* joints, uint32_t* out_index);

/* Uploads the built scenes to device 0 (Simulation::refreshSpatialData, Simulation.cpp:561-579,
 * runs once here) and returns the world. The builder may be destroyed afterwards. */
int agx_world_create(agx_builder* b, const agx_settings* settings, agx_world** out);
int agx_world_destroy(agx_world* w);
const char* agx_world_last_error(agx_world* w);

int agx_world_set_settings(agx_world* w, const agx_settings* s);   /* getSimulationSettings() = */
int agx_world_get_settings(agx_world* w, agx_settings* s);
int agx_world_counts(agx_world* w, agx_counts* out);

/* Simulation::update (Simulation.cpp:32-74): fixed-clock sub-steps, then the post-update pose
 * pass. Returns the number of physics steps taken through *out_steps (may be NULL). */
int agx_world_update(agx_world* w, float dt, uint32_t* out_steps);
/* n_steps physics steps of 'dt' then the post-update pose pass (update() with a clock that
 * yields exactly n_steps). Asynchronous on the world's stream; call agx_world_sync to wait. */
int agx_world_step(agx_world* w, float dt, uint32_t n_steps);
/* One phase of a physics step (tests and tools drive a step phase by phase). */
int agx_world_run_phase(agx_world* w, uint32_t phase, float dt);
int agx_world_post_update(agx_world* w);
int agx_world_sync(agx_world* w);
/* Captures agx_world_step(w, dt, 1) as a HIP graph and replays it n_steps times. */
int agx_world_step_graph(agx_world* w, float dt, uint32_t n_steps);

/* Read-outs: ObjectAPI::getBodies/getColliders viewers (ObjectAPI.h:96-101),
 * Simulation::getBodyCollisionData (Simulation.h:186), getBroadPhaseCollisionDetector
 * (Simulation.h:183), islands() (Simulation.h:203), getDebugData (Simulation.h:189). */
int agx_world_read_body_f32(agx_world* w, uint32_t field, float* out, uint32_t n);

/* Deferred modifications: ObjectAPI::setBodyPosition / setBodyRotation / setBodyVelocity /
 * setBodyAngularVelocity / setBodyMass / setBodyInertia (SimulationAPIs/ObjectAPI.h:76-81),
 * queued for body index `body` and applied in call order at the start of the next
 * agx_world_update / _step / _step_graph / run_phase(INTEGRATE), or by agx_world_flush
 * (ObjectManager::flushCommands, SimulationImpl/ObjectManager.cpp:638-865):
 *   POSITION: (a, b) = the body origin; the centre of mass is placed at origin + R(rotation) local COM;
 *   ROTATION: a, wrapped to [0, 2 pi); cos/sin refreshed;
 *   VELOCITY: (a, b); ANGULAR_VELOCITY: a;
 *   MASS: a > 0 on a dynamic body (the inertia scales with the mass; moving a body between the
 *         static and dynamic groups is not supported and is refused);
 *   INERTIA: max(0, a) on a dynamic body (ignored on a static one).
 * As in the reference, collider poses follow at the next step's pose pass, so that step's CCD
 * sweep starts from the pose before the edit. */
#define AGX_SET_POSITION 0
#define AGX_SET_ROTATION 1
#define AGX_SET_VELOCITY 2
#define AGX_SET_ANGULAR_VELOCITY 3
#define AGX_SET_MASS 4
#define AGX_SET_INERTIA 5
int agx_world_queue_body(agx_world* w, uint32_t kind, uint32_t body, float a, float b);

/* NarrowPhaseCollisionDetector::findCollisions on injected pairs
 * (Narrow/NarrowPhaseCollisionDetector.cpp:101-151, :638-767): the manifolds of collider pairs
 * (pairs_a[i], pairs_b[i]) at the world's current collider poses, in canonical A/B with contact
 * ids, warm-started from the previous step's manifolds, in pair order. Between steps only; the
 * next step's broad and narrow phases are not affected. AGX_ERR_CAPACITY when more than cap
 * manifolds (*count receives the total). */
int agx_world_narrowphase(agx_world* w, const uint32_t* pairs_a, const uint32_t* pairs_b, uint64_t n, agx_manifold* out,
                          uint64_t cap, uint64_t* count);

/* BodyCollisionSolver::solveCollisions on injected manifolds (Solving/BodyCollisionSolver.cpp:229-980:
 * constraint data, warm start, velocity and position sweeps with the world's settings): the n
 * manifolds (colliders a, b; points, normal, depth, contact ids and starting impulses, as
 * agx_world_read_manifolds gives them) are solved against the world's current body state, which
 * the call updates (velocities, positions). order (n entries, a permutation of 0 .. n-1, or NULL for
 * the world's own colouring) is the sequential solve order, e.g. the reference planner's waves
 * (SolvingPlanner.cpp:141-310): it runs level by level, a manifold's level being one more than the
 * highest level before it at either of its dynamic bodies, which keeps every body's sequence of
 * updates and so the sequential result; AGX_ERR_CAPACITY when the order needs more than
 * AGX_MAX_COLOURS levels. out (n entries) receives the solved manifolds. Between steps only; the world's step state (warm-start source, next step) is not
 * changed, but the last step's pair and manifold counts in agx_world_read_stats are replaced. */
#define AGX_MAX_COLOURS 129
int agx_world_solve_contacts(agx_world* w, const agx_manifold* manifolds, uint64_t n, const uint32_t* order,
                             agx_manifold* out);
int agx_world_flush(agx_world* w);
int agx_world_write_body_f32(agx_world* w, uint32_t field, const float* in, uint32_t n);
int agx_world_read_collider_f32(agx_world* w, uint32_t field, float* out, uint32_t n);
int agx_world_read_collider_u32(agx_world* w, uint32_t field, uint32_t* out, uint32_t n);
int agx_world_read_pairs(agx_world* w, uint32_t* a, uint32_t* b, uint64_t cap, uint64_t* count);
/* Sorted keys of the last broad-phase tree build (BvhTree::computeMortonCodes and
 * sortByMortonCodes, Broad/BvhTree.cpp:252-354): which = 0 the dynamic colliders' tree, 1 the
 * static one; n must be that tree's collider count. keys[i] = scene << morton_bits | Morton code of
 * the AABB centre of collider colliders[i] (morton_bits / 2 bits per axis, quantised to the scene's
 * bounds; the reference uses 16 bits per axis and one scene), ascending, equal keys in ascending
 * collider order. Either array may be NULL. */
int agx_world_read_broadphase_order(agx_world* w, uint32_t which, uint32_t* keys, uint32_t* colliders, uint32_t n,
                                    uint32_t* morton_bits);
/* which = 0: manifolds of the current step (pre-solve after NARROW, solved after SOLVE);
 * which = 1: the previous step's manifolds (the warm-start source). */
int agx_world_read_manifolds(agx_world* w, uint32_t which, agx_manifold* out, uint64_t cap, uint64_t* count);
int agx_world_write_manifolds(agx_world* w, uint32_t which, const agx_manifold* in, uint64_t count);
/* Solve order of the last SOLVE phase: order[c] = manifold solved at position c; constraints
 * [colour_start[k], colour_start[k+1]) share colour k and no dynamic body. */
int agx_world_read_solve_order(agx_world* w, uint32_t* order, uint64_t cap, uint32_t* colour_start,
                               uint32_t colour_cap, uint32_t* colours);
int agx_world_read_islands(agx_world* w, uint32_t* island_id, uint32_t n, uint32_t* island_count);
int agx_world_read_stats(agx_world* w, agx_stats* out);
/* Velocity sweeps each scene ran in the last step solved scene by scene (stats.solved_per_scene);
 * n must be the scene count. No counterpart in the reference (one simulation per scene there). */
int agx_world_read_scene_sweeps(agx_world* w, uint32_t* out, uint32_t n);
/* Each scene's relative residual (the norm of its impulse change over the norm of its impulses)
 * after its last velocity sweep in the last step solved scene by scene with track_residual set;
 * n must be the scene count. A scene has met residual_target when its value is at most the target. */
int agx_world_read_scene_residual(agx_world* w, float* out, uint32_t n);

/* Joints of the world (all scenes, scene by scene in creation order) and their solve order: order
 * receives joint indices type by type, colour by colour; colour_start[k] .. colour_start[k+1] is
 * colour k, type_colour[t] .. type_colour[t+1] the colours of type t (type_colour has
 * AGX_JOINT_TYPES + 1 entries). NULL arrays are skipped; *joints and *colours receive the counts. */
int agx_world_read_joint_order(agx_world* w, uint32_t* order, uint32_t cap, uint32_t* colour_start, uint32_t colour_cap,
                               uint32_t* type_colour, uint32_t* joints, uint32_t* colours);
/* Joint records as built (n must be the joint count; body indices are world body indices) and the
 * motors' accumulated impulses (the warm-start state, MotorConstraintSystem::Data::accumulatedImpulse;
 * 0 for other types). */
int agx_world_read_joints(agx_world* w, agx_joint_desc* out, float* motor_impulse, uint32_t n);

/* ---- collision events: BodyCollisionSolver::computeConstraintData's event recording
 * (Solving/BodyCollisionSolver.cpp:300-388), Simulation::setColliderCollisionCallback
 * (Simulation.cpp:152-172) and dispatchCollisionEvents (Simulation.cpp:174-180). While
 * settings.collision_events is set and some collider has a callback id, every step's constraint
 * preparation computes each manifold's impact impulse (the most, over its contact points, of the
 * approach speed times the normal effective mass, before the warm start) and records one event per
 * side whose collider has a callback id and a threshold at most that impulse (impulse > 0), side A
 * before side B, manifolds in ascending index order. Events of all physics steps of one
 * agx_world_update / _step / _step_graph call are kept in step order (the device keeps up to
 * settings.event_capacity of them); the next such call starts a new list. The callback registry
 * and the calls themselves live with the caller (host/GpuSimulation.h runs them after update()). */
typedef struct agx_collision_event {
    uint32_t self, other;             /* collider indices: the opted-in side, the other side */
    uint32_t self_body, other_body;   /* their bodies */
    float point_x, point_y;           /* contact point (mean of the manifold's two points) */
    float normal_x, normal_y;         /* from self towards other */
    float impulse;                    /* impact impulse */
    uint32_t callback_id;             /* self's callback id */
    uint32_t step;                    /* physics step of the call that recorded it (0, 1, ...) */
    uint32_t reserved;
} agx_collision_event;
#define AGX_MAX_CALLBACK_ID 65535u
/* callback_id 0 clears the collider's callback (its threshold returns to "never"); otherwise
 * callback_id <= AGX_MAX_CALLBACK_ID and threshold >= 0 (ColliderSoA callbackId/eventThreshold). */
int agx_world_set_collider_callback(agx_world* w, uint32_t collider, uint32_t callback_id, float threshold);
/* Events of the last update/step call in record order; *count receives how many were recorded
 * (events beyond event_capacity are counted but not kept: *dropped receives that number, may be
 * NULL). AGX_ERR_CAPACITY when more events were kept than cap. */
int agx_world_read_events(agx_world* w, agx_collision_event* out, uint64_t cap, uint64_t* count, uint64_t* dropped);

/* ---- sleeping: ObjectManager::updateSleeping / sleepBody / wakeSleepIslands
 * (SimulationImpl/ObjectManager.cpp:1314-1491), wakeIslandsOfContacts (ObjectManager.h:139-156),
 * NarrowPhaseCollisionDetector::sleepContactData / wakeContactData
 * (Narrow/NarrowPhaseCollisionDetector.cpp:299-414) and Simulation::wakeIslandsOfEditedStatics
 * (Simulation.cpp:463-486). Sleeping is on once a step runs with a finite
 * settings.sleep_time_threshold. After each step's islands, every awake dynamic body's timer grows
 * by dt while |v| + |w| x (bound radius) <= sleep_speed_threshold and restarts at 0 otherwise; an
 * island whose bodies all have timers >= sleep_time_threshold falls asleep, and its manifolds move
 * to a stash. Sleeping bodies are not integrated and pair only with awake dynamic bodies; a manifold
 * that touches one wakes its whole island before the solve, with the island's stashed manifolds
 * appended to the step's. A deferred body command wakes its body's island at the next flush; an
 * edited static body wakes the islands resting on it. Bodies keep their indices (the reference moves
 * a sleeping body into a Sleeping group of its arrays): sleep_island[b] is 0 for an awake body and
 * otherwise its island's sleep id, 2 + the island's smallest body index. */
int agx_world_read_sleep(agx_world* w, uint32_t* sleep_island, float* timer, uint32_t n);
/* The stash: manifolds of sleeping islands, with each one's sleep island id (either array may be
 * NULL; both NULL: *count only). AGX_ERR_CAPACITY when more than cap (*count receives the total). */
int agx_world_read_stash(agx_world* w, agx_manifold* out, uint32_t* island, uint64_t cap, uint64_t* count);

/* ---- world snapshots. The reference has no on-disk or in-memory state format. A snapshot holds
 * everything a later step reads: body state (poses, velocities, cos/sin, masses and inertias as
 * edited), collider poses, AABBs and CCD sweep origins, island ids, the previous step's manifolds
 * (the warm-start source, its hash rebuilt on load), the device counters (statistics, gate), the
 * per-scene sweep counts, collider callbacks, sleep ids and timers with the stash, the settings, the
 * fixed clock's accumulator and the queued body commands. Loading it into a world built from the same scenes (same bodies, colliders,
 * shapes and materials; checked by a fingerprint) makes that world step to the same bits as the world
 * that saved it. The network history of the learned warm start is not kept (its next step runs
 * without one). Format: little-endian, a 64-byte header ("AGXSNAP1", version, counts, fingerprint,
 * payload size) followed by tagged sections (u32 tag, u32 reserved, u64 bytes, data padded to 16). */
int agx_world_snapshot_size(agx_world* w, uint64_t* bytes);
int agx_world_save(agx_world* w, void* buffer, uint64_t cap, uint64_t* written);
int agx_world_load(agx_world* w, const void* buffer, uint64_t bytes);

/* GPU time per step segment, measured with HIP events on the world's stream while profiling
 * is on. Segments: see agx_profile_segment_name. agx_world_read_profile waits for the stream,
 * returns per-segment total milliseconds and launch counts since the last read, and resets. */
int agx_world_set_profiling(agx_world* w, uint32_t enabled);
int agx_world_read_profile(agx_world* w, double* ms_total, uint32_t* launches, uint32_t cap, uint32_t* segments);
const char* agx_profile_segment_name(uint32_t segment);

/* Launch-shape tuning. The reference partitions work by fixed sizes (COLLISION_COUNT_PER_WORKER,
 * Solving/BodyCollisionSolver.cpp:211-212; DETERMINISTIC_CHUNK_SIZE,
 * CollisionDetection/Narrow/NarrowPhaseCollisionDetector.h:153). Here the launch shape of the path the
 * world's steps run is measured instead: for a world solved with the dataflow kernels, their
 * co-resident blocks per CU (1, 2, 4 ... up to the maximum); for a world solved scene by scene
 * (solve_per_scene, every scene in LDS), how many of the step's heaviest scenes whole blocks take
 * before the rest go to lane groups (half, one, one and a half and two per block).
 * agx_world_autotune times graph-replayed steps for each candidate (two interleaved rounds, which
 * advances the simulation) and keeps the fastest, per sweep run when a residual target makes the
 * sweep count vary. Every shape steps to the same bits. Results are cached in the process per GPU,
 * world size class and solver settings, and in the file named by the environment variable
 * AGX_TUNE_CACHE when it is set; a cached shape is applied without stepping (from_cache = 1). */
typedef struct agx_tuning {
    uint32_t solve_blocks_per_cu;       /* dataflow sweep kernel */
    uint32_t colour_blocks_per_cu;      /* dataflow colouring kernel */
    float ms_per_step;                  /* mean step time measured with this shape (autotune only) */
    uint32_t from_cache;
    uint32_t max_solve_blocks_per_cu, max_colour_blocks_per_cu;   /* limits (get_tuning, autotune) */
    uint32_t scene_whole;               /* per-scene solve: heaviest scenes whole blocks take (get_tuning,
                                           autotune; 0 when the world cannot solve scene by scene);
                                           set_tuning applies it unless it is 0 */
} agx_tuning;
int agx_world_get_tuning(agx_world* w, agx_tuning* out);
int agx_world_set_tuning(agx_world* w, const agx_tuning* in);
int agx_world_autotune(agx_world* w, float dt, uint32_t steps_per_trial, agx_tuning* out);

/* Standalone broad phase over injected AABBs (BroadPhaseCollisionDetector::findCollisions,
 * Broad/BroadPhaseCollisionDetector.cpp:109-244): aabb holds n x (min_x, min_y, max_x, max_y),
 * owner[i] the owning body, is_dynamic[i] 1 for dynamic colliders, scene[i] (may be NULL) the
 * scene of each collider. Writes pairs (dynamic x dynamic of different owners, dynamic x static)
 * of strictly overlapping AABBs in one scene to pairs_a/pairs_b; *count receives the total, and
 * AGX_ERR_CAPACITY is returned when it exceeds cap. */
int agx_broadphase_pairs(const float* aabb, const uint32_t* owner, const uint8_t* is_dynamic,
                         const uint32_t* scene, uint32_t n, uint32_t* pairs_a, uint32_t* pairs_b,
                         uint64_t cap, uint64_t* count);

/* No counterpart in the reference. For callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc): agx_broadphase_pairs over DEVICE buffers (aabb 16-byte
 * aligned; scene may be NULL for one scene; num_scenes must exceed every scene index) writing the
 * same pairs, in the same order, to the device arrays pairs_a/pairs_b; *count is a host pointer.
 * Every intermediate array lives in one static device array of the library (about 36 MB per
 * device, held until the process ends); the call runs on the null stream, synchronizes, and
 * allocates nothing. Limits: n <= AGX_NOALLOC_MAX_COLLIDERS, num_scenes <= AGX_NOALLOC_MAX_SCENES,
 * at most AGX_NOALLOC_MAX_PAIRS pairs (AGX_ERR_CAPACITY beyond). Calls from one process run one
 * at a time (a mutex); each call starts from cleared counters whatever the previous call left. */
#define AGX_NOALLOC_MAX_COLLIDERS 65536u
#define AGX_NOALLOC_MAX_SCENES 1024u
#define AGX_NOALLOC_MAX_PAIRS 1048576u
int agx_broadphase_pairs_noalloc(const float* aabb, const uint32_t* owner, const uint8_t* is_dynamic,
                                 const uint32_t* scene, uint32_t n, uint32_t num_scenes, uint32_t* pairs_a,
                                 uint32_t* pairs_b, uint64_t cap, uint64_t* count);

/* ---- learned warm start. No counterpart in the reference, which warm-starts every contact point
 * from the cached accumulators of the previous step (NarrowPhaseCollisionDetector.cpp:753-765,
 * BodyCollisionSolver.cpp:416-490). A network 24 -> 128 -> 128 -> 2 (ReLU, bf16 weights and
 * activations, fp32 accumulation) maps per-point contact features to corrections of those cached
 * accumulators; see k_learn.hip for the features and physicssimulation-agony_amd/csrc/k_mlp.hip for
 * the forward, backward and Adam kernels.
 * Parameters in host layout (AGX_MLP_PARAMS floats): W1[128][24], b1[128], W2[128][128], b2[128],
 * W3[2][128], b3[2]; layer l computes h = W h_prev + b. */
#define AGX_MLP_INPUTS 24
#define AGX_MLP_HIDDEN 128
#define AGX_MLP_OUTPUTS 2
#define AGX_MLP_PARAMS 19970
#define AGX_MLP_BATCH_MAX 8192

/* He-initialised network (seeded) with a device buffer for sample_capacity training rows. */
int agx_mlp_create(uint32_t seed, uint64_t sample_capacity, agx_mlp** out);
int agx_mlp_destroy(agx_mlp* m);
int agx_mlp_get_params(agx_mlp* m, float* out, uint64_t cap);
int agx_mlp_set_params(agx_mlp* m, const float* in, uint64_t n);   /* also resets the optimiser */
/* x: n x 24 features (rounded to bf16 on upload), y: n x 2 outputs. */
int agx_mlp_forward(agx_mlp* m, const float* x, uint64_t n, float* y);
/* Gradient of the mean squared error (1/n) sum_rows sum_outputs (y - t)^2 over one batch of n rows
 * (a positive multiple of 16, at most AGX_MLP_BATCH_MAX), in host parameter layout. */
int agx_mlp_gradients(agx_mlp* m, const float* x, const float* t, uint64_t n, float* grad, float* loss);
/* Training rows: appended by agx_mlp_add_samples or by a world in AGX_LEARN_COLLECT mode. */
int agx_mlp_add_samples(agx_mlp* m, const float* x, const float* t, uint64_t n);
int agx_mlp_sample_count(agx_mlp* m, uint64_t* count);
int agx_mlp_read_samples(agx_mlp* m, float* x, float* t, uint64_t offset, uint64_t n);
int agx_mlp_clear_samples(agx_mlp* m);
/* Adam (0.9, 0.999, 1e-8) over `iterations` random batches of the stored rows, all on the GPU;
 * *mean_loss receives the mean batch loss. */
int agx_mlp_train(agx_mlp* m, uint32_t iterations, uint32_t batch, float learning_rate, uint32_t seed, float* mean_loss);

/* Modes of agx_world_set_learned (bit flags). PREDICT: warm start from the network, subject to the
 * residual gate unless UNGATED; COLLECT: append every contact point of each step to the network's
 * training rows (features, and as targets the corrections that reach the solver's final
 * accumulators). The world keeps a reference to the network, which must outlive it. */
#define AGX_LEARN_OFF 0u
#define AGX_LEARN_PREDICT 1u
#define AGX_LEARN_COLLECT 2u
#define AGX_LEARN_UNGATED 4u
/* Apply the network only at contact points without a cached impulse (new contacts); points with one
   keep the cached warm start. With AGX_LEARN_COLLECT, only such points give training rows. */
#define AGX_LEARN_NEW_CONTACTS 8u
int agx_world_set_learned(agx_world* w, agx_mlp* mlp, uint32_t mode);

/* Residual gate: exponential averages (weight 1/4 per step) of the cost of a step with the cached
 * (cost_cached) and the learned warm start (cost_learned). Cost: log10 of the relative residual
 * after the 8th velocity sweep (after the last one when fewer run). The cheaper method runs; the other runs once as a
 * probe every `period` steps (16 at first, doubled after each losing probe up to 1024, reset by a
 * winning one) and its cost replaces that method's average; a learned step that cost more than the
 * cached average is followed by a cached step. agx_gate_update is the same update on the host (no
 * device work). */
typedef struct agx_gate {
    float cost_cached, cost_learned;
    uint32_t use_learned;          /* method of the next step */
    uint32_t since_probe;
    uint32_t steps_cached, steps_learned;
    uint32_t period;               /* steps between probes (0 before the first update) */
    uint32_t probing;              /* the next step is a probe */
} agx_gate;
int agx_gate_update(agx_gate* g, uint32_t used_learned, float cost);
int agx_world_read_gate(agx_world* w, agx_gate* out);
int agx_world_reset_gate(agx_world* w);

#ifdef __cplusplus
}
#endif

#endif /* AGX_H */
