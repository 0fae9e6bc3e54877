// Steps a small pile through the C++ mirror (GpuSimulation.h) the way a PS_AGONY2D program drives
// Simulation: create a material and bodies, call update(), read manifolds, islands and bodies.
// Exit status 0 when the run is sane (bodies stay finite and inside the box, contacts form).
#include "GpuSimulation.h"

#include <cmath>
#include <cstdio>

using namespace PS_AGONY2D_GPU;

int main() {
    Simulation sim;
    auto& settings = sim.getSimulationSettings();
    settings.velocity_iterations = 8;
    settings.update_interval = 1.0f / 60.0f;
    const auto mat = sim.createMaterial({ 0.0f, 0.6f, 0.6f });
    if (!mat) return 2;
    auto& objects = sim.objects();
    // static box of 20 x 20 m
    const Vec2 walls[4][2] = { { { 0, -0.5f }, { 21, 1 } }, { { 0, 20.5f }, { 21, 1 } },
                               { { -10.5f, 10 }, { 1, 21 } }, { { 10.5f, 10 }, { 1, 21 } } };
    for (const auto& w : walls) {
        BoxCreateParams p;
        p.base.position = w[0];
        p.base.materialIndex = *mat;
        p.size = w[1];
        if (!objects.createBox(p)) return 2;
    }
    int made = 0, capsules = 0;
    for (int iy = 0; iy < 12; iy++)
        for (int ix = 0; ix < 16; ix++) {
            BodyCreateParams b;
            b.position = { -8.5f + 1.1f * ix + 0.05f * (iy & 1), 1.0f + 1.1f * iy };
            b.rotation = 0.3f * (float)(ix + iy);
            b.mass = 1.0f;
            b.materialIndex = *mat;
            bool ok;
            switch ((ix + iy) % 3) {
            case 0:
                if (ix % 4 == 0) {   // a capsule given by its end points, tilted like the rotation would
                    CapsuleCreateParams p;
                    p.base = b;
                    p.radius = 0.2f;
                    p.start = Vec2{ b.position.x - 0.3f * std::cos(b.rotation), b.position.y - 0.3f * std::sin(b.rotation) };
                    p.end = Vec2{ b.position.x + 0.3f * std::cos(b.rotation), b.position.y + 0.3f * std::sin(b.rotation) };
                    ok = objects.createCapsule(p).has_value();
                    capsules += ok ? 1 : 0;
                } else {
                    CircleCreateParams p; p.base = b; p.radius = 0.45f; ok = objects.createCircle(p).has_value();
                }
                break;
            case 1: { BoxCreateParams p; p.base = b; p.size = { 0.8f, 0.6f }; ok = objects.createBox(p).has_value(); break; }
            default: {
                PolygonCreateParams p;
                p.base = b;
                for (int k = 0; k < 5; k++) p.localVertices.push_back({ 0.45f * std::cos(1.2566f * k), 0.45f * std::sin(1.2566f * k) });
                ok = objects.createPolygon(p).has_value();
            }
            }
            made += ok ? 1 : 0;
        }
    // an invalid polygon (collinear) is rejected like the reference does
    PolygonCreateParams bad;
    bad.base.mass = 1.0f;
    bad.localVertices = { { 0, 0 }, { 1, 0 }, { 2, 0 } };
    const bool rejected = !objects.createPolygon(bad).has_value();

    // collision callbacks (Simulation::setColliderCollisionCallback) on the first 40 dynamic colliders
    struct Tally { uint32_t events = 0, badNormals = 0; float maxImpulse = 0; } tally;
    const auto cb = sim.createCollisionCallback(
        [](const CollisionEvent& e, Simulation&, void* user) {
            Tally* t = static_cast<Tally*>(user);
            t->events++;
            t->maxImpulse = std::fmax(t->maxImpulse, e.impulse);
            t->badNormals += std::fabs(e.normal.x * e.normal.x + e.normal.y * e.normal.y - 1.0f) > 1e-4f ? 1u : 0u;
        },
        &tally);
    bool callbacksSet = cb.has_value() && !sim.setColliderCollisionCallback(5, 77, 0.0f);   // 77 is not alive
    for (ObjectDenseIndex c = 4; c < 44; c++) callbacksSet = callbacksSet && sim.setColliderCollisionCallback(c, *cb, 0.0f);

    // a spinning motor on one body and a spring between two others (ConstraintAPI)
    auto& joints = sim.constraints();
    const bool jointsMade = joints.createMotor({ 150, 3.0f, 50.0f }).has_value() &&
                            joints.createSpring({ 151, 152, { 0, 0 }, { 0, 0 }, 1.2f, 30.0f, 0.5f }).has_value() &&
                            !joints.createMotor({ 100000, 1.0f, 1.0f }).has_value();   // no such body

    // a body with two colliders (ObjectAPI::createBody + createCircleCollider / createBoxCollider)
    const auto dumbbell = objects.createBody({ { -6.0f, 16.0f }, {}, 0.0f, 0.0f, 2.0f, 1.2f, {} });
    const bool multiMade = dumbbell && objects.createCircleCollider({ { *dumbbell, { 0.8f, 0.0f } }, 0.4f }).has_value() &&
                           objects.createBoxCollider({ { *dumbbell, { -0.8f, 0.0f } }, { 0.6f, 0.6f } }).has_value() &&
                           !objects.createCircleCollider({ { 1000000u, {} }, 0.4f }).has_value();   // no such body
    // a setter before the upload is applied at the first update
    const bool queuedEarly = objects.setBodyVelocity(10, { 0.0f, 2.0f });
    uint32_t steps = 0;
    for (int frame = 0; frame < 180; frame++) steps += sim.update(1.0f / 60.0f);
    // deferred setters: lift one body above the pile and spin it, double another's mass
    const ObjectDenseIndex lifted = 100, heavy = 120;
    const bool queued = objects.setBodyPosition(lifted, { 0.0f, 18.0f }) && objects.setBodyAngularVelocity(lifted, 4.0f) &&
                        objects.setBodyMass(heavy, 2.0f) && !objects.setBodyMass(0, 1.0f);   // a wall stays static
    steps += sim.update(1.0f / 60.0f);
    const auto liftedData = objects.getBodyData(lifted);
    const auto heavyData = objects.getBodyData(heavy);
    const bool applied = liftedData && liftedData->centerOfMass.y > 16.0f && liftedData->angularVelocity != 0.0f &&
                         heavyData && heavyData->mass == 2.0f && heavyData->invMass == 0.5f;
    // a body created after the upload: the world is rebuilt around it and it falls into the box
    const uint32_t before = sim.counts().bodies;
    CircleCreateParams lateParams;
    lateParams.base.position = { 0.5f, 19.0f };
    lateParams.base.mass = 1.0f;
    lateParams.radius = 0.4f;
    const auto late = objects.createCircle(lateParams);
    const bool lateMade = late && *late == before && sim.counts().bodies == before + 1 &&
                          !sim.createMaterial({ 0.0f, 0.5f, 0.5f }).has_value() &&   // materials only before the upload
                          objects.destroyBody(before - 1) && sim.counts().bodies == before &&   // the dumbbell; the late body moves into its place
                          !objects.destroyBody(1000000u);
    for (int frame = 0; frame < 60; frame++) steps += sim.update(1.0f / 60.0f);
    const auto lateData = late ? objects.getBodyData(before - 1) : std::nullopt;
    const bool lateFell = lateData && lateData->centerOfMass.y < 18.0f;
    const ColliderArrays cols = objects.getColliders();
    const PolygonArrays polys = objects.getPolygons();
    uint32_t polygonColliders = 0, capsuleColliders = 0;
    bool aabbsContain = true;
    for (size_t c = 0; c < cols.worldX.size(); c++) {
        polygonColliders += cols.shapeType[c] == AGX_SHAPE_POLYGON ? 1u : 0u;
        capsuleColliders += cols.shapeType[c] == AGX_SHAPE_CAPSULE ? 1u : 0u;
        aabbsContain = aabbsContain && cols.aabbMinX[c] <= cols.worldX[c] && cols.worldX[c] <= cols.aabbMaxX[c] &&
                       cols.aabbMinY[c] <= cols.worldY[c] && cols.worldY[c] <= cols.aabbMaxY[c];
    }
    const bool viewers = polygonColliders == 64 && capsules > 0 && capsuleColliders == (uint32_t)capsules &&
                         polys.count.size() == cols.worldX.size() && polys.vertices.size() == 64 * 5 && aabbsContain &&
                         cols.worldX.size() == (size_t)before;   // the dumbbell (two colliders) was removed, the late body added
    const auto manifolds = sim.getBodyCollisionData();
    const auto [ids, islands] = sim.islands();
    const BodyArrays bodies = objects.getBodies();
    bool finite = true, inside = true;
    for (size_t i = 0; i < bodies.comX.size(); i++) {
        finite = finite && std::isfinite(bodies.comX[i]) && std::isfinite(bodies.comY[i]);
        inside = inside && std::fabs(bodies.comX[i]) < 11.0f && bodies.comY[i] > -1.0f && bodies.comY[i] < 21.0f;
    }
    const agx_stats st = sim.getStats();
    std::printf("bodies %d (capsules %d) steps %u manifolds %zu islands %u colours %u overflow %u rejected %d setters %d/%d/%d viewers %d"
                " events %u (max impulse %.3f) created after upload %d/%d\n",
                made, capsules, steps, manifolds.size(), islands, st.colours, st.overflow, rejected ? 1 : 0, queuedEarly ? 1 : 0,
                queued ? 1 : 0, applied ? 1 : 0, viewers ? 1 : 0, tally.events, tally.maxImpulse, lateMade ? 1 : 0,
                lateFell ? 1 : 0);
    const bool ok = made == 192 && steps == 241 && finite && inside && rejected && manifolds.size() > 100 && islands > 0 &&
                    st.overflow == 0 && queuedEarly && queued && applied && viewers && callbacksSet && tally.events > 0 &&
                    tally.badNormals == 0 && tally.maxImpulse > 0.0f && jointsMade && lateMade && lateFell && multiMade;
    return ok ? 0 : 1;
}
