// Oracle sleeping and waking (test infrastructure only; see oracle.h). The reference moves a
// sleeping body into a Sleeping group of its dense arrays (ObjectManager.cpp:1286-1361); here a body
// sleeps in place and carries its sleep island's id, which gives the same step: sleeping bodies are
// not integrated, pair only with awake dynamic bodies, wake with their whole island on contact, and
// their islands' manifolds wait in the stash until then.
#include "oracle.h"

#include <algorithm>
#include <cmath>

namespace orc {

namespace {

constexpr uint32_t FIRST_SLEEP_ISLAND_ID = 2;   // SoA/BodySoA.h:45

void ensure(World& w) {
    if (w.sleepId.size() != w.bodyCount()) {
        w.sleepId.resize(w.bodyCount(), 0u);
        w.sleepTimer.resize(w.bodyCount(), 0.0f);
    }
}

// wakeOne for every body of a flagged island (ObjectManager.cpp:1344-1361: timer 0), then
// wakeContactData (NarrowPhaseCollisionDetector.cpp:348-414) with the stash in its own order. Its
// match against `into` takes stored impulses for pairs generated this frame; a woken body slept
// through this frame's broad phase (or the previous one, for `prevManifolds`), and a stashed manifold
// joins two bodies of its island or one and a static body, so no stashed pair is in `into` and every
// restored manifold is appended.
void wakeFlagged(World& w, const std::vector<uint8_t>& woken, std::vector<Manifold>& into) {
    for (uint32_t b = 0; b < w.bodyCount(); b++) {
        const uint32_t id = w.sleepId[b];
        if (id != 0 && woken[id - FIRST_SLEEP_ISLAND_ID]) {
            w.sleepId[b] = 0;
            w.sleepTimer[b] = 0.0f;
        }
    }
    std::vector<Manifold> keep;
    std::vector<uint32_t> keepIsland;
    for (size_t s = 0; s < w.stash.size(); s++) {
        if (woken[w.stashIsland[s] - FIRST_SLEEP_ISLAND_ID]) {
            into.push_back(w.stash[s]);
        } else {
            keep.push_back(w.stash[s]);
            keepIsland.push_back(w.stashIsland[s]);
        }
    }
    w.stash.swap(keep);
    w.stashIsland.swap(keepIsland);
}

}  // namespace

bool asleep(const World& w, uint32_t b) { return b < w.sleepId.size() && w.sleepId[b] != 0; }

void wakeContacts(World& w, std::vector<Manifold>& cur) {
    ensure(w);
    std::vector<uint8_t> woken(w.bodyCount(), 0);
    bool any = false;
    for (const Manifold& d : cur) {
        const uint32_t bodies[2] = { w.colBody[d.a], w.colBody[d.b] };
        for (uint32_t b : bodies)
            if (w.sleepId[b] != 0) {
                woken[w.sleepId[b] - FIRST_SLEEP_ISLAND_ID] = 1;
                any = true;
            }
    }
    if (any) wakeFlagged(w, woken, cur);
}

void updateSleeping(World& w, float dt, std::vector<Manifold>& cur) {
    ensure(w);
    const uint32_t K = w.islandCount;
    std::vector<uint32_t> flag(K, 1u), minBody(K, UINT32_MAX);
    const float speedThreshold = w.settings.sleepSpeedThreshold, timeThreshold = w.settings.sleepTimeThreshold;
    for (uint32_t b = w.staticBodies; b < w.bodyCount(); b++) {
        if (w.sleepId[b] != 0) continue;
        const float speed = std::sqrt(w.velX[b] * w.velX[b] + w.velY[b] * w.velY[b]) + std::fabs(w.angVel[b]) * w.boundRadius[b];
        const bool quiet = speed <= speedThreshold;
        const float timer = (w.sleepTimer[b] + dt) * float(quiet);
        w.sleepTimer[b] = timer;
        const uint32_t isl = w.islandId[b];
        if (isl >= K) continue;   // woken after the islands were built: stays awake this step
        flag[isl] &= uint32_t(quiet) & uint32_t(timer >= timeThreshold);
        minBody[isl] = std::min(minBody[isl], b);
    }
    bool slept = false;
    for (uint32_t b = w.staticBodies; b < w.bodyCount(); b++) {
        const uint32_t isl = w.islandId[b];
        if (w.sleepId[b] != 0 || isl >= K || !flag[isl]) continue;
        w.sleepId[b] = FIRST_SLEEP_ISLAND_ID + minBody[isl];   // sleepBody (:1314-1342)
        w.sleepTimer[b] = 0.0f;
        slept = true;
    }
    if (!slept) return;
    // sleepContactData: a stable split, each stashed manifold filed under A's island, else B's
    std::vector<Manifold> kept;
    for (const Manifold& d : cur) {
        const uint32_t bA = w.colBody[d.a], bB = w.colBody[d.b];
        const uint32_t id = w.sleepId[bA] != 0 ? w.sleepId[bA] : w.sleepId[bB];
        if (id != 0) {
            w.stash.push_back(d);
            w.stashIsland.push_back(id);
        } else {
            kept.push_back(d);
        }
    }
    cur.swap(kept);
}

void wakeForCommands(World& w, const std::vector<uint32_t>& commandedBodies, const std::vector<uint32_t>& editedStatic) {
    ensure(w);
    std::vector<uint8_t> woken(w.bodyCount(), 0);
    bool any = false;
    for (uint32_t b : commandedBodies)
        if (w.sleepId[b] != 0) {
            woken[w.sleepId[b] - FIRST_SLEEP_ISLAND_ID] = 1;
            any = true;
        }
    if (!editedStatic.empty()) {
        std::vector<uint8_t> edited(w.colliderCount(), 0);
        for (uint32_t b : editedStatic)
            for (uint32_t c : w.bodyColliders[b]) edited[c] = 1;
        // takeContactsOfColliders (NarrowPhaseCollisionDetector.cpp:416-457)
        std::vector<Manifold> keep;
        std::vector<uint32_t> keepIsland;
        for (size_t s = 0; s < w.stash.size(); s++) {
            const Manifold& d = w.stash[s];
            if (edited[d.a] || edited[d.b]) {
                woken[w.stashIsland[s] - FIRST_SLEEP_ISLAND_ID] = 1;
                any = true;
                continue;
            }
            keep.push_back(d);
            keepIsland.push_back(w.stashIsland[s]);
        }
        w.stash.swap(keep);
        w.stashIsland.swap(keepIsland);
        std::vector<uint32_t> partners;
        for (const Manifold& d : w.prevManifolds)
            if (edited[d.a] != edited[d.b]) partners.push_back(w.colBody[edited[d.a] ? d.b : d.a]);
        if (any) wakeFlagged(w, woken, w.prevManifolds);
        any = false;
        for (uint32_t b : partners)   // Simulation.cpp:478-485
            if (b >= w.staticBodies && w.sleepId[b] == 0) w.sleepTimer[b] = 0.0f;
    }
    if (any) wakeFlagged(w, woken, w.prevManifolds);
}

}  // namespace orc
