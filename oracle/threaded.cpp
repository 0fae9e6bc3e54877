// Threaded CPU step (test infrastructure only; the CPU baseline's structure): the reference's
// threading of one physics step on a pool of pinned std::threads with the caller as worker 0
// (Threading/Threading.h:10-40, SolverBase.cpp:107):
// - narrow phase: 1,024-pair chunks claimed by the participants and merged in chunk order, above
//   6,000 pairs, at most one participant per 3,000 pairs (NarrowPhaseCollisionDetector.cpp:119-145,
//   :519-636; Narrow/Constants.h:17-19);
// - solver: the greedy cache-line-aware wave plan (SolvingPlanner.cpp:141-310) for one worker per
//   3,320 manifolds above 6,640 (BodyCollisionSolver.cpp:205-228), constraint data gathered in the
//   plan's order, then warm start and sweeps as threaded waves with a barrier per wave
//   (SolverBase.cpp:15-116, BodyCollisionSolver.cpp:982-1031);
// - broad phase, CCD, integration and islands single-threaded (Simulation.cpp:413).
// Passes of one wave share no body group, so the result equals the sequential solve in the plan's
// flat order bit for bit.
#include "oracle.h"
#include "solver_internal.h"

#include <pthread.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <thread>
#include <vector>

namespace orc {

struct Pool {
    std::vector<std::thread> threads;
    std::atomic<uint32_t> generation{ 0 };
    std::atomic<uint32_t> running{ 0 };
    const std::function<void(uint32_t)>* job = nullptr;
    uint32_t active = 0;
    bool stop = false;

    std::vector<int> cpus;   // the CPUs the calling thread may run on, in order
    cpu_set_t callerMask;    // the caller's own affinity, restored when the pool ends
    bool callerMaskValid = false;

    explicit Pool(uint32_t participants) {
        CPU_ZERO(&callerMask);
        callerMaskValid = pthread_getaffinity_np(pthread_self(), sizeof callerMask, &callerMask) == 0;
        if (callerMaskValid)
            for (int c = 0; c < CPU_SETSIZE; c++)
                if (CPU_ISSET(c, &callerMask)) cpus.push_back(c);
        pin(pthread_self(), 0);
        for (uint32_t i = 1; i < participants; i++) {
            threads.emplace_back([this, i] { loop(i); });
            pin(threads.back().native_handle(), i);
        }
    }
    ~Pool() {
        stop = true;
        generation.fetch_add(1, std::memory_order_release);
        generation.notify_all();
        for (auto& t : threads) t.join();
        if (callerMaskValid) pthread_setaffinity_np(pthread_self(), sizeof callerMask, &callerMask);
    }
    uint32_t size() const { return (uint32_t)threads.size() + 1; }

    void pin(pthread_t t, uint32_t participant) const {
        if (cpus.empty()) return;
        cpu_set_t set;
        CPU_ZERO(&set);
        CPU_SET(cpus[participant % cpus.size()], &set);
        pthread_setaffinity_np(t, sizeof set, &set);
    }

    void loop(uint32_t index) {
        uint32_t seen = 0;
        while (true) {
            for (int spin = 0; spin < 4000 && generation.load(std::memory_order_acquire) == seen; spin++)
                __builtin_ia32_pause();
            generation.wait(seen, std::memory_order_acquire);
            seen = generation.load(std::memory_order_acquire);
            if (stop) return;
            if (index < active) (*job)(index);
            if (running.fetch_sub(1, std::memory_order_acq_rel) == 1) running.notify_one();
        }
    }

    // fn(i) for participants i in [0, n), n <= size(); the caller runs i = 0.
    void run(uint32_t n, const std::function<void(uint32_t)>& fn) {
        n = std::max(1u, std::min(n, size()));
        if (n == 1) { fn(0); return; }
        job = &fn;
        active = n;
        running.store((uint32_t)threads.size(), std::memory_order_release);
        generation.fetch_add(1, std::memory_order_release);
        generation.notify_all();
        fn(0);
        for (uint32_t r; (r = running.load(std::memory_order_acquire)) != 0;) running.wait(r, std::memory_order_acquire);
    }
};

Pool* poolCreate(uint32_t participants) { return new Pool(std::max(1u, participants)); }
void poolDestroy(Pool* p) { delete p; }
uint32_t poolSize(const Pool* p) { return p ? p->size() : 1u; }

namespace {

constexpr size_t CHUNK_PAIRS = 1024;              // DETERMINISTIC_CHUNK_SIZE
constexpr size_t PAIRS_PER_WORKER = 3000;         // Narrow/Constants.h:17
constexpr size_t MULTITHREAD_PAIR_THRESHOLD = 6000;
constexpr size_t COLLISIONS_PER_WORKER = 830 * 4; // BodyCollisionSolver.cpp:213-214
constexpr size_t MIN_COLLISIONS_FOR_THREADING = COLLISIONS_PER_WORKER * 2;
constexpr uint8_t UNOWNED = 0xFF;

// Ticket barrier of SolverBase::runThreadedWaves: tick t runs wave t mod W on every participant;
// the last to finish a tick opens the next one.
void threadedWaves(Pool& pool, uint32_t participants, uint32_t waves, uint32_t ticks,
                   const std::function<void(uint32_t wave, uint32_t worker, uint32_t tick)>& task) {
    if (waves == 0 || ticks == 0) return;
    std::atomic<uint32_t> ticket{ 0 }, notDone{ participants };
    pool.run(participants, [&](uint32_t worker) {
        uint32_t local = 0;
        while (local < ticks) {
            task(local % waves, worker, local);
            if (notDone.fetch_sub(1, std::memory_order_acq_rel) == 1) {
                notDone.store(participants, std::memory_order_relaxed);
                ticket.fetch_add(1, std::memory_order_release);
                ticket.notify_all();
                local++;
            } else {
                for (int spin = 0; spin < 2000 && ticket.load(std::memory_order_acquire) == local; spin++)
                    __builtin_ia32_pause();
                if (ticket.load(std::memory_order_acquire) == local) ticket.wait(local, std::memory_order_acquire);
                local = ticket.load(std::memory_order_acquire);
            }
        }
    });
}

}  // namespace

// SolvingPlanner::planGreedy<true> with the cache-line groups of planCacheLineAwareExecution
// (bodies >> 4: 16 four-byte body values per 64-byte line): per wave, each contact goes to the
// worker that owns both body groups, or claims free groups for the least-loaded worker (the lowest
// owner when one group is free); contacts whose groups two workers own wait for a later wave.
// Static bodies are never claimed (a contact with one is keyed by its dynamic body).
WavePlan planWaves(const World& w, const std::vector<Manifold>& ms, uint32_t workers) {
    WavePlan plan;
    plan.workers = workers;
    const uint32_t shift = 4;
    const uint32_t n = w.bodyCount(), s = w.staticBodies;
    const size_t M = ms.size();
    if (M == 0 || workers == 0 || n == s) return plan;
    const uint32_t base = s >> shift;
    const uint32_t groups = ((n - 1) >> shift) - base + 1;
    std::vector<uint32_t> ga(M), gb(M);
    for (size_t i = 0; i < M; i++) {
        const uint32_t a = w.colBody[ms[i].a], b = w.colBody[ms[i].b];
        const bool sa = a < s, sb = b < s;
        ga[i] = ((sa ? b : a) >> shift) - base;
        gb[i] = ((sb ? a : b) >> shift) - base;
    }
    std::vector<uint8_t> owner(groups);
    std::vector<size_t> load(workers);
    std::vector<std::vector<uint32_t>> queue(workers);
    std::vector<uint32_t> pending, half, next;
    pending.reserve(M);
    for (size_t i = 0; i < M; i++) pending.push_back((uint32_t)i);
    while (!pending.empty()) {
        std::fill(owner.begin(), owner.end(), UNOWNED);
        std::fill(load.begin(), load.end(), 0);
        for (auto& q : queue) q.clear();
        half.clear();
        next.clear();
        for (uint32_t i : pending) {
            const uint8_t oa = owner[ga[i]], ob = owner[gb[i]];
            if (oa == ob) {
                uint32_t k = oa;
                if (oa == UNOWNED) {
                    k = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
                    owner[ga[i]] = owner[gb[i]] = (uint8_t)k;
                }
                queue[k].push_back(i);
                load[k]++;
            } else {
                (oa == UNOWNED || ob == UNOWNED ? half : next).push_back(i);
            }
        }
        for (uint32_t i : half) {
            const uint8_t oa = owner[ga[i]], ob = owner[gb[i]];
            if (oa != UNOWNED && ob != UNOWNED && oa != ob) { next.push_back(i); continue; }
            const uint8_t k = std::min(oa, ob);
            queue[k].push_back(i);
            load[k]++;
            owner[ga[i]] = owner[gb[i]] = k;
        }
        for (uint32_t k = 0; k < workers; k++) {
            plan.passStart.push_back((uint32_t)plan.flat.size());
            plan.passSize.push_back((uint32_t)queue[k].size());
            plan.flat.insert(plan.flat.end(), queue[k].begin(), queue[k].end());
        }
        pending.swap(next);
    }
    return plan;
}

// BodyCollisionSolver::solveCollisions (:38-117): sequential below the threading threshold, else
// the plan's waves on the pool.
void solveThreaded(World& w, std::vector<Manifold>& ms, Pool& pool, uint32_t vIt, uint32_t pIt) {
    const size_t M = ms.size();
    const size_t needed = M < MIN_COLLISIONS_FOR_THREADING ? 0 : M / COLLISIONS_PER_WORKER;
    const uint32_t workers = (uint32_t)std::min<size_t>(pool.size(), needed);
    if (workers <= 1) {
        solve(w, ms, nullptr, vIt, pIt, nullptr);
        return;
    }
    const WavePlan plan = planWaves(w, ms, workers);
    Solver sv(w, ms);
    sv.prepare(plan.flat.data(), 0, M);
    const uint32_t waves = (uint32_t)(plan.passStart.size() / workers);
    threadedWaves(pool, workers, waves, waves, [&](uint32_t wave, uint32_t worker, uint32_t) {
        const size_t p = (size_t)wave * workers + worker;
        sv.warmStart(plan.passStart[p], plan.passStart[p] + plan.passSize[p]);
    });
    const uint32_t positionStart = waves * vIt;
    threadedWaves(pool, workers, waves, waves * (vIt + pIt), [&](uint32_t wave, uint32_t worker, uint32_t tick) {
        const size_t p = (size_t)wave * workers + worker;
        const size_t lo = plan.passStart[p], hi = lo + plan.passSize[p];
        if (lo == hi) return;
        if (tick >= positionStart) sv.position(lo, hi);
        else sv.velocity(lo, hi);
    });
    sv.finish();
}

// NarrowPhaseCollisionDetector::findCollisions (:101-151) with its multi-threaded path.
void narrowPhaseThreaded(World& w, const uint32_t* pa, const uint32_t* pb, size_t n, std::vector<Manifold>& out, Pool& pool) {
    const uint32_t participants =
        n > MULTITHREAD_PAIR_THRESHOLD ? (uint32_t)std::max<size_t>(std::min<size_t>(n / PAIRS_PER_WORKER, pool.size()), 1) : 1;
    if (participants <= 1) {
        narrowPhase(w, pa, pb, n, out);
        return;
    }
    const size_t chunks = (n + CHUNK_PAIRS - 1) / CHUNK_PAIRS;
    std::vector<std::vector<Manifold>> results(chunks);
    std::atomic<size_t> nextChunk{ 0 };
    pool.run((uint32_t)std::min<size_t>(participants, chunks), [&](uint32_t) {
        for (size_t c; (c = nextChunk.fetch_add(1, std::memory_order_relaxed)) < chunks;) {
            const size_t lo = c * CHUNK_PAIRS, hi = std::min(n, lo + CHUNK_PAIRS);
            narrowPhase(w, pa + lo, pb + lo, hi - lo, results[c]);
        }
    });
    out.clear();
    for (const auto& r : results) out.insert(out.end(), r.begin(), r.end());
}

}  // namespace orc
