// Autotuner: times the candidate kernel schedules of a grid size on the caller's own
// workload (one density! per repetition) and keeps the fastest in the process-wide
// schedule table that the launchers consult.  Candidates: for 3-D grids the xy-slab form
// (generic, compile-time with 256 or 512 threads, shuffle-exchange) x the z-pencil form
// (generic, compile-time LDS, register-resident, two threads per column), and in batches
// the number of streams per lane group; for 1-D/2-D grids the generic or compile-time
// planar kernel (2-D planes with 256, 512 or 1024 threads).
#include <algorithm>
#include <functional>
#include <limits>

#include "pf_batch.h"

namespace pf {

static double time_ms(hipStream_t st, const std::function<void()>& run, int reps) {
    hipEvent_t a, b;
    PF_HIP(hipEventCreate(&a));
    PF_HIP(hipEventCreate(&b));
    run();                                       // warm-up (and any rebuild)
    PF_HIP(hipStreamSynchronize(st));
    PF_HIP(hipEventRecord(a, st));
    for (int r = 0; r < reps; ++r) run();
    PF_HIP(hipEventRecord(b, st));
    PF_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    PF_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

static std::vector<Schedule> candidates(int elem, const int* dims, bool batch) {
    std::vector<Schedule> out;
    const int nz = dims[2];
    if (nz > 1) {
        // the pair pencil exists for FP32 z lines of 64, the shuffle-exchange slab for FP32
        // 64 x 64 planes, the one-row LDS pencils (256 and 512 threads) for 64 x 64 x 64, the
        // 512-thread LDS pencil for 128 x 128 x 128, and the 1024-thread slab and pencil for
        // FP32 128 x 128 x 128
        const bool f32 = elem == 4;
        const bool p64 = f32 && dims[0] == 64 && dims[1] == 64;
        const bool p128 = f32 && dims[0] == 128 && dims[1] == 128;
        std::vector<int> pencils = {0, 1, 2};
        if (f32 && nz == 64) pencils.push_back(3);
        if (nz == 64 && dims[0] == 64) {
            pencils.push_back(4);
            pencils.push_back(5);
        }
        if ((nz == 64 || nz == 128) && dims[0] == nz) pencils.push_back(6);
        if (nz == 128 && dims[0] == 128) {
            pencils.push_back(4);
            if (f32) pencils.push_back(5);
        }
        std::vector<int> slabs = {0, 1, 2};
        if (p64) slabs.push_back(3);
        if (p128) slabs.push_back(4);
        for (int sl : slabs)
            for (int pe : pencils)
                for (int sp : batch ? std::vector<int>{1, 2} : std::vector<int>{-1}) {
                    Schedule c;
                    c.slab = sl;
                    c.pencil = pe;
                    c.split = sp;
                    out.push_back(c);
                }
    } else {
        for (int pl : {0, 1, 2, 3}) {
            Schedule c;
            c.planar = pl;
            out.push_back(c);
        }
    }
    return out;
}

// Tune the schedule of (elem, dims) with `run` timing one density!; returns the best.
// Every candidate is timed once; the four fastest are then timed again in interleaved
// rounds and ranked by their best time, so that a few percent of run-to-run noise does not
// decide between close candidates.
static Schedule tune(int elem, const int* dims, hipStream_t st, bool batch,
                     const std::function<void()>& run, const std::function<void()>& changed,
                     int reps, double* best_ms) {
    Schedule& sc = schedule(elem, dims[0], dims[1], dims[2]);
    const std::vector<Schedule> cand = candidates(elem, dims, batch);
    std::vector<double> t(cand.size(), std::numeric_limits<double>::infinity());
    auto measure = [&](size_t i) {
        sc = cand[i];
        changed();
        t[i] = std::min(t[i], time_ms(st, run, reps));
    };
    for (size_t i = 0; i < cand.size(); ++i) measure(i);
    std::vector<size_t> order(cand.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return t[x] < t[y]; });
    order.resize(std::min<size_t>(order.size(), 4));
    if (order.size() > 1)
        for (int round = 0; round < 2; ++round)
            for (size_t i : order) measure(i);
    size_t bi = order.empty() ? 0 : order[0];
    for (size_t i : order)
        if (t[i] < t[bi]) bi = i;
    const Schedule best = cand.empty() ? sc : cand[bi];
    sc = best;
    changed();
    if (best_ms) *best_ms = cand.empty() ? 0.0 : t[bi];
    return best;
}

Schedule autotune_system(System& s, int reps, double* best_ms) {
    const int elem = s.precision ? 4 : 8;
    return tune(elem, s.dims, s.stream, false, [&] { s.density(); }, [] {}, reps, best_ms);
}

void autotune_batch(Batch& b, int reps, std::vector<std::pair<std::array<int, 4>, Schedule>>& out) {
    std::vector<int> all(b.sys.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
    const int elem = b.precision ? 4 : 8;
    std::vector<std::array<int, 3>> sizes;
    for (System* s : b.sys) {
        std::array<int, 3> d = {s->dims[0], s->dims[1], s->dims[2]};
        if (std::find(sizes.begin(), sizes.end(), d) == sizes.end()) sizes.push_back(d);
    }
    // timed as the batch runs: captured and replayed when graphs are on (the first run of
    // each candidate captures)
    auto changed = [&] {
        PF_HIP(hipStreamSynchronize(b.stream));
        b.drop_graphs();
        b.eng->invalidate();
    };
    for (const auto& d : sizes) {
        double ms = 0.0;
        const Schedule s = tune(elem, d.data(), b.stream, true, [&] { b.density(all); }, changed,
                                reps, &ms);
        out.push_back({{elem, d[0], d[1], d[2]}, s});
    }
    changed();
}

}  // namespace pf
