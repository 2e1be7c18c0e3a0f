// rem_interleave_check.cpp -- CPU model check of the concurrent union-find
// unions used on the device, at the granularity of single shared-memory
// operations.  Each simulated lane runs its unions as a state machine whose
// every step is exactly one load or one atomicMin on the shared parent array
// (sequentially consistent, as LDS within a workgroup); a random scheduler
// interleaves the steps of all lanes.  After all lanes finish, the forest's
// partition must equal the union of the initial forest and all requested
// pairs, and parents must never exceed their node.
//
// Variants:
//   lds     the tile kernels' loop: lds_rem_step (Rem's algorithm with
//           splicing, uf_device.hpp) until done, then both ORIGINAL nodes are
//           pointed at the meeting node (atomicMin).
//   global  the global-forest unite that was tried and removed (Rem's
//           splicing, then the meeting node chased to its root, then the
//           starting nodes pointed at that root, guarded by their parents
//           as first read), and its "linked" result used for component counts.
//   find    the shipped global unite: read-only finds to both roots, link the
//           larger root under the smaller by atomicMin, retry.
// Usage: rem_interleave_check [trials]; prints a counterexample on failure.
// Exit code 0 = every variant behaved as expected (see main).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

namespace {

constexpr int kMaxN = 16;

// The shared parent array.  Sequentially consistent by default (LDS).  With
// `stale`, a load may return any earlier value of its location that is not
// older than what the loading lane has already observed there (per-location
// coherence only, what relaxed atomic loads of global memory guarantee);
// atomicMin always acts on the latest value.
struct Shared {
    uint32_t par[kMaxN];
    int n;
    bool stale = false;
    std::mt19937_64 *rng = nullptr;
    std::vector<uint32_t> hist[kMaxN];   // every value the location has held, oldest first
    void init(const uint32_t *p, int n_)
    {
        n = n_;
        for (int i = 0; i < n; ++i) {
            par[i] = p[i];
            hist[i].assign(1, p[i]);
        }
    }
    uint32_t load(uint32_t i, size_t &seen)
    {
        size_t v = hist[i].size() - 1;
        if (stale && v > seen) v = seen + (*rng)() % (v - seen + 1);
        if (v < seen) v = seen;
        seen = v;
        return hist[i][v];
    }
    uint32_t amin(uint32_t i, uint32_t v, size_t &seen)
    {
        const uint32_t old = par[i];
        if (v < old) {
            par[i] = v;
            hist[i].push_back(v);
        }
        seen = hist[i].size() - 1;
        return old;
    }
};

enum Variant { kLds, kGlobal, kFind };

// One lane running a list of unions, one shared-memory operation per step().
struct Lane {
    std::vector<std::pair<uint32_t, uint32_t>> jobs;
    size_t next = 0;
    int pc = 0;   // 0 = start next job
    uint32_t oa = 0, ob = 0, x = 0, y = 0, px = 0, py = 0, top = 0, v0x = 0, v0y = 0, chase = 0;
    int links = 0;   // "linked" results (global variant) / root links (find variant)
    bool done() const { return next >= jobs.size() && pc == 0; }

    // find variant state
    uint32_t fa = 0, fb = 0, fcur = 0;
    size_t seen[kMaxN] = {};   // newest version of each location this lane has observed
    uint32_t ld(Shared &s, uint32_t i) { return s.load(i, seen[i]); }
    uint32_t mn(Shared &s, uint32_t i, uint32_t v) { return s.amin(i, v, seen[i]); }

    void step(Shared &s, Variant var)
    {
        if (pc == 0) {
            oa = x = jobs[next].first;
            ob = y = jobs[next].second;
            ++next;
            pc = var == kGlobal ? 10 : (var == kFind ? 30 : 1);
        }
        switch (pc) {
        // ---- lds: lds_rem_step loop, then compress both original nodes to the meeting node
        case 1: px = ld(s, x); pc = 2; return;
        case 2:
            py = ld(s, y);
            if (px == py) { top = px; pc = 4; return; }
            if (px < py) { std::swap(x, y); std::swap(px, py); }
            pc = 3;
            return;
        case 3: {
            const uint32_t old = mn(s, x, py);
            if (old == x) { top = py; pc = 4; return; }
            x = old;
            pc = 1;
            return;
        }
        case 4: mn(s, oa, top); pc = 5; return;
        case 5: mn(s, ob, top); pc = 0; return;

        // ---- global (removed): initial loads, splicing loop, chase, guarded compression
        case 10: v0x = ld(s, oa); pc = 11; return;
        case 11: v0y = ld(s, ob); px = v0x; py = v0y; pc = 12; [[fallthrough]];
        case 12:
            if (px == py) { top = px; pc = 15; return; }
            if (px < py) { std::swap(x, y); std::swap(px, py); }
            pc = 13;
            return;
        case 13: {
            const uint32_t old = mn(s, x, py);
            if (old == x) { top = py; ++links; pc = 15; return; }
            x = old;
            pc = 14;
            return;
        }
        case 14: px = ld(s, x); pc = 16; return;
        case 16: py = ld(s, y); pc = 12; return;
        case 15: chase = ld(s, top); pc = 17; return;   // chase top to its root
        case 17:
            if (chase != top) { top = chase; pc = 15; return; }
            pc = 18;
            return;
        case 18: if (v0x > top) mn(s, oa, top); pc = 19; return;
        case 19: if (v0y > top) mn(s, ob, top); pc = 0; return;

        // ---- find: read-only finds (one load per step), link the larger root
        case 30: fcur = x; pc = 31; [[fallthrough]];
        case 31: {
            const uint32_t p = ld(s, fcur);
            if (p != fcur) { fcur = p; return; }
            fa = fcur;
            fcur = y;
            pc = 32;
            return;
        }
        case 32: {
            const uint32_t p = ld(s, fcur);
            if (p != fcur) { fcur = p; return; }
            fb = fcur;
            if (fa == fb) { pc = 0; return; }
            if (fa < fb) std::swap(fa, fb);
            pc = 33;
            return;
        }
        case 33: {
            const uint32_t old = mn(s, fa, fb);
            if (old == fa) { ++links; pc = 0; return; }
            x = old;   // fa was linked meanwhile: unite old and fb again
            y = fb;
            pc = 30;
            return;
        }
        default: std::abort();
        }
    }
};

uint32_t root_of(const uint32_t *p, uint32_t i)
{
    while (p[i] != i) i = p[i];
    return i;
}

struct Result {
    bool partition_ok = true, invariant_ok = true, count_ok = true, terminated = true;
};

Result run_trial(std::mt19937_64 &rng, Variant var, bool stale, int n, int lanes, int jobs_per_lane,
                 std::string *trace)
{
    Shared s{};
    // initial forest with parent <= self (tile-local structure)
    uint32_t init[kMaxN];
    for (int i = 0; i < n; ++i) init[i] = (i > 0 && rng() % 3 == 0) ? (uint32_t)(rng() % i) : (uint32_t)i;
    s.init(init, n);
    s.stale = stale;
    s.rng = &rng;
    // reference partition
    uint32_t ref[kMaxN];
    std::copy(s.par, s.par + n, ref);
    auto rfind = [&](uint32_t a) {
        while (ref[a] != a) a = ref[a];
        return a;
    };
    std::vector<Lane> L(lanes);
    int roots0 = 0;
    for (int i = 0; i < n; ++i) roots0 += init[i] == (uint32_t)i;
    for (auto &l : L)
        for (int j = 0; j < jobs_per_lane; ++j) {
            uint32_t a = rng() % n, b = rng() % n;
            if (var != kFind) {   // lanes unite nodes, which the tile kernels and merges start from
                a = root_of(init, a);
                b = root_of(init, b);
            }
            l.jobs.push_back({a, b});
            uint32_t ra = rfind(a), rb = rfind(b);
            if (ra != rb) ref[std::max(ra, rb)] = std::min(ra, rb);
        }
    Result r;
    long steps = 0;
    for (;;) {
        std::vector<int> live;
        for (int i = 0; i < lanes; ++i)
            if (!L[i].done()) live.push_back(i);
        if (live.empty()) break;
        const int k = live[rng() % live.size()];
        if (trace) {
            char buf[96];
            std::snprintf(buf, sizeof buf, "lane %d pc %d x %u y %u | ", k, L[k].pc, L[k].x, L[k].y);
            *trace += buf;
        }
        L[k].step(s, var);
        for (int i = 0; i < n; ++i)
            if (s.par[i] > (uint32_t)i) r.invariant_ok = false;
        if (++steps > 100000) {
            r.terminated = false;
            break;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if ((root_of(s.par, i) == root_of(s.par, j)) != (rfind(i) == rfind(j))) r.partition_ok = false;
    if (var != kLds) {
        int links = 0, roots1 = 0;
        for (auto &l : L) links += l.links;
        for (int i = 0; i < n; ++i) roots1 += s.par[i] == (uint32_t)i;
        r.count_ok = roots0 - links == roots1;
    }
    return r;
}

}  // namespace

int main(int argc, char **argv)
{
    const long trials = argc > 1 ? std::atol(argv[1]) : 300000;
    const char *names[] = {"lds", "global", "find"};
    long fails[2][3][4] = {};
    for (int st = 0; st < 2; ++st)
        for (int v = 0; v < 3; ++v) {
            std::string example;
            std::mt19937_64 rng(1234 + v + 10 * st);
            for (long t = 0; t < trials; ++t) {
                const int n = 4 + (int)(rng() % 9), lanes = 2 + (int)(rng() % 5), jobs = 1 + (int)(rng() % 3);
                std::mt19937_64 keep = rng;
                const Result r = run_trial(rng, (Variant)v, st != 0, n, lanes, jobs, nullptr);
                const bool bad[4] = {!r.partition_ok, !r.invariant_ok, !r.count_ok, !r.terminated};
                for (int k = 0; k < 4; ++k) fails[st][v][k] += bad[k];
                if ((bad[0] || bad[1] || bad[2] || bad[3]) && example.empty()) {
                    std::string tr;
                    run_trial(keep, (Variant)v, st != 0, n, lanes, jobs, &tr);
                    example = "n=" + std::to_string(n) + " lanes=" + std::to_string(lanes) + ": " + tr.substr(0, 1200);
                }
            }
            std::printf("%-6s %-9s trials %ld: lost unions %ld, parent>self %ld, link count wrong %ld, "
                        "no termination %ld\n", names[v], st ? "stale" : "seq-cst", trials, fails[st][v][0],
                        fails[st][v][1], fails[st][v][2], fails[st][v][3]);
            if (!example.empty()) std::printf("  first counterexample: %s\n", example.c_str());
        }
    // required: the LDS loop exact under sequential consistency (LDS), the
    // shipped global unite exact also with stale loads (global memory)
    const bool lds_ok = !fails[0][0][0] && !fails[0][0][1] && !fails[0][0][3];
    const bool find_ok = !fails[0][2][0] && !fails[0][2][1] && !fails[0][2][2] && !fails[0][2][3] &&
                         !fails[1][2][0] && !fails[1][2][1] && !fails[1][2][2] && !fails[1][2][3];
    std::printf("shipped variants %s\n", lds_ok && find_ok ? "exact" : "FAILED");
    return lds_ok && find_ok ? 0 : 1;
}
