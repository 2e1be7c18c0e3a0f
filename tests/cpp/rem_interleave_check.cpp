// rem_interleave_check.cpp -- CPU model check of the concurrent union-find
// unions used on the device, at the granularity of single shared-memory
// operations.  Each simulated lane runs its unions as a state machine whose
// every step is exactly one load or one atomicMin on the shared parent array;
// a random scheduler interleaves the steps of all lanes.  After all lanes
// finish, the forest's partition must equal the union of the initial forest
// and all requested pairs, parents must never exceed their node, every lane
// must terminate, and (global unites) the number of unions reporting a link
// must equal the drop in the number of roots (the component count relies on it).
//
// Memory models for loads (atomicMin always acts on the latest value):
//   seq-cst    every load returns the latest value (LDS within a workgroup);
//   coherent   a load may return an earlier value of its location, but never
//              one older than the lane has already seen there (relaxed atomic
//              loads of global memory);
//   cached     as coherent for device-coherent loads; a plain load may return
//              ANY earlier value of its location (a line kept in a non-coherent
//              cache for the whole kernel, after the lane saw newer values).
//
// Variants:
//   lds     the tile kernels' loop: lds_rem_step (Rem's algorithm with
//           splicing, uf_device.hpp) until done.
//   lds-top the same, but a union whose node is the previous union's same
//           node starts there from the node where the previous union met.
//   lds-meet  the same, then both ORIGINAL nodes pointed at the node where the
//           sets met (removed: it can cut a set off, see the recorded schedule
//           below; random trials rarely reach it).
//   rem     the global-forest unite that was tried and removed (Rem's
//           splicing, the meeting node chased to its root, the starting nodes
//           pointed at it when their first-read parents are larger).
//   find    the global unite before cached loads: finds that chase by
//           device-coherent loads and point their start at the root found;
//           link the larger root under the smaller by atomicMin; a root that
//           was linked meanwhile is found again.
//   cfind   the merges' unite (PackedForest::unite_cached): find chasing by
//           plain (cacheable) loads; the root check is a device-coherent load,
//           and a node found not to be a root continues from its fresh parent
//           instead of being found again (which could see the same stale root
//           forever).
// Trials: random forests (parent <= self) of 4..12 nodes with 2..6 lanes of
// 1..3 unions, and of 16..44 nodes with 8..24 lanes of 1..2 unions; the LDS
// loops' unions start at any node (a run's node may have been moved by
// another lane's union before its own contact is processed), and half of a
// lane's later unions share a node with its union before.
// Usage: rem_interleave_check [trials]; exit code 0 = the variants the device
// uses are exact under their memory model (lds: seq-cst; cfind: all three)
// and lds-meet loses a union on the recorded schedule.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

namespace {

constexpr int kMaxN = 48;
enum Model { kSeqCst, kCoherent, kCached };
enum Kind { kPlain, kFresh };

struct Shared {
    uint32_t par[kMaxN];
    int n = 0;
    Model model = kSeqCst;
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
    uint32_t load(uint32_t i, size_t &seen, Kind kind)
    {
        const size_t last = hist[i].size() - 1;
        size_t v = last;
        if (model == kCached && kind == kPlain) {
            v = (*rng)() % (last + 1);   // any earlier value; does not move `seen`
            return hist[i][v];
        }
        if (model != kSeqCst && last > seen) v = seen + (*rng)() % (last - seen + 1);
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

enum Variant { kLds, kLdsTop, kLdsMeet, kRem, kFind, kCFind, kVariants };
const char *kNames[kVariants] = {"lds", "lds-top", "lds-meet", "rem", "find", "cfind"};

// One lane running a list of unions, one shared-memory operation per step().
struct Lane {
    std::vector<std::pair<uint32_t, uint32_t>> jobs;
    size_t next = 0;
    int pc = 0;   // 0 = start the next job
    uint32_t oa = 0, ob = 0, x = 0, y = 0, px = 0, py = 0, top = 0, v0x = 0, v0y = 0, chase = 0;
    uint32_t last_oa = ~0u, last_ob = ~0u, last_top = 0;   // lds-top: the previous union's nodes and meeting node
    // find sub-machine: chases from fstart, result in froot; `ret` resumes the caller
    uint32_t fstart = 0, fcur = 0, fv0 = 0, fv = 0, froot = 0, fa = 0, fb = 0;
    int ret = 0;
    int links = 0;
    size_t seen[kMaxN] = {};
    bool done() const { return next >= jobs.size() && pc == 0; }
    uint32_t ld(Shared &s, uint32_t i, Kind k) { return s.load(i, seen[i], k); }
    // end of an LDS union: lds-meet goes on to its last steps; lds-top keeps
    // the meeting node for the next union
    int done_lds(Variant var)
    {
        if (var == kLdsMeet) return 4;
        last_oa = oa;
        last_ob = ob;
        last_top = top;
        return 0;
    }
    uint32_t mn(Shared &s, uint32_t i, uint32_t v) { return s.amin(i, v, seen[i]); }

    void step(Shared &s, Variant var)
    {
        const Kind chase_kind = var == kCFind ? kPlain : kFresh;
        if (pc == 0) {
            oa = x = jobs[next].first;
            ob = y = jobs[next].second;
            ++next;
            if (var == kLdsTop) {   // a node shared with the previous union: start from where it met
                if (oa == last_oa) x = last_top;
                if (ob == last_ob) y = last_top;
            }
            pc = var == kRem ? 10 : (var == kLds || var == kLdsTop || var == kLdsMeet ? 1 : 30);
        }
        switch (pc) {
        // ---- lds
        case 1: px = ld(s, x, kFresh); pc = 2; return;
        case 2:
            py = ld(s, y, kFresh);
            if (px == py) { top = px; pc = done_lds(var); return; }
            if (px < py) { std::swap(x, y); std::swap(px, py); }
            pc = 3;
            return;
        case 3: {
            const uint32_t old = mn(s, x, py);
            if (old == x) { top = py; pc = done_lds(var); return; }
            x = old;
            pc = 1;
            return;
        }
        case 4: mn(s, oa, top); pc = 5; return;   // lds-meet only
        case 5: mn(s, ob, top); pc = 0; return;

        // ---- rem (removed global variant)
        case 10: v0x = ld(s, oa, kFresh); pc = 11; return;
        case 11: v0y = ld(s, ob, kFresh); px = v0x; py = v0y; pc = 12; [[fallthrough]];
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
        case 14: px = ld(s, x, kFresh); pc = 16; return;
        case 16: py = ld(s, y, kFresh); pc = 12; return;
        case 15: chase = ld(s, top, kFresh); pc = 17; return;
        case 17:
            if (chase != top) { top = chase; pc = 15; return; }
            pc = 18;
            return;
        case 18: if (v0x > top) mn(s, oa, top); pc = 19; return;
        case 19: if (v0y > top) mn(s, ob, top); pc = 0; return;

        // ---- find / cfind: unite(x, y)
        case 30: fstart = x; ret = 31; pc = 50; return step(s, var);   // a = find(x)
        case 31: fa = froot; fstart = y; ret = 32; pc = 50; return step(s, var);   // b = find(y)
        case 32:
            fb = froot;
            if (fa == fb) { pc = 0; return; }
            if (fa < fb) std::swap(fa, fb);
            pc = 33;
            return;
        case 33: {   // is a still a root?  (device-coherent load)
            const uint32_t va = ld(s, fa, kFresh);
            if (va != fa) {
                if (var == kCFind) { x = va; y = fb; pc = 30; return; }   // continue from the fresh parent
                x = fa; y = fb; pc = 30;   // find a again
                return;
            }
            pc = 34;
            return;
        }
        case 34: {
            const uint32_t old = mn(s, fa, fb);
            if (old == fa) { ++links; pc = 0; return; }
            x = old;
            y = fb;
            pc = 30;
            return;
        }
        // find(fstart) as PackedForest::find: v0 = w[x0]; chase x -> p -> gp;
        // then point x0 at the root unless it is x0 or its first-read parent
        case 50: fcur = fstart; fv0 = fv = ld(s, fstart, chase_kind); pc = 51; return;
        case 51: {
            const uint32_t p = fv;
            if (p == fcur) { froot = fcur; pc = 54; return; }
            const uint32_t vp = ld(s, p, chase_kind);
            if (vp == p) { froot = p; pc = 54; return; }
            fcur = vp;
            pc = 52;
            return;
        }
        case 52: fv = ld(s, fcur, chase_kind); pc = 51; return;
        case 54:
            if (froot != fstart && froot != fv0) mn(s, fstart, froot);
            pc = ret;
            return;
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

Result run_trial(std::mt19937_64 &rng, Variant var, Model model, int n, int lanes, int jobs_per_lane,
                 std::string *trace)
{
    Shared s{};
    // initial forest with parent <= self (tile-local structure)
    uint32_t init[kMaxN];
    for (int i = 0; i < n; ++i) init[i] = (i > 0 && rng() % 3 == 0) ? (uint32_t)(rng() % i) : (uint32_t)i;
    s.init(init, n);
    s.model = model;
    s.rng = &rng;
    uint32_t ref[kMaxN];
    std::copy(init, init + n, ref);
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
            if (var == kRem || (var != kLds && var != kLdsTop && var != kLdsMeet && rng() % 2)) {   // unions start at tile-local roots
                a = root_of(init, a);
                b = root_of(init, b);
            }
            // a lane's contacts often share a node with the one before (a run
            // touching several runs of the band above, or the reverse)
            if (j > 0 && (var == kLds || var == kLdsTop || var == kLdsMeet) && rng() % 2) {
                if (rng() % 2) a = l.jobs.back().first;
                else b = l.jobs.back().second;
            }
            l.jobs.push_back({a, b});
            const uint32_t ra = rfind(a), rb = rfind(b);
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
        if (++steps > 200000) {
            r.terminated = false;
            break;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if ((root_of(s.par, i) == root_of(s.par, j)) != (rfind(i) == rfind(j))) r.partition_ok = false;
    if (var != kLds && var != kLdsTop && var != kLdsMeet) {
        int links = 0, roots1 = 0;
        for (auto &l : L) links += l.links;
        for (int i = 0; i < n; ++i) roots1 += s.par[i] == (uint32_t)i;
        r.count_ok = roots0 - links == roots1;
    }
    return r;
}

// A schedule on which lds-meet loses a union (23 nodes, 12 lanes of one union
// each; step k runs lane kSchedule[k]).  Lane 0 (13,7) stands at (5,11), the
// only pending join of the set rooted at 1 with the set rooted at 0, when lane
// 1 (5,17), done (its splice put 5 under 1, then 2 and 17 both had parent 0),
// points its original node 5 at 0: lane 0 then finds 5 and 11 together and
// stops, and the set rooted at 1 stays apart.
const uint32_t kReplayForest[23] = {0, 1, 2, 0, 1, 2, 4, 5, 8, 6, 6, 9, 11, 13, 1, 9, 14, 13, 16, 19, 14, 21, 11};
const uint32_t kReplayPairs[12][2] = {{13, 7}, {5, 17}, {5, 12}, {12, 5},  {14, 17}, {11, 17},
                                      {21, 1}, {1, 20}, {19, 5}, {11, 3}, {17, 17}, {17, 22}};
const int kSchedule[] = {11, 2, 9, 3, 0, 3, 8, 9, 6, 6, 5, 8, 0, 11, 4, 11, 3, 10, 1, 3, 3, 10, 3, 11, 8, 5, 7,
                         11, 5, 4, 4, 4, 8, 8, 9, 1, 6, 3, 3, 6, 11, 6, 0, 11, 9, 4, 7, 0, 5, 4, 1, 5, 4, 9,
                         11, 3, 0, 5, 7, 4, 2, 3, 3, 5, 9, 4, 3, 2, 3, 7, 4, 1, 4, 10, 4, 4, 0, 7, 1, 7, 7, 10,
                         0, 9, 1, 1, 0, 2, 9, 0, 0, 0, 0, 0};

// Runs the lanes of the recorded case on the schedule (then any lane still
// running to its end); true when the final partition is the expected one.
bool replay(Variant var)
{
    constexpr int n = 23, lanes = 12;
    Shared s{};
    std::mt19937_64 rng(1);
    s.init(kReplayForest, n);
    s.rng = &rng;
    uint32_t ref[n];
    std::copy(kReplayForest, kReplayForest + n, ref);
    auto rfind = [&](uint32_t a) {
        while (ref[a] != a) a = ref[a];
        return a;
    };
    std::vector<Lane> L(lanes);
    for (int i = 0; i < lanes; ++i) {
        L[i].jobs.push_back({kReplayPairs[i][0], kReplayPairs[i][1]});
        const uint32_t ra = rfind(kReplayPairs[i][0]), rb = rfind(kReplayPairs[i][1]);
        if (ra != rb) ref[std::max(ra, rb)] = std::min(ra, rb);
    }
    for (const int k : kSchedule)
        if (!L[k].done()) L[k].step(s, var);
    for (auto &l : L)
        while (!l.done()) l.step(s, var);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if ((root_of(s.par, i) == root_of(s.par, j)) != (rfind(i) == rfind(j))) return false;
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    const long trials = argc > 1 ? std::atol(argv[1]) : 300000;
    const char *models[] = {"seq-cst", "coherent", "cached"};
    long fails[3][kVariants][4] = {};
    for (int m = 0; m < 3; ++m)
        for (int v = 0; v < kVariants; ++v) {
            std::string example;
            std::mt19937_64 rng(1234 + v + 10 * m);
            for (long t = 0; t < trials; ++t) {
                const bool large = t % 4 == 3;
                const int n = large ? 16 + (int)(rng() % 29) : 4 + (int)(rng() % 9);
                const int lanes = large ? 8 + (int)(rng() % 17) : 2 + (int)(rng() % 5);
                const int jobs = large ? 1 + (int)(rng() % 2) : 1 + (int)(rng() % 3);
                std::mt19937_64 keep = rng;
                const Result r = run_trial(rng, (Variant)v, (Model)m, n, lanes, jobs, nullptr);
                const bool bad[4] = {!r.partition_ok, !r.invariant_ok, !r.count_ok, !r.terminated};
                for (int k = 0; k < 4; ++k) fails[m][v][k] += bad[k];
                if ((bad[0] || bad[1] || bad[2] || bad[3]) && example.empty()) {
                    std::string tr;
                    run_trial(keep, (Variant)v, (Model)m, n, lanes, jobs, &tr);
                    example = "n=" + std::to_string(n) + " lanes=" + std::to_string(lanes) + ": " + tr.substr(0, 1200);
                }
            }
            std::printf("%-5s %-8s trials %ld: lost unions %ld, parent>self %ld, link count wrong %ld, "
                        "no termination %ld\n", kNames[v], models[m], trials, fails[m][v][0], fails[m][v][1],
                        fails[m][v][2], fails[m][v][3]);
            if (!example.empty()) std::printf("  first counterexample: %s\n", example.c_str());
        }
    auto exact = [&](int m, int v) {
        return !fails[m][v][0] && !fails[m][v][1] && !fails[m][v][2] && !fails[m][v][3];
    };
    // the device's variants under the memory each one runs on
    const bool meet_lost = !replay(kLdsMeet), lds_kept = replay(kLds);
    std::printf("recorded schedule: lds-meet %s, lds %s\n", meet_lost ? "loses a union" : "exact",
                lds_kept ? "exact" : "loses a union");
    const bool ok = !fails[kSeqCst][kLds][0] && !fails[kSeqCst][kLds][1] && !fails[kSeqCst][kLds][3] &&
                    !fails[kSeqCst][kLdsTop][0] && !fails[kSeqCst][kLdsTop][1] && !fails[kSeqCst][kLdsTop][3] &&
                    exact(kSeqCst, kCFind) && exact(kCoherent, kCFind) && exact(kCached, kCFind) && meet_lost &&
                    lds_kept;
    std::printf("device variants %s\n", ok ? "exact" : "FAILED");
    return ok ? 0 : 1;
}
