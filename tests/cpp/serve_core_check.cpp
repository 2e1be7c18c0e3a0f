// serve_core_check.cpp -- CPU test of the server's scheduler (serve_core.hpp)
// with a simulated device backend: completion statuses under injected faults,
// launch errors and completion-query errors, admission by arena footprint,
// fusing of consecutive fusable requests, and arena accounting.
// Build: g++ -std=c++17 -pthread -I<repo> serve_core_check.cpp; exit code 0 = pass.
#include <atomic>
#include <cstdio>
#include <map>
#include <set>
#include <string>

#include "labeling_amd/csrc/serve_core.hpp"

using namespace lbl::serve;

namespace {

int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            ++failures;                                    \
            std::printf("FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
            std::printf(__VA_ARGS__);                      \
            std::printf("\n");                             \
        }                                                  \
    } while (0)

// Completes each launch after a few polls; count of a request = its ticket * 10.
class FakeBackend final : public ServeBackend {
public:
    explicit FakeBackend(uint32_t slots) : polls_(slots, 0), tickets_(slots), qerr_(slots, 0), cap_(slots, false) {}
    std::mutex mu;
    std::set<uint64_t> fail_launch;   // launches of groups holding one of these tickets fail
    std::set<uint64_t> fail_query;    // completion queries of such groups fail
    std::set<uint64_t> iter_cap_tk;   // such groups report the iteration cap
    std::vector<std::vector<uint64_t>> groups;   // tickets of each launched group
    uint64_t inflight = 0, max_inflight = 0, cap_bytes = 0;
    std::map<int, uint64_t> slot_bytes;
    int injected = 0;

    uint64_t group_extra(uint32_t n) const override { return 256ull * n; }
    int launch(Group &g, bool fail_kernels) override
    {
        std::lock_guard<std::mutex> lk(mu);
        std::vector<uint64_t> tk;
        bool fail = false, qfail = false, cap = false;
        for (Req *r : g.reqs) {
            tk.push_back(r->ticket);
            fail |= fail_launch.count(r->ticket) != 0;
            qfail |= fail_query.count(r->ticket) != 0;
            cap |= iter_cap_tk.count(r->ticket) != 0;
        }
        groups.push_back(tk);
        slot_bytes[g.slot] = g.bytes;
        inflight += g.bytes;
        max_inflight = std::max(max_inflight, inflight);
        polls_[g.slot] = 3;
        tickets_[g.slot] = tk;
        qerr_[g.slot] = qfail ? 77 : 0;
        cap_[g.slot] = cap;
        if (fail_kernels) {
            ++injected;
            return 55;
        }
        return fail ? 66 : 0;
    }
    int query(int slot) override
    {
        std::lock_guard<std::mutex> lk(mu);
        if (polls_[slot] > 0) {
            --polls_[slot];
            return kNotReady;
        }
        inflight -= slot_bytes[slot];
        slot_bytes[slot] = 0;
        return qerr_[slot];
    }
    void wait(int slot) override
    {
        std::lock_guard<std::mutex> lk(mu);
        polls_[slot] = 0;
        inflight -= slot_bytes[slot];
        slot_bytes[slot] = 0;
    }
    uint32_t count(int slot, uint32_t i) const override { return (uint32_t)(tickets_[slot][i] * 10); }
    bool iter_cap(int slot, uint32_t) const override { return cap_[slot]; }
    std::string describe(int err) const override { return "fake error " + std::to_string(err); }

private:
    std::vector<int> polls_;
    std::vector<std::vector<uint64_t>> tickets_;
    std::vector<int> qerr_;
    std::vector<bool> cap_;
};

Req *make(uint64_t bytes, bool fusable)
{
    Req *r = new Req;
    r->bytes = bytes;
    r->fusable = fusable;
    return r;
}

std::map<uint64_t, lbl_completion> collect(Scheduler &s, size_t n)
{
    std::map<uint64_t, lbl_completion> out;
    s.drain();
    lbl_completion buf[64];
    uint32_t k = 0;
    do {
        s.poll(buf, 64, &k);
        for (uint32_t i = 0; i < k; ++i) out[buf[i].ticket] = buf[i];
    } while (k);
    CHECK(out.size() == n, "%zu completions, expected %zu", out.size(), n);
    return out;
}

// Every request completes OK with its count; fused groups hold at most
// max_batch consecutive fusable requests; the arena is whole again.
void test_plain_and_fusing()
{
    FakeBackend be(2);
    Scheduler s(be, 2, 1 << 20, 4);
    s.start();
    std::vector<uint64_t> tk;
    for (int i = 0; i < 10; ++i) {
        uint64_t t = 0;
        CHECK(s.submit(make(1000, i != 6), &t) == LBL_OK, "submit %d", i);   // request 6 is not fusable
        tk.push_back(t);
    }
    auto done = collect(s, 10);
    for (uint64_t t : tk) {
        CHECK(done[t].status == LBL_OK, "ticket %llu status %d", (unsigned long long)t, done[t].status);
        CHECK(done[t].n_components == t * 10, "ticket %llu count %u", (unsigned long long)t, done[t].n_components);
        CHECK(done[t].t_submit <= done[t].t_start && done[t].t_start <= done[t].t_done, "times");
    }
    size_t seen = 0;
    for (const auto &g : be.groups) {
        CHECK(g.size() <= 4, "group of %zu > max_batch", g.size());
        for (size_t i = 1; i < g.size(); ++i) CHECK(g[i] == g[i - 1] + 1, "fused requests are consecutive");
        for (uint64_t t : g) CHECK(t != tk[6] || g.size() == 1, "a non-fusable request runs alone");
        seen += g.size();
    }
    CHECK(seen == 10, "every request launched once (%zu)", seen);
    CHECK(s.arena_free() == s.arena_capacity(), "arena released: %llu of %llu", (unsigned long long)s.arena_free(),
          (unsigned long long)s.arena_capacity());
    s.stop();
}

// Injected faults: the next group fails at its kernel launch; its requests
// complete with LBL_EDEVICE, later groups run normally.
void test_injected_fault()
{
    FakeBackend be(1);
    Scheduler s(be, 1, 1 << 20, 3);
    s.inject_faults(1);
    s.start();
    std::vector<uint64_t> tk;
    for (int i = 0; i < 6; ++i) {
        uint64_t t = 0;
        CHECK(s.submit(make(500, true), &t) == LBL_OK, "submit");
        tk.push_back(t);
    }
    auto done = collect(s, 6);
    CHECK(be.injected == 1, "one injected launch failure (%d)", be.injected);
    const auto &first = be.groups.at(0);
    for (uint64_t t : tk) {
        const bool in_first = std::find(first.begin(), first.end(), t) != first.end();
        CHECK(done[t].status == (in_first ? LBL_EDEVICE : LBL_OK), "ticket %llu status %d", (unsigned long long)t,
              done[t].status);
        if (in_first) CHECK(done[t].n_components == 0, "failed request reports no count");
    }
    CHECK(s.last_enqueue_error() == "fake error 55", "message: %s", s.last_enqueue_error().c_str());
    CHECK(s.fatal().empty(), "an enqueue failure is not fatal");
    CHECK(s.arena_free() == s.arena_capacity(), "arena released after a failed group");
    s.stop();
}

// A launch error of the backend itself and a completion-query error.
void test_launch_and_query_errors()
{
    FakeBackend be(2);
    be.fail_launch.insert(3);
    be.iter_cap_tk.insert(5);
    Scheduler s(be, 2, 1 << 20, 1);
    s.start();
    for (int i = 0; i < 6; ++i) {
        uint64_t t = 0;
        CHECK(s.submit(make(100, false), &t) == LBL_OK, "submit");
    }
    auto done = collect(s, 6);
    for (uint64_t t = 1; t <= 6; ++t) {
        const lbl_status want = t == 3 ? LBL_EDEVICE : (t == 5 ? LBL_EITER : LBL_OK);
        CHECK(done[t].status == want, "ticket %llu status %d want %d", (unsigned long long)t, done[t].status, want);
    }
    s.stop();

    FakeBackend be2(1);
    be2.fail_query.insert(2);
    Scheduler s2(be2, 1, 1 << 20, 1);
    s2.start();
    uint64_t t = 0;
    CHECK(s2.submit(make(100, false), &t) == LBL_OK && t == 1, "submit 1");
    CHECK(s2.submit(make(100, false), &t) == LBL_OK && t == 2, "submit 2");
    s2.drain();
    lbl_completion buf[4];
    uint32_t k = 0;
    CHECK(s2.poll(buf, 4, &k) == LBL_EDEVICE, "poll reports the device failure");
    CHECK(k == 2, "both completed (%u)", k);
    for (uint32_t i = 0; i < k; ++i)
        CHECK(buf[i].status == (buf[i].ticket == 2 ? LBL_EDEVICE : LBL_OK), "ticket %llu status %d",
              (unsigned long long)buf[i].ticket, buf[i].status);
    CHECK(s2.submit(make(100, false), &t) == LBL_EDEVICE, "no submissions after a device failure");
    CHECK(!s2.fatal().empty(), "fatal message set");
    s2.stop();
}

// Admission by footprint: a request that can never fit is refused; the bytes in
// flight never exceed the arena; requests wait for memory in FIFO order.
void test_admission()
{
    FakeBackend be(4);
    const uint64_t cap = 10000;
    Scheduler s(be, 4, cap, 8);
    s.start();
    uint64_t t = 0;
    CHECK(s.submit(make(cap, false), &t) == LBL_ENOMEM, "too large for the arena");
    for (int i = 0; i < 12; ++i) CHECK(s.submit(make(3000, i % 3 != 0), &t) == LBL_OK, "submit %d", i);
    auto done = collect(s, 12);
    for (const auto &kv : done) CHECK(kv.second.status == LBL_OK, "status");
    CHECK(be.max_inflight <= cap, "in flight %llu > arena %llu", (unsigned long long)be.max_inflight,
          (unsigned long long)cap);
    uint64_t prev = 0;
    for (const auto &g : be.groups) {
        CHECK(g.front() > prev, "FIFO order of admission");
        prev = g.back();
    }
    CHECK(s.arena_free() == cap, "arena released");
    s.stop();
}

void test_arena()
{
    Arena a;
    a.init(1000);
    uint64_t o1, o2, o3, o4;
    CHECK(a.alloc(300, o1) && o1 == 0, "first fit");
    CHECK(a.alloc(300, o2) && o2 == 300, "second");
    CHECK(a.alloc(300, o3) && o3 == 600, "third");
    CHECK(!a.alloc(200, o4), "no room");
    a.release(o2, 300);
    CHECK(a.largest() == 300, "hole");
    a.release(o1, 300);
    CHECK(a.largest() == 600, "coalesced left");
    a.release(o3, 300);
    CHECK(a.largest() == 1000 && a.free_bytes() == 1000, "coalesced whole");
}

}  // namespace

int main()
{
    test_arena();
    test_plain_and_fusing();
    test_injected_fault();
    test_launch_and_query_errors();
    test_admission();
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
