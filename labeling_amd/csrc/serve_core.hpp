// serve_core.hpp -- the device-independent half of the continuous-batching
// server: request queue, admission by device footprint, grouping of fusable
// requests, completion bookkeeping and error reporting.  The device half
// (copies, kernels, completion events) sits behind ServeBackend; serve.cpp
// implements it with HIP streams, and a CPU test drives this core with a
// simulated backend.  No reference counterpart: the reference labels one image
// per Label() call (main.cpp:309-340).
//
// One scheduler thread per server.  Each loop iteration
//   1. retires groups whose slot reports completion (their arena space and
//      slot are released, one lbl_completion per request),
//   2. admits requests from the FIFO while a slot is free and the arena has
//      room: a run of consecutive fusable requests (up to max_batch) becomes
//      one group, any other request is a group of its own.
// Status of a finished request: LBL_EDEVICE if anything of its group failed to
// enqueue or its completion query failed, else LBL_EITER if a kernel reported
// the iteration cap, else LBL_OK.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/labeling_amd.h"

namespace lbl {
namespace serve {

// First-fit allocator over [0, cap) with coalescing of freed neighbours.
class Arena {
public:
    void init(uint64_t cap)
    {
        free_.clear();
        cap_ = cap;
        if (cap) free_[0] = cap;
    }
    uint64_t largest() const
    {
        uint64_t m = 0;
        for (const auto &kv : free_) m = std::max(m, kv.second);
        return m;
    }
    uint64_t free_bytes() const
    {
        uint64_t s = 0;
        for (const auto &kv : free_) s += kv.second;
        return s;
    }
    uint64_t capacity() const { return cap_; }
    bool alloc(uint64_t bytes, uint64_t &off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rem = it->second - bytes, noff = it->first + bytes;
            free_.erase(it);
            if (rem) free_[noff] = rem;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        if (!bytes) return;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<uint64_t, uint64_t> free_;   // offset -> length
    uint64_t cap_ = 0;
};

struct Req {
    virtual ~Req() = default;   // backends extend it with their launch settings
    uint64_t ticket = 0;
    lbl_request r{};
    lbl_label_mode mode = LBL_LABELS_ROOT;   // the context's label mode at submission
    uint64_t bytes = 0;                      // device footprint (input + labels + workspace)
    bool fusable = false;                    // may share one launch with neighbouring fusable requests
    double t_submit = 0, t_start = 0;
};

struct Group {
    std::vector<Req *> reqs;
    int slot = -1;
    uint64_t off = 0, bytes = 0;   // arena range
    int err = 0;                   // first enqueue error (backend code), 0 = none
};

// The device half.  Error codes are the backend's own (nonzero = failure).
class ServeBackend {
public:
    static constexpr int kNotReady = -1;
    virtual ~ServeBackend() = default;
    // Called once on the scheduler thread before anything else.
    virtual void bind_thread() {}
    // Arena bytes a group of n requests needs beyond the requests' own bytes.
    virtual uint64_t group_extra(uint32_t n) const = 0;
    // Enqueue the whole group on slot g.slot (copies in, kernels, copies out,
    // completion marker).  fail_kernels: fault injection, the kernel launches
    // fail before they are enqueued (what a failing launch returns).  Returns
    // 0 or the first error; nothing more of the group is enqueued after it.
    virtual int launch(Group &g, bool fail_kernels) = 0;
    // 0 when the slot's work has finished, kNotReady, or an error code.
    virtual int query(int slot) = 0;
    // Block until everything enqueued on the slot has finished.
    virtual void wait(int slot) = 0;
    // After completion: component count of request i, and whether a kernel of
    // the group reported the iteration cap.
    virtual uint32_t count(int slot, uint32_t i) const = 0;
    virtual bool iter_cap(int slot, uint32_t n) const = 0;
    // Text of a backend error code.
    virtual std::string describe(int err) const = 0;
};

inline lbl_status completion_status(int enqueue_err, int query_err, bool iter_cap)
{
    if (enqueue_err || query_err) return LBL_EDEVICE;
    return iter_cap ? LBL_EITER : LBL_OK;
}

class Scheduler {
public:
    Scheduler(ServeBackend &be, uint32_t slots, uint64_t arena_bytes, uint32_t max_batch)
        : be_(be), slots_(slots, nullptr), max_batch_(max_batch), t0_(std::chrono::steady_clock::now())
    {
        arena_.init(arena_bytes);
    }
    ~Scheduler() { stop(); }

    void start()
    {
        th_ = std::thread([this] { loop(); });
    }
    // Stop the scheduler thread once every queued request has finished.
    void stop()
    {
        if (!th_.joinable()) return;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        th_.join();
    }

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }
    uint64_t arena_capacity() const { return arena_.capacity(); }
    uint64_t arena_free() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return arena_.free_bytes();
    }

    // Queue a request (ownership passes to the scheduler).  LBL_ENOMEM when it
    // can never fit the arena, LBL_EDEVICE once the device has failed.
    lbl_status submit(Req *r, uint64_t *ticket)
    {
        if (r->bytes + be_.group_extra(1) > arena_.capacity()) {
            delete r;
            return LBL_ENOMEM;
        }
        std::lock_guard<std::mutex> lk(mu_);
        if (!fatal_.empty()) {
            delete r;
            return LBL_EDEVICE;
        }
        r->ticket = next_ticket_++;
        r->t_submit = now();
        *ticket = r->ticket;
        pending_.push_back(r);
        ++submitted_;
        cv_work_.notify_one();
        return LBL_OK;
    }

    // Move up to cap completions into out; LBL_EDEVICE once a completion query
    // has failed (the device is then unusable).
    lbl_status poll(lbl_completion *out, uint32_t cap, uint32_t *n)
    {
        std::lock_guard<std::mutex> lk(mu_);
        const uint32_t k = (uint32_t)std::min<size_t>(cap, finished_.size());
        std::copy(finished_.begin(), finished_.begin() + k, out);
        finished_.erase(finished_.begin(), finished_.begin() + k);
        *n = k;
        return fatal_.empty() ? LBL_OK : LBL_EDEVICE;
    }

    lbl_status drain()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [this] { return completed_ == submitted_; });
        return fatal_.empty() ? LBL_OK : LBL_EDEVICE;
    }

    // Fault injection: the next n groups launched fail at their kernel launch.
    void inject_faults(uint32_t n)
    {
        std::lock_guard<std::mutex> lk(mu_);
        inject_ += n;
    }

    std::string fatal() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return fatal_;
    }
    // Message of the most recent group that failed to enqueue ("" if none).
    std::string last_enqueue_error() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return last_enqueue_error_;
    }

    // Requests still owned by the scheduler (queued or in flight); for cleanup
    // after stop().
    void release_all()
    {
        for (Req *r : pending_) delete r;
        pending_.clear();
        for (Group *&g : slots_) {
            if (!g) continue;
            for (Req *r : g->reqs) delete r;
            delete g;
            g = nullptr;
        }
    }

private:
    bool admit_one()
    {
        if (pending_.empty()) return false;
        int slot = -1;
        for (size_t i = 0; i < slots_.size(); ++i)
            if (!slots_[i]) {
                slot = (int)i;
                break;
            }
        if (slot < 0) return false;
        const uint64_t room = arena_.largest();
        Group *g = new Group;
        uint64_t bytes = 0;
        const bool fuse = pending_.front()->fusable;
        for (size_t k = 0; k < pending_.size(); ++k) {
            Req *r = pending_[k];
            if (k > 0 && (!fuse || !r->fusable || g->reqs.size() >= max_batch_)) break;
            const uint64_t nb = bytes + r->bytes;
            if (nb + be_.group_extra((uint32_t)g->reqs.size() + 1) > room) break;
            bytes = nb;
            g->reqs.push_back(r);
        }
        if (g->reqs.empty()) {
            delete g;
            return false;   // the head of the queue waits for memory
        }
        g->bytes = bytes + be_.group_extra((uint32_t)g->reqs.size());
        if (!arena_.alloc(g->bytes, g->off)) {
            delete g;
            return false;
        }
        pending_.erase(pending_.begin(), pending_.begin() + (long)g->reqs.size());
        g->slot = slot;
        slots_[slot] = g;
        const double t = now();
        for (Req *r : g->reqs) r->t_start = t;
        const bool fail = inject_ > 0;
        if (fail) --inject_;
        g->err = be_.launch(*g, fail);
        if (g->err) {
            // nothing more of the group was enqueued: wait for what was, then
            // retire it at once with its error
            last_enqueue_error_ = be_.describe(g->err);
            be_.wait(slot);
            retire(slot, 0);
        }
        return true;
    }

    void retire(int slot, int query_err)
    {
        Group *g = slots_[slot];
        const double t = now();
        const uint32_t n = (uint32_t)g->reqs.size();
        const bool ok_run = !g->err && !query_err;
        const bool cap = ok_run && be_.iter_cap(slot, n);
        const lbl_status st = completion_status(g->err, query_err, cap);
        for (uint32_t i = 0; i < n; ++i) {
            Req *r = g->reqs[i];
            finished_.push_back(lbl_completion{r->ticket, st, ok_run ? be_.count(slot, i) : 0u, r->t_submit,
                                               r->t_start, t});
            delete r;
        }
        completed_ += n;
        arena_.release(g->off, g->bytes);
        slots_[slot] = nullptr;
        delete g;
        cv_done_.notify_all();
    }

    void loop()
    {
        be_.bind_thread();
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            bool progressed = false;
            for (size_t s = 0; s < slots_.size(); ++s) {
                if (!slots_[s]) continue;
                const int q = be_.query((int)s);
                if (q == ServeBackend::kNotReady) continue;
                if (q != 0 && fatal_.empty()) fatal_ = be_.describe(q);
                retire((int)s, q);
                progressed = true;
            }
            while (admit_one()) progressed = true;
            bool busy = false;
            for (Group *g : slots_) busy |= g != nullptr;
            if (stop_ && pending_.empty() && !busy) break;
            if (!progressed) {
                if (!busy && pending_.empty()) {
                    cv_work_.wait(lk);
                } else {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                    lk.lock();
                }
            }
        }
    }

    ServeBackend &be_;
    std::vector<Group *> slots_;   // group in flight per slot
    uint32_t max_batch_;
    std::chrono::steady_clock::time_point t0_;
    Arena arena_;
    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Req *> pending_;
    std::vector<lbl_completion> finished_;
    uint64_t next_ticket_ = 1, submitted_ = 0, completed_ = 0;
    uint32_t inject_ = 0;
    bool stop_ = false;
    std::string fatal_, last_enqueue_error_;
    std::thread th_;
};

}  // namespace serve
}  // namespace lbl
