// Claim counters of a plan, handed to launches so that no two launches that may run at
// the same time draw blocks from the same counters.
//
// A persistent kernel draws its blocks from a counter in device memory (claim_block in
// kernels.hpp) and hands it back at zero when the last ticket is drawn.  Two launches
// drawing from one counter at once would skip or repeat blocks, so a plan owns three
// regions of kSlots counter sets (kClaimWords unsigned each, one set per launch) and
// assigns them as follows:
//
//   region 0  while every launch of the plan so far went to one stream.  A stream runs its
//             launches one after another, so when launch n starts, launch n - kSlots (the
//             last user of slot n % kSlots) has finished.  No event, no wait.
//   region 1  from the first launch on a second stream on, for the plan's lifetime.  Each
//             launch records an event after itself; the next user of its slot makes its own
//             stream wait for that event (a device-side wait, skipped when the event has
//             already completed).  Launches of region 0 that may still run never share a
//             counter with these, and no stream but the caller's is touched.
//   region 2  launches captured into a graph: each capture keeps a counter of its own for
//             the plan's lifetime (a graph exec runs one replay at a time).  After kSlots
//             captures the next one is refused.
//
// RT supplies the runtime calls (HIP in the library; a recording fake in the CPU test):
//   Stream, Event, Error, ok, busy (the error returned when region 2 is exhausted),
//   capturing(s, &yes), event_create(&e), event_done(e), stream_wait(s, e),
//   event_record(e, s), event_destroy(e).
#pragma once
#include <stddef.h>

#include <mutex>

#include "claim_layout.hpp"

namespace fx {

template <class RT>
class ClaimSlotsT {
public:
    static constexpr int kSlots = 64;
    static constexpr int kLine = kClaimWords;  // unsigned per launch (claim_layout.hpp)
    static constexpr size_t bytes() { return sizeof(unsigned) * 3 * kSlots * kLine; }

    ClaimSlotsT() = default;
    ClaimSlotsT(const ClaimSlotsT&) = delete;
    ClaimSlotsT& operator=(const ClaimSlotsT&) = delete;
    ~ClaimSlotsT() { release_events(); }

    // `dev`: bytes() of zeroed device memory owned by the caller.
    void attach(unsigned* dev) { dev_ = dev; }
    unsigned* base() const { return dev_; }

    // Run launch(counter) for one launch on stream s, under the plan's lock, with a counter
    // that no launch which may still be drawing shares.  Returns launch's status or the
    // runtime error that prevented the launch.
    template <class Launch>
    typename RT::Error run(typename RT::Stream s, Launch&& launch)
    {
        std::lock_guard<std::mutex> g(mu_);
        bool cap = false;
        typename RT::Error e = RT::capturing(s, &cap);
        if (e != RT::ok) return e;
        if (cap) {
            if (ncaptured_ >= kSlots) return RT::busy;
            return launch(line(2, ncaptured_++));
        }
        if (!multi_) {
            if (!have_first_) {
                have_first_ = true;
                first_ = s;
            }
            if (s == first_) return launch(line(0, (int)(next0_++ % kSlots)));
            multi_ = true;
        }
        const int i = (int)(next1_++ % kSlots);
        if (!made_[i]) {
            e = RT::event_create(&ev_[i]);
            if (e != RT::ok) return e;
            made_[i] = true;
        } else if (recorded_[i] && !RT::event_done(ev_[i])) {
            e = RT::stream_wait(s, ev_[i]);
            if (e != RT::ok) return e;
        }
        recorded_[i] = false;
        e = launch(line(1, i));
        if (e != RT::ok) return e;
        e = RT::event_record(ev_[i], s);
        recorded_[i] = e == RT::ok;
        return e;
    }

    // For tests: which region the plan is in (0 one stream, 1 several) and captures taken.
    int region() const { return multi_ ? 1 : 0; }
    int captured() const { return ncaptured_; }

    void release_events()
    {
        for (int i = 0; i < kSlots; ++i)
            if (made_[i]) {
                RT::event_destroy(ev_[i]);
                made_[i] = recorded_[i] = false;
            }
    }

private:
    unsigned* line(int region, int i) const
    {
        return dev_ + ((size_t)region * kSlots + (size_t)i) * kLine;
    }

    std::mutex mu_;
    unsigned* dev_ = nullptr;
    bool have_first_ = false, multi_ = false;
    typename RT::Stream first_{};
    unsigned long next0_ = 0, next1_ = 0;
    int ncaptured_ = 0;
    typename RT::Event ev_[kSlots]{};
    bool made_[kSlots]{}, recorded_[kSlots]{};
};

}  // namespace fx
