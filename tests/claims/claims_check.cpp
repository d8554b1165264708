// TEST INFRASTRUCTURE: ClaimSlotsT (opt-fft_amd/csrc/claims.hpp) over a recording fake
// runtime.  A fake launch is "drawing" from launch until its event completes; the checker
// fails if a launch is handed a counter that a still-drawing launch holds, unless the
// launch's stream was made to wait for that launch first.  Prints "OK" or the failure.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "claims.hpp"

struct Fake {
    static int capturing_stream;                // stream id in capture, or -1
    static std::map<int, bool> done;            // event id -> completed
    static std::map<int, std::set<int>> waits;  // stream -> events it waited for
    static int next_event, records, creates;
    static int fail_record;
};
int Fake::capturing_stream = -1;
std::map<int, bool> Fake::done;
std::map<int, std::set<int>> Fake::waits;
int Fake::next_event = 1, Fake::records = 0, Fake::creates = 0, Fake::fail_record = 0;

struct RT {
    using Stream = int;
    using Event = int;
    using Error = int;
    static constexpr int ok = 0;
    static constexpr int busy = 7;
    static int capturing(int s, bool* yes) { *yes = s == Fake::capturing_stream; return 0; }
    static int event_create(int* e) { *e = Fake::next_event++; Fake::creates++; Fake::done[*e] = false; return 0; }
    static bool event_done(int e) { return Fake::done[e]; }
    static int stream_wait(int s, int e) { Fake::waits[s].insert(e); return 0; }
    static int event_record(int e, int) { if (Fake::fail_record) return 3; Fake::done[e] = false; Fake::records++; return 0; }
    static void event_destroy(int) {}
};

static unsigned g_dev[fx::ClaimSlotsT<RT>::bytes() / sizeof(unsigned)];
static int g_fail = 0;
#define CHECK(c, ...) do { if (!(c)) { printf("FAIL line %d: ", __LINE__); printf(__VA_ARGS__); printf("\n"); g_fail = 1; } } while (0)

int main()
{
    using CS = fx::ClaimSlotsT<RT>;
    auto idx = [](unsigned* w) { return (int)((w - g_dev) / CS::kLine); };
    {   // one stream: region 0 round-robin, no events, no waits
        CS c;
        c.attach(g_dev);
        for (int n = 0; n < 200; ++n) {
            int got = -1;
            CHECK(c.run(5, [&](unsigned* w) { got = idx(w); return 0; }) == 0, "run");
            CHECK(got == n % 64, "one stream: launch %d got line %d", n, got);
        }
        CHECK(Fake::creates == 0 && Fake::records == 0 && Fake::waits.empty(), "events on one stream");
        CHECK(c.region() == 0, "region");
    }
    {   // two streams: region 1 from the switch on; the 65th launch since the switch reuses
        // slot 0 and must wait for the 1st unless that one has completed
        CS c;
        c.attach(g_dev);
        int got = -1;
        c.run(1, [&](unsigned* w) { got = idx(w); return 0; });
        CHECK(got == 0, "first");
        std::vector<int> evs;
        for (int n = 0; n < 64; ++n) {
            c.run(2 + n % 3, [&](unsigned* w) { got = idx(w); return 0; });
            CHECK(got == 64 + n, "switch: launch %d got line %d", n, got);
        }
        CHECK(c.region() == 1 && Fake::creates == 64 && Fake::records == 64, "events made");
        // event ids 1..64 belong to slots 0..63; slot 0's launch still running
        c.run(9, [&](unsigned* w) { got = idx(w); return 0; });
        CHECK(got == 64, "wrap: got line %d", got);
        CHECK(Fake::waits[9].count(1) == 1, "65th launch did not wait for the 1st");
        // slot 1's launch has completed: no wait
        Fake::done[2] = true;
        c.run(10, [&](unsigned* w) { got = idx(w); return 0; });
        CHECK(got == 65 && Fake::waits[10].empty(), "completed slot waited");
        // the plan stays in region 1 even on the first stream again
        c.run(1, [&](unsigned* w) { got = idx(w); return 0; });
        CHECK(got == 66, "back on stream 1 got line %d", got);
        // slots 3..63 once more, then a failed launch on slot 0: it records nothing, and the
        // next user of slot 0 does not wait
        int recs = Fake::records;
        for (int n = 0; n < 61; ++n) c.run(2, [&](unsigned*) { return 0; });
        CHECK(c.run(3, [&](unsigned* w) { got = idx(w); return 4; }) == 4, "launch error");
        CHECK(got == 64 && Fake::records == recs + 61, "failed launch: line %d", got);
        Fake::waits.clear();
        for (int n = 0; n < 63; ++n) c.run(4, [&](unsigned*) { return 0; });
        c.run(11, [&](unsigned* w) { got = idx(w); return 0; });
        CHECK(got == 64 && Fake::waits[11].empty(), "waited on an unrecorded slot");
    }
    {   // captures: region 2, one counter each, never reused; the 65th is refused
        CS c;
        c.attach(g_dev);
        Fake::capturing_stream = 3;
        std::set<int> seen;
        for (int n = 0; n < 64; ++n) {
            int got = -1;
            CHECK(c.run(3, [&](unsigned* w) { got = idx(w); return 0; }) == 0, "capture");
            CHECK(got == 128 + n && !seen.count(got), "capture %d got line %d", n, got);
            seen.insert(got);
        }
        bool called = false;
        CHECK(c.run(3, [&](unsigned*) { called = true; return 0; }) == RT::busy && !called,
              "65th capture not refused");
        Fake::capturing_stream = -1;
        int got = -1;
        c.run(3, [&](unsigned* w) { got = idx(w); return 0; });
        CHECK(got == 0 && c.captured() == 64, "after capture");
    }
    if (!g_fail) printf("OK\n");
    return g_fail;
}
