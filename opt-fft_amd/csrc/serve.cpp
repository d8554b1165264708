// Continuous-batching request server (include/fft3d_serve.h).
//
// One scheduler thread owns the device side: admission against the byte budget, packing of
// fused-geometry requests into mixed launches (MultiDescriptor + captured hipGraph per
// slot), plan launches for the pencil geometries, and retirement by event polling.  Other
// threads only append to the incoming queue and wait on completion flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "fft3d_mi355x.h"
#include "fft3d_serve.h"
#include "multi_tasks.hpp"
#include "plan_internal.hpp"

using namespace fx;

namespace {

double steady_now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr int kMaxStreams = 16;
constexpr int kMaxBatch = 512;
constexpr size_t kAlign = 256;

// First-fit sub-allocator over one device allocation.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    std::map<size_t, size_t> free_;  // offset -> size

    bool init(size_t bytes)
    {
        cap = bytes;
        if (hipMalloc((void**)&base, bytes) != hipSuccess) return false;
        free_[0] = bytes;
        return true;
    }
    void* alloc(size_t n)
    {
        n = (n + kAlign - 1) / kAlign * kAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            return base + off;
        }
        return nullptr;
    }
    void release(void* p, size_t n)
    {
        n = (n + kAlign - 1) / kAlign * kAlign;
        size_t off = (size_t)((char*)p - base);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }
};

struct Req {
    fft3d_srv_request r;
    long id;
    int code;          // mixed-launch geometry code, or -1 (plan path)
    size_t io_bytes;   // in + out
    size_t out_bytes;
    void* out;         // where the result goes
    bool arena_out;
    double t_arrive = 0, t_done = 0, t_ready = 0;
    std::atomic<int> done{0};
};

// A reusable mixed launch: device descriptor, pinned staging, claim counters (zeroed once,
// left at zero by every launch), captured graph.
struct Slot {
    int prec;
    void* dev = nullptr;
    void* host = nullptr;
    unsigned* work = nullptr;
    size_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool busy = false;
};

struct Launch {
    hipEvent_t ev;
    int stream;
    int slot;  // -1 for a plan launch
    std::vector<Req*> reqs;
};

}  // namespace

struct fft3d_srv {
    fft3d_srv_config cfg;
    int device = 0;
    Arena arena;
    hipStream_t streams[kMaxStreams] = {};
    bool stream_busy[kMaxStreams] = {};
    std::vector<hipEvent_t> event_pool;
    void* bank[2] = {nullptr, nullptr};
    const MultiDesc* mk[2] = {nullptr, nullptr};
    std::vector<Slot> slots;
    std::map<long, fft3d_mx_plan*> plans;  // key: geometry/precision/direction

    std::mutex mu;
    std::condition_variable cv_in, cv_done;
    std::deque<Req*> incoming;
    std::vector<std::unique_ptr<Req>> all;
    long submitted = 0;  // guarded by mu
    bool stop = false;
    std::thread sched;

    // scheduler-owned
    std::deque<Req*> pending, ready;
    std::vector<Launch> inflight;
    size_t inflight_bytes = 0;
    fft3d_srv_stats st{};
    fft3d_srv_stats st_pub{};  // snapshot of st published under mu
    long items_in_mixed = 0;
    double busy_since = -1;
};

static long plan_key(int n0, int n1, int n2, int prec, int dir)
{
    return ((((long)n0 * 64 + n1) * 64 + n2) * 2 + prec) * 2 + (dir > 0);
}

static size_t vol_bytes(const fft3d_srv_request& r)
{
    return (size_t)r.n0 * r.n1 * r.n2 * (r.precision == FFT3D_MX_C128 ? 16 : 8);
}

// Launch a mixed batch of fused requests on a free slot of their precision.
static int launch_mixed(fft3d_srv* s, std::vector<Req*>& reqs, int stream)
{
    const int prec = reqs[0]->r.precision;
    Slot* slot = nullptr;
    int si = -1;
    for (size_t i = 0; i < s->slots.size(); ++i)
        if (s->slots[i].prec == prec && !s->slots[i].busy) {
            slot = &s->slots[i];
            si = (int)i;
            break;
        }
    if (!slot) return 0;
    const MultiDesc& d = *s->mk[prec];
    int code[kMultiMaxItems], batch[kMultiMaxItems];
    const int n = (int)reqs.size();
    for (int i = 0; i < n; ++i) {
        code[i] = reqs[i]->code;
        batch[i] = reqs[i]->r.batch;
    }
    std::vector<int2> tasks;
    multi_tasks(d, code, batch, n, elem_bytes(prec), tasks);
    auto fill = [&](auto* desc) {
        using T = std::remove_reference_t<decltype(desc->args.item[0].scale)>;
        std::memset(&desc->args, 0, sizeof desc->args);
        for (int i = 0; i < n; ++i) {
            auto& it = desc->args.item[i];
            it.in = (const cx<T>*)reqs[i]->r.in;
            it.out = (cx<T>*)reqs[i]->out;
            it.nvol = reqs[i]->r.batch;
            it.code = code[i];
            it.conj = reqs[i]->r.direction > 0;
            it.scale = (T)1;
        }
        desc->ntasks = (int)tasks.size();
        std::memcpy(desc->tasks, tasks.data(), tasks.size() * sizeof(int2));
        return offsetof(std::remove_reference_t<decltype(*desc)>, tasks) + tasks.size() * sizeof(int2);
    };
    const size_t bytes = prec == FFT3D_MX_C128 ? fill((MultiDescriptor<double>*)slot->host)
                                               : fill((MultiDescriptor<float>*)slot->host);
    hipStream_t st = s->streams[stream];
    if (hipMemcpyAsync(slot->dev, slot->host, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
        return -1;
    if (hipGraphLaunch(slot->exec, st) != hipSuccess) return -1;
    Launch L{s->event_pool.back(), stream, si, reqs};
    s->event_pool.pop_back();
    (void)hipEventRecord(L.ev, st);
    slot->busy = true;
    s->stream_busy[stream] = true;
    s->inflight.push_back(L);
    s->st.launches++;
    s->st.mixed_launches++;
    s->items_in_mixed += n;
    return 1;
}

static bool launch_plan(fft3d_srv* s, Req* q, int stream, bool sync)
{
    const fft3d_srv_request& r = q->r;
    auto it = s->plans.find(plan_key(r.n0, r.n1, r.n2, r.precision, r.direction));
    if (it == s->plans.end()) return false;
    hipStream_t st = s->streams[stream];
    if (fft3d_mx_execute_n(it->second, r.in, q->out, r.batch, st) != FFT3D_MX_OK) return false;
    Launch L{s->event_pool.back(), stream, -1, {q}};
    s->event_pool.pop_back();
    (void)hipEventRecord(L.ev, st);
    if (sync) (void)hipEventSynchronize(L.ev);
    s->stream_busy[stream] = true;
    s->inflight.push_back(L);
    s->st.launches++;
    s->st.plan_launches++;
    return true;
}

// Copy the scheduler's counters to the snapshot fft3d_srv_get_stats reads (caller holds mu).
static void publish_locked(fft3d_srv* s)
{
    s->st_pub = s->st;
    s->st_pub.submitted = s->submitted;
    s->st_pub.mean_items_per_mixed_launch =
        s->st.mixed_launches ? (double)s->items_in_mixed / (double)s->st.mixed_launches : 0.0;
}

static void publish(fft3d_srv* s)
{
    std::lock_guard<std::mutex> lk(s->mu);
    publish_locked(s);
}

// t < 0 marks a request that could not be launched.  The waiter is released by the caller
// (release()) once the statistics that count this completion are published.
static void finish(fft3d_srv* s, Req* q, double t)
{
    q->t_done = t;
    s->inflight_bytes -= q->io_bytes;
    if (q->arena_out) s->arena.release(q->out, q->out_bytes);
    s->st.completed++;
}

// Publish the statistics that count these completions and mark them done, both under mu, so
// a waiter that tested its flag under mu before the mark is already waiting on cv_done when
// the notification comes.
static void release(fft3d_srv* s, const std::vector<Req*>& reqs)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        publish_locked(s);
        for (Req* q : reqs) q->done.store(1, std::memory_order_release);
    }
    s->cv_done.notify_all();
}

// Requests that could not be launched complete at once with a failure mark.
static void fail_reqs(fft3d_srv* s, const std::vector<Req*>& reqs)
{
    for (Req* q : reqs) finish(s, q, -1.0);
    release(s, reqs);
}

static void retire(fft3d_srv* s)
{
    std::vector<Req*> done;
    for (size_t i = 0; i < s->inflight.size();) {
        Launch& L = s->inflight[i];
        if (hipEventQuery(L.ev) != hipSuccess) {
            ++i;
            continue;
        }
        const double t = steady_now();
        for (Req* q : L.reqs) {
            finish(s, q, t);
            done.push_back(q);
        }
        if (L.slot >= 0) s->slots[L.slot].busy = false;
        s->stream_busy[L.stream] = false;
        s->event_pool.push_back(L.ev);
        s->inflight[i] = s->inflight.back();
        s->inflight.pop_back();
    }
    if (done.empty()) return;
    release(s, done);
}

static int free_stream(fft3d_srv* s)
{
    for (int i = 0; i < s->cfg.streams; ++i)
        if (!s->stream_busy[i]) return i;
    return -1;
}

static void admit(fft3d_srv* s)
{
    while (!s->pending.empty()) {
        Req* q = s->pending.front();
        if (s->inflight_bytes + q->io_bytes > s->cfg.budget_bytes && s->inflight_bytes > 0) break;
        if (!q->r.out) {
            void* p = s->arena.alloc(q->out_bytes);
            if (!p) break;
            q->out = p;
            q->arena_out = true;
        }
        s->inflight_bytes += q->io_bytes;
        s->st.peak_inflight_bytes = std::max(s->st.peak_inflight_bytes, s->inflight_bytes);
        s->pending.pop_front();
        q->t_ready = steady_now();
        s->ready.push_back(q);
    }
}

static void dispatch(fft3d_srv* s)
{
    while (!s->ready.empty()) {
        const int stream = free_stream(s);
        if (stream < 0) return;
        Req* head = s->ready.front();
        if (s->cfg.batched != FFT3D_SRV_BATCHED || head->code < 0) {
            // plain modes, and the plan-path geometries of the batched mode
            const bool sync = s->cfg.batched == FFT3D_SRV_PLAIN_SYNC;
            s->ready.pop_front();
            if (!launch_plan(s, head, stream, sync)) fail_reqs(s, {head});
            if (sync) retire(s);
            continue;
        }
        // A partial mixed batch stays open for up to window_us after its oldest request was
        // admitted; meanwhile plan-path requests behind it still launch.
        if (s->cfg.window_us > 0) {
            int n = 0;
            for (Req* q : s->ready)
                if (q->code >= 0 && q->r.precision == head->r.precision && ++n >= s->cfg.max_items)
                    break;
            if (n < s->cfg.max_items && steady_now() - head->t_ready < s->cfg.window_us * 1e-6) {
                auto it = std::find_if(s->ready.begin(), s->ready.end(),
                                       [](const Req* q) { return q->code < 0; });
                if (it == s->ready.end()) return;
                Req* q = *it;
                s->ready.erase(it);
                if (!launch_plan(s, q, stream, false)) fail_reqs(s, {q});
                continue;
            }
        }
        // Pack ready fused requests of the head's precision, oldest first.
        std::vector<Req*> pack;
        for (auto it = s->ready.begin(); it != s->ready.end() && (int)pack.size() < s->cfg.max_items;) {
            if ((*it)->code >= 0 && (*it)->r.precision == head->r.precision) {
                pack.push_back(*it);
                it = s->ready.erase(it);
            } else {
                ++it;
            }
        }
        const int rc = launch_mixed(s, pack, stream);
        if (rc < 0) {
            fail_reqs(s, pack);
        } else if (rc == 0) {
            // no free slot: put them back in order and wait for a retirement
            for (auto it = pack.rbegin(); it != pack.rend(); ++it) s->ready.push_front(*it);
            return;
        }
    }
}

static void scheduler(fft3d_srv* s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->incoming.empty() && s->pending.empty() && s->ready.empty() &&
                s->inflight.empty()) {
                if (s->stop) return;
                s->cv_in.wait_for(lk, std::chrono::milliseconds(2));
            }
            while (!s->incoming.empty()) {
                s->pending.push_back(s->incoming.front());
                s->incoming.pop_front();
            }
        }
        retire(s);
        admit(s);
        dispatch(s);
        const double t = steady_now();
        if (!s->inflight.empty() && s->busy_since < 0) s->busy_since = t;
        if (s->inflight.empty() && s->busy_since >= 0) {
            s->st.busy_seconds += t - s->busy_since;
            s->busy_since = -1;
        }
        publish(s);
        if (!s->inflight.empty()) std::this_thread::yield();
    }
}

static bool make_slots(fft3d_srv* s, int prec, int count)
{
    const MultiDesc* d = s->mk[prec];
    const size_t max_tasks = (size_t)kMultiMaxItems * kMaxBatch;
    const size_t bytes = (prec == FFT3D_MX_C128 ? sizeof(MultiDescriptor<double>)
                                                : sizeof(MultiDescriptor<float>)) +
                         max_tasks * sizeof(int2);
    hipStream_t cap;
    if (hipStreamCreate(&cap) != hipSuccess) return false;
    const int grid = d->resident_grid();
    bool ok = true;
    for (int i = 0; i < count && ok; ++i) {
        Slot sl;
        sl.prec = prec;
        sl.bytes = bytes;
        ok = hipMalloc(&sl.dev, bytes) == hipSuccess &&
             hipHostMalloc(&sl.host, bytes, hipHostMallocDefault) == hipSuccess &&
             hipMalloc((void**)&sl.work, kClaimWords * sizeof(unsigned)) == hipSuccess &&
             hipMemset(sl.work, 0, kClaimWords * sizeof(unsigned)) == hipSuccess;
        if (ok) {
            // An empty descriptor until the first refresh: every workgroup exits at once.
            std::memset(sl.host, 0, bytes);
            ok = hipMemcpy(sl.dev, sl.host, bytes, hipMemcpyHostToDevice) == hipSuccess;
        }
        ok = ok && hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            const hipError_t le = d->launch_dev(sl.dev, s->bank[prec], sl.work, grid, cap);
            const hipError_t ce = hipStreamEndCapture(cap, &sl.graph);
            ok = le == hipSuccess && ce == hipSuccess &&
                 hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0) == hipSuccess;
        }
        s->slots.push_back(sl);
    }
    (void)hipStreamDestroy(cap);
    return ok;
}

static void destroy_all(fft3d_srv* s)
{
    for (Slot& sl : s->slots) {
        if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
        if (sl.graph) (void)hipGraphDestroy(sl.graph);
        (void)hipFree(sl.dev);
        (void)hipFree(sl.work);
        if (sl.host) (void)hipHostFree(sl.host);
    }
    for (auto& kv : s->plans) fft3d_mx_destroy(kv.second);
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    for (int i = 0; i < kMaxStreams; ++i)
        if (s->streams[i]) (void)hipStreamDestroy(s->streams[i]);
    for (void* b : s->bank) (void)hipFree(b);
    (void)hipFree(s->arena.base);
}

extern "C" double fft3d_srv_now(void) { return steady_now(); }

extern "C" fft3d_srv* fft3d_srv_create(const fft3d_srv_config* cfg)
{
    if (!cfg || cfg->budget_bytes < (1u << 20) || cfg->streams < 1 || cfg->streams > kMaxStreams ||
        cfg->max_items < 1 || cfg->max_items > kMultiMaxItems || !(cfg->window_us >= 0) ||
        cfg->window_us > 1e6 || cfg->batched < FFT3D_SRV_PLAIN_SYNC ||
        cfg->batched > FFT3D_SRV_PLAIN_ASYNC) {
        fail(FFT3D_MX_EINVAL, "fft3d_srv_create: bad configuration");
        return nullptr;
    }
    fft3d_srv* s = new (std::nothrow) fft3d_srv();
    if (!s) return nullptr;
    s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    bool ok = s->arena.init(cfg->budget_bytes);
    for (int i = 0; i < cfg->streams && ok; ++i) ok = hipStreamCreate(&s->streams[i]) == hipSuccess;
    for (int i = 0; i < 64 && ok; ++i) {
        hipEvent_t e;
        ok = hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (ok) s->event_pool.push_back(e);
    }
    for (int prec = 0; prec < 2 && ok; ++prec) {
        std::vector<unsigned char> img;
        bank_host_image(prec, img);
        ok = hipMalloc(&s->bank[prec], img.size()) == hipSuccess &&
             hipMemcpy(s->bank[prec], img.data(), img.size(), hipMemcpyHostToDevice) == hipSuccess;
        for (const MultiDesc& d : multi_registry())
            if (d.prec == prec && d.algo == 0 && d.launch_dev && !s->mk[prec]) s->mk[prec] = &d;
        ok = ok && s->mk[prec] && make_slots(s, prec, cfg->streams);
    }
    // Plans for every geometry, both precisions and directions (the plain modes use them
    // for all requests, the batched mode for the pencil geometries).
    static const int geo[][3] = {{6, 6, 6}, {8, 8, 8}, {12, 12, 12}, {17, 17, 17}, {8, 8, 16},
                                 {24, 24, 24}, {36, 36, 36}, {24, 24, 48}};
    for (const auto& g : geo)
        for (int prec = 0; prec < 2 && ok; ++prec)
            for (int dir = -1; dir <= 1 && ok; dir += 2) {
                if (cfg->batched == FFT3D_SRV_BATCHED && multi_code(g[0], g[1], g[2]) >= 0)
                    continue;
                fft3d_mx_plan* p = fft3d_mx_create(g[0], g[1], g[2], kMaxBatch, dir, prec, 0);
                ok = p != nullptr;
                if (ok) s->plans[plan_key(g[0], g[1], g[2], prec, dir)] = p;
            }
    if (!ok) {
        if (!*last_error()) fail(FFT3D_MX_EHIP, "fft3d_srv_create: device set-up failed");
        destroy_all(s);
        delete s;
        return nullptr;
    }
    s->sched = std::thread(scheduler, s);
    return s;
}

extern "C" long fft3d_srv_submit(fft3d_srv* s, const fft3d_srv_request* r, double arrival)
{
    if (!s || !r) return fail(FFT3D_MX_EINVAL, "fft3d_srv_submit: NULL argument");
    if (r->batch < 1 || r->batch > kMaxBatch || !r->in ||
        (r->direction != -1 && r->direction != 1) || (r->precision != 0 && r->precision != 1))
        return fail(FFT3D_MX_EINVAL, "fft3d_srv_submit: bad request");
    if (!elem_aligned(r->in, r->precision) || (r->out && !elem_aligned(r->out, r->precision)))
        return fail(FFT3D_MX_EINVAL, "fft3d_srv_submit: buffers must be aligned to %zu bytes",
                    elem_bytes(r->precision));
    if (!fft3d_mx_supports(r->n0, r->n1, r->n2, r->precision))
        return fail(FFT3D_MX_EUNSUPPORTED, "fft3d_srv_submit: %dx%dx%d not supported", r->n0,
                    r->n1, r->n2);
    std::unique_ptr<Req> q(new (std::nothrow) Req());
    if (!q) return fail(FFT3D_MX_ENOMEM, "fft3d_srv_submit: out of host memory");
    q->r = *r;
    q->code = multi_code(r->n0, r->n1, r->n2);
    q->out_bytes = vol_bytes(*r) * (size_t)r->batch;
    q->io_bytes = 2 * q->out_bytes;
    if (q->io_bytes > s->cfg.budget_bytes)
        return fail(FFT3D_MX_EINVAL, "fft3d_srv_submit: request larger than the budget");
    q->out = r->out;
    q->arena_out = false;
    q->t_arrive = arrival > 0 ? arrival : steady_now();
    Req* raw = q.get();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        raw->id = (long)s->all.size();
        s->all.push_back(std::move(q));
        s->incoming.push_back(raw);
        s->submitted++;
    }
    s->cv_in.notify_one();
    return raw->id;
}

extern "C" int fft3d_srv_wait(fft3d_srv* s, long id, double* latency)
{
    if (!s) return fail(FFT3D_MX_EINVAL, "fft3d_srv_wait: NULL server");
    Req* q;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (id < 0 || id >= (long)s->all.size()) return fail(FFT3D_MX_EINVAL, "fft3d_srv_wait: bad id");
        q = s->all[id].get();
    }
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [q] { return q->done.load(std::memory_order_acquire) != 0; });
    }
    if (q->t_done < 0) return fail(FFT3D_MX_EHIP, "fft3d_srv_wait: request %ld failed to launch", id);
    if (latency) *latency = q->t_done - q->t_arrive;
    return FFT3D_MX_OK;
}

extern "C" int fft3d_srv_run_trace(fft3d_srv* s, const fft3d_srv_request* reqs,
                                   const double* arrival, int n, double* latency)
{
    if (!s || !reqs || !arrival || n < 0) return fail(FFT3D_MX_EINVAL, "fft3d_srv_run_trace: bad arguments");
    std::vector<long> ids(n);
    const double t0 = steady_now() + 1e-3;
    for (int i = 0; i < n; ++i) {
        const double when = t0 + arrival[i];
        for (;;) {
            const double now = steady_now();
            if (now >= when) break;
            if (when - now > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
        ids[i] = fft3d_srv_submit(s, &reqs[i], when);
        if (ids[i] < 0) return (int)ids[i];
    }
    int rc = FFT3D_MX_OK;
    for (int i = 0; i < n; ++i) {
        double lat = 0;
        const int r = fft3d_srv_wait(s, ids[i], &lat);
        if (r != FFT3D_MX_OK) rc = r;
        if (latency) latency[i] = lat;
    }
    return rc;
}

extern "C" int fft3d_srv_get_stats(const fft3d_srv* s, fft3d_srv_stats* out)
{
    if (!s || !out) return fail(FFT3D_MX_EINVAL, "fft3d_srv_get_stats: NULL argument");
    fft3d_srv* m = const_cast<fft3d_srv*>(s);
    std::lock_guard<std::mutex> lk(m->mu);
    *out = s->st_pub;
    return FFT3D_MX_OK;
}

extern "C" void fft3d_srv_destroy(fft3d_srv* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_in.notify_all();
    if (s->sched.joinable()) s->sched.join();
    (void)hipDeviceSynchronize();
    destroy_all(s);
    delete s;
}
