// serve.hip -- the serving form: continuous batching of sort requests.
//
// NWChem calls tce_sort_4/6 once per tile, one after another (trunk/
// makelib.py:64-95 is a synchronous subroutine).  Here requests arrive
// over time and are served by an engine thread that
//   * moves requests whose arrival time has passed into a FIFO;
//   * admits them, in order, while the bytes of admitted-but-unfinished
//     requests (input + output footprint) fit the HBM budget;
//   * holds back a request whose output overlaps the output or input of an
//     earlier unfinished request (or whose input overlaps such an output),
//     so writes to one tensor happen in arrival order -- (b + x) + y and
//     (b + y) + x differ in floating point;
//   * launches each admitted group as ONE batched persistent kernel on a
//     free slot (its own stream), so groups overlap on the device while
//     later requests keep arriving and being admitted;
//   * retires groups when their completion event fires, freeing budget.
// Per-request arrival, launch and completion times give p50/p99 latency.
// A fake executor (SPG_SERVER_FAKE) replaces the device with a bandwidth
// model so the scheduling logic can be tested without a GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "spg_internal.hpp"

namespace spg {
namespace {

struct ServeReq {
    const double* in = nullptr;
    double* out = nullptr;
    bool own_out = false;  // output carved from the server's arena
    uint64_t arena_off = 0;
    Plan P;                // batched plan, or the per-call plan of a big request
    bool big = false;      // at least kBatchBigElems elements: its own launch
    int64_t foot = 0;      // bytes counted against the budget (in + out)
    int64_t alg = 0;       // algorithmic HBM bytes 8*N*(2+acc)
    double arrival = 0, admit = -1, done = -1;
    int32_t batch = -1;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<int64_t> reqs;
    char* d_buf = nullptr;  // work list on the device
    char* h_buf = nullptr;  // pinned staging copy
    size_t cap = 0;
    double fake_done = 0;
};

struct Range {
    uintptr_t lo, hi;
};

// What a launch needs of one admitted request, copied under the server's
// lock at admission: the engine never reads the request list unlocked
// (spg_server_submit may grow it at any time).
struct GroupItem {
    Plan P;
    const double* in;
    double* out;
    bool big;  // launched on its own per-call plan
};

bool hits(const std::vector<Range>& v, uintptr_t lo, uintptr_t hi)
{
    for (const Range& r : v)
        if (lo < r.hi && r.lo < hi) return true;
    return false;
}

}  // namespace
}  // namespace spg

using namespace spg;

struct spg_server {
    int64_t budget = 0;
    int nslots = 1;
    int max_batch = 1024;
    bool fake = false;
    double fake_gbs = 4000.0, fake_lat_us = 5.0;
    std::mutex mu;
    std::condition_variable cv;    // completions (spg_server_wait)
    std::condition_variable wake;  // the engine: a submission or stop
    std::vector<ServeReq> reqs;    // read and written only under mu
    // ids not yet arrived, earliest arrival first
    std::priority_queue<std::pair<double, int64_t>, std::vector<std::pair<double, int64_t>>,
                        std::greater<std::pair<double, int64_t>>>
        future;
    std::deque<int64_t> pending;
    int64_t in_flight = 0, peak_in_flight = 0;
    int64_t completed = 0;
    int nbatches = 0;
    std::vector<Slot> slots;
    std::thread th;
    std::atomic<bool> stop{false};
    bool started = false;
    std::chrono::steady_clock::time_point t0;
    int dev = 0;
    int error = SPG_OK;
    // arena of server-allocated outputs (first fit over [0, budget))
    char* arena = nullptr;
    std::map<uint64_t, uint64_t> free_list;  // offset -> size
    size_t work_cap = 0;

    double now_us() const
    {
        return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

int sfail(int rc, const char* msg)
{
    set_error(rc, msg);
    return rc;
}

bool arena_alloc(spg_server* s, uint64_t bytes, uint64_t* off)
{
    bytes = (bytes + 255) & ~(uint64_t)255;
    for (auto it = s->free_list.begin(); it != s->free_list.end(); ++it) {
        if (it->second >= bytes) {
            *off = it->first;
            const uint64_t rest = it->second - bytes, at = it->first + bytes;
            s->free_list.erase(it);
            if (rest) s->free_list[at] = rest;
            return true;
        }
    }
    return false;
}

void arena_free(spg_server* s, uint64_t off, uint64_t bytes)
{
    bytes = (bytes + 255) & ~(uint64_t)255;
    auto it = s->free_list.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != s->free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        s->free_list.erase(nx);
    }
    if (it != s->free_list.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            s->free_list.erase(it);
        }
    }
}

size_t work_bytes(size_t ntasks)
{
    // the tasks plus a typical chunk count; slots grow on demand
    return sizeof(BatchTask) * ntasks + sizeof(BatchChunk) * ntasks * 16 + 256;
}

int launch_small(Slot& sl, const std::vector<BatchTask>& tasks, const std::vector<BatchChunk>& chunks, int G,
                 uint32_t half, int paths);

// Launches a group on slot `sl` (device executor).
int launch_group(Slot& sl, const std::vector<GroupItem>& items)
{
    std::vector<BatchTask> tasks;
    tasks.reserve(items.size());
    for (const GroupItem& it : items) {
        if (it.big) continue;
        BatchTask t;
        if (make_batch_task(it.P, it.in, it.out, &t) != SPG_OK) return SPG_EINVAL;
        tasks.push_back(t);
    }
    std::vector<BatchChunk> chunks;
    uint32_t half = 1, tot = 0;
    int G = 1, paths = 0;
    int rc = schedule_batch(tasks.data(), tasks.size(), 0, kServeFlags, chunks, &half, &tot, &G, &paths);
    if (rc != SPG_OK) return rc;
    if (tot) {
        rc = launch_small(sl, tasks, chunks, G, half, paths);
        if (rc != SPG_OK) return rc;
    }
    // requests that fill the device on their own, on their per-call plans
    for (const GroupItem& it : items)
        if (it.big && (rc = launch_plan(it.P, it.in, it.out, sl.stream)) != SPG_OK) return rc;
    if (hipEventRecord(sl.ev, sl.stream) != hipSuccess) return SPG_EHIP;
    return SPG_OK;
}

// The group's small requests as one persistent batched launch on the slot.
int launch_small(Slot& sl, const std::vector<BatchTask>& tasks, const std::vector<BatchChunk>& chunks, int G,
                 uint32_t half, int paths)
{
    const size_t need = sizeof(BatchTask) * tasks.size() + sizeof(BatchChunk) * chunks.size();
    if (need > sl.cap) {
        // grow this slot's buffers (the previous group on it has finished)
        if (sl.d_buf) (void)hipFree(sl.d_buf);
        if (sl.h_buf) (void)hipHostFree(sl.h_buf);
        sl.d_buf = sl.h_buf = nullptr;
        sl.cap = 0;
        const size_t cap = need + need / 2;
        if (hipMalloc(&sl.d_buf, cap) != hipSuccess || hipHostMalloc(&sl.h_buf, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return SPG_ENOMEM;
        }
        sl.cap = cap;
    }
    char* h = sl.h_buf;
    std::memcpy(h, tasks.data(), sizeof(BatchTask) * tasks.size());
    const size_t o_ch = sizeof(BatchTask) * tasks.size();
    std::memcpy(h + o_ch, chunks.data(), sizeof(BatchChunk) * chunks.size());
    if (hipMemcpyAsync(sl.d_buf, h, need, hipMemcpyHostToDevice, sl.stream) != hipSuccess) return SPG_EHIP;
    const hipError_t e = launch_batch_kernel(reinterpret_cast<BatchTask*>(sl.d_buf),
                                             reinterpret_cast<BatchChunk*>(sl.d_buf + o_ch),
                                             (uint32_t)chunks.size(), G, half, kServeFlags, paths,
                                             sl.stream);
    if (e != hipSuccess) return SPG_EHIP;
    return SPG_OK;
}

// While groups are on the device the engine polls their completion events
// (hipEventQuery), yielding between polls, for up to kSpinUs after the last
// progress; then it sleeps kNapUs between polls.  With nothing on the device
// and nothing arrived it blocks until the next arrival time, a submission or
// stop: an idle server takes no host core.
constexpr double kSpinUs = 200.0;
constexpr double kNapUs = 50.0;

void engine(spg_server* s)
{
    if (!s->fake) (void)hipSetDevice(s->dev);
    double last_progress = s->now_us();
    while (!s->stop.load(std::memory_order_relaxed)) {
        bool did = false;
        // 1. retire finished groups
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            bool fin;
            if (s->fake) {
                fin = s->now_us() >= sl.fake_done;
            } else {
                const hipError_t q = hipEventQuery(sl.ev);
                fin = q == hipSuccess;
                if (q != hipSuccess && q != hipErrorNotReady) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->error = SPG_EHIP;
                    fin = true;
                }
            }
            if (!fin) continue;
            const double t = s->now_us();
            std::lock_guard<std::mutex> lk(s->mu);
            for (int64_t id : sl.reqs) {
                ServeReq& r = s->reqs[id];
                r.done = t;
                s->in_flight -= r.foot;
                if (r.own_out) arena_free(s, r.arena_off, (uint64_t)r.P.nelem * sizeof(double));
                ++s->completed;
            }
            sl.reqs.clear();
            sl.busy = false;
            did = true;
            s->cv.notify_all();
        }
        // 2. arrivals and admission
        Slot* free_slot = nullptr;
        for (Slot& sl : s->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        std::vector<int64_t> group;
        std::vector<GroupItem> items;
        int64_t group_alg = 0;
        double t_admit = 0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            const double now = s->now_us();
            while (!s->future.empty() && s->future.top().first <= now) {
                s->pending.push_back(s->future.top().second);
                s->future.pop();
            }
            if (free_slot && !s->pending.empty()) {
                std::vector<Range> busy_out, busy_in;
                for (const Slot& sl : s->slots)
                    for (int64_t id : sl.reqs) {
                        const ServeReq& r = s->reqs[id];
                        const uintptr_t len = (uintptr_t)r.P.nelem * sizeof(double);
                        if (!r.own_out) busy_out.push_back({(uintptr_t)r.out, (uintptr_t)r.out + len});
                        busy_in.push_back({(uintptr_t)r.in, (uintptr_t)r.in + len});
                    }
                std::deque<int64_t> keep;
                bool full = false;
                for (int64_t id : s->pending) {
                    ServeReq& r = s->reqs[id];
                    if (full) {
                        keep.push_back(id);
                        continue;
                    }
                    const uintptr_t len = (uintptr_t)r.P.nelem * sizeof(double);
                    const uintptr_t ilo = (uintptr_t)r.in, ihi = ilo + len;
                    const uintptr_t olo = (uintptr_t)r.out, ohi = olo + len;
                    bool conflict = hits(busy_out, ilo, ihi);
                    if (!r.own_out) conflict = conflict || hits(busy_out, olo, ohi) || hits(busy_in, olo, ohi);
                    if (conflict) {
                        // wait behind the earlier request on the same data
                        if (!r.own_out) busy_out.push_back({olo, ohi});
                        busy_in.push_back({ilo, ihi});
                        keep.push_back(id);
                        continue;
                    }
                    if (s->in_flight + r.foot > s->budget) {
                        full = true;  // strict FIFO on the budget: no overtaking
                        keep.push_back(id);
                        continue;
                    }
                    if (r.own_out && !arena_alloc(s, (uint64_t)r.P.nelem * sizeof(double), &r.arena_off)) {
                        full = true;
                        keep.push_back(id);
                        continue;
                    }
                    s->in_flight += r.foot;
                    s->peak_in_flight = std::max(s->peak_in_flight, s->in_flight);
                    if (!r.own_out) busy_out.push_back({olo, ohi});
                    busy_in.push_back({ilo, ihi});
                    group.push_back(id);
                    if ((int)group.size() >= s->max_batch) full = true;
                }
                s->pending.swap(keep);
                t_admit = now;
                if (!group.empty()) {
                    items.reserve(group.size());
                    for (int64_t id : group) {
                        ServeReq& r = s->reqs[id];
                        r.admit = t_admit;
                        r.batch = s->nbatches;
                        group_alg += r.alg;
                        double* out = r.own_out ? reinterpret_cast<double*>(s->arena + r.arena_off) : r.out;
                        items.push_back(GroupItem{r.P, r.in, out, r.big});
                    }
                    ++s->nbatches;
                }
            }
        }
        if (!group.empty()) {
            free_slot->reqs = group;
            free_slot->busy = true;
            if (s->fake) {
                free_slot->fake_done =
                    s->now_us() + s->fake_lat_us + (double)group_alg / (s->fake_gbs * 1e3);
            } else {
                const int rc = launch_group(*free_slot, items);
                if (rc != SPG_OK) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->error = rc;
                    // the event of the previous use is complete: the group
                    // retires on the next pass and the error is reported
                    (void)hipEventRecord(free_slot->ev, free_slot->stream);
                }
            }
            did = true;
        }
        const double now = s->now_us();
        if (did) {
            last_progress = now;
            continue;
        }
        bool in_flight = false;
        for (const Slot& sl : s->slots) in_flight = in_flight || sl.busy;
        if (in_flight && now - last_progress < kSpinUs) {
            std::this_thread::yield();
            continue;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop.load(std::memory_order_relaxed)) break;
        // sleep until the next arrival (or a nap while groups run; the fake
        // executor's groups end at known times)
        double until = in_flight ? now + kNapUs : -1.0;
        if (s->fake)
            for (const Slot& sl : s->slots)
                if (sl.busy) until = std::min(until, sl.fake_done);
        if (!s->future.empty()) {
            const double next = s->future.top().first;
            until = until < 0 ? next : std::min(until, next);
        }
        if (!s->pending.empty() && !in_flight) until = now;  // admit again at once
        if (until < 0) {
            s->wake.wait(lk);
        } else if (until > now) {
            s->wake.wait_until(lk, s->t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                           std::chrono::duration<double, std::micro>(until)));
        }
    }
}

}  // namespace

extern "C" {

int spg_server_create(int64_t budget_bytes, int slots, int max_batch, int flags, spg_server** out)
{
    if (!out || budget_bytes <= 0 || slots < 1 || slots > 16 || max_batch < 1)
        return sfail(SPG_EINVAL, "invalid server parameters");
    spg_server* s = new spg_server();
    s->budget = budget_bytes;
    s->nslots = slots;
    s->max_batch = max_batch;
    s->fake = (flags & SPG_SERVER_FAKE) != 0;
    s->slots.resize(slots);
    s->free_list[0] = (uint64_t)budget_bytes;
    if (!s->fake) {
        if (hipGetDevice(&s->dev) != hipSuccess) {
            delete s;
            return sfail(SPG_EHIP, "no device");
        }
        const size_t cap = work_bytes((size_t)max_batch);
        for (Slot& sl : s->slots) {
            if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
                hipMalloc(&sl.d_buf, cap) != hipSuccess ||
                hipHostMalloc(&sl.h_buf, cap, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                spg_server_destroy(s);
                return sfail(SPG_ENOMEM, "server slot allocation failed");
            }
            sl.cap = cap;
        }
    }
    *out = s;
    return SPG_OK;
}

int spg_server_set_fake_rate(spg_server* s, double gbs, double latency_us)
{
    if (!s || gbs <= 0 || latency_us < 0) return sfail(SPG_EINVAL, "invalid rate");
    s->fake_gbs = gbs;
    s->fake_lat_us = latency_us;
    return SPG_OK;
}

int spg_server_submit(spg_server* s, const double* unsorted, double* sorted, int rank,
                      const int* dims, const int* perm, double factor, double beta,
                      double arrival_us, int64_t* id)
{
    if (!s) return sfail(SPG_EINVAL, "null server");
    ServeReq r;
    int rc = batch_plan(rank, dims, perm, factor, beta, &r.P);
    if (rc != SPG_OK) return sfail(rc, spg_strerror(rc));
    if (!unsorted && r.P.nelem) return sfail(SPG_EINVAL, "null input");
    const int64_t bytes = r.P.nelem * (int64_t)sizeof(double);
    r.in = unsorted;
    r.out = sorted;
    r.own_out = sorted == nullptr;
    if (!r.own_out && r.P.nelem && (uintptr_t)unsorted < (uintptr_t)sorted + bytes &&
        (uintptr_t)sorted < (uintptr_t)unsorted + bytes)
        return sfail(SPG_EALIAS, "unsorted and sorted overlap");
    if (r.P.nelem >= kBatchBigElems) {
        // the per-call plan; its launch geometry is fixed on the engine's
        // device at launch.  A server-allocated output is 256-byte aligned
        // in the arena, so only the input's alignment can rule out 16-byte
        // accesses there.
        r.big = true;
        const uintptr_t a = reinterpret_cast<uintptr_t>(unsorted) | (r.own_out ? 0 : reinterpret_cast<uintptr_t>(sorted));
        rc = plan_for_alignment((a & 15u) == 0, rank, dims, perm, factor, beta, nullptr, &r.P, false);
        if (rc != SPG_OK) return sfail(rc, spg_strerror(rc));
    }
    r.foot = 2 * bytes;
    r.alg = bytes * (r.P.mode == MODE_SORT ? 2 : 3);
    if (r.foot > s->budget) return sfail(SPG_EBUDGET, "request larger than the HBM budget");
    std::lock_guard<std::mutex> lk(s->mu);
    if (r.own_out && !s->fake && !s->arena) {
        // the first request without an output allocates the arena, under the
        // lock: concurrent submissions allocate it once, and the engine reads
        // it only under the lock (at admission)
        // on the server's device; the caller's current device is restored
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(s->dev) != hipSuccess)
            return sfail(SPG_EHIP, "cannot select the server's device");
        const hipError_t e = hipMalloc(&s->arena, (size_t)s->budget);
        (void)hipSetDevice(cur);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            s->arena = nullptr;
            return sfail(SPG_ENOMEM, "arena allocation failed");
        }
    }
    r.arrival = arrival_us >= 0 ? arrival_us : (s->started ? s->now_us() : 0.0);
    const int64_t k = (int64_t)s->reqs.size();
    s->reqs.push_back(r);
    if (r.P.nelem == 0) {
        s->reqs[k].admit = s->reqs[k].done = r.arrival;
        ++s->completed;
    } else {
        s->future.push({r.arrival, k});
    }
    if (id) *id = k;
    s->wake.notify_all();
    return SPG_OK;
}

int spg_server_start(spg_server* s)
{
    if (!s) return sfail(SPG_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return sfail(SPG_ESTATE, "server already started");
    s->t0 = std::chrono::steady_clock::now();
    s->started = true;
    s->th = std::thread(engine, s);
    return SPG_OK;
}

int spg_server_wait(spg_server* s, double timeout_s)
{
    if (!s) return sfail(SPG_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->started) return sfail(SPG_ESTATE, "server not started");
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (s->completed < (int64_t)s->reqs.size() && s->error == SPG_OK) {
        if (s->cv.wait_until(lk, until) == std::cv_status::timeout &&
            s->completed < (int64_t)s->reqs.size())
            return sfail(SPG_ESTATE, "timed out waiting for requests");
    }
    if (s->error != SPG_OK) return sfail(s->error, "a batch failed");
    return SPG_OK;
}

int spg_server_results(spg_server* s, int64_t n, double* arrival_us, double* admit_us,
                       double* done_us, int32_t* batch)
{
    if (!s) return sfail(SPG_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (n > (int64_t)s->reqs.size()) return sfail(SPG_EINVAL, "more results than requests");
    for (int64_t i = 0; i < n; ++i) {
        const ServeReq& r = s->reqs[i];
        if (arrival_us) arrival_us[i] = r.arrival;
        if (admit_us) admit_us[i] = r.admit;
        if (done_us) done_us[i] = r.done;
        if (batch) batch[i] = r.batch;
    }
    return SPG_OK;
}

int spg_server_stats(spg_server* s, char* buf, int buflen)
{
    if (!s || !buf) return sfail(SPG_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> lat;
    double t_first = 1e300, t_last = 0;
    int64_t alg = 0;
    for (const ServeReq& r : s->reqs) {
        if (r.done < 0) continue;
        lat.push_back(r.done - r.arrival);
        t_first = std::min(t_first, r.arrival);
        t_last = std::max(t_last, r.done);
        alg += r.alg;
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        size_t i = (size_t)std::min<double>((double)lat.size() - 1, q * (double)(lat.size() - 1) + 0.5);
        return lat[i];
    };
    const double span = t_last > t_first ? t_last - t_first : 0.0;
    char tmp[512];
    std::snprintf(tmp, sizeof tmp,
                  "{\"requests\":%zu,\"completed\":%lld,\"batches\":%d,\"p50_us\":%.2f,"
                  "\"p90_us\":%.2f,\"p99_us\":%.2f,\"max_us\":%.2f,\"span_us\":%.2f,"
                  "\"alg_bytes\":%lld,\"gbs\":%.2f,\"peak_in_flight\":%lld,\"budget\":%lld,"
                  "\"slots\":%d,\"error\":%d}",
                  s->reqs.size(), (long long)s->completed, s->nbatches, pct(0.5), pct(0.9), pct(0.99),
                  lat.empty() ? 0.0 : lat.back(), span, (long long)alg,
                  span > 0 ? (double)alg / span / 1e3 : 0.0, (long long)s->peak_in_flight,
                  (long long)s->budget, s->nslots, s->error);
    if (buflen <= (int)std::strlen(tmp)) return sfail(SPG_EINVAL, "buffer too small");
    std::strcpy(buf, tmp);
    return SPG_OK;
}

int spg_server_destroy(spg_server* s)
{
    if (!s) return SPG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop.store(true);
    }
    s->wake.notify_all();
    if (s->th.joinable()) s->th.join();
    if (!s->fake) {
        for (Slot& sl : s->slots) {
            if (sl.stream) (void)hipStreamSynchronize(sl.stream);
            if (sl.ev) (void)hipEventDestroy(sl.ev);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
            if (sl.d_buf) (void)hipFree(sl.d_buf);
            if (sl.h_buf) (void)hipHostFree(sl.h_buf);
        }
        if (s->arena) (void)hipFree(s->arena);
    }
    delete s;
    return SPG_OK;
}

}  // extern "C"
