// Erasure-coding service over the batched kernel: continuous batching with
// HBM-budget admission, per-slot streams and a device plan cache.
//
// Requests queue in arrival order. A scheduler thread loops:
//   1. retire slots whose launch has finished (hipEventQuery): stamp the
//      completion time of their requests, release their budget and plans;
//   2. if a slot is free, admit queued requests of the head request's field
//      in order while the admitted-but-unfinished bytes stay within the HBM
//      budget and the slot has room, resolve their coefficient plans through
//      the cache, upload the slot's descriptors from pinned staging and
//      launch the persistent kernel on the slot's stream.
// Launches on different slots overlap on the device, so requests that arrive
// while others run start as soon as a slot and budget are free.
//
// The device work (launch a slot's batch, poll its end) goes through an
// executor: the HIP one below, or one the caller supplies, over which the
// same scheduler runs with no device at all. Requests still queued can be
// cancelled or time out; a batch whose launch or completion fails fails
// exactly its own requests, and the loop goes on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "fgf_amd.h"
#include "field.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace {

using Clock = std::chrono::steady_clock;

int sfail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int sfail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    fgf::set_last_error(buf);
    return FGF_EINVAL;
}

int sdev(const char *what, hipError_t e) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    fgf::set_last_error(m.c_str());
    return FGF_EDEVICE;
}

struct Req {
    uint64_t ticket;
    fgf_request r;
    std::vector<uint8_t *> shards;
    std::vector<uint32_t> erased;
    uint64_t bytes;
    Clock::time_point submitted;
};

// Plan cache key: field, encode/reconstruct, k, m and the erased set as a
// bitmask over the k + m <= 256 shards.
struct PlanKey {
    uint32_t head;        // field | reconstruct << 8 | k << 9 (k <= 64)
    uint32_t m;
    uint64_t erased[4];
    bool operator==(const PlanKey &o) const {
        return head == o.head && m == o.m && erased[0] == o.erased[0] && erased[1] == o.erased[1] &&
               erased[2] == o.erased[2] && erased[3] == o.erased[3];
    }
};

struct PlanKeyHash {
    size_t operator()(const PlanKey &k) const {
        uint64_t h = ((uint64_t)k.head << 32 | k.m) * 0x9E3779B97F4A7C15ull;
        for (uint64_t e : k.erased) h = (h ^ e) * 0x9E3779B97F4A7C15ull;
        return (size_t)(h ^ (h >> 29));
    }
};

struct PlanEntry {
    fgf_plan *plan = nullptr;          // device tables (HIP executor only)
    std::vector<uint64_t> vals;        // k x rows coefficients, (t, j) at t * rows + j
    uint32_t rows = 0;
    uint64_t bytes = 0;                // device table bytes the entry stands for
    uint32_t inflight = 0;
    std::list<PlanKey>::iterator lru;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    // One descriptor block per launch, uploaded with one copy: the shard
    // pointer table, the stripe descriptors and the work items, each at a
    // 256-byte aligned offset (device memory, with pinned host staging).
    size_t cap_blob = 0;
    uint8_t *d_blob = nullptr, *h_blob = nullptr;
    uint32_t *d_counter = nullptr;
    std::vector<Req> reqs;
    std::vector<PlanEntry *> plans;
    uint64_t bytes = 0;
    // A caller-supplied executor's batch points into these until the slot
    // is polled finished.
    std::vector<fgf_request> creqs;
    std::vector<uint64_t> tickets;
    std::vector<uint32_t> rows, idx;   // idx: each request's sources, then its targets
    std::vector<const uint32_t *> srcp, tgtp;
    std::vector<const uint64_t *> valp;
};

}  // namespace

struct fgf_server {
    fgf_server_config cfg{};
    int device = 0;
    int cus = 256;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    // Queued requests per field, each in arrival (ticket) order; the head
    // request overall is the front with the smallest ticket.
    std::deque<Req> queues[fgf::FIELD_COUNT];
    size_t queued = 0;
    bool stop = false;
    uint64_t next_ticket = 0;
    uint64_t inflight_bytes = 0;
    std::vector<double> latency;   // per ticket, -1 while pending
    std::vector<int> status;
    fgf_server_stats stats{};
    std::unordered_map<PlanKey, PlanEntry, PlanKeyHash> cache;
    std::list<PlanKey> lru;
    uint64_t plan_bytes = 0;
    bool ext = false;                  // a caller-supplied executor
    fgf_server_executor exec{};
    std::thread worker;
};

namespace {

using namespace fgf;

// Host matrices for the systematic code [I_k ; C], C[i][j] = 1/(i ^ (m + j)).
uint64_t cauchy(int f, uint32_t m, uint32_t i, uint32_t j) {
    return HostField::inv(f, (uint64_t)(i ^ (m + j)));
}

std::vector<uint64_t> generator_row(int f, uint32_t k, uint32_t m, uint32_t shard) {
    std::vector<uint64_t> row(k, 0);
    if (shard < k)
        row[shard] = 1;
    else
        for (uint32_t j = 0; j < k; ++j) row[j] = cauchy(f, m, shard - k, j);
    return row;
}

// Gauss-Jordan inverse over the field; false if singular.
bool invert(int f, std::vector<uint64_t> a, uint32_t n, std::vector<uint64_t> &inv) {
    inv.assign((size_t)n * n, 0);
    for (uint32_t i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1;
    for (uint32_t c = 0; c < n; ++c) {
        uint32_t p = c;
        while (p < n && a[(size_t)p * n + c] == 0) ++p;
        if (p == n) return false;
        if (p != c)
            for (uint32_t x = 0; x < n; ++x) {
                std::swap(a[(size_t)c * n + x], a[(size_t)p * n + x]);
                std::swap(inv[(size_t)c * n + x], inv[(size_t)p * n + x]);
            }
        const uint64_t s = HostField::inv(f, a[(size_t)c * n + c]);
        for (uint32_t x = 0; x < n; ++x) {
            a[(size_t)c * n + x] = HostField::mul(f, a[(size_t)c * n + x], s);
            inv[(size_t)c * n + x] = HostField::mul(f, inv[(size_t)c * n + x], s);
        }
        for (uint32_t r = 0; r < n; ++r) {
            const uint64_t g = a[(size_t)r * n + c];
            if (r == c || g == 0) continue;
            for (uint32_t x = 0; x < n; ++x) {
                a[(size_t)r * n + x] ^= HostField::mul(f, g, a[(size_t)c * n + x]);
                inv[(size_t)r * n + x] ^= HostField::mul(f, g, inv[(size_t)c * n + x]);
            }
        }
    }
    return true;
}

// Survivors: the first k shards that are not erased.
std::vector<uint32_t> survivors_of(const Req &q) {
    std::vector<uint32_t> s;
    for (uint32_t i = 0; i < q.r.k + q.r.m && s.size() < q.r.k; ++i)
        if (std::find(q.erased.begin(), q.erased.end(), i) == q.erased.end()) s.push_back(i);
    return s;
}

PlanKey plan_key(const Req &q) {
    PlanKey k{(uint32_t)q.r.field | (q.r.reconstruct ? 1u : 0u) << 8 | q.r.k << 9, q.r.m, {0, 0, 0, 0}};
    for (uint32_t e : q.erased) k.erased[e >> 6] |= 1ull << (e & 63);
    return k;
}

// Plan values (sources, rows) for a request: encode rows are the parity
// shards; decode rows are G[e] * S^-1 over the k survivors S.
bool plan_values(const Req &q, std::vector<uint64_t> &vals, uint32_t &rows) {
    const int f = q.r.field;
    const uint32_t k = q.r.k, m = q.r.m;
    if (!q.r.reconstruct) {
        rows = m;
        vals.resize((size_t)k * m);
        for (uint32_t t = 0; t < k; ++t)
            for (uint32_t j = 0; j < m; ++j) vals[(size_t)t * m + j] = cauchy(f, m, j, t);
        return true;
    }
    const std::vector<uint32_t> surv = survivors_of(q);
    std::vector<uint64_t> s((size_t)k * k), sinv;
    for (uint32_t r = 0; r < k; ++r) {
        const std::vector<uint64_t> g = generator_row(f, k, m, surv[r]);
        std::copy(g.begin(), g.end(), s.begin() + (size_t)r * k);
    }
    if (!invert(f, s, k, sinv)) return false;
    rows = (uint32_t)q.erased.size();
    vals.assign((size_t)k * rows, 0);
    for (uint32_t j = 0; j < rows; ++j) {
        const std::vector<uint64_t> g = generator_row(f, k, m, q.erased[j]);
        for (uint32_t t = 0; t < k; ++t) {
            uint64_t acc = 0;
            for (uint32_t x = 0; x < k; ++x) acc ^= HostField::mul(f, g[x], sinv[(size_t)x * k + t]);
            vals[(size_t)t * rows + j] = acc;
        }
    }
    return true;
}

// Device bytes of a plan's tables: entry_dwords(B) per coefficient.
uint64_t plan_table_bytes(int f, uint32_t k, uint32_t rows) {
    return (uint64_t)k * rows * entry_dwords(field_bytes(f)) * sizeof(uint32_t);
}

void evict(fgf_server *s, std::list<PlanKey>::reverse_iterator victim) {
    auto ce = s->cache.find(*victim);
    fgf_plan_free(ce->second.plan);
    s->plan_bytes -= ce->second.bytes;
    s->lru.erase(std::next(victim).base());
    s->cache.erase(ce);
    ++s->stats.plan_evictions;
    s->stats.plan_cache_bytes = s->plan_bytes;
}

// The request's plan from the cache, or built and cached (evicting
// least-recently-used plans that no launch is reading). FGF_OK with *out
// set; 1 when the plan fits the byte cap only after plans in flight are
// released (the request waits); negative when the request fails: FGF_EINVAL
// for a singular decode matrix, FGF_ENOMEM when the plan exceeds the cap on
// its own or its device tables cannot be allocated.
int acquire_plan(fgf_server *s, const Req &q, PlanEntry **out) {
    const PlanKey key = plan_key(q);
    auto it = s->cache.find(key);
    if (it != s->cache.end()) {
        s->lru.splice(s->lru.begin(), s->lru, it->second.lru);
        ++s->stats.plan_hits;
        ++it->second.inflight;
        *out = &it->second;
        return FGF_OK;
    }
    std::vector<uint64_t> vals;
    uint32_t rows = 0;
    if (!plan_values(q, vals, rows)) return FGF_EINVAL;
    const uint64_t bytes = plan_table_bytes(q.r.field, q.r.k, rows);
    const uint64_t cap = s->cfg.plan_cache_bytes;
    if (cap && bytes > cap) return FGF_ENOMEM;
    auto over = [&] {
        return s->cache.size() >= s->cfg.plan_cache_entries || (cap && s->plan_bytes + bytes > cap);
    };
    while (over()) {
        auto victim = std::find_if(s->lru.rbegin(), s->lru.rend(),
                                   [&](const PlanKey &k2) { return s->cache[k2].inflight == 0; });
        if (victim == s->lru.rend()) break;
        evict(s, victim);
    }
    // Every remaining plan is in flight: wait for their launches (the entry
    // count may run over instead, as it always could).
    if (cap && s->plan_bytes + bytes > cap) return 1;
    fgf_plan *p = nullptr;
    if (!s->ext) {
        const int nb = field_bytes(q.r.field);
        std::vector<uint8_t> raw(vals.size() * nb);
        for (size_t i = 0; i < vals.size(); ++i)
            for (int b = 0; b < nb; ++b) raw[i * nb + b] = (uint8_t)(vals[i] >> (8 * b));
        const int rc = fgf_plan_matrix((fgf_field)q.r.field, q.r.k, rows, raw.data(), vals.size(), &p);
        if (rc != FGF_OK) return rc == FGF_EINVAL ? FGF_EINVAL : FGF_ENOMEM;
    }
    ++s->stats.plan_misses;
    s->lru.push_front(key);
    PlanEntry &e = s->cache[key];
    e.plan = p;
    e.vals = std::move(vals);
    e.rows = rows;
    e.bytes = bytes;
    e.inflight = 1;
    e.lru = s->lru.begin();
    s->plan_bytes += bytes;
    s->stats.plan_cache_bytes = s->plan_bytes;
    *out = &e;
    return FGF_OK;
}

template <typename T>
hipError_t grow(T *&dev, T *&host, size_t &cap, size_t need) {
    if (need <= cap) return hipSuccess;
    size_t n = std::max<size_t>(need, cap * 2);
    if (dev) (void)hipFree(dev);
    if (host) (void)hipHostFree(host);
    dev = nullptr;
    host = nullptr;
    hipError_t e = hipMalloc((void **)&dev, n * sizeof(T));
    if (e == hipSuccess) e = hipHostMalloc((void **)&host, n * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) cap = n;
    return e;
}

void finish_slot(fgf_server *s, Slot &sl, int status) {
    const Clock::time_point now = Clock::now();
    for (Req &q : sl.reqs) {
        s->latency[q.ticket] = std::chrono::duration<double>(now - q.submitted).count();
        s->status[q.ticket] = status;
        if (status == 0) {
            ++s->stats.completed;
            s->stats.bytes += q.bytes;
        } else {
            ++s->stats.failed;
        }
    }
    for (PlanEntry *p : sl.plans) --p->inflight;
    s->inflight_bytes -= sl.bytes;
    sl.reqs.clear();
    sl.plans.clear();
    sl.bytes = 0;
    sl.busy = false;
    s->cv_done.notify_all();
}

constexpr size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Bytes of a slot's descriptor block for the given counts.
size_t blob_bytes(size_t nptrs, size_t nstripes, size_t nitems) {
    return align256(nptrs * sizeof(void *)) + align256(nstripes * sizeof(StripeDesc)) +
           nitems * sizeof(WorkItem);
}

// Kernel shape of a launch for one field: the configured dwords and rows
// per item where the element width allows them, else the width's own
// (4- and 8-byte towers take 8 dwords and 4 or 2 rows per item).
void tile_for(const fgf_server *s, int nb, int &dwords, int &rows) {
    dwords = std::max(s->cfg.dwords, batch_min_dwords(nb));
    rows = batch_rows_allowed(nb, s->cfg.rows_per_item) ? s->cfg.rows_per_item : batch_default_rows(nb);
}

// Builds and launches one slot's batch from the admitted requests.
int launch_slot(fgf_server *s, Slot &sl) {
    const int f = sl.reqs.front().r.field;
    const int nb = field_bytes(f);
    const size_t edw = entry_dwords(nb);
    int dwords = 0, rows_per_item = 0;
    tile_for(s, nb, dwords, rows_per_item);
    // Stripe geometry first (the work items need only lengths and counts),
    // then the block is sized, then the device pointers are filled in.
    std::vector<StripeDesc> stripes(sl.reqs.size());
    size_t nptrs = 0;
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        const Req &q = sl.reqs[i];
        const fgf_plan *plan = sl.plans[i]->plan;
        StripeDesc &d = stripes[i];
        d.row_len = q.r.row_len;
        d.k = (uint32_t)plan->rows;
        d.m = (uint32_t)plan->cols;
        d.tab = plan->d_tab;
        d.tab_stride_t = (uint32_t)(plan->cols * edw);
        d.tab_stride_j = (uint32_t)edw;
        d.overwrite = 1;
        d.pad = 0;
        nptrs += plan->rows + plan->cols;
    }
    // Item size: a launch too small to give every resident wave one wave
    // tile uses items of one tile per wave, spreading a lone request over as
    // many waves as it has tiles (its latency is one wave's serial work);
    // larger launches use the configured multi-tile items.
    const uint64_t tile_bytes = 256u * (uint64_t)dwords;
    uint64_t tiles = 0;
    for (const StripeDesc &d : stripes)
        tiles += (d.row_len + tile_bytes - 1) / tile_bytes *
                 ((d.m + (uint32_t)rows_per_item - 1) / (uint32_t)rows_per_item);
    const int subtiles =
        tiles <= (uint64_t)s->cus * (uint64_t)s->cfg.waves_per_cu ? 1 : s->cfg.subtiles;
    std::vector<WorkItem> items = build_work_items(stripes, dwords, rows_per_item, subtiles, 1);
    const size_t off_str = align256(nptrs * sizeof(void *));
    const size_t off_items = off_str + align256(stripes.size() * sizeof(StripeDesc));
    const size_t total = blob_bytes(nptrs, stripes.size(), items.size());
    hipError_t e = grow(sl.d_blob, sl.h_blob, sl.cap_blob, total);
    if (e != hipSuccess) return sdev("slot buffers", e);
    const void **h_ptrs = (const void **)sl.h_blob;
    const void **d_ptrs = (const void **)sl.d_blob;
    size_t p = 0;
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        const Req &q = sl.reqs[i];
        StripeDesc &d = stripes[i];
        const size_t p0 = p;
        if (!q.r.reconstruct) {
            for (uint32_t t = 0; t < q.r.k; ++t) h_ptrs[p++] = q.shards[t];
            for (uint32_t j = 0; j < q.r.m; ++j) h_ptrs[p++] = q.shards[q.r.k + j];
        } else {
            for (uint32_t t : survivors_of(q)) h_ptrs[p++] = q.shards[t];
            for (uint32_t e2 : q.erased) h_ptrs[p++] = q.shards[e2];
        }
        d.srcs = (const uint8_t *const *)(d_ptrs + p0);
        d.rows = (uint8_t *const *)(d_ptrs + p0 + d.k);
    }
    std::memcpy(sl.h_blob + off_str, stripes.data(), stripes.size() * sizeof(StripeDesc));
    std::memcpy(sl.h_blob + off_items, items.data(), items.size() * sizeof(WorkItem));
    if ((e = hipMemcpyAsync(sl.d_blob, sl.h_blob, total, hipMemcpyHostToDevice, sl.stream)))
        return sdev("descriptor upload", e);
    const uint32_t wpb = batch_block_threads() / 64;
    const uint64_t want = (uint64_t)s->cus * s->cfg.waves_per_cu;
    const uint64_t waves = std::max<uint64_t>(1, std::min<uint64_t>(want, items.size() * wpb));
    const uint32_t blocks = (uint32_t)((waves + wpb - 1) / wpb);
    BatchArgs a{(const StripeDesc *)(sl.d_blob + off_str), (const WorkItem *)(sl.d_blob + off_items),
                (uint32_t)items.size(), blocks * wpb, (uint32_t)subtiles, sl.d_counter};
    if ((e = launch_batch(f, dwords, rows_per_item, a, blocks, sl.stream)) ||
        (e = hipEventRecord(sl.done, sl.stream)))
        return sdev("batch launch", e);
    sl.busy = true;
    return FGF_OK;
}

// Ends a request that never reached a launch (lock held; the caller takes
// it off its queue).
void end_unlaunched(fgf_server *s, const Req &q, int status) {
    s->latency[q.ticket] = std::chrono::duration<double>(Clock::now() - q.submitted).count();
    s->status[q.ticket] = status;
    if (status == FGF_ECANCELED)
        ++s->stats.cancelled;
    else if (status == FGF_ETIMEDOUT)
        ++s->stats.timed_out;
    else
        ++s->stats.failed;
    --s->queued;
    s->cv_done.notify_all();
}

// Hands a slot's admitted requests to the caller-supplied executor.
int launch_slot_ext(fgf_server *s, Slot &sl) {
    const size_t n = sl.reqs.size();
    sl.creqs.resize(n);
    sl.tickets.resize(n);
    sl.rows.resize(n);
    sl.srcp.resize(n);
    sl.tgtp.resize(n);
    sl.valp.resize(n);
    sl.idx.clear();
    for (size_t i = 0; i < n; ++i) {
        const Req &q = sl.reqs[i];
        if (!q.r.reconstruct) {
            for (uint32_t t = 0; t < q.r.k + q.r.m; ++t) sl.idx.push_back(t);
        } else {
            for (uint32_t t : survivors_of(q)) sl.idx.push_back(t);
            for (uint32_t e : q.erased) sl.idx.push_back(e);
        }
    }
    size_t off = 0;
    for (size_t i = 0; i < n; ++i) {
        const Req &q = sl.reqs[i];
        const PlanEntry *pe = sl.plans[i];
        sl.creqs[i] = q.r;
        sl.creqs[i].shards = q.shards.data();
        sl.creqs[i].erased = q.erased.empty() ? nullptr : q.erased.data();
        sl.tickets[i] = q.ticket;
        sl.rows[i] = pe->rows;
        sl.srcp[i] = sl.idx.data() + off;
        sl.tgtp[i] = sl.idx.data() + off + q.r.k;
        sl.valp[i] = pe->vals.data();
        off += q.r.k + pe->rows;
    }
    const fgf_server_batch b{(uint32_t)(&sl - s->slots.data()), (fgf_field)sl.reqs.front().r.field, n,
                             sl.creqs.data(), sl.tickets.data(), sl.rows.data(), sl.srcp.data(),
                             sl.tgtp.data(), sl.valp.data()};
    const int rc = s->exec.launch(s->exec.ctx, &b);
    if (rc > 0) return sfail("executor launch returned %d", rc);
    if (rc == FGF_OK) sl.busy = true;
    return rc;
}

// 1 while the slot's batch runs, FGF_OK when it has finished, negative when
// it failed (lock held).
int poll_slot(fgf_server *s, Slot &sl) {
    if (s->ext) {
        const int r = s->exec.poll(s->exec.ctx, (uint32_t)(&sl - s->slots.data()));
        return r > 0 ? 1 : r;
    }
    const hipError_t q = hipEventQuery(sl.done);
    return q == hipSuccess ? FGF_OK : q == hipErrorNotReady ? 1 : FGF_EDEVICE;
}

// Queued requests older than the queue time-out (lock held). Each field's
// queue is in submission order, so only fronts can have expired first.
void expire_queued(fgf_server *s) {
    if (s->cfg.queue_timeout_s <= 0) return;
    const Clock::time_point now = Clock::now();
    for (std::deque<Req> &fq : s->queues)
        while (!fq.empty() &&
               std::chrono::duration<double>(now - fq.front().submitted).count() > s->cfg.queue_timeout_s) {
            end_unlaunched(s, fq.front(), FGF_ETIMEDOUT);
            fq.pop_front();
        }
}

void scheduler(fgf_server *s) {
    if (!s->ext) (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool any_busy = false;
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            const int r = poll_slot(s, sl);
            if (r == 1)
                any_busy = true;
            else
                finish_slot(s, sl, r);
        }
        expire_queued(s);
        Slot *free_slot = nullptr;
        for (Slot &sl : s->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        bool launched = false;
        if (free_slot && s->queued) {
            // The head request's field: the queue front with the oldest ticket.
            int field = -1;
            for (int f = 0; f < FIELD_COUNT; ++f)
                if (!s->queues[f].empty() &&
                    (field < 0 || s->queues[f].front().ticket < s->queues[field].front().ticket))
                    field = f;
            std::deque<Req> &fq = s->queues[field];
            while (!fq.empty() && free_slot->reqs.size() < s->cfg.max_batch_stripes) {
                Req &q = fq.front();
                if (s->inflight_bytes + q.bytes > s->cfg.hbm_budget_bytes) {
                    if (free_slot->reqs.empty() && !any_busy) {
                        // Larger than the whole budget: can never be admitted.
                        end_unlaunched(s, q, FGF_ENOMEM);
                        fq.pop_front();
                        continue;
                    }
                    ++s->stats.admission_waits;
                    break;  // in-order admission: wait for budget
                }
                PlanEntry *pe = nullptr;
                const int prc = acquire_plan(s, q, &pe);
                if (prc == 1) {
                    // The plan cache's byte cap is held by plans in flight
                    // (this slot's too, launched below): wait for them.
                    ++s->stats.admission_waits;
                    break;
                }
                if (prc != FGF_OK) {
                    end_unlaunched(s, q, prc);
                    fq.pop_front();
                    continue;
                }
                s->inflight_bytes += q.bytes;
                free_slot->bytes += q.bytes;
                free_slot->plans.push_back(pe);
                free_slot->reqs.push_back(std::move(q));
                fq.pop_front();
                --s->queued;
            }
            s->stats.max_inflight_bytes = std::max(s->stats.max_inflight_bytes, s->inflight_bytes);
            if (!free_slot->reqs.empty()) {
                // Only this thread touches slots: submissions proceed meanwhile.
                lk.unlock();
                const int rc = s->ext ? launch_slot_ext(s, *free_slot) : launch_slot(s, *free_slot);
                lk.lock();
                if (rc != FGF_OK) {
                    finish_slot(s, *free_slot, rc < 0 ? rc : FGF_EDEVICE);
                } else {
                    ++s->stats.launches;
                    s->stats.stripes_launched += free_slot->reqs.size();
                }
                launched = true;
            }
        }
        if (s->stop && !s->queued && !any_busy && !launched) break;
        if (!launched) {
            if (any_busy && !s->ext) {
                // A launch is in flight: poll its completion event at once
                // again (a timed wait would oversleep by tens of us). The
                // scheduler thread spins while the device works.
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            } else if (any_busy) {
                s->cv_work.wait_for(lk, std::chrono::microseconds(200));
            } else {
                s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
    s->cv_done.notify_all();
}

}  // namespace

extern "C" {

int fgf_server_create(const fgf_server_config *cfg, fgf_server **out) {
    if (!cfg || !out) return sfail("fgf_server_create: null argument");
    if (cfg->slots < 1 || cfg->slots > 16) return sfail("fgf_server_create: slots must be in 1..16");
    if (cfg->max_batch_stripes < 1) return sfail("fgf_server_create: max_batch_stripes must be >= 1");
    if (cfg->plan_cache_entries < 1) return sfail("fgf_server_create: plan_cache_entries must be >= 1");
    if (cfg->dwords != 4 && cfg->dwords != 8) return sfail("fgf_server_create: dwords must be 4 or 8");
    if (cfg->rows_per_item != 4 && cfg->rows_per_item != 8)
        return sfail("fgf_server_create: rows_per_item must be 4 or 8");
    if (cfg->subtiles < 1 || cfg->subtiles > 64) return sfail("fgf_server_create: subtiles must be in 1..64");
    if (cfg->waves_per_cu < 1 || cfg->waves_per_cu > 32)
        return sfail("fgf_server_create: waves_per_cu must be in 1..32");
    if (!(cfg->queue_timeout_s >= 0) || cfg->queue_timeout_s > 1e9)
        return sfail("fgf_server_create: queue_timeout_s must be a number of seconds >= 0");
    if (cfg->executor && (!cfg->executor->launch || !cfg->executor->poll))
        return sfail("fgf_server_create: executor needs both launch and poll");
    auto *s = new (std::nothrow) fgf_server;
    if (!s) return sfail("out of host memory");
    s->cfg = *cfg;
    s->cfg.executor = nullptr;
    s->slots.resize(cfg->slots);
    if (cfg->executor) {
        // No device resources: the executor does the device work.
        s->ext = true;
        s->exec = *cfg->executor;
        s->worker = std::thread(scheduler, s);
        *out = s;
        return FGF_OK;
    }
    (void)hipGetDevice(&s->device);
    (void)hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device);
    // Descriptor blocks sized for a full batch up front: growing one later
    // frees memory, and hipFree waits for every stream of the device.
    const size_t want = blob_bytes((size_t)cfg->max_batch_stripes * 32, cfg->max_batch_stripes,
                                   (size_t)cfg->max_batch_stripes * 128);
    for (Slot &sl : s->slots) {
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) ||
            (e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) ||
            (e = hipMalloc((void **)&sl.d_counter, 2 * sizeof(uint32_t))) ||
            (e = hipMemset(sl.d_counter, 0, 2 * sizeof(uint32_t))) ||
            (e = grow(sl.d_blob, sl.h_blob, sl.cap_blob, want))) {
            fgf_server_destroy(s);
            return sdev("fgf_server_create", e);
        }
    }
    s->worker = std::thread(scheduler, s);
    *out = s;
    return FGF_OK;
}

}  // extern "C"

namespace {

// Validates a request and builds its queue entry (no lock held).
int make_req(fgf_server *s, const fgf_request *r, Req &q) {
    if (!s || !r) return sfail("fgf_server_submit: null argument");
    const int f = r->field;
    if (!fgf::valid_field(f)) return sfail("fgf_server_submit: unknown field %d", f);
    if (r->k < 1 || r->m < 1 || r->k + r->m > 256)
        return sfail("fgf_server_submit: k = %u, m = %u outside 1 <= k, 1 <= m, k + m <= 256", r->k, r->m);
    if (r->k > 64) return sfail("fgf_server_submit: k = %u exceeds 64 sources", r->k);
    if (!r->shards) return sfail("fgf_server_submit: shards is null");
    if (r->row_len == 0 || r->row_len % 16)
        return sfail("fgf_server_submit: row_len %zu is not a nonzero multiple of 16", r->row_len);
    for (uint32_t i = 0; i < r->k + r->m; ++i)
        if ((uintptr_t)r->shards[i] & 15u) return sfail("fgf_server_submit: shard %u is not 16-byte aligned", i);
    // The shards of a stripe are distinct buffers: an overlap would race the
    // kernel's stores with its loads.
    std::vector<uintptr_t> spans(r->k + r->m);
    for (uint32_t i = 0; i < r->k + r->m; ++i) spans[i] = (uintptr_t)r->shards[i];
    std::sort(spans.begin(), spans.end());
    for (size_t i = 1; i < spans.size(); ++i)
        if (spans[i] - spans[i - 1] < r->row_len)
            return sfail("fgf_server_submit: shards overlap for %zu-byte rows", r->row_len);
    q.r = *r;
    q.shards.assign(r->shards, r->shards + r->k + r->m);
    if (r->reconstruct) {
        if (r->nerased < 1 || r->nerased > r->m || !r->erased)
            return sfail("fgf_server_submit: reconstruct needs 1..m erased shards");
        q.erased.assign(r->erased, r->erased + r->nerased);
        std::sort(q.erased.begin(), q.erased.end());
        for (size_t i = 0; i < q.erased.size(); ++i)
            if (q.erased[i] >= r->k + r->m || (i && q.erased[i] == q.erased[i - 1]))
                return sfail("fgf_server_submit: erased shard list is invalid");
    }
    q.r.shards = nullptr;
    q.r.erased = nullptr;
    q.bytes = (uint64_t)(r->k + r->m) * r->row_len;
    return FGF_OK;
}

// Enqueues built requests in order under one lock; each gets its ticket
// and its submission time.
void enqueue(fgf_server *s, Req *qs, size_t n, uint64_t *tickets) {
    const Clock::time_point now = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (size_t i = 0; i < n; ++i) {
            Req &q = qs[i];
            q.submitted = now;
            q.ticket = s->next_ticket++;
            s->latency.push_back(-1.0);
            s->status.push_back(0);
            ++s->stats.submitted;
            if (tickets) tickets[i] = q.ticket;
            s->queues[q.r.field].push_back(std::move(q));
            ++s->queued;
        }
    }
    s->cv_work.notify_one();
}

}  // namespace

extern "C" {

int fgf_server_submit(fgf_server *s, const fgf_request *r, uint64_t *ticket) {
    Req q;
    const int rc = make_req(s, r, q);
    if (rc != FGF_OK) return rc;
    enqueue(s, &q, 1, ticket);
    return FGF_OK;
}

int fgf_server_drain(fgf_server *s) {
    if (!s) return sfail("fgf_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        const fgf_server_stats &st = s->stats;
        return st.completed + st.failed + st.cancelled + st.timed_out == s->next_ticket;
    });
    return s->stats.failed ? FGF_EDEVICE : s->stats.timed_out ? FGF_ETIMEDOUT : FGF_OK;
}

int fgf_server_cancel(fgf_server *s, uint64_t ticket) {
    if (!s) return sfail("fgf_server_cancel: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket >= s->next_ticket)
        return sfail("fgf_server_cancel: unknown ticket %llu", (unsigned long long)ticket);
    for (std::deque<Req> &fq : s->queues)
        for (auto it = fq.begin(); it != fq.end(); ++it)
            if (it->ticket == ticket) {
                end_unlaunched(s, *it, FGF_ECANCELED);
                fq.erase(it);
                return FGF_OK;
            }
    return sfail("fgf_server_cancel: ticket %llu is not queued (already launched or finished)",
                 (unsigned long long)ticket);
}

int fgf_server_wait(fgf_server *s, uint64_t ticket, double timeout_s, int *status) {
    if (!s) return sfail("fgf_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket >= s->latency.size())
        return sfail("fgf_server_wait: unknown ticket %llu", (unsigned long long)ticket);
    const auto done = [&] { return s->latency[ticket] >= 0; };
    if (timeout_s < 0 || timeout_s > 1e9) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        fgf::set_last_error("fgf_server_wait: timed out");
        return FGF_ETIMEDOUT;
    }
    if (status) *status = s->status[ticket];
    return FGF_OK;
}

int fgf_server_result(fgf_server *s, uint64_t ticket, double *latency_s, int *status) {
    if (!s) return sfail("fgf_server_result: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket >= s->latency.size()) return sfail("fgf_server_result: unknown ticket %llu", (unsigned long long)ticket);
    if (latency_s) *latency_s = s->latency[ticket];
    if (status) *status = s->status[ticket];
    return FGF_OK;
}

int fgf_server_stats_get(fgf_server *s, fgf_server_stats *out) {
    if (!s || !out) return sfail("fgf_server_stats_get: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return FGF_OK;
}

int fgf_server_replay(fgf_server *s, const fgf_request *reqs, const double *arrival, size_t n,
                      double *latency_s, double *wall_s) {
    if (!s || (n && (!reqs || !arrival))) return sfail("fgf_server_replay: null argument");
    std::vector<uint64_t> tickets(n);
    std::vector<double> lateness(n, 0.0);
    int rc = FGF_OK;
    const Clock::time_point t0 = Clock::now();
    auto due_at = [&](size_t i) {
        return t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival[i]));
    };
    std::thread client([&] {
        std::vector<Req> group;
        std::vector<size_t> index;
        std::vector<uint64_t> got;
        for (size_t i = 0; i < n;) {
            while (Clock::now() < due_at(i)) {
            }
            // Every request whose arrival time has passed goes in one group.
            const Clock::time_point now = Clock::now();
            group.clear();
            index.clear();
            for (; i < n && due_at(i) <= now; ++i) {
                lateness[i] = std::chrono::duration<double>(now - due_at(i)).count();
                Req q;
                const int r = make_req(s, &reqs[i], q);
                if (r != FGF_OK) {
                    rc = r;
                    tickets[i] = UINT64_MAX;
                    continue;
                }
                group.push_back(std::move(q));
                index.push_back(i);
            }
            got.resize(group.size());
            enqueue(s, group.data(), group.size(), got.data());
            for (size_t g = 0; g < index.size(); ++g) tickets[index[g]] = got[g];
        }
    });
    client.join();
    const int d = fgf_server_drain(s);
    if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    if (latency_s) {
        std::lock_guard<std::mutex> lk(s->mu);
        for (size_t i = 0; i < n; ++i) {
            if (tickets[i] == UINT64_MAX) {
                latency_s[i] = -1;
                continue;
            }
            // From the scheduled arrival: includes any lateness of the client.
            latency_s[i] = s->latency[tickets[i]] + lateness[i];
        }
    }
    return rc != FGF_OK ? rc : d;
}

void fgf_server_destroy(fgf_server *s) {
    if (!s) return;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.d_blob) (void)hipFree(sl.d_blob);
        if (sl.h_blob) (void)hipHostFree(sl.h_blob);
        if (sl.d_counter) (void)hipFree(sl.d_counter);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (auto &kv : s->cache) fgf_plan_free(kv.second.plan);
    delete s;
}

}  // extern "C"
