// Concurrency test of the serving loop through its executor seam, on the
// CPU: four client threads each submit 150 requests (encodes and
// reconstructs of varying k, m over host memory), cancel some, wait on each
// (some with a short time-out) and read the counters, while the scheduler
// thread admits, batches, times out and retires them. The executor finishes
// a batch a few hundred microseconds after its launch, fails every 7th
// launch and every 11th completion. At the end every request must be
// accounted for exactly once (completed + failed + cancelled + timed out =
// submitted).
//
//   serve_threads [batch_us [queue_timeout_s]]   defaults 200 and 0.02
//   serve_threads 3000 0.002                      makes queue time-outs fire
//
// Built plainly (make serve_threads) it checks the counters; built as
// serve_threads_tsan it runs the library's host code under ThreadSanitizer.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "fgf_amd.h"

static int slow_us = 200;
struct Exec {
    std::mutex mu;
    std::vector<std::chrono::steady_clock::time_point> due;
    std::vector<int> fail;
    std::atomic<int> launches{0};
};

static int ex_launch(void *ctx, const fgf_server_batch *b) {
    Exec *e = (Exec *)ctx;
    int n = ++e->launches;
    // Touch the request data the way an executor would.
    size_t sum = 0;
    for (size_t i = 0; i < b->n; ++i) {
        sum += b->reqs[i].k + b->rows[i] + b->sources[i][0] + b->targets[i][0] + (size_t)b->values[i][0];
        for (uint32_t j = 0; j < b->rows[i]; ++j) memset(b->reqs[i].shards[b->targets[i][j]], (int)(sum & 0xff), b->reqs[i].row_len);
    }
    if (n % 7 == 0) return FGF_EDEVICE;
    std::lock_guard<std::mutex> lk(e->mu);
    e->due[b->slot] = std::chrono::steady_clock::now() + std::chrono::microseconds(slow_us + (n % 5) * 100);
    e->fail[b->slot] = (n % 11 == 0);
    return FGF_OK;
}

static int ex_poll(void *ctx, uint32_t slot) {
    Exec *e = (Exec *)ctx;
    std::lock_guard<std::mutex> lk(e->mu);
    if (std::chrono::steady_clock::now() < e->due[slot]) return 1;
    return e->fail[slot] ? FGF_EDEVICE : FGF_OK;
}

int main(int argc, char **argv) {
    if (argc > 1) slow_us = atoi(argv[1]);
    Exec ex;
    ex.due.resize(3);
    ex.fail.resize(3);
    fgf_server_executor exe{&ex, ex_launch, ex_poll};
    fgf_server_config cfg{};
    cfg.hbm_budget_bytes = 1 << 20;
    cfg.slots = 3;
    cfg.max_batch_stripes = 4;
    cfg.plan_cache_entries = 5;
    cfg.dwords = 8;
    cfg.rows_per_item = 8;
    cfg.subtiles = 1;
    cfg.waves_per_cu = 8;
    cfg.queue_timeout_s = argc > 2 ? atof(argv[2]) : 0.02;
    cfg.plan_cache_bytes = 4000;
    cfg.executor = &exe;
    fgf_server *s = nullptr;
    if (fgf_server_create(&cfg, &s) != FGF_OK) { printf("create failed: %s\n", fgf_last_error()); return 1; }
    const size_t L = 256;
    std::atomic<int> bad{0};
    auto client = [&](int id) {
        std::vector<std::vector<uint8_t>> mem;
        for (int r = 0; r < 150; ++r) {
            const uint32_t k = 2 + (r + id) % 9, m = 1 + (r * 3 + id) % 4;
            std::vector<uint8_t> buf((k + m) * L + 16);
            uint8_t *base = (uint8_t *)(((uintptr_t)buf.data() + 15) & ~(uintptr_t)15);
            std::vector<uint8_t *> shards(k + m);
            for (uint32_t i = 0; i < k + m; ++i) shards[i] = base + i * L;
            uint32_t erased[1] = {(uint32_t)(r % (k + m))};
            fgf_request q{FGF_GF8D, (r % 3) == 0, k, m, shards.data(), L, erased, (r % 3) == 0 ? 1u : 0u};
            uint64_t t = 0;
            if (fgf_server_submit(s, &q, &t) != FGF_OK) { ++bad; continue; }
            if (r % 5 == 1) fgf_server_cancel(s, t);           // may or may not still be queued
            int st = 0;
            fgf_server_wait(s, t, r % 4 == 0 ? 0.0005 : -1.0, &st);
            fgf_server_stats stats;
            fgf_server_stats_get(s, &stats);
            mem.push_back(std::move(buf));                      // shards stay valid until the end
        }
        fgf_server_drain(s);
    };
    std::vector<std::thread> th;
    for (int i = 0; i < 4; ++i) th.emplace_back(client, i);
    for (auto &t : th) t.join();
    fgf_server_drain(s);
    fgf_server_stats st;
    fgf_server_stats_get(s, &st);
    printf("submitted %llu completed %llu failed %llu cancelled %llu timed_out %llu launches %llu evictions %llu bad %d\n",
           (unsigned long long)st.submitted, (unsigned long long)st.completed, (unsigned long long)st.failed,
           (unsigned long long)st.cancelled, (unsigned long long)st.timed_out, (unsigned long long)st.launches,
           (unsigned long long)st.plan_evictions, bad.load());
    const bool ok = st.completed + st.failed + st.cancelled + st.timed_out == st.submitted && st.submitted == 600;
    fgf_server_destroy(s);
    printf(ok ? "counts consistent\n" : "COUNTS INCONSISTENT\n");
    return ok ? 0 : 1;
}
