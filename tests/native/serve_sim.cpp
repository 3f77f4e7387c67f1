// CPU check of the serving loop's threading (gpupi_amd/csrc/serve.hip built
// against host stand-ins for HIP and the library): many random mixes of pi
// and multiply jobs, submitted from one thread while another collects
// results, under generous and tight budgets.  Every job must finish exactly
// once with its own result, admitted jobs must never exceed the budget, and
// no scenario may stall (watchdog).  Exit status 0 = all checks passed.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../gpupi_amd/csrc/serve.hip"

// ------------------------------------------------ stand-ins (host memory)
struct BatchMPContext {
    int id;
};
static std::atomic<uint64_t> g_in_flight_bytes{0};
static thread_local std::mt19937 t_rng(std::hash<std::thread::id>()(std::this_thread::get_id()));
static void jitter(int max_us) { std::this_thread::sleep_for(std::chrono::microseconds(t_rng() % (max_us + 1))); }

hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipMalloc(void **p, size_t bytes) { *p = malloc(bytes ? bytes : 1); return *p ? hipSuccess : hipErrorOutOfMemory; }
hipError_t hipFree(void *p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void *d, const void *s, size_t b, hipMemcpyKind) { memcpy(d, s, b); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t b, hipMemcpyKind, hipStream_t) { memcpy(d, s, b); return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s) { *s = (hipStream_t)malloc(1); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { free(s); return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { jitter(200); return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
// the pool: host memory, with live and peak byte counts
static std::mutex g_pool_mu;
static std::map<void *, size_t> g_pool_live;
static uint64_t g_pool_used = 0, g_pool_high = 0;
hipError_t hipMemPoolCreate(hipMemPool_t *pool, const hipMemPoolProps *) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool_used = g_pool_high = 0;
    *pool = (hipMemPool_t)malloc(1);
    return hipSuccess;
}
hipError_t hipMemPoolDestroy(hipMemPool_t pool) { free(pool); return hipSuccess; }
hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void *) { return hipSuccess; }
hipError_t hipMemPoolGetAttribute(hipMemPool_t, hipMemPoolAttr attr, void *v) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    *(uint64_t *)v = attr == hipMemPoolAttrUsedMemHigh || attr == hipMemPoolAttrReservedMemHigh ? g_pool_high : 0;
    return hipSuccess;
}
hipError_t hipMallocFromPoolAsync(void **p, size_t bytes, hipMemPool_t, hipStream_t) {
    *p = malloc(bytes ? bytes : 1);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool_live[*p] = bytes;
    g_pool_used += bytes;
    if (g_pool_used > g_pool_high) g_pool_high = g_pool_used;
    return hipSuccess;
}
hipError_t hipFreeAsync(void *p, hipStream_t) {
    if (p == nullptr) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_pool_live.find(p);
    if (it != g_pool_live.end()) {
        g_pool_used -= it->second;
        g_pool_live.erase(it);
    }
    free(p);
    return hipSuccess;
}
hipError_t gpupi_set_memory_pool(BatchMPContext *, hipMemPool_t) { return hipSuccess; }
BatchMPContext *batch_mp_init(void) { jitter(2000); return new BatchMPContext{1}; }
void batch_mp_destroy(BatchMPContext *c) { delete c; }
hipError_t batch_mp_set_stream(BatchMPContext *, hipStream_t) { return hipSuccess; }
// the pi result is pool memory in the server (freed with hipFreeAsync)
static void pool_track(void *p, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool_live[p] = bytes;
    g_pool_used += bytes;
    if (g_pool_used > g_pool_high) g_pool_high = g_pool_used;
}
hipError_t gpupi_pi_run(BatchMPContext *, uint32_t digits, int, GpupiPiResult *r) {
    jitter(3000);
    memset(r, 0, sizeof(*r));
    r->ret_length = digits / 9 + 2;
    r->ret = (uint32_t *)malloc(4 * r->ret_length);
    for (uint32_t i = 0; i < r->ret_length; ++i) r->ret[i] = digits * 7 + i;
    r->digits = (char *)malloc(digits + 1);
    for (uint32_t i = 0; i < digits; ++i) r->digits[i] = (char)('0' + (digits + i) % 10);
    pool_track(r->ret, 4 * r->ret_length);
    pool_track(r->digits, digits + 1);
    return hipSuccess;
}
void gpupi_pi_free(GpupiPiResult *r) { free(r->ret); free(r->digits); }
uint32_t gpupi_pi_groups_for_digits(uint32_t d) {
    uint32_t n = 1;
    while (n * 17 * 14.3 < d) n *= 2;
    return n;
}
static std::atomic<int> g_max_pi_batch{0};
hipError_t gpupi_pi_run_batch(BatchMPContext *c, const uint32_t *digits, int count, int prec, GpupiPiResult *r) {
    for (int i = 0; i < count; ++i) {
        if (gpupi_pi_groups_for_digits(digits[i]) != gpupi_pi_groups_for_digits(digits[0])) return hipErrorInvalidValue;
        gpupi_pi_run(c, digits[i], prec, &r[i]);
    }
    int seen = g_max_pi_batch.load();
    while (count > seen && !g_max_pi_batch.compare_exchange_weak(seen, count)) {
    }
    return hipSuccess;
}
hipError_t gpupi_cache_trim(BatchMPContext *) { jitter(200); return hipSuccess; }
hipError_t gpupi_mul_ragged(BatchMPContext *, const GpupiMulItem *items, uint32_t count) {
    jitter(2000);
    for (uint32_t i = 0; i < count; ++i)  // a stand-in product: C[k] = A[0] + B[0] + k
        for (uint32_t k = 0; k < items[i].L_a + items[i].L_b; ++k)
            items[i].C[k] = (items[i].L_a ? items[i].A[0] : 0) + (items[i].L_b ? items[i].B[0] : 0) + k;
    return hipSuccess;
}

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { if (fails++ < 20) { printf("FAIL: " __VA_ARGS__); printf("\n"); } } } while (0)

// The serialized case: a budget that holds exactly one job at a time.
static void serialized(int pi_workers) {
    const uint64_t budget = gpupi_serve_job_bytes(GPUPI_JOB_PI, 5000, 0);
    GpupiServeConfig cfg{budget, (uint32_t)pi_workers, 0};
    GpupiServer *s = gpupi_server_create(&cfg);
    std::vector<std::vector<uint32_t>> A(3, std::vector<uint32_t>(1u << 19, 1)), C(3, std::vector<uint32_t>(1u << 20));
    std::vector<int64_t> ids;
    ids.push_back(gpupi_server_submit_pi(s, 5000, nullptr));
    for (int i = 0; i < 3; ++i)
        ids.push_back(gpupi_server_submit_mul(s, A[i].data(), 1u << 19, A[i].data(), 1u << 19, C[i].data()));
    for (int64_t id : ids) EXPECT(id >= 0, "serialized submit");
    std::vector<uint32_t> big(1u << 22, 1);
    std::vector<uint32_t> bigc(1u << 23);
    EXPECT(gpupi_server_submit_mul(s, big.data(), 1u << 22, big.data(), 1u << 22, bigc.data()) == -1, "oversize refused");
    int got = 0;
    GpupiJobResult r[8];
    std::vector<GpupiJobResult> all;
    while (got < 4) {
        const int n = gpupi_server_results(s, r, 8, 1);
        for (int i = 0; i < n; ++i) all.push_back(r[i]);
        got += n;
        EXPECT(n > 0, "results returned nothing while jobs were pending");
        if (n == 0) break;
    }
    gpupi_server_destroy(s);
    std::sort(all.begin(), all.end(), [](auto &a, auto &b) { return a.start_ms < b.start_ms; });
    for (size_t i = 1; i < all.size(); ++i) EXPECT(all[i].start_ms >= all[i - 1].finish_ms, "one job at a time");
}

static void scenario(int seed, uint64_t budget, int pi_workers, int jobs) {
    std::mt19937 rng(seed);
    GpupiServeConfig cfg{budget, (uint32_t)pi_workers, 8, 1 + (uint32_t)(seed % 5)};
    GpupiServer *s = gpupi_server_create(&cfg);
    struct Mul {
        std::vector<uint32_t> A, B, C;
    };
    std::map<int64_t, int> kind;
    std::map<int64_t, Mul> muls;
    std::map<int64_t, uint32_t> digits;
    std::mutex mu;
    std::atomic<int> got{0};
    std::vector<GpupiJobResult> results;
    std::atomic<bool> submitted{false};
    std::thread collector([&] {
        GpupiJobResult r[16];
        for (;;) {
            const int n = gpupi_server_results(s, r, 16, 1);
            std::lock_guard<std::mutex> lk(mu);
            for (int i = 0; i < n; ++i) results.push_back(r[i]);
            if (submitted && (int)results.size() == (int)kind.size()) break;
            if (n == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
    });
    for (int j = 0; j < jobs; ++j) {
        int64_t id;
        if (rng() % 10 < 4) {
            const uint32_t d = 100 + rng() % 5000;
            std::lock_guard<std::mutex> lk(mu);
            id = gpupi_server_submit_pi(s, d, nullptr);
            if (id >= 0) {
                kind[id] = 0;
                digits[id] = d;
            }
        } else {
            Mul m;
            m.A.assign(1 + rng() % 3000, rng());
            m.B.assign(1 + rng() % 3000, rng());
            m.C.assign(m.A.size() + m.B.size(), 0);
            std::lock_guard<std::mutex> lk(mu);
            Mul &mm = muls[-1 - j] = std::move(m);
            id = gpupi_server_submit_mul(s, mm.A.data(), (uint32_t)mm.A.size(), mm.B.data(), (uint32_t)mm.B.size(),
                                         mm.C.data());
            if (id >= 0) {
                kind[id] = 1;
                muls[id] = std::move(mm);
            }
            muls.erase(-1 - j);
        }
        EXPECT(id >= 0, "submit refused");
        if (rng() % 3 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 1500));
    }
    submitted = true;
    collector.join();
    uint64_t used_high = 0, reserved_high = 0;
    EXPECT(gpupi_server_memory(s, &used_high, &reserved_high) == hipSuccess, "memory query");
    gpupi_server_destroy(s);
    // the one pool never held more than the budget
    EXPECT(used_high <= budget, "pool peak %llu > budget %llu", (unsigned long long)used_high,
           (unsigned long long)budget);
    std::map<int64_t, int> seen;
    for (auto &r : results) {
        seen[r.id]++;
        EXPECT(r.status == 0, "status");
        EXPECT(r.kind != 0 || (r.batch_size >= 1 && r.batch_size <= cfg.pi_batch_max), "pi batch size");
        EXPECT(r.submit_ms <= r.start_ms && r.start_ms <= r.finish_ms, "timestamps");
        EXPECT(r.kind == kind[r.id], "kind");
        if (r.kind == 0) {
            EXPECT(r.ret_length == digits[r.id] / 9 + 2, "pi result belongs to its job");
        } else {
            const Mul &m = muls[r.id];
            bool ok = true;
            for (size_t k = 0; k < m.C.size(); ++k) ok = ok && m.C[k] == m.A[0] + m.B[0] + (uint32_t)k;
            EXPECT(ok, "multiply result belongs to its job");
        }
    }
    EXPECT(seen.size() == kind.size(), "every job finished (%zu of %zu)", seen.size(), kind.size());
    for (auto &p : seen) EXPECT(p.second == 1, "job %lld returned %d times", (long long)p.first, p.second);
    // the reservations held at any instant never exceed the budget
    std::vector<std::pair<double, int64_t>> ev;
    for (auto &r : results) {
        ev.push_back({r.start_ms, (int64_t)r.reserved_bytes});
        ev.push_back({r.finish_ms, -(int64_t)r.reserved_bytes});
    }
    std::sort(ev.begin(), ev.end(), [](auto &a, auto &b) { return a.first < b.first || (a.first == b.first && a.second < b.second); });
    int64_t cur = 0, peak = 0;
    for (auto &e : ev) peak = std::max(peak, cur += e.second);
    EXPECT((uint64_t)peak <= budget, "budget exceeded: %lld > %llu", (long long)peak, (unsigned long long)budget);
}

int main() {
    std::atomic<bool> finished{false};
    std::thread watchdog([&] {
        for (int i = 0; i < 600 && !finished; ++i) std::this_thread::sleep_for(std::chrono::milliseconds(100));
        if (!finished) {
            printf("serve sim: stalled (watchdog)\n");
            fflush(stdout);
            std::_Exit(2);
        }
    });
    const uint64_t tight = gpupi_serve_job_bytes(GPUPI_JOB_PI, 5100, 0);
    for (int w = 1; w <= 3; ++w) serialized(w);
    for (int seed = 0; seed < 40; ++seed) {
        scenario(seed, 200ull << 30, 1 + seed % 4, 60);
        scenario(1000 + seed, tight, 1 + seed % 3, 30);
    }
    // pi jobs of one tree size did share trees somewhere in the mixes
    EXPECT(g_max_pi_batch.load() > 1, "no pi batch formed");
    finished = true;
    watchdog.join();
    printf(fails ? "serve sim: %d failures\n" : "serve sim: ok\n", fails);
    return fails ? 1 : 0;
}
