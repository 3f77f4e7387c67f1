// Serving loop (gpupi_serve.h): FCFS admission against an HBM budget, pi
// jobs on worker threads with their own contexts and streams, multiply jobs
// batched continuously into shared ragged launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "batch_arith.h"
#include "gpupi_pi.h"
#include "gpupi_serve.h"

namespace {

#define GPUPI_SERVE_RETURN_IF(x)                  \
    do {                                          \
        const hipError_t err_ = (x);              \
        if (err_ != hipSuccess) return err_;      \
    } while (0)

constexpr uint64_t kHashMod = 1000000007ull;

uint32_t ret_hash(const uint32_t *v, uint32_t n) {  // pi_bs_gpu.cu:852-858
    const uint64_t base = (1ull << 32) % kHashMod;
    uint64_t h = 0;
    for (int64_t i = (int64_t)n - 1; i >= 0; --i) h = (h * base + v[i]) % kHashMod;
    return (uint32_t)h;
}

uint32_t dec_hash(const char *d, uint32_t n) {  // pi_bs_gpu.cu:860-866
    uint64_t h = 3;
    for (uint32_t i = 0; i < n; ++i) h = (h * 10u + (uint32_t)(d[i] - '0')) % kHashMod;
    return (uint32_t)h;
}

struct Job {
    int64_t id = 0;
    int kind = GPUPI_JOB_PI;
    uint32_t digits = 0;
    char *digits_out = nullptr;
    const uint32_t *A = nullptr, *B = nullptr;
    uint32_t La = 0, Lb = 0;
    uint32_t *C_out = nullptr;
    uint64_t bytes = 0;
    bool worker_slot = false;  // a pi job admitted with a worker of its own
    GpupiJobResult res{};
};

}  // namespace

struct GpupiServer {
    GpupiServeConfig cfg{};
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_pi, cv_mul, cv_done;
    std::deque<std::unique_ptr<Job>> waiting;   // submitted, not admitted (arrival order)
    std::deque<std::unique_ptr<Job>> pi_ready;  // admitted pi jobs
    std::deque<std::unique_ptr<Job>> mul_ready; // admitted multiply jobs
    std::deque<GpupiJobResult> done;
    uint64_t in_use = 0;
    int free_pi = 0;
    int64_t next_id = 0;
    int64_t pending = 0;  // submitted and not yet returned as finished
    bool stop = false;
    std::vector<std::thread> threads;
    // one stream-ordered memory pool for every worker: device memory of a
    // finished job is reused by the next without hipMalloc / hipFree (which
    // wait for the whole device); memory above the HBM budget is released
    hipMemPool_t pool = nullptr;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    // Admit what the rule allows; called with mu held.
    void schedule() {
        std::vector<uint64_t> bytes;
        std::vector<int32_t> kind;
        for (auto &j : waiting) {
            bytes.push_back(j->bytes);
            kind.push_back(j->kind);
        }
        std::vector<int> pick(waiting.size());
        const int n = gpupi_serve_admit(bytes.data(), kind.data(), (int)waiting.size(), cfg.hbm_budget_bytes, in_use,
                                        free_pi, pick.data());
        if (n == 0) return;
        std::vector<bool> taken(waiting.size(), false);
        const double t = now_ms();
        for (int i = 0; i < n; ++i) {
            std::unique_ptr<Job> &j = waiting[pick[i]];
            taken[pick[i]] = true;
            in_use += j->bytes;
            j->res.start_ms = t;
            if (j->kind == GPUPI_JOB_PI) {
                --free_pi;
                j->worker_slot = true;
                pi_ready.push_back(std::move(j));
            } else {
                mul_ready.push_back(std::move(j));
            }
        }
        std::deque<std::unique_ptr<Job>> rest;
        for (size_t i = 0; i < waiting.size(); ++i)
            if (!taken[i]) rest.push_back(std::move(waiting[i]));
        waiting.swap(rest);
        cv_pi.notify_all();
        cv_mul.notify_all();
    }
    void finish(std::unique_ptr<Job> j, int status) {
        std::lock_guard<std::mutex> lk(mu);
        j->res.status = status;
        j->res.finish_ms = now_ms();
        in_use -= j->bytes;
        if (j->worker_slot) ++free_pi;
        done.push_back(j->res);
        schedule();
        cv_done.notify_all();
    }
    void pi_worker();
    void mul_worker();
};

void GpupiServer::pi_worker() {
    BatchMPContext *ctx = batch_mp_init();
    hipStream_t stream = nullptr;
    if (ctx && hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess) {
        (void)batch_mp_set_stream(ctx, stream);
        (void)gpupi_set_memory_pool(ctx, pool);
    }
    for (;;) {
        // the front job and the admitted pi jobs waiting for a worker whose
        // tree has as many groups: one shared tree
        std::vector<std::unique_ptr<Job>> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_pi.wait(lk, [&] { return stop || !pi_ready.empty(); });
            if (pi_ready.empty()) break;
            batch.push_back(std::move(pi_ready.front()));
            pi_ready.pop_front();
            // jobs not yet admitted because every worker is busy join it, on
            // the budget alone: same tree size, in arrival order up to the
            // first job that does not fit, a fair share (one worker's part of
            // those jobs) so that the other workers' pipelines still overlap
            const uint32_t groups = gpupi_pi_groups_for_digits(batch[0]->digits);
            uint64_t room = cfg.hbm_budget_bytes - in_use;
            size_t same = 0;
            for (const auto &w : waiting) {
                if (w->bytes > room) break;
                room -= w->bytes;
                if (w->kind == GPUPI_JOB_PI && gpupi_pi_groups_for_digits(w->digits) == groups) ++same;
            }
            const size_t share = (same + cfg.pi_workers - 1) / cfg.pi_workers;
            const size_t want = std::min<size_t>(cfg.pi_batch_max, 1 + share);
            const double t = now_ms();
            for (auto it = waiting.begin(); it != waiting.end() && batch.size() < want;) {
                Job &w = **it;
                if (in_use + w.bytes > cfg.hbm_budget_bytes) break;
                if (w.kind == GPUPI_JOB_PI && gpupi_pi_groups_for_digits(w.digits) == groups) {
                    in_use += w.bytes;
                    w.res.start_ms = t;
                    batch.push_back(std::move(*it));
                    it = waiting.erase(it);
                } else {
                    ++it;
                }
            }
        }
        const size_t n = batch.size();
        std::vector<uint32_t> digits(n);
        for (size_t i = 0; i < n; ++i) digits[i] = batch[i]->digits;
        std::vector<GpupiPiResult> r(n);
        int status = ctx && stream ? hipSuccess : hipErrorNotInitialized;
        if (status == hipSuccess)
            status = gpupi_pi_run_batch(ctx, digits.data(), (int)n, GPUPI_PREC_REQUESTED, r.data());
        for (size_t i = 0; i < n; ++i) {
            Job &j = *batch[i];
            int st = status;
            if (st == hipSuccess) {
                std::vector<uint32_t> ret(r[i].ret_length);
                std::vector<char> own;
                char *dig = j.digits_out;
                if (dig == nullptr) {
                    own.resize(j.digits);
                    dig = own.data();
                }
                // copies and frees in this worker's stream only: nothing
                // waits for the other workers' work
                st = hipMemcpyAsync(ret.data(), r[i].ret, ret.size() * 4, hipMemcpyDeviceToHost, stream);
                if (st == hipSuccess && j.digits)
                    st = hipMemcpyAsync(dig, r[i].digits, j.digits, hipMemcpyDeviceToHost, stream);
                (void)hipFreeAsync(r[i].ret, stream);
                (void)hipFreeAsync(r[i].digits, stream);
                const hipError_t synced = hipStreamSynchronize(stream);
                if (st == hipSuccess) st = synced;
                if (st == hipSuccess) {
                    j.res.ret_length = r[i].ret_length;
                    j.res.ret_hash = ret_hash(ret.data(), r[i].ret_length);
                    j.res.dec_hash = dec_hash(dig, j.digits);
                }
            }
            j.res.batch_size = (uint32_t)n;
            finish(std::move(batch[i]), st);
        }
    }
    if (ctx) batch_mp_destroy(ctx);
    if (stream) (void)hipStreamDestroy(stream);
}

void GpupiServer::mul_worker() {
    BatchMPContext *ctx = batch_mp_init();
    hipStream_t stream = nullptr;
    if (ctx && hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess) {
        (void)batch_mp_set_stream(ctx, stream);
        (void)gpupi_set_memory_pool(ctx, pool);
    }
    for (;;) {
        std::vector<std::unique_ptr<Job>> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_mul.wait(lk, [&] { return stop || !mul_ready.empty(); });
            if (mul_ready.empty()) break;
            while (!mul_ready.empty() && batch.size() < cfg.mul_batch_max) {
                batch.push_back(std::move(mul_ready.front()));
                mul_ready.pop_front();
            }
        }
        // one device arena per batch: A | B | C of every job
        size_t words = 0;
        for (auto &j : batch) words += 2 * ((size_t)j->La + j->Lb);
        uint32_t *arena = nullptr;
        int status = ctx && stream ? hipSuccess : hipErrorNotInitialized;
        if (status == hipSuccess)
            status = hipMallocFromPoolAsync(reinterpret_cast<void **>(&arena), std::max<size_t>(words, 1) * 4, pool,
                                            stream);
        std::vector<GpupiMulItem> items(batch.size());
        size_t off = 0;
        for (size_t i = 0; i < batch.size() && status == hipSuccess; ++i) {
            Job &j = *batch[i];
            items[i] = GpupiMulItem{arena + off, arena + off + j.La, arena + off + j.La + j.Lb, j.La, j.Lb};
            if (j.La) status = hipMemcpyAsync(arena + off, j.A, (size_t)j.La * 4, hipMemcpyHostToDevice, stream);
            if (status == hipSuccess && j.Lb)
                status = hipMemcpyAsync(arena + off + j.La, j.B, (size_t)j.Lb * 4, hipMemcpyHostToDevice, stream);
            off += 2 * ((size_t)j.La + j.Lb);
        }
        if (status == hipSuccess) status = gpupi_mul_ragged(ctx, items.data(), (uint32_t)items.size());
        for (size_t i = 0; i < batch.size() && status == hipSuccess; ++i) {
            const size_t n = (size_t)batch[i]->La + batch[i]->Lb;
            if (n) status = hipMemcpyAsync(batch[i]->C_out, items[i].C, n * 4, hipMemcpyDeviceToHost, stream);
        }
        if (arena) (void)hipFreeAsync(arena, stream);
        if (stream) {
            const hipError_t synced = hipStreamSynchronize(stream);
            if (status == hipSuccess) status = synced;
        }
        const uint32_t nb = (uint32_t)batch.size();
        for (auto &j : batch) {
            j->res.batch_size = nb;
            finish(std::move(j), status);
        }
    }
    if (ctx) batch_mp_destroy(ctx);
    if (stream) (void)hipStreamDestroy(stream);
}

extern "C" {

uint64_t gpupi_serve_job_bytes(int kind, uint32_t a, uint32_t b) {
    if (kind == GPUPI_JOB_PI) {
        // the pipeline's peak: the largest product's four transform planes
        // (32 B per coefficient, K <= 2^ceil(log2(0.83 d / 4) + 1)) plus the
        // tree and inverse buffers; 24 B per digit covers both with margin
        return (64ull << 20) + 24ull * a;
    }
    const uint64_t L = (uint64_t)a + b;
    uint64_t K = 16;
    while (K < L) K <<= 1;
    // operands + product (8 B per limb) + ragged NTT scratch (4 planes of
    // 8 B, staging 4 B, spill and flags: < 37 B per coefficient)
    return 8 * L + 37 * K + 4096;
}

int gpupi_serve_admit(const uint64_t *bytes, const int32_t *kind, int n, uint64_t budget, uint64_t in_use,
                      int free_pi_workers, int *out) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (in_use + bytes[i] > budget) break;  // FCFS: the head waits for memory, so does the rest
        if (kind[i] == GPUPI_JOB_PI) {
            if (free_pi_workers == 0) continue;  // fits but no worker: later jobs may go ahead
            --free_pi_workers;
        }
        in_use += bytes[i];
        out[m++] = i;
    }
    return m;
}

GpupiServer *gpupi_server_create(const GpupiServeConfig *cfg) {
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) return nullptr;
    GpupiServer *s = new (std::nothrow) GpupiServer();
    if (s == nullptr) return nullptr;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.hbm_budget_bytes == 0) s->cfg.hbm_budget_bytes = 200ull * 1000 * 1000 * 1000;
    if (s->cfg.pi_workers == 0) s->cfg.pi_workers = 4;
    if (s->cfg.mul_batch_max == 0) s->cfg.mul_batch_max = 512;
    if (s->cfg.pi_batch_max == 0) s->cfg.pi_batch_max = 8;
    s->t0 = std::chrono::steady_clock::now();
    int device = 0;
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    uint64_t threshold = s->cfg.hbm_budget_bytes;
    if (hipGetDevice(&device) != hipSuccess || (props.location.id = device, hipMemPoolCreate(&s->pool, &props)) != hipSuccess ||
        hipMemPoolSetAttribute(s->pool, hipMemPoolAttrReleaseThreshold, &threshold) != hipSuccess) {
        if (s->pool) (void)hipMemPoolDestroy(s->pool);
        delete s;
        return nullptr;
    }
    s->free_pi = (int)s->cfg.pi_workers;
    for (uint32_t i = 0; i < s->cfg.pi_workers; ++i) s->threads.emplace_back([s] { s->pi_worker(); });
    s->threads.emplace_back([s] { s->mul_worker(); });
    return s;
}

static int64_t submit(GpupiServer *s, std::unique_ptr<Job> j) {
    if (j->bytes > s->cfg.hbm_budget_bytes) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = s->next_id++;
    j->res.id = j->id;
    j->res.kind = j->kind;
    j->res.submit_ms = s->now_ms();
    j->res.reserved_bytes = j->bytes;
    const int64_t id = j->id;
    ++s->pending;
    s->waiting.push_back(std::move(j));
    s->schedule();
    return id;
}

int64_t gpupi_server_submit_pi(GpupiServer *s, uint32_t digits, char *digits_out) {
    if (s == nullptr) return -1;
    std::unique_ptr<Job> j(new (std::nothrow) Job());
    if (!j) return -1;
    j->kind = GPUPI_JOB_PI;
    j->digits = digits;
    j->digits_out = digits_out;
    j->bytes = gpupi_serve_job_bytes(GPUPI_JOB_PI, digits, 0);
    return submit(s, std::move(j));
}

int64_t gpupi_server_submit_mul(GpupiServer *s, const uint32_t *A, uint32_t L_a, const uint32_t *B, uint32_t L_b,
                                uint32_t *C_out) {
    if (s == nullptr || (L_a && A == nullptr) || (L_b && B == nullptr) || ((L_a || L_b) && C_out == nullptr) ||
        (uint64_t)L_a + L_b > 0xffffffffull)
        return -1;
    std::unique_ptr<Job> j(new (std::nothrow) Job());
    if (!j) return -1;
    j->kind = GPUPI_JOB_MUL;
    j->A = A;
    j->B = B;
    j->La = L_a;
    j->Lb = L_b;
    j->C_out = C_out;
    j->bytes = gpupi_serve_job_bytes(GPUPI_JOB_MUL, L_a, L_b);
    return submit(s, std::move(j));
}

int gpupi_server_results(GpupiServer *s, GpupiJobResult *out, int max, int wait) {
    if (s == nullptr || out == nullptr || max <= 0) return 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (wait) s->cv_done.wait(lk, [&] { return !s->done.empty() || s->pending == 0; });
    int n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
        --s->pending;
    }
    return n;
}

void gpupi_server_drain(GpupiServer *s) {
    if (s == nullptr) return;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        return s->waiting.empty() && s->pi_ready.empty() && s->mul_ready.empty() &&
               (int64_t)s->done.size() == s->pending;
    });
}

void gpupi_server_destroy(GpupiServer *s) {
    if (s == nullptr) return;
    gpupi_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_pi.notify_all();
    s->cv_mul.notify_all();
    for (auto &t : s->threads) t.join();
    if (s->pool) (void)hipMemPoolDestroy(s->pool);
    delete s;
}

hipError_t gpupi_server_memory(GpupiServer *s, uint64_t *used_high, uint64_t *reserved_high) {
    if (s == nullptr || used_high == nullptr || reserved_high == nullptr) return hipErrorInvalidValue;
    GPUPI_SERVE_RETURN_IF(hipMemPoolGetAttribute(s->pool, hipMemPoolAttrUsedMemHigh, used_high));
    return hipMemPoolGetAttribute(s->pool, hipMemPoolAttrReservedMemHigh, reserved_high);
}

}  // extern "C"
