// serve.hip -- continuous-batching request server over the mixed-precision
// solvers (magma_amd_server_* in magma_amd.h).
//
// Requests (an n x n system with nrhs right-hand sides, LU or SPD) are
// submitted from any host thread.  A scheduler thread admits them into device
// memory under a fixed budget, groups whatever is pending into a batch as soon
// as one of its HIP streams is free, and runs the batch there:
//
//   copy     each request's matrix into server memory (the client's buffers
//            are only read), or, for a repeat request, the resident copy;
//   mixed    FP32 factorization (skipped when a resident factorization is
//            reused) and FP64 refinement, one workgroup per request, largest
//            requests first, one launch per size class;
//   fallback the FP64 factorization and solve for requests that did not
//            converge;
//   results  info / iter / backward error copied back, an event recorded.
//
// Several batches are in flight at once on different streams, so requests that
// arrive while others run are admitted as soon as a stream and memory are free
// and their work overlaps on the device.  Each request is processed by its own
// workgroup with the same code whatever else is in the batch, so its result
// does not depend on the mix.
//
// Memory: one device arena of the budget size, carved into power-of-two blocks
// with per-size free lists.  A request holds its matrix copy and its factor
// storage while in flight; a request with a factor key keeps both resident
// after it completes (the factor is reused by later requests with the same
// key and order).  Resident factorizations are evicted least recently used
// first when an admission needs room.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "magma_amd.h"
#include "mixed.h"
#include "serve_cache.h"

namespace magma_amd {
namespace {

constexpr int kStreams = 4;

using serve::Arena;
using serve::FactorCache;
using serve::Grant;

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    magma_amd_request_t r;
    long long ticket;
    double t_submit;
};

// Batch descriptor: one block of arrays, identical in pinned host memory and
// on the device, copied host -> device with one transfer.
struct BatchArrays {
    double **A, **B, **X;
    float **F;
    int **ipiv;       // FP32 / fallback pivots, into ipivbuf
    int *ipivbuf;     // max_batch x kMixedMaxN
    const double **Asrc;
    int *n, *nrhs, *lda_src, *ldb, *ldx, *reuse, *info, *iter, *index, *fb;
    char *kind;
    double *berr;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char *host = nullptr, *dev = nullptr;   // descriptor blocks
    size_t bytes = 0;
    BatchArrays h, d;                       // views into host / device blocks
    std::vector<Request> reqs;
    std::vector<Grant> grants;              // per request: its blocks (and whether a resident served it)
    double t_start = 0.0;
    bool busy = false;
};

size_t carve(char *base, size_t cap, BatchArrays &a)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off += round_up(bytes, 64);
        return p;
    };
    a.A = reinterpret_cast<double **>(take(sizeof(void *) * cap));
    a.B = reinterpret_cast<double **>(take(sizeof(void *) * cap));
    a.X = reinterpret_cast<double **>(take(sizeof(void *) * cap));
    a.F = reinterpret_cast<float **>(take(sizeof(void *) * cap));
    a.ipiv = reinterpret_cast<int **>(take(sizeof(void *) * cap));
    a.Asrc = reinterpret_cast<const double **>(take(sizeof(void *) * cap));
    a.n = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.nrhs = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.lda_src = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.ldb = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.ldx = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.reuse = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.index = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.kind = take(cap);
    // results block (copied back): info, iter, berr, then fallback counters/lists
    a.info = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.iter = reinterpret_cast<int *>(take(sizeof(int) * cap));
    a.berr = reinterpret_cast<double *>(take(sizeof(double) * cap));
    a.fb = reinterpret_cast<int *>(take(sizeof(int) * (2 * cap + 2)));
    // device-only scratch (never copied)
    a.ipivbuf = reinterpret_cast<int *>(take(sizeof(int) * cap * kMixedMaxN));
    return off;
}

// request matrices into server memory (ld n)
__global__ void copy_matrices_kernel(const double *const *src, const int *ld_src, double *const *dst,
                                     const int *n, int count)
{
    for (int t = blockIdx.x; t < count; t += gridDim.x) {
        const int m = n[t];
        const double *s = src[t];
        double *d = dst[t];
        const int ld = ld_src[t];
        for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
            const int i = e % m, j = e / m;
            d[i + (size_t)j * m] = s[i + (size_t)j * ld];
        }
    }
}

__global__ void zero_ints_kernel(int *p, int count)
{
    for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0;
}

}  // namespace
}  // namespace magma_amd

using namespace magma_amd;

struct magma_amd_server {
    int device = 0;
    int max_batch = 0;
    char *arena_mem = nullptr;              // the budget, one device allocation
    Arena arena;
    FactorCache cache{arena};
    Slot slots[kStreams];
    std::thread worker;
    std::mutex mu;                          // guards submitted, done, counters
    std::condition_variable cv_work, cv_done;
    std::deque<Request> submitted;
    std::deque<magma_amd_completion_t> done;
    long long next_ticket = 0;
    long long outstanding = 0;              // submitted but not completed
    std::atomic<bool> stop{false};
    // scheduler-thread state (pending, arena, cache)
    std::deque<Request> pending;
    magma_amd_server_stats_t stats{};
    std::mutex stats_mu;

    static size_t matrix_bytes(int n) { return sizeof(double) * (size_t)n * n; }
    static size_t factor_bytes(int n) { return sizeof(float) * mixed_factor_floats(n); }

    void complete(Slot &s)
    {
        const double t = now_s();
        const int cnt = (int)s.reqs.size();
        std::vector<magma_amd_completion_t> out(cnt);
        for (int i = 0; i < cnt; ++i) {
            const Request &q = s.reqs[i];
            const magma_amd_request_t &r = q.r;
            const bool reused = s.grants[i].hit;
            const int iter = s.h.iter[i], info = s.h.info[i];
            cache.retire(r.factor_key, r.n, (int)r.kind, matrix_bytes(r.n), factor_bytes(r.n), s.grants[i],
                         iter >= 0 && info == 0);
            magma_amd_completion_t &c = out[i];
            c.ticket = q.ticket;
            c.info = info;
            c.iter = iter;
            c.reused = reused ? 1 : 0;
            c.berr = s.h.berr[i];
            c.t_submit = q.t_submit;
            c.t_start = s.t_start;
            c.t_done = t;
        }
        {
            std::lock_guard<std::mutex> g(stats_mu);
            stats.completed += cnt;
            stats.batches++;
            stats.bytes_in_use = arena.used();
            stats.resident = (magma_int_t)cache.resident();
            stats.evictions = cache.evictions();
        }
        s.reqs.clear();
        s.grants.clear();
        s.busy = false;
        {
            std::lock_guard<std::mutex> g(mu);
            for (auto &c : out) done.push_back(c);
            outstanding -= cnt;
        }
        cv_done.notify_all();
    }

    // forms a batch from the pending requests and launches it on slot s;
    // returns false when nothing could be admitted
    bool launch(Slot &s)
    {
        s.reqs.clear();
        s.grants.clear();
        while (!pending.empty() && (int)s.reqs.size() < max_batch) {
            Request &q = pending.front();
            const magma_amd_request_t &r = q.r;
            Grant g;
            if (!cache.admit(r.factor_key, r.n, (int)r.kind, matrix_bytes(r.n), factor_bytes(r.n), g)) break;
            const int i = (int)s.reqs.size();
            s.h.A[i] = static_cast<double *>(g.A);
            s.h.F[i] = static_cast<float *>(g.F);
            s.h.ipiv[i] = s.d.ipivbuf + (size_t)i * kMixedMaxN;
            s.h.B[i] = const_cast<double *>(r.dB);
            s.h.X[i] = r.dX;
            s.h.n[i] = r.n;
            s.h.nrhs[i] = r.nrhs;
            s.h.ldb[i] = r.lddb;
            s.h.ldx[i] = r.lddx;
            s.h.kind[i] = r.kind == MAGMA_AMD_SOLVE_SPD ? 'C' : 'L';
            s.h.reuse[i] = g.hit ? 1 : 0;
            s.h.Asrc[i] = g.hit ? static_cast<const double *>(g.Asrc) : r.dA;
            s.h.lda_src[i] = g.hit ? r.n : r.ldda;
            s.grants.push_back(g);
            s.reqs.push_back(q);
            pending.pop_front();
        }
        const int cnt = (int)s.reqs.size();
        if (cnt == 0) return false;
        // largest systems first (each size class is one persistent launch)
        std::vector<int> order(cnt);
        for (int i = 0; i < cnt; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return s.h.n[a] > s.h.n[b]; });
        int n_large = 0, max_n = 0, max_small = 0;
        for (int i = 0; i < cnt; ++i) {
            s.h.index[i] = order[i];
            const int ni = s.h.n[order[i]];
            if (ni > 256) ++n_large;
            else max_small = std::max(max_small, ni);
            max_n = std::max(max_n, ni);
        }
        hipStream_t st = s.stream;
        s.t_start = now_s();
        // descriptors: everything before the results block
        const size_t desc_bytes = reinterpret_cast<char *>(s.h.info) - s.host;
        (void)hipMemcpyAsync(s.dev, s.host, desc_bytes, hipMemcpyHostToDevice, st);
        copy_matrices_kernel<<<std::min(cnt, 1024), 256, 0, st>>>(s.d.Asrc, s.d.lda_src, s.d.A, s.d.n, cnt);
        zero_ints_kernel<<<1, 64, 0, st>>>(s.d.fb, 2);
        MixedSet g;
        g.n_arr = s.d.n; g.nrhs_arr = s.d.nrhs; g.lda_arr = s.d.n; g.ldb_arr = s.d.ldb; g.ldx_arr = s.d.ldx;
        g.kind_arr = s.d.kind;
        g.A = s.d.A; g.B = s.d.B; g.X = s.d.X; g.F = s.d.F; g.ipiv = s.d.ipiv;
        g.info = s.d.info; g.iter = s.d.iter; g.berr = s.d.berr;
        g.reuse = s.d.reuse;
        g.fb_count = s.d.fb; g.fbc_count = s.d.fb + 1;
        g.fb_list = s.d.fb + 2; g.fbc_list = s.d.fb + 2 + max_batch;
        if (n_large > 0) {
            MixedSet gl = g;
            gl.index = s.d.index;
            gl.count = n_large;
            (void)launch_mixed(gl, max_n, mixed_plan(gl.count, max_n, 0), st);
        }
        if (cnt > n_large) {
            MixedSet gs = g;
            gs.index = s.d.index + n_large;
            gs.count = cnt - n_large;
            (void)launch_mixed(gs, max_small, mixed_plan(gs.count, max_small, 0), st);
        }
        g.count = cnt;
        (void)launch_mixed_fallback(g, max_n, 16, st);
        // results: info, iter, berr (contiguous)
        const size_t res_off = reinterpret_cast<char *>(s.h.info) - s.host;
        const size_t res_bytes = reinterpret_cast<char *>(s.h.fb) - reinterpret_cast<char *>(s.h.info);
        (void)hipMemcpyAsync(s.host + res_off, s.dev + res_off, res_bytes, hipMemcpyDeviceToHost, st);
        (void)hipEventRecord(s.done, st);
        s.busy = true;
        {
            std::lock_guard<std::mutex> g2(stats_mu);
            stats.bytes_in_use = arena.used();
            stats.bytes_peak = std::max(stats.bytes_peak, arena.used());
            stats.max_batch_seen = std::max<magma_int_t>(stats.max_batch_seen, cnt);
        }
        return true;
    }

    void run()
    {
        (void)hipSetDevice(device);
        while (true) {
            bool stopping = stop.load();
            {
                std::unique_lock<std::mutex> g(mu);
                while (!submitted.empty()) {
                    pending.push_back(submitted.front());
                    submitted.pop_front();
                }
            }
            bool any_busy = false;
            for (Slot &s : slots) {
                if (s.busy && hipEventQuery(s.done) == hipSuccess) complete(s);
                any_busy |= s.busy;
            }
            bool launched = false;
            if (!pending.empty()) {
                for (Slot &s : slots) {
                    if (!s.busy && !pending.empty()) launched |= launch(s);
                }
            }
            if (stopping && pending.empty() && !any_busy) break;
            if (!launched) {
                std::unique_lock<std::mutex> g(mu);
                if (submitted.empty() && !stop.load())
                    cv_work.wait_for(g, std::chrono::microseconds(any_busy || !pending.empty() ? 20 : 2000));
            }
        }
    }
};

extern "C" {

magma_int_t magma_amd_server_create(magma_device_t device, size_t budget_bytes, magma_int_t max_batch,
                                    magma_amd_server_t *server)
{
    if (!server) return -4;
    *server = nullptr;
    if (max_batch <= 0) return -3;
    if (hipSetDevice(device) != hipSuccess) return -1;
    magma_amd_server *s = new magma_amd_server();
    s->device = device;
    s->max_batch = max_batch;
    if (hipMalloc(reinterpret_cast<void **>(&s->arena_mem), std::max<size_t>(budget_bytes, 1)) != hipSuccess) {
        delete s;
        return MAGMA_ERR_DEVICE_ALLOC;
    }
    s->arena.init(s->arena_mem, budget_bytes);
    for (Slot &sl : s->slots) {
        BatchArrays tmp;
        sl.bytes = carve(nullptr, max_batch, tmp);
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&sl.host), sl.bytes, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&sl.dev), sl.bytes) != hipSuccess) {
            magma_amd_server_destroy(s);
            return MAGMA_ERR_DEVICE_ALLOC;
        }
        carve(sl.host, max_batch, sl.h);
        carve(sl.dev, max_batch, sl.d);
    }
    s->stats.budget_bytes = budget_bytes;
    s->worker = std::thread([s]() { s->run(); });
    *server = s;
    return MAGMA_SUCCESS;
}

long long magma_amd_server_submit(magma_amd_server_t s, const magma_amd_request_t *req)
{
    if (!s || !req) return -1;
    const magma_amd_request_t &r = *req;
    if (r.kind != MAGMA_AMD_SOLVE_LU && r.kind != MAGMA_AMD_SOLVE_SPD) return -1;
    if (r.n < 1 || r.n > kMixedMaxN) return -2;
    if (r.nrhs < 1) return -3;
    if (!r.dA || r.ldda < r.n) return -4;
    if (!r.dB || r.lddb < r.n) return -5;
    if (!r.dX || r.lddx < r.n) return -6;
    const size_t need = Arena::block(magma_amd_server::matrix_bytes(r.n)) +
                        Arena::block(magma_amd_server::factor_bytes(r.n));
    if (need > s->arena.cap()) return MAGMA_ERR_DEVICE_ALLOC;
    long long t;
    {
        std::lock_guard<std::mutex> g(s->mu);
        t = s->next_ticket++;
        s->submitted.push_back(Request{r, t, now_s()});
        s->outstanding++;
    }
    {
        std::lock_guard<std::mutex> g(s->stats_mu);
        s->stats.submitted++;
    }
    s->cv_work.notify_one();
    return t;
}

magma_int_t magma_amd_server_poll(magma_amd_server_t s, magma_amd_completion_t *out, magma_int_t max_out)
{
    if (!s || (!out && max_out > 0)) return -1;
    std::lock_guard<std::mutex> g(s->mu);
    magma_int_t k = 0;
    while (k < max_out && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    return k;
}

magma_int_t magma_amd_server_wait(magma_amd_server_t s, magma_amd_completion_t *out, magma_int_t max_out,
                                  double timeout_s)
{
    if (!s || (!out && max_out > 0)) return -1;
    std::unique_lock<std::mutex> g(s->mu);
    if (s->done.empty() && s->outstanding > 0)
        s->cv_done.wait_for(g, std::chrono::duration<double>(timeout_s),
                            [s]() { return !s->done.empty() || s->outstanding == 0; });
    magma_int_t k = 0;
    while (k < max_out && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    return k;
}

magma_int_t magma_amd_server_drain(magma_amd_server_t s)
{
    if (!s) return -1;
    std::unique_lock<std::mutex> g(s->mu);
    s->cv_done.wait(g, [s]() { return s->outstanding == 0; });
    return MAGMA_SUCCESS;
}

magma_int_t magma_amd_server_stats(magma_amd_server_t s, magma_amd_server_stats_t *st)
{
    if (!s || !st) return -1;
    std::lock_guard<std::mutex> g(s->stats_mu);
    *st = s->stats;
    return MAGMA_SUCCESS;
}

magma_int_t magma_amd_server_run_trace(magma_amd_server_t s, const magma_amd_request_t *reqs,
                                       const double *arrival_s, magma_int_t count, magma_amd_completion_t *out)
{
    if (!s || count < 0 || (count > 0 && (!reqs || !arrival_s || !out))) return -1;
    std::vector<long long> tickets(count);
    std::unordered_map<long long, magma_int_t> pos;
    pos.reserve(count * 2 + 1);
    magma_int_t got = 0;
    std::vector<magma_amd_completion_t> buf(1024);
    auto collect = [&](magma_int_t k) {
        for (magma_int_t j = 0; j < k; ++j) {
            auto it = pos.find(buf[j].ticket);
            if (it != pos.end()) {
                out[it->second] = buf[j];
                ++got;
            }
        }
    };
    const double t0 = now_s();
    for (magma_int_t i = 0; i < count; ++i) {
        const double due = t0 + arrival_s[i];
        while (now_s() < due) {
            const magma_int_t k = magma_amd_server_poll(s, buf.data(), (magma_int_t)buf.size());
            if (k > 0) collect(k);
        }
        const long long t = magma_amd_server_submit(s, &reqs[i]);
        if (t < 0) return (magma_int_t)t;
        tickets[i] = t;
        pos[t] = i;
    }
    while (got < count) {
        const magma_int_t k = magma_amd_server_wait(s, buf.data(), (magma_int_t)buf.size(), 1.0);
        collect(k);
    }
    return MAGMA_SUCCESS;
}

void magma_amd_server_destroy(magma_amd_server_t s)
{
    if (!s) return;
    if (s->worker.joinable()) {
        s->stop.store(true);
        s->cv_work.notify_all();
        s->worker.join();
    }
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.dev) (void)hipFree(sl.dev);
    }
    if (s->arena_mem) (void)hipFree(s->arena_mem);
    delete s;
}

}  // extern "C"
