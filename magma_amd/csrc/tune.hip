// tune.hip -- the autotuner and the variant table it fills.
//
// For every kernel family that has more than one implementation (tune.h) and
// every size class it serves, magma_amd_autotune times each variant on the
// queue's device with synthetic matrices of that class and keeps the fastest.
// The table can be saved to and loaded from a text file ("family class
// variant" per line), so a deployment tunes once per device.  Untuned entries
// use variant 0.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "magma_amd.h"
#include "mixed.h"
#include "tune.h"
#include "workqueue.h"

namespace magma_amd {

magma_int_t launch_gemm_tiles(const GemmSet &g, const int *prefix, long max_tiles, int tile, hipStream_t st);

namespace {

int g_table[kTuneFamilies][kTuneClasses];
bool g_init = false;

void init_table()
{
    if (g_init) return;
    for (auto &row : g_table)
        for (int &v : row) v = -1;
    g_init = true;
}

const int kVariants[kTuneFamilies] = {5, 5, 4};

// variants a size class can use: the register-resident LU (3, 4) only up to order 128
int class_variants(int f, int c) { return (f == kTuneGetrf && c > 4) ? 3 : kVariants[f]; }

__global__ void fill_kernel(double *p, size_t count, unsigned seed)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
        unsigned long long z = (i + 1) * 0x9E3779B97F4A7C15ull + seed;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        p[i] = 2.0 * ((double)(z >> 11) * 0x1p-53) - 1.0;
    }
}

__global__ void ptr_kernel(double **ptrs, double *base, size_t stride, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) ptrs[i] = base + stride * i;
}

__global__ void iptr_kernel(int **ptrs, int *base, size_t stride, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) ptrs[i] = base + stride * i;
}

__global__ void sizes_kernel(int *n, int count, int lo, int hi, unsigned seed)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        unsigned long long z = (i + 1) * 0x9E3779B97F4A7C15ull + seed;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * 0x1p-53;
        int v = (int)floor(exp(log((double)lo) + u * (log((double)hi + 1) - log((double)lo))));
        n[i] = v < lo ? lo : (v > hi ? hi : v);
    }
    if (i == count) n[i] = 0;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <class T> T *get() { return static_cast<T *>(p); }
    bool alloc(size_t b)
    {
        const hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            std::fprintf(stderr, "magma_amd_autotune: hipMalloc(%zu) failed: %s\n", b, hipGetErrorString(e));
            p = nullptr;
            return false;
        }
        return true;
    }
};

// median-free: best of `reps` timings of fn() (each preceded by prep())
template <class Prep, class Fn>
float best_ms(hipStream_t st, int reps, Prep prep, Fn fn)
{
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    float best = 1e30f;
    prep();
    fn();                                  // warm-up
    for (int r = 0; r < reps; ++r) {
        prep();
        (void)hipEventRecord(a, st);
        fn();
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        best = ms < best ? ms : best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

}  // namespace

int size_class_of(int d)
{
    int c = 0;
    while (c < kTuneClasses - 1 && kBucketLimit[c] < d) ++c;
    return c;
}

int tune_variant_count(TuneFamily f) { return (f >= 0 && f < kTuneFamilies) ? kVariants[f] : 0; }

int tuned_variant(TuneFamily f, int c)
{
    if (!g_init || f < 0 || f >= kTuneFamilies || c < 0 || c >= kTuneClasses) return 0;
    const int v = g_table[f][c];
    return (v < 0 || v >= class_variants(f, c)) ? 0 : v;
}

}  // namespace magma_amd

using namespace magma_amd;

namespace {

magma_int_t tune_getrf(hipStream_t st, FILE *log)
{
    const int classes[] = {4, 5, 6};
    const int dims[] = {96, 192, 384};
    const int counts[] = {4096, 2048, 512};
    const int nv[] = {5, 3, 3};            // the register-resident variants cover n <= 128 only
    for (int t = 0; t < 3; ++t) {
        const int n = dims[t], count = counts[t];
        const size_t elems = (size_t)n * n * count;
        DevBuf src, work, ptrs, piv, pivp, info;
        if (!src.alloc(elems * 8) || !work.alloc(elems * 8) || !ptrs.alloc(8 * (size_t)count) ||
            !piv.alloc(4 * (size_t)n * count) || !pivp.alloc(8 * (size_t)count) || !info.alloc(4 * (size_t)count))
            return MAGMA_ERR_DEVICE_ALLOC;
        fill_kernel<<<1024, 256, 0, st>>>(src.get<double>(), elems, 11u + t);
        ptr_kernel<<<ceil_div(count, 256), 256, 0, st>>>(ptrs.get<double *>(), work.get<double>(), (size_t)n * n, count);
        iptr_kernel<<<ceil_div(count, 256), 256, 0, st>>>(pivp.get<int *>(), piv.get<int>(), n, count);
        MatSet s;
        s.m = n; s.n = n; s.lda = n;
        s.A = ptrs.get<double *>(); s.ipiv = pivp.get<int *>(); s.info = info.get<int>();
        s.count = count;
        float best = 1e30f;
        int bestv = 0;
        for (int v = 0; v < nv[t]; ++v) {
            g_table[kTuneGetrf][classes[t]] = v;
            const float ms = best_ms(
                st, 3, [&]() { (void)hipMemcpyAsync(work.p, src.p, elems * 8, hipMemcpyDeviceToDevice, st); },
                [&]() { (void)launch_getrf(s, n, n, false, st); });
            if (log) std::fprintf(log, "autotune getrf n=%d count=%d variant %d: %.3f ms\n", n, count, v, ms);
            if (ms < best) {
                best = ms;
                bestv = v;
            }
        }
        g_table[kTuneGetrf][classes[t]] = bestv;
    }
    return MAGMA_SUCCESS;
}

magma_int_t tune_mixed(hipStream_t st, FILE *log)
{
    const int classes[] = {4, 5, 6};
    const int dims[] = {96, 192, 384};
    const int counts[] = {4096, 1024, 256};
    const int nv[] = {5, 5, 2};
    for (int t = 0; t < 3; ++t) {
        const int n = dims[t], count = counts[t];
        const size_t elems = (size_t)n * n * count;
        DevBuf a, b, x, pa, pb, px, w, it, inf, fb;
        if (!a.alloc(elems * 8) || !b.alloc((size_t)n * count * 8) || !x.alloc((size_t)n * count * 8) ||
            !pa.alloc(8 * (size_t)count) || !pb.alloc(8 * (size_t)count) || !px.alloc(8 * (size_t)count) ||
            !it.alloc(4 * (size_t)count) || !inf.alloc(4 * (size_t)count) || !fb.alloc(4 * (size_t)(count + 2)))
            return MAGMA_ERR_DEVICE_ALLOC;
        // factor slots and status entries for the largest plan of the variants
        size_t wbytes = 0;
        for (int v = 0; v < nv[t]; ++v) {
            g_table[kTuneMixed][classes[t]] = v;
            wbytes = std::max(wbytes, mixed_workspace_bytes(mixed_plan(count, n, 1), n, true));
        }
        if (!w.alloc(wbytes)) return MAGMA_ERR_DEVICE_ALLOC;
        fill_kernel<<<1024, 256, 0, st>>>(a.get<double>(), elems, 21u + t);
        fill_kernel<<<256, 256, 0, st>>>(b.get<double>(), (size_t)n * count, 31u + t);
        ptr_kernel<<<ceil_div(count, 256), 256, 0, st>>>(pa.get<double *>(), a.get<double>(), (size_t)n * n, count);
        ptr_kernel<<<ceil_div(count, 256), 256, 0, st>>>(pb.get<double *>(), b.get<double>(), (size_t)n, count);
        ptr_kernel<<<ceil_div(count, 256), 256, 0, st>>>(px.get<double *>(), x.get<double>(), (size_t)n, count);
        MixedSet g;
        g.n = n; g.nrhs = 1; g.lda = n; g.ldb = n; g.ldx = n;
        g.A = pa.get<double *>(); g.B = pb.get<double *>(); g.X = px.get<double *>();
        g.info = inf.get<int>(); g.iter = it.get<int>();
        g.fb_count = fb.get<int>(); g.fb_list = fb.get<int>() + 1;
        g.count = count;
        float best = 1e30f;
        int bestv = 0;
        for (int v = 0; v < nv[t]; ++v) {
            g_table[kTuneMixed][classes[t]] = v;
            const MixedPlan plan = mixed_plan(count, n, 1);
            mixed_workspace_view(g, plan, n, true, w.get<char>());
            const float ms = best_ms(st, 3, [&]() {}, [&]() { (void)launch_mixed(g, n, plan, st); });
            if (log) std::fprintf(log, "autotune mixed n=%d count=%d variant %d: %.3f ms\n", n, count, v, ms);
            if (ms < best) {
                best = ms;
                bestv = v;
            }
        }
        g_table[kTuneMixed][classes[t]] = bestv;
    }
    return MAGMA_SUCCESS;
}

// vbatched DGEMM on a log-uniform [8,256] size mix (the class used for
// calls whose sizes are not known on the host), and fixed sizes 8..256
magma_int_t tune_gemm(magma_queue_t queue, FILE *log)
{
    hipStream_t st = stream_of(queue);
    const int count = 20000;
    DevBuf nn, mem, pa, pb, pc, prefix;
    if (!nn.alloc(4 * (size_t)(count + 1)) || !mem.alloc(8ull * 3 * 256 * 256 * 2048) ||
        !pa.alloc(8 * (size_t)count) || !pb.alloc(8 * (size_t)count) || !pc.alloc(8 * (size_t)count) ||
        !prefix.alloc(4 * (size_t)(count + 1)))
        return MAGMA_ERR_DEVICE_ALLOC;
    sizes_kernel<<<ceil_div(count + 1, 256), 256, 0, st>>>(nn.get<int>(), count, 8, 256, 41u);
    // operands: matrices of up to 256 x 256 cycling over a 2048-matrix pool
    fill_kernel<<<1024, 256, 0, st>>>(mem.get<double>(), 3ull * 256 * 256 * 2048, 43u);
    const size_t st2 = 256 * 256;
    double *base = mem.get<double>();
    std::vector<double *> ha(count), hb(count), hc(count);
    for (int i = 0; i < count; ++i) {
        ha[i] = base + st2 * (i % 2048);
        hb[i] = base + st2 * (2048 + i % 2048);
        hc[i] = base + st2 * (4096 + i % 2048);
    }
    (void)hipMemcpyAsync(pa.p, ha.data(), 8 * (size_t)count, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(pb.p, hb.data(), 8 * (size_t)count, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(pc.p, hc.data(), 8 * (size_t)count, hipMemcpyHostToDevice, st);
    (void)hipStreamSynchronize(st);
    GemmSet g;
    g.m_arr = nn.get<int>(); g.n_arr = nn.get<int>(); g.k_arr = nn.get<int>();
    g.lda_arr = nn.get<int>(); g.ldb_arr = nn.get<int>(); g.ldc_arr = nn.get<int>();
    g.alpha = 1.0; g.beta = 1.0;
    g.A = pa.get<double *>(); g.B = pb.get<double *>(); g.C = pc.get<double *>();
    g.count = count;
    float best = 1e30f;
    int bestv = 0;
    for (int v = 0; v < kVariants[kTuneGemm]; ++v) {
        const int tile = gemm_tile_of(v);
        if (build_tile_prefix(g.m_arr, g.n_arr, count, tile, tile, prefix.get<int>(), st) != MAGMA_SUCCESS)
            return MAGMA_ERR_UNKNOWN;
        const long tiles = (long)ceil_div(256, tile) * ceil_div(256, tile) * count;
        const float ms = best_ms(st, 3, [&]() {},
                                 [&]() { (void)launch_gemm_tiles(g, prefix.get<int>(), tiles, tile, st); });
        if (log) std::fprintf(log, "autotune gemm vbatched [8,256] x %d tile %d: %.3f ms\n", count, tile, ms);
        if (ms < best) {
            best = ms;
            bestv = v;
        }
    }
    g_table[kTuneGemm][kTuneClasses - 1] = bestv;
    // fixed sizes
    const int dims[] = {8, 16, 32, 64, 128, 256};
    for (int t = 0; t < 6; ++t) {
        const int n = dims[t];
        const int cnt = (int)std::min<long>(20000, (64L << 20) / ((long)n * n * 8));
        GemmSet f = g;
        f.m_arr = f.n_arr = f.k_arr = f.lda_arr = f.ldb_arr = f.ldc_arr = nullptr;
        f.m = f.n = f.k = f.lda = f.ldb = f.ldc = n;
        f.count = std::min(cnt, count);
        best = 1e30f;
        bestv = 0;
        for (int v = 0; v < kVariants[kTuneGemm]; ++v) {
            const int tile = gemm_tile_of(v);
            const long tiles = (long)ceil_div(n, tile) * ceil_div(n, tile) * f.count;
            const float ms = best_ms(st, 3, [&]() {}, [&]() { (void)launch_gemm_tiles(f, nullptr, tiles, tile, st); });
            if (log) std::fprintf(log, "autotune gemm n=%d x %d tile %d: %.3f ms\n", n, f.count, tile, ms);
            if (ms < best) {
                best = ms;
                bestv = v;
            }
        }
        g_table[kTuneGemm][size_class_of(n)] = bestv;
    }
    return MAGMA_SUCCESS;
}

}  // namespace

extern "C" {

magma_int_t magma_amd_autotune(magma_queue_t queue, magma_int_t flags)
{
    init_table();
    FILE *log = (flags & MAGMA_AMD_TUNE_VERBOSE) ? stderr : nullptr;
    hipStream_t st = stream_of(queue);
    magma_int_t r = MAGMA_SUCCESS;
    const bool all = (flags & (MAGMA_AMD_TUNE_GETRF | MAGMA_AMD_TUNE_MIXED | MAGMA_AMD_TUNE_GEMM)) == 0;
    if (r == MAGMA_SUCCESS && (all || (flags & MAGMA_AMD_TUNE_GETRF))) r = tune_getrf(st, log);
    if (r == MAGMA_SUCCESS && (all || (flags & MAGMA_AMD_TUNE_MIXED))) r = tune_mixed(st, log);
    if (r == MAGMA_SUCCESS && (all || (flags & MAGMA_AMD_TUNE_GEMM))) r = tune_gemm(queue, log);
    (void)hipStreamSynchronize(st);
    return r;
}

magma_int_t magma_amd_tune_get(magma_int_t family, magma_int_t size_class)
{
    if (family < 0 || family >= kTuneFamilies || size_class < 0 || size_class >= kTuneClasses) return -1;
    init_table();
    return g_table[family][size_class];
}

magma_int_t magma_amd_tune_set(magma_int_t family, magma_int_t size_class, magma_int_t variant)
{
    if (family < 0 || family >= kTuneFamilies) return -1;
    if (size_class < 0 || size_class >= kTuneClasses) return -2;
    if (variant < -1 || variant >= class_variants(family, size_class)) return -3;
    init_table();
    g_table[family][size_class] = variant;
    return MAGMA_SUCCESS;
}

magma_int_t magma_amd_tune_save(const char *path)
{
    init_table();
    FILE *f = std::fopen(path, "w");
    if (!f) return MAGMA_ERR_UNKNOWN;
    std::fprintf(f, "# magma_amd tuning table: family size_class variant\n");
    for (int fam = 0; fam < kTuneFamilies; ++fam)
        for (int c = 0; c < kTuneClasses; ++c)
            if (g_table[fam][c] >= 0) std::fprintf(f, "%d %d %d\n", fam, c, g_table[fam][c]);
    std::fclose(f);
    return MAGMA_SUCCESS;
}

magma_int_t magma_amd_tune_load(const char *path)
{
    init_table();
    FILE *f = std::fopen(path, "r");
    if (!f) return MAGMA_ERR_UNKNOWN;
    char line[256];
    magma_int_t bad = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int fam, c, v;
        if (std::sscanf(line, "%d %d %d", &fam, &c, &v) != 3) continue;
        if (magma_amd_tune_set(fam, c, v) != MAGMA_SUCCESS) ++bad;
    }
    std::fclose(f);
    return bad ? MAGMA_ERR_ILLEGAL_VALUE : MAGMA_SUCCESS;
}

}  // extern "C"
