// mixed.hip -- batched and variable-size mixed-precision solves with the
// semantics of dsgesv / dsposv (src/zcgesv_gpu.cpp:143-365, src/zcposv_gpu.cpp):
// factor in FP32, refine in FP64, fall back to FP64 when refinement fails.
//
// One 256-thread workgroup per matrix (persistent over the work list), n <= 512.
//
// Factorization (left-looking by block columns of NB, 32-wide chunks):
//   * rows stay in the input's original order for the whole factorization;
//     a pivot is recorded (pivot order p, pivot position of each row) instead
//     of interchanging rows, so no row is ever moved in memory.  LAPACK-form
//     ipiv is produced on the fly from each row's current position, and the
//     pivot search breaks ties by current position as idamax does.
//   * a block column (all n rows x NB columns, FP32) lives in LDS.  It is read
//     once from the FP64 input, each column scaled by a power of two so that
//     its largest entry is in [0.5, 1) (Cholesky: the symmetric scaling
//     d_i = 2^-e_i with a_ii 2^-2e_i in [0.5, 2)), then updated by every
//     previous chunk of the factor (read back from the factor storage in
//     HBM), then factored chunk by chunk, then written out once.
//   * applying a chunk: the 32 pivot rows get the unit-lower triangular
//     solve (LU), then every row still to be pivoted gets C -= L U on
//     v_mfma_f32_16x16x16f16 with both operands split hi + lo in FP16 (three
//     products, FP32 accumulation), computed transposed so that the rows of
//     C are the fast MFMA index.  Rows already pivoted get a zero L operand.
//   * the chunk panel (32 columns, all rows) is factored in registers, one
//     workgroup barrier per column: wavefront arg-max by DPP, each
//     wavefront publishes its best row in an LDS slot, and every thread picks
//     the winning slot after the barrier.
//
// Refinement (all right-hand sides of a group at once, as dsgesv does):
//   x = 0, r = b; repeat: d = (LU)^-1 r in FP32 (blocked sweeps over the pivot
//   order: one wavefront per right-hand side solves each 32x32 diagonal block
//   with lane broadcasts, the workgroup applies the block to the other rows),
//   x += d in FP64, r = b - A x in FP64, stop when every column satisfies
//   ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n) (src/zcgesv_gpu.cpp:266-279),
//   at most MAGMA_MIXED_ITERMAX correction sweeps.
//
// Matrices whose FP32 factorization meets a zero pivot (-3), whose entries
// cannot be represented (-2) or that do not converge (-(ITERMAX+1)) are
// appended to a device list, and the plain FP64 kernels factor exactly those.
#include <cfloat>
#include <climits>
#include <cstdio>

#include <algorithm>
#include <type_traits>

#include "magma_amd.h"
#include "mixed.h"
#include "tune.h"
#include "wave_reduce.h"

namespace magma_amd {

magma_int_t launch_getrf_small_v(const MatSet &s, int max_mn_dim, bool solve, bool var, hipStream_t st);
magma_int_t launch_potrf_small_v(const MatSet &s, int max_n, bool solve, bool var, hipStream_t st);
magma_int_t launch_getrf_mfma(const MatSet &s, int max_dim, hipStream_t st);
magma_int_t launch_getrf_blocked(const MatSet &s, int max_m, hipStream_t st);
magma_int_t launch_potrf_blocked(const MatSet &s, int max_n, hipStream_t st);
magma_int_t zero_info(int *info, int count, hipStream_t st);

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kT = 256;
constexpr int kResidLoads = 32;     // loads of A in flight per thread in the FP64 residual
constexpr int kWaves = kT / 64;
constexpr int CH = 32;          // chunk width
constexpr int kUPad = CH + 4;   // staging row (halves): 72-byte stride, conflict-free b64 reads

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

struct alignas(16) FSlot {
    unsigned key;     // pivot_key of the wavefront's best row
    int cur;
    int row;
    int pad;
    float val[CH];    // 16-byte aligned: written and read four at a time
};

template <int NMAX, int NB, int RG>
struct MxSmem {
    union {
        struct {
            float blk[NB][NMAX + 1];          // block column, [column][original row]
            _Float16 Uh[NB][kUPad];           // -U (LU) or -L^T (Cholesky) of the chunk, split
            _Float16 Ul[NB][kUPad];
            // L11 transposed: LT[k][r] = L(r, k) for r > k, zero elsewhere
            // (r up to 2 CH - 2, so column k from row k + 1 on is one run of
            // CH - 1 entries at constant offsets); the odd pitch keeps the
            // transposing stores free of bank conflicts
            float LT[CH][2 * CH - 1];
        } f;
        struct {
            double x[RG][NMAX];
            double r[RG][NMAX];
            float v[RG][NMAX];
            float z[RG][CH];
            float T[CH][CH + 1];              // diagonal block of the sweep
            double part[kWaves][RG][2];
        } s;
    } u;
    int p[NMAX];          // pivot position -> original row
    int pivpos[NMAX];     // original row -> pivot position (INT_MAX while unpivoted)
    int cur[NMAX];        // original row -> current position (LAPACK interchange order)
    int cexp[NMAX];       // scaling exponents
    FSlot slot[2][kWaves];
    float ccol[2][CH];    // Cholesky: the published column of the diagonal block
    double red[kWaves];
    int flag;
};

// ---- wavefront reductions (DPP within rows of 16, then the 4 row leaders) ----
template <int CTRL>
__device__ __forceinline__ int dpp(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false); }

__device__ __forceinline__ double dmax_nan(double a, double b) { return (a > b || a != a) ? a : b; }

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = dpp<CTRL>((int)b), hi = dpp<CTRL>((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double readlane_d(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// wave-uniform max (NaN wins)
__device__ __forceinline__ double wave_max(double v)
{
    v = dmax_nan(v, dpp_d<0xB1>(v));
    v = dmax_nan(v, dpp_d<0x4E>(v));
    v = dmax_nan(v, dpp_d<0x141>(v));
    v = dmax_nan(v, dpp_d<0x140>(v));
    double m = readlane_d(v, 0);
#pragma unroll
    for (int r = 1; r < 4; ++r) m = dmax_nan(m, readlane_d(v, 16 * r));
    return m;
}

template <int I, int N, class Fn>
__device__ __forceinline__ void static_for(Fn &&fn)
{
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}

// Lanes 0-7 of each row of 16 get lane G's value, lanes 8-15 lane 8 + G's
// (row_newbcast: every lane of a row reads the given lane of its row).
template <int G>
__device__ __forceinline__ float bcast8(float v, int lane)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + G, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x158 + G, 0xF, 0xF, false);
    return __int_as_float((lane & 8) ? hi : lo);
}

// raw buffer resource over [p, p + bytes) (gfx9 descriptor word 3)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ void split(float v, _Float16 &h, _Float16 &l)
{
    h = (_Float16)v;
    l = (_Float16)(v - (float)h);
}

// ===========================================================================
template <int NMAX, int NB, int RG>
struct Mixed {
    using Smem = MxSmem<NMAX, NB, RG>;
    static constexpr int RPT = NMAX / kT;     // panel rows per thread
    static constexpr int RPL = NMAX / 64;     // block-load rows per lane

    // ---- block column load -------------------------------------------------
    // LU: per-column power-of-two scaling.  Cholesky: symmetric scaling with
    // the row exponents already in sm.cexp.  Sets sm.flag on a non-finite entry.
    static __device__ __forceinline__ void load_block(Smem &sm, const double *A, int lda, int n, int c0, int nb, bool chol)
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        // LC columns per wavefront in flight at once (rows lane + 64 r)
        constexpr int LC = RPL <= 4 ? 4 : 2;
#pragma unroll 1
        for (int cg = w; cg < nb; cg += LC * kWaves) {
            double vv[LC][RPL];
#pragma unroll
            for (int u = 0; u < LC; ++u) {
                const int cl = c0 + min(cg + kWaves * u, nb - 1);
                const double *col = A + (size_t)cl * lda;
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    const int o = lane + 64 * r;
                    // Cholesky reads the lower triangle only (the strict upper
                    // triangle may hold anything, as in LAPACK)
                    const bool in = o < n && (!chol || o >= cl);
                    vv[u][r] = in ? col[min(o, n - 1)] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < LC; ++u) {
            const int cc = cg + kWaves * u;
            if (cc >= nb) break;
            const int c = c0 + cc;
            double v[RPL];
            double mx = 0.0;
            bool bad = false;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                v[r] = vv[u][r];
                mx = dmax_nan(mx, fabs(v[r]));
                bad |= !__builtin_isfinite(v[r]);
            }
            int e = 0;
            if (!chol) {
                mx = wave_max(mx);
                if (mx > 0.0 && __builtin_isfinite(mx)) frexp(mx, &e);
            } else {
                e = sm.cexp[c];
            }
            const bool anybad = __builtin_amdgcn_ballot_w64(bad) != 0;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int o = lane + 64 * r;
                const int eo = chol ? e + ((o < n) ? sm.cexp[o] : 0) : e;
                sm.u.f.blk[cc][o] = (float)ldexp(v[r], -eo);
            }
            if (lane == 0) {
                if (!chol) sm.cexp[c] = e;
                if (anybad) sm.flag = 1;
            }
            }
        }
    }

    // ---- chunk application ---------------------------------------------------
    // Applies factor chunk [kq, kq+kb) to block columns [cs, ce).  The chunk's L
    // comes from the factor storage Fg (global, ld n) or from block column kl0
    // of the LDS block.  LU: triangular solve on the chunk's pivot rows, then
    // C -= L U on rows not yet pivoted.  Cholesky: C -= L L^T on rows >= rlo.
    static __device__ __forceinline__ void apply_chunk(Smem &sm, int n, int c0, int kq, int kb, int cs, int ce, const float *Fg,
                                       int kl0, bool chol, int rlo)
    {
        const int tid = threadIdx.x;
        if (!chol) {
            for (int e = tid; e < (2 * CH - 1) * CH; e += kT) {
                const int i = e / CH, i2 = e % CH;      // row i (< 2 CH - 1), column i2
                float v = 0.0f;
                if (i2 < i && i < kb)
                    v = Fg ? Fg[mixed_diag_offset(n) + 32 * (size_t)kq + CH * i + i2] : sm.u.f.blk[kl0 + i2][sm.p[kq + i]];
                sm.u.f.LT[i2][i] = v;
            }
            __syncthreads();
            // U of the chunk's pivot rows: eight threads per column, four rows
            // each, all 32 steps unrolled.  At step k the row-k value of a
            // column is final; it is broadcast to the column's other seven
            // lanes by DPP (row_newbcast within each row of 16 lanes, which
            // holds two columns), and every row r > k takes fma(L(r,k), -x_k).
            // LT is zero at r <= k and r >= kb, so those rows are unchanged.
            // sm.p is a valid row index at every position, so loads need no guard.
            const int lane = tid & 63, g = tid & 7;
#pragma unroll 1
            for (int cb = cs; cb < ce; cb += kT / 8) {
                const int cc = cb + (tid >> 3);
                const int ccl = cc < ce ? cc : cs;
                float x[4], lc[4], ln[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float v = sm.u.f.blk[ccl][sm.p[kq + 4 * g + j]];
                    x[j] = (4 * g + j < kb) ? v : 0.0f;
                    lc[j] = sm.u.f.LT[0][4 * g + j];
                }
                static_for<0, CH>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
#pragma unroll
                    for (int j = 0; j < 4; ++j) ln[j] = sm.u.f.LT[k + 1 < CH ? k + 1 : CH - 1][4 * g + j];
                    const float xk = bcast8<(k >> 2)>(x[k & 3], lane);
#pragma unroll
                    for (int j = 0; j < 4; ++j) x[j] = fmaf(lc[j], -xk, x[j]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) lc[j] = ln[j];
                });
                if (cc < ce) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = 4 * g + j;
                        if (r < kb) sm.u.f.blk[cc][sm.p[kq + r]] = x[j];
                        split((r < kb) ? -x[j] : 0.0f, sm.u.f.Uh[cc][r], sm.u.f.Ul[cc][r]);
                    }
                }
            }
        } else {
            const int w = ce - cs;
            for (int e = tid; e < w * CH; e += kT) {
                const int cc = cs + e % w, k = e / w;
                float v = 0.0f;
                if (k < kb) v = Fg ? Fg[(c0 + cc) + (size_t)(kq + k) * n] : sm.u.f.blk[kl0 + k][c0 + cc];
                split(-v, sm.u.f.Uh[cc][k], sm.u.f.Ul[cc][k]);
            }
        }
        __syncthreads();
        gemm(sm, n, kq, kb, cs, ce, Fg, kl0, rlo);
        __syncthreads();
    }

    // blk[c][o] -= sum_k L(o, kq+k) * U(k, c) for c in [cs, ce) and rows o >= rlo
    // with pivpos[o] >= kq + kb; -U^T is staged in Uh/Ul.
    static __device__ __forceinline__ void gemm(Smem &sm, int n, int kq, int kb, int cs, int ce, const float *Fg, int kl0, int rlo)
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int li = lane & 15, lk = lane >> 4;
        const int ostart = (rlo / 16) * 16;
        if (Fg) {
            gemm_stored(sm, n, kq, kb, cs, ce, Fg, rlo);
            return;
        }
#pragma unroll 1
        for (int o0 = ostart + 16 * w; o0 < n; o0 += 16 * kWaves) {
            const int o = o0 + li;
            const bool act = o < n && o >= rlo && sm.pivpos[o] >= kq + kb;
            half4 lh[2], ll[2];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int k = 16 * ks + 4 * lk + t;
                    float v = 0.0f;
                    if (act && k < kb) v = Fg ? Fg[o + (size_t)(kq + k) * n] : sm.u.f.blk[kl0 + k][o];
                    _Float16 h, l;
                    split(v, h, l);
                    lh[ks][t] = h;
                    ll[ks][t] = l;
                }
            }
#pragma unroll 1
            for (int cb = cs; cb < ce; cb += 16) {
                float4v acc;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int c = cb + 4 * lk + i;
                    acc[i] = (c < ce) ? sm.u.f.blk[c][o] : 0.0f;
                }
#pragma unroll
                for (int ks = 0; ks < (CH / 16); ++ks) {
                    const half4 uh = *reinterpret_cast<const half4 *>(&sm.u.f.Uh[cb + li][16 * ks + 4 * lk]);
                    const half4 ul = *reinterpret_cast<const half4 *>(&sm.u.f.Ul[cb + li][16 * ks + 4 * lk]);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ul, lh[ks], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(uh, ll[ks], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(uh, lh[ks], acc, 0, 0, 0);
                }
                if (act) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int c = cb + 4 * lk + i;
                        if (c < ce) sm.u.f.blk[c][o] = acc[i];
                    }
                }
            }
        }
    }

    // The same product with L from the factor storage: every L value of the
    // wavefront's row blocks is loaded first (one round of memory latency per
    // chunk instead of one per row block), by buffer loads whose column part
    // is a scalar offset.  Loads past column n - 1 (a partial last chunk) stay
    // inside the factor storage and are discarded.
    static constexpr int RB = NMAX / (16 * kWaves);     // row blocks per wavefront
    static __device__ __forceinline__ void gemm_stored(Smem &sm, int n, int kq, int kb, int cs, int ce, const float *Fg,
                                                       int rlo)
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int li = lane & 15, lk = lane >> 4;
        const int ob = (rlo / 16) * 16 + 16 * w;
        const __amdgpu_buffer_rsrc_t Fr = rsrc(Fg, 4u * (unsigned)mixed_factor_floats(n));
        float lv[RB][8];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const int o = min(ob + 16 * kWaves * rb + li, n - 1);
            const int vo = 4 * (o + 4 * lk * n);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    lv[rb][4 * ks + t] =
                        __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(Fr, vo, 4 * n * (kq + 16 * ks + t), 0));
        }
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const int o0 = ob + 16 * kWaves * rb;
            if (o0 >= n) break;
            const int o = o0 + li;
            const bool act = o < n && o >= rlo && sm.pivpos[o] >= kq + kb;
            half4 lh[2], ll[2];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int k = 16 * ks + 4 * lk + t;
                    const float v = (act && k < kb) ? lv[rb][4 * ks + t] : 0.0f;
                    _Float16 h, l;
                    split(v, h, l);
                    lh[ks][t] = h;
                    ll[ks][t] = l;
                }
            }
#pragma unroll 1
            for (int cb = cs; cb < ce; cb += 16) {
                float4v acc;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int c = cb + 4 * lk + i;
                    acc[i] = (c < ce) ? sm.u.f.blk[c][o] : 0.0f;
                }
#pragma unroll
                for (int ks = 0; ks < (CH / 16); ++ks) {
                    const half4 uh = *reinterpret_cast<const half4 *>(&sm.u.f.Uh[cb + li][16 * ks + 4 * lk]);
                    const half4 ul = *reinterpret_cast<const half4 *>(&sm.u.f.Ul[cb + li][16 * ks + 4 * lk]);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ul, lh[ks], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(uh, ll[ks], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(uh, lh[ks], acc, 0, 0, 0);
                }
                if (act) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int c = cb + 4 * lk + i;
                        if (c < ce) sm.u.f.blk[c][o] = acc[i];
                    }
                }
            }
        }
    }

    // ---- LU chunk panel: block columns [s0, s0+kb), global columns c0+s0.. ----
    // Returns nonzero on a zero or NaN pivot.
    static __device__ __forceinline__ int panel_lu(Smem &sm, int n, int c0, int s0, int kb, int *ipiv)
    {
        const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
        float a[RPT][CH];
        int cur[RPT];
        bool act[RPT];
        int bad = 0;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int o = tid + kT * r;
            act[r] = o < n && sm.pivpos[o] == INT_MAX;
            cur[r] = sm.cur[o];
#pragma unroll
            for (int cc = 0; cc < CH; ++cc) a[r][cc] = (cc < kb) ? sm.u.f.blk[s0 + cc][o] : 0.0f;
        }
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
            if (jj < kb) {
                const int gpos = c0 + s0 + jj;
                unsigned key = 0;                    // ordered integer key, 0: no candidate
                int kc = INT_MAX, own = -1;
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    if (act[r]) {
                        const float v = a[r][jj];
                        const unsigned k2 = pivot_key((v != v) ? __builtin_huge_valf() : fabsf(v));
                        const bool take = (k2 > key) | ((k2 == key) & (cur[r] < kc));
                        key = take ? k2 : key;
                        kc = take ? cur[r] : kc;
                        own = take ? r : own;
                    }
                }
                unsigned wk = key;
                int wc = kc;
                wave_argmax_key(wk, wc);
                FSlot &sl = sm.slot[jj & 1][w];
                if (own >= 0 && kc == wc) {
                    // 16-byte stores from the quad holding column jj on (the
                    // entries left of jj are written too, and unread)
                    float4 *dst = reinterpret_cast<float4 *>(sl.val);
#pragma unroll
                    for (int cc = jj & ~3; cc < CH; cc += 4) {
                        float v[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            v[e] = a[0][cc + e];
#pragma unroll
                            for (int r = 1; r < RPT; ++r) v[e] = (own == r) ? a[r][cc + e] : v[e];
                        }
                        dst[cc / 4] = make_float4(v[0], v[1], v[2], v[3]);
                    }
                    sl.row = tid + kT * own;
                }
                if (lane == 0) {
                    sl.key = wk;
                    sl.cur = wc;
                }
                __syncthreads();
                // each slot's pivot value is read with the keys, so the
                // reciprocal starts while the winner's row is in flight
                unsigned ck[kWaves];
                int cc4[kWaves];
                float cv[kWaves];
#pragma unroll
                for (int q = 0; q < kWaves; ++q) {
                    ck[q] = sm.slot[jj & 1][q].key;
                    cc4[q] = sm.slot[jj & 1][q].cur;
                    cv[q] = sm.slot[jj & 1][q].val[jj];
                }
                asm volatile("" ::: "memory");
                unsigned bk = ck[0];
                int bc = cc4[0], bw = 0;
                float piv = cv[0];
#pragma unroll
                for (int q = 1; q < kWaves; ++q) {
                    const bool take = (ck[q] > bk) | ((ck[q] == bk) & (cc4[q] < bc));
                    bk = take ? ck[q] : bk;
                    bc = take ? cc4[q] : bc;
                    bw = take ? q : bw;
                    piv = take ? cv[q] : piv;
                }
                const float rc = 1.0f / piv;
                const FSlot &ps = sm.slot[jj & 1][bw];
                const int prow = ps.row;
                float u[CH];
#pragma unroll
                for (int cc = jj & ~3; cc < CH; cc += 4) {
                    const float4 t = reinterpret_cast<const float4 *>(ps.val)[cc / 4];
                    u[cc] = t.x;
                    u[cc + 1] = t.y;
                    u[cc + 2] = t.z;
                    u[cc + 3] = t.w;
                }
                asm volatile("" ::: "memory");
                if (!(piv != 0.0f) || piv != piv) bad = 1;
                if (tid == 0) {
                    sm.p[gpos] = prow;
                    if (ipiv) ipiv[gpos] = bc + 1;
                }
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int o = tid + kT * r;
                    const bool isp = (o == prow);
                    if (isp) {
                        act[r] = false;
                        sm.pivpos[o] = gpos;
                    }
                    cur[r] = isp ? gpos : ((cur[r] == gpos) ? bc : cur[r]);
                    if (act[r] && !bad) {
                        const float l = a[r][jj] * rc;
                        a[r][jj] = l;
#pragma unroll
                        for (int cc = jj + 1; cc < CH; ++cc) a[r][cc] = fmaf(l, -u[cc], a[r][cc]);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int o = tid + kT * r;
#pragma unroll
            for (int cc = 0; cc < CH; ++cc)
                if (cc < kb) sm.u.f.blk[s0 + cc][o] = a[r][cc];
            sm.cur[o] = cur[r];
        }
        __syncthreads();
        return bad;
    }

    // ---- Cholesky chunk panel: columns [s0, s0+kb) of the block, rows >= c0+s0 ----
    static __device__ __forceinline__ int panel_chol(Smem &sm, int n, int c0, int s0, int kb)
    {
        const int tid = threadIdx.x;
        const int d0 = c0 + s0;          // first row / column of the diagonal block
        float a[RPT][CH];
        int bad = 0;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int o = tid + kT * r;
#pragma unroll
            for (int cc = 0; cc < CH; ++cc) a[r][cc] = (cc < kb && o >= d0 && o < n) ? sm.u.f.blk[s0 + cc][o] : 0.0f;
        }
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
            if (jj < kb) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int i = tid + kT * r - d0;
                    if (i >= jj && i < kb) sm.ccol[jj & 1][i] = a[r][jj];
                }
                __syncthreads();
                float cv[CH];
#pragma unroll
                for (int cc = jj; cc < CH; ++cc) cv[cc] = sm.ccol[jj & 1][cc];
                asm volatile("" ::: "memory");
                const float d = cv[jj];
                if (!(d > 0.0f)) bad = 1;
                const float ljj = sqrtf(d);
                const float rl = 1.0f / ljj;
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int o = tid + kT * r;
                    if (o == d0 + jj) a[r][jj] = ljj;
                    if (o > d0 + jj && o < n && !bad) {
                        const float l = a[r][jj] * rl;
                        a[r][jj] = l;
#pragma unroll
                        for (int cc = jj + 1; cc < CH; ++cc)
                            if (cc < kb) a[r][cc] = fmaf(-l, cv[cc] * rl, a[r][cc]);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int o = tid + kT * r;
            if (o >= d0 && o < n) {
#pragma unroll
                for (int cc = 0; cc < CH; ++cc)
                    if (cc < kb) sm.u.f.blk[s0 + cc][o] = a[r][cc];
            }
        }
        __syncthreads();
        return bad;
    }

    // ---- block column to the factor storage ----------------------------------
    static __device__ __forceinline__ void write_block(Smem &sm, float *F, int n, int c0, int nb, bool chol)
    {
        const int tid = threadIdx.x;
        if (!chol) {
            for (int e = tid; e < nb * n; e += kT) {
                const int cc = e / n, o = e % n;
                F[o + (size_t)(c0 + cc) * n] = sm.u.f.blk[cc][o];
            }
            // the diagonal blocks of the block column's chunks, rows in pivot order
            float *Fd = F + mixed_diag_offset(n) + 32 * (size_t)c0;
            const int nbr = (nb + CH - 1) / CH * CH;
            for (int e = tid; e < nbr * CH; e += kT) {
                const int s0 = (e / (CH * CH)) * CH, i = (e / CH) % CH, k = e % CH;
                float v = 0.0f;
                if (s0 + i < nb && s0 + k < nb) v = sm.u.f.blk[s0 + k][sm.p[c0 + s0 + i]];
                Fd[e] = v;
            }
        } else {
            // lower part in place; the transpose into the upper part (used by the
            // backward sweep as an upper-triangular factor)
            for (int e = tid; e < nb * n; e += kT) {
                const int cc = e / n, o = e % n;
                if (o >= c0 + cc) F[o + (size_t)(c0 + cc) * n] = sm.u.f.blk[cc][o];
            }
            for (int e = tid; e < nb * n; e += kT) {
                const int cc = e % nb, o = e / nb;
                if (o > c0 + cc) F[(c0 + cc) + (size_t)o * n] = sm.u.f.blk[cc][o];
            }
        }
    }

    // ---- the whole FP32 factorization; returns the status code --------------
    static __device__ __forceinline__ int factor(Smem &sm, const double *A, int lda, int n, float *F, int *ipiv, bool chol)
    {
        const int tid = threadIdx.x;
        for (int o = tid; o < NMAX; o += kT) {
            sm.pivpos[o] = chol ? o : ((o < n) ? INT_MAX : -1);
            sm.p[o] = o;
            sm.cur[o] = o;
        }
        if (tid == 0) sm.flag = 0;
        __syncthreads();
        if (chol) {
            // symmetric scaling from the diagonal: a_ii 2^-2e_i in [0.5, 2)
            for (int o = tid; o < n; o += kT) {
                const double d = A[o + (size_t)o * lda];
                int e = 0;
                if (d > 0.0 && __builtin_isfinite(d)) {
                    frexp(d, &e);
                    e >>= 1;
                } else {
                    // bit 0: not finite, bit 1: not positive; a matrix with
                    // both kinds reports the first (iter -2) whichever
                    // thread stores last
                    atomicOr(&sm.flag, (d != d || !__builtin_isfinite(d)) ? 1 : 2);
                }
                sm.cexp[o] = e;
            }
        }
        __syncthreads();
        const int f0 = sm.flag;
        __syncthreads();
        if (f0) return (f0 & 1) ? -2 : -3;
        int status = 0;
#pragma unroll 1
        for (int c0 = 0; c0 < n; c0 += NB) {
            const int nb = min(NB, n - c0);
            load_block(sm, A, lda, n, c0, nb, chol);
            __syncthreads();
            if (sm.flag) {
                status = -2;
                break;
            }
            const int rlo = chol ? c0 : 0;
#pragma unroll 1
            for (int kq = 0; kq < c0; kq += CH) apply_chunk(sm, n, c0, kq, CH, 0, nb, F, -1, chol, rlo);
#pragma unroll 1
            for (int s0 = 0; s0 < nb; s0 += CH) {
                const int kb = min(CH, nb - s0);
                const int bad = chol ? panel_chol(sm, n, c0, s0, kb) : panel_lu(sm, n, c0, s0, kb, ipiv);
                if (bad && status == 0) status = -3;
                if (s0 + kb < nb) apply_chunk(sm, n, c0, c0 + s0, kb, s0 + kb, nb, nullptr, s0, chol,
                                                     chol ? c0 + s0 + kb : 0);
            }
            write_block(sm, F, n, c0, nb, chol);
            __syncthreads();
        }
        int *Fexp = reinterpret_cast<int *>(F + (size_t)n * n);
        for (int o = tid; o < n; o += kT) {
            Fexp[o] = sm.cexp[o];
            Fexp[n + o] = sm.p[o];
        }
        return status;
    }

    // pivot order and exponents of a stored factorization
    static __device__ __forceinline__ void load_factor_meta(Smem &sm, const float *F, int n, bool chol)
    {
        const int tid = threadIdx.x;
        const int *Fexp = reinterpret_cast<const int *>(F + (size_t)n * n);
        for (int o = tid; o < n; o += kT) {
            sm.cexp[o] = Fexp[o];
            const int po = chol ? o : Fexp[n + o];
            sm.p[o] = po;
            sm.pivpos[po] = o;
        }
        __syncthreads();
    }

    // ---- correction solve: v (FP32, original-row order) := (LU)^-1 v ---------
    // forward sweep with L (unit for LU), backward sweep with U (for Cholesky
    // the transposed factor stored above the diagonal)
    static __device__ __forceinline__ void sweeps(Smem &sm, const float *F, int n, int nr, bool chol)
    {
        const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
        // forward
#pragma unroll 1
        for (int kq = 0; kq < n; kq += CH) {
            const int kb = min(CH, n - kq);
            for (int e = tid; e < CH * CH; e += kT) {
                // consecutive threads: down a column (Cholesky) / along a block row (LU)
                const int i = chol ? e % CH : e / CH, k = chol ? e / CH : e % CH;
                float v = 0.0f;
                if (i < kb && k < kb && k <= i)
                    v = chol ? F[(kq + i) + (size_t)(kq + k) * n] : F[mixed_diag_offset(n) + 32 * (size_t)kq + CH * i + k];
                sm.u.s.T[i][k] = v;
            }
            __syncthreads();
            // unit-diagonal form: lane i holds row i of the block divided by its
            // diagonal (Cholesky; LU's L is unit) with the diagonal itself
            // zeroed, and v_i divided the same way, so each step is one
            // broadcast and one fma with no lane conditions.  Rows and
            // columns past kb are zero in T, and v is zero there.
            const int li = lane & (CH - 1);
            const float tdiag = sm.u.s.T[li][li];
            const float rd = (chol && lane < kb) ? 1.0f / tdiag : 1.0f;
            float trow[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) trow[k] = (k == li) ? 0.0f : sm.u.s.T[li][k] * rd;
            asm volatile("" ::: "memory");
            for (int g = w; g < nr; g += kWaves) {
                const float v0 = sm.u.s.v[g][sm.p[min(kq + lane, NMAX - 1)]];
                float vi = (lane < kb) ? v0 * rd : 0.0f;
#pragma unroll
                for (int k = 0; k < CH; ++k) {
                    const float zk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vi), k));
                    vi = fmaf(-trow[k], zk, vi);
                }
                if (lane < kb) {
                    sm.u.s.z[g][lane] = vi;
                    sm.u.s.v[g][sm.p[kq + lane]] = vi;
                }
            }
            __syncthreads();
            if (kq + kb < n) {
                for (int o = tid; o < n; o += kT) {
                    if (sm.pivpos[o] >= kq + kb) {
                        float f[CH];
#pragma unroll
                        for (int k = 0; k < CH; ++k) f[k] = (k < kb) ? F[o + (size_t)(kq + k) * n] : 0.0f;
                        for (int g = 0; g < nr; ++g) {
                            float acc = sm.u.s.v[g][o];
#pragma unroll
                            for (int k = 0; k < CH; ++k) acc = fmaf(-f[k], sm.u.s.z[g][k], acc);
                            sm.u.s.v[g][o] = acc;
                        }
                    }
                }
            }
            __syncthreads();
        }
        // backward, over the same 32-column chunks as the factorization
#pragma unroll 1
        for (int kq = ((n - 1) / CH) * CH; kq >= 0; kq -= CH) {
            const int kb = min(CH, n - kq);
            for (int e = tid; e < CH * CH; e += kT) {
                // consecutive threads: down a column (Cholesky) / along a block row (LU)
                const int i = chol ? e % CH : e / CH, k = chol ? e / CH : e % CH;
                float v = 0.0f;
                if (i < kb && k < kb && k >= i)
                    v = chol ? F[(kq + i) + (size_t)(kq + k) * n] : F[mixed_diag_offset(n) + 32 * (size_t)kq + CH * i + k];
                sm.u.s.T[i][k] = v;
            }
            __syncthreads();
            // unit-diagonal form as in the forward sweep (U and L^T have
            // non-unit diagonals in both kinds)
            const int li = lane & (CH - 1);
            const float tdiag = sm.u.s.T[li][li];
            const float rd = (lane < kb) ? 1.0f / tdiag : 1.0f;
            float trow[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) trow[k] = (k == li) ? 0.0f : sm.u.s.T[li][k] * rd;
            asm volatile("" ::: "memory");
            for (int g = w; g < nr; g += kWaves) {
                const float v0 = sm.u.s.v[g][sm.p[min(kq + lane, NMAX - 1)]];
                float vi = (lane < kb) ? v0 * rd : 0.0f;
#pragma unroll
                for (int k = CH - 1; k >= 0; --k) {
                    const float yk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vi), k));
                    vi = fmaf(-trow[k], yk, vi);
                }
                if (lane < kb) {
                    sm.u.s.z[g][lane] = vi;
                    sm.u.s.v[g][sm.p[kq + lane]] = vi;
                }
            }
            __syncthreads();
            if (kq > 0) {
                for (int o = tid; o < n; o += kT) {
                    if (sm.pivpos[o] < kq) {
                        float f[CH];
#pragma unroll
                        for (int k = 0; k < CH; ++k) f[k] = (k < kb) ? F[o + (size_t)(kq + k) * n] : 0.0f;
                        for (int g = 0; g < nr; ++g) {
                            float acc = sm.u.s.v[g][o];
#pragma unroll
                            for (int k = 0; k < CH; ++k) acc = fmaf(-f[k], sm.u.s.z[g][k], acc);
                            sm.u.s.v[g][o] = acc;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    // ---- refinement of one group of right-hand sides; returns iters or status --
    static __device__ __forceinline__ int refine(Smem &sm, const double *A, int lda, const double *B, int ldb, double *X, int ldx,
                                 const float *F, int n, int nr, bool chol, double &anrm, double &berr)
    {
        const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
        const double eps = 0x1p-53;
        for (int e = tid; e < RG * NMAX; e += kT) {
            const int g = e / NMAX, o = e % NMAX;
            sm.u.s.x[g][o] = 0.0;
            sm.u.s.r[g][o] = (g < nr && o < n) ? B[o + (size_t)g * ldb] : 0.0;
        }
        __syncthreads();
        int result = -(MAGMA_MIXED_ITERMAX + 1);
#pragma unroll 1
        for (int it = 0; it <= MAGMA_MIXED_ITERMAX; ++it) {
            // v = r (Cholesky: D r), FP32
            for (int e = tid; e < nr * NMAX; e += kT) {
                const int g = e / NMAX, o = e % NMAX;
                float v = 0.0f;
                if (o < n) v = (float)(chol ? ldexp(sm.u.s.r[g][o], -sm.cexp[o]) : sm.u.s.r[g][o]);
                sm.u.s.v[g][o] = v;
            }
            __syncthreads();
            sweeps(sm, F, n, nr, chol);
            // x += D y, with y_c held at row p(c)
            for (int e = tid; e < nr * NMAX; e += kT) {
                const int g = e / NMAX, c = e % NMAX;
                if (c < n) sm.u.s.x[g][c] += ldexp((double)sm.u.s.v[g][sm.p[c]], -sm.cexp[c]);
            }
            __syncthreads();
            // r = b - A x (FP64); ||A||_inf on the first pass
            double rmax[RG], xmax[RG];
#pragma unroll
            for (int g = 0; g < RG; ++g) {
                rmax[g] = 0.0;
                xmax[g] = 0.0;
            }
            double rs_max = 0.0;
            for (int o = tid; o < n; o += kT) {
                double acc[RG];
#pragma unroll
                for (int g = 0; g < RG; ++g) acc[g] = (g < nr) ? B[o + (size_t)g * ldb] : 0.0;
                double rs = 0.0;
                const double *arow = A + o;
                // Cholesky reads the lower triangle only, as dsymv does in
                // dsposv's residual: A(o, c) for c > o is A(c, o)
                const double *acol = A + (size_t)o * lda;
                // kResidLoads independent loads in flight per step
#pragma unroll 1
                for (int c0 = 0; c0 < n; c0 += kResidLoads) {
                    double a8[kResidLoads];
#pragma unroll
                    for (int t = 0; t < kResidLoads; ++t) {
                        const int c = c0 + t;
                        a8[t] = (c < n) ? ((chol && c > o) ? acol[c] : arow[(size_t)c * lda]) : 0.0;
                    }
#pragma unroll
                    for (int t = 0; t < kResidLoads; ++t) {
                        const int c = min(c0 + t, NMAX - 1);
                        rs += fabs(a8[t]);
#pragma unroll
                        for (int g = 0; g < RG; ++g)
                            if (g < nr) acc[g] = fma(-a8[t], sm.u.s.x[g][c], acc[g]);
                    }
                }
                rs_max = dmax_nan(rs_max, rs);
#pragma unroll
                for (int g = 0; g < RG; ++g) {
                    if (g < nr) {
                        sm.u.s.r[g][o] = acc[g];
                        rmax[g] = dmax_nan(rmax[g], fabs(acc[g]));
                        xmax[g] = dmax_nan(xmax[g], fabs(sm.u.s.x[g][o]));
                    }
                }
            }
            // workgroup maxima
            rs_max = wave_max(rs_max);
#pragma unroll
            for (int g = 0; g < RG; ++g) {
                if (g < nr) {
                    rmax[g] = wave_max(rmax[g]);
                    xmax[g] = wave_max(xmax[g]);
                }
            }
            if (lane == 0) {
                sm.red[w] = rs_max;
#pragma unroll
                for (int g = 0; g < RG; ++g) {
                    sm.u.s.part[w][g][0] = rmax[g];
                    sm.u.s.part[w][g][1] = xmax[g];
                }
            }
            __syncthreads();
            if (it == 0) {
                double m = sm.red[0];
#pragma unroll
                for (int q = 1; q < kWaves; ++q) m = dmax_nan(m, sm.red[q]);
                anrm = dmax_nan(anrm, m);
            }
            const double cte = anrm * eps * sqrt((double)n);
            bool conv = true;
            double be = 0.0;
            for (int g = 0; g < nr; ++g) {
                double rn = sm.u.s.part[0][g][0], xn = sm.u.s.part[0][g][1];
#pragma unroll
                for (int q = 1; q < kWaves; ++q) {
                    rn = dmax_nan(rn, sm.u.s.part[q][g][0]);
                    xn = dmax_nan(xn, sm.u.s.part[q][g][1]);
                }
                if (!(rn <= xn * cte)) conv = false;
                be = dmax_nan(be, rn / (anrm * xn));
            }
            __syncthreads();
            berr = be;
            if (conv) {
                result = it;
                break;
            }
        }
        if (result >= 0) {
            for (int e = tid; e < nr * n; e += kT) {
                const int g = e / n, o = e % n;
                X[o + (size_t)g * ldx] = sm.u.s.x[g][o];
            }
        }
        __syncthreads();
        return result;
    }
};

// OCC: wavefronts per SIMD the register allocation is held to (workgroups
// per CU when the LDS allows it)
template <int NMAX, int NB, int RG, int OCC>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(OCC))) mixed_kernel(MixedSet g)
{
    using M = Mixed<NMAX, NB, RG>;
    __shared__ typename M::Smem sm;
    const int tid = threadIdx.x;
    for (int item = blockIdx.x; item < g.count; item += gridDim.x) {
        const int id = g.index ? g.index[item] : item;
        const int n = pick(g.n_arr, g.n, id);
        const int nrhs = pick(g.nrhs_arr, g.nrhs, id);
        const int lda = pick(g.lda_arr, g.lda, id);
        const int ldb = pick(g.ldb_arr, g.ldb, id);
        const int ldx = pick(g.ldx_arr, g.ldx, id);
        const bool chol = (g.kind_arr ? g.kind_arr[id] : g.kind) == 'C';
        if (n <= 0 || n > NMAX) {
            if (tid == 0) {
                if (g.iter) g.iter[id] = 0;
                if (g.info) g.info[id] = 0;
            }
            continue;
        }
        // Every matrix, vector and factor pointer is read through a global
        // (address space 1) pointer type and converted back, so the compiler
        // emits global rather than flat loads and stores for them: a flat
        // access also counts against the LDS wait counter, and every LDS wait
        // of the sweeps, panels and residual would drain the global loads in
        // flight.
        float *F = g.F ? (float *)((GFloatPtr *)g.F)[id] : (float *)(GFloatPtr)(g.Wslot + g.slot_floats * blockIdx.x);
        const double *A = (const double *)((GDoublePtr *)g.A)[id];
        double *Bm = (double *)((GDoublePtr *)g.B)[id];
        double *Xm = (double *)((GDoublePtr *)g.X)[id];
        int status = 0;
        if (g.reuse && g.reuse[id]) {
            M::load_factor_meta(sm, F, n, chol);
        } else {
            status = M::factor(sm, A, lda, n, F, (!chol && g.ipiv) ? (int *)((GIntPtr *)g.ipiv)[id] : nullptr, chol);
            __syncthreads();
        }
        int iters = 0;
        double anrm = 0.0, berr = 0.0;
#pragma unroll 1
        for (int g0 = 0; g0 < nrhs && status == 0; g0 += RG) {
            const int nr = min(RG, nrhs - g0);
            double be = 0.0;
            const int r = M::refine(sm, A, lda, Bm + (size_t)g0 * ldb, ldb, Xm + (size_t)g0 * ldx, ldx,
                                    F, n, nr, chol, anrm, be);
            if (r < 0) status = r;
            else iters = max(iters, r);
            berr = dmax_nan(berr, be);
        }
        if (tid == 0) {
            if (status != 0) {
                if (chol) g.fbc_list[atomicAdd(g.fbc_count, 1)] = id;
                else g.fb_list[atomicAdd(g.fb_count, 1)] = id;
            }
            if (g.iter) g.iter[id] = status == 0 ? iters : status;
            if (g.info) g.info[id] = 0;
            if (g.berr) g.berr[id] = berr;
        }
        __syncthreads();
    }
}

// X <- B for the listed matrices (the FP64 solve then works in place on X)
__global__ void copy_rhs_kernel(MixedSet g, const int *list, const int *count)
{
    const int cnt = *count;
    for (int t = blockIdx.x; t < cnt; t += gridDim.x) {
        const int id = list[t];
        const int n = pick(g.n_arr, g.n, id), nrhs = pick(g.nrhs_arr, g.nrhs, id);
        const int ldb = pick(g.ldb_arr, g.ldb, id), ldx = pick(g.ldx_arr, g.ldx, id);
        for (int e = threadIdx.x; e < n * nrhs; e += blockDim.x) {
            const int i = e % n, c = e / n;
            g.X[id][i + (size_t)c * ldx] = g.B[id][i + (size_t)c * ldb];
        }
    }
}

__global__ void zero_counts_kernel(int *a, int *b)
{
    if (threadIdx.x == 0) {
        if (a) *a = 0;
        if (b) *b = 0;
    }
}

__global__ void fill_fallback_all_kernel(int *list, int *count, int *iter, int batch)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < batch) {
        list[i] = i;
        if (iter) iter[i] = -1;
    }
    if (i == 0) *count = batch;
}

}  // namespace

// Workgroups resident per CU for a kernel at 256 threads (LDS and registers).
template <class K>
static int resident_per_cu(K kernel)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kT, 0) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

// Kernel variants (block-column width NB, right-hand-side group RG):
//   n <= 256: 0 (NB 128, RG 16), 1 (NB 64, RG 8), 2 (NB 32, RG 8, two
//             workgroups per CU), 3 (as 2 but one workgroup per CU: the
//             workgroups' matrices and factors then stay within the
//             memory-side cache between their passes), 4 (as 2 with registers
//             held to three workgroups per CU)
//   n <= 512: 0 (NB 64, RG 8), 1 (NB 32, RG 8)
// The variant is chosen per size class of max_n (classes 4, 5, 6: orders up
// to 128, 256, 512); the two smaller classes share the NMAX = 256 kernels.
static int mixed_variant(int max_n)
{
    return tuned_variant(kTuneMixed, size_class_of(max_n <= 128 ? 128 : (max_n <= 256 ? 256 : 512)));
}

// Workgroups resident per CU of variant v for orders up to max_n.
static int mixed_per_cu(int v, int max_n)
{
    if (max_n <= 256) {
        return v == 0 ? resident_per_cu(mixed_kernel<256, 128, 16, 1>)
             : v == 1 ? resident_per_cu(mixed_kernel<256, 64, 8, 1>)
             : v == 2 ? resident_per_cu(mixed_kernel<256, 32, 8, 2>)
             : v == 3 ? 1 : resident_per_cu(mixed_kernel<256, 32, 8, 3>);
    }
    return v == 0 ? resident_per_cu(mixed_kernel<512, 64, 8, 1>) : resident_per_cu(mixed_kernel<512, 32, 8, 1>);
}

MixedPlan mixed_plan(int count, int max_n, int max_nrhs)
{
    (void)max_nrhs;
    MixedPlan p;
    p.variant = mixed_variant(max_n);
    p.grid = persistent_grid(count, mixed_per_cu(p.variant, max_n));
    return p;
}

size_t mixed_workspace_bytes(const MixedPlan &p, int max_n, bool slots)
{
    return slots ? round_up(sizeof(float) * mixed_factor_floats(max_n) * (size_t)p.grid, 256) : 0;
}

void mixed_workspace_view(MixedSet &g, const MixedPlan &p, int max_n, bool slots, void *ws)
{
    (void)p;
    if (slots) {
        g.Wslot = static_cast<float *>(ws);
        g.slot_floats = mixed_factor_floats(max_n);
    }
}

magma_int_t launch_mixed(const MixedSet &g, int max_n, const MixedPlan &plan, hipStream_t st)
{
    if (g.count <= 0) return MAGMA_SUCCESS;
    if (max_n > kMixedMaxN) return MAGMA_ERR_NOT_SUPPORTED;
    const int grid = plan.grid, v = plan.variant;
    if (grid <= 0) return MAGMA_ERR_ILLEGAL_VALUE;
    if (max_n <= 256) {
        if (v == 0) mixed_kernel<256, 128, 16, 1><<<grid, kT, 0, st>>>(g);
        else if (v == 1) mixed_kernel<256, 64, 8, 1><<<grid, kT, 0, st>>>(g);
        else if (v == 4) mixed_kernel<256, 32, 8, 3><<<grid, kT, 0, st>>>(g);
        else mixed_kernel<256, 32, 8, 2><<<grid, kT, 0, st>>>(g);
    } else {
        if (v == 0) mixed_kernel<512, 64, 8, 1><<<grid, kT, 0, st>>>(g);
        else mixed_kernel<512, 32, 8, 1><<<grid, kT, 0, st>>>(g);
    }
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t launch_copy_rhs(const MixedSet &g, const int *list, const int *count, hipStream_t st)
{
    copy_rhs_kernel<<<persistent_grid(g.count, 4), 256, 0, st>>>(g, list, count);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t launch_zero_counts(int *a, int *b, hipStream_t st)
{
    zero_counts_kernel<<<1, 64, 0, st>>>(a, b);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t launch_mixed_fallback(const MixedSet &g, int max_n, int max_nrhs, hipStream_t st)
{
    (void)max_nrhs;
    const bool var = g.n_arr != nullptr;
    for (int pass = 0; pass < 2; ++pass) {
        const bool chol = pass == 1;
        int *list = chol ? g.fbc_list : g.fb_list;
        int *count = chol ? g.fbc_count : g.fb_count;
        if (!list || !count) continue;
        copy_rhs_kernel<<<persistent_grid(g.count, 4), 256, 0, st>>>(g, list, count);
        MAGMA_LAUNCH_CHECK();
        MatSet s;
        s.m = g.n; s.n = g.n; s.lda = g.lda;
        s.m_arr = g.n_arr; s.n_arr = g.n_arr; s.lda_arr = g.lda_arr;
        s.A = g.A; s.ipiv = g.ipiv; s.info = g.info;
        s.B = g.X; s.nrhs = g.nrhs; s.ldb = g.ldx;
        s.nrhs_arr = g.nrhs_arr; s.ldb_arr = g.ldx_arr;
        s.index = list; s.dev_count = count; s.count = g.count;
        magma_int_t r;
        if (!chol) {
            if (max_n <= 64) {
                r = launch_getrf_small_v(s, max_n, true, var, st);
            } else {
                r = (max_n <= 512) ? launch_getrf_mfma(s, max_n, st) : launch_getrf_blocked(s, max_n, st);
                if (r == MAGMA_SUCCESS) r = launch_getrs(s, 'N', max_n, 0, st);
            }
        } else {
            if (max_n <= 64) {
                r = launch_potrf_small_v(s, max_n, true, var, st);
            } else {
                r = launch_potrf_blocked(s, max_n, st);
                if (r == MAGMA_SUCCESS) r = launch_potrs(s, 'L', max_n, 0, st);
            }
        }
        if (r != MAGMA_SUCCESS) return r;
    }
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd

using namespace magma_amd;

namespace {

// shared driver of the fixed-size entry points
magma_int_t mixed_batched(char kind, bool upper, magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                          magma_int_t **dipiv_array, double **dB_array, magma_int_t lddb, double **dX_array,
                          magma_int_t lddx, magma_int_t *dinfo_array, magma_int_t *diter_array,
                          magma_int_t batchCount, magma_queue_t queue)
{
    hipStream_t st = stream_of(queue);
    const bool mixed_ok = n <= kMixedMaxN;
    const MixedPlan plan = mixed_ok ? mixed_plan(batchCount, n, nrhs) : MixedPlan();
    const size_t wbytes = mixed_ok ? mixed_workspace_bytes(plan, n, true) : 0;
    const size_t lbytes = sizeof(int) * (size_t)(batchCount + 2);
    char *ws = static_cast<char *>(queue_workspace(queue, wbytes + lbytes));
    if (!ws) return MAGMA_ERR_DEVICE_ALLOC;
    int *fb_count = reinterpret_cast<int *>(ws + wbytes);
    int *fb_list = fb_count + 1;
    zero_counts_kernel<<<1, 64, 0, st>>>(fb_count, nullptr);
    MAGMA_LAUNCH_CHECK();
    MixedSet g;
    g.n = n; g.nrhs = nrhs; g.lda = ldda; g.ldb = lddb; g.ldx = lddx;
    g.kind = kind;
    g.A = dA_array; g.B = dB_array; g.X = dX_array;
    if (mixed_ok) mixed_workspace_view(g, plan, n, true, ws);
    g.ipiv = dipiv_array; g.info = dinfo_array; g.iter = diter_array;
    if (kind == 'C') {
        g.fbc_list = fb_list; g.fbc_count = fb_count;
    } else {
        g.fb_list = fb_list; g.fb_count = fb_count;
    }
    g.count = batchCount;
    // Upper: the lower kernels on the swapped triangles, swapped back at the end
    // (A is unchanged on success; on fallback its upper triangle holds U)
    MatSet tri;
    tri.n = n; tri.m = n; tri.lda = ldda;
    tri.A = dA_array;
    tri.count = batchCount;
    if (upper) {
        magma_int_t r = launch_swap_triangles(tri, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    if (mixed_ok) {
        magma_int_t r = launch_mixed(g, n, plan, st);
        if (r != MAGMA_SUCCESS) return r;
    } else {
        // n above the mixed kernels' range: every matrix takes the FP64 path
        // (dsgesv's iter = -1, "implementation-specific reasons")
        fill_fallback_all_kernel<<<ceil_div(batchCount, 256), 256, 0, st>>>(fb_list, fb_count, diter_array,
                                                                             batchCount);
        MAGMA_LAUNCH_CHECK();
    }
    magma_int_t r = launch_mixed_fallback(g, n, nrhs, st);
    if (r == MAGMA_SUCCESS && upper) r = launch_swap_triangles(tri, st);
    if (r != MAGMA_SUCCESS) return r;
    return sync_unless_async(queue);
}

}  // namespace

extern "C" magma_int_t magma_dsgesv_batched(magma_int_t n, magma_int_t nrhs,
                                            double **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                            double **dB_array, magma_int_t lddb,
                                            double **dX_array, magma_int_t lddx,
                                            magma_int_t *dinfo_array, magma_int_t *diter_array,
                                            magma_int_t batchCount, magma_queue_t queue)
{
    // argument checks in the order of magma_dsgesv_gpu (src/zcgesv_gpu.cpp:163-180)
    magma_int_t info = 0;
    if (n < 0) info = -1;
    else if (nrhs < 0) info = -2;
    else if (ldda < std::max(1, n)) info = -4;
    else if (lddb < std::max(1, n)) info = -7;
    else if (lddx < std::max(1, n)) info = -9;
    else if (batchCount < 0) info = -12;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) {
        // the quick return leaves iter = 0 and info = 0 for every matrix (src/zcgesv_gpu.cpp:174-193)
        const hipStream_t st = stream_of(queue);
        magma_int_t e = zero_info(dinfo_array, batchCount, st);
        if (e == MAGMA_SUCCESS) e = zero_info(diter_array, batchCount, st);
        return e != MAGMA_SUCCESS ? e : sync_unless_async(queue);
    }
    return mixed_batched('L', false, n, nrhs, dA_array, ldda, dipiv_array, dB_array, lddb, dX_array, lddx, dinfo_array,
                         diter_array, batchCount, queue);
}

extern "C" magma_int_t magma_dsposv_batched(magma_uplo_t uplo, magma_int_t n, magma_int_t nrhs,
                                            double **dA_array, magma_int_t ldda,
                                            double **dB_array, magma_int_t lddb,
                                            double **dX_array, magma_int_t lddx,
                                            magma_int_t *dinfo_array, magma_int_t *diter_array,
                                            magma_int_t batchCount, magma_queue_t queue)
{
    // argument checks and codes of magma_dsposv_gpu (src/zcposv_gpu.cpp:163-174),
    // which numbers n, nrhs and ldda as if uplo were not an argument
    magma_int_t info = 0;
    if (uplo != MagmaLower && uplo != MagmaUpper) info = -1;
    else if (n < 0) info = -1;
    else if (nrhs < 0) info = -2;
    else if (ldda < std::max(1, n)) info = -4;
    else if (lddb < std::max(1, n)) info = -7;
    else if (lddx < std::max(1, n)) info = -9;
    else if (batchCount < 0) info = -12;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) {
        // the quick return leaves iter = 0 and info = 0 for every matrix (src/zcposv_gpu.cpp:163-182)
        const hipStream_t st = stream_of(queue);
        magma_int_t e = zero_info(dinfo_array, batchCount, st);
        if (e == MAGMA_SUCCESS) e = zero_info(diter_array, batchCount, st);
        return e != MAGMA_SUCCESS ? e : sync_unless_async(queue);
    }
    return mixed_batched('C', uplo == MagmaUpper, n, nrhs, dA_array, ldda, nullptr, dB_array, lddb, dX_array, lddx, dinfo_array,
                         diter_array, batchCount, queue);
}
