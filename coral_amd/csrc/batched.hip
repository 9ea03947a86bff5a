// Batched uneven S/DGEMM on gfx950: thousands of independent problems of
// different shapes in ONE persistent launch that pulls work from a device queue.
//
// No reference counterpart: in CORAL the equivalent is a loop of sgemm/dgemm calls
// (coral-aarch64/src/level3/sgemm.rs:106, dgemm.rs:107), each with its own packing
// buffers.  Here:
//   1. batched_setup_kernel (one workgroup) reads the per-problem dims from device
//      memory, picks a tile kind per problem from the (m, n, k) bucket table the
//      autotuner fills, writes the exclusive prefix sum of tile counts and resets
//      the queue head -- with kernel stores only, so the sequence can be captured
//      in a hipGraph and replayed;
//   2. batched_gemm_kernel runs ~8 workgroups per CU; each repeatedly takes the
//      next work item with one atomicAdd, binary-searches the prefix sums for its
//      problem, computes one C tile with 16x16 MFMAs (FP32 or FP64) from LDS-staged
//      operand slices read straight from the caller's column-major matrices (with
//      bounds checks; no packing pass), and applies alpha/beta.  Every workgroup
//      leaves when the queue is exhausted.
// Tile kinds: 0 = 32x32 C tile (one 16x16 MFMA tile per wavefront), 1 = 64x64
// (2x2 per wavefront).
#include "coral_common.hpp"
#include "mfma.hpp"

#include <mutex>
#include <string.h>
#include <vector>

namespace coral {

int autotune_version();  // autotune.hip

namespace batched {
using namespace coral::gemm;

constexpr int kNB = 16;  // log2 buckets per dimension (dims up to 2^15; larger clamp)
constexpr int kTableSize = 2 * 4 * kNB * kNB * kNB;  // [dtype][op pair][bm][bn][bk]
__device__ int8_t g_kind_table[kTableSize];

__host__ __device__ inline int bucket_of(int x) {
  int b = 0;
  while (b < kNB - 1 && (1 << b) < x) ++b;
  return b;
}
__host__ __device__ inline int default_kind(int m, int n, int) { return (m <= 32 && n <= 32) ? 0 : 1; }

struct BatchArgs {
  int64_t batch;
  const int32_t *m, *n, *k, *lda, *ldb, *ldc;
  const void *const *a;
  const void *const *b;
  void *const *c;
  double alpha, beta;
  int op_a, op_b;
  int64_t *offsets;            // [batch + 1] prefix sums of tile counts
  int8_t *kinds;               // [batch]
  unsigned long long *head;    // queue head
  int *err;                    // the stream's error word
  int table_base;              // offset of this dtype / op pair in g_kind_table
};

// A problem is skipped (no tiles, C untouched) and flagged in the error word when a
// dimension is negative or a leading dimension is below max(1, rows), the checks
// the single-problem routine makes (gemm.hip check_gemm).
__device__ __forceinline__ bool problem_valid(const BatchArgs &g, int64_t p) {
  const int64_t m = g.m[p], n = g.n[p], k = g.k[p];
  if (m < 0 || n < 0 || k < 0) return false;
  const int64_t arows = g.op_a == CORAL_NO_TRANSPOSE ? m : k;
  const int64_t brows = g.op_b == CORAL_NO_TRANSPOSE ? k : n;
  return g.lda[p] >= imax64(1, arows) && g.ldb[p] >= imax64(1, brows) && g.ldc[p] >= imax64(1, m);
}

__global__ __launch_bounds__(256) void batched_setup_kernel(BatchArgs g) {
  __shared__ int64_t sums[256];
  const int t = threadIdx.x;
  const int64_t per = (g.batch + 255) / 256;
  const int64_t p0 = t * per, p1 = p0 + per < g.batch ? p0 + per : g.batch;
  int64_t local = 0;
  bool bad = false;
  for (int64_t p = p0; p < p1; ++p) {
    const int m = g.m[p], n = g.n[p], k = g.k[p];
    int kind = 1;
    int64_t tiles = 0;
    if (!problem_valid(g, p)) {
      bad = true;
    } else if (m > 0 && n > 0) {
      const int idx = g.table_base + (bucket_of(m) * kNB + bucket_of(n)) * kNB + bucket_of(k > 0 ? k : 1);
      kind = g_kind_table[idx];
      if (kind < 0 || kind > 1) kind = default_kind(m, n, k);
      const int64_t ts = kind == 0 ? 32 : 64;
      tiles = ((m + ts - 1) / ts) * ((n + ts - 1) / ts);
    }
    g.kinds[p] = (int8_t)kind;
    g.offsets[p] = tiles;  // count for now
    local += tiles;
  }
  if (bad) atomicOr(g.err, (int)ERRF_PROBLEM_ARG);
  sums[t] = local;
  __syncthreads();
  // inclusive scan of the 256 per-thread totals (Hillis-Steele)
  for (int off = 1; off < 256; off <<= 1) {
    int64_t v = t >= off ? sums[t - off] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  int64_t run = t > 0 ? sums[t - 1] : 0;
  for (int64_t p = p0; p < p1; ++p) {
    const int64_t c = g.offsets[p];
    g.offsets[p] = run;
    run += c;
  }
  if (t == 255) {
    g.offsets[g.batch] = sums[255];
    *g.head = 0ull;
  }
}

// KS KF-deep slices of op(A) rows [r0, r0+TS) (or op(B) columns) in registers.
// load() reads them from indices clamped into the operand, so every load is
// unconditional and all are in flight at once; store() writes them to LDS in
// fragment order, element (r, kk) of slice s at s*TS*KF +
// ((r/16)*64 + (r%16) + 16*(kk/VE))*VE + kk%VE, with zeros outside the operand.
template <class TR, int TS, int KS>
struct Slices {
  using T = typename TR::T;
  static constexpr int VE = TR::VE, KF = TR::KF, COUNT = TS * KF, PER = COUNT / 256;
  static_assert(COUNT % 256 == 0, "a slice splits evenly over the workgroup");
  T v[KS][PER];
  __device__ __forceinline__ void coords(int u, bool kcontig, int &r, int &kk) const {
    // neighbouring threads take neighbouring addresses of the source layout
    const int e = threadIdx.x + 256 * u;
    r = kcontig ? e / KF : e % TS;
    kk = kcontig ? e % KF : e / TS;
  }
  __device__ __forceinline__ void load(const T *p, int64_t ld, bool kcontig, int64_t r0, int64_t rmax, int64_t k0,
                                       int64_t kmax) {
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        int r, kk;
        coords(u, kcontig, r, kk);
        const int64_t gr = r0 + r < rmax ? r0 + r : rmax - 1;
        const int64_t gk = k0 + s * KF + kk < kmax ? k0 + s * KF + kk : kmax - 1;
        v[s][u] = kcontig ? p[gk + gr * ld] : p[gr + gk * ld];
      }
  }
  __device__ __forceinline__ void store(T *lds, bool kcontig, int64_t r0, int64_t rmax, int64_t k0,
                                        int64_t kmax) const {
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        int r, kk;
        coords(u, kcontig, r, kk);
        const bool in = r0 + r < rmax && k0 + s * KF + kk < kmax;
        lds[s * COUNT + ((r >> 4) * 64 + (r & 15) + 16 * (kk / VE)) * VE + (kk % VE)] = in ? v[s][u] : T(0);
      }
  }
};

// k slices per pipeline stage: the next stage's loads are in flight while this
// stage's MFMAs run
constexpr int kBatchKS = 2;

template <class TR, int TS>
__device__ __forceinline__ void compute_tile(const BatchArgs &g, int64_t p, int64_t t, v4u *ldsA, v4u *ldsB) {
  using T = typename TR::T;
  using acc_t = typename TR::acc_t;
  constexpr int F = TS / 32;  // 16x16 tiles per wavefront per dimension
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m = g.m[p], n = g.n[p], K = g.k[p] > 0 ? g.k[p] : 0;
  const int64_t lda = g.lda[p], ldb = g.ldb[p], ldc = g.ldc[p];
  const int64_t tiles_m = (m + TS - 1) / TS;
  const int64_t r0 = (t % tiles_m) * TS, c0 = (t / tiles_m) * TS;
  const T *A = reinterpret_cast<const T *>(g.a[p]);
  const T *B = reinterpret_cast<const T *>(g.b[p]);
  T *C = reinterpret_cast<T *>(g.c[p]);
  acc_t acc[F][F];
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j) acc[i][j] = acc_t{};
  // op(A) row-contiguous for N; op(B) column index contiguous for T
  const bool a_kcontig = g.op_a != CORAL_NO_TRANSPOSE;
  const bool b_kcontig = g.op_b == CORAL_NO_TRANSPOSE;
  const int64_t kend = g.alpha == 0.0 ? 0 : K;  // alpha == 0: A and B are not read
  constexpr int KS = kBatchKS;
  constexpr int64_t KB = (int64_t)KS * TR::KF;
  constexpr int SLICE_V = TS * TR::KF / TR::VE;  // 16-byte vectors per slice
  Slices<TR, TS, KS> sa, sb;
  if (kend > 0) {
    sa.load(A, lda, a_kcontig, r0, m, 0, K);
    sb.load(B, ldb, b_kcontig, c0, n, 0, K);
  }
  for (int64_t k0 = 0; k0 < kend; k0 += KB) {
    sa.store(reinterpret_cast<T *>(ldsA), a_kcontig, r0, m, k0, K);
    sb.store(reinterpret_cast<T *>(ldsB), b_kcontig, c0, n, k0, K);
    __syncthreads();
    if (k0 + KB < kend) {
      sa.load(A, lda, a_kcontig, r0, m, k0 + KB, K);
      sb.load(B, ldb, b_kcontig, c0, n, k0 + KB, K);
    }
#pragma unroll
    for (int sl = 0; sl < KS; ++sl) {  // slices past K hold zeros in both operands
      v4u fa[F], fb[F];
#pragma unroll
      for (int i = 0; i < F; ++i) fa[i] = ldsA[sl * SLICE_V + (wm * F + i) * 64 + lane];
#pragma unroll
      for (int j = 0; j < F; ++j) fb[j] = ldsB[sl * SLICE_V + (wn * F + j) * 64 + lane];
#pragma unroll
      for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j) TR::mma(acc[i][j], &fa[i], &fb[j]);
    }
    __syncthreads();
  }
  const T al = (T)g.alpha, be = (T)g.beta;
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j) {
      const int64_t col = c0 + (wn * F + j) * 16 + (lane & 15);
      const int64_t row0 = r0 + (wm * F + i) * 16 + (TR::ROWS_INTERLEAVED ? (lane >> 4) : 4 * (lane >> 4));
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = row0 + (TR::ROWS_INTERLEAVED ? 4 * r : r);
        if (row < m && col < n) {
          T *cp = C + row + col * ldc;
          T o = al * acc[i][j][r];
          if (be != T(0)) o += be * (*cp);
          *cp = o;
        }
      }
    }
}

template <class TR>
__global__ __launch_bounds__(256) void batched_gemm_kernel(BatchArgs g) {
  __shared__ v4u ldsA[kBatchKS * 4 * 64], ldsB[kBatchKS * 4 * 64];
  __shared__ int64_t s_item;
  const int64_t total = g.offsets[g.batch];
  for (;;) {
    if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(g.head, 1ull);
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();  // s_item is rewritten at the top of the next iteration
    if (item >= total) return;  // uniform: the whole workgroup leaves
    // problem p with offsets[p] <= item < offsets[p + 1] (first such p)
    int64_t lo = 0, hi = g.batch - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (g.offsets[mid] <= item) lo = mid; else hi = mid - 1;
    }
    // the last p with offsets[p] <= item: empty problems (equal offsets) are passed over
    const int64_t p = lo;
    const int64_t t = item - g.offsets[p];
    if (g.kinds[p] == 0) compute_tile<TR, 32>(g, p, t, ldsA, ldsB);
    else compute_tile<TR, 64>(g, p, t, ldsA, ldsB);
  }
}

// host-side kind table (mirrors g_kind_table), rebuilt when the autotuner changes
static std::mutex g_mu;
static int g_table_version = -1;

static int upload_table(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  const int v = autotune_version();
  if (v == g_table_version) return CORAL_OK;
  std::vector<int8_t> host(kTableSize);
  const char *names[2] = {"sgemm_batched", "dgemm_batched"};
  for (int d = 0; d < 2; ++d)
    for (int ops = 0; ops < 4; ++ops)
      for (int bm = 0; bm < kNB; ++bm)
        for (int bn = 0; bn < kNB; ++bn)
          for (int bk = 0; bk < kNB; ++bk) {
            const int idx = (((d * 4 + ops) * kNB + bm) * kNB + bn) * kNB + bk;
            int kind = tuned_tile(names[d], ops >> 1, ops & 1, int64_t(1) << bm, int64_t(1) << bn, int64_t(1) << bk);
            if (kind < 0 || kind > 1) kind = default_kind(1 << bm, 1 << bn, 1 << bk);
            host[idx] = (int8_t)kind;
          }
  CORAL_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(g_kind_table), host.data(), kTableSize, 0, hipMemcpyHostToDevice, s));
  CORAL_HIP(hipStreamSynchronize(s));
  g_table_version = v;
  return CORAL_OK;
}

// CORAL-style argument order: op_a=1 op_b=2 batch=3 m=4 n=5 k=6 alpha=7 a=8 lda=9 b=10 ldb=11 beta=12 c=13 ldc=14
template <class TR>
static int batched_impl(int dt, int32_t op_a, int32_t op_b, int64_t batch, const int32_t *m, const int32_t *n,
                        const int32_t *k, double alpha, const void *const *a, const int32_t *lda,
                        const void *const *b, const int32_t *ldb, double beta, void *const *c,
                        const int32_t *ldc, coral_stream_t stream) {
  if (!valid_op(op_a)) return -1;
  if (!valid_op(op_b)) return -2;
  if (batch < 0 || batch > (int64_t)1 << 30) return -3;
  if (batch == 0) return CORAL_OK;
  if (!m) return -4;
  if (!n) return -5;
  if (!k) return -6;
  if (!a) return -8;
  if (!lda) return -9;
  if (!b) return -10;
  if (!ldb) return -11;
  if (!c) return -13;
  if (!ldc) return -14;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  CORAL_HIP(hipStreamIsCapturing(s, &cap));
  if (cap == hipStreamCaptureStatusNone) CORAL_TRY(upload_table(s));
  void *ws = nullptr;
  int *err = nullptr;
  const size_t off_bytes = sizeof(int64_t) * (size_t)(batch + 1);
  const size_t bytes = off_bytes + (size_t)batch + 64;
  CORAL_TRY(workspace(s, WS_BATCH, bytes + 256, &ws));
  CORAL_TRY(error_word(s, &err));
  BatchArgs g;
  g.batch = batch;
  g.m = m; g.n = n; g.k = k; g.lda = lda; g.ldb = ldb; g.ldc = ldc;
  g.a = a; g.b = b; g.c = c;
  g.alpha = alpha;
  g.beta = beta;
  g.op_a = op_a == CORAL_CONJUGATE_TRANSPOSE ? CORAL_TRANSPOSE : op_a;  // real: C == T
  g.op_b = op_b == CORAL_CONJUGATE_TRANSPOSE ? CORAL_TRANSPOSE : op_b;
  g.offsets = (int64_t *)ws;
  g.kinds = (int8_t *)ws + off_bytes;
  g.head = (unsigned long long *)((char *)ws + ((off_bytes + batch + 63) & ~size_t(63)));
  g.err = err;
  g.table_base = ((dt * 4) + (g.op_a != 0) * 2 + (g.op_b != 0)) * kNB * kNB * kNB;
  hipLaunchKernelGGL(batched_setup_kernel, dim3(1), dim3(256), 0, s, g);
  CORAL_LAUNCH_CHECK();
  const int grid = 8 * num_cus();
  const int slot = timing_begin(TK_GEMM, s);
  hipLaunchKernelGGL(batched_gemm_kernel<TR>, dim3(grid), dim3(256), 0, s, g);
  CORAL_LAUNCH_CHECK();
  timing_end(slot, s);
  return CORAL_OK;
}

}  // namespace batched

int batched_menu_size(const char *routine) {
  return (!strcmp(routine, "sgemm_batched") || !strcmp(routine, "dgemm_batched")) ? 2 : 0;
}

}  // namespace coral

using namespace coral;
using namespace coral::batched;

extern "C" {

int coral_sgemm_batched(int32_t op_a, int32_t op_b, int64_t batch, const int32_t *m, const int32_t *n,
                        const int32_t *k, float alpha, const float *const *a, const int32_t *lda,
                        const float *const *b, const int32_t *ldb, float beta, float *const *c, const int32_t *ldc,
                        coral_stream_t stream) {
  return batched_impl<TF32>(0, op_a, op_b, batch, m, n, k, alpha, (const void *const *)a, lda,
                            (const void *const *)b, ldb, beta, (void *const *)c, ldc, stream);
}

int coral_dgemm_batched(int32_t op_a, int32_t op_b, int64_t batch, const int32_t *m, const int32_t *n,
                        const int32_t *k, double alpha, const double *const *a, const int32_t *lda,
                        const double *const *b, const int32_t *ldb, double beta, double *const *c,
                        const int32_t *ldc, coral_stream_t stream) {
  return batched_impl<TF64>(1, op_a, op_b, batch, m, n, k, alpha, (const void *const *)a, lda,
                            (const void *const *)b, ldb, beta, (void *const *)c, ldc, stream);
}

}  // extern "C"
