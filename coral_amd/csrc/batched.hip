// Batched uneven S/DGEMM on gfx950: thousands of independent problems of
// different shapes in ONE persistent launch that pulls work from a device queue.
//
// No reference counterpart: in CORAL the equivalent is a loop of sgemm/dgemm calls
// (coral-aarch64/src/level3/sgemm.rs:106, dgemm.rs:107), each with its own packing
// buffers.  Here:
//   1. batched_setup_kernel (one workgroup of 1024 threads) reads the per-problem
//      dims from device memory, checks them, picks a tile kind per problem from the
//      (m, n, k) bucket table the autotuner fills, writes the exclusive prefix sum of
//      tile counts (int64) and resets the queue head -- with kernel stores only, so
//      the sequence can be captured in a hipGraph and replayed;
//   2. batched_wave_kernel: every wavefront is an independent worker (no barriers,
//      no LDS).  It takes a chunk of consecutive work items with one atomicAdd
//      (the next chunk's atomic is in flight while the current chunk runs), finds
//      each item's problem with a 64-way search of the prefix sums (one load per lane
//      per level), and computes one C tile of 16x16 FP32 or FP64 MFMAs with its
//      operand fragments loaded straight from the caller's column-major matrices
//      into registers, one k stage ahead of the MFMAs.  Resident waves hide each
//      other's memory latency; the grid is exactly the resident occupancy.
// Tile kinds (one wavefront's C tile): 0 = 16x16, 1 = 32x32, 2 = 64x32 (rows x
// columns), 3 = 32x64.
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
constexpr int kKinds = 4;
// Problem p belongs to queue p % kXcd; the waves of workgroup w take items from queue
// w % kXcd first.  Workgroups are dispatched to the chip's 8 XCDs round-robin, so a
// problem's tiles run on one XCD and its operands are fetched into that XCD's L2
// once instead of into all eight.  A wave whose queue is exhausted takes items from
// the other queues.
constexpr int kXcd = 8;
constexpr int kHeadStride = 16;  // 128 bytes between queue heads
__host__ __device__ inline int kind_rows(int kind) { return kind == 0 ? 16 : (kind == 2 ? 64 : 32); }
__host__ __device__ inline int kind_cols(int kind) { return kind == 0 ? 16 : (kind == 3 ? 64 : 32); }
// Without a tuned entry: 16x16 tiles for problems within 16x16, 64x32 / 32x64 wave
// tiles once both sides reach 64 (more MFMAs per fragment load), 32x32 otherwise.
__host__ __device__ inline int default_kind(int m, int n, int) {
  if (m <= 16 && n <= 16) return 0;
  if (m >= 64 && n >= 64) return m >= n ? 2 : 3;
  return 1;
}

struct BatchArgs {
  int64_t batch;
  const int32_t *m, *n, *k, *lda, *ldb, *ldc;
  const void *const *a;
  const void *const *b;
  void *const *c;
  double alpha, beta;
  int op_a, op_b;
  int64_t *offsets;            // [batch + kXcd]: per-queue prefix sums of tile counts, then queue totals
  int8_t *kinds;               // [batch]
  unsigned long long *head;    // kXcd queue heads, kHeadStride apart
  int *err;                    // the stream's error word
  int table_base;              // offset of this dtype / op pair in g_kind_table
};

// A problem is skipped (no tiles, C untouched) and flagged in the error word when a
// dimension is negative or a leading dimension is below max(1, rows), the checks
// the single-problem routine makes (gemm.hip check_gemm), or when lda or ldb exceeds
// kMaxLd: the tile kernel addresses A and B with 32-bit per-lane byte offsets of up to
// 15 ld + 15 elements.
constexpr int64_t kMaxLd = int64_t(1) << 24;
__device__ __forceinline__ bool problem_valid(const BatchArgs &g, int64_t p) {
  const int64_t m = g.m[p], n = g.n[p], k = g.k[p];
  if (m < 0 || n < 0 || k < 0) return false;
  const int64_t arows = g.op_a == CORAL_NO_TRANSPOSE ? m : k;
  const int64_t brows = g.op_b == CORAL_NO_TRANSPOSE ? k : n;
  const int64_t lda = g.lda[p], ldb = g.ldb[p];
  return lda >= imax64(1, arows) && ldb >= imax64(1, brows) && g.ldc[p] >= imax64(1, m) && lda <= kMaxLd &&
         ldb <= kMaxLd;
}

// Rounds of 1024 consecutive problems (coalesced reads of the dims): each thread
// takes one problem, and a segmented inclusive scan (the problems of one queue are
// the lanes l with equal l % kXcd; then across the 16 waves) turns the tile counts
// into per-queue exclusive prefix sums carried across rounds.
__global__ __launch_bounds__(1024) void batched_setup_kernel(BatchArgs g) {
  static_assert(64 % kXcd == 0 && 1024 % kXcd == 0, "queues are lane classes");
  __shared__ int64_t wsum[16][kXcd];
  __shared__ int64_t s_run[kXcd];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, q = lane % kXcd;
  if (t < kXcd) s_run[t] = 0;
  bool bad = false;
  __syncthreads();
  for (int64_t base = 0; base < g.batch; base += 1024) {
    const int64_t p = base + t;  // p % kXcd == q
    int64_t tiles = 0;
    if (p < g.batch) {
      int kind = 1;
      if (!problem_valid(g, p)) {
        bad = true;
      } else {
        const int m = g.m[p], n = g.n[p], k = g.k[p];
        if (m > 0 && n > 0) {
          const int idx = g.table_base + (bucket_of(m) * kNB + bucket_of(n)) * kNB + bucket_of(k > 0 ? k : 1);
          kind = g_kind_table[idx];
          if (kind < 0 || kind >= kKinds) kind = default_kind(m, n, k);
          const int64_t tr = kind_rows(kind), tc = kind_cols(kind);
          tiles = ((m + tr - 1) / tr) * ((n + tc - 1) / tc);
        }
      }
      g.kinds[p] = (int8_t)kind;
    }
    int64_t v = tiles;  // inclusive scan over this wave's lanes of queue q
#pragma unroll
    for (int off = kXcd; off < 64; off <<= 1) {
      const int64_t u = __shfl_up(v, off, 64);
      if (lane >= off) v += u;
    }
    if (lane >= 64 - kXcd) wsum[w][q] = v;  // the wave's total for queue q
    __syncthreads();
    if (t < kXcd) {  // inclusive scan over the waves, one queue per thread
      int64_t run = 0;
      for (int i = 0; i < 16; ++i) {
        run += wsum[i][t];
        wsum[i][t] = run;
      }
    }
    __syncthreads();
    if (p < g.batch) g.offsets[p] = s_run[q] + (w > 0 ? wsum[w - 1][q] : 0) + v - tiles;
    __syncthreads();  // every thread has read s_run and wsum
    if (t < kXcd) s_run[t] += wsum[15][t];
    __syncthreads();
  }
  if (bad) atomicOr(g.err, (int)ERRF_PROBLEM_ARG);
  if (t < kXcd) {
    g.offsets[g.batch + t] = s_run[t];
    g.head[t * kHeadStride] = 0ull;
  }
}

// Work items are taken kChunk at a time (consecutive items: usually the tiles of one
// problem), which keeps the queue head's atomics far below the L2's rate.
constexpr int kChunk = 4;
// k slices (of TR::KF each) per pipeline stage: two (32 k values for FP32, 16 for
// FP64) measured faster than one for both precisions on the 10,000-problem workload
constexpr int kStageSlices = 2;

// lane 0's value as a wave-uniform (scalar register) value.  Values the wave loads
// from the descriptor arrays are the same in every lane, but the compiler cannot
// prove it (the arrays may alias C), so they are made uniform explicitly; loop
// bounds, branches and base addresses then live in scalar registers.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int32_t uniform32(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename P> __device__ __forceinline__ P *uniform_ptr(P *v) {
  return reinterpret_cast<P *>(uniform64(reinterpret_cast<uintptr_t>(v)));
}

// Scalar (constant address space) load of a descriptor entry: the arrays are only
// read while the tile kernel runs, and the index is wave-uniform.
template <typename T> __device__ __forceinline__ T sload(const T *p, int64_t i) {
  typedef const __attribute__((address_space(4))) T *cptr;
  return ((cptr)(uintptr_t)p)[i];
}

// Queue q holds problems q, q + kXcd, q + 2 kXcd, ...: the j-th of them has prefix sum
// off[q + kXcd j].  Returns the largest j < nq with off[q + kXcd j] <= item (item is
// below the queue's total): a 64-way search, each level one load per lane and a
// ballot.  That problem has tiles, so problems with none are passed over.
__device__ __forceinline__ int64_t find_problem(const int64_t *__restrict__ off, int q, int64_t nq, int64_t item,
                                                int lane) {
  int64_t lo = 0, hi = nq;
  while (hi - lo > 1) {
    const int64_t step = (hi - lo + 63) >> 6;
    const int64_t idx = lo + (int64_t)lane * step;
    const bool le = idx < hi && off[q + kXcd * idx] <= item;
    const uint64_t mask = __ballot(le);  // lanes below a set lane are set too; lane 0 always is
    const int top = 63 - __builtin_clzll(mask);
    lo += (int64_t)top * step;
    hi = lo + step < hi ? lo + step : hi;
  }
  return lo;
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Raw buffer descriptor over [base, base + bytes) (bytes clamped to [0, 2^32 - 1]):
// loads at or past the end return zeros (the gfx9-family untyped-buffer format word).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const char *base, int64_t bytes) {
  const int64_t nb = bytes <= 0 ? 0 : (bytes > (int64_t)0xFFFFFFFF ? (int64_t)0xFFFFFFFF : bytes);
  return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)(uint32_t)nb, 0x00020000);
}

// One element (4 or 8 bytes) at byte offset voff of a buffer, as raw bits.
template <typename T>
__device__ __forceinline__ T buf_elem(__amdgpu_buffer_rsrc_t r, int32_t voff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
  }
}

// The operand fragments of one wave: F 16-row (A) or 16-column (B) fragments starting
// at operand row r0 (for B: column), streamed in k.  Element (r, k) of the operand is
// at byte r * ld * es + k * es when k is contiguous (KC) and r * es + k * ld * es
// otherwise.  Loads go through raw buffer descriptors rebased at each k slice, so
// per-lane offsets are 32-bit and fixed for the whole tile, and everything past the
// operand's last element reads as zero: rows/columns past the operand and, for !KC,
// k past K need no clamping.  (Rows past m or columns past n inside the operand only
// feed results that are not stored.)  For KC the k values past K inside the operand
// are zeroed where the fragment is used (tail_mask).
//
// Lane l holds row r0 + 16 f + l % 16 and k values VE (l / 16) + t, t < VE, of
// fragment f (the MFMA fragment order of mfma.hpp).
template <class TR, bool KC, int F>
struct OperandStream {
  using T = typename TR::T;
  static constexpr int VE = TR::VE, KF = TR::KF, ES = (int)sizeof(T);
  const char *base;  // byte address of (r0, k = 0)
  int64_t extent;    // bytes from base to the end of the operand (<= 0: nothing left)
  int64_t ld;        // elements
  int32_t voff[KC ? 1 : VE];

  __device__ __forceinline__ void init(const void *p, int64_t rows, int64_t K, int64_t ld_, int64_t r0, int lane) {
    ld = ld_;
    const int li = lane & 15, grp = lane >> 4;
    // KC: rows of length K, ld apart; !KC: columns of length rows, ld apart
    const int64_t total = KC ? (rows - 1) * ld + K : (K - 1) * ld + rows;
    const int64_t start = KC ? r0 * ld : r0;
    base = reinterpret_cast<const char *>(p) + start * ES;
    extent = (total - start) * ES;
    if constexpr (KC) {
      voff[0] = (int32_t)(((int64_t)li * ld + VE * grp) * ES);
    } else {
#pragma unroll
      for (int t = 0; t < VE; ++t) voff[t] = (int32_t)((li + (int64_t)(VE * grp + t) * ld) * ES);
    }
  }

  // the F fragments of the k slice starting at k0 (full: k0 + KF <= K)
  __device__ __forceinline__ void load(v4u (&out)[F], int64_t k0, bool full) const {
    if constexpr (KC) {
#pragma unroll
      for (int f = 0; f < F; ++f) {
        // one descriptor per fragment: its 16 rows start 16 f rows further on
        const int64_t off = ((int64_t)16 * f * ld + k0) * ES;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(base + off, extent - off);
        if (full) {
          out[f] = __builtin_bit_cast(v4u, __builtin_amdgcn_raw_buffer_load_b128(r, voff[0], 0, 0));
        } else {
          T v[VE];
#pragma unroll
          for (int t = 0; t < VE; ++t) v[t] = buf_elem<T>(r, voff[0] + t * ES);
          __builtin_memcpy(&out[f], &v[0], 16);
        }
      }
    } else {
      (void)full;
      const int64_t off = k0 * ld * ES;
      const __amdgpu_buffer_rsrc_t r = make_rsrc(base + off, extent - off);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        T v[VE];
#pragma unroll
        for (int t = 0; t < VE; ++t) v[t] = buf_elem<T>(r, voff[t] + 16 * f * ES);
        __builtin_memcpy(&out[f], &v[0], 16);
      }
    }
  }
};

// Zero the values of a fragment vector whose k (kb + t) is at or past K.
template <class TR>
__device__ __forceinline__ v4u tail_mask(v4u f, int64_t kb, int64_t K) {
  using T = typename TR::T;
  constexpr int VE = TR::VE;
  T v[VE];
  __builtin_memcpy(&v[0], &f, 16);
#pragma unroll
  for (int t = 0; t < VE; ++t) v[t] = kb + t < K ? v[t] : T(0);
  __builtin_memcpy(&f, &v[0], 16);
  return f;
}

struct Problem {
  int64_t m, n, K, lda, ldb, ldc;
  const void *a, *b;
  void *c;
};

// One wavefront computes the FM x FN block of 16x16 tiles at tile index t of problem
// P: C := alpha * op(A) op(B) + beta * C on its valid part.
template <class TR, bool AKC, bool BKC, int FM, int FN>
__device__ __forceinline__ void wave_tile(const BatchArgs &g, const Problem &P, int64_t t, int lane) {
  using T = typename TR::T;
  using acc_t = typename TR::acc_t;
  constexpr int KF = TR::KF, VE = TR::VE, KS = kStageSlices;
  constexpr int64_t KST = (int64_t)KF * KS;  // k values per stage
  const int li = lane & 15, grp = lane >> 4;
  const int64_t tiles_m = (P.m + 16 * FM - 1) / (16 * FM);
  const int64_t r0 = (t % tiles_m) * (16 * FM), c0 = (t / tiles_m) * (16 * FN);
  // op(A) rows are k-contiguous for Transpose (AKC); op(B) columns for NoTranspose (BKC)
  OperandStream<TR, AKC, FM> sa;
  OperandStream<TR, BKC, FN> sb;
  const int64_t K = g.alpha == 0.0 ? 0 : P.K;  // alpha == 0: A and B are not read
  if (K > 0) {
    sa.init(P.a, P.m, K, P.lda, r0, lane);
    sb.init(P.b, P.n, K, P.ldb, c0, lane);
  }
  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};
  const int64_t nst = (K + KST - 1) / KST;
  v4u fa0[KS][FM], fb0[KS][FN], fa1[KS][FM], fb1[KS][FN];
  auto load = [&](int64_t st, v4u(&fa)[KS][FM], v4u(&fb)[KS][FN]) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int64_t k0 = st * KST + s * KF;
      const bool full = k0 + KF <= K;
      sa.load(fa[s], k0, full);
      sb.load(fb[s], k0, full);
    }
  };
  auto compute = [&](int64_t st, v4u(&fa)[KS][FM], v4u(&fb)[KS][FN]) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int64_t k0 = st * KST + s * KF;
      if (s > 0 && k0 >= K) break;  // slices past K: nothing to add
      if (k0 + KF > K) {            // the last slice: zero the k-contiguous values past K
        const int64_t kb = k0 + VE * grp;
        if constexpr (AKC) {
#pragma unroll
          for (int i = 0; i < FM; ++i) fa[s][i] = tail_mask<TR>(fa[s][i], kb, K);
        }
        if constexpr (BKC) {
#pragma unroll
          for (int j = 0; j < FN; ++j) fb[s][j] = tail_mask<TR>(fb[s][j], kb, K);
        }
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) TR::mma(acc[i][j], &fa[s][i], &fb[s][j]);
    }
  };
  if (nst > 0) load(0, fa0, fb0);
  for (int64_t st = 0; st < nst; st += 2) {
    if (st + 1 < nst) load(st + 1, fa1, fb1);
    compute(st, fa0, fb0);
    if (st + 1 < nst) {
      if (st + 2 < nst) load(st + 2, fa0, fb0);
      compute(st + 1, fa1, fb1);
    }
  }
  const T al = (T)g.alpha, be = (T)g.beta;
  __attribute__((address_space(1))) T *C = (__attribute__((address_space(1))) T *)reinterpret_cast<T *>(P.c);
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t col = c0 + 16 * j + li;
      // 16x16 accumulator: lane l holds column l % 16; register r holds row
      // 4 * (l / 16) + r (f32 MFMA) or l / 16 + 4 * r (f64 MFMA)
      const int64_t row0 = r0 + 16 * i + (TR::ROWS_INTERLEAVED ? grp : 4 * grp);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = row0 + (TR::ROWS_INTERLEAVED ? 4 * r : r);
        if (row < P.m && col < P.n) {
          __attribute__((address_space(1))) T *cp = C + row + col * P.ldc;
          T o = al * acc[i][j][r];
          if (be != T(0)) o += be * (*cp);
          *cp = o;
        }
      }
    }
}

template <class TR, bool AKC, bool BKC>
__global__ __launch_bounds__(256) void batched_wave_kernel(BatchArgs g) {
  const int lane = threadIdx.x & 63;
  const int home = blockIdx.x % kXcd;
  for (int qi = 0; qi < kXcd; ++qi) {  // the home queue, then the others
    const int q = (home + qi) % kXcd;
    const int64_t nq = (g.batch - q + kXcd - 1) / kXcd;  // problems in queue q
    if (nq <= 0) continue;
    const int64_t total = sload(g.offsets, g.batch + q);
    unsigned long long *head = g.head + q * kHeadStride;
    unsigned long long chunk = 0;
    if (lane == 0) chunk = atomicAdd(head, 1ull);
    chunk = uniform64(chunk);
    for (;;) {
      const int64_t item0 = (int64_t)chunk * kChunk;
      if (item0 >= total) break;  // uniform over the wave: this queue is exhausted
      unsigned long long next = 0;
      if (lane == 0) next = atomicAdd(head, 1ull);  // in flight while this chunk runs
      int64_t j = (int64_t)uniform64((uint64_t)find_problem(g.offsets, q, nq, item0, lane));
      int64_t j_end = j + 1 < nq ? sload(g.offsets, q + kXcd * (j + 1)) : total;
      for (int c = 0; c < kChunk; ++c) {
        const int64_t item = item0 + c;
        if (item >= total) break;
        while (item >= j_end) {  // the next problem of the queue with tiles
          ++j;
          j_end = j + 1 < nq ? sload(g.offsets, q + kXcd * (j + 1)) : total;
        }
        const int64_t p = q + kXcd * j;
        Problem P;
        P.m = sload(g.m, p);
        P.n = sload(g.n, p);
        P.K = sload(g.k, p);
        P.lda = sload(g.lda, p);
        P.ldb = sload(g.ldb, p);
        P.ldc = sload(g.ldc, p);
        P.a = sload(g.a, p);
        P.b = sload(g.b, p);
        P.c = sload(g.c, p);
        const int64_t t = item - sload(g.offsets, p);
        switch (sload(g.kinds, p)) {
          case 0: wave_tile<TR, AKC, BKC, 1, 1>(g, P, t, lane); break;
          case 1: wave_tile<TR, AKC, BKC, 2, 2>(g, P, t, lane); break;
          case 2: wave_tile<TR, AKC, BKC, 4, 2>(g, P, t, lane); break;
          default: wave_tile<TR, AKC, BKC, 2, 4>(g, P, t, lane); break;
        }
      }
      chunk = uniform64(next);
    }
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
            if (kind < 0 || kind >= kKinds) kind = default_kind(1 << bm, 1 << bn, 1 << bk);
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
  const size_t off_bytes = sizeof(int64_t) * (size_t)(batch + kXcd);
  const size_t head_at = (off_bytes + (size_t)batch + 127) & ~size_t(127);
  const size_t bytes = head_at + sizeof(unsigned long long) * kXcd * kHeadStride;
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
  g.head = (unsigned long long *)((char *)ws + head_at);
  g.err = err;
  g.table_base = ((dt * 4) + (g.op_a != 0) * 2 + (g.op_b != 0)) * kNB * kNB * kNB;
  hipLaunchKernelGGL(batched_setup_kernel, dim3(1), dim3(1024), 0, s, g);
  CORAL_LAUNCH_CHECK();
  // op(A) rows k-contiguous (Transpose) / op(B) columns k-contiguous (NoTranspose)
  const int layout = (g.op_a != CORAL_NO_TRANSPOSE ? 2 : 0) + (g.op_b == CORAL_NO_TRANSPOSE ? 1 : 0);
  void (*kern)(BatchArgs) = layout == 0 ? batched_wave_kernel<TR, false, false>
                          : layout == 1 ? batched_wave_kernel<TR, false, true>
                          : layout == 2 ? batched_wave_kernel<TR, true, false>
                                        : batched_wave_kernel<TR, true, true>;
  // persistent: exactly as many workgroups as can be resident at once
  static int per_cu[2][4] = {};
  if (per_cu[dt][layout] == 0) {
    int nb = 0;
    CORAL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, 0));
    per_cu[dt][layout] = nb > 0 ? nb : 1;
  }
  const int grid = per_cu[dt][layout] * num_cus();
  const int slot = timing_begin(TK_GEMM, s);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, g);
  CORAL_LAUNCH_CHECK();
  timing_end(slot, s);
  return CORAL_OK;
}

}  // namespace batched

int batched_menu_size(const char *routine) {
  return (!strcmp(routine, "sgemm_batched") || !strcmp(routine, "dgemm_batched")) ? batched::kKinds : 0;
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
