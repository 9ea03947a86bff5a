// Batched uneven S/DGEMM on gfx950: thousands of independent problems of
// different shapes in ONE persistent launch that pulls work from a device queue.
//
// No reference counterpart: in CORAL the equivalent is a loop of sgemm/dgemm calls
// (coral-aarch64/src/level3/sgemm.rs:106, dgemm.rs:107), each with its own packing
// buffers.  Here:
//   1. batched_count_kernel (a thread per problem) reads the per-problem dims from
//      device memory, checks them, picks a tile kind per problem from the (m, n, k)
//      bucket table the autotuner fills and writes its tile count; batched_scan_kernel
//      (one workgroup of 1024 threads) turns the counts into per-queue exclusive
//      prefix sums (int64) and resets the queue heads -- with kernel stores only, so
//      the sequence can be captured in a hipGraph and replayed;
//   2. batched_wg_kernel: persistent workgroups of 2 x 2 wavefronts.  A workgroup
//      takes one work item (a C tile) at a time with an atomicAdd on its queue head,
//      issued a tile ahead; it finds the item's problem with a 64-way search of the
//      prefix sums (the next 64 problems first) between the current tile's last MFMA
//      and its stores, and computes the tile with 16x16 FP32 or FP64 MFMAs.  The operands' k slices go
//      from the caller's column-major matrices straight into LDS with buffer-to-LDS
//      DMA (one 1 KiB instruction per 16 rows/columns of a slice, split over the four
//      waves), through a three-slot ring that keeps two slices in flight while the
//      MFMAs consume the third; one LDS-only barrier per slice.  The grid is exactly
//      the resident occupancy.
// Tile kinds (the workgroup's C tile, rows x columns): 0 = 32x32, 1 = 64x64,
// 2 = 128x64, 3 = 64x128.
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

// The smallest b with 2^b >= x, at most kNB - 1 (0 for x <= 1).
__host__ __device__ inline int bucket_of(int x) {
  if (x <= 1) return 0;
  const int b = 32 - __builtin_clz((unsigned)(x - 1));
  return b < kNB - 1 ? b : kNB - 1;
}
constexpr int kKinds = 4;
// Problem p belongs to queue p % kXcd; workgroup w takes items from queue w % kXcd
// first.  Workgroups are dispatched to the chip's 8 XCDs round-robin, so a problem's
// tiles run on one XCD and its operands are fetched into that XCD's L2 once instead
// of into all eight.  A workgroup whose queue is exhausted takes items from the
// other queues.
constexpr int kXcd = 8;
constexpr int kHeadStride = 16;  // 128 bytes between queue heads
__host__ __device__ inline int kind_rows(int kind) { return kind == 0 ? 32 : (kind == 2 ? 128 : 64); }
__host__ __device__ inline int kind_cols(int kind) { return kind == 0 ? 32 : (kind == 3 ? 128 : 64); }
// log2 of kind_rows / kind_cols (tile counts by shifts: a 64-bit division is a long
// software routine on the GPU)
__host__ __device__ inline int kind_rows_log2(int kind) { return kind == 0 ? 5 : (kind == 2 ? 7 : 6); }
__host__ __device__ inline int kind_cols_log2(int kind) { return kind == 0 ? 5 : (kind == 3 ? 7 : 6); }
// Without a tuned entry: 32x32 tiles for problems within 32x32, 128x64 / 64x128 once
// the long side reaches 128 and the short one 64 (fewer operand bytes per MFMA),
// 64x64 otherwise.
__host__ __device__ inline int default_kind(int m, int n, int) {
  if (m <= 32 && n <= 32) return 0;
  if (m >= 128 && n >= 64 && m >= n) return 2;
  if (n >= 128 && m >= 64) return 3;
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
__device__ __forceinline__ bool problem_valid(const BatchArgs &g, int64_t m, int64_t n, int64_t k, int64_t lda,
                                              int64_t ldb, int64_t ldc) {
  if (m < 0 || n < 0 || k < 0) return false;
  const int64_t arows = g.op_a == CORAL_NO_TRANSPOSE ? m : k;
  const int64_t brows = g.op_b == CORAL_NO_TRANSPOSE ? k : n;
  return lda >= imax64(1, arows) && ldb >= imax64(1, brows) && ldc >= imax64(1, m) && lda <= kMaxLd && ldb <= kMaxLd;
}

// Setup, in two launches:
//  1. batched_count_kernel, one thread per problem: the problem is checked, gets its
//     tile kind, and its tile count is written to offsets[p].
//  2. batched_scan_kernel, one workgroup of 1024 threads: queue q's problems q,
//     q + kXcd, ... are split into 128 contiguous shares, one per thread of the queue's
//     128 threads (two wavefronts).  Each thread sums its share's counts (kShare loads
//     in flight at a time), the 128 share totals are scanned once (wave shuffles, then
//     the two wavefronts' totals through LDS), and each thread writes the exclusive
//     prefix sums of its share over the counts (from registers when the share fits in
//     one chunk of kShare, else re-reading them).  It also resets the queue heads.
constexpr int kSetupThreads = 1024;
constexpr int kQueueThreads = kSetupThreads / kXcd;  // 128: two wavefronts per queue
constexpr int kShare = 16;                            // counts one thread holds at once

struct ProblemDims {
  int32_t m, n, k, lda, ldb, ldc;
};
// Loads from an index clamped into the batch, so no load sits behind a branch.
__device__ __forceinline__ ProblemDims load_dims(const BatchArgs &g, int64_t p) {
  const int64_t pc = p < g.batch ? p : g.batch - 1;
  return ProblemDims{g.m[pc], g.n[pc], g.k[pc], g.lda[pc], g.ldb[pc], g.ldc[pc]};
}

__device__ __forceinline__ int64_t problem_tiles(const BatchArgs &g, int64_t p, const ProblemDims &d, bool &bad) {
  const bool in = p < g.batch;
  const bool valid = problem_valid(g, d.m, d.n, d.k, d.lda, d.ldb, d.ldc);
  bad |= in && !valid;
  const bool use = in && valid && d.m > 0 && d.n > 0;
  int idx = g.table_base + (bucket_of(d.m) * kNB + bucket_of(d.n)) * kNB + bucket_of(d.k > 0 ? d.k : 1);
  idx = use ? idx : g.table_base;  // an index inside the table either way
  int kind = g_kind_table[idx];
  if (kind < 0 || kind >= kKinds) kind = default_kind(d.m, d.n, d.k);
  kind = use ? kind : 1;
  if (in) g.kinds[p] = (int8_t)kind;
  const int64_t rt = ((int64_t)d.m + kind_rows(kind) - 1) >> kind_rows_log2(kind);
  const int64_t ct = ((int64_t)d.n + kind_cols(kind) - 1) >> kind_cols_log2(kind);
  return use ? rt * ct : 0;
}

__global__ __launch_bounds__(256) void batched_count_kernel(BatchArgs g) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool bad = false;
  const ProblemDims d = load_dims(g, p);
  const int64_t tiles = problem_tiles(g, p, d, bad);
  if (p < g.batch) g.offsets[p] = tiles;
  if (bad) atomicOr(g.err, (int)ERRF_PROBLEM_ARG);
}

__global__ __launch_bounds__(kSetupThreads) void batched_scan_kernel(BatchArgs g) {
  static_assert(kQueueThreads == 128, "two wavefronts per queue");
  __shared__ int64_t wtot[kXcd][2];  // [queue][wavefront of the queue]: inclusive share totals
  const int t = threadIdx.x, lane = t & 63;
  const int q = t / kQueueThreads, c = t % kQueueThreads, hw = c / 64;  // queue, share, its wavefront
  const int64_t nq = g.batch > q ? (g.batch - q + kXcd - 1) / kXcd : 0;  // problems in queue q
  const int64_t len = (nq + kQueueThreads - 1) / kQueueThreads;          // share length (uniform per queue)
  const int64_t j0 = (int64_t)c * len, j1 = j0 + len < nq ? j0 + len : nq;
  auto at = [&](int64_t j) -> int64_t & { return g.offsets[q + (int64_t)kXcd * j]; };
  int64_t v[kShare];
  int64_t sum = 0;
  for (int64_t cb = 0; cb < len; cb += kShare) {  // pass A: the share's total
#pragma unroll
    for (int i = 0; i < kShare; ++i) {
      const int64_t j = j0 + cb + i;
      v[i] = (cb + i < len && j < j1) ? at(j) : 0;
    }
#pragma unroll
    for (int i = 0; i < kShare; ++i) sum += v[i];
  }
  int64_t incl = sum;  // inclusive scan of the share totals over the queue's 128 threads
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t u = __shfl_up(incl, off, 64);
    if (lane >= off) incl += u;
  }
  if (lane == 63) wtot[q][hw] = incl;
  __syncthreads();
  int64_t run = (hw ? wtot[q][0] : 0) + incl - sum;  // exclusive prefix of the share
  if (len <= kShare) {  // pass B from registers (v holds the whole share)
#pragma unroll
    for (int i = 0; i < kShare; ++i) {
      const int64_t j = j0 + i;
      if (i < len && j < j1) at(j) = run;
      run += v[i];
    }
  } else {  // pass B re-reading the counts, chunk by chunk
    for (int64_t cb = 0; cb < len; cb += kShare) {
#pragma unroll
      for (int i = 0; i < kShare; ++i) {
        const int64_t j = j0 + cb + i;
        v[i] = (cb + i < len && j < j1) ? at(j) : 0;
      }
#pragma unroll
      for (int i = 0; i < kShare; ++i) {
        const int64_t j = j0 + cb + i;
        if (cb + i < len && j < j1) at(j) = run;
        run += v[i];
      }
    }
  }
  if (c == 0) {
    g.offsets[g.batch + q] = wtot[q][0] + wtot[q][1];
    g.head[q * kHeadStride] = 0ull;
  }
}

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
// off[q + kXcd j].  Returns the largest j in [lo, hi) with off[q + kXcd j] <= item
// (off[q + kXcd lo] <= item < the queue's total): a 64-way search, each level one load
// per lane and a ballot.  That problem has tiles, so problems with none are passed
// over.
__device__ __forceinline__ int64_t find_problem(const int64_t *__restrict__ off, int q, int64_t lo, int64_t hi,
                                                int64_t item, int lane) {
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

__device__ __forceinline__ int64_t readlane64(int64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(uint64_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)v >> 32), l);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

struct Problem {
  int64_t m, n, K, lda, ldb, ldc;
  const void *a, *b;
  void *c;
};

// A work item's place: problem j of its queue (problem index q + kXcd j), whose
// tiles are the items [start, end), with that problem's descriptor and tile kind.
struct Cursor {
  int64_t j, start, end;
  Problem P;
  int kind;
};

// The cursor of `item` of queue q, given the cursor `from` of an earlier item (items
// only grow, so the problem is from's or a later one; from.j = -1, from.end = 0 for
// none).  The next 64 problems are tried first (one load per lane), then the rest
// of the queue is searched.
__device__ __forceinline__ Cursor locate(const BatchArgs &g, int q, int64_t nq, int64_t total, int64_t item,
                                         const Cursor &from, int lane) {
  Cursor c = from;
  if (item < from.end) return c;
  const int64_t lo = from.j + 1;  // off[q + kXcd lo] = from.end <= item
  const int64_t idx = lo + lane;
  const int64_t v = idx < nq ? g.offsets[q + kXcd * idx] : INT64_MAX;
  const uint64_t mask = __ballot(v <= item);
  const int top = 63 - __builtin_clzll(mask);
  if (top < 63 || lo + 64 >= nq) {
    c.j = lo + top;
    c.start = readlane64(v, top);
    c.end = lo + top + 1 < nq ? readlane64(v, top + 1) : total;
  } else {
    c.j = (int64_t)uniform64((uint64_t)find_problem(g.offsets, q, lo + 63, nq, item, lane));
    c.start = sload(g.offsets, q + kXcd * c.j);
    c.end = c.j + 1 < nq ? sload(g.offsets, q + kXcd * (c.j + 1)) : total;
  }
  const int64_t p = q + kXcd * c.j;
  c.P.m = sload(g.m, p);
  c.P.n = sload(g.n, p);
  c.P.K = sload(g.k, p);
  c.P.lda = sload(g.lda, p);
  c.P.ldb = sload(g.ldb, p);
  c.P.ldc = sload(g.ldc, p);
  c.P.a = sload(g.a, p);
  c.P.b = sload(g.b, p);
  c.P.c = sload(g.c, p);
  c.kind = sload(g.kinds, p);
  return c;
}

// Raw buffer descriptor over [base, base + bytes) (bytes clamped to [0, 2^32 - 1]):
// loads of dwords at or past the end return zeros (the gfx9-family untyped-buffer
// format word).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const char *base, int64_t bytes) {
  const int64_t nb = bytes <= 0 ? 0 : (bytes > (int64_t)0xFFFFFFFF ? (int64_t)0xFFFFFFFF : bytes);
  return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)(uint32_t)nb, 0x00020000);
}

typedef __attribute__((address_space(3))) void *lds_ptr;

// Buffer-to-LDS DMA of 16 bytes per lane (a non-template wrapper: the builtin's
// operand checks then run once, outside any template instantiation).
__device__ __forceinline__ void dma16(__amdgpu_buffer_rsrc_t r, char *dst, int32_t voff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_ptr)dst, 16, voff, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Operand staging.  A k slice of KF values of one operand row (of op(A)) or column
// (of op(B)) is 64 bytes for both precisions (16 floats, 8 doubles).  A tile's slice
// of R rows/columns is R * 64 bytes, moved global -> LDS by R / 16 buffer-to-LDS DMA
// instructions of 1 KiB (64 lanes x 16 bytes; lane l lands at the instruction's LDS
// base + 16 l, and each lane chooses which 16 source bytes it fetches).  Two LDS
// layouts, one per operand storage order:
//
//  !KC (the operand's rows are contiguous: element (r, k) at r + k ld): the slice is
//      KF columns of R elements, column-major (quad L = k * R * es / 16 + r * es / 16),
//      quad L stored at LDS quad position (L & ~15) | ((L & 15) ^ h(k)) with
//      h(k) = es * ((k / VE) % (16 / es)): within every 256 bytes (all 64 banks) the
//      16 / es column groups of one fragment read (lanes of one k offset t read
//      column VE * grp + t, rows 16 i .. 16 i + 15) land on disjoint banks.  A fragment
//      read is VE loads of one element.
//  KC (k contiguous: element (r, k) at r ld + k): instruction j carries rows
//      16 j .. 16 j + 15, each as four 16-byte quads; row c's quad q sits at LDS
//      position 4 c + (q ^ ((c / 4) & 3)) (x 16 bytes), so the 16 lanes of one
//      fragment read (16 rows, the same quad) cover every bank pair once.  A fragment
//      read is one 16-byte load.
//
// Every DMA goes through a raw buffer descriptor rebased at the slice, so lane
// offsets are 32-bit and bytes past the operand's last element read as zero (rows
// past m / columns past n inside the operand only feed results that are not
// stored).  The buffer unit range-checks every dword of a multi-dword load on its own
// (a quad that straddles the operand's end keeps its dwords before the end), so the
// last k slice of a tile, which can run past K, needs no other care for !KC operands
// (columns past K lie past the operand's end); for KC operands the values past K
// that lie inside the operand are zeroed after the fragment read (tail_mask).
// ---------------------------------------------------------------------------
constexpr int kBlock = 1024;              // LDS bytes per DMA instruction
constexpr int kNS = 3;                    // LDS ring: slices in flight while one is used
constexpr int kMaxRows = 192;             // BM + BN of the largest kind
constexpr int kSlot = (kMaxRows / 16) * kBlock;

// Byte offset (32-bit, from the slice base at (r0, k0)) of quad `chunk` of DMA
// instruction j for operand storage KC / !KC, rows R.
template <int ES, int VE>
__device__ __forceinline__ int nkc_swizzle(int k) {
  return ES * ((k / VE) % (16 / ES));
}

template <bool KC, int R, int ES, int VE>
__device__ __forceinline__ int32_t quad_offset(int j, int chunk, int64_t ld) {
  if constexpr (KC) {
    const int c = chunk >> 2;                     // row within the instruction's 16
    const int q = (chunk & 3) ^ ((c >> 2) & 3);   // quad held at this LDS position
    return (int32_t)(((int64_t)c * ld) * ES + q * 16);  // relative to row 16 j
  } else {
    constexpr int QC = R * ES / 16;               // quads per slice column
    const int pos = 64 * j + chunk;               // LDS quad position
    const int u = pos >> 4;                       // 256-byte unit
    // the column group, hence the swizzle, is fixed per unit: 16 u / (VE QC) = k / VE
    const int h = ES * (((16 * u) / (VE * QC)) % (16 / ES));
    const int L = (pos & ~15) | ((pos & 15) ^ h);  // natural quad index held here
    const int k = L / QC, r = (L % QC) * (16 / ES);
    return (int32_t)(((int64_t)r + (int64_t)k * ld) * ES);
  }
}

// Issues DMA instruction j (rows / columns 16 j .. 16 j + 15) of one k slice (k0) of
// one operand into that operand's LDS slice region.
template <bool KC, int R, int ES, int VE>
__device__ __forceinline__ void dma_one(const char *base, int64_t extent, int64_t ld, int64_t k0, char *region,
                                        int j, int lane) {
  // descriptor for this instruction: the slice (!KC), or its 16 rows (KC)
  const int64_t off = KC ? ((int64_t)16 * j * ld + k0) * ES : k0 * ld * ES;
  const __amdgpu_buffer_rsrc_t r = make_rsrc(base + off, extent - off);
  dma16(r, region + kBlock * j, quad_offset<KC, R, ES, VE>(j, lane, ld));
}

// The fragment vector (VE consecutive k values at VE * grp) of operand row rho (a
// multiple of 16 plus li) from one LDS slice region.
template <class TR, bool KC, int R>
__device__ __forceinline__ v4u read_frag(const char *region, int rho, int li, int grp) {
  using T = typename TR::T;
  constexpr int VE = TR::VE, ES = (int)sizeof(T);
  if constexpr (KC) {
    const int pos = 4 * li + (grp ^ ((li >> 2) & 3));
    return *reinterpret_cast<const v4u *>(region + 1024 * (rho >> 4) + 16 * pos);
  } else {
    T v[VE];
#pragma unroll
    for (int t = 0; t < VE; ++t) {
      const int k = VE * grp + t;
      const int b = (k * R + rho) * ES;  // natural byte offset in the column-major slice
      const int L = b >> 4;
      const int pos = (L & ~15) | ((L & 15) ^ nkc_swizzle<ES, VE>(k));
      v[t] = *reinterpret_cast<const T *>(region + 16 * pos + (b & 15));
    }
    v4u out;
    __builtin_memcpy(&out, &v[0], 16);
    return out;
  }
}

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

// Workgroup barrier that orders LDS only: this wave's LDS reads and writes (and scalar
// loads) are complete, then s_barrier.  __syncthreads() would also wait for every
// vector memory operation in flight (its release fence), i.e. for the DMAs of the
// slices still being prefetched.  The LDS DMA of a slice is waited for explicitly
// (wait_vmcnt) before the barrier that publishes it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// s_waitcnt vmcnt(0) with the other counters not waited for (gfx9 encoding: vmcnt
// bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
constexpr int kWaitVm0 = (0x7 << 4) | (0xF << 8);

// Waits until at most N vector memory operations of this wave are in flight.
template <int N> __device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N <= 3, "vmcnt immediates used here");
  if constexpr (N == 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  else if constexpr (N == 1) asm volatile("s_waitcnt vmcnt(1)" ::: "memory");
  else if constexpr (N == 2) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
  else asm volatile("s_waitcnt vmcnt(3)" ::: "memory");
}

// One workgroup (2 x 2 waves) computes the BM x BN tile t of problem P:
// C := alpha * op(A) op(B) + beta * C on its valid part.  The k slices stream through
// a three-slot LDS ring (three arrays), two slices ahead of the MFMAs; one barrier
// per slice.  look() runs between the last MFMA and the stores of C (it finds the
// next tile: its loads then wait for nothing but themselves).  The stores are not
// waited for: the next tile's first DMAs go out behind them, and the first slice's
// wait covers both.
template <class TR, bool AKC, bool BKC, int BM, int BN, class Look>
__device__ __forceinline__ void wg_tile(const BatchArgs &g, const Problem &P, int64_t t, char *ring0, char *ring1,
                                        char *ring2, Look &look) {
  using T = typename TR::T;
  using acc_t = typename TR::acc_t;
  constexpr int KF = TR::KF, VE = TR::VE, ES = (int)sizeof(T);
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int NIA = BM / 16, NIB = BN / 16, NI = NIA + NIB;
  static_assert(NI % 4 == 0 && NI <= kMaxRows / 16, "DMA instructions split evenly over the waves");
  constexpr int D = NI / 4;  // DMA instructions per wave per slice
  constexpr int A_BYTES = NIA * kBlock;
  const int lane = threadIdx.x & 63;
  const int wave = uniform32((int)(threadIdx.x >> 6));
  const int wm = wave & 1, wn = wave >> 1, li = lane & 15, grp = lane >> 4;
  const int64_t tiles_m = (P.m + BM - 1) / BM;
  const int64_t r0 = (t % tiles_m) * BM, c0 = (t / tiles_m) * BN;
  const int64_t K = g.alpha == 0.0 ? 0 : P.K;  // alpha == 0: A and B are not read
  const int64_t nst = (K + KF - 1) / KF;
  // this wave's 16 x 16 fragments that hold rows < m (columns < n): the others are
  // not computed
  const int fmv = (int)imin64(FM, imax64(0, (P.m - r0 - wm * WM + 15) / 16));
  const int fnv = (int)imin64(FN, imax64(0, (P.n - c0 - wn * WN + 15) / 16));

  // operand bases at (r0, k = 0) and bytes to the operand's end
  // KC: rows of length K, ld apart; !KC: columns of length rows, ld apart
  const int64_t a_total = AKC ? (P.m - 1) * P.lda + K : (K - 1) * P.lda + P.m;
  const int64_t a_start = AKC ? r0 * P.lda : r0;
  const char *a_base = reinterpret_cast<const char *>(P.a) + a_start * ES;
  const int64_t a_ext = (a_total - a_start) * ES;
  const int64_t b_total = BKC ? (P.n - 1) * P.ldb + K : (K - 1) * P.ldb + P.n;
  const int64_t b_start = BKC ? c0 * P.ldb : c0;
  const char *b_base = reinterpret_cast<const char *>(P.b) + b_start * ES;
  const int64_t b_ext = (b_total - b_start) * ES;

  // this wave's DMA instructions of slice s_ into slot: instruction g of the
  // combined (A then B) list of NI goes to wave g % 4
  auto issue = [&](char *slot, int64_t s_) {
    const int64_t k0_ = s_ * KF;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const int gi = wave + 4 * i;
      if (gi < NIA) dma_one<AKC, BM, ES, VE>(a_base, a_ext, P.lda, k0_, slot, gi, lane);
      else dma_one<BKC, BN, ES, VE>(b_base, b_ext, P.ldb, k0_, slot + A_BYTES, gi - NIA, lane);
    }
  };

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};

  // slice s uses slot s % 3 and, after the barrier, refills slot (s + 2) % 3 (the one
  // slice s - 1 used) with slice s + 2
  auto step = [&](const char *cur, char *refill, int64_t s_) {
    // this wave's DMAs of slice s_ are done once at most those of slice s_ + 1 are
    // pending (vector memory operations of one kind complete in issue order)
    if (s_ + 1 < nst) wait_vmcnt<D>();
    else wait_vmcnt<0>();
    lds_barrier();  // every wave's slice s_ is in LDS; every wave is done with slice s_ - 1
    if (s_ + 2 < nst) issue(refill, s_ + 2);
    const int64_t k0 = s_ * KF;
    v4u fa[FM], fb[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) fa[i] = read_frag<TR, AKC, BM>(cur, wm * WM + 16 * i + li, li, grp);
#pragma unroll
    for (int j = 0; j < FN; ++j) fb[j] = read_frag<TR, BKC, BN>(cur + A_BYTES, wn * WN + 16 * j + li, li, grp);
    if (k0 + KF > K) {  // the last slice: zero the k-contiguous values past K
      const int64_t kb = k0 + VE * grp;
      if constexpr (AKC) {
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = tail_mask<TR>(fa[i], kb, K);
      }
      if constexpr (BKC) {
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = tail_mask<TR>(fb[j], kb, K);
      }
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        if (i < fmv && j < fnv) TR::mma(acc[i][j], &fb[j], &fa[i]);  // C^T tile
  };

  // The vector memory counter also holds the previous tile's C stores and the queue
  // atomic, all issued before these DMAs: a wait for "at most the DMAs of slice s + 1
  // pending" still implies slice s is done, since DMAs complete in issue order.
  if (nst > 0) issue(ring0, 0);
  if (nst > 1) issue(ring1, 1);
  // unrolled by the ring size so that every LDS access names its slot's own array
  for (int64_t s = 0; s < nst; s += kNS) {
    step(ring0, ring2, s);
    if (s + 1 < nst) step(ring1, ring0, s + 1);
    if (s + 2 < nst) step(ring2, ring1, s + 2);
  }
  look();
  // The MFMAs compute the transposed tile (B fragment as the first operand), so
  // lane l holds C row l % 16 and register r holds column 4 * (l / 16) + r (f32 MFMA)
  // or l / 16 + 4 * r (f64 MFMA): a store instruction writes 16 consecutive rows of
  // four columns, i.e. four contiguous 64- or 128-byte runs.  beta == 0 and beta != 0
  // are separate paths: in the first no load of C exists whose destination registers
  // a store would have to wait for (which, with stores in flight, would be a wait for
  // all of them); in the second the four old values of a fragment are loaded
  // before its stores.
  const T al = (T)g.alpha, be = (T)g.beta;
  __attribute__((address_space(1))) T *C = (__attribute__((address_space(1))) T *)reinterpret_cast<T *>(P.c);
  const int64_t row_base = r0 + wm * WM + li;
  const int64_t col_base = c0 + wn * WN + (TR::ROWS_INTERLEAVED ? grp : 4 * grp);
  auto cidx = [&](int i, int j, int r, bool &in) {
    const int64_t row = row_base + 16 * i;
    const int64_t col = col_base + 16 * j + (TR::ROWS_INTERLEAVED ? 4 * r : r);
    in = row < P.m && col < P.n;
    return row + col * P.ldc;
  };
  if (be == T(0)) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        if (!(i < fmv && j < fnv)) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          bool in;
          const int64_t x = cidx(i, j, r, in);
          if (in) C[x] = al * acc[i][j][r];
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        if (!(i < fmv && j < fnv)) continue;
        T old[4];
        bool in[4];
        int64_t x[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          x[r] = cidx(i, j, r, in[r]);
          old[r] = in[r] ? C[x[r]] : T(0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T o = al * acc[i][j][r];  // the single-problem routine's epilogue, operation for operation
          o += be * old[r];
          if (in[r]) C[x[r]] = o;
        }
      }
    // an s_waitcnt the compiler sees: after the join no load of this branch is still
    // pending in its bookkeeping, so the code that follows has no waits inserted for it
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
  }
  lds_barrier();  // the ring's slots are free for the next tile
}

template <class TR, bool AKC, bool BKC>
__global__ __launch_bounds__(256, 4) void batched_wg_kernel(BatchArgs g) {
  // the LDS ring: three separate arrays, so that the compiler can tell a read of one
  // slot from a DMA into another and does not wait on that DMA
  __shared__ __attribute__((aligned(16))) char ring0[kSlot];
  __shared__ __attribute__((aligned(16))) char ring1[kSlot];
  __shared__ __attribute__((aligned(16))) char ring2[kSlot];
  __shared__ unsigned long long s_item;
  const int t = threadIdx.x, lane = t & 63;
  const int home = blockIdx.x % kXcd;
  for (int qi = 0; qi < kXcd; ++qi) {  // the home queue, then the others
    const int q = (home + qi) % kXcd;
    const int64_t nq = (g.batch - q + kXcd - 1) / kXcd;  // problems in queue q
    if (nq <= 0) continue;
    const int64_t total = sload(g.offsets, g.batch + q);
    unsigned long long *head = g.head + q * kHeadStride;
    if (t == 0) s_item = atomicAdd(head, 1ull);
    __syncthreads();
    int64_t item = (int64_t)uniform64(s_item);
    lds_barrier();  // every thread has read s_item
    if (item >= total) continue;  // uniform over the workgroup: this queue is exhausted
    Cursor none;
    none.j = -1;
    none.start = none.end = 0;
    Cursor cur = locate(g, q, nq, total, item, none, lane);
    // thread 0: the queue atomic for the item after the next one to be read, issued
    // once the search's loads are done; its value is published one tile later
    unsigned long long ahead = 0;
    if (t == 0) ahead = atomicAdd(head, 1ull);
    for (;;) {
      Cursor nx = cur;
      int64_t next = total;
      auto look = [&]() {
        if (t == 0) s_item = ahead;
        lds_barrier();  // s_item is published (and every read of its previous value is done)
        next = (int64_t)uniform64(s_item);
        if (next < total) {
          nx = locate(g, q, nq, total, next, cur, lane);
          if (t == 0) ahead = atomicAdd(head, 1ull);
        }
      };
      const int64_t tt = item - cur.start;
      switch (cur.kind) {
        case 0: wg_tile<TR, AKC, BKC, 32, 32>(g, cur.P, tt, ring0, ring1, ring2, look); break;
        case 1: wg_tile<TR, AKC, BKC, 64, 64>(g, cur.P, tt, ring0, ring1, ring2, look); break;
        case 2: wg_tile<TR, AKC, BKC, 128, 64>(g, cur.P, tt, ring0, ring1, ring2, look); break;
        default: wg_tile<TR, AKC, BKC, 64, 128>(g, cur.P, tt, ring0, ring1, ring2, look); break;
      }
      if (next >= total) break;  // uniform: this queue is exhausted
      cur = nx;
      item = next;
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
  hipLaunchKernelGGL(batched_count_kernel, dim3((unsigned)cdiv(batch, 256)), dim3(256), 0, s, g);
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(batched_scan_kernel, dim3(1), dim3(kSetupThreads), 0, s, g);
  CORAL_LAUNCH_CHECK();
  // op(A) rows k-contiguous (Transpose) / op(B) columns k-contiguous (NoTranspose)
  const int layout = (g.op_a != CORAL_NO_TRANSPOSE ? 2 : 0) + (g.op_b == CORAL_NO_TRANSPOSE ? 1 : 0);
  void (*kern)(BatchArgs) = layout == 0 ? batched_wg_kernel<TR, false, false>
                          : layout == 1 ? batched_wg_kernel<TR, false, true>
                          : layout == 2 ? batched_wg_kernel<TR, true, false>
                                        : batched_wg_kernel<TR, true, true>;
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
