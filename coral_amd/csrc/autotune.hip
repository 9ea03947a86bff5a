// Autotuner: per (routine, op pair, size bucket) tile choice, persisted as text.
//
// CORAL fixes its blocking at compile time (MC/NC/KC = 384/576/256 and MR x NR = 8 x 12
// for SGEMM, coral-aarch64/src/level3/sgemm.rs:92-94, f32_packers.rs:1-2).  On the GPU
// the best macro-tile depends on the problem shape relative to the 256 CUs, so the
// choice is made per bucket at run time: a bucket is (ceil(log2 m), ceil(log2 n),
// ceil(log2 k)).  Lookups fall back to a shape heuristic when no entry exists.
#include "coral_common.hpp"

#include <dlfcn.h>

#include <map>
#include <mutex>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

namespace coral {

static std::mutex g_tune_mu;
static std::map<std::string, int> g_table;
static thread_local int g_forced_tile = -1;
static int g_version = 0;  // bumped on every table change

// Changes whenever a lookup could return something new (table edits or a forced tile).
int autotune_version() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_version * 16 + (g_forced_tile + 1);
}

static int bucket(int64_t x) {
  int b = 0;
  while (b < 62 && (int64_t(1) << b) < x) ++b;
  return b;
}

static std::string key_of(const char *routine, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  char buf[160];
  snprintf(buf, sizeof(buf), "%s %d %d %d %d %d", routine, opa, opb, bucket(m), bucket(n), bucket(k));
  return buf;
}

int tile_count(const char *routine) {
  if (!routine) return 0;
  int g = gemm_menu_size(routine);
  if (g) return g;
  return batched_menu_size(routine);
}

int tuned_tile(const char *routine, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  if (g_forced_tile >= 0) return g_forced_tile;
  static const char *env = getenv("CORAL_TILE");
  if (env && *env) return atoi(env);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_table.find(key_of(routine, opa, opb, m, n, k));
  return it == g_table.end() ? -1 : it->second;
}

struct ForcedTile {
  explicit ForcedTile(int t) { g_forced_tile = t; }
  ~ForcedTile() { g_forced_tile = -1; }
};

// Batched buckets: a homogeneous batch of P problems of shape (m, n, k) (shared A and
// B, distinct C blocks) is timed with each tile kind forced; the winner is stored for
// the (m, n, k) bucket.
static int tune_batched(const char *routine, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                        int32_t *chosen) {
  const bool dbl = routine[0] == 'd';
  const size_t es = dbl ? 8 : 4;
  int64_t P = ((int64_t)1 << 24) / (m * n);
  P = P < 16 ? 16 : (P > 4096 ? 4096 : P);
  const int64_t lda = op_a == CORAL_NO_TRANSPOSE ? m : k, ldb = op_b == CORAL_NO_TRANSPOSE ? k : n;
  void *a = nullptr, *b = nullptr, *c = nullptr, *dims = nullptr, *ptrs = nullptr;
  CORAL_HIP(hipMalloc(&a, es * (size_t)(m * k)));
  CORAL_HIP(hipMalloc(&b, es * (size_t)(k * n)));
  CORAL_HIP(hipMalloc(&c, es * (size_t)(m * n * P)));
  CORAL_HIP(hipMalloc(&dims, sizeof(int32_t) * 6 * (size_t)P));
  CORAL_HIP(hipMalloc(&ptrs, sizeof(void *) * 3 * (size_t)P));
  std::vector<int32_t> hd(6 * (size_t)P);
  std::vector<void *> hp(3 * (size_t)P);
  for (int64_t p = 0; p < P; ++p) {
    hd[p] = (int32_t)m; hd[P + p] = (int32_t)n; hd[2 * P + p] = (int32_t)k;
    hd[3 * P + p] = (int32_t)lda; hd[4 * P + p] = (int32_t)ldb; hd[5 * P + p] = (int32_t)m;
    hp[p] = a; hp[P + p] = b; hp[2 * P + p] = (char *)c + es * (size_t)(m * n * p);
  }
  CORAL_HIP(hipMemcpy(dims, hd.data(), hd.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  CORAL_HIP(hipMemcpy(ptrs, hp.data(), hp.size() * sizeof(void *), hipMemcpyHostToDevice));
  if (dbl) {
    coral_fill_random_f64((double *)a, m * k, 11, 0, nullptr);
    coral_fill_random_f64((double *)b, k * n, 12, 0, nullptr);
  } else {
    coral_fill_random_f32((float *)a, m * k, 11, 0, nullptr);
    coral_fill_random_f32((float *)b, k * n, 12, 0, nullptr);
  }
  const int32_t *dm = (const int32_t *)dims;
  void **pp = (void **)ptrs;
  auto run_once = [&]() -> int {
    if (dbl)
      return coral_dgemm_batched(op_a, op_b, P, dm, dm + P, dm + 2 * P, 1.0, (const double *const *)pp, dm + 3 * P,
                                 (const double *const *)(pp + P), dm + 4 * P, 0.0, (double *const *)(pp + 2 * P),
                                 dm + 5 * P, nullptr);
    return coral_sgemm_batched(op_a, op_b, P, dm, dm + P, dm + 2 * P, 1.0f, (const float *const *)pp, dm + 3 * P,
                               (const float *const *)(pp + P), dm + 4 * P, 0.0f, (float *const *)(pp + 2 * P),
                               dm + 5 * P, nullptr);
  };
  hipEvent_t e0, e1;
  CORAL_HIP(hipEventCreate(&e0));
  CORAL_HIP(hipEventCreate(&e1));
  int best = -1, status = CORAL_OK;
  float best_ms = 1e30f;
  for (int t = 0; t < batched_menu_size(routine) && status == CORAL_OK; ++t) {
    ForcedTile ft(t);
    status = run_once();
    if (status != CORAL_OK) break;
    CORAL_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < 3 && status == CORAL_OK; ++r) status = run_once();
    CORAL_HIP(hipEventRecord(e1, nullptr));
    CORAL_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CORAL_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) { best_ms = ms; best = t; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(c); (void)hipFree(dims); (void)hipFree(ptrs);
  if (status != CORAL_OK) return status;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_table[key_of(routine, op_a == CORAL_CONJUGATE_TRANSPOSE ? 1 : op_a, op_b == CORAL_CONJUGATE_TRANSPOSE ? 1 : op_b,
                   m, n, k)] = best;
    ++g_version;
  }
  if (chosen) *chosen = best;
  return CORAL_OK;
}

}  // namespace coral

using namespace coral;

namespace coral {
// The table the library starts with: the file CORAL_AUTOTUNE_TABLE names ("none":
// no table), else autotune_table.txt beside libcoralhip.so when it exists.
int autotune_load_default() {
  const char *env = getenv("CORAL_AUTOTUNE_TABLE");
  std::string path;
  if (env && *env) {
    if (!strcmp(env, "none")) return CORAL_OK;
    path = env;
  } else {
    Dl_info info;
    if (!dladdr((void *)&autotune_load_default, &info) || !info.dli_fname) return CORAL_OK;
    path = info.dli_fname;
    const size_t slash = path.rfind('/');
    path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/autotune_table.txt";
    FILE *probe = fopen(path.c_str(), "r");
    if (!probe) return CORAL_OK;  // no table shipped: heuristics
    fclose(probe);
  }
  return coral_autotune_load(path.c_str());
}
}  // namespace coral

extern "C" {

int coral_autotune_get(const char *routine, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                       int32_t *tile) {
  if (!routine || tile_count(routine) == 0) return -1;
  if (!tile) return -7;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_table.find(key_of(routine, op_a, op_b, m, n, k));
  *tile = it == g_table.end() ? -1 : it->second;
  return CORAL_OK;
}

int coral_autotune_set(const char *routine, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                       int32_t tile) {
  if (!routine || tile_count(routine) == 0) return -1;
  if (tile < 0 || tile >= tile_count(routine)) return -7;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_table[key_of(routine, op_a, op_b, m, n, k)] = tile;
  ++g_version;
  return CORAL_OK;
}

int coral_autotune_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_table.clear();
  ++g_version;
  return CORAL_OK;
}

int coral_autotune_save(const char *path) {
  if (!path) return -1;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  FILE *f = fopen(path, "w");
  if (!f) { set_last_error("cannot write %s", path); return CORAL_ERR_IO; }
  fprintf(f, "# coral autotune v1: routine op_a op_b log2m log2n log2k tile\n");
  for (auto &kv : g_table) fprintf(f, "%s %d\n", kv.first.c_str(), kv.second);
  fclose(f);
  return CORAL_OK;
}

int coral_autotune_load(const char *path) {
  if (!path) return -1;
  FILE *f = fopen(path, "r");
  if (!f) { set_last_error("cannot read %s", path); return CORAL_ERR_IO; }
  char line[256];
  std::map<std::string, int> loaded;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char routine[64];
    int opa, opb, bm, bn, bk, tile;
    if (sscanf(line, "%63s %d %d %d %d %d %d", routine, &opa, &opb, &bm, &bn, &bk, &tile) != 7) {
      fclose(f);
      set_last_error("malformed autotune line: %s", line);
      return CORAL_ERR_IO;
    }
    if (tile_count(routine) == 0 || tile < 0 || tile >= tile_count(routine)) continue;
    char key[160];
    snprintf(key, sizeof(key), "%s %d %d %d %d %d", routine, opa, opb, bm, bn, bk);
    loaded[key] = tile;
  }
  fclose(f);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto &kv : loaded) g_table[kv.first] = kv.second;
  ++g_version;
  return CORAL_OK;
}

// Times every tile of the routine's menu on (m, n, k) with seeded data and stores the
// fastest for the bucket.  Real routines use one float/double per element, complex
// routines two.
int coral_autotune_run(const char *routine, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                       int32_t *chosen_tile) {
  if (!routine) return -1;
  const int ntiles = tile_count(routine);
  if (ntiles == 0) return -1;
  if (!valid_op(op_a)) return -2;
  if (!valid_op(op_b)) return -3;
  if (m <= 0) return -4;
  if (n <= 0) return -5;
  if (k <= 0) return -6;
  CORAL_TRY(ensure_init());
  if (strstr(routine, "batched")) return tune_batched(routine, op_a, op_b, m, n, k, chosen_tile);
  const bool dbl = routine[0] == 'd' || routine[0] == 'z';
  const bool cpx = routine[0] == 'c' || routine[0] == 'z';
  const size_t es = (dbl ? 8 : 4) * (cpx ? 2 : 1);
  const int64_t lda = op_a == CORAL_NO_TRANSPOSE ? m : k;
  const int64_t ldb = op_b == CORAL_NO_TRANSPOSE ? k : n;
  void *a = nullptr, *b = nullptr, *c = nullptr;
  CORAL_HIP(hipMalloc(&a, es * (size_t)m * (size_t)k));
  CORAL_HIP(hipMalloc(&b, es * (size_t)k * (size_t)n));
  CORAL_HIP(hipMalloc(&c, es * (size_t)m * (size_t)n));
  const int64_t mult = cpx ? 2 : 1;
  if (dbl) {
    coral_fill_random_f64((double *)a, mult * m * k, 11, 0, nullptr);
    coral_fill_random_f64((double *)b, mult * k * n, 12, 0, nullptr);
  } else {
    coral_fill_random_f32((float *)a, mult * m * k, 11, 0, nullptr);
    coral_fill_random_f32((float *)b, mult * k * n, 12, 0, nullptr);
  }
  hipEvent_t e0, e1;
  CORAL_HIP(hipEventCreate(&e0));
  CORAL_HIP(hipEventCreate(&e1));
  const float one_f[2] = {1.f, 0.f}, zero_f[2] = {0.f, 0.f};
  const double one_d[2] = {1.0, 0.0}, zero_d[2] = {0.0, 0.0};
  auto run_once = [&]() -> int {
    if (!strcmp(routine, "sgemm"))
      return coral_sgemm(op_a, op_b, m, n, k, 1.f, (float *)a, lda, (float *)b, ldb, 0.f, (float *)c, m, nullptr);
    if (!strcmp(routine, "dgemm"))
      return coral_dgemm(op_a, op_b, m, n, k, 1.0, (double *)a, lda, (double *)b, ldb, 0.0, (double *)c, m, nullptr);
    if (!strcmp(routine, "sgemm_bf16x3"))
      return coral_sgemm_bf16x3(op_a, op_b, m, n, k, 1.f, (float *)a, lda, (float *)b, ldb, 0.f, (float *)c, m,
                                nullptr);
    if (!strcmp(routine, "sgemm_fp16x3"))
      return coral_sgemm_fp16x3(op_a, op_b, m, n, k, 1.f, (float *)a, lda, (float *)b, ldb, 0.f, (float *)c, m,
                                nullptr);
    if (!strcmp(routine, "cgemm_3m_fp16x3"))
      return coral_cgemm_3m_fp16x3(op_a, op_b, m, n, k, one_f, (float *)a, lda, (float *)b, ldb, zero_f, (float *)c,
                                   m, nullptr);
    if (!strcmp(routine, "cgemm"))
      return coral_cgemm(op_a, op_b, m, n, k, one_f, (float *)a, lda, (float *)b, ldb, zero_f, (float *)c, m, nullptr);
    if (!strcmp(routine, "zgemm"))
      return coral_zgemm(op_a, op_b, m, n, k, one_d, (double *)a, lda, (double *)b, ldb, zero_d, (double *)c, m,
                         nullptr);
    return coral_cgemm_3m_bf16x3(op_a, op_b, m, n, k, one_f, (float *)a, lda, (float *)b, ldb, zero_f, (float *)c, m,
                                 nullptr);
  };
  int best = -1;
  float best_ms = 1e30f;
  int status = CORAL_OK;
  for (int t = 0; t < ntiles && status == CORAL_OK; ++t) {
    ForcedTile ft(t);
    status = run_once();  // warm-up (also sizes the workspace)
    if (status != CORAL_OK) break;
    CORAL_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < 3 && status == CORAL_OK; ++r) status = run_once();
    CORAL_HIP(hipEventRecord(e1, nullptr));
    CORAL_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CORAL_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) { best_ms = ms; best = t; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(c);
  if (status != CORAL_OK) return status;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_table[key_of(routine, op_a, op_b, m, n, k)] = best;
    ++g_version;
  }
  if (chosen_tile) *chosen_tile = best;
  return CORAL_OK;
}

}  // extern "C"
