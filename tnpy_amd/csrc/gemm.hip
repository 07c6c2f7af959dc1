// Strided, batched GEMM on CDNA4 matrix cores, the contraction engine of
// every tensor operation in the library (H_eff matvec stages, environment
// updates, QR trailing updates, gauge absorption): host side. The kernel
// template and tile configurations are in gemm_kernel.h; this file holds the
// split-K rule, the choice between typed and generic staging, the autotuner
// and the tile table.
#include "gemm_kernel.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <mutex>
#include <unordered_map>


namespace tnx {

// C = alpha * sum_z partial[bat][z] + beta * C   (fixed summation order)
__global__ void splitk_reduce_kernel(const double* __restrict__ P, int splits, int M, int N,
                                     void* C, int tc, int64_t ldc, int64_t sC, double alpha,
                                     double beta, int nb2, int64_t sC2, const int* bidx) {
  const int bat = blockIdx.y;  // launched index (partials); logical below
  const int lb = bidx ? bidx[bat] : bat;
  const int b1 = lb / nb2, b2 = lb % nb2;
  const int64_t MN = (int64_t)M * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < MN;
       e += (int64_t)gridDim.x * blockDim.x) {
    // in split order; eight loads in flight at a time
    const double* pe = P + (int64_t)bat * splits * MN + e;
    double acc = 0.0;
    int z = 0;
    for (; z + 8 <= splits; z += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = pe[(int64_t)(z + u) * MN];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; z < splits; ++z) acc += pe[(int64_t)z * MN];
    const int64_t m = e / N, n = e % N;
    const int64_t off = (int64_t)b1 * sC + (int64_t)b2 * sC2 + m * ldc + n;
    double val = alpha * acc;
    if (tc == 0) {
      double* c = reinterpret_cast<double*>(C) + off;
      if (beta != 0.0) val += beta * *c;
      *c = val;
    } else {
      float* c = reinterpret_cast<float*>(C) + off;
      if (beta != 0.0) val += beta * (double)*c;
      *c = (float)val;
    }
  }
}

// ---------------------------------------------------------------------------
// Tile configurations, split-K rule and the runtime autotuner.
//
// A config only changes how output tiles are cut and scheduled (results are
// bitwise identical whichever config runs). Split-K does change the rounding
// and is therefore decided by a fixed rule on the shape, never by timing.
// ---------------------------------------------------------------------------
template <int PREC>
static void launch_index(int ci, const GemmDesc& d, const GemmArgs& g, int splits, bool typed,
                         hipStream_t s) {
  if (PREC == 0) gemm_launch_f64(ci, d, g, splits, typed, s);
  else if (PREC == 1) gemm_launch_f32(ci, d, g, splits, typed, s);
  else gemm_launch_bf16(ci, d, g, splits, typed, s);
}

// One operand of the typed staging: contiguous extent `ext` (elements),
// leading dimension, batch strides and base all in whole 16-byte vectors,
// one batch entry (rows x ld, contiguous extent on the last row) < 2 GiB.
static bool typed_operand_ok(const void* p, DT t, int64_t ld, int64_t s1, int64_t s2, int nb2,
                             int64_t ext, int64_t rows, int bk_rows) {
  const int64_t es = t == DT::F64 ? 8 : 4, v = 16 / es;
  if (((uintptr_t)p & 15) != 0 || ext % v != 0 || ld % v != 0) return false;
  if ((s1 * es) % 16 != 0 || (nb2 > 1 && (s2 * es) % 16 != 0)) return false;
  const int64_t span = ((rows - 1) * ld + ext) * es;
  const int64_t lim = ((int64_t)1 << 31) - 64;
  return span < lim && (int64_t)bk_rows * ld * es < lim;
}

bool gemm_typed_ok(const GemmDesc& d, int kc) {
  const bool f64f64 = d.ta == DT::F64 && d.tb == DT::F64;
  const bool f64f32 = d.ta == DT::F64 && d.tb == DT::F32;
  const bool f32f64 = d.ta == DT::F32 && d.tb == DT::F64;
  const bool f32f32 = d.ta == DT::F32 && d.tb == DT::F32;
  if (!(f64f64 || f64f32 || (d.prec != Prec::F64 && (f32f64 || f32f32)))) return false;
  const int nb2 = d.nb2 > 0 ? d.nb2 : 1;
  // the K range of a split is a whole number of vectors when K-contiguous
  if (d.a_kcontig && kc % (d.ta == DT::F64 ? 2 : 4) != 0) return false;
  if (!d.b_ncontig && kc % (d.tb == DT::F64 ? 2 : 4) != 0) return false;
  const bool a_ok = d.a_kcontig
      ? typed_operand_ok(d.A, d.ta, d.lda, d.sA, d.sA2, nb2, d.K, d.M, 32)
      : typed_operand_ok(d.A, d.ta, d.lda, d.sA, d.sA2, nb2, d.M, d.K, 32);
  const bool b_ok = d.b_ncontig
      ? typed_operand_ok(d.B, d.tb, d.ldb, d.sB, d.sB2, nb2, d.N, d.K, 32)
      : typed_operand_ok(d.B, d.tb, d.ldb, d.sB, d.sB2, nb2, d.K, d.N, 32);
  return a_ok && b_ok;
}

// Deterministic split-K rule: split only when the output is small (fewer
// than ~512 tiles of 128x128), each split keeps at least 64 of K (two k
// steps of the 64x64 tile), and the scratch fits. Skinny outputs with a
// long K (the 32-row V^T A products of the blocked QR) otherwise run on a
// few workgroups, each looping over all of K.
static int split_rule(const GemmDesc& d) {
  if (!d.splitk_ws) return 1;
  const int64_t elems = (int64_t)d.M * d.N * d.batch;
  // about two 256x128 output tiles per CU after splitting
  const int64_t target = (int64_t)2 * 256 * 128 * 128;
  if (elems >= target) return 1;
  int sp = 1;
  while (sp < 32 && elems * sp * 2 <= target && (int64_t)d.K / (sp * 2) >= 64) sp *= 2;
  while (sp > 1 && (size_t)sp * elems * sizeof(double) > d.splitk_ws_bytes) sp /= 2;
  return sp;
}

// Heuristic when not autotuning: fewest full-chip waves, weighted by the
// per-CU efficiency of the tile size (larger tiles reuse operands more).
template <int PREC>
static int heuristic_cfg(const GemmDesc& d, int splits) {
  int best = 0;
  double best_t = 1e300;
  for (int i = 0; i < Cfgs<PREC>::n; ++i) {
    const Cfg& c = Cfgs<PREC>::c[i];
    const int64_t blocks = (int64_t)ceil_div(d.M, c.bm) * ceil_div(d.N, c.bn) * d.batch * splits;
    const int per_cu = (c.wgm * c.wgn <= 4) ? 2 : 1;
    const double waves = std::ceil((double)blocks / (256.0 * per_cu));
    const double eff = (c.bm * c.bn >= 256 * 128) ? 0.8 : (c.bm * c.bn >= 128 * 128 ? 0.6 : 0.45);
    const double t = waves * (double)c.bm * c.bn / (per_cu * eff);
    if (t < best_t - 1e-9) { best_t = t; best = i; }
  }
  return best;
}

struct TuneKey {
  int prec, M, N, K, batch, ak, bn, ta, tb, tc, splits;
  bool operator==(const TuneKey& o) const {
    return prec == o.prec && M == o.M && N == o.N && K == o.K && batch == o.batch && ak == o.ak &&
           bn == o.bn && ta == o.ta && tb == o.tb && tc == o.tc && splits == o.splits;
  }
};
struct TuneHash {
  size_t operator()(const TuneKey& k) const {
    size_t h = 1469598103934665603ull;
    for (int v : {k.prec, k.M, k.N, k.K, k.batch, k.ak, k.bn, k.ta, k.tb, k.tc, k.splits})
      h = (h ^ (size_t)(unsigned)v) * 1099511628211ull;
    return h;
  }
};

static std::mutex g_tune_mu;
static std::unordered_map<TuneKey, int, TuneHash> g_tune;
// Inline tuning mode: 0 (default) takes the table entry or the heuristic and
// never allocates or synchronises; 1 times the configurations of a shape the
// table lacks (allocates a scratch output and waits on the stream).
static std::atomic<int> g_autotune{0};
static std::once_flag g_env_once;

static int tune_load_locked(const char* path);

// TNX_AUTOTUNE=1 turns the inline tuning on; TNX_TUNE_CACHE=<file> loads a
// tile table. Read once, before the first GEMM or tuning call.
static void init_from_env() {
  std::call_once(g_env_once, [] {
    const char* e = getenv("TNX_AUTOTUNE");
    if (e && (e[0] == '1' || e[0] == '2')) g_autotune.store(e[0] - '0');
    const char* cache = getenv("TNX_TUNE_CACHE");
    if (cache && cache[0]) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      tune_load_locked(cache);
    }
  });
}

static thread_local int g_no_tune_depth = 0;

NoTuneScope::NoTuneScope() { ++g_no_tune_depth; }
NoTuneScope::~NoTuneScope() { --g_no_tune_depth; }

static bool autotune_on() {
  init_from_env();
  return g_no_tune_depth == 0 && g_autotune.load() >= 1;
}

// Tile table file: a comment line, then one line per GEMM shape key,
// "prec M N K batch a_kcontig b_ncontig ta tb tc splits config".
static const char* kTuneHeader = "# tnx gemm tile table v1: prec M N K batch ak bn ta tb tc splits cfg";

static int tune_load_locked(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[256];
  int n = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    TuneKey k;
    int cfg;
    if (sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d", &k.prec, &k.M, &k.N, &k.K, &k.batch,
               &k.ak, &k.bn, &k.ta, &k.tb, &k.tc, &k.splits, &cfg) != 12)
      continue;
    const int ncfg = k.prec == 0 ? Cfgs<0>::n : (k.prec == 1 ? Cfgs<1>::n : Cfgs<2>::n);
    if (k.prec < 0 || k.prec > 2 || cfg < 0 || cfg >= ncfg) continue;
    g_tune[k] = cfg;
    ++n;
  }
  fclose(f);
  return n;
}

template <int PREC>
static int run_prec(const GemmDesc& d, hipStream_t s) {
  GemmArgs g;
  g.M = d.M; g.N = d.N; g.K = d.K;
  g.A = d.A; g.lda = d.lda; g.sA = d.sA; g.ta = (int)d.ta;
  g.B = d.B; g.ldb = d.ldb; g.sB = d.sB; g.tb = (int)d.tb;
  g.C = d.C; g.ldc = d.ldc; g.sC = d.sC; g.tc = (int)d.tc;
  g.alpha = d.alpha; g.beta = d.beta;
  g.tiles_n = 1;
  g.group_m = 0;
  g.nb2 = d.nb2 > 0 ? d.nb2 : 1;
  g.sA2 = d.sA2; g.sB2 = d.sB2; g.sC2 = d.sC2;
  g.bidx = d.bidx;
  const int splits = split_rule(d);
  g.kc = d.K;
  g.partial = nullptr;
  if (splits > 1) {
    g.kc = ((d.K + splits - 1) / splits + 31) / 32 * 32;
    g.partial = d.splitk_ws;
  }
  const int nsplit = splits > 1 ? ceil_div(d.K, g.kc) : 1;
  const bool typed = gemm_typed_ok(d, g.kc);
  const TuneKey key{PREC, d.M, d.N, d.K, d.batch, d.a_kcontig, d.b_ncontig, (int)d.ta, (int)d.tb,
                    (int)d.tc, nsplit};
  int ci = -1;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) ci = it->second;
  }
  if (ci < 0) {
    ci = heuristic_cfg<PREC>(d, nsplit);
    const int64_t work = (int64_t)d.M * d.N * d.K * d.batch;
    // mode 1 tunes shapes of at least 2^28 multiply-adds, mode 2 every shape
    const int64_t min_work = g_autotune.load() >= 2 ? 0 : ((int64_t)1 << 28);
    if (autotune_on() && work >= min_work && !d.bidx) {
      // Time every config once on a scratch output (the caller's C is not
      // touched), keep the fastest. Runs once per shape key.
      // Split-K launches only write the partial-sum scratch (which the real
      // launch below overwrites), so they are timed in place; others write a
      // scratch output.
      GemmArgs gt = g;
      void* scratch = nullptr;
      const int nb2 = g.nb2;
      const size_t cbytes = ((size_t)((d.batch - 1) / nb2) * d.sC + (size_t)((d.batch - 1) % nb2) * d.sC2 +
                             (size_t)(d.M - 1) * d.ldc + d.N) * (d.tc == DT::F64 ? 8 : 4);
      if (g.partial || hipMalloc(&scratch, cbytes) == hipSuccess) {
        if (scratch) gt.C = scratch;
        gt.beta = 0.0;
        // A failure anywhere in the timing keeps the heuristic's choice.
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
          float best = 1e30f;
          int pick = ci;
          bool ok = true;
          for (int i = 0; ok && i < Cfgs<PREC>::n; ++i) {
            launch_index<PREC>(i, d, gt, nsplit, typed, s);  // warm
            float ms = 0.f;
            ok = hipEventRecord(e0, s) == hipSuccess;
            launch_index<PREC>(i, d, gt, nsplit, typed, s);
            launch_index<PREC>(i, d, gt, nsplit, typed, s);
            ok = ok && hipEventRecord(e1, s) == hipSuccess &&
                 hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (ok && ms < best) { best = ms; pick = i; }
          }
          if (ok) ci = pick;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamSynchronize(s);
        if (scratch) (void)hipFree(scratch);
        (void)hipGetLastError();
      }
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = ci;
  }
  launch_index<PREC>(ci, d, g, nsplit, typed, s);
  TNX_LAUNCH_CHECK();
  if (g.partial) {
    const int64_t MN = (int64_t)d.M * d.N;
    const int rb = (int)std::min<int64_t>((MN + 255) / 256, 2048);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(rb, d.batch), dim3(256), 0, s, g.partial,
                       nsplit, d.M, d.N, d.C, (int)d.tc, d.ldc, d.sC, d.alpha, d.beta,
                       g.nb2, g.sC2, g.bidx);
    TNX_LAUNCH_CHECK();
  }
  return TNX_OK;
}

int gemm(const GemmDesc& d, hipStream_t s) {
  TNX_CHECK_ARG(d.M >= 0 && d.N >= 0 && d.K >= 0 && d.batch >= 0, "gemm: negative size");
  if (d.M == 0 || d.N == 0 || d.batch == 0) return TNX_OK;
  TNX_CHECK_ARG(d.A && d.B && d.C, "gemm: null operand");
  TNX_CHECK_ARG(d.batch <= 65535, "gemm: batch %d exceeds 65535", d.batch);
  switch (d.prec) {
    case Prec::F64: return run_prec<0>(d, s);
    case Prec::F32: return run_prec<1>(d, s);
    case Prec::BF16X3: return run_prec<2>(d, s);
  }
  set_error("gemm: unknown precision");
  return TNX_ERR_INVALID;
}

void set_autotune(int on) {
  init_from_env();
  g_autotune.store(on <= 0 ? 0 : (on >= 2 ? 2 : 1));
}

void tune_clear() {
  init_from_env();
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
}

int tune_load(const char* path) {
  init_from_env();  // resolve the environment first so it cannot override later
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const int n = tune_load_locked(path);
  if (n < 0) set_error("tnx_tune_load: cannot read %s", path);
  return n < 0 ? TNX_ERR_INVALID : n;
}

int tune_save(const char* path) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  FILE* f = fopen(path, "w");
  if (!f) {
    set_error("tnx_tune_save: cannot write %s", path);
    return TNX_ERR_INVALID;
  }
  fprintf(f, "%s\n", kTuneHeader);
  int n = 0;
  for (const auto& kv : g_tune) {
    const TuneKey& k = kv.first;
    fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d %d\n", k.prec, k.M, k.N, k.K, k.batch, k.ak, k.bn,
            k.ta, k.tb, k.tc, k.splits, kv.second);
    ++n;
  }
  fclose(f);
  return n;
}

}  // namespace tnx

extern "C" int tnx_set_autotune(int on) {
  tnx::set_autotune(on);
  return TNX_OK;
}

extern "C" int tnx_tune_clear(void) {
  tnx::tune_clear();
  return TNX_OK;
}

extern "C" int tnx_tune_load(const char* path) {
  if (!path) {
    tnx::set_error("tnx_tune_load: null path");
    return TNX_ERR_INVALID;
  }
  return tnx::tune_load(path);
}

extern "C" int tnx_tune_save(const char* path) {
  if (!path) {
    tnx::set_error("tnx_tune_save: null path");
    return TNX_ERR_INVALID;
  }
  return tnx::tune_save(path);
}

extern "C" int tnx_gemm_f64(int M, int N, int K, int batch, double alpha,
                            const double* A, int64_t lda, int64_t strideA, int a_kcontig,
                            const double* B, int64_t ldb, int64_t strideB, int b_ncontig,
                            double beta, double* C, int64_t ldc, int64_t strideC,
                            int precision, tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(precision >= TNX_F64 && precision <= TNX_BF16X3, "tnx_gemm_f64: bad precision %d", precision);
  GemmDesc d;
  d.M = M; d.N = N; d.K = K; d.batch = batch;
  d.A = A; d.ta = DT::F64; d.lda = lda; d.sA = strideA; d.a_kcontig = a_kcontig != 0;
  d.B = B; d.tb = DT::F64; d.ldb = ldb; d.sB = strideB; d.b_ncontig = b_ncontig != 0;
  d.C = C; d.tc = DT::F64; d.ldc = ldc; d.sC = strideC;
  d.alpha = alpha; d.beta = beta; d.prec = (Prec)precision;
  return gemm(d, (hipStream_t)stream);
}

extern "C" int tnx_gemm(int M, int N, int K, int batch, double alpha, const void* A, int a_f32,
                        int64_t lda, int64_t strideA, int a_kcontig, const void* B, int b_f32,
                        int64_t ldb, int64_t strideB, int b_ncontig, double beta, void* C,
                        int c_f32, int64_t ldc, int64_t strideC, int precision,
                        tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(precision >= TNX_F64 && precision <= TNX_BF16X3, "tnx_gemm: bad precision %d", precision);
  GemmDesc d;
  d.M = M; d.N = N; d.K = K; d.batch = batch;
  d.A = A; d.ta = a_f32 ? DT::F32 : DT::F64; d.lda = lda; d.sA = strideA; d.a_kcontig = a_kcontig != 0;
  d.B = B; d.tb = b_f32 ? DT::F32 : DT::F64; d.ldb = ldb; d.sB = strideB; d.b_ncontig = b_ncontig != 0;
  d.C = C; d.tc = c_f32 ? DT::F32 : DT::F64; d.ldc = ldc; d.sC = strideC;
  d.alpha = alpha; d.beta = beta; d.prec = (Prec)precision;
  return gemm(d, (hipStream_t)stream);
}
