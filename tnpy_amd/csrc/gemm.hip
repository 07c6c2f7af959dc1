// Strided, batched GEMM on CDNA4 matrix cores, the contraction engine of
// every tensor operation in the library (H_eff matvec stages, environment
// updates, QR trailing updates, gauge absorption).
//
// One kernel template, three arithmetics:
//   F64    : v_mfma_f64_16x16x4_f64, FP64 LDS tiles [BK][BM]
//   F32    : v_mfma_f32_16x16x4_f32, FP32 LDS tiles [BK][BM]
//   BF16X3 : v_mfma_f32_16x16x32_bf16 three times per tile pair,
//            a*b ~ ahi*bhi + ahi*blo + alo*bhi with a = ahi + alo (bf16),
//            FP32 accumulate, LDS tiles [BM][BK] so each lane reads its
//            8 consecutive k values with one 16-byte LDS read.
// Tile configurations (block tile BM x BN, WGM x WGN wavefronts of 64 lanes,
// K-tile depth BK) are listed in Cfgs below; the autotuner picks one per
// shape. Each wave owns a (BM/WGM) x (BN/WGN) sub-tile of 16x16 MFMA
// accumulators. Operands may be FP64 or FP32 in HBM (runtime, wave-uniform
// branch); they are converted while staging the tile in LDS. Global loads of
// tile t+1 are issued before the MFMAs of tile t and their results are first
// used after them; LDS is double-buffered (one barrier per k-tile).
#include "tnx_common.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <unordered_map>


namespace tnx {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct GemmArgs {
  int M, N, K;
  const void* A; int64_t lda, sA; int ta;
  const void* B; int64_t ldb, sB; int tb;
  void* C; int64_t ldc, sC; int tc;
  double alpha, beta;
  int tiles_n;  // number of tiles along N
  int kc;       // K extent of one split (== K when not split)
  double* partial;  // split-K partial sums [batch][split][M][N], or null
  int nb2;          // second batch level (1 = none)
  int64_t sA2, sB2, sC2;
  const int* bidx;  // launched -> logical batch index, or null
};

template <int PREC> struct PT;
template <> struct PT<0> { using lds_t = double; using acc_t = f64x4; };
template <> struct PT<1> { using lds_t = float;  using acc_t = f32x4; };
template <> struct PT<2> { using lds_t = __bf16; using acc_t = f32x4; };

// WGM x WGN wavefronts per block; each wave owns a (BM/WGM) x (BN/WGN) tile.
template <int PREC, int BM, int BN, bool A_K, bool B_N, int WGM, int WGN, int BK>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_kernel(GemmArgs g) {
  using lds_t = typename PT<PREC>::lds_t;
  using acc_t = typename PT<PREC>::acc_t;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int EA = BM * BK / NT;
  constexpr int EB = BK * BN / NT;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(EA * NT == BM * BK && EB * NT == BK * BN, "tile not divisible by the block");
  // LDS geometry. F64/F32: [BK][BM + PADM]. BF16X3: [BM][BK + PADK] for
  // hi and lo parts (the k index is innermost so a lane's 8 values are
  // one 16-byte read).
  constexpr int PADM = (PREC == 0) ? 2 : 4;
  constexpr int PADK = 8;
  constexpr int A_ELEMS = (PREC == 2) ? 2 * BM * (BK + PADK) : BK * (BM + PADM);
  constexpr int B_ELEMS = (PREC == 2) ? 2 * BN * (BK + PADK) : BK * (BN + PADM);
  // 16-byte alignment: the BF16X3 path reads 8 bf16 values per lane at once
  __shared__ __attribute__((aligned(16))) lds_t As[2][A_ELEMS];
  __shared__ __attribute__((aligned(16))) lds_t Bs[2][B_ELEMS];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM;
  const int wn0 = (wave % WGN) * WN;

  const int bat = g.bidx ? g.bidx[blockIdx.y] : (int)blockIdx.y;
  const int tile = blockIdx.x;
  const int split = blockIdx.z;
  const int kbeg = split * g.kc;
  const int kend = min(g.K, kbeg + g.kc);
  const int m0 = (tile / g.tiles_n) * BM;
  const int n0 = (tile % g.tiles_n) * BN;

  const int b1 = bat / g.nb2, b2 = bat % g.nb2;
  const char* Ab = (const char*)g.A + ((size_t)b1 * g.sA + (size_t)b2 * g.sA2) * (g.ta == 0 ? 8 : 4);
  const char* Bb = (const char*)g.B + ((size_t)b1 * g.sB + (size_t)b2 * g.sB2) * (g.tb == 0 ? 8 : 4);

  // Staging registers hold the raw loaded bits (a double, or a float in the
  // low word) so that no instruction uses a global load's result until
  // store_tiles, after the MFMAs of the current k-block: the loads of
  // block k+1 stay in flight under the compute of block k.
  uint64_t ua[EA], ub[EB];

  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + e * NT;
      int mi, ki;
      if (A_K) { ki = idx % BK; mi = idx / BK; } else { mi = idx % BM; ki = idx / BM; }
      const int m = m0 + mi, k = k0 + ki;
      uint64_t v = 0;
      if (m < g.M && k < kend) {
        const int64_t off = A_K ? (int64_t)m * g.lda + k : (int64_t)k * g.lda + m;
        if (g.ta == 0) v = reinterpret_cast<const uint64_t*>(Ab)[off];
        else v = reinterpret_cast<const uint32_t*>(Ab)[off];
      }
      ua[e] = v;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + e * NT;
      int ni, ki;
      if (B_N) { ni = idx % BN; ki = idx / BN; } else { ki = idx % BK; ni = idx / BK; }
      const int n = n0 + ni, k = k0 + ki;
      uint64_t v = 0;
      if (n < g.N && k < kend) {
        const int64_t off = B_N ? (int64_t)k * g.ldb + n : (int64_t)n * g.ldb + k;
        if (g.tb == 0) v = reinterpret_cast<const uint64_t*>(Bb)[off];
        else v = reinterpret_cast<const uint32_t*>(Bb)[off];
      }
      ub[e] = v;
    }
  };
  auto as_double = [](uint64_t v, int t) -> double {
    return t == 0 ? __longlong_as_double((long long)v) : (double)__uint_as_float((uint32_t)v);
  };
  auto as_float = [](uint64_t v, int t) -> float {
    return t == 0 ? (float)__longlong_as_double((long long)v) : __uint_as_float((uint32_t)v);
  };

  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + e * NT;
      int mi, ki;
      if (A_K) { ki = idx % BK; mi = idx / BK; } else { mi = idx % BM; ki = idx / BM; }
      if constexpr (PREC == 0) {
        As[buf][ki * (BM + PADM) + mi] = as_double(ua[e], g.ta);
      } else if constexpr (PREC == 1) {
        As[buf][ki * (BM + PADM) + mi] = as_float(ua[e], g.ta);
      } else {
        const float x = as_float(ua[e], g.ta);
        const __bf16 hi = (__bf16)x;
        const __bf16 lo = (__bf16)(x - (float)hi);
        As[buf][mi * (BK + PADK) + ki] = hi;
        As[buf][BM * (BK + PADK) + mi * (BK + PADK) + ki] = lo;
      }
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + e * NT;
      int ni, ki;
      if (B_N) { ni = idx % BN; ki = idx / BN; } else { ki = idx % BK; ni = idx / BK; }
      if constexpr (PREC == 0) {
        Bs[buf][ki * (BN + PADM) + ni] = as_double(ub[e], g.tb);
      } else if constexpr (PREC == 1) {
        Bs[buf][ki * (BN + PADM) + ni] = as_float(ub[e], g.tb);
      } else {
        const float x = as_float(ub[e], g.tb);
        const __bf16 hi = (__bf16)x;
        const __bf16 lo = (__bf16)(x - (float)hi);
        Bs[buf][ni * (BK + PADK) + ki] = hi;
        Bs[buf][BN * (BK + PADK) + ni * (BK + PADK) + ki] = lo;
      }
    }
  };

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int nk = (kend - kbeg + BK - 1) / BK;
  load_tiles(kbeg);
  store_tiles(0);
  __syncthreads();

  const int lr = lane & 15;   // row (A) / column (B) inside a 16x16 tile
  const int lk = lane >> 4;   // k group of the lane

  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load_tiles(kbeg + (kt + 1) * BK);
    const lds_t* as = As[buf];
    const lds_t* bs = Bs[buf];
    if constexpr (PREC == 0 || PREC == 1) {
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        lds_t a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = as[(kk + lk) * (BM + PADM) + wm0 + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = bs[(kk + lk) * (BN + PADM) + wn0 + j * 16 + lr];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if constexpr (PREC == 0)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
          }
      }
    } else {
      bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
      constexpr int LO_A = BM * (BK + PADK);
      constexpr int LO_B = BN * (BK + PADK);
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = wm0 + i * 16 + lr;
        ah[i] = *reinterpret_cast<const bf16x8*>(&as[r * (BK + PADK) + lk * 8]);
        al[i] = *reinterpret_cast<const bf16x8*>(&as[LO_A + r * (BK + PADK) + lk * 8]);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = wn0 + j * 16 + lr;
        bh[j] = *reinterpret_cast<const bf16x8*>(&bs[c * (BK + PADK) + lk * 8]);
        bl[j] = *reinterpret_cast<const bf16x8*>(&bs[LO_B + c * (BK + PADK) + lk * 8]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nk) store_tiles(buf ^ 1);
    __syncthreads();
  }

  // Epilogue. 16x16 accumulator tile, lane group lk = lane/16, item v:
  //   f32 MFMAs: row = 4*lk + v ;  f64 MFMA: row = lk + 4*v ; column = lane%16
  auto row_of = [&](int v) { return (PREC == 0) ? (lk + 4 * v) : (lk * 4 + v); };
  if (g.partial) {
    double* P = g.partial + ((size_t)blockIdx.y * gridDim.z + split) * (size_t)g.M * g.N;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + wm0 + i * 16 + row_of(v);
          const int n = n0 + wn0 + j * 16 + lr;
          if (m < g.M && n < g.N) P[(size_t)m * g.N + n] = (double)acc[i][j][v];
        }
    return;
  }
  char* Cb = (char*)g.C + ((size_t)b1 * g.sC + (size_t)b2 * g.sC2) * (g.tc == 0 ? 8 : 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm0 + i * 16 + row_of(v);
        const int n = n0 + wn0 + j * 16 + lr;
        if (m < g.M && n < g.N) {
          const int64_t off = (int64_t)m * g.ldc + n;
          double val = g.alpha * (double)acc[i][j][v];
          if (g.tc == 0) {
            double* c = reinterpret_cast<double*>(Cb) + off;
            if (g.beta != 0.0) val += g.beta * *c;
            *c = val;
          } else {
            float* c = reinterpret_cast<float*>(Cb) + off;
            if (g.beta != 0.0) val += g.beta * (double)*c;
            *c = (float)val;
          }
        }
      }
}

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
    double acc = 0.0;
    for (int z = 0; z < splits; ++z) acc += P[((int64_t)bat * splits + z) * MN + e];
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
// A config only changes how output tiles are cut and scheduled: every output
// element still accumulates its K range in order, so results are bitwise
// identical whichever config runs. Split-K does change the rounding and is
// therefore decided by a fixed rule on the shape, never by timing.
// ---------------------------------------------------------------------------
struct Cfg { int bm, bn, wgm, wgn, bk; };

template <int PREC> struct Cfgs;
template <> struct Cfgs<0> {
  static constexpr int n = 6;
  static constexpr Cfg c[n] = {{64, 64, 2, 2, 32}, {128, 128, 2, 2, 32}, {256, 128, 4, 2, 16},
                               {128, 256, 2, 4, 16}, {256, 160, 4, 2, 16}, {160, 256, 2, 4, 16}};
};
template <> struct Cfgs<1> {
  static constexpr int n = 6;
  static constexpr Cfg c[n] = {{64, 64, 2, 2, 32}, {128, 128, 2, 2, 32}, {256, 128, 4, 2, 32},
                               {128, 256, 2, 4, 32}, {256, 160, 4, 2, 32}, {160, 256, 2, 4, 32}};
};
template <> struct Cfgs<2> {
  static constexpr int n = 6;
  static constexpr Cfg c[n] = {{64, 64, 2, 2, 32}, {128, 128, 2, 2, 32}, {256, 128, 4, 2, 32},
                               {128, 256, 2, 4, 32}, {256, 160, 4, 2, 32}, {160, 256, 2, 4, 32}};
};
constexpr Cfg Cfgs<0>::c[];
constexpr Cfg Cfgs<1>::c[];
constexpr Cfg Cfgs<2>::c[];

template <int PREC, int BM, int BN, int WGM, int WGN, int BK>
static void launch_cfg(const GemmDesc& d, const GemmArgs& g0, int splits, hipStream_t s) {
  GemmArgs g = g0;
  g.tiles_n = ceil_div(d.N, BN);
  const int tiles = ceil_div(d.M, BM) * g.tiles_n;
  dim3 grid(tiles, d.batch, splits);
  dim3 block(64 * WGM * WGN);
  if (d.a_kcontig && d.b_ncontig)
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, true, true, WGM, WGN, BK>), grid, block, 0, s, g);
  else if (d.a_kcontig && !d.b_ncontig)
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, true, false, WGM, WGN, BK>), grid, block, 0, s, g);
  else if (!d.a_kcontig && d.b_ncontig)
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, false, true, WGM, WGN, BK>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, false, false, WGM, WGN, BK>), grid, block, 0, s, g);
}

template <int PREC>
static void launch_index(int ci, const GemmDesc& d, const GemmArgs& g, int splits, hipStream_t s) {
#define TNX_CFG(I)                                                                              \
  if (ci == I && I < Cfgs<PREC>::n) {                                                           \
    constexpr Cfg c = Cfgs<PREC>::c[I < Cfgs<PREC>::n ? I : 0];                                 \
    launch_cfg<PREC, c.bm, c.bn, c.wgm, c.wgn, c.bk>(d, g, splits, s);                          \
    return;                                                                                     \
  }
  TNX_CFG(0) TNX_CFG(1) TNX_CFG(2) TNX_CFG(3) TNX_CFG(4) TNX_CFG(5) TNX_CFG(6)
#undef TNX_CFG
}

// Deterministic split-K rule: split only when the output is small (fewer
// than ~512 tiles of 128x128) and K is long, and the scratch fits.
static int split_rule(const GemmDesc& d) {
  if (!d.splitk_ws) return 1;
  const int64_t elems = (int64_t)d.M * d.N * d.batch;
  // about two 256x128 output tiles per CU after splitting
  const int64_t target = (int64_t)2 * 256 * 128 * 128;
  if (elems >= target || d.K < 1024) return 1;
  int sp = 1;
  while (sp < 32 && elems * sp * 2 <= target && (int64_t)d.K / (sp * 2) >= 256) sp *= 2;
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
static int g_autotune = -1;  // -1: from TNX_AUTOTUNE (default on)

static int tune_load_locked(const char* path);

static bool autotune_on() {
  if (g_autotune < 0) {
    const char* e = getenv("TNX_AUTOTUNE");
    g_autotune = (e && e[0] == '0') ? 0 : 1;
    // a saved tile table named by TNX_TUNE_CACHE is loaded at first use
    const char* cache = getenv("TNX_TUNE_CACHE");
    if (cache && cache[0]) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      tune_load_locked(cache);
    }
  }
  return g_autotune == 1;
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
    if (autotune_on() && work >= ((int64_t)1 << 28) && !d.bidx) {
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
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        float best = 1e30f;
        for (int i = 0; i < Cfgs<PREC>::n; ++i) {
          launch_index<PREC>(i, d, gt, nsplit, s);  // warm
          hipEventRecord(e0, s);
          launch_index<PREC>(i, d, gt, nsplit, s);
          launch_index<PREC>(i, d, gt, nsplit, s);
          hipEventRecord(e1, s);
          hipEventSynchronize(e1);
          float ms = 0.f;
          hipEventElapsedTime(&ms, e0, e1);
          if (ms < best) { best = ms; ci = i; }
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        hipStreamSynchronize(s);
        if (scratch) hipFree(scratch);
        hipGetLastError();
      }
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = ci;
  }
  launch_index<PREC>(ci, d, g, nsplit, s);
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
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_autotune = on ? 1 : 0;
  if (!on) g_tune.clear();
}

int tune_load(const char* path) {
  autotune_on();  // resolve the environment first so it cannot override later
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
