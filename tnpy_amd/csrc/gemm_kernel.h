// Strided, batched GEMM on CDNA4 matrix cores: the kernel template, its tile
// configurations and the per-configuration launchers. The host side (split-K
// rule, autotuner, tile table) is in gemm.hip; each arithmetic's kernels are
// instantiated in a file of their own (gemm_f64.hip, gemm_f32.hip,
// gemm_bf16.hip) so the three compile in parallel.
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
// accumulators. Global loads of tile t+1 are issued before the MFMAs of tile
// t and their results are first used after them; LDS is double-buffered (one
// barrier per k-tile).
//
// Two ways to stage a tile from HBM into LDS, chosen per launch:
//   * typed (TA, TB = double or float, fixed at compile time): 16-byte
//     buffer loads along each operand's contiguous dimension. Every vector's
//     byte offset is computed once, before the k loop; the k step is a
//     scalar offset. Vectors outside the matrix get an offset beyond the
//     buffer's range, which the hardware turns into zeros, so the loop has
//     no bounds branches. Needs the contiguous extent, the leading dimension
//     and the base addresses in whole 16-byte vectors, and one batch entry
//     of each operand below 2 GiB.
//   * generic (TA = TB = void): element loads, operand types read at run time
//     and bounds-checked per element. Every other shape.
// Both stage the same values in the same LDS places, so the two give bitwise
// identical results.
#pragma once

#include "tnx_common.h"

#include <cmath>
#include <type_traits>

namespace tnx {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
  int M, N, K;
  const void* A; int64_t lda, sA; int ta;
  const void* B; int64_t ldb, sB; int tb;
  void* C; int64_t ldc, sC; int tc;
  double alpha, beta;
  int tiles_n;  // number of tiles along N
  int group_m;  // tile rows per group of the XCD-aware tile order (0: row-major order)
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

// Offset with the top bit set: beyond the 0x7fffffff-byte range of every
// buffer resource below, so a load from it returns zero without an access.
constexpr unsigned kGemmOob = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t gemm_rsrc(const void* base) {
  // raw buffer (stride 0), range 0x7fffffff bytes, gfx9 data format word
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff,
                                           0x00020000);
}

// element j of a 16-byte vector holding doubles or floats
template <typename T>
__device__ __forceinline__ T vec_elem(const u32x4& v, int j) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long lo = v[2 * j], hi = v[2 * j + 1];
    return __longlong_as_double((long long)((hi << 32) | lo));
  } else {
    return __uint_as_float(v[j]);
  }
}

// WGM x WGN wavefronts per block; each wave owns a (BM/WGM) x (BN/WGN) tile.
template <int PREC, int BM, int BN, bool A_K, bool B_N, int WGM, int WGN, int BK, typename TA,
          typename TB>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_kernel(GemmArgs g) {
  using lds_t = typename PT<PREC>::lds_t;
  using acc_t = typename PT<PREC>::acc_t;
  constexpr bool TYPED = !std::is_void<TA>::value;
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
  // 16-byte alignment: vector LDS stores and the BF16X3 16-byte reads
  __shared__ __attribute__((aligned(16))) lds_t As[2][A_ELEMS];
  __shared__ __attribute__((aligned(16))) lds_t Bs[2][B_ELEMS];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / WGN) * WM;
  const int wn0 = (wave % WGN) * WN;

  const int bat = g.bidx ? g.bidx[blockIdx.y] : (int)blockIdx.y;
  int tile = blockIdx.x;
  if (g.group_m > 0) {
    // XCD-aware tile order. Consecutive workgroups are dispatched to the 8
    // XCDs in turn, each XCD with its own L2: workgroup x takes place x / 8
    // of a contiguous range of tile ids kept for XCD x % 8. Tile ids run
    // through groups of group_m tile rows, column by column inside a group,
    // so an XCD's range is a compact block of tiles that share A row panels
    // and B column panels in its L2. A bijection on [0, gridDim.x) for any
    // grid size; each tile is computed exactly as in row-major order.
    const int T = gridDim.x, x = blockIdx.x & 7, per = T >> 3, rem = T & 7;
    const int lin = x * per + min(x, rem) + (blockIdx.x >> 3);
    const int gw = g.group_m * g.tiles_n;
    const int grp = lin / gw;
    const int fm = grp * g.group_m;
    const int gs = min(T / g.tiles_n - fm, g.group_m);
    const int w = lin - grp * gw;
    tile = (fm + w % gs) * g.tiles_n + w / gs;
  }
  const int split = blockIdx.z;
  const int kbeg = split * g.kc;
  const int kend = min(g.K, kbeg + g.kc);
  const int m0 = (tile / g.tiles_n) * BM;
  const int n0 = (tile % g.tiles_n) * BN;

  const int b1 = bat / g.nb2, b2 = bat % g.nb2;
  const char* Ab = (const char*)g.A + ((size_t)b1 * g.sA + (size_t)b2 * g.sA2) * (g.ta == 0 ? 8 : 4);
  const char* Bb = (const char*)g.B + ((size_t)b1 * g.sB + (size_t)b2 * g.sB2) * (g.tb == 0 ? 8 : 4);

  // LDS writes of one staged element (row/column inside the tile, value).
  auto put_a = [&](int buf, int mi, int ki, float xf, double xd) {
    if constexpr (PREC == 0) {
      As[buf][ki * (BM + PADM) + mi] = xd;
    } else if constexpr (PREC == 1) {
      As[buf][ki * (BM + PADM) + mi] = xf;
    } else {
      // The split starts from the FP32 value: the empty asm keeps the
      // compiler from folding double->float->bf16 into one direct rounding
      // (which the generic staging, reading the type at run time, never
      // gets), so both stagings split identically.
      asm("" : "+v"(xf));
      const __bf16 hi = (__bf16)xf;
      const __bf16 lo = (__bf16)(xf - (float)hi);
      As[buf][mi * (BK + PADK) + ki] = hi;
      As[buf][BM * (BK + PADK) + mi * (BK + PADK) + ki] = lo;
    }
  };
  auto put_b = [&](int buf, int ni, int ki, float xf, double xd) {
    if constexpr (PREC == 0) {
      Bs[buf][ki * (BN + PADM) + ni] = xd;
    } else if constexpr (PREC == 1) {
      Bs[buf][ki * (BN + PADM) + ni] = xf;
    } else {
      asm("" : "+v"(xf));
      const __bf16 hi = (__bf16)xf;
      const __bf16 lo = (__bf16)(xf - (float)hi);
      Bs[buf][ni * (BK + PADK) + ki] = hi;
      Bs[buf][BN * (BK + PADK) + ni * (BK + PADK) + ki] = lo;
    }
  };

  // ---- typed staging: 16-byte vectors along the contiguous dimension ----
  using TA_ = typename std::conditional<TYPED, TA, double>::type;
  using TB_ = typename std::conditional<TYPED, TB, double>::type;
  constexpr int VA = 16 / (int)sizeof(TA_), VB = 16 / (int)sizeof(TB_);
  constexpr int CA = A_K ? BK : BM;  // contiguous extent of the A tile
  constexpr int CB = B_N ? BN : BK;  // contiguous extent of the B tile
  static_assert(!TYPED || (CA % VA == 0 && CB % VB == 0), "tile not a whole number of vectors");
  constexpr int NVA = BM * BK / VA, NVB = BK * BN / VB;
  constexpr int EVA = TYPED ? (NVA + NT - 1) / NT : 1;
  constexpr int EVB = TYPED ? (NVB + NT - 1) / NT : 1;
  // tile coordinates of vector e of this thread (first element)
  auto vec_a = [&](int e, int& mi, int& ki) {
    const int idx = tid + e * NT;
    const int cv = idx % (CA / VA), r = idx / (CA / VA);
    if (A_K) { ki = cv * VA; mi = r; } else { mi = cv * VA; ki = r; }
  };
  auto vec_b = [&](int e, int& ni, int& ki) {
    const int idx = tid + e * NT;
    const int cv = idx % (CB / VB), r = idx / (CB / VB);
    if (B_N) { ni = cv * VB; ki = r; } else { ki = cv * VB; ni = r; }
  };
  auto live_a = [&](int e) { return NVA % NT == 0 || tid + e * NT < NVA; };
  auto live_b = [&](int e) { return NVB % NT == 0 || tid + e * NT < NVB; };
  unsigned oa[EVA], ob[EVB];
  int kva[EVA], kvb[EVB];
  u32x4 va[EVA], vb[EVB];
  __amdgpu_buffer_rsrc_t ra, rb;
  int stepA = 0, stepB = 0;  // bytes per k-tile
  if constexpr (TYPED) {
    ra = gemm_rsrc(Ab);
    rb = gemm_rsrc(Bb);
    stepA = (A_K ? BK : BK * (int)g.lda) * (int)sizeof(TA);
    stepB = (B_N ? BK * (int)g.ldb : BK) * (int)sizeof(TB);
#pragma unroll
    for (int e = 0; e < EVA; ++e) {
      int mi, ki;
      vec_a(e, mi, ki);
      const int m = m0 + mi, k = kbeg + ki;
      const int64_t off = A_K ? (int64_t)m * g.lda + k : (int64_t)k * g.lda + m;
      oa[e] = (live_a(e) && m < g.M) ? (unsigned)(off * (int64_t)sizeof(TA)) : kGemmOob;
      kva[e] = ki;
    }
#pragma unroll
    for (int e = 0; e < EVB; ++e) {
      int ni, ki;
      vec_b(e, ni, ki);
      const int n = n0 + ni, k = kbeg + ki;
      const int64_t off = B_N ? (int64_t)k * g.ldb + n : (int64_t)n * g.ldb + k;
      ob[e] = (live_b(e) && n < g.N) ? (unsigned)(off * (int64_t)sizeof(TB)) : kGemmOob;
      kvb[e] = ki;
    }
  }

  // ---- generic staging: raw element bits, converted at the LDS store ----
  // Staging registers hold the raw loaded bits (a double, or a float in the
  // low word) so that no instruction uses a global load's result until
  // store_tiles, after the MFMAs of the current k-block: the loads of
  // block k+1 stay in flight under the compute of block k.
  uint64_t ua[TYPED ? 1 : EA], ub[TYPED ? 1 : EB];

  auto load_tiles = [&](int t) {
    const int k0 = kbeg + t * BK;
    if constexpr (TYPED) {
      const int lim = kend - k0;  // k extent left in this split
#pragma unroll
      for (int e = 0; e < EVA; ++e) {
        const unsigned off = kva[e] < lim ? oa[e] : kGemmOob;
        va[e] = __builtin_amdgcn_raw_buffer_load_b128(ra, (int)off, t * stepA, 0);
      }
#pragma unroll
      for (int e = 0; e < EVB; ++e) {
        const unsigned off = kvb[e] < lim ? ob[e] : kGemmOob;
        vb[e] = __builtin_amdgcn_raw_buffer_load_b128(rb, (int)off, t * stepB, 0);
      }
    } else {
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
    }
  };
  auto as_double = [](uint64_t v, int t) -> double {
    return t == 0 ? __longlong_as_double((long long)v) : (double)__uint_as_float((uint32_t)v);
  };
  auto as_float = [](uint64_t v, int t) -> float {
    return t == 0 ? (float)__longlong_as_double((long long)v) : __uint_as_float((uint32_t)v);
  };

  auto store_tiles = [&](int buf) {
    if constexpr (TYPED) {
#pragma unroll
      for (int e = 0; e < EVA; ++e) {
        if (!live_a(e)) continue;
        int mi, ki;
        vec_a(e, mi, ki);
        if constexpr (PREC != 2 && !A_K) {
          // VA consecutive m of one k row: one vector LDS store
          lds_t* p = &As[buf][ki * (BM + PADM) + mi];
          if constexpr (PREC == 0 && VA == 2) {
            *reinterpret_cast<f64x2*>(p) = f64x2{(double)vec_elem<TA>(va[e], 0), (double)vec_elem<TA>(va[e], 1)};
          } else if constexpr (PREC == 1 && VA == 2) {
            *reinterpret_cast<f32x2*>(p) = f32x2{(float)vec_elem<TA>(va[e], 0), (float)vec_elem<TA>(va[e], 1)};
          } else if constexpr (PREC == 1 && VA == 4) {
            *reinterpret_cast<f32x4*>(p) = f32x4{vec_elem<float>(va[e], 0), vec_elem<float>(va[e], 1),
                                                 vec_elem<float>(va[e], 2), vec_elem<float>(va[e], 3)};
          } else {
#pragma unroll
            for (int j = 0; j < VA; ++j) p[j] = (lds_t)vec_elem<TA>(va[e], j);
          }
        } else {
#pragma unroll
          for (int j = 0; j < VA; ++j) {
            const TA x = vec_elem<TA>(va[e], j);
            if (A_K) put_a(buf, mi, ki + j, (float)x, (double)x);
            else put_a(buf, mi + j, ki, (float)x, (double)x);
          }
        }
      }
#pragma unroll
      for (int e = 0; e < EVB; ++e) {
        if (!live_b(e)) continue;
        int ni, ki;
        vec_b(e, ni, ki);
        if constexpr (PREC != 2 && B_N) {
          lds_t* p = &Bs[buf][ki * (BN + PADM) + ni];
          if constexpr (PREC == 0 && VB == 2) {
            *reinterpret_cast<f64x2*>(p) = f64x2{(double)vec_elem<TB>(vb[e], 0), (double)vec_elem<TB>(vb[e], 1)};
          } else if constexpr (PREC == 1 && VB == 2) {
            *reinterpret_cast<f32x2*>(p) = f32x2{(float)vec_elem<TB>(vb[e], 0), (float)vec_elem<TB>(vb[e], 1)};
          } else if constexpr (PREC == 1 && VB == 4) {
            *reinterpret_cast<f32x4*>(p) = f32x4{vec_elem<float>(vb[e], 0), vec_elem<float>(vb[e], 1),
                                                 vec_elem<float>(vb[e], 2), vec_elem<float>(vb[e], 3)};
          } else {
#pragma unroll
            for (int j = 0; j < VB; ++j) p[j] = (lds_t)vec_elem<TB>(vb[e], j);
          }
        } else {
#pragma unroll
          for (int j = 0; j < VB; ++j) {
            const TB x = vec_elem<TB>(vb[e], j);
            if (B_N) put_b(buf, ni + j, ki, (float)x, (double)x);
            else put_b(buf, ni, ki + j, (float)x, (double)x);
          }
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int idx = tid + e * NT;
        int mi, ki;
        if (A_K) { ki = idx % BK; mi = idx / BK; } else { mi = idx % BM; ki = idx / BM; }
        put_a(buf, mi, ki, as_float(ua[e], g.ta), PREC == 0 ? as_double(ua[e], g.ta) : 0.0);
      }
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int idx = tid + e * NT;
        int ni, ki;
        if (B_N) { ni = idx % BN; ki = idx / BN; } else { ki = idx % BK; ni = idx / BK; }
        put_b(buf, ni, ki, as_float(ub[e], g.tb), PREC == 0 ? as_double(ub[e], g.tb) : 0.0);
      }
    }
  };

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int nk = (kend - kbeg + BK - 1) / BK;
  const int lr = lane & 15;   // row (A) / column (B) inside a 16x16 tile
  const int lk = lane >> 4;   // k group of the lane
  //   f32 MFMAs: row = 4*lk + v ;  f64 MFMA: row = lk + 4*v ; column = lane%16
  auto row_of = [&](int v) { return (PREC == 0) ? (lk + 4 * v) : (lk * 4 + v); };
  char* Cb = (char*)g.C + ((size_t)b1 * g.sC + (size_t)b2 * g.sC2) * (g.tc == 0 ? 8 : 4);
  auto load_c = [&](int i, int j, int v) -> double {
    const int m = m0 + wm0 + i * 16 + row_of(v);
    const int n = n0 + wn0 + j * 16 + lr;
    if (!(m < g.M && n < g.N)) return 0.0;
    const int64_t off = (int64_t)m * g.ldc + n;
    return g.tc == 0 ? reinterpret_cast<const double*>(Cb)[off]
                     : (double)reinterpret_cast<const float*>(Cb)[off];
  };
  // Small wave tiles (at most 4 accumulator tiles, the 64x64 block) load
  // their old C values before the first operand tiles, so with beta != 0
  // the two reads overlap instead of the epilogue waiting for a second
  // round trip (the rank-32 and rank-128 updates of the blocked QR: one or
  // a few k-tiles).
  constexpr bool PRE_C = TM * TN <= 4;
  const bool pre_c = PRE_C && g.beta != 0.0 && !g.partial;
  double cpre[PRE_C ? TM : 1][PRE_C ? TN : 1][4];
  if constexpr (PRE_C) {
    if (pre_c) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) cpre[i][j][v] = load_c(i, j, v);
    }
  }
  load_tiles(0);
  store_tiles(0);
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load_tiles(kt + 1);
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

  // Epilogue. 16x16 accumulator tile, lane group lk = lane/16, item v
  // (rows as in row_of above).
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
  // One 16-row band of the wave's tile at a time: with beta != 0 all of the
  // band's old C values are loaded first (or were, before the k loop), then
  // combined and stored (a load placed after the previous element's store
  // would wait for it, since the compiler cannot tell the addresses apart).
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    double cv[TN][4];
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        cv[j][v] = 0.0;
        if constexpr (PRE_C) {
          if (pre_c) cv[j][v] = cpre[i][j][v];
          else if (g.beta != 0.0) cv[j][v] = load_c(i, j, v);
        } else {
          if (g.beta != 0.0) cv[j][v] = load_c(i, j, v);
        }
      }
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm0 + i * 16 + row_of(v);
        const int n = n0 + wn0 + j * 16 + lr;
        if (m < g.M && n < g.N) {
          const int64_t off = (int64_t)m * g.ldc + n;
          double val = g.alpha * (double)acc[i][j][v];
          if (g.beta != 0.0) val += g.beta * cv[j][v];
          if (g.tc == 0) reinterpret_cast<double*>(Cb)[off] = val;
          else reinterpret_cast<float*>(Cb)[off] = (float)val;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Tile configurations. A config only changes how output tiles are cut and
// scheduled: every output element still accumulates its K range in order,
// so results are bitwise identical whichever config runs.
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

// Launch of config `ci` of arithmetic PREC. `typed` selects the typed
// staging (the caller has checked gemm_typed_ok); defined per arithmetic in
// gemm_f64.hip, gemm_f32.hip and gemm_bf16.hip.
void gemm_launch_f64(int ci, const GemmDesc& d, const GemmArgs& g, int splits, bool typed, hipStream_t s);
void gemm_launch_f32(int ci, const GemmDesc& d, const GemmArgs& g, int splits, bool typed, hipStream_t s);
void gemm_launch_bf16(int ci, const GemmDesc& d, const GemmArgs& g, int splits, bool typed, hipStream_t s);

// True when the typed staging applies: the operand types are an
// instantiated pair, each operand's contiguous extent, leading dimension and
// base/batch offsets are whole 16-byte vectors, the split length is a whole
// number of vectors, and one batch entry of each operand spans < 2 GiB.
bool gemm_typed_ok(const GemmDesc& d, int kc);

// Tile rows per group of the XCD-aware order: an XCD's ceil(tiles / 8)
// tiles as close to the block a x b that minimises the panels it reads
// (a BM + b BN with a b fixed: a = sqrt(tiles/8 * BN / BM)). Row-major
// order (0) below 64 tiles, and where the batch or split index makes
// blockIdx.x % 8 differ from the XCD (a grid.x not a multiple of 8 with
// more than one batch entry or split).
inline int gemm_group_m(int tiles_m, int tiles_n, int bm, int bn, int yz) {
  const int tiles = tiles_m * tiles_n;
  if (tiles < 64 || (yz > 1 && tiles % 8 != 0)) return 0;
  const double per = (double)((tiles + 7) / 8);
  const int a = (int)(std::sqrt(per * bn / bm) + 0.5);
  return std::max(1, std::min(a, tiles_m));
}

template <int PREC, int BM, int BN, int WGM, int WGN, int BK, typename TA, typename TB>
static void gemm_launch_cfg(const GemmDesc& d, const GemmArgs& g0, int splits, hipStream_t s) {
  GemmArgs g = g0;
  g.tiles_n = ceil_div(d.N, BN);
  const int tiles_m = ceil_div(d.M, BM);
  const int tiles = tiles_m * g.tiles_n;
  g.group_m = gemm_group_m(tiles_m, g.tiles_n, BM, BN, d.batch * splits);
  dim3 grid(tiles, d.batch, splits);
  dim3 block(64 * WGM * WGN);
  if (d.a_kcontig && d.b_ncontig)
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, true, true, WGM, WGN, BK, TA, TB>), grid, block, 0, s, g);
  else if (d.a_kcontig && !d.b_ncontig)
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, true, false, WGM, WGN, BK, TA, TB>), grid, block, 0, s, g);
  else if (!d.a_kcontig && d.b_ncontig)
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, false, true, WGM, WGN, BK, TA, TB>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, false, false, WGM, WGN, BK, TA, TB>), grid, block, 0, s, g);
}

// Config index -> launch, for one operand-type pair (void, void = generic).
template <int PREC, typename TA, typename TB>
static void gemm_launch_index(int ci, const GemmDesc& d, const GemmArgs& g, int splits, hipStream_t s) {
#define TNX_CFG(I)                                                                              \
  if (ci == I && I < Cfgs<PREC>::n) {                                                           \
    constexpr Cfg c = Cfgs<PREC>::c[I < Cfgs<PREC>::n ? I : 0];                                 \
    gemm_launch_cfg<PREC, c.bm, c.bn, c.wgm, c.wgn, c.bk, TA, TB>(d, g, splits, s);             \
    return;                                                                                     \
  }
  TNX_CFG(0) TNX_CFG(1) TNX_CFG(2) TNX_CFG(3) TNX_CFG(4) TNX_CFG(5)
#undef TNX_CFG
}

}  // namespace tnx
