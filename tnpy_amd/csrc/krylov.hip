// Lowest eigenpair of the one-site effective Hamiltonian on the device, and
// the small dense symmetric eigensolver.
//
// Replaces primme.eigsh(which='SA', k=1) (tnpy/linalg.py:64-87, called at
// tnpy/finite_dmrg.py:111) and np.linalg.eigh below exact_solver_dim
// (tnpy/linalg.py:42-61, tnpy/finite_dmrg.py:108-110).
//
// Krylov method: Lanczos with full reorthogonalisation (classical
// Gram-Schmidt, two passes) and thick restart keeping one Ritz vector. The
// projected matrix stays tridiagonal; its lowest pair is found by one
// wavefront (64-way multisection on Sturm counts, then inverse iteration).
// All Krylov scalars live in device memory; the host reads one small status
// block per iteration to decide convergence / restart.
//
// Mixed precision: the Krylov matvecs run in FP32 or BF16x3; every cycle
// starts from an FP64 matvec on the current Ritz vector (Rayleigh quotient
// and residual in FP64), and the next cycle's basis is [y, r/|r|]. The
// returned eigenvalue, residual and H*y are FP64.
//
// Vector kernels are HBM bound: a CGS pass over J basis vectors of length N
// reads (J+1)*N*8 bytes (+N*8 written when it updates w).
#include "tnx_common.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

namespace tnx {

int heff1_matvec(const double* L, const double* W, const double* R, const double* x,
                 double* y, int chil, int Dl, int Dr, int d, int chir, int prec,
                 void* ws, size_t ws_bytes, hipStream_t s);
size_t heff1_ws_bytes(int chil, int Dl, int Dr, int d, int chir);
int heff2_matvec(const double* L, const double* W1, const double* W2, const double* R,
                 const void* x, int x_f32, double* y, int chil, int Dl, int Dm, int Dr, int d,
                 int chir, int prec, void* ws, size_t ws_bytes, hipStream_t s);
int heff1_matvec_b(const void* L, const double* W, const void* R, const void* x, int x_f32,
                   double* y, int chil, int Dl, int Dr, int d, int chir, int prec, int batch,
                   int count, const int* bidx, int64_t sx, int64_t sy, void* ws, size_t ws_bytes,
                   hipStream_t s, int env_f32);
size_t heff2_ws_bytes(int chil, int Dl, int Dm, int Dr, int d, int chir);

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Problem map of a batched launch: launched problem i = blockIdx.y works on
// logical problem idx[i] (or i when idx is null), whose work arrays sit at
// +problem * sp doubles from the base pointers. One problem: idx null, sp 0.
struct PMap {
  const int* idx;
  int64_t sp;
  __device__ __forceinline__ int pid() const { return idx ? idx[blockIdx.y] : (int)blockIdx.y; }
  __device__ __forceinline__ int64_t off() const { return (int64_t)pid() * sp; }
};
static const PMap kOne{nullptr, 0};

constexpr int kMaxBasis = 64;        // Lanczos vectors kept before restart
constexpr int kPartStride = kMaxBasis + 2;
constexpr int kCgsBlocks = 512;      // blocks of a CGS pass (partials rows)

// One classical Gram-Schmidt pass over basis V[0..J-1] (row stride ldv):
//   if hsub:  w -= sum_j V[j] * hsub[j]               (in place)
//   partial[blk][j] = sum_i V[j][i] w[i]   (j < J, if DOTS)
//   partial[blk][JMAX... stored at index J] = sum_i w[i]^2
template <int JMAX, bool DOTS, typename VT>
__global__ __launch_bounds__(256) void cgs_pass_kernel(const VT* __restrict__ V,
                                                       int64_t ldv, int J,
                                                       double* __restrict__ w,
                                                       const double* __restrict__ hsub,
                                                       double* __restrict__ partials,
                                                       int64_t N, PMap pm) {
  {
    const int64_t o = pm.off();
    V += o; w += o; partials += o;
    if (hsub) hsub += o;
  }
  __shared__ double hs[JMAX];
  __shared__ double red[4][JMAX + 1];
  if (threadIdx.x < JMAX) hs[threadIdx.x] = (hsub && (int)threadIdx.x < J) ? hsub[threadIdx.x] : 0.0;
  __syncthreads();
  double acc[JMAX];
#pragma unroll
  for (int j = 0; j < JMAX; ++j) acc[j] = 0.0;
  double nrm = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if constexpr (std::is_same<VT, double>::value) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
      double vv[JMAX];
#pragma unroll
      for (int j = 0; j < JMAX; ++j) vv[j] = (j < J) ? V[(int64_t)j * ldv + i] : 0.0;
      double wi = w[i];
      if (hsub) {
#pragma unroll
        for (int j = 0; j < JMAX; ++j) wi -= vv[j] * hs[j];
        w[i] = wi;
      }
      if (DOTS) {
#pragma unroll
        for (int j = 0; j < JMAX; ++j) acc[j] += vv[j] * wi;
      }
      nrm += wi * wi;
    }
  } else {
    // FP32 basis: unconditional loads (rows past J re-read row J-1, a cache
    // hit) and the widening after all of them. A guarded load with its
    // conversion in the same branch waits for its data before the next
    // load issues.
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
      float raw[JMAX];
#pragma unroll
      for (int j = 0; j < JMAX; ++j) raw[j] = V[(int64_t)(j < J ? j : J - 1) * ldv + i];
      double vv[JMAX];
#pragma unroll
      for (int j = 0; j < JMAX; ++j) vv[j] = (j < J) ? (double)raw[j] : 0.0;
      double wi = w[i];
      if (hsub) {
#pragma unroll
        for (int j = 0; j < JMAX; ++j) wi -= vv[j] * hs[j];
        w[i] = wi;
      }
      if (DOTS) {
#pragma unroll
        for (int j = 0; j < JMAX; ++j) acc[j] += vv[j] * wi;
      }
      nrm += wi * wi;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (DOTS) {
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
      const double t = wave_sum(acc[j]);
      if (lane == 0) red[wv][j] = t;
    }
  }
  {
    const double t = wave_sum(nrm);
    if (lane == 0) red[wv][JMAX] = t;
  }
  __syncthreads();
  double* out = partials + (int64_t)blockIdx.x * kPartStride;
  if ((int)threadIdx.x < J && DOTS)
    out[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
  if (threadIdx.x == 0) out[J] = red[0][JMAX] + red[1][JMAX] + red[2][JMAX] + red[3][JMAX];
}

// The same pass with R lanes per group of E elements (FP64 basis: R = 2,
// E = 2; FP32: R = 4, E = 4; one 16-byte load per basis row): lane
// h = lane % R of a group reads basis rows [h*JMAX/R, (h+1)*JMAX/R) of its
// group; the subtraction's partial sums are combined across the group by
// shuffles, and each lane accumulates dot products for its own rows
// (reduced over the lanes with its h). Every load moves 16 bytes and a lane
// holds 1/R of the rows, against 4- or 8-byte loads of all rows for one
// lane per element. Needs N and ldv multiples of E and V, w 16-byte aligned
// (cgs_pass checks).
__device__ __forceinline__ double vcomp(const double2& v, int q) { return q == 0 ? v.x : v.y; }
__device__ __forceinline__ double vcomp(const float4& v, int q) {
  return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}

template <int JMAX, bool DOTS, typename VT>
__global__ __launch_bounds__(256) void cgs_pair_kernel(const VT* __restrict__ V, int64_t ldv,
                                                       int J, double* __restrict__ w,
                                                       const double* __restrict__ hsub,
                                                       double* __restrict__ partials, int64_t N,
                                                       PMap pm) {
  using PV = typename std::conditional<std::is_same<VT, float>::value, float4, double2>::type;
  constexpr int E = 16 / sizeof(VT);
  constexpr int R = E;
  constexpr int JH = JMAX / R;
  {
    const int64_t o = pm.off();
    V += o; w += o; partials += o;
    if (hsub) hsub += o;
  }
  __shared__ double hs[JMAX];
  __shared__ double red[4][JMAX + 1];
  if (threadIdx.x < JMAX) hs[threadIdx.x] = (hsub && (int)threadIdx.x < J) ? hsub[threadIdx.x] : 0.0;
  __syncthreads();
  const int h = threadIdx.x & (R - 1);
  const int r0 = h * JH;
  double acc[JH];
#pragma unroll
  for (int j = 0; j < JH; ++j) acc[j] = 0.0;
  double nrm = 0.0;
  const PV* Vp = reinterpret_cast<const PV*>(V);
  double2* w2 = reinterpret_cast<double2*>(w);
  const int64_t NE = N / E, ldE = ldv / E;
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) / R;
  for (int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / R; e < NE; e += stride) {
    PV raw[JH];
#pragma unroll
    for (int jj = 0; jj < JH; ++jj) {
      const int j = r0 + jj;
      raw[jj] = Vp[(int64_t)(j < J ? j : J - 1) * ldE + e];
    }
    double wv[E];
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
      const double2 t = w2[e * (E / 2) + q];
      wv[2 * q] = t.x;
      wv[2 * q + 1] = t.y;
    }
    if (hsub) {
      double sm[E];
#pragma unroll
      for (int q = 0; q < E; ++q) sm[q] = 0.0;
#pragma unroll
      for (int jj = 0; jj < JH; ++jj) {
        const double hj = hs[r0 + jj];
#pragma unroll
        for (int q = 0; q < E; ++q) sm[q] += vcomp(raw[jj], q) * hj;
      }
#pragma unroll
      for (int o = 1; o < R; o <<= 1)
#pragma unroll
        for (int q = 0; q < E; ++q) sm[q] += __shfl_xor(sm[q], o, 64);
#pragma unroll
      for (int q = 0; q < E; ++q) wv[q] -= sm[q];
      if (h == 0) {
#pragma unroll
        for (int q = 0; q < E / 2; ++q) w2[e * (E / 2) + q] = make_double2(wv[2 * q], wv[2 * q + 1]);
      }
    }
    if (DOTS) {
#pragma unroll
      for (int jj = 0; jj < JH; ++jj) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < E; ++q) t += vcomp(raw[jj], q) * wv[q];
        acc[jj] += t;
      }
    }
    if (h == 0) {
#pragma unroll
      for (int q = 0; q < E; ++q) nrm += wv[q] * wv[q];
    }
  }
  const int lane = threadIdx.x & 63, wv_ = threadIdx.x >> 6;
  if (DOTS) {
#pragma unroll
    for (int jj = 0; jj < JH; ++jj) {
      double t = acc[jj];
#pragma unroll
      for (int o = 32; o >= R; o >>= 1) t += __shfl_xor(t, o, 64);  // lanes with one h
      if (lane < R) red[wv_][lane * JH + jj] = t;
    }
  }
  {
    const double t = wave_sum(nrm);
    if (lane == 0) red[wv_][JMAX] = t;
  }
  __syncthreads();
  double* out = partials + (int64_t)blockIdx.x * kPartStride;
  if ((int)threadIdx.x < J && DOTS)
    out[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
  if (threadIdx.x == 0) out[J] = red[0][JMAX] + red[1][JMAX] + red[2][JMAX] + red[3][JMAX];
}

// out[j] = sum over blocks of partials[blk][j], j <= J. One workgroup per
// j; a fixed strided split and a fixed tree, so the sum is deterministic.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* partials, int nblk,
                                                              int J, double* out, PMap pm) {
  {
    const int64_t o = pm.off();
    partials += o; out += o;
  }
  __shared__ double red[4];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partials[(int64_t)b * kPartStride + j];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = (red[0] + red[1]) + (red[2] + red[3]);
}

// One CGS pass over a basis stored in VT (double, or float for the basis of
// a low-precision Krylov cycle); w and all sums are FP64.
template <typename VT>
static int cgs_pass(const VT* V, int64_t ldv, int J, double* w, const double* hsub,
                    bool dots, double* partials, double* out, int64_t N, hipStream_t s,
                    PMap pm = kOne, int count = 1) {
  const int nblk = (int)std::min<int64_t>(count > 1 ? std::max(kCgsBlocks / count, 8) : kCgsBlocks,
                                          (N + 255) / 256);
  // Passes with dot products take the pair kernel (measured faster); the
  // subtract-and-norm pass keeps one lane per element (the pair form is
  // slower there). The pair kernel needs lengths in whole 16-byte groups and
  // 16-byte aligned pointers (offsets of batched problems are multiples of
  // 256 bytes).
  constexpr int E = 16 / sizeof(VT);
  const bool pair = dots && J >= 2 && N % E == 0 && ldv % E == 0 && ((uintptr_t)V % 16) == 0 &&
                    ((uintptr_t)w % 16) == 0;
  auto go = [&](auto jtag, auto dtag) {
    constexpr int JM = decltype(jtag)::value;
    constexpr bool DT_ = decltype(dtag)::value;
    if (pair)
      hipLaunchKernelGGL((cgs_pair_kernel<JM, DT_, VT>), dim3(nblk, count), dim3(256), 0, s, V,
                         ldv, J, w, hsub, partials, N, pm);
    else
      hipLaunchKernelGGL((cgs_pass_kernel<JM, DT_, VT>), dim3(nblk, count), dim3(256), 0, s, V, ldv,
                         J, w, hsub, partials, N, pm);
  };
  using T8 = std::integral_constant<int, 8>;
  using T16 = std::integral_constant<int, 16>;
  using T32 = std::integral_constant<int, 32>;
  using T64 = std::integral_constant<int, 64>;
  using BT = std::integral_constant<bool, true>;
  using BF = std::integral_constant<bool, false>;
  TNX_CHECK_ARG(J >= 0 && J <= kMaxBasis, "cgs_pass: J=%d out of range", J);
  if (J <= 8) { if (dots) go(T8{}, BT{}); else go(T8{}, BF{}); }
  else if (J <= 16) { if (dots) go(T16{}, BT{}); else go(T16{}, BF{}); }
  else if (J <= 32) { if (dots) go(T32{}, BT{}); else go(T32{}, BF{}); }
  else { if (dots) go(T64{}, BT{}); else go(T64{}, BF{}); }
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(J + 1, count), dim3(256), 0, s, partials, nblk,
                     J, out, pm);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

// out[i] = sum_j c[j] V[j][i]   (Ritz vector from the basis)
template <typename VT>
__global__ void vcomb_kernel(const VT* __restrict__ V, int64_t ldv, int J,
                             const double* __restrict__ c, double* __restrict__ out, int64_t N,
                             PMap pm) {
  {
    const int64_t o = pm.off();
    V += o; c += o; out += o;
  }
  __shared__ double cs[kMaxBasis + 1];
  if ((int)threadIdx.x < J) cs[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < J; ++j) acc += cs[j] * (double)V[(int64_t)j * ldv + i];
    out[i] = acc;
  }
}

// out = x / sqrt(*nrm2)  (unit vector; zero if the norm underflows)
template <typename OT>
__global__ void normalize_kernel(const double* __restrict__ x, const double* __restrict__ nrm2,
                                 OT* __restrict__ out, int64_t N, PMap pm) {
  {
    const int64_t o = pm.off();
    x += o; nrm2 += o; out += o;
  }
  const double n2 = *nrm2;
  const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (OT)(x[i] * inv);
}

// out = (double)x
__global__ void to_f64_kernel(const float* __restrict__ x, double* __restrict__ out, int64_t N) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (double)x[i];
}

// out = (float)x (the first vector of a low-precision Krylov basis)
__global__ void to_f32_kernel(const double* __restrict__ x, float* __restrict__ out, int64_t N) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)x[i];
}

// out = (*a) * x + (*b) * z, scalars read from device memory. Batched:
// a, x, b, z move with the problem map, out by so doubles per problem.
__global__ void axpbz_dev_kernel(const double* __restrict__ a, const double* __restrict__ x,
                                 const double* __restrict__ b, const double* __restrict__ z,
                                 double* __restrict__ out, int64_t N, PMap pm, int64_t so) {
  {
    const int64_t o = pm.off();
    a += o; x += o;
    if (b) b += o;
    if (z) z += o;
    out += (int64_t)pm.pid() * so;
  }
  const double av = *a, bv = b ? *b : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = av * x[i] + (z ? bv * z[i] : 0.0);
}

__global__ void axpy_kernel(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
__global__ void scal_kernel(int64_t n, double a, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

static int vec_blocks(int64_t N, int count = 1) {
  return (int)std::min<int64_t>((N + 255) / 256, count > 1 ? std::max(4096 / count, 8) : 4096);
}

// out_p = in_p for the mapped problems: in moves with the map (sp), out by so.
__global__ void copy_map_kernel(const double* __restrict__ in, double* __restrict__ out,
                                int64_t N, PMap pm, int64_t sin_, int64_t so) {
  const int64_t p = pm.pid();
  in += p * sin_;
  out += p * so;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

// ---------------------------------------------------------------------------
// Symmetric tridiagonal lowest eigenpair, one wavefront.
// diag a[0..n-1], off-diagonal b[0..n-2]. Writes the unit eigenvector to
// s[0..n-1] and status: st[0] = theta, st[1] = |bres * s[n-1]| (Lanczos
// residual estimate, bres read from *bres if given), st[2] = max row sum
// (estimate of ||T||, a lower bound of ||H||).
// ---------------------------------------------------------------------------
constexpr int kTriMax = 256;

// Number of eigenvalues of the tridiagonal (a, b) below x: sign changes
// of the leading principal minors p_k = det(T_k - x I) (p_0 = 1), by the
// three-term recurrence p_{k+1} = (a_k - x) p_k - b_{k-1}^2 p_{k-1}. It has
// no division (two dependent FMAs per step, against a division per step for
// the LDL^T pivots); every 4 steps both carried minors are rescaled by a
// power of two so they neither overflow nor underflow. An exact zero minor
// takes the sign opposite to its predecessor (the pivot convention q = -0).
__device__ int sturm_count(const double* a, const double* b2, int n, double x) {
  double pm = 1.0, p = a[0] - x;
  bool neg_prev = false;
  int cnt = 0;
  for (int i = 0;; ++i) {
    const bool neg = (p < 0.0) || (p == 0.0 && !neg_prev);
    cnt += (neg != neg_prev);
    neg_prev = neg;
    if (i + 1 >= n) break;
    const double pn = fma(a[i + 1] - x, p, -b2[i] * pm);
    pm = p;
    p = pn;
    if ((i & 3) == 3) {
      const int e = __builtin_amdgcn_frexp_exp(p == 0.0 ? pm : p);
      p = ldexp(p, -e);
      pm = ldexp(pm, -e);
    }
  }
  return cnt;
}

// Lane i's value of a wave-distributed double, as a wave-uniform scalar
// (v_readlane: no LDS round trip). i must be wave-uniform.
__device__ __forceinline__ double lane_val(double v, int i) {
  const long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), i);
  const unsigned hi = __builtin_amdgcn_readlane((int)(u >> 32), i);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ int lane_int(int v, int i) { return __builtin_amdgcn_readlane(v, i); }

// sturm_count with row k of the tridiagonal held by lane k (n <= 64): the
// same recurrence, rescaling and zero convention, unrolled by four so the
// rescale is not a per-step select, and the sign test without branches
// (the loop issues few instructions per step; one wave runs it alone).
__device__ int sturm_count_reg(double ar, double b2r, int n, double x) {
  double pm = 1.0, p = lane_val(ar, 0) - x;
  bool neg_prev = false;
  int cnt = 0;
  auto sign_step = [&](double v) {
    const bool neg = (v < 0.0) | ((v == 0.0) & !neg_prev);
    cnt += (int)(neg ^ neg_prev);
    neg_prev = neg;
  };
  auto rec = [&](int i) {  // p_{i+1} from p_i, p_{i-1}
    const double pn = fma(lane_val(ar, i + 1) - x, p, -lane_val(b2r, i) * pm);
    pm = p;
    p = pn;
  };
  sign_step(p);
  int i = 0;
  for (; i + 4 < n; i += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      rec(i + u);
      sign_step(p);
    }
    const int e = __builtin_amdgcn_frexp_exp(p == 0.0 ? pm : p);
    p = ldexp(p, -e);
    pm = ldexp(pm, -e);
  }
  for (; i + 1 < n; ++i) {
    rec(i);
    sign_step(p);
  }
  return cnt;
}

// The lowest eigenpair for n <= 64 (every Lanczos projection): the same
// steps as the LDS form in tridiag_lowest_kernel, with row k of T and of the
// factors in lane k's registers. The serial recurrences (Sturm counts,
// factorisation, the two solves) run on all lanes as wave-uniform scalar
// chains that read a row with v_readlane and write lane k's value with a
// lane select, so no step waits on LDS.
__device__ void tridiag_lowest_reg(const double* __restrict__ ad, const double* __restrict__ bd,
                                   int n, bool has_bres, double bres_v, bool fin, double a_last,
                                   double* __restrict__ s, double* st, int hinted) {
  const int lane = threadIdx.x;
  // the previous call's theta (st[0]), read before this call overwrites it
  const double hint_in = hinted ? st[0] : 0.0;
  const double ar = lane < n ? ((fin && lane == n - 1) ? a_last : ad[lane]) : 0.0;
  const double br = lane < n - 1 ? bd[lane] : 0.0;
  const double b2r = br * br;
  // Gershgorin bounds and ||T|| estimate
  double lo, hi, tnorm;
  {
    // the shuffle runs on every lane (a lane that is switched off reads as
    // zero to the lanes that shuffle from it); lane 0 then drops its value
    const double bup = __shfl_up(br, 1, 64);
    const double bl = lane > 0 ? fabs(bup) : 0.0;
    const double r = bl + (lane < n - 1 ? fabs(br) : 0.0);
    lo = lane < n ? ar - r : 1e308;
    hi = lane < n ? ar : 1e308;
    tnorm = lane < n ? fabs(ar) + r : 0.0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmin(hi, __shfl_xor(hi, o, 64));
    tnorm = fmax(tnorm, __shfl_xor(tnorm, o, 64));
  }
  const double pivmin = fmax(tnorm, 1e-300) * 1e-300 + 1e-300;
  const double eps = 2.220446049250313e-16;
  // Multisection. The 64 points of a round may be any increasing sequence
  // in (lo, hi): the first with a count >= 1 and its predecessor bracket the
  // lowest eigenvalue whatever the points are. Each round spaces them
  // evenly, except a first round given a hint h (the lowest Ritz value of
  // the previous Lanczos step, an upper bound of this one by interlacing,
  // usually just above it): there they approach h geometrically from below,
  // h - (h - lo) q^(k+1), so a small move below h is bracketed to within a
  // fraction of its size in one round. A wrong hint costs one round, never
  // the result.
  double h = hint_in;
  if (!(h > lo && h <= hi)) h = hi;
  const double dmin = 4.0 * eps * fabs(h) + pivmin;
  bool geo = hinted && (h - lo) > 4.0 * dmin;
  for (int it = 0; it < 13; ++it) {
    if (hi - lo <= 4.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    double x;
    if (geo) {
      const double d = h - lo;
      const double lq = log2(dmin / d) / 64.0;  // log2 q, q < 1
      x = h - d * exp2(lq * (double)(lane + 1));
      x = fmin(fmax(x, lo), h);
      geo = false;
    } else {
      x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    }
    const int c = sturm_count_reg(ar, b2r, n, x);
    const unsigned long long mask = __ballot(c >= 1);
    if (mask == 0ull) {
      lo = __shfl(x, 63, 64);
    } else {
      const int f = __ffsll((long long)mask) - 1;
      const double xf = __shfl(x, f, 64);
      const double xp = __shfl(x, f > 0 ? f - 1 : 0, 64);
      if (f > 0) lo = xp;
      hi = xf;
    }
  }
  const double theta = 0.5 * (lo + hi);
  const double tiny = eps * fmax(tnorm, 1e-300);
  // Gaussian elimination with partial pivoting of T - theta I (factors once)
  double u0r = 0.0, u1r = 0.0, u2r = 0.0, fmr = 0.0;
  int swr = 0;
  {
    double c0 = lane_val(ar, 0) - theta, c1 = lane_val(br, 0), c2 = 0.0;
    for (int i = 0; i < n - 1; ++i) {
      const double d0 = lane_val(br, i), d1 = lane_val(ar, i + 1) - theta;
      const double d2 = (i + 1 < n - 1) ? lane_val(br, i + 1) : 0.0;
      double e0, e1, e2, f;
      int sw;
      if (fabs(c0) >= fabs(d0)) {
        if (c0 == 0.0) c0 = tiny;
        f = d0 / c0;
        e0 = c0; e1 = c1; e2 = c2; sw = 0;
        c0 = d1 - f * c1; c1 = d2 - f * c2; c2 = 0.0;
      } else {
        f = c0 / d0;
        e0 = d0; e1 = d1; e2 = d2; sw = 1;
        const double n0 = c1 - f * d1, n1 = c2 - f * d2;
        c0 = n0; c1 = n1; c2 = 0.0;
      }
      if (lane == i) { u0r = e0; u1r = e1; u2r = e2; fmr = f; swr = sw; }
    }
    if (c0 == 0.0) c0 = tiny;
    if (lane == n - 1) { u0r = c0; u1r = 0.0; u2r = 0.0; }
  }
  const double rinvr = lane < n ? 1.0 / u0r : 0.0;
  double zr = lane < n ? 1.0 + 0.01 * (double)((lane * 7919) % 13) : 0.0;
  for (int rep = 0; rep < 2; ++rep) {
    // forward: apply the row swaps and multipliers to the right-hand side
    double rc = lane_val(zr, 0);
    for (int i = 0; i < n - 1; ++i) {
      const double rd = lane_val(zr, i + 1);
      const double f = lane_val(fmr, i);
      double zi;
      if (lane_int(swr, i)) { zi = rd; rc = rc - f * rd; }
      else { zi = rc; rc = rd - f * rc; }
      if (lane == i) zr = zi;
    }
    // back substitution, carrying z[i+1] and z[i+2]
    double z1 = rc * lane_val(rinvr, n - 1);
    if (lane == n - 1) zr = z1;
    double z2 = 0.0;
    if (n >= 2) {
      const double zn2 = (lane_val(zr, n - 2) - lane_val(u1r, n - 2) * z1) * lane_val(rinvr, n - 2);
      if (lane == n - 2) zr = zn2;
      z2 = z1;
      z1 = zn2;
    }
    for (int i = n - 3; i >= 0; --i) {
      const double zi =
          (lane_val(zr, i) - lane_val(u1r, i) * z1 - lane_val(u2r, i) * z2) * lane_val(rinvr, i);
      if (lane == i) zr = zi;
      z2 = z1;
      z1 = zi;
    }
    const double inv = 1.0 / sqrt(wave_sum(zr * zr));
    zr *= inv;
  }
  // sign: largest |component| positive, the first such index on ties
  double mx = fabs(zr);
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  const unsigned long long at = __ballot(lane < n && fabs(zr) == mx);
  const int im = __ffsll((long long)at) - 1;
  const double sg = lane_val(zr, im) < 0.0 ? -1.0 : 1.0;
  if (lane == 0) {
    st[0] = theta;
    st[1] = has_bres ? fabs(bres_v * lane_val(zr, n - 1)) : 0.0;
    st[2] = tnorm;
  }
  if (lane < n) s[lane] = sg * zr;
}

// The end of a Lanczos step, folded into the solve that follows it: from
// the step's CGS sums, alpha[j] = h1[j] + h2[j] and beta[j] = sqrt(h3[j+1])
// (j = n - 1, the new row; beta[j] is also the residual coupling), written to
// alpha, beta and the dense T. h1 == nullptr: no step to finish.
struct TriFinish {
  const double* h1; const double* h2; const double* h3;
  double* alpha; double* beta; double* T;
};

// hinted: st[0] holds the lowest eigenvalue of the previous, one row
// smaller projection of the same Lanczos run (a search hint only).
// With a step to finish (f.h1), ad is f.alpha and the new row's alpha and
// the coupling come from the sums directly (every lane computes them; lane
// 0 stores them for the later users), never through memory in this kernel.
// Device-decided convergence for a batched run that reads the statuses only
// every few steps (ctol >= 0; negative switches it off): st[12] keeps the running ||T|| estimate (the
// criterion's anorm); once est <= ctol * anorm, st[10] = 1 and st[11] = n,
// and the problem's later calls return at once, leaving its Ritz pair
// (st[0..2], s) and the basis rows up to n as they were at convergence.
__device__ __forceinline__ void chunk_mark(double* st, int n, double ctol) {
  if (ctol >= 0.0 && threadIdx.x == 0) {
    const double an = fmax(st[12], st[2]);
    st[12] = an;
    if (st[1] <= ctol * an) {
      st[10] = 1.0;
      st[11] = (double)n;
    }
  }
}

__global__ __launch_bounds__(64) void tridiag_lowest_kernel(const double* __restrict__ ad,
                                                            const double* __restrict__ bd, int n,
                                                            const double* __restrict__ bres,
                                                            double* __restrict__ s,
                                                            double* st, PMap pm, int hinted,
                                                            TriFinish f, double ctol) {
  const int64_t o = pm.off();
  ad += o; bd += o; s += o; st += o;
  if (bres) bres += o;
  if (ctol >= 0.0 && st[10] != 0.0) return;
  const bool fin = f.h1 != nullptr;
  double a_last = 0.0, b_new = 0.0;
  if (fin) {
    const int j = n - 1;
    a_last = f.h1[o + j] + f.h2[o + j];
    b_new = sqrt(fmax(f.h3[o + j + 1], 0.0));
    if (threadIdx.x == 0) {
      f.alpha[o + j] = a_last;
      f.beta[o + j] = b_new;
      double* T = f.T + o;
      T[j * kMaxBasis + j] = a_last;
      if (j + 1 < kMaxBasis) {
        T[j * kMaxBasis + j + 1] = b_new;
        T[(j + 1) * kMaxBasis + j] = b_new;
      }
    }
  }
  const bool has_bres = fin || bres != nullptr;
  const double bres_v = fin ? b_new : (bres ? *bres : 0.0);
  if (n >= 2 && n <= 64) {
    tridiag_lowest_reg(ad, bd, n, has_bres, bres_v, fin, a_last, s, st, hinted);
    chunk_mark(st, n, ctol);
    return;
  }
  __shared__ double a[kTriMax], b[kTriMax], b2[kTriMax];
  __shared__ double u0[kTriMax], u1[kTriMax], u2[kTriMax], z[kTriMax];
  __shared__ double fm[kTriMax], rinv[kTriMax];
  __shared__ int swp[kTriMax];
  __shared__ double shr[4];
  const int lane = threadIdx.x;
  for (int i = lane; i < n; i += 64) {
    a[i] = (fin && i == n - 1) ? a_last : ad[i];
    const double bi = (i < n - 1) ? bd[i] : 0.0;
    b[i] = bi;
    b2[i] = bi * bi;
  }
  __syncthreads();
  if (n == 1) {
    if (lane == 0) {
      s[0] = 1.0;
      st[0] = a[0];
      st[1] = has_bres ? fabs(bres_v) : 0.0;
      st[2] = fabs(a[0]);
    }
    chunk_mark(st, n, ctol);
    return;
  }
  // Gershgorin bounds (lanes over rows, then wave min/max).
  double lo = 1e308, hi = 1e308, tnorm = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < n - 1 ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmin(hi, a[i]);
    tnorm = fmax(tnorm, fabs(a[i]) + r);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmin(hi, __shfl_xor(hi, o, 64));
    tnorm = fmax(tnorm, __shfl_xor(tnorm, o, 64));
  }
  const double pivmin = fmax(tnorm, 1e-300) * 1e-300 + 1e-300;
  const double eps = 2.220446049250313e-16;
  // 64-way multisection: each lane evaluates the Sturm count at one point.
  for (int it = 0; it < 12; ++it) {
    if (hi - lo <= 4.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    const int c = sturm_count(a, b2, n, x);
    const unsigned long long mask = __ballot(c >= 1);
    if (mask == 0ull) {
      lo = __shfl(x, 63, 64);
    } else {
      const int f = __ffsll((long long)mask) - 1;
      const double xf = __shfl(x, f, 64);
      const double xp = __shfl(x, f > 0 ? f - 1 : 0, 64);
      if (f > 0) lo = xp;
      hi = xf;
    }
  }
  const double theta = 0.5 * (lo + hi);
  // Inverse iteration on (T - theta I): Gaussian elimination with partial
  // pivoting (lane 0; zero pivots replaced by eps*||T||) factors once,
  // keeping multipliers and row swaps; the two refinement passes then only
  // apply them to the right-hand side and back-substitute with reciprocal
  // pivots computed by all lanes.
  const double tiny = eps * fmax(tnorm, 1e-300);
  for (int i = lane; i < n; i += 64) z[i] = 1.0 + 0.01 * (double)((i * 7919) % 13);
  if (lane == 0) {
    // current pivot row: (c0, c1, c2) at columns i, i+1, i+2
    double c0 = a[0] - theta, c1 = b[0], c2 = 0.0;
    for (int i = 0; i < n - 1; ++i) {
      const double d0 = b[i], d1 = a[i + 1] - theta, d2 = (i + 1 < n - 1) ? b[i + 1] : 0.0;
      if (fabs(c0) >= fabs(d0)) {
        if (c0 == 0.0) c0 = tiny;
        const double f = d0 / c0;
        u0[i] = c0; u1[i] = c1; u2[i] = c2; fm[i] = f; swp[i] = 0;
        c0 = d1 - f * c1; c1 = d2 - f * c2; c2 = 0.0;
      } else {
        const double f = c0 / d0;
        u0[i] = d0; u1[i] = d1; u2[i] = d2; fm[i] = f; swp[i] = 1;
        const double n0 = c1 - f * d1, n1 = c2 - f * d2;
        c0 = n0; c1 = n1; c2 = 0.0;
      }
    }
    if (c0 == 0.0) c0 = tiny;
    u0[n - 1] = c0; u1[n - 1] = 0.0; u2[n - 1] = 0.0;
  }
  __syncthreads();
  for (int i = lane; i < n; i += 64) rinv[i] = 1.0 / u0[i];
  __syncthreads();
  for (int rep = 0; rep < 2; ++rep) {
    if (lane == 0) {
      double rc = z[0];
      for (int i = 0; i < n - 1; ++i) {
        const double rd = z[i + 1];
        if (swp[i]) { z[i] = rd; rc = rc - fm[i] * rd; }
        else { z[i] = rc; rc = rd - fm[i] * rc; }
      }
      z[n - 1] = rc * rinv[n - 1];
      if (n >= 2) z[n - 2] = (z[n - 2] - u1[n - 2] * z[n - 1]) * rinv[n - 2];
      for (int i = n - 3; i >= 0; --i) z[i] = (z[i] - u1[i] * z[i + 1] - u2[i] * z[i + 2]) * rinv[i];
    }
    __syncthreads();
    double nn = 0.0;
    for (int i = lane; i < n; i += 64) nn += z[i] * z[i];
    const double inv = 1.0 / sqrt(wave_sum(nn));
    for (int i = lane; i < n; i += 64) z[i] *= inv;
    __syncthreads();
  }
  if (lane == 0) {
    // fix the sign: largest component positive (deterministic output)
    int im = 0;
    for (int i = 1; i < n; ++i) if (fabs(z[i]) > fabs(z[im])) im = i;
    shr[0] = z[im] < 0.0 ? -1.0 : 1.0;
    st[0] = theta;
    st[1] = has_bres ? fabs(bres_v * z[n - 1]) : 0.0;
    st[2] = tnorm;
  }
  __syncthreads();
  const double sg = shr[0];
  for (int i = lane; i < n; i += 64) s[i] = sg * z[i];
  chunk_mark(st, n, ctol);
}

// ---------------------------------------------------------------------------
// Dense symmetric tridiagonalisation (one workgroup) and back-transform.
// A: n x n symmetric in global workspace (overwritten), n <= 256.
// Householder reflectors H_k = I - tau_k v_k v_k^T act on indices k+1..n-1.
// ---------------------------------------------------------------------------
__device__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

__global__ __launch_bounds__(256) void sytrd_kernel(double* __restrict__ A, int n,
                                                   double* __restrict__ Vst,
                                                   double* __restrict__ tau,
                                                   double* __restrict__ dg,
                                                   double* __restrict__ eg, PMap pm) {
  {
    const int64_t o = pm.off();
    A += o; Vst += o; tau += o; dg += o; eg += o;
  }
  __shared__ double v[kTriMax], p[kTriMax];
  __shared__ double red[8];
  const int t = threadIdx.x;
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    const double alpha = A[(int64_t)(k + 1) * n + k];
    double part = 0.0;
    for (int i = 1 + t; i < m; i += blockDim.x) {
      const double x = A[(int64_t)(k + 1 + i) * n + k];
      part += x * x;
    }
    const double sigma = block_sum(part, red);
    double tk, beta;
    if (sigma == 0.0) {
      tk = 0.0; beta = alpha;
      for (int i = t; i < m; i += blockDim.x) v[i] = (i == 0) ? 1.0 : 0.0;
    } else {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tk = (beta - alpha) / beta;
      const double sc = 1.0 / (alpha - beta);
      for (int i = t; i < m; i += blockDim.x) v[i] = (i == 0) ? 1.0 : A[(int64_t)(k + 1 + i) * n + k] * sc;
    }
    __syncthreads();
    for (int i = t; i < m; i += blockDim.x) Vst[(int64_t)k * n + i] = v[i];
    if (t == 0) { tau[k] = tk; eg[k] = beta; }
    if (tk != 0.0) {
      // p = tau * A22 v
      for (int i = t; i < m; i += blockDim.x) {
        const double* row = A + (int64_t)(k + 1 + i) * n + (k + 1);
        double acc = 0.0;
        for (int j = 0; j < m; ++j) acc += row[j] * v[j];
        p[i] = tk * acc;
      }
      __syncthreads();
      double pv = 0.0;
      for (int i = t; i < m; i += blockDim.x) pv += p[i] * v[i];
      const double K = 0.5 * tk * block_sum(pv, red);
      for (int i = t; i < m; i += blockDim.x) p[i] -= K * v[i];  // p <- w
      __syncthreads();
      for (int i = t; i < m; i += blockDim.x) {
        double* row = A + (int64_t)(k + 1 + i) * n + (k + 1);
        const double vi = v[i], wi = p[i];
        for (int j = 0; j < m; ++j) row[j] -= vi * p[j] + wi * v[j];
      }
    }
    __syncthreads();
  }
  for (int i = t; i < n; i += blockDim.x) dg[i] = A[(int64_t)i * n + i];
  if (t == 0 && n >= 2) eg[n - 2] = A[(int64_t)(n - 1) * n + (n - 2)];
}

__global__ __launch_bounds__(256) void ormtr_kernel(const double* __restrict__ Vst,
                                                   const double* __restrict__ tau, int n,
                                                   const double* __restrict__ zin,
                                                   double* __restrict__ x, PMap pm, int64_t so) {
  {
    const int64_t o = pm.off();
    Vst += o; tau += o; zin += o;
    x += (int64_t)pm.pid() * so;
  }
  __shared__ double xs[kTriMax];
  __shared__ double red[8];
  const int t = threadIdx.x;
  for (int i = t; i < n; i += blockDim.x) xs[i] = zin[i];
  __syncthreads();
  for (int k = n - 3; k >= 0; --k) {
    const double tk = tau[k];
    if (tk == 0.0) continue;
    const int m = n - k - 1;
    double part = 0.0;
    for (int i = t; i < m; i += blockDim.x) part += Vst[(int64_t)k * n + i] * xs[k + 1 + i];
    const double dot = block_sum(part, red);
    for (int i = t; i < m; i += blockDim.x) xs[k + 1 + i] -= tk * Vst[(int64_t)k * n + i] * dot;
    __syncthreads();
  }
  for (int i = t; i < n; i += blockDim.x) x[i] = xs[i];
}

size_t syev_ws_bytes(int n) {
  const size_t nn = (size_t)n * n;
  return (2 * nn + 6 * (size_t)n + 16) * 8 + 256;
}

// Lowest eigenpair of a dense symmetric n x n matrix (device). eval written
// to *eval_dev (device) or copied to host if eval_host is given.
int syev_lowest(const double* M, int n, double* evec, double* eval_host, void* ws,
                size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(n >= 1 && n <= kTriMax, "syev_lowest: n=%d outside [1, %d]", n, kTriMax);
  TNX_CHECK_ARG(M && evec, "syev_lowest: null pointer");
  TNX_CHECK_ARG(ws && ws_bytes >= syev_ws_bytes(n), "syev_lowest: workspace too small");
  double* A = (double*)ws;
  double* Vst = A + (size_t)n * n;
  double* tau = Vst + (size_t)n * n;
  double* dg = tau + n;
  double* eg = dg + n;
  double* z = eg + n;
  double* st = z + n + 8;
  TNX_HIP(hipMemcpyAsync(A, M, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  if (n == 1) {
    TNX_TRY(fill_f64(evec, 1, 1.0, s));
    if (eval_host) {
      TNX_HIP(hipMemcpyAsync(eval_host, M, sizeof(double), hipMemcpyDeviceToHost, s));
      TNX_HIP(hipStreamSynchronize(s));
    }
    return TNX_OK;
  }
  hipLaunchKernelGGL(sytrd_kernel, dim3(1), dim3(256), 0, s, A, n, Vst, tau, dg, eg, kOne);
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(tridiag_lowest_kernel, dim3(1), dim3(64), 0, s, dg, eg, n,
                     (const double*)nullptr, z, st, kOne, 0, TriFinish{}, -1.0);
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(ormtr_kernel, dim3(1), dim3(256), 0, s, Vst, tau, n, z, evec, kOne,
                     (int64_t)0);
  TNX_LAUNCH_CHECK();
  if (eval_host) {
    TNX_HIP(hipMemcpyAsync(eval_host, st, sizeof(double), hipMemcpyDeviceToHost, s));
    TNX_HIP(hipStreamSynchronize(s));
  }
  return TNX_OK;
}

// ---------------------------------------------------------------------------
// All eigenpairs of a symmetric tridiagonal matrix (n <= 64): implicit QL
// with Wilkinson shifts, one wavefront. Every lane runs the same scalar
// recurrence on the LDS copies of the diagonal / off-diagonal and writes the
// same values to them, so each lane only ever reads its own writes and the
// recurrence needs no barrier; lane k owns row k of the eigenvector matrix Z
// and applies each Givens rotation to it. Eigenvalues ascending to w, eigenvectors as columns
// of S (row stride ldS). Used at thick restarts, where the projected matrix
// is tridiagonal.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void tridiag_all_kernel(const double* __restrict__ ad,
                                                         const double* __restrict__ bd, int n,
                                                         double* __restrict__ w,
                                                         double* __restrict__ S, int ldS,
                                                         PMap pm) {
  {
    const int64_t o = pm.off();
    ad += o; bd += o; w += o; S += o;
  }
  // Implicit QL with eigenvectors (n <= 64). d[i], e[i] live in lane i's
  // registers: the rotation chain runs on all lanes as wave-uniform scalars,
  // reading a row with v_readlane and writing lane i's value with a lane
  // select, so no step waits on LDS. Row k of the eigenvector matrix Z is
  // private to lane k (in LDS); within a sweep the lane carries the freshly
  // rotated element Z[k][i] in a register to the next step, so its LDS row
  // sees only reads of untouched columns and stores nobody waits for.
  __shared__ double Z[64][65];
  __shared__ double dsh[64];
  const int k = threadIdx.x;
  double dr = k < n ? ad[k] : 0.0;
  double er = k < n - 1 ? bd[k] : 0.0;
  if (k < n)
    for (int c = 0; c < n; ++c) Z[k][c] = (k == c) ? 1.0 : 0.0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 60; ++iter) {
      // m = first index >= l with a negligible e[m] (n - 1 if none)
      const double dnext = __shfl_down(dr, 1, 64);  // every lane takes part
      const bool negl = k >= l && k < n - 1 && fabs(er) <= eps * (fabs(dr) + fabs(dnext));
      const unsigned long long mk = __ballot(negl);
      const int m = mk ? __ffsll((long long)mk) - 1 : n - 1;
      if (m == l) break;
      const double el = lane_val(er, l), dl = lane_val(dr, l);
      double g = (lane_val(dr, l + 1) - dl) / (2.0 * el);
      double r = hypot(g, 1.0);
      g = lane_val(dr, m) - dl + el / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double sn = 1.0, cs = 1.0, p = 0.0;
      bool early = false;
      double zc = 0.0;  // rotated Z[k][hidx], not yet stored
      int hidx = -1;    // -1: nothing held
      for (int i = m - 1; i >= l; --i) {
        const double ei = lane_val(er, i);
        double f = sn * ei;
        const double b = cs * ei;
        // |f|, |g| <= ||T||: no scaling needed for the rotation length
        r = sqrt(fma(f, f, g * g));
        if (k == i + 1) er = r;
        if (r == 0.0) {
          if (k == i + 1) dr -= p;
          if (k == m) er = 0.0;
          early = true;
          break;
        }
        const double ir = 1.0 / r;
        sn = f * ir;
        cs = g * ir;
        g = lane_val(dr, i + 1) - p;
        r = (lane_val(dr, i) - g) * sn + 2.0 * cs * b;
        p = sn * r;
        if (k == i + 1) dr = g + p;
        g = cs * r - b;
        if (k < n) {
          const double zi1 = hidx == i + 1 ? zc : Z[k][i + 1];
          const double zi = Z[k][i];
          Z[k][i + 1] = sn * zi + cs * zi1;
          zc = cs * zi - sn * zi1;
        }
        hidx = i;
      }
      // store the held element: Z[k][l] after a full sweep, Z[k][i+1] when
      // the sweep stopped at step i
      if (k < n && hidx >= 0) Z[k][hidx] = zc;
      if (early) continue;
      if (k == l) { dr -= p; er = g; }
      if (k == m) er = 0.0;
    }
  }
  if (k < n) dsh[k] = dr;
  __syncthreads();
  if (k < n) {
    const double dk = dr;
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (dsh[j] < dk) || (dsh[j] == dk && j < k);
    w[rank] = dk;
    for (int row = 0; row < n; ++row) S[row * ldS + rank] = Z[row][k];
  }
}

// ---------------------------------------------------------------------------
// Lanczos driver
// ---------------------------------------------------------------------------
struct KWork {
  double* V;      // (m+1) x N basis
  double* Y;      // keep x N restart vectors
  double* w;      // N
  double* y;      // N
  void* mv;       // matvec workspace
  size_t mv_bytes;
  double* partials;
  double* h1; double* h2; double* h3;  // reductions (kPartStride each)
  double* alpha; double* beta;         // tridiagonal part, kMaxBasis + 2 each
  double* T;                           // dense projected matrix, kMaxBasis^2
  double* S;                           // its eigenvectors, kMaxBasis^2
  double* Cm;                          // restart combination, kMaxBasis^2
  double* theta;                       // its eigenvalues, kMaxBasis
  double* s;                           // Ritz coefficients of the lowest pair
  double* st;                          // status (16)
  double* scal;                        // scratch scalars (16)
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static int keep_of(int m, int keep) {
  int k = keep > 0 ? keep : m / 4 + 2;
  return std::max(1, std::min(k, m - 2));
}

static size_t eigsh_ws_bytes_n(size_t N, size_t mvb, int max_basis) {
  const int m = std::max(3, std::min(max_basis, kMaxBasis));
  size_t b = 0;
  b += align256((size_t)(m + 1) * N * 8);
  b += align256((size_t)(m - 2) * N * 8);
  b += 2 * align256(N * 8);
  b += align256(mvb);
  b += align256((size_t)kCgsBlocks * kPartStride * 8);
  b += 3 * align256(kPartStride * 8);
  b += 3 * align256((kMaxBasis + 2) * 8);
  b += 3 * align256((size_t)kMaxBasis * kMaxBasis * 8);
  b += align256((kMaxBasis + 2) * 8);
  b += 2 * align256(16 * 8);
  return b + 256;
}

// Single-problem solver: the Lanczos block plus the FP64 correction block
// of the mixed-precision path (W = H V for up to m vectors, and the small
// arrays of its dense Rayleigh-Ritz).
static size_t gd_ws_bytes(size_t N, int m) {
  return align256((size_t)m * N * 8) + align256((size_t)8 * (kMaxBasis + 8) * 8);
}

static size_t eigsh_single_ws_bytes(size_t N, size_t mvb, int max_basis, size_t env32_bytes = 0) {
  const int m = std::max(3, std::min(max_basis, kMaxBasis));
  return eigsh_ws_bytes_n(N, mvb, max_basis) + gd_ws_bytes(N, m) + align256(env32_bytes);
}

// FP32 copies of L and R for the low-precision cycles of a one-site solve
static size_t env32_bytes_of(int chil, int Dl, int Dr, int chir) {
  return align256((size_t)chil * Dl * chil * 4) + align256((size_t)chir * Dr * chir * 4);
}

size_t eigsh_ws_bytes(int chil, int Dl, int Dr, int d, int chir, int max_basis) {
  return eigsh_single_ws_bytes((size_t)chil * d * chir, heff1_ws_bytes(chil, Dl, Dr, d, chir),
                               max_basis, env32_bytes_of(chil, Dl, Dr, chir));
}

size_t eigsh2_ws_bytes(int chil, int Dl, int Dm, int Dr, int d, int chir, int max_basis) {
  return eigsh_single_ws_bytes((size_t)chil * d * d * chir,
                               heff2_ws_bytes(chil, Dl, Dm, Dr, d, chir), max_basis);
}

static void carve(KWork& k, void* ws, int m, size_t N, size_t mvb) {
  char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  auto take = [&](size_t bytes) { char* r = p; p += align256(bytes); return r; };
  k.V = (double*)take((size_t)(m + 1) * N * 8);
  k.Y = (double*)take((size_t)(m - 2) * N * 8);
  k.w = (double*)take(N * 8);
  k.y = (double*)take(N * 8);
  k.mv = take(mvb); k.mv_bytes = mvb;
  k.partials = (double*)take((size_t)kCgsBlocks * kPartStride * 8);
  k.h1 = (double*)take(kPartStride * 8);
  k.h2 = (double*)take(kPartStride * 8);
  k.h3 = (double*)take(kPartStride * 8);
  k.alpha = (double*)take((kMaxBasis + 2) * 8);
  k.beta = (double*)take((kMaxBasis + 2) * 8);
  k.s = (double*)take((kMaxBasis + 2) * 8);
  k.T = (double*)take((size_t)kMaxBasis * kMaxBasis * 8);
  k.S = (double*)take((size_t)kMaxBasis * kMaxBasis * 8);
  k.Cm = (double*)take((size_t)kMaxBasis * kMaxBasis * 8);
  k.theta = (double*)take((kMaxBasis + 2) * 8);
  k.st = (double*)take(16 * 8);
  k.scal = (double*)take(16 * 8);
}

// H y = theta y + (beta_j s_j) V[j+1] for the lowest Ritz pair:
// scal[0] = theta, scal[1] = beta_j * s_j
__global__ void ritz_scalars_kernel(const double* st, const double* s, const double* beta, int j,
                                    double* scal, PMap pm) {
  {
    const int64_t o = pm.off();
    st += o; s += o; beta += o; scal += o;
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    scal[0] = st[0];
    scal[1] = beta[j] * s[j];
  }
}

// Thick restart that keeps the projected matrix tridiagonal. The kept Ritz
// pairs (theta_i, Y_i = V S[:, i]), i < k, satisfy H Y_i = theta_i Y_i +
// c_i v_next with c_i = beta_j S[j][i]. A Lanczos run on diag(theta) started
// from c / |c| gives an orthonormal Q (k x k') in which diag(theta) is
// tridiagonal and only its first vector couples to v_next (with |c|). The new
// basis is the reversed Q columns, so the coupling sits on the last kept
// vector and the Lanczos recurrence continues from V[k'] = v_next.
// Writes C (J x k', ld kMaxBasis) with V_new = V_old C, the new alpha/beta
// (beta[k'-1] = |c|), the dense T, and k' to *kp_out. One wavefront.
__global__ __launch_bounds__(64) void tr_restart_kernel(const double* __restrict__ theta,
                                                        const double* __restrict__ S,
                                                        const double* __restrict__ beta_in, int j,
                                                        int k, int J, double* __restrict__ Cm,
                                                        double* __restrict__ alpha,
                                                        double* __restrict__ beta,
                                                        double* __restrict__ T, int* kp_out,
                                                        PMap pm) {
  {
    const int64_t o = pm.off();
    theta += o; S += o; beta_in += o; Cm += o; alpha += o; beta += o; T += o;
    kp_out = reinterpret_cast<int*>(reinterpret_cast<double*>(kp_out) + o);
  }
  __shared__ double q[kMaxBasis][kMaxBasis];
  __shared__ double a[kMaxBasis], b[kMaxBasis];
  const int l = threadIdx.x;  // component index of the k-vectors
  const double bj = beta_in[j];
  const double cl = (l < k) ? bj * S[j * kMaxBasis + l] : 0.0;
  const double th = (l < k) ? theta[l] : 0.0;
  double tmax = fabs(th);
  for (int o = 32; o > 0; o >>= 1) tmax = fmax(tmax, __shfl_xor(tmax, o, 64));
  const double cn = sqrt(wave_sum(cl * cl));
  double qi = (cn > 0.0) ? cl / cn : (l == 0 ? 1.0 : 0.0);
  if (l < kMaxBasis) q[0][l] = (l < k) ? qi : 0.0;
  int kp = 1;
  double qprev = 0.0, bprev = 0.0;
  for (int i = 0; i < k; ++i) {
    double w = th * qi;
    const double ai = wave_sum(qi * w);
    w -= ai * qi + bprev * qprev;
    for (int pass = 0; pass < 2; ++pass)
      for (int t = 0; t <= i; ++t) {
        const double h = wave_sum(q[t][l] * w);
        w -= h * q[t][l];
      }
    if (l < k) {} else w = 0.0;
    if (l == 0) a[i] = ai;
    const double bn = sqrt(wave_sum(w * w));
    kp = i + 1;
    if (i == k - 1) break;
    if (bn > 1e-13 * fmax(tmax, 1e-300)) {
      if (l == 0) b[i] = bn;
      qprev = qi;
      bprev = bn;
      qi = w / bn;
    } else {
      // span(q[0..i]) is invariant under diag(theta): continue in its
      // complement from the unit vector e_t with the largest component
      // there (zero coupling b[i]), so exactly k vectors are kept.
      double best = -1.0, qb = 0.0;
      for (int t = 0; t < k; ++t) {
        double u = (l == t) ? 1.0 : 0.0;
        for (int pass = 0; pass < 2; ++pass)
          for (int r = 0; r <= i; ++r) {
            const double h = wave_sum(q[r][l] * u);
            u -= h * q[r][l];
          }
        if (l >= k) u = 0.0;
        const double un = sqrt(wave_sum(u * u));
        if (un > best) { best = un; qb = u / un; }
      }
      if (l == 0) b[i] = 0.0;
      qprev = 0.0;
      bprev = 0.0;
      qi = qb;
    }
    q[i + 1][l] = qi;
  }
  __syncthreads();
  // reversed order: new index r holds q[kp-1-r]
  for (int r = 0; r < kp; ++r) {
    const int i = kp - 1 - r;
    for (int row = l; row < J; row += 64) {
      double acc = 0.0;
      for (int t = 0; t < k; ++t) acc += S[row * kMaxBasis + t] * q[i][t];
      Cm[row * kMaxBasis + r] = acc;
    }
  }
  for (int e = l; e < kMaxBasis * kMaxBasis; e += 64) T[e] = 0.0;
  __syncthreads();
  if (l == 0) {
    for (int r = 0; r < kp; ++r) {
      const int i = kp - 1 - r;
      alpha[r] = a[i];
      T[r * kMaxBasis + r] = a[i];
      const double br = (r < kp - 1) ? b[i - 1] : cn;
      beta[r] = br;
      T[r * kMaxBasis + r + 1] = br;
      T[(r + 1) * kMaxBasis + r] = br;
    }
    *kp_out = kp;
  }
}

// Outer (FP64) step scalars: theta = y.(Hy) = h1[0] + h2[0] (both CGS
// passes), rho^2 = h3[1] (norm of r). Resets the projected matrix to the
// two-vector basis [y, r/rho].
__global__ void outer_scalars_kernel(const double* h1, const double* h2, const double* h3,
                                     double* alpha, double* beta, double* scal, double* st,
                                     double* T, PMap pm) {
  {
    const int64_t o = pm.off();
    h1 += o; h2 += o; h3 += o; alpha += o; beta += o; scal += o; st += o; T += o;
  }
  const double th = h1[0] + h2[0];
  const double rho = sqrt(fmax(h3[1], 0.0));
  for (int e = threadIdx.x; e < kMaxBasis * kMaxBasis; e += blockDim.x) {
    const int i = e / kMaxBasis, c = e % kMaxBasis;
    T[e] = (i == 0 && c == 0) ? th : ((i + c == 1) ? rho : 0.0);
  }
  if (threadIdx.x == 0) {
    alpha[0] = th; beta[0] = rho;
    scal[0] = th; scal[2] = -th; scal[3] = 1.0;
    st[4] = th; st[5] = rho;
  }
}

// The operator of one eigensolve: the one-site H_eff (W2 == nullptr) or
// the two-site H_eff (W, W2 with middle MPO bond Dm).
struct EigCtx {
  const double *L, *W, *R;
  const double* W2 = nullptr;
  int chil, Dl, Dm = 0, Dr, d, chir;
  int64_t N;
  hipStream_t s;
  KWork k;
  int matvecs = 0, matvecs64 = 0;
  double* Wd = nullptr;   // FP64 correction: H V (m x N)
  double* gsm = nullptr;  // its small arrays (8 x (kMaxBasis + 8))
  // one-site problems: L and R rounded to FP32 at most once per solve (at
  // its first low-precision cycle), read by the low-precision matvecs (the
  // same values their tile staging would round to, half the bytes, no
  // conversion per matvec); null until converted and for FP64 solves
  const float* L32 = nullptr;
  const float* R32 = nullptr;
  float* L32_buf = nullptr;  // their workspace (set for low-precision solves)
  float* R32_buf = nullptr;
  size_t env32_bytes = 0;  // workspace reserved for them
};

static int apply_h(EigCtx& c, const void* x, int x_f32, double* y, int prec) {
  c.matvecs++;
  if (prec == TNX_F64) c.matvecs64++;
  if (c.W2)
    return heff2_matvec(c.L, c.W, c.W2, c.R, x, x_f32, y, c.chil, c.Dl, c.Dm, c.Dr, c.d, c.chir,
                        prec, c.k.mv, c.k.mv_bytes, c.s);
  if (prec != TNX_F64 && c.L32)
    return heff1_matvec_b(c.L32, c.W, c.R32, x, x_f32, y, c.chil, c.Dl, c.Dr, c.d, c.chir, prec, 1,
                          1, nullptr, c.N, c.N, c.k.mv, c.k.mv_bytes, c.s, 1);
  return heff1_matvec_b(c.L, c.W, c.R, x, x_f32, y, c.chil, c.Dl, c.Dr, c.d, c.chir, prec, 1, 1,
                        nullptr, c.N, c.N, c.k.mv, c.k.mv_bytes, c.s, 0);
}

static int read_status(EigCtx& c, double* host, int n) {
  TNX_HIP(hipMemcpyAsync(host, c.k.st, sizeof(double) * n, hipMemcpyDeviceToHost, c.s));
  TNX_HIP(hipStreamSynchronize(c.s));
  return TNX_OK;
}

// FP64 outer step on the unit vector y (in k.y): w = H y, theta = y.w,
// r = w - theta y (in k.w), rho = |r|. Sets alpha[0] = theta, beta[0] = rho
// and T = [[theta, rho], [rho, .]]. Returns theta, rho on the host.
static int outer_step(EigCtx& c, double* theta, double* rho) {
  KWork& k = c.k;
  const int64_t N = c.N;
  TNX_TRY(apply_h(c, k.y, 0, k.w, TNX_F64));
  TNX_TRY(cgs_pass(k.y, N, 1, k.w, nullptr, true, k.partials, k.h1, N, c.s));
  TNX_TRY(cgs_pass(k.y, N, 1, k.w, k.h1, true, k.partials, k.h2, N, c.s));
  TNX_TRY(cgs_pass(k.y, N, 1, k.w, k.h2, false, k.partials, k.h3, N, c.s));
  hipLaunchKernelGGL(outer_scalars_kernel, dim3(1), dim3(256), 0, c.s, k.h1, k.h2, k.h3,
                     k.alpha, k.beta, k.scal, k.st, k.T, kOne);
  TNX_LAUNCH_CHECK();
  double h[8];
  TNX_TRY(read_status(c, h, 8));
  *theta = h[4];
  *rho = h[5];
  return TNX_OK;
}

static int eigsh_run(EigCtx& c, const double* v0, double* vec, double* hvec,
                     const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                     size_t ws_bytes, size_t mvb);

// ---------------------------------------------------------------------------
// FP64 correction after a low-precision Lanczos cycle (mixed precision).
// A cycle that has reached its arithmetic's floor holds, in its kept Ritz
// vectors, the cluster of low eigenvalues its Krylov space has resolved. A
// fresh FP64 Lanczos from [y, r] would have to resolve that cluster again.
// Instead: take the `keep` lowest Ritz vectors of the cycle, orthonormalise
// them in FP64, apply H to each in FP64 (W = H V), and run a subspace
// expansion with an exact dense projection G = V^T H V: each step takes
// the lowest Ritz pair of G, its FP64 residual r = W s - theta V s (exact,
// no recurrence estimate), stops when |r| <= tol ||H||_est, else adds
// r (orthogonalised) to V and H r to W. (Without a preconditioner this
// spans the same space as Lanczos from the kept block, but keeps the block
// exact in FP64.)
// ---------------------------------------------------------------------------

// z = V s, hz = W s, r = hz - theta z (theta = *st), J vectors each
__global__ void gd_residual_kernel(const double* __restrict__ V, const double* __restrict__ Wm,
                                   int64_t ld, int J, const double* __restrict__ sv,
                                   const double* __restrict__ st, double* __restrict__ z,
                                   double* __restrict__ hz, double* __restrict__ r, int64_t N) {
  __shared__ double cs[kMaxBasis + 1];
  if ((int)threadIdx.x < J) cs[threadIdx.x] = sv[threadIdx.x];
  __syncthreads();
  const double th = st[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j < J; ++j) {
      a += cs[j] * V[(int64_t)j * ld + i];
      b += cs[j] * Wm[(int64_t)j * ld + i];
    }
    z[i] = a;
    hz[i] = b;
    r[i] = b - th * a;
  }
}

// column and row J of the dense projected matrix: G[i][J] = G[J][i] = h[i]
__global__ void gd_gcol_kernel(const double* __restrict__ h, int J, double* __restrict__ G) {
  const int i = threadIdx.x;
  if (i <= J) {
    G[i * kMaxBasis + J] = h[i];
    G[J * kMaxBasis + i] = h[i];
  }
}

// compact symmetric copy A[J][J] of G (row stride kMaxBasis)
__global__ void gd_compact_kernel(const double* __restrict__ G, int J, double* __restrict__ A) {
  for (int e = threadIdx.x; e < J * J; e += blockDim.x) {
    const int i = e / J, j = e % J;
    A[e] = 0.5 * (G[i * kMaxBasis + j] + G[j * kMaxBasis + i]);
  }
}

// Returns TNX_OK; *done says whether the criterion was met (then k.y = z,
// hvec = H z, *theta and *rho set). If not, k.y holds the best Ritz vector
// and the caller continues with FP64 Lanczos from it.
static int gd_correction(EigCtx& c, const float* V32, int J32, int keep, int m, double tol,
                         double* anorm, int max_mv, double* hvec, double* theta, double* rho,
                         bool* done) {
  KWork& k = c.k;
  hipStream_t s = c.s;
  const int64_t N = c.N;
  const int nb = vec_blocks(N);
  *done = false;
  // kept Ritz values of the cycle (ascending); the block is the cluster
  // within 0.1% of ||H||_est of the lowest (at least 2, at most `keep`): a
  // gapped spectrum needs a couple of FP64 matvecs here, a clustered one
  // keeps the whole cluster
  hipLaunchKernelGGL(tridiag_all_kernel, dim3(1), dim3(64), 0, s, k.alpha, k.beta, J32, k.theta,
                     k.S, kMaxBasis, kOne);
  TNX_LAUNCH_CHECK();
  std::vector<double> th(J32);
  TNX_HIP(hipMemcpyAsync(th.data(), k.theta, sizeof(double) * J32, hipMemcpyDeviceToHost, s));
  TNX_HIP(hipStreamSynchronize(s));
  int kk = 1;
  while (kk < std::min(keep, J32) && th[kk] - th[0] <= 1e-3 * *anorm) ++kk;
  kk = std::min(std::max(kk, 2), std::min(keep, J32));
  double* A = k.S;        // compact copy of G for the dense solve
  double* Vst = k.Cm;     // its Householder vectors
  double* tau = c.gsm;
  double* dg = tau + (kMaxBasis + 8);
  double* eg = dg + (kMaxBasis + 8);
  double* zt = eg + (kMaxBasis + 8);
  double* stg = zt + (kMaxBasis + 8);
  double* G = k.T;
  // kept Ritz vectors of the low-precision cycle, in FP64: Y = V32 S[:, :kk]
  {
    GemmDesc g{};
    g.M = kk; g.N = (int)N; g.K = J32; g.batch = 1;
    g.A = k.S; g.ta = DT::F64; g.lda = kMaxBasis; g.sA = 0; g.a_kcontig = false;
    g.B = V32; g.tb = DT::F32; g.ldb = N; g.sB = 0; g.b_ncontig = true;
    g.C = k.Y; g.tc = DT::F64; g.ldc = N; g.sC = 0;
    g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
    TNX_TRY(gemm(g, s));
  }
  double* V = k.V;  // the FP64 basis of the correction (V32 is no longer needed)
  TNX_HIP(hipMemcpyAsync(V, k.Y, sizeof(double) * kk * N, hipMemcpyDeviceToDevice, s));
  // orthonormalise (CGS2) and apply H in FP64, building G column by column
  for (int i = 0; i < kk; ++i) {
    double* vi = V + (int64_t)i * N;
    if (i > 0) {
      TNX_TRY(cgs_pass(V, N, i, vi, nullptr, true, k.partials, k.h1, N, s));
      TNX_TRY(cgs_pass(V, N, i, vi, k.h1, true, k.partials, k.h2, N, s));
    }
    TNX_TRY(cgs_pass(V, N, i, vi, i > 0 ? k.h2 : nullptr, false, k.partials, k.h3, N, s));
    hipLaunchKernelGGL(normalize_kernel<double>, dim3(nb), dim3(256), 0, s, vi, k.h3 + i, vi, N,
                       kOne);
    TNX_LAUNCH_CHECK();
    double* wi = c.Wd + (int64_t)i * N;
    TNX_TRY(apply_h(c, vi, 0, wi, TNX_F64));
    TNX_TRY(cgs_pass(V, N, i + 1, wi, nullptr, true, k.partials, k.h1, N, s));
    hipLaunchKernelGGL(gd_gcol_kernel, dim3(1), dim3(kMaxBasis + 1), 0, s, k.h1, i, G);
    TNX_LAUNCH_CHECK();
  }
  int J = kk;
  while (true) {
    // lowest Ritz pair of G (J x J): Householder tridiagonalisation,
    // multisection + inverse iteration, back-transform into k.s
    hipLaunchKernelGGL(gd_compact_kernel, dim3(1), dim3(256), 0, s, G, J, A);
    TNX_LAUNCH_CHECK();
    if (J == 1) {
      TNX_TRY(fill_f64(k.s, 1, 1.0, s));
      TNX_HIP(hipMemcpyAsync(stg, A, sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
      hipLaunchKernelGGL(sytrd_kernel, dim3(1), dim3(256), 0, s, A, J, Vst, tau, dg, eg, kOne);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(tridiag_lowest_kernel, dim3(1), dim3(64), 0, s, dg, eg, J,
                         (const double*)nullptr, zt, stg, kOne, 0, TriFinish{}, -1.0);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(ormtr_kernel, dim3(1), dim3(256), 0, s, Vst, tau, J, zt, k.s, kOne,
                         (int64_t)0);
      TNX_LAUNCH_CHECK();
    }
    // z = V s -> k.y, H z = W s -> k.w, r -> V[J] (the next basis slot)
    double* r = V + (int64_t)J * N;
    hipLaunchKernelGGL(gd_residual_kernel, dim3(nb), dim3(256), 0, s, V, c.Wd, N, J, k.s, stg,
                       k.y, k.w, r, N);
    TNX_LAUNCH_CHECK();
    TNX_TRY(cgs_pass(r, N, 0, r, nullptr, false, k.partials, k.h3, N, s));
    double h[2];
    TNX_HIP(hipMemcpyAsync(&h[0], stg, sizeof(double), hipMemcpyDeviceToHost, s));
    TNX_HIP(hipMemcpyAsync(&h[1], k.h3, sizeof(double), hipMemcpyDeviceToHost, s));
    TNX_HIP(hipStreamSynchronize(s));
    *theta = h[0];
    *rho = std::sqrt(std::max(h[1], 0.0));
    *anorm = std::max(*anorm, std::fabs(h[0]));
    if (*rho <= tol * *anorm || *rho == 0.0) {
      if (hvec) TNX_TRY(copy_f64(k.w, hvec, N, s));
      *done = true;
      return TNX_OK;
    }
    if (J >= m || c.matvecs >= max_mv) return TNX_OK;
    // expand: r orthogonalised against V, normalised, and H r
    TNX_TRY(cgs_pass(V, N, J, r, nullptr, true, k.partials, k.h1, N, s));
    TNX_TRY(cgs_pass(V, N, J, r, k.h1, true, k.partials, k.h2, N, s));
    TNX_TRY(cgs_pass(V, N, J, r, k.h2, false, k.partials, k.h3, N, s));
    hipLaunchKernelGGL(normalize_kernel<double>, dim3(nb), dim3(256), 0, s, r, k.h3 + J, r, N,
                       kOne);
    TNX_LAUNCH_CHECK();
    double* wj = c.Wd + (int64_t)J * N;
    TNX_TRY(apply_h(c, r, 0, wj, TNX_F64));
    TNX_TRY(cgs_pass(V, N, J + 1, wj, nullptr, true, k.partials, k.h1, N, s));
    hipLaunchKernelGGL(gd_gcol_kernel, dim3(1), dim3(kMaxBasis + 1), 0, s, k.h1, J, G);
    TNX_LAUNCH_CHECK();
    ++J;
  }
}

int eigsh_heff1(const double* L, const double* W, const double* R, const double* v0,
                double* vec, double* hvec, int chil, int Dl, int Dr, int d, int chir,
                const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(opts && info, "eigsh: null opts/info");
  TNX_CHECK_ARG(L && W && R && v0 && vec, "eigsh: null pointer");
  TNX_CHECK_ARG(chil > 0 && Dl > 0 && Dr > 0 && d > 0 && chir > 0, "eigsh: bad dims");
  EigCtx c;
  c.L = L; c.W = W; c.R = R;
  c.chil = chil; c.Dl = Dl; c.Dr = Dr; c.d = d; c.chir = chir;
  c.N = (int64_t)chil * d * chir;
  c.s = s;
  c.env32_bytes = env32_bytes_of(chil, Dl, Dr, chir);
  return eigsh_run(c, v0, vec, hvec, opts, info, ws, ws_bytes, heff1_ws_bytes(chil, Dl, Dr, d, chir));
}

int eigsh_heff2(const double* L, const double* W1, const double* W2, const double* R,
                const double* v0, double* vec, double* hvec, int chil, int Dl, int Dm, int Dr,
                int d, int chir, const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(opts && info, "eigsh2: null opts/info");
  TNX_CHECK_ARG(L && W1 && W2 && R && v0 && vec, "eigsh2: null pointer");
  TNX_CHECK_ARG(chil > 0 && Dl > 0 && Dm > 0 && Dr > 0 && d > 0 && chir > 0, "eigsh2: bad dims");
  EigCtx c;
  c.L = L; c.W = W1; c.W2 = W2; c.R = R;
  c.chil = chil; c.Dl = Dl; c.Dm = Dm; c.Dr = Dr; c.d = d; c.chir = chir;
  c.N = (int64_t)chil * d * d * chir;
  c.s = s;
  return eigsh_run(c, v0, vec, hvec, opts, info, ws, ws_bytes,
                   heff2_ws_bytes(chil, Dl, Dm, Dr, d, chir));
}

static int eigsh_run(EigCtx& c, const double* v0, double* vec, double* hvec,
                     const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                     size_t ws_bytes, size_t mvb) {
  TNX_CHECK_ARG(opts->precision >= TNX_F64 && opts->precision <= TNX_BF16X3, "eigsh: bad precision");
  TNX_CHECK_ARG(opts->tol >= 0.0, "eigsh: negative tol");
  const int m = std::max(3, std::min(opts->max_basis, kMaxBasis));
  TNX_CHECK_ARG(ws && ws_bytes >= eigsh_single_ws_bytes((size_t)c.N, mvb, m, c.env32_bytes),
                "eigsh: workspace too small");
  hipStream_t s = c.s;
  carve(c.k, ws, m, (size_t)c.N, mvb);
  {
    // the correction block follows the Lanczos block
    char* gp = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255) +
               (eigsh_ws_bytes_n((size_t)c.N, mvb, m) - 256);
    c.Wd = (double*)gp;
    c.gsm = (double*)(gp + align256((size_t)m * c.N * 8));
    if (c.env32_bytes && !c.W2 && opts->precision != TNX_F64) {
      // converted at the first low-precision cycle (a solve that converges
      // on its first FP64 step never reads them)
      c.L32_buf = (float*)(gp + gd_ws_bytes((size_t)c.N, m));
      c.R32_buf = (float*)((char*)c.L32_buf + align256((size_t)c.chil * c.Dl * c.chil * 4));
    }
  }
  auto env_to_f32 = [&]() -> int {
    if (c.L32 || !c.L32_buf) return TNX_OK;
    const int64_t nl = (int64_t)c.chil * c.Dl * c.chil, nr = (int64_t)c.chir * c.Dr * c.chir;
    hipLaunchKernelGGL(to_f32_kernel, dim3(vec_blocks(nl)), dim3(256), 0, s, c.L, c.L32_buf, nl);
    TNX_LAUNCH_CHECK();
    hipLaunchKernelGGL(to_f32_kernel, dim3(vec_blocks(nr)), dim3(256), 0, s, c.R, c.R32_buf, nr);
    TNX_LAUNCH_CHECK();
    c.L32 = c.L32_buf;
    c.R32 = c.R32_buf;
    return TNX_OK;
  };
  KWork& k = c.k;
  const int64_t N = c.N;
  const int prec = opts->precision;
  const int max_mv = opts->max_matvecs > 0 ? opts->max_matvecs : 10000;
  const double tol = opts->tol;
  const int nb = vec_blocks(N);
  // Basis size cannot exceed the dimension (Krylov space exhausts at N).
  const int mb = (int)std::min<int64_t>(m, std::max<int64_t>(N, 1));
  const int keep = keep_of(mb, opts->keep);

  // y = v0 / |v0|
  TNX_TRY(cgs_pass(v0, N, 0, const_cast<double*>(v0), nullptr, false, k.partials, k.h3, N, s));
  hipLaunchKernelGGL(normalize_kernel<double>, dim3(nb), dim3(256), 0, s, v0, k.h3, k.y, N, kOne);
  TNX_LAUNCH_CHECK();

  double theta = 0.0, rho = 0.0, anorm = 0.0;
  int restarts = 0;
  bool converged = false;
  bool have_hy = false;  // k.w holds r = H y - theta y for the final y
  // Krylov matvec precision of the current cycle: starts at `prec`, and
  // escalates to FP64 when an FP64-corrected cycle fails to cut the true
  // residual by 4x (the low-precision operator cannot resolve the local
  // spectrum, e.g. clusters narrower than its rounding).
  int cprec = prec;
  double rho_prev = 0.0;
  bool first_cycle = true;
  // Step pipelining: the next step's matvec and CGS passes (they need only
  // V[j+1]) are queued before the host reads step j's status, so the GPU
  // works while the host decides. When step j ends the cycle, that queued
  // step is discarded (its buffers are rewritten before any later use) and
  // not counted.
  StepEvent sev;
  PinnedLease pin_lease(64);
  double* pin = pin_lease.get();
  const bool pipelined = sev.e != nullptr && pin != nullptr;

  while (true) {
    TNX_TRY(outer_step(c, &theta, &rho));
    have_hy = true;
    anorm = std::max(anorm, std::fabs(theta));
    if (cprec != TNX_F64 && rho_prev > 0.0 && rho > 0.25 * rho_prev) cprec = TNX_F64;
    // After the first low-precision cycle (run to its arithmetic's floor), a
    // tolerance below that floor is reached in FP64: a further low-precision
    // cycle would restart the Krylov space for a target it cannot resolve.
    if (cprec != TNX_F64 && !first_cycle && tol < 8.0 * (cprec == TNX_F32 ? 5.96e-8 : 3.9e-6))
      cprec = TNX_F64;
    rho_prev = rho;
    const bool mixed = cprec != TNX_F64;
    if (rho <= tol * anorm || rho == 0.0 || N == 1) {
      converged = true;
      break;
    }
    if (c.matvecs >= max_mv) break;
    // Inner Lanczos cycle in `cprec`, from column j = 1 on V[0] = y,
    // V[1] = r / rho. Inner target. FP64: the final tolerance. Low
    // precision: the first cycle runs the full thick-restart Lanczos down to
    // the rounding floor of its arithmetic (it keeps its restart subspace
    // the whole way); later cycles start from [y, r/|r|] and only need to
    // cut the FP64 residual. A low-precision cycle keeps its basis in FP32:
    // it resolves nothing below FP32 rounding, and the CGS passes and the
    // matvec read half the bytes.
    const double rho_outer = rho;
    const double unit_lp = cprec == TNX_F32 ? 5.96e-8 : 3.9e-6;
    // A first low-precision cycle whose target is below its arithmetic's
    // floor runs on its own (low-precision operator) residual estimate down
    // to tol: the residual it reaches against H stays at the floor, but its
    // Krylov space keeps growing into the low spectrum (near-degenerate
    // states below the first converged Ritz value appear late), and its kept
    // block then goes to the FP64 correction.
    const bool gd_after = mixed && first_cycle && tol < 8.0 * unit_lp;
    const double inner_tol =
        !mixed ? tol * anorm
               : (first_cycle ? (gd_after ? tol : std::max(tol, 4.0 * unit_lp)) * anorm
                              : std::max(tol * anorm, std::min(1e-3 * rho_outer, 0.5 * rho_outer)));
    first_cycle = false;
    have_hy = false;  // the inner cycle reuses k.w
    int last_J = 0;   // Lanczos basis size when the cycle stopped
    auto cycle = [&](auto* V, auto* Yb) -> int {
      using VT = std::remove_pointer_t<decltype(V)>;
      constexpr bool f32 = std::is_same<VT, float>::value;
      if constexpr (f32) {
        hipLaunchKernelGGL(to_f32_kernel, dim3(nb), dim3(256), 0, s, k.y, V, N);
      } else {
        TNX_TRY(copy_f64(k.y, V, N, s));
      }
      hipLaunchKernelGGL(normalize_kernel<VT>, dim3(nb), dim3(256), 0, s, k.w, k.h3 + 1, V + N, N,
                         kOne);
      TNX_LAUNCH_CHECK();
      // Step jj's matvec and CGS passes: w = H V[jj] orthogonalised against
      // V[0..jj] (sums in h1, h2, h3).
      auto front = [&](int jj) -> int {
        if constexpr (f32) {
          if (c.L32) {
            // FP32 basis row and FP32 environments: the matvec's GEMMs stage
            // both operands with 16-byte loads of floats
            TNX_TRY(apply_h(c, V + (int64_t)jj * N, 1, k.w, cprec));
          } else {
            // FP64 environments: widen V[jj] into k.y (free until the
            // cycle's Ritz vector is formed) for 16-byte FP64 tile loads
            hipLaunchKernelGGL(to_f64_kernel, dim3(nb), dim3(256), 0, s, V + (int64_t)jj * N, k.y, N);
            TNX_LAUNCH_CHECK();
            TNX_TRY(apply_h(c, k.y, 0, k.w, cprec));
          }
        } else {
          TNX_TRY(apply_h(c, V + (int64_t)jj * N, 0, k.w, cprec));
        }
        const int JJ = jj + 1;
        TNX_TRY(cgs_pass(V, N, JJ, k.w, nullptr, true, k.partials, k.h1, N, s));
        TNX_TRY(cgs_pass(V, N, JJ, k.w, k.h1, true, k.partials, k.h2, N, s));
        TNX_TRY(cgs_pass(V, N, JJ, k.w, k.h2, false, k.partials, k.h3, N, s));
        return TNX_OK;
      };
      int j = 1;
      bool queued = false;  // step j's front already launched
      while (true) {
        const int J = j + 1;
        if (!queued) TNX_TRY(front(j));
        const int mv_now = c.matvecs, mv64_now = c.matvecs64;
        hipLaunchKernelGGL(tridiag_lowest_kernel, dim3(1), dim3(64), 0, s, k.alpha, k.beta, J,
                           k.beta + j, k.s, k.st, kOne, 1,
                           TriFinish{k.h1, k.h2, k.h3, k.alpha, k.beta, k.T}, -1.0);
        TNX_LAUNCH_CHECK();
        // V[j+1] = w / beta[j]
        hipLaunchKernelGGL(normalize_kernel<VT>, dim3(nb), dim3(256), 0, s, k.w, k.h3 + J,
                           V + (int64_t)(j + 1) * N, N, kOne);
        TNX_LAUNCH_CHECK();
        double h[4];
        // the next step goes ahead only where this one, not ending the
        // cycle, would be followed by it (no restart, not exhausted)
        const bool ahead = pipelined && J < mb && J < N && mv_now < max_mv;
        if (ahead) {
          TNX_HIP(hipMemcpyAsync(pin, k.st, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
          TNX_HIP(hipEventRecord(sev.e, s));
          TNX_TRY(front(j + 1));
          TNX_HIP(hipEventSynchronize(sev.e));
          for (int q = 0; q < 3; ++q) h[q] = pin[q];
        } else {
          TNX_TRY(read_status(c, h, 3));
        }
        const double theta_in = h[0];
        const double est = h[1];
        anorm = std::max(anorm, h[2]);
        const bool exhausted = (J >= N);
        // FP64 cycles test against the current ||H|| estimate (it grows with
        // the Ritz values seen); low-precision cycles against their fixed target
        const double itol = mixed ? inner_tol : tol * anorm;
        if (est <= itol || exhausted || mv_now >= max_mv) {
          if (ahead) {  // the queued step is not used
            c.matvecs = mv_now;
            c.matvecs64 = mv64_now;
          }
          last_J = J;
          // Ritz vector y = V[0..j] s
          hipLaunchKernelGGL(vcomb_kernel<VT>, dim3(nb), dim3(256), 0, s, V, N, J, k.s, k.y, N,
                             kOne);
          TNX_LAUNCH_CHECK();
          if constexpr (f32) {
            // an FP32 basis is orthonormal only to FP32 rounding: renormalise
            // y in FP64 before the FP64 step takes its Rayleigh quotient
            TNX_TRY(cgs_pass(k.y, N, 0, k.y, nullptr, false, k.partials, k.h3, N, s));
            hipLaunchKernelGGL(normalize_kernel<double>, dim3(nb), dim3(256), 0, s, k.y, k.h3, k.y,
                               N, kOne);
            TNX_LAUNCH_CHECK();
          }
          if constexpr (!f32) {
            if (!mixed && (est <= tol * anorm || exhausted)) {
              // FP64 Lanczos: H y = theta y + beta_j s_j V[j+1]
              hipLaunchKernelGGL(ritz_scalars_kernel, dim3(1), dim3(1), 0, s, k.st, k.s, k.beta, j,
                                 k.scal, kOne);
              TNX_LAUNCH_CHECK();
              if (hvec) {
                hipLaunchKernelGGL(axpbz_dev_kernel, dim3(nb), dim3(256), 0, s, k.scal, k.y,
                                   exhausted ? nullptr : k.scal + 1, V + (int64_t)(j + 1) * N,
                                   hvec, N, kOne, (int64_t)0);
                TNX_LAUNCH_CHECK();
              }
              theta = theta_in;
              rho = est;
              converged = true;
            }
          }
          return TNX_OK;
        }
        if (J >= mb) {
          // Thick restart keeping the `keep` lowest Ritz pairs in a basis
          // where the projected matrix stays tridiagonal.
          hipLaunchKernelGGL(tridiag_all_kernel, dim3(1), dim3(64), 0, s, k.alpha, k.beta, J,
                             k.theta, k.S, kMaxBasis, kOne);
          TNX_LAUNCH_CHECK();
          int* kp_dev = reinterpret_cast<int*>(k.st + 8);
          hipLaunchKernelGGL(tr_restart_kernel, dim3(1), dim3(64), 0, s, k.theta, k.S, k.beta, j,
                             keep, J, k.Cm, k.alpha, k.beta, k.T, kp_dev, kOne);
          TNX_LAUNCH_CHECK();
          int kp = 0;
          TNX_HIP(hipMemcpyAsync(&kp, kp_dev, sizeof(int), hipMemcpyDeviceToHost, s));
          TNX_HIP(hipStreamSynchronize(s));
          TNX_CHECK_ARG(kp == keep, "eigsh: restart kept %d of %d vectors", kp, keep);
          const DT vt = f32 ? DT::F32 : DT::F64;
          GemmDesc g{};
          g.M = kp; g.N = (int)N; g.K = J; g.batch = 1;
          g.A = k.Cm; g.ta = DT::F64; g.lda = kMaxBasis; g.sA = 0; g.a_kcontig = false;
          g.B = V; g.tb = vt; g.ldb = N; g.sB = 0; g.b_ncontig = true;
          g.C = Yb; g.tc = vt; g.ldc = N; g.sC = 0;
          g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
          TNX_TRY(gemm(g, s));
          TNX_HIP(hipMemcpyAsync(V, Yb, sizeof(VT) * kp * N, hipMemcpyDeviceToDevice, s));
          TNX_HIP(hipMemcpyAsync(V + (int64_t)kp * N, V + (int64_t)(j + 1) * N, sizeof(VT) * N,
                                 hipMemcpyDeviceToDevice, s));
          restarts++;
          j = kp;
          queued = false;
        } else {
          j++;
          queued = ahead;
        }
      }
    };
    if (mixed && N % 2 == 0) {
      TNX_TRY(env_to_f32());
      TNX_TRY(cycle(reinterpret_cast<float*>(k.V), reinterpret_cast<float*>(k.Y)));
    } else {
      TNX_TRY(cycle(k.V, k.Y));
    }
    if (gd_after && N % 2 == 0 && last_J > 0 && c.matvecs < max_mv) {
      bool done = false;
      TNX_TRY(gd_correction(c, reinterpret_cast<const float*>(k.V), last_J, keep, mb, tol, &anorm,
                            max_mv, hvec, &theta, &rho, &done));
      if (done) {
        converged = true;
        have_hy = false;  // outputs: k.y, and hvec written by the correction
        break;
      }
      cprec = TNX_F64;  // continue with FP64 Lanczos from the best vector
    }
    if (converged) break;
    if (c.matvecs >= max_mv) break;
    // loop: next outer step re-evaluates y in FP64 (mixed precision, or
    // FP64 when the matvec limit cut the cycle short)
  }

  // Outputs. After an outer step k.w holds r = H y - theta y.
  if (have_hy) {
    TNX_TRY(copy_f64(k.y, vec, N, s));
    if (hvec) {
      // H y = r + theta y  (scal[3] = 1, scal[0] = theta)
      hipLaunchKernelGGL(axpbz_dev_kernel, dim3(nb), dim3(256), 0, s, k.scal + 3, k.w, k.scal,
                         k.y, hvec, N, kOne, (int64_t)0);
      TNX_LAUNCH_CHECK();
    }
  } else {
    // FP64 Lanczos exit: hvec was written from the Lanczos relation.
    TNX_TRY(copy_f64(k.y, vec, N, s));
  }
  TNX_HIP(hipStreamSynchronize(s));
  info->eval = theta;
  info->resnorm = rho;
  info->anorm = anorm;
  info->matvecs = c.matvecs;
  info->matvecs_f64 = c.matvecs64;
  info->restarts = restarts;
  info->converged = converged ? 1 : 0;
  return converged ? TNX_OK : TNX_ERR_NOCONV;
}

// ---------------------------------------------------------------------------
// Batched eigensolves: `batch` independent one-site problems of one shape
// (stacks of L, W, R, v0; outputs vec, hvec stacked). FP64 thick-restart
// Lanczos in lockstep: every problem is at the same Lanczos step (restarts
// keep exactly `keep` vectors), each problem's Krylov scalars are its own,
// and a problem leaves the launched set (device index list) once its
// residual estimate meets tol * ||H||_est. One launch per kernel per step
// covers all problems still iterating.
// ---------------------------------------------------------------------------
size_t heff1_ws_bytes_b(int chil, int Dl, int Dr, int d, int chir, int batch);

// out[i * nst + t] = st[first + t] of launched problem i
__global__ void gather_status_kernel(const double* __restrict__ st, int first, int nst,
                                     double* __restrict__ out, PMap pm) {
  const int t = threadIdx.x;
  if (t < nst) out[(int64_t)blockIdx.y * nst + t] = st[pm.off() + first + t];
}

// Device-decided convergence (see chunk_mark): clear the flags and start
// the running ||T|| estimate at |theta| of the start vector (st[4]).
__global__ void chunk_init_kernel(double* st, PMap pm) {
  st += pm.off();
  if (threadIdx.x == 0) {
    st[10] = 0.0;
    st[11] = 0.0;
    st[12] = fabs(st[4]);
  }
}

// Results of the problems that finished within a run of steps: with Jc
// the problem's row count at convergence (st[11]; Jl for one that did not
// converge), vec = y = V[0..Jc) s and hvec = theta y + beta[Jc-1] s[Jc-1]
// V[Jc] (the beta term dropped for a problem stopped by an exhausted
// space). vec/hvec of logical problem p at p * N.
__global__ void chunk_finalize_kernel(const double* __restrict__ V, int64_t N,
                                      const double* __restrict__ s, const double* __restrict__ st,
                                      const double* __restrict__ beta, int Jl, int exhausted,
                                      double* __restrict__ vec, double* __restrict__ hvec, PMap pm) {
  const int64_t o = pm.off();
  V += o; s += o; st += o; beta += o;
  const int64_t po = (int64_t)pm.pid() * N;
  __shared__ double cs[kMaxBasis + 1];
  const bool done = st[10] != 0.0;
  const int Jc = done ? (int)st[11] : Jl;
  if ((int)threadIdx.x < Jc) cs[threadIdx.x] = s[threadIdx.x];
  __syncthreads();
  const double th = st[0];
  const double c = (!done && exhausted) ? 0.0 : beta[Jc - 1] * s[Jc - 1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int t = 0; t < Jc; ++t) acc += cs[t] * V[(int64_t)t * N + i];
    vec[po + i] = acc;
    if (hvec) hvec[po + i] = th * acc + c * V[(int64_t)Jc * N + i];
  }
}

static size_t eigsh_b_problem_bytes(size_t N, int m) { return align256(eigsh_ws_bytes_n(N, 0, m)); }

size_t eigsh_b_ws_bytes(int chil, int Dl, int Dr, int d, int chir, int max_basis, int batch) {
  const int m = std::max(3, std::min(max_basis, kMaxBasis));
  const size_t N = (size_t)chil * d * chir;
  return align256(heff1_ws_bytes_b(chil, Dl, Dr, d, chir, batch)) +
         2 * align256((size_t)batch * sizeof(int)) + align256((size_t)batch * 16 * 8) +
         eigsh_b_problem_bytes(N, m) * (size_t)batch + 256;
}

int eigsh_heff1_b(const double* L, const double* W, const double* R, const double* v0,
                  double* vec, double* hvec, int chil, int Dl, int Dr, int d, int chir,
                  int batch, const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                  size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(opts && info, "eigsh_batched: null opts/info");
  TNX_CHECK_ARG(L && W && R && v0 && vec && hvec, "eigsh_batched: null pointer");
  TNX_CHECK_ARG(chil > 0 && Dl > 0 && Dr > 0 && d > 0 && chir > 0, "eigsh_batched: bad dims");
  TNX_CHECK_ARG(batch >= 1 && batch <= 65535, "eigsh_batched: bad batch %d", batch);
  TNX_CHECK_ARG(opts->tol >= 0.0, "eigsh_batched: negative tol");
  if (opts->precision != TNX_F64) {
    set_error("eigsh_batched: only FP64 (precision 0) is supported, got %d", opts->precision);
    return TNX_ERR_UNSUPPORTED;
  }
  const int m = std::max(3, std::min(opts->max_basis, kMaxBasis));
  const int64_t N = (int64_t)chil * d * chir;
  TNX_CHECK_ARG(ws && ws_bytes >= eigsh_b_ws_bytes(chil, Dl, Dr, d, chir, m, batch),
                "eigsh_batched: workspace too small");
  const int max_mv = opts->max_matvecs > 0 ? opts->max_matvecs : 10000;
  const double tol = opts->tol;
  const int mb = (int)std::min<int64_t>(m, std::max<int64_t>(N, 1));
  const int keep = keep_of(mb, opts->keep);

  char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  void* mv = p;
  const size_t mvb = heff1_ws_bytes_b(chil, Dl, Dr, d, chir, batch);
  p += align256(mvb);
  int* idx = (int*)p; p += align256((size_t)batch * sizeof(int));
  int* fidx = (int*)p; p += align256((size_t)batch * sizeof(int));
  double* stat = (double*)p; p += align256((size_t)batch * 16 * 8);
  const size_t P = eigsh_b_problem_bytes((size_t)N, m);
  const int64_t sp = (int64_t)(P / 8);
  KWork k;
  carve(k, p, m, (size_t)N, 0);  // problem 0; problem q at + q * sp doubles

  std::vector<int> act(batch), fin;
  for (int b = 0; b < batch; ++b) act[b] = b;
  std::vector<double> anorm(batch, 0.0), hs((size_t)batch * 16);
  for (int b = 0; b < batch; ++b) info[b] = tnx_eigsh_info_t{0.0, 0.0, 0.0, 0, 0, 0, 0};
  int matvecs = 0, restarts = 0;
  bool any_fail = false;

  auto upload = [&](int* dst, const std::vector<int>& v) -> int {
    TNX_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, s));
    return TNX_OK;
  };
  auto status = [&](int first, int nst, int cnt) -> int {
    hipLaunchKernelGGL(gather_status_kernel, dim3(1, cnt), dim3(64), 0, s, k.st, first, nst, stat,
                       PMap{idx, sp});
    TNX_LAUNCH_CHECK();
    TNX_HIP(hipMemcpyAsync(hs.data(), stat, sizeof(double) * nst * cnt, hipMemcpyDeviceToHost, s));
    TNX_HIP(hipStreamSynchronize(s));
    return TNX_OK;
  };
  auto matvec = [&](const double* x, double* y, int cnt) -> int {
    matvecs++;
    return heff1_matvec_b(L, W, R, x, 0, y, chil, Dl, Dr, d, chir, TNX_F64, batch, cnt, idx, sp,
                          sp, mv, mvb, s, 0);
  };

  TNX_TRY(upload(idx, act));
  int cnt = batch;
  PMap pm{idx, sp};
  // y = v0 / |v0|
  hipLaunchKernelGGL(copy_map_kernel, dim3(vec_blocks(N, cnt), cnt), dim3(256), 0, s, v0, k.y, N,
                     pm, N, sp);
  TNX_LAUNCH_CHECK();
  TNX_TRY(cgs_pass(k.y, N, 0, k.y, nullptr, false, k.partials, k.h3, N, s, pm, cnt));
  hipLaunchKernelGGL(normalize_kernel<double>, dim3(vec_blocks(N, cnt), cnt), dim3(256), 0, s, k.y, k.h3,
                     k.y, N, pm);
  TNX_LAUNCH_CHECK();
  // FP64 step on the start vectors: theta = y.Hy, r = Hy - theta y
  TNX_TRY(matvec(k.y, k.w, cnt));
  TNX_TRY(cgs_pass(k.y, N, 1, k.w, nullptr, true, k.partials, k.h1, N, s, pm, cnt));
  TNX_TRY(cgs_pass(k.y, N, 1, k.w, k.h1, true, k.partials, k.h2, N, s, pm, cnt));
  TNX_TRY(cgs_pass(k.y, N, 1, k.w, k.h2, false, k.partials, k.h3, N, s, pm, cnt));
  hipLaunchKernelGGL(outer_scalars_kernel, dim3(1, cnt), dim3(256), 0, s, k.h1, k.h2, k.h3,
                     k.alpha, k.beta, k.scal, k.st, k.T, pm);
  TNX_LAUNCH_CHECK();
  TNX_TRY(status(4, 2, cnt));
  {
    std::vector<int> rest;
    fin.clear();
    for (int i = 0; i < cnt; ++i) {
      const int b = act[i];
      const double th = hs[2 * i], rho = hs[2 * i + 1];
      anorm[b] = std::max(anorm[b], std::fabs(th));
      info[b].eval = th;
      info[b].resnorm = rho;
      info[b].matvecs = info[b].matvecs_f64 = matvecs;
      if (rho <= tol * anorm[b] || rho == 0.0 || N == 1) {
        fin.push_back(b);
        info[b].converged = 1;
      } else {
        rest.push_back(b);
      }
    }
    if (!fin.empty()) {
      // converged start vectors: vec = y, hvec = r + theta y
      TNX_TRY(upload(fidx, fin));
      const int fc = (int)fin.size();
      PMap pf{fidx, sp};
      hipLaunchKernelGGL(copy_map_kernel, dim3(vec_blocks(N, fc), fc), dim3(256), 0, s, k.y, vec,
                         N, pf, sp, N);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(axpbz_dev_kernel, dim3(vec_blocks(N, fc), fc), dim3(256), 0, s,
                         k.scal + 3, k.w, k.scal, k.y, hvec, N, pf, N);
      TNX_LAUNCH_CHECK();
    }
    act.swap(rest);
  }
  cnt = (int)act.size();
  if (cnt > 0) {
    TNX_TRY(upload(idx, act));
    // V[0] = y, V[1] = r / rho
    hipLaunchKernelGGL(copy_map_kernel, dim3(vec_blocks(N, cnt), cnt), dim3(256), 0, s, k.y, k.V,
                       N, pm, sp, sp);
    TNX_LAUNCH_CHECK();
    hipLaunchKernelGGL(normalize_kernel<double>, dim3(vec_blocks(N, cnt), cnt), dim3(256), 0, s, k.w,
                       k.h3 + 1, k.V + N, N, pm);
    TNX_LAUNCH_CHECK();
  }
  // Step pipelining as in the single-problem solver: the next step's
  // matvec and CGS passes for every active problem are queued before the
  // host reads this step's statuses; problems that finish here drop theirs
  // (not counted), the others take it as their next step.
  StepEvent sev;
  PinnedLease pin_lease((size_t)3 * batch);
  double* pin = pin_lease.get();
  const bool pipelined = sev.e != nullptr && pin != nullptr;
  auto front = [&](int jj) -> int {
    TNX_TRY(matvec(k.V + (int64_t)jj * N, k.w, cnt));
    const int JJ = jj + 1;
    TNX_TRY(cgs_pass(k.V, N, JJ, k.w, nullptr, true, k.partials, k.h1, N, s, pm, cnt));
    TNX_TRY(cgs_pass(k.V, N, JJ, k.w, k.h1, true, k.partials, k.h2, N, s, pm, cnt));
    TNX_TRY(cgs_pass(k.V, N, JJ, k.w, k.h2, false, k.partials, k.h3, N, s, pm, cnt));
    return TNX_OK;
  };
  // Device-decided convergence: with steps_per_check K > 1 the host queues
  // up to K lockstep steps (bounded by the next restart, an exhausted space
  // and the matvec limit) without reading statuses. Each problem's
  // tridiagonal solve tests the criterion itself and freezes the problem's
  // Ritz pair when it is met (chunk_mark), so a finished problem keeps its
  // result while the others step on. One status read per run of steps then
  // retires the finished problems (their Ritz vectors from the basis rows up
  // to their own convergence) and restarts the rest when the basis is full.
  const int K = opts->steps_per_check > 1 ? std::min(opts->steps_per_check, kMaxBasis) : 1;
  if (K > 1 && cnt > 0) {
    hipLaunchKernelGGL(chunk_init_kernel, dim3(1, cnt), dim3(64), 0, s, k.st, pm);
    TNX_LAUNCH_CHECK();
    int jc = 1;
    while (cnt > 0) {
      const int nb = vec_blocks(N, cnt);
      const int mv0 = matvecs;
      int steps = std::min(K, mb - jc);
      steps = (int)std::min<int64_t>(steps, N - jc);
      steps = std::max(0, std::min(steps, max_mv - matvecs));
      for (int t = 0; t < steps; ++t) {
        const int jj = jc + t, JJ = jj + 1;
        TNX_TRY(front(jj));
        hipLaunchKernelGGL(tridiag_lowest_kernel, dim3(1, cnt), dim3(64), 0, s, k.alpha, k.beta, JJ,
                           k.beta + jj, k.s, k.st, pm, 1,
                           TriFinish{k.h1, k.h2, k.h3, k.alpha, k.beta, k.T}, tol);
        TNX_LAUNCH_CHECK();
        hipLaunchKernelGGL(normalize_kernel<double>, dim3(nb, cnt), dim3(256), 0, s, k.w,
                           k.h3 + JJ, k.V + (int64_t)(jj + 1) * N, N, pm);
        TNX_LAUNCH_CHECK();
      }
      const int Jl = jc + steps;  // basis rows after the last step
      TNX_TRY(status(0, 13, cnt));
      const bool exhausted = Jl >= N;
      const bool limit = matvecs >= max_mv;
      std::vector<int> rest;
      fin.clear();
      for (int i = 0; i < cnt; ++i) {
        const int b = act[i];
        const double* h = &hs[(size_t)13 * i];
        const bool done = h[10] != 0.0;
        anorm[b] = std::max(anorm[b], h[12]);
        if (done || exhausted || limit) {
          const int Jc = done ? (int)h[11] : Jl;
          fin.push_back(b);
          info[b].eval = h[0];
          info[b].resnorm = h[1];
          info[b].matvecs = info[b].matvecs_f64 = mv0 + (Jc - jc);
          info[b].restarts = restarts;
          info[b].converged = (done || exhausted) ? 1 : 0;
          any_fail |= !(done || exhausted);
        } else {
          rest.push_back(b);
        }
      }
      if (!fin.empty()) {
        TNX_TRY(upload(fidx, fin));
        const int fc = (int)fin.size();
        hipLaunchKernelGGL(chunk_finalize_kernel, dim3(vec_blocks(N, fc), fc), dim3(256), 0, s,
                           k.V, N, k.s, k.st, k.beta, Jl, exhausted ? 1 : 0, vec, hvec,
                           PMap{fidx, sp});
        TNX_LAUNCH_CHECK();
        act.swap(rest);
        cnt = (int)act.size();
        if (cnt == 0) break;
        TNX_TRY(upload(idx, act));
      }
      if (Jl >= mb) {
        const int jl = Jl - 1;
        hipLaunchKernelGGL(tridiag_all_kernel, dim3(1, cnt), dim3(64), 0, s, k.alpha, k.beta, Jl,
                           k.theta, k.S, kMaxBasis, pm);
        TNX_LAUNCH_CHECK();
        hipLaunchKernelGGL(tr_restart_kernel, dim3(1, cnt), dim3(64), 0, s, k.theta, k.S, k.beta, jl,
                           keep, Jl, k.Cm, k.alpha, k.beta, k.T,
                           reinterpret_cast<int*>(k.st + 8), pm);
        TNX_LAUNCH_CHECK();
        GemmDesc g{};
        g.M = keep; g.N = (int)N; g.K = Jl; g.batch = cnt; g.bidx = idx;
        g.A = k.Cm; g.ta = DT::F64; g.lda = kMaxBasis; g.sA = sp; g.a_kcontig = false;
        g.B = k.V; g.tb = DT::F64; g.ldb = N; g.sB = sp; g.b_ncontig = true;
        g.C = k.Y; g.tc = DT::F64; g.ldc = N; g.sC = sp;
        g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
        TNX_TRY(gemm(g, s));
        hipLaunchKernelGGL(copy_map_kernel, dim3(vec_blocks((int64_t)keep * N, cnt), cnt),
                           dim3(256), 0, s, k.Y, k.V, (int64_t)keep * N, pm, sp, sp);
        TNX_LAUNCH_CHECK();
        hipLaunchKernelGGL(copy_map_kernel, dim3(vec_blocks(N, cnt), cnt), dim3(256), 0, s,
                           k.V + (int64_t)Jl * N, k.V + (int64_t)keep * N, N, pm, sp, sp);
        TNX_LAUNCH_CHECK();
        restarts++;
        jc = keep;
      } else {
        jc = Jl;
      }
    }
    TNX_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < batch; ++b) info[b].anorm = anorm[b];
    if (any_fail) {
      set_error("eigsh_batched: matvec limit %d reached before convergence", max_mv);
      return TNX_ERR_NOCONV;
    }
    return TNX_OK;
  }
  int j = 1;
  bool queued = false;  // step j's front already launched for the active set
  while (cnt > 0) {
    const int nb = vec_blocks(N, cnt);
    const int J = j + 1;
    if (!queued) TNX_TRY(front(j));
    const int mv_now = matvecs;
    hipLaunchKernelGGL(tridiag_lowest_kernel, dim3(1, cnt), dim3(64), 0, s, k.alpha, k.beta, J,
                       k.beta + j, k.s, k.st, pm, 1,
                       TriFinish{k.h1, k.h2, k.h3, k.alpha, k.beta, k.T}, -1.0);
    TNX_LAUNCH_CHECK();
    hipLaunchKernelGGL(normalize_kernel<double>, dim3(nb, cnt), dim3(256), 0, s, k.w, k.h3 + J,
                       k.V + (int64_t)(j + 1) * N, N, pm);
    TNX_LAUNCH_CHECK();
    const bool ahead = pipelined && J < mb && J < N && mv_now < max_mv;
    if (ahead) {
      hipLaunchKernelGGL(gather_status_kernel, dim3(1, cnt), dim3(64), 0, s, k.st, 0, 3, stat, pm);
      TNX_LAUNCH_CHECK();
      TNX_HIP(hipMemcpyAsync(pin, stat, sizeof(double) * 3 * cnt, hipMemcpyDeviceToHost, s));
      TNX_HIP(hipEventRecord(sev.e, s));
      TNX_TRY(front(j + 1));
      TNX_HIP(hipEventSynchronize(sev.e));
      std::copy(pin, pin + (size_t)3 * cnt, hs.begin());
    } else {
      TNX_TRY(status(0, 3, cnt));
    }
    const bool exhausted = (J >= N);
    const bool limit = mv_now >= max_mv;
    std::vector<int> rest;
    fin.clear();
    for (int i = 0; i < cnt; ++i) {
      const int b = act[i];
      const double th = hs[3 * i], est = hs[3 * i + 1];
      anorm[b] = std::max(anorm[b], hs[3 * i + 2]);
      const bool conv = est <= tol * anorm[b] || exhausted;
      if (conv || limit) {
        fin.push_back(b);
        info[b].eval = th;
        info[b].resnorm = est;
        info[b].matvecs = info[b].matvecs_f64 = mv_now;
        info[b].restarts = restarts;
        info[b].converged = conv ? 1 : 0;
        any_fail |= !conv;
      } else {
        rest.push_back(b);
      }
    }
    if (!fin.empty()) {
      // Ritz vector y = V s; H y = theta y + beta_j s_j V[j+1]
      TNX_TRY(upload(fidx, fin));
      const int fc = (int)fin.size();
      const int nf = vec_blocks(N, fc);
      PMap pf{fidx, sp};
      hipLaunchKernelGGL(vcomb_kernel<double>, dim3(nf, fc), dim3(256), 0, s, k.V, N, J, k.s, k.y, N, pf);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(ritz_scalars_kernel, dim3(1, fc), dim3(1), 0, s, k.st, k.s, k.beta, j,
                         k.scal, pf);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(axpbz_dev_kernel, dim3(nf, fc), dim3(256), 0, s, k.scal, k.y,
                         exhausted ? nullptr : k.scal + 1, k.V + (int64_t)(j + 1) * N, hvec, N,
                         pf, N);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(copy_map_kernel, dim3(nf, fc), dim3(256), 0, s, k.y, vec, N, pf, sp, N);
      TNX_LAUNCH_CHECK();
      act.swap(rest);
      cnt = (int)act.size();
      if (cnt == 0) break;
      TNX_TRY(upload(idx, act));
    }
    if (J >= mb) {
      hipLaunchKernelGGL(tridiag_all_kernel, dim3(1, cnt), dim3(64), 0, s, k.alpha, k.beta, J,
                         k.theta, k.S, kMaxBasis, pm);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(tr_restart_kernel, dim3(1, cnt), dim3(64), 0, s, k.theta, k.S, k.beta, j,
                         keep, J, k.Cm, k.alpha, k.beta, k.T,
                         reinterpret_cast<int*>(k.st + 8), pm);
      TNX_LAUNCH_CHECK();
      // Y = C^T V (keep x N per problem), then V[0..keep) = Y, V[keep] = V[j+1]
      GemmDesc g{};
      g.M = keep; g.N = (int)N; g.K = J; g.batch = cnt; g.bidx = idx;
      g.A = k.Cm; g.ta = DT::F64; g.lda = kMaxBasis; g.sA = sp; g.a_kcontig = false;
      g.B = k.V; g.tb = DT::F64; g.ldb = N; g.sB = sp; g.b_ncontig = true;
      g.C = k.Y; g.tc = DT::F64; g.ldc = N; g.sC = sp;
      g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
      TNX_TRY(gemm(g, s));
      hipLaunchKernelGGL(copy_map_kernel, dim3(vec_blocks((int64_t)keep * N, cnt), cnt),
                         dim3(256), 0, s, k.Y, k.V, (int64_t)keep * N, pm, sp, sp);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(copy_map_kernel, dim3(nb, cnt), dim3(256), 0, s,
                         k.V + (int64_t)(j + 1) * N, k.V + (int64_t)keep * N, N, pm, sp, sp);
      TNX_LAUNCH_CHECK();
      restarts++;
      j = keep;
      queued = false;
    } else {
      j++;
      queued = ahead;
    }
  }
  TNX_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < batch; ++b) info[b].anorm = anorm[b];
  if (any_fail) {
    set_error("eigsh_batched: matvec limit %d reached before convergence", max_mv);
    return TNX_ERR_NOCONV;
  }
  return TNX_OK;
}

// Lowest eigenpairs of `batch` dense symmetric n x n matrices (stacked).
size_t syev_b_ws_bytes(int n, int batch) {
  return align256((size_t)batch * 8) + align256(syev_ws_bytes(n)) * (size_t)batch + 256;
}

int syev_lowest_b(const double* M, int n, int batch, double* evec, double* eval_host, void* ws,
                  size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(n >= 1 && n <= kTriMax, "syev_batched: n=%d outside [1, %d]", n, kTriMax);
  TNX_CHECK_ARG(batch >= 1 && batch <= 65535, "syev_batched: bad batch %d", batch);
  TNX_CHECK_ARG(M && evec && eval_host, "syev_batched: null pointer");
  TNX_CHECK_ARG(ws && ws_bytes >= syev_b_ws_bytes(n, batch), "syev_batched: workspace too small");
  char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* ev = (double*)p; p += align256((size_t)batch * 8);
  const int64_t nn = (int64_t)n * n;
  if (n == 1) {
    TNX_TRY(fill_f64(evec, batch, 1.0, s));
    TNX_HIP(hipMemcpyAsync(eval_host, M, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
    TNX_HIP(hipStreamSynchronize(s));
    return TNX_OK;
  }
  const int64_t sp = (int64_t)(align256(syev_ws_bytes(n)) / 8);
  double* A = (double*)p;
  double* Vst = A + nn;
  double* tau = Vst + nn;
  double* dg = tau + n;
  double* eg = dg + n;
  double* z = eg + n;
  double* st = z + n + 8;
  const PMap pm{nullptr, sp};
  hipLaunchKernelGGL(copy_map_kernel, dim3(vec_blocks(nn, batch), batch), dim3(256), 0, s, M, A,
                     nn, pm, nn, sp);
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(sytrd_kernel, dim3(1, batch), dim3(256), 0, s, A, n, Vst, tau, dg, eg, pm);
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(tridiag_lowest_kernel, dim3(1, batch), dim3(64), 0, s, dg, eg, n,
                     (const double*)nullptr, z, st, pm, 0, TriFinish{}, -1.0);
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(ormtr_kernel, dim3(1, batch), dim3(256), 0, s, Vst, tau, n, z, evec, pm,
                     (int64_t)n);
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(gather_status_kernel, dim3(1, batch), dim3(64), 0, s, st, 0, 1, ev, pm);
  TNX_LAUNCH_CHECK();
  TNX_HIP(hipMemcpyAsync(eval_host, ev, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
  TNX_HIP(hipStreamSynchronize(s));
  return TNX_OK;
}

}  // namespace tnx

// ---- C-ABI -----------------------------------------------------------------
extern "C" {

size_t tnx_syev_lowest_workspace_bytes(int n) { return tnx::syev_ws_bytes(n); }

int tnx_syev_lowest(const double* M, int n, double* evec, double* eval_host, void* ws,
                    size_t ws_bytes, tnx_stream_t stream) {
  TNX_CHECK_ARG(eval_host, "tnx_syev_lowest: null eval_host");
  return tnx::syev_lowest(M, n, evec, eval_host, ws, ws_bytes, (hipStream_t)stream);
}

size_t tnx_eigsh_heff1_workspace_bytes(int chil, int Dl, int Dr, int d, int chir,
                                       int max_basis) {
  return tnx::eigsh_ws_bytes(chil, Dl, Dr, d, chir, max_basis);
}

size_t tnx_eigsh_heff2_workspace_bytes(int chil, int Dl, int Dm, int Dr, int d, int chir,
                                       int max_basis) {
  return tnx::eigsh2_ws_bytes(chil, Dl, Dm, Dr, d, chir, max_basis);
}

int tnx_eigsh_heff2(const double* L, const double* W1, const double* W2, const double* R,
                    const double* v0, double* vec, double* hvec, int chil, int Dl, int Dm,
                    int Dr, int d, int chir, const tnx_eigsh_opts_t* opts,
                    tnx_eigsh_info_t* info, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::eigsh_heff2(L, W1, W2, R, v0, vec, hvec, chil, Dl, Dm, Dr, d, chir, opts, info,
                          ws, ws_bytes, (hipStream_t)stream);
}

int tnx_eigsh_heff1(const double* L, const double* W, const double* R, const double* v0,
                    double* vec, double* hvec, int chil, int Dl, int Dr, int d, int chir,
                    const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                    size_t ws_bytes, tnx_stream_t stream) {
  return tnx::eigsh_heff1(L, W, R, v0, vec, hvec, chil, Dl, Dr, d, chir, opts, info, ws,
                          ws_bytes, (hipStream_t)stream);
}

int tnx_axpy(int64_t n, double alpha, const double* x, double* y, tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(n >= 0, "tnx_axpy: negative n");
  if (n == 0) return TNX_OK;
  TNX_CHECK_ARG(x && y, "tnx_axpy: null pointer");
  hipLaunchKernelGGL(axpy_kernel, dim3(vec_blocks(n)), dim3(256), 0, (hipStream_t)stream, n,
                     alpha, x, y);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

int tnx_scal(int64_t n, double alpha, double* x, tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(n >= 0, "tnx_scal: negative n");
  if (n == 0) return TNX_OK;
  TNX_CHECK_ARG(x, "tnx_scal: null pointer");
  hipLaunchKernelGGL(scal_kernel, dim3(vec_blocks(n)), dim3(256), 0, (hipStream_t)stream, n,
                     alpha, x);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

size_t tnx_dot_workspace_bytes(int64_t n) {
  (void)n;
  return (size_t)(tnx::kCgsBlocks + 2) * tnx::kPartStride * 8 + 256;
}

int tnx_dot(int64_t n, const double* x, const double* y, double* out_host, void* ws,
            size_t ws_bytes, tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(n >= 1 && x && y && out_host, "tnx_dot: bad arguments");
  TNX_CHECK_ARG(ws && ws_bytes >= tnx_dot_workspace_bytes(n), "tnx_dot: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  double* partials = (double*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* out = partials + (size_t)kCgsBlocks * kPartStride;
  // one pass with V = x (J = 1), w = y: out[0] = x.y
  TNX_TRY(cgs_pass(x, n, 1, const_cast<double*>(y), nullptr, true, partials, out, n, s));
  TNX_HIP(hipMemcpyAsync(out_host, out, sizeof(double), hipMemcpyDeviceToHost, s));
  TNX_HIP(hipStreamSynchronize(s));
  return TNX_OK;
}

size_t tnx_syev_lowest_batched_workspace_bytes(int n, int batch) {
  return tnx::syev_b_ws_bytes(n, batch);
}

int tnx_syev_lowest_batched(const double* M, int n, int batch, double* evec,
                            double* eval_host, void* ws, size_t ws_bytes,
                            tnx_stream_t stream) {
  return tnx::syev_lowest_b(M, n, batch, evec, eval_host, ws, ws_bytes, (hipStream_t)stream);
}

size_t tnx_eigsh_heff1_batched_workspace_bytes(int chil, int Dl, int Dr, int d, int chir,
                                               int max_basis, int batch) {
  return tnx::eigsh_b_ws_bytes(chil, Dl, Dr, d, chir, max_basis, batch);
}

int tnx_eigsh_heff1_batched(const double* L, const double* W, const double* R,
                            const double* v0, double* vec, double* hvec, int chil,
                            int Dl, int Dr, int d, int chir, int batch,
                            const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info,
                            void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::eigsh_heff1_b(L, W, R, v0, vec, hvec, chil, Dl, Dr, d, chir, batch, opts, info, ws,
                            ws_bytes, (hipStream_t)stream);
}

}  // extern "C"
