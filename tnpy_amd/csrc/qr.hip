// Blocked Householder QR and the untruncated gauge moves of one-site DMRG.
//
// The reference moves the orthogonality centre with an untruncated SVD
// (tnpy/matrix_product_state.py:187-225 via tnpy/linalg.py:9-23): site <- U,
// neighbour <- (s V^T) * neighbour. Any factorisation psi = Q X with Q an
// isometry leaves the state unchanged and the site left-orthonormal; Q R is
// the cheapest. Householder (not Cholesky) keeps Q orthonormal when psi is
// rank deficient (product-like states, e.g. the ferromagnetic XXZ chain).
//
// Layout: A is row-major m x n (m >= n). Panels of NB columns are factored
// by one workgroup that holds its rows in registers (two block reductions
// per column: the column norm, then the reflector products with the other
// panel columns fused with the V^T v Gram entries that build the compact-WY
// factor T). Trailing updates and the explicit Q are GEMMs on MFMA:
//   A_trail -= V (T^T (V^T A_trail)),   Q_sub -= V (T (V^T Q_sub)).
#include "tnx_common.h"

#include <algorithm>
#include <vector>

namespace tnx {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum K per-thread values over the block; result in out[0..K-1] (LDS),
// visible to all threads on return.
template <int NT, int K>
__device__ __forceinline__ void block_reduce(const double (&v)[K], double* red, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int NW = NT / 64;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double t = wsum(v[k]);
    if (lane == 0) red[wv * K + k] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x < K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w * K + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Factor the panel A[jb:m, jb:jb+b] (b <= NB). Writes R and the reflectors
// back to A, the explicit V (unit diagonal, zeros above) to Vall at the same
// positions, and the b x b upper-triangular T (row-major, ld NB) to Tp.
template <int NT, int RPT, int NB>
__global__ __launch_bounds__(NT) void qr_panel_kernel(double* __restrict__ A, int m, int n,
                                                     int jb, int b, double* __restrict__ Vall,
                                                     double* __restrict__ Tp) {
  __shared__ double red[(NT / 64) * NB];
  __shared__ double dots[NB];
  __shared__ double sig[1];
  __shared__ double alpha_s[1];
  __shared__ double T[NB * NB];
  const int tid = threadIdx.x;
  const int mp = m - jb;
  double P[RPT][NB];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int rho = tid + NT * r;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      P[r][c] = (rho < mp && c < b) ? A[(int64_t)(jb + rho) * n + jb + c] : 0.0;
  }
  for (int t = tid; t < NB * NB; t += NT) T[t] = 0.0;

  // The column loop is a runtime loop; every access to P uses a compile-time
  // column index (select on c == j) so the panel stays in registers.
  for (int j = 0; j < b; ++j) {
    // --- column norm below the diagonal, and alpha = P[j][j]
    double part[1] = {0.0};
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int rho = tid + NT * r;
      double pj = 0.0;
#pragma unroll
      for (int c = 0; c < NB; ++c) pj = (c == j) ? P[r][c] : pj;
      if (rho > j && rho < mp) part[0] += pj * pj;
      if (rho == j) alpha_s[0] = pj;
    }
    block_reduce<NT, 1>(part, red, sig);
    const double sigma = sig[0];
    const double alpha = alpha_s[0];
    double tau, beta, scale;
    if (sigma == 0.0) {
      tau = 0.0; beta = alpha; scale = 0.0;
    } else {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    // v_j: 1 at rho == j, P*scale below (kept in column j); R diag = beta
    double vj[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int rho = tid + NT * r;
      const bool below = rho > j && rho < mp;
      double pj = 0.0;
#pragma unroll
      for (int c = 0; c < NB; ++c) pj = (c == j) ? P[r][c] : pj;
      vj[r] = below ? pj * scale : (rho == j ? 1.0 : 0.0);
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c == j) P[r][c] = below ? vj[r] : (rho == j ? beta : P[r][c]);
    }
    // --- one reduction: c > j: v_j . P[:, c] (reflector application);
    //     c < j: v_j . v_c (Gram entries for T). Below the diagonal of a
    //     column c < j, P[.][c] holds v_c, and rho >= j > c covers the
    //     whole support of v_j.
    {
      // streamed: one wave reduction per column keeps a single partial live
      const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        double pd = 0.0;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const int rho = tid + NT * r;
          if (c != j && c < b && rho >= j && rho < mp) pd += vj[r] * P[r][c];
        }
        pd = wsum(pd);
        if (lane == 0) red[wv * NB + c] = pd;
      }
      __syncthreads();
      if (tid < NB) {
        double acc = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) acc += red[w * NB + tid];
        dots[tid] = acc;
      }
      __syncthreads();
    }
    if (tau != 0.0) {
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c > j && c < b) {
          const double f = tau * dots[c];
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            const int rho = tid + NT * r;
            if (rho >= j && rho < mp) P[r][c] -= f * vj[r];
          }
        }
      }
    }
    // T[0:j, j] = -tau T[0:j, 0:j] y,  y_i = v_i . v_j ;  T[j][j] = tau
    if (tid == 0) {
      for (int i = 0; i < j; ++i) {
        double acc = 0.0;
        for (int l = i; l < j; ++l) acc += T[i * NB + l] * dots[l];
        T[i * NB + j] = -tau * acc;
      }
      T[j * NB + j] = tau;
    }
    __syncthreads();
  }
  // write back R / reflectors, explicit V, and T
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int rho = tid + NT * r;
    if (rho < mp) {
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c < b) {
          const int64_t off = (int64_t)(jb + rho) * n + jb + c;
          A[off] = P[r][c];
          Vall[off] = rho < c ? 0.0 : (rho == c ? 1.0 : P[r][c]);
        }
      }
    }
  }
  for (int t = tid; t < NB * NB; t += NT) Tp[t] = T[t];
}

// R[i][c] = A[i][c] for c >= i, else 0   (n x n, row-major)
__global__ void extract_r_kernel(const double* A, int n, double* R) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, c = e % n;
    R[e] = c >= i ? A[i * n + c] : 0.0;
  }
}

// A (m x n) <- [I_n ; 0]
__global__ void eye_kernel(double* A, int m, int n) {
  const int64_t mn = (int64_t)m * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, c = e % n;
    A[e] = (i == c) ? 1.0 : 0.0;
  }
}

// Sign convention of every QR here: R with a non-negative diagonal (a
// column j with R_jj < 0 flips column j of Q and row j of R), so the TSQR
// panels, the one-workgroup panels and the LDS-resident batched QR return
// the same factorisation of a full-rank matrix.
__global__ void qr_signs_kernel(const double* __restrict__ R, int n, double* __restrict__ sg) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    sg[j] = R[(int64_t)j * n + j] < 0.0 ? -1.0 : 1.0;
}

__global__ void qr_apply_signs_kernel(double* __restrict__ Q, int m, double* __restrict__ R, int n,
                                      const double* __restrict__ sg) {
  const int64_t mn = (int64_t)m * n, nn = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn + nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e < mn) {
      Q[e] *= sg[e % n];
    } else {
      const int64_t f = e - mn;
      R[f] *= sg[f / n];
    }
  }
}

static int qr_fix_signs(double* Q, int m, double* R, int n, double* sg, hipStream_t s) {
  hipLaunchKernelGGL(qr_signs_kernel, dim3((n + 255) / 256), dim3(256), 0, s, R, n, sg);
  TNX_LAUNCH_CHECK();
  const int64_t tot = (int64_t)m * n + (int64_t)n * n;
  const int nb = (int)std::min<int64_t>((tot + 255) / 256, 8192);
  hipLaunchKernelGGL(qr_apply_signs_kernel, dim3(nb), dim3(256), 0, s, Q, m, R, n, sg);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

struct QrCfg { int nt, rpt, nb; };

static bool qr_config(int m, QrCfg* c) {
  if (m <= 512) { *c = {512, 1, 16}; return true; }
  if (m <= 1024) { *c = {512, 2, 16}; return true; }
  if (m <= 2048) { *c = {512, 4, 16}; return true; }
  if (m <= 4096) { *c = {1024, 4, 8}; return true; }
  if (m <= 8192) { *c = {1024, 8, 4}; return true; }
  return false;
}

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- TSQR panels ---------------------------------------------------------------
// A panel of b <= kTsNB columns and mp rows is factored without a long
// serial chain on one workgroup:
//   leaf  (one workgroup per block of h..2h-1 rows): Householder QR of the
//         block in LDS, its R_i and explicit thin Q_i;
//   root  (one workgroup): Householder QR of the stacked R_i, giving the
//         panel's R and the explicit Q_top, then the Householder vectors of
//         the whole panel are rebuilt from its explicit Q (Ballard, Demmel,
//         Grigori, Jacquelin, Nguyen, Solomonik, "Reconstructing Householder
//         vectors from tall-skinny QR", 2014): with Q1 the top b x b block of
//         Q, the LU factorisation without pivoting Q1 - S = L1 U, where the
//         signs S_ii = -sign of the current pivot make |U_ii| >= 1, gives
//         V = [L1; Q2 U^-1], T = -U S L1^-T and R = S R_tsqr, so that the
//         panel equals (I - V T V^T)[R; 0] in the compact-WY form the
//         trailing updates use;
//   apply (one workgroup per leaf): the explicit Q rows Q_i Q_top,i, then
//         V rows (Q2 U^-1), V T^T and V T for the two-GEMM trailing update
//         A_trail -= (V T^T)(V^T A_trail) and the backward formation of Q,
//         Q_sub -= (V T)(V^T Q_sub).
constexpr int kTsNB = 32;         // panel width
constexpr int kTsLd = kTsNB + 1;  // LDS row stride
constexpr int kTsMaxRows = 511;   // rows of one leaf (h <= 256, last leaf < 2h)
constexpr int kTsMaxLeaves = 16;  // stacked R_i in the root: 16 * 32 rows

struct TsPlan { int h, nleaf; };

static bool ts_plan(int mp, TsPlan* p) {
  const int h = mp > kTsMaxLeaves * 128 + 127 ? 256 : 128;
  const int nleaf = std::max(1, mp / h);
  if (nleaf > kTsMaxLeaves || mp - (nleaf - 1) * h > kTsMaxRows) return false;
  *p = {h, nleaf};
  return true;
}

template <int NT>
__device__ __forceinline__ double ts_block_sum(double v, double* red) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wsum(v);
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) t += red[w];
  return t;
}

// Householder QR of the m x n (m >= n) matrix held in LDS (row stride ld):
// R on and above the diagonal, the reflector tails below it, tau_s[j].
template <int NT>
__device__ void lds_householder(double* S, int ld, int m, int n, double* tau_s, double* red) {
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = 0; j < n; ++j) {
    double part = 0.0;
    for (int i = j + 1 + tid; i < m; i += NT) part += S[i * ld + j] * S[i * ld + j];
    const double sigma = ts_block_sum<NT>(part, red);
    const double alpha = S[j * ld + j];
    double tj = 0.0, beta = alpha, scale = 0.0;
    if (sigma != 0.0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < m; i += NT) S[i * ld + j] *= scale;
    if (tid == 0) { S[j * ld + j] = beta; tau_s[j] = tj; }
    __syncthreads();
    if (tj != 0.0) {
      for (int c = j + 1 + wave; c < n; c += NW) {
        double acc = 0.0;
        for (int i = j + lane; i < m; i += 64) acc += (i == j ? 1.0 : S[i * ld + j]) * S[i * ld + c];
        const double w = tj * wsum(acc);
        for (int i = j + lane; i < m; i += 64) S[i * ld + c] -= (i == j ? 1.0 : S[i * ld + j]) * w;
      }
    }
    __syncthreads();
  }
}

// Explicit thin Q = H_0 ... H_{n-1} [I; 0], formed in place over the output
// of lds_householder (whose R is overwritten).
template <int NT>
__device__ void lds_form_q(double* S, int ld, int m, int n, const double* tau_s) {
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = n - 1; j >= 0; --j) {
    const double tj = tau_s[j];
    if (tj != 0.0) {
      for (int c = j + 1 + wave; c < n; c += NW) {
        double acc = 0.0;
        for (int i = j + lane; i < m; i += 64) acc += (i == j ? 1.0 : S[i * ld + j]) * S[i * ld + c];
        const double w = tj * wsum(acc);
        for (int i = j + lane; i < m; i += 64) S[i * ld + c] -= (i == j ? 1.0 : S[i * ld + j]) * w;
      }
    }
    __syncthreads();
    for (int i = tid; i < m; i += NT)
      S[i * ld + j] = (i < j) ? 0.0 : (i == j ? 1.0 - tj : -tj * S[i * ld + j]);
    __syncthreads();
  }
}

// Leaf: rows [leaf*h, leaf*h + rows) of the panel A[jb:, jb:jb+b].
// Writes R_i (kTsNB x kTsNB, zero padded) to Rst and the explicit Q_i rows to
// Qst (mp x kTsNB, panel-relative rows).
__global__ __launch_bounds__(256) void tsqr_leaf_kernel(const double* __restrict__ A, int n, int jb,
                                                       int mp, int b, int h, int nleaf,
                                                       double* __restrict__ Rst,
                                                       double* __restrict__ Qst,
                                                       const int* __restrict__ flag) {
  if (flag && *flag == 0) return;  // panel already factored by the Cholesky-QR path
  __shared__ double S[kTsMaxRows * kTsLd];
  __shared__ double tau_s[kTsNB];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const int leaf = blockIdx.x;
  const int r0 = leaf * h;
  const int rows = leaf == nleaf - 1 ? mp - r0 : h;
  for (int e = tid; e < rows * b; e += 256) {
    const int i = e / b, c = e - i * b;
    S[i * kTsLd + c] = A[(int64_t)(jb + r0 + i) * n + jb + c];
  }
  __syncthreads();
  lds_householder<256>(S, kTsLd, rows, b, tau_s, red);
  double* Rp = Rst + (size_t)leaf * kTsNB * kTsNB;
  for (int e = tid; e < kTsNB * kTsNB; e += 256) {
    const int i = e / kTsNB, c = e - i * kTsNB;
    Rp[e] = (i < b && c < b && c >= i) ? S[i * kTsLd + c] : 0.0;
  }
  __syncthreads();
  lds_form_q<256>(S, kTsLd, rows, b, tau_s);
  for (int e = tid; e < rows * kTsNB; e += 256) {
    const int i = e / kTsNB, c = e - i * kTsNB;
    Qst[(int64_t)(r0 + i) * kTsNB + c] = c < b ? S[i * kTsLd + c] : 0.0;
  }
}

// Root: QR of the stacked R_i, explicit Q_top, then the reconstruction of the
// panel's Householder form (see above). Writes the panel's R (= S R_tsqr)
// into A's panel rows on and above the diagonal, L1 into Vall's panel rows
// 0..b-1, T, U and Q_top for the apply kernel.
__global__ __launch_bounds__(512) void tsqr_root_kernel(const double* __restrict__ Rst, int nleaf,
                                                       int b, const double* __restrict__ Qst,
                                                       double* __restrict__ A, int n, int jb,
                                                       double* __restrict__ Vall,
                                                       double* __restrict__ Tst,
                                                       double* __restrict__ Qtop,
                                                       double* __restrict__ Ust,
                                                       const int* __restrict__ flag,
                                                       const double* __restrict__ Rbuf) {
  if (flag && *flag == 0) {
    // the Cholesky-QR factorisation was accepted: its R (S R, b x b) into A
    for (int e = threadIdx.x; e < b * b; e += 512) {
      const int i = e / b, c = e - i * b;
      if (c >= i) A[(int64_t)(jb + i) * n + jb + c] = Rbuf[i * kTsNB + c];
    }
    return;
  }
  __shared__ double S[kTsMaxLeaves * kTsNB * kTsLd];
  __shared__ double Rb[kTsNB * kTsLd];
  __shared__ double B[kTsNB * kTsLd];
  __shared__ double Li[kTsNB * kTsLd];
  __shared__ double tau_s[kTsNB];
  __shared__ double sg[kTsNB];
  __shared__ double red[8];
  const int tid = threadIdx.x;
  const int rows = nleaf * b;
  for (int e = tid; e < rows * b; e += 512) {
    const int r = e / b, c = e - r * b;
    const int k = r / b, i = r - k * b;
    S[r * kTsLd + c] = Rst[(size_t)k * kTsNB * kTsNB + i * kTsNB + c];
  }
  __syncthreads();
  lds_householder<512>(S, kTsLd, rows, b, tau_s, red);
  for (int e = tid; e < b * b; e += 512) {
    const int i = e / b, c = e - i * b;
    Rb[i * kTsLd + c] = c >= i ? S[i * kTsLd + c] : 0.0;
  }
  __syncthreads();
  lds_form_q<512>(S, kTsLd, rows, b, tau_s);
  for (int e = tid; e < rows * kTsNB; e += 512) {
    const int r = e / kTsNB, c = e - r * kTsNB;
    Qtop[e] = c < b ? S[r * kTsLd + c] : 0.0;
  }
  // Q1 = (leaf 0's Q rows 0..b-1) x (Q_top rows 0..b-1)
  for (int e = tid; e < b * b; e += 512) {
    const int i = e / b, c = e - i * b;
    double acc = 0.0;
    for (int k = 0; k < b; ++k) acc += Qst[i * kTsNB + k] * S[k * kTsLd + c];
    B[i * kTsLd + c] = acc;
  }
  __syncthreads();
  // LU of Q1 - S without pivoting, S chosen at each pivot: |U_ii| >= 1
  for (int i = 0; i < b; ++i) {
    if (tid == 0) {
      const double a = B[i * kTsLd + i];
      const double sgn = a >= 0.0 ? -1.0 : 1.0;
      B[i * kTsLd + i] = a - sgn;
      sg[i] = sgn;
    }
    __syncthreads();
    const double piv = B[i * kTsLd + i];
    for (int r = i + 1 + tid; r < b; r += 512) B[r * kTsLd + i] /= piv;
    __syncthreads();
    const int w = b - 1 - i;
    for (int e = tid; e < w * w; e += 512) {
      const int r = i + 1 + e / w, c = i + 1 + e % w;
      B[r * kTsLd + c] -= B[r * kTsLd + i] * B[i * kTsLd + c];
    }
    __syncthreads();
  }
  // Li = L1^{-1} (unit lower), one column per thread
  if (tid < b) {
    const int c = tid;
    for (int r = 0; r < b; ++r) {
      double x;
      if (r < c) x = 0.0;
      else if (r == c) x = 1.0;
      else {
        double acc = 0.0;
        for (int k = c; k < r; ++k) acc += B[r * kTsLd + k] * Li[k * kTsLd + c];
        x = -acc;
      }
      Li[r * kTsLd + c] = x;
    }
  }
  __syncthreads();
  // T = -U S L1^-T ; U ; L1 -> Vall ; R -> A
  for (int e = tid; e < kTsNB * kTsNB; e += 512) {
    const int i = e / kTsNB, j = e - i * kTsNB;
    double t = 0.0, u = 0.0;
    if (i < b && j < b && i <= j) {
      for (int k = i; k <= j; ++k) t += B[i * kTsLd + k] * sg[k] * Li[j * kTsLd + k];
      u = B[i * kTsLd + j];
    }
    Tst[e] = -t;
    Ust[e] = u;
  }
  for (int e = tid; e < b * b; e += 512) {
    const int i = e / b, c = e - i * b;
    const int64_t off = (int64_t)(jb + i) * n + jb + c;
    Vall[off] = c < i ? B[i * kTsLd + c] : (c == i ? 1.0 : 0.0);
    if (c >= i) A[off] = sg[i] * Rb[i * kTsLd + c];
  }
}

// Apply: V rows below the top b rows = Q_i Q_top,i U^-1 = Q_i M with
// M = Q_top,i U^-1 solved once per workgroup (the top rows are L1, from the
// root); Y1 = V T^T, Y2 = V T. One panel row per thread.
__global__ __launch_bounds__(256) void tsqr_apply_kernel(const double* __restrict__ Qst,
                                                        const double* __restrict__ Qtop,
                                                        const double* __restrict__ Ust,
                                                        const double* __restrict__ Tst, int mp,
                                                        int b, int h, int nleaf, int n, int jb,
                                                        double* __restrict__ Vall,
                                                        double* __restrict__ Y1,
                                                        double* __restrict__ Y2,
                                                        const int* __restrict__ flag) {
  if (flag && *flag == 0) return;
  __shared__ double M[kTsNB * kTsNB];
  __shared__ double U[kTsNB * kTsNB];
  __shared__ double T[kTsNB * kTsNB];
  const int tid = threadIdx.x;
  const int leaf = blockIdx.x;
  const int r0 = leaf * h;
  const int rows = leaf == nleaf - 1 ? mp - r0 : h;
  for (int e = tid; e < kTsNB * kTsNB; e += 256) {
    const int k = e / kTsNB;
    M[e] = k < b ? Qtop[(size_t)(leaf * b + k) * kTsNB + (e - k * kTsNB)] : 0.0;
    U[e] = Ust[e];
    T[e] = Tst[e];
  }
  __syncthreads();
  // row k of M: M[k] U = Q_top,i[k]  (U upper triangular, |U_cc| >= 1)
  if (tid < b) {
    const int k = tid;
    for (int c = 0; c < b; ++c) {
      double acc = M[k * kTsNB + c];
      for (int j = 0; j < c; ++j) acc -= M[k * kTsNB + j] * U[j * kTsNB + c];
      M[k * kTsNB + c] = acc / U[c * kTsNB + c];
    }
  }
  __syncthreads();
  for (int i = tid; i < rows; i += 256) {
    const int pr = r0 + i;
    const int64_t off = (int64_t)(jb + pr) * n + jb;
    double v[kTsNB];
    if (pr < b) {
#pragma unroll
      for (int c = 0; c < kTsNB; ++c) v[c] = c < b ? Vall[off + c] : 0.0;
    } else {
#pragma unroll
      for (int c = 0; c < kTsNB; ++c) v[c] = 0.0;
      const double* qrow = Qst + (int64_t)pr * kTsNB;
      for (int k = 0; k < b; ++k) {
        const double a = qrow[k];
#pragma unroll
        for (int c = 0; c < kTsNB; ++c) v[c] += a * M[k * kTsNB + c];
      }
#pragma unroll
      for (int c = 0; c < kTsNB; ++c)
        if (c < b) Vall[off + c] = v[c];
    }
    for (int c = 0; c < b; ++c) {
      double y1 = 0.0, y2 = 0.0;
#pragma unroll
      for (int k = 0; k < kTsNB; ++k) {
        y1 += v[k] * T[c * kTsNB + k];  // (V T^T)[r][c] = sum_k V[r][k] T[c][k]
        y2 += v[k] * T[k * kTsNB + c];  // (V T)[r][c]   = sum_k V[r][k] T[k][c]
      }
      Y1[off + c] = y1;
      Y2[off + c] = y2;
    }
  }
}

// ---- Cholesky-QR panels ----------------------------------------------------------
// The panel P (mp x b, b <= kTsNB) is first factored by CholeskyQR2
// (Fukaya, Nakatsukasa, Yanagisawa, Yamamoto, "CholeskyQR2: a simple and
// communication-avoiding algorithm for computing a tall-skinny QR
// factorization", 2014): R1^T R1 = P^T P, Q1 = P R1^-1, R2^T R2 = Q1^T Q1,
// Q = Q1 R2^-1, R = R2 R1; Q is orthonormal and P - Q R at rounding level
// for cond(P) below about 1e7. Three kernels of many workgroups (kCqRB panel
// rows each) per panel; in the first two each workgroup writes the Gram
// partial of its rows, and the workgroup that arrives last (atomic ticket
// after a release fence) sums the partials in a fixed order and does the
// b x b work once, in one wave with a row per lane (no barriers in the
// serial steps): the Cholesky factor and its inverse, and after the second
// pass R and the Householder reconstruction of the panel from the explicit
// Q as the TSQR root does (LU of Q1 - S). The third kernel forms the rows
// of Q, V, V T^T and V T.
// The factorisation is accepted when both Cholesky pivots are positive and
// finite, the second Gram is within 0.1 of the identity, and every row of
// P - Q R is within 2e-14 ||P||_F (CholeskyQR keeps Q orthonormal for
// panels with exactly dependent columns but can lose R's accuracy there;
// accepted panels measure below 1e-15 ||P||_F). Otherwise the panel's flag
// is set and the TSQR kernels factor it: they return at once for a panel
// whose flag is clear, the root then only copying the accepted R into A.
constexpr int kCqRB = 64;         // panel rows per workgroup
constexpr int kCqLd = kTsNB + 1;  // LDS row stride
constexpr int kCqG = kTsNB * kTsNB;

// Small per-QR buffers of the Cholesky-QR panels (device memory).
struct CqBuf {
  double* Qb;    // mp x kTsNB: the panel's rows after the first pass (Q1)
  double* Gp;    // nblk x kCqG Gram partials (lower triangles)
  double* Li;    // the current pass's L^-1
  double* L1s;   // Cholesky factor of the first pass (R1 = L1^T)
  double* Rs;    // R = R2 R1 (upper)
  double* Rbuf;  // S R, the Householder-form R the TSQR root copies into A
  double* L1f;   // unit lower factor of the LU of Q1 - S
  double* Ui;    // U^-1
  double* T;     // compact-WY T
  double* pnorm; // ||P||_F
  int* cnt;      // arrival ticket, zero between kernels
};

// True in the workgroup that finishes last; its reads then see every other
// workgroup's writes made before their arrival.
__device__ __forceinline__ bool cq_arrive_last(int* cnt) {
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(cnt, 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// G = sum of nblk partial Grams, lower triangle, in the fixed order k = 0,
// 1, ... per entry; the 528 lower entries are spread over the threads
// (compact index p = i (i + 1) / 2 + c) with 16 partials in flight per
// entry. Entries outside b x b are the identity's, the upper triangle zero.
__device__ __forceinline__ void cq_gram_reduce(const double* __restrict__ Gp, int nblk, int b, double* G) {
  constexpr int kLow = kTsNB * (kTsNB + 1) / 2;  // 528
  int e[3];
  bool on[3];
  double s[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int p = threadIdx.x + 256 * q;
    int i = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    if ((i + 1) * (i + 2) / 2 <= p) ++i;
    if (i * (i + 1) / 2 > p) --i;
    const int c = p - i * (i + 1) / 2;
    e[q] = p < kLow ? i * kTsNB + c : -1;
    on[q] = p < kLow && i < b;
    s[q] = 0.0;
  }
  for (int k = 0; k < nblk; k += 16) {
    double v[3][16];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int u = 0; u < 16; ++u)
        v[q][u] = on[q] && k + u < nblk ? Gp[(size_t)(k + u) * kCqG + e[q]] : 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int u = 0; u < 16; ++u) s[q] += v[q][u];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (e[q] < 0) continue;
    const int i = e[q] >> 5, c = e[q] & 31;
    G[i * kCqLd + c] = on[q] ? s[q] : (i == c ? 1.0 : 0.0);
  }
  for (int x = threadIdx.x; x < kCqG; x += 256) {
    const int i = x >> 5, c = x & 31;
    if (c > i) G[i * kCqLd + c] = 0.0;
  }
  __syncthreads();
}

// LU of B - S without pivoting (B in LDS; S_jj = -sign of the pivot, so
// |U_jj| >= 1): the unit lower factor to Lf, U to U (zeros elsewhere), S to
// sg, 1/U_jj to rdU. By all 256 threads of the workgroup, right looking:
// thread t keeps row t / 8, columns 4 (t % 8) .. + 3 in registers; the
// pivot row and pivot column of step j are published in LDS by the threads
// that own them, double-buffered by the step's parity, with one barrier per
// step. (Measured faster than one wave with a row per lane, whose pivot row
// leaves a single lane one value at a time; for the Cholesky below, whose
// column is published by all lanes at once, the one-wave form is faster.)
// pub: 4 * kTsNB doubles, 16-byte aligned.

typedef double qr_f64x2 __attribute__((ext_vector_type(2)));

// entry q of a thread's four registers, q known only at run time
__device__ __forceinline__ double pick4(const double (&a)[4], int q) {
  return q == 0 ? a[0] : (q == 1 ? a[1] : (q == 2 ? a[2] : a[3]));
}

__device__ __forceinline__ void blk_lu(const double* B, double* Lf, double* U, double* sg,
                                       double* rdU, double* pub) {
  const int t = threadIdx.x, i = t >> 3, c0 = (t & 7) * 4;
  double a[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) a[q] = B[i * kCqLd + c0 + q];
  double* rowb = pub;               // [2][kTsNB]: row j of the step
  double* colb = pub + 2 * kTsNB;   // [2][kTsNB]: column j of the step
  if (i == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) rowb[c0 + q] = a[q];
  }
  if (c0 == 0) colb[i] = a[0];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTsNB; ++j) {
    const double* rb = rowb + (j & 1) * kTsNB;
    const double* cl = colb + (j & 1) * kTsNB;
    const double av = rb[j];
    const double sgn = av >= 0.0 ? -1.0 : 1.0;
    const double piv = av - sgn;
    const double rp = rcp_nr(piv);
    const double l = i > j ? cl[i] * rp : 0.0;
    const qr_f64x2 u0 = *reinterpret_cast<const qr_f64x2*>(rb + c0);
    const qr_f64x2 u1 = *reinterpret_cast<const qr_f64x2*>(rb + c0 + 2);
    const double uc[4] = {u0[0], u0[1], u1[0], u1[1]};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = c0 + q;
      if (i > j && c > j) a[q] = fma(-l, uc[q], a[q]);
      else if (i > j && c == j) a[q] = l;
      else if (i == j && c == j) a[q] = piv;
    }
    if (j + 1 < kTsNB) {
      double* rn = rowb + ((j + 1) & 1) * kTsNB;
      double* cn = colb + ((j + 1) & 1) * kTsNB;
      if (i == j + 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) rn[c0 + q] = a[q];
      }
      const int qn = j + 1 - c0;
      if (i > j + 1 && qn >= 0 && qn < 4) cn[i] = pick4(a, qn);
    }
    if (t == 0) { sg[j] = sgn; rdU[j] = rp; }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = c0 + q;
    Lf[i * kCqLd + c] = c < i ? a[q] : (c == i ? 1.0 : 0.0);
    U[i * kCqLd + c] = c >= i ? a[q] : 0.0;
  }
}

// Lane i's value of a wave-distributed double as a wave-uniform scalar
// (v_readlane); i must be wave-uniform.
__device__ __forceinline__ double qr_lane_val(double v, int i) {
  const long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), i);
  const unsigned hi = __builtin_amdgcn_readlane((int)(u >> 32), i);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Orders one wave's LDS writes before its later reads of other lanes'
// values (LDS serves a wave's instructions in order; this keeps the
// compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cholesky G = L L^T (lower triangle of G in LDS) by the calling wave: row i
// in lane i's registers (lanes >= kTsNB mirror row kTsNB - 1). The pivot is
// read with v_readlane; column j of L goes to the LDS buffer cb and every
// lane reads the entries it needs. L (zeros above the diagonal) to Lo,
// 1/L_jj to rd. Returns false when a pivot is not positive and finite (the
// same in every lane).
__device__ __forceinline__ bool wv_cholesky(const double* G, double* Lo, double* rd, double* cb) {
  const int lane = threadIdx.x & 63;
  const int i = lane < kTsNB ? lane : kTsNB - 1;
  double g[kTsNB];
#pragma unroll
  for (int c = 0; c < kTsNB; ++c) g[c] = c <= i ? G[i * kCqLd + c] : 0.0;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < kTsNB; ++j) {
    const double dj = qr_lane_val(g[j], j);
    ok = ok && dj > 0.0 && dj <= 1.7976931348623157e308;
    const double rs = rsqrt_nr(dj);
    const double l = i > j ? g[j] * rs : (i == j ? dj * rs : 0.0);
    g[j] = l;
    if (lane < kTsNB) cb[lane] = l;
    wave_lds_sync();
#pragma unroll
    for (int c = j + 1; c < kTsNB; ++c) g[c] = fma(-l, cb[c], g[c]);
    wave_lds_sync();
    if (lane == 0) rd[j] = rs;
  }
  if (lane < kTsNB) {
#pragma unroll
    for (int c = 0; c < kTsNB; ++c) Lo[i * kCqLd + c] = c <= i ? g[c] : 0.0;
  }
  return ok;
}

// Inverse of a 32 x 32 triangular matrix in LDS (identity outside its
// leading b x b block) by all 256 threads of the workgroup: the four
// diagonal 8 x 8 blocks by forward substitution (one column per thread),
// then the blocks below them by distance from the diagonal,
// X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj (through the LDS scratch Z), two
// barriers per distance. The dependent chains are 8 long instead of 32.
// TRANS: L is read transposed and X written transposed, so an upper
// triangular U gives X = U^-1. rd: the reciprocal diagonal, or null for a
// unit diagonal. Zeros on the other side of the diagonal.
template <bool TRANS>
__device__ __forceinline__ void blk_tri_inverse(const double* L, const double* rd, double* X,
                                                double* Z) {
  const int t = threadIdx.x;
  auto lv = [&](int r, int c) { return TRANS ? L[c * kCqLd + r] : L[r * kCqLd + c]; };
  auto xi = [&](int r, int c) { return TRANS ? c * kCqLd + r : r * kCqLd + c; };
  for (int e = t; e < kCqG; e += 256) X[(e >> 5) * kCqLd + (e & 31)] = 0.0;
  __syncthreads();
  if (t < kTsNB) {
    const int q8 = (t >> 3) * 8, cl = t & 7;
    double x[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      double acc = r == cl ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < r; ++k) acc = fma(-lv(q8 + r, q8 + k), x[k], acc);
      x[r] = r < cl ? 0.0 : (rd ? acc * rd[q8 + r] : acc);
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) X[xi(q8 + r, q8 + cl)] = x[r];
  }
  __syncthreads();
  for (int dist = 1; dist < 4; ++dist) {
    const int nb = 4 - dist;  // blocks (j + dist, j), j < nb
    for (int e = t; e < nb * 64; e += 256) {
      const int j = e >> 6, rr = (e >> 3) & 7, cc = e & 7, i = j + dist;
      double acc = 0.0;
      for (int k = 8 * j; k < 8 * i; ++k) acc = fma(lv(8 * i + rr, k), X[xi(k, 8 * j + cc)], acc);
      Z[(8 * i + rr) * kCqLd + 8 * j + cc] = acc;
    }
    __syncthreads();
    for (int e = t; e < nb * 64; e += 256) {
      const int j = e >> 6, rr = (e >> 3) & 7, cc = e & 7, i = j + dist;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        acc = fma(X[xi(8 * i + rr, 8 * i + k)], Z[(8 * i + k) * kCqLd + 8 * j + cc], acc);
      X[xi(8 * i + rr, 8 * j + cc)] = -acc;
    }
    __syncthreads();
  }
}

typedef double qr_f64x4 __attribute__((ext_vector_type(4)));

// One 16 x 16 tile of a product on the matrix cores: sum over k < K of
// fa(m, k) fb(k, n) for the tile's rows m and columns n (0..15), by the
// calling wave. The result's element v sits at row lk + 4 v, column lr
// (lr = lane % 16, lk = lane / 16).
template <int K, typename FA, typename FB>
__device__ __forceinline__ qr_f64x4 qr_mfma_tile(FA fa, FB fb) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  qr_f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < K; kk += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa(lr, kk + lk), fb(kk + lk, lr), acc, 0, 0, 0);
  return acc;
}

// The same with K (a multiple of 4) known only at run time.
template <typename FA, typename FB>
__device__ __forceinline__ qr_f64x4 qr_mfma_tile_rt(FA fa, FB fb, int K) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  qr_f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int kk = 0; kk < K; kk += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa(lr, kk + lk), fb(kk + lk, lr), acc, 0, 0, 0);
  return acc;
}

// Y[r] = X[r] Li^T for nrows (a multiple of 16) rows in LDS, columns past b
// zero; one 16 x 16 tile per wave at a time.
__device__ __forceinline__ void cq_rows_transform(const double* X, const double* Li, double* Y, int nrows, int b) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  for (int tile = wave; tile < (nrows / 16) * 2; tile += 4) {
    const int r0 = (tile >> 1) * 16, c0 = (tile & 1) * 16;
    const qr_f64x4 y = qr_mfma_tile<kTsNB>(
        [&](int m, int k) { return X[(r0 + m) * kCqLd + k]; },
        [&](int k, int n) { return Li[(c0 + n) * kCqLd + k]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = c0 + lr;
      Y[(r0 + lk + 4 * v) * kCqLd + c] = c < b ? y[v] : 0.0;
    }
  }
}

// Partial Gram Y^T Y (lower triangle) of the kCqRB rows in LDS (rows past
// the panel are zero): waves 0..2 take the tiles (0,0), (1,0), (1,1).
__device__ __forceinline__ void cq_gram_partial(const double* Y, double* __restrict__ Gout) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  if (wave > 2) return;
  const int i0 = wave == 0 ? 0 : 16, c0 = wave == 2 ? 16 : 0;
  const qr_f64x4 g = qr_mfma_tile<kCqRB>(
      [&](int m, int k) { return Y[k * kCqLd + i0 + m]; },
      [&](int k, int n) { return Y[k * kCqLd + c0 + n]; });
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + lk + 4 * v, c = c0 + lr;
    if (c <= i) Gout[(size_t)blockIdx.x * kCqG + i * kTsNB + c] = g[v];
  }
}

// 32 x 32 blocks between LDS (row stride kCqLd) and global memory, by the
// 256 threads of the workgroup, four entries each. cq_fetch loads a block
// into registers and cq_put stores it to LDS, so the loads of several
// blocks can be in flight together (a strided loop that stores each value
// as it arrives waits for every load in turn).
constexpr int kCqPer = kCqG / 256;

__device__ __forceinline__ void cq_store(const double* S, double* __restrict__ g) {
#pragma unroll
  for (int u = 0; u < kCqPer; ++u) {
    const int e = threadIdx.x + 256 * u;
    g[e] = S[(e >> 5) * kCqLd + (e & 31)];
  }
}

struct CqRegs {
  double v[kCqPer];
};

__device__ __forceinline__ CqRegs cq_fetch(const double* __restrict__ g) {
  CqRegs r;
#pragma unroll
  for (int u = 0; u < kCqPer; ++u) r.v[u] = g[threadIdx.x + 256 * u];
  return r;
}

__device__ __forceinline__ void cq_put(const CqRegs& r, double* S) {
#pragma unroll
  for (int u = 0; u < kCqPer; ++u) {
    const int e = threadIdx.x + 256 * u;
    S[(e >> 5) * kCqLd + (e & 31)] = r.v[u];
  }
}

__global__ void cq_init_kernel(int* cnt) { *cnt = 0; }

// PASS 0: Gram of P; the last workgroup factors it (L1, L1^-1) and sets the
// flag's first value. PASS 1: Q1 = P R1^-1 to Qb and its Gram; the last
// workgroup checks it against the identity, factors it (L2, L2^-1), forms
// R = (L1 L2)^T and rebuilds the Householder form from the top b rows of
// Q = Q1 R2^-1: LU of Q1 - S (S_ii = -sign of the pivot, |U_ii| >= 1),
// V = [L1; Q2' U^-1], T = -U S L1^-T.
template <int PASS>
__global__ __launch_bounds__(256) void cq_pass_kernel(const double* __restrict__ A, int n, int jb,
                                                     int mp, int b, CqBuf q, int nblk,
                                                     int* __restrict__ flag) {
  __shared__ double X[kCqRB * kCqLd];
  __shared__ double Y[kCqRB * kCqLd];
  __shared__ double G[kTsNB * kCqLd];
  __shared__ double Lo[kTsNB * kCqLd];
  __shared__ double Li[kTsNB * kCqLd];
  __shared__ double rd[kTsNB];
  __shared__ __attribute__((aligned(16))) double cb[4 * kTsNB];
  __shared__ int okf;
  const int tid = threadIdx.x;
  if (PASS > 0 && *flag) return;
  const int r0 = blockIdx.x * kCqRB;
  {
    // the panel's rows (and in the second pass L1^-1), all loads in flight
    // before the first LDS store
    constexpr int UX = kCqRB * kTsNB / 256;
    double xv[UX];
#pragma unroll
    for (int u = 0; u < UX; ++u) {
      const int e = tid + 256 * u, i = e >> 5, c = e & 31;
      xv[u] = r0 + i < mp && c < b ? A[(int64_t)(jb + r0 + i) * n + jb + c] : 0.0;
    }
    CqRegs li;
    if (PASS > 0) li = cq_fetch(q.Li);
#pragma unroll
    for (int u = 0; u < UX; ++u) {
      const int e = tid + 256 * u;
      X[(e >> 5) * kCqLd + (e & 31)] = xv[u];
    }
    if (PASS > 0) cq_put(li, Li);
  }
  __syncthreads();
  const double* Yr = X;
  if (PASS > 0) {
    cq_rows_transform(X, Li, Y, kCqRB, b);
    __syncthreads();
    for (int e = tid; e < kCqRB * kTsNB; e += 256) {
      const int i = e >> 5, c = e & 31;
      if (r0 + i < mp) q.Qb[(size_t)(r0 + i) * kTsNB + c] = Y[i * kCqLd + c];
    }
    Yr = Y;
  }
  cq_gram_partial(Yr, q.Gp);
  if (!cq_arrive_last(q.cnt)) return;
  // the reconstruction's inputs (L1 of the first pass, the top b rows of
  // Q1), loaded now so that they arrive during the reduction and Cholesky
  double pre_l1[kCqG / 256], pre_q1[kCqG / 256];
  if (PASS > 0) {
#pragma unroll
    for (int u = 0; u < kCqG / 256; ++u) {
      const int e = tid + 256 * u, i = e >> 5, c = e & 31;
      pre_l1[u] = q.L1s[e];
      pre_q1[u] = i < b && c < b ? q.Qb[(size_t)i * kTsNB + c] : 0.0;
    }
  }
  if (tid == 0) okf = 1;
  cq_gram_reduce(q.Gp, nblk, b, G);
  if (PASS == 0) {
    if (tid == 0) {
      double tr = 0.0;
      for (int i = 0; i < b; ++i) tr += G[i * kCqLd + i];
      *q.pnorm = sqrt(tr);
    }
  } else {
    for (int e = tid; e < kCqG; e += 256) {
      const int i = e >> 5, c = e & 31;
      if (c <= i && i < b && !(fabs(G[i * kCqLd + c] - (i == c ? 1.0 : 0.0)) <= 0.1)) okf = 0;
    }
    __syncthreads();
  }
  if (tid < 64) {
    const bool ok = wv_cholesky(G, Lo, rd, cb);
    if (tid == 0 && !ok) okf = 0;
  }
  __syncthreads();
  if (!okf) {
    if (tid == 0) { *flag = 1; *q.cnt = 0; }
    return;
  }
  blk_tri_inverse<false>(Lo, rd, Li, Y);  // Y is free: this block's rows are done
  if (PASS == 0) {
    cq_store(Li, q.Li);
    cq_store(Lo, q.L1s);
    if (tid == 0) { *flag = 0; *q.cnt = 0; }
    return;
  }
  // ---- after the second factor: R, then the reconstruction from Q1
  __shared__ double Ma[kTsNB * kCqLd];
  __shared__ double Mb[kTsNB * kCqLd];
  __shared__ double B[kTsNB * kCqLd];
  __shared__ double U[kTsNB * kCqLd];
  __shared__ double Lf[kTsNB * kCqLd];
  __shared__ double sg[kTsNB];
  __shared__ double rdU[kTsNB];
  // L1 -> Ma; the top b rows of Q1 (written by the first workgroup) -> X
  // rows 0..31
#pragma unroll
  for (int u = 0; u < kCqG / 256; ++u) {
    const int e = tid + 256 * u, i = e >> 5, c = e & 31;
    Ma[i * kCqLd + c] = pre_l1[u];
    X[i * kCqLd + c] = pre_q1[u];
  }
  __syncthreads();
  {  // R[i][c] = (L1 L2)[c][i] -> Mb (upper): one 16 x 16 tile per wave
    const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int c0 = (wave >> 1) * 16, i0 = (wave & 1) * 16;
    const qr_f64x4 r = qr_mfma_tile<kTsNB>(
        [&](int m, int k) { return Ma[(c0 + m) * kCqLd + k]; },
        [&](int k, int nn) { return Lo[k * kCqLd + i0 + nn]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = c0 + lk + 4 * v, i = i0 + lr;
      Mb[i * kCqLd + c] = i < b && c < b && i <= c ? r[v] : 0.0;
    }
  }
  cq_rows_transform(X, Li, B, kTsNB, b);  // B = the top b x b of Q
  __syncthreads();
  cq_store(Mb, q.Rs);
  cq_store(Li, q.Li);
  blk_lu(B, Lf, U, sg, rdU, cb);
  __syncthreads();
  blk_tri_inverse<false>(Lf, nullptr, Ma, Y);  // Ma = L1f^-1
  blk_tri_inverse<true>(U, rdU, G, Y);         // G = U^-1
  cq_store(Lf, q.L1f);
  cq_store(G, q.Ui);
  {  // T = -U S L1f^-T: one 16 x 16 tile per wave
    const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int i0 = (wave >> 1) * 16, j0 = (wave & 1) * 16;
    const qr_f64x4 t = qr_mfma_tile<kTsNB>(
        [&](int m, int k) { return U[(i0 + m) * kCqLd + k] * sg[k]; },
        [&](int k, int nn) { return Ma[(j0 + nn) * kCqLd + k]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + lk + 4 * v, j = j0 + lr;
      q.T[i * kTsNB + j] = i < b && j < b && i <= j ? -t[v] : 0.0;
    }
  }
  for (int e = tid; e < kCqG; e += 256) {
    const int i = e >> 5, j = e & 31;
    q.Rbuf[e] = i < b && j < b && j >= i ? sg[i] * Mb[i * kCqLd + j] : 0.0;
  }
  if (tid == 0) *q.cnt = 0;
}

// Rows of Q = Q1 R2^-1, the check of P - Q R, and the rows of V, Y1 = V T^T
// and Y2 = V T (V's top b rows are L1 of the LU), kCqRows rows per
// workgroup. A workgroup whose rows fail the check sets the flag; with
// put_r the workgroups of the top rows then write the accepted R into A
// (the TSQR root does that when its kernels follow).
constexpr int kCqRows = 16;

__global__ __launch_bounds__(256) void cq_rows_kernel(double* __restrict__ A, int n, int jb,
                                                     int mp, int b, CqBuf q,
                                                     double* __restrict__ Vall,
                                                     double* __restrict__ Y1,
                                                     double* __restrict__ Y2,
                                                     int* __restrict__ flag, int put_r) {
  __shared__ double X[kCqRows * kCqLd];
  __shared__ double Qr[kCqRows * kCqLd];
  __shared__ double Li[kTsNB * kCqLd];
  __shared__ double Ui[kTsNB * kCqLd];
  __shared__ double T[kTsNB * kCqLd];
  __shared__ double Lf[kTsNB * kCqLd];
  __shared__ double Rs[kTsNB * kCqLd];
  __shared__ double Vt[kCqRows * kCqLd];
  __shared__ double Pr[kCqRows * kCqLd];
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (*flag) return;
  const int r0 = blockIdx.x * kCqRows;
  if (tid == 0) bad = 0;
  {
    // five 32 x 32 factors, the rows of Q1 and, for the check, of P: all
    // loads in flight before the first LDS store
    const CqRegs li = cq_fetch(q.Li), ui = cq_fetch(q.Ui), tt = cq_fetch(q.T),
                 lf = cq_fetch(q.L1f), rs = cq_fetch(q.Rs);
    constexpr int UR = kCqRows * kTsNB / 256;
    double xv[UR], pv[UR];
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const int e = tid + 256 * u, i = e >> 5, c = e & 31;
      const bool in = r0 + i < mp && c < b;
      xv[u] = in ? q.Qb[(size_t)(r0 + i) * kTsNB + c] : 0.0;
      pv[u] = in ? A[(int64_t)(jb + r0 + i) * n + jb + c] : 0.0;
    }
    cq_put(li, Li);
    cq_put(ui, Ui);
    cq_put(tt, T);
    cq_put(lf, Lf);
    cq_put(rs, Rs);
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const int e = tid + 256 * u;
      X[(e >> 5) * kCqLd + (e & 31)] = xv[u];
      Pr[(e >> 5) * kCqLd + (e & 31)] = pv[u];
    }
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  if (wave < 2) {  // Q = X Li^T, one 16-column tile per wave
    const int c0 = wave * 16;
    const qr_f64x4 y = qr_mfma_tile<kTsNB>(
        [&](int m, int k) { return X[m * kCqLd + k]; },
        [&](int k, int nn) { return Li[(c0 + nn) * kCqLd + k]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = c0 + lr;
      Qr[(lk + 4 * v) * kCqLd + c] = c < b ? y[v] : 0.0;
    }
  }
  __syncthreads();
  if (wave < 2) {  // |P - Q R| against 2e-14 ||P||_F
    const int c0 = wave * 16;
    const double thr = 2e-14 * *q.pnorm;
    const qr_f64x4 d = qr_mfma_tile<kTsNB>(
        [&](int m, int k) { return Qr[m * kCqLd + k]; },
        [&](int k, int nn) { return Rs[k * kCqLd + c0 + nn]; });
    bool fail = false;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = lk + 4 * v, c = c0 + lr;
      if (c < b && r0 + r < mp && !(fabs(Pr[r * kCqLd + c] - d[v]) <= thr)) fail = true;
    }
    if (fail) bad = 1;
  } else {  // V rows: L1f for the top b rows, Q Ui below
    const int c0 = (wave - 2) * 16;
    const qr_f64x4 w = qr_mfma_tile<kTsNB>(
        [&](int m, int k) { return Qr[m * kCqLd + k]; },
        [&](int k, int nn) { return Ui[k * kCqLd + c0 + nn]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = lk + 4 * v, c = c0 + lr, pr = r0 + r;
      double val = 0.0;
      if (c < b) val = pr < b ? (c < pr ? Lf[pr * kCqLd + c] : (c == pr ? 1.0 : 0.0)) : w[v];
      Vt[r * kCqLd + c] = val;
    }
  }
  __syncthreads();
  if (bad) {
    if (tid == 0) *flag = 1;
    return;
  }
  if (put_r && r0 < b) {
    // the accepted R (S R) into the panel's top rows; this workgroup has
    // read those rows already and no other workgroup reads them
    for (int e = tid; e < kCqRows * kTsNB; e += 256) {
      const int i = e >> 5, c = e & 31, pr = r0 + i;
      if (pr < b && c < b && c >= pr) A[(int64_t)(jb + pr) * n + jb + c] = q.Rbuf[pr * kTsNB + c];
    }
  }
  {  // Y1 = V T^T (waves 0, 1), Y2 = V T (waves 2, 3); V itself with Y1
    const int c0 = (wave & 1) * 16;
    const bool second = wave >= 2;
    const qr_f64x4 y = qr_mfma_tile<kTsNB>(
        [&](int m, int k) { return Vt[m * kCqLd + k]; },
        [&](int k, int nn) { return second ? T[k * kCqLd + c0 + nn] : T[(c0 + nn) * kCqLd + k]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = lk + 4 * v, c = c0 + lr, pr = r0 + r;
      if (pr < mp && c < b) {
        const int64_t off = (int64_t)(jb + pr) * n + jb + c;
        if (second) {
          Y2[off] = y[v];
        } else {
          Y1[off] = y[v];
          Vall[off] = Vt[r * kCqLd + c];
        }
      }
    }
  }
}

// Compact-WY factor of four consecutive panels (kQrGroup = 128 columns):
// H_0 H_1 H_2 H_3 = I - V T V^T with T upper triangular, its diagonal
// blocks the panels' T_j and above them, column block by column block,
// T[0:32j, j] = -T[0:32j, 0:32j] (V_{0:j}^T V_j) T_j. G = V^T V (ld 128),
// Tp the four panel factors (32 x 32 each, consecutive). One workgroup;
// T and the product Z = (V_{0:j}^T V_j) T_j stay in LDS (rows padded by
// one against bank conflicts), 16 x 16 MFMA tiles, one per wave at a time.
constexpr int kQrGroup = 4 * kTsNB;
constexpr int kQrGLd = kQrGroup + 1;
__global__ __launch_bounds__(256) void qr_group_t_kernel(const double* __restrict__ G,
                                                         const double* __restrict__ Tp,
                                                         double* __restrict__ TG) {
  // one workgroup per group: its V^T V, its four panel T's, its T_G
  G += (size_t)blockIdx.x * kQrGroup * kQrGroup;
  Tp += (size_t)blockIdx.x * (kQrGroup / kTsNB) * kCqG;
  TG += (size_t)blockIdx.x * kQrGroup * kQrGroup;
  __shared__ double Ts[kQrGroup * kQrGLd];
  __shared__ double Z[3 * kTsNB * kCqLd];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  {
    // the four diagonal blocks: 16 values per thread, loads issued together
    double tv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) tv[u] = Tp[tid + 256 * u];
    for (int e = tid; e < kQrGroup * kQrGLd; e += 256) Ts[e] = 0.0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = tid + 256 * u, blk = e / kCqG, i = (e % kCqG) / kTsNB, c = e % kTsNB;
      Ts[(blk * kTsNB + i) * kQrGLd + blk * kTsNB + c] = tv[u];
    }
  }
  __syncthreads();
  for (int j = 1; j < 4; ++j) {
    const int rows = kTsNB * j, cj = kTsNB * j;
    // Z = G[0:rows, cj:cj+32] T_j
    for (int tile = wave; tile < (rows / 16) * 2; tile += 4) {
      const int r0 = (tile >> 1) * 16, c0 = (tile & 1) * 16;
      const qr_f64x4 z = qr_mfma_tile<kTsNB>(
          [&](int m, int k) { return G[(r0 + m) * kQrGroup + cj + k]; },
          [&](int k, int nn) { return Ts[(cj + k) * kQrGLd + cj + c0 + nn]; });
#pragma unroll
      for (int v = 0; v < 4; ++v) Z[(r0 + lk + 4 * v) * kCqLd + c0 + lr] = z[v];
    }
    __syncthreads();
    // T[0:rows, cj:cj+32] = -T[0:rows, 0:rows] Z
    for (int tile = wave; tile < (rows / 16) * 2; tile += 4) {
      const int r0 = (tile >> 1) * 16, c0 = (tile & 1) * 16;
      const qr_f64x4 t = qr_mfma_tile_rt(
          [&](int m, int k) { return Ts[(r0 + m) * kQrGLd + k]; },
          [&](int k, int nn) { return Z[k * kCqLd + c0 + nn]; }, rows);
#pragma unroll
      for (int v = 0; v < 4; ++v) Ts[(r0 + lk + 4 * v) * kQrGLd + cj + c0 + lr] = -t[v];
    }
    __syncthreads();
  }
  for (int e = tid; e < kQrGroup * kQrGroup; e += 256)
    TG[e] = Ts[(e / kQrGroup) * kQrGLd + e % kQrGroup];
}

__global__ void fill_zero_kernel(double* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

static bool ts_fits(int m, int n) {
  TsPlan p;
  return m >= n && ts_plan(m, &p);  // the first panel has the most rows
}

static size_t ts_ws_bytes(int m, int n) {
  size_t b = 3 * al((size_t)m * n * 8);                  // Vall, Y1 = V T^T, Y2 = V T
  b += al((size_t)m * kTsNB * 8);                        // leaf explicit Q rows
  b += al((size_t)kTsMaxLeaves * kTsNB * kTsNB * 8);     // leaf R factors
  b += al((size_t)kTsMaxLeaves * kTsNB * kTsNB * 8);     // Q_top
  b += 2 * al((size_t)kTsNB * kTsNB * 8);                // T, U
  b += al((size_t)kTsNB * n * 8);                        // W
  b += al((size_t)32 * kTsNB * n * 8);                   // split-K partials
  b += al((size_t)n * 8);                                // R diagonal signs
  b += al((size_t)m * kTsNB * 8);                        // Cholesky-QR panel rows
  b += al((size_t)((m + kCqRB - 1) / kCqRB) * kCqG * 8);  // Gram partials
  b += 7 * al((size_t)kCqG * 8);                         // Li, L1, R, S R, L1f, U^-1, T
  b += al(8) + al(sizeof(int));                          // ||P||_F, arrival ticket
  b += al((size_t)((n + kTsNB - 1) / kTsNB) * sizeof(int));    // per-panel fallback flags
  b += al((size_t)m * n * 8);                            // copy of A for a refactorisation
  b += al((size_t)((n + kTsNB - 1) / kTsNB) * kCqG * 8);  // every panel's T
  b += 2 * al((size_t)(n / kQrGroup + 1) * kQrGroup * kQrGroup * 8);  // every group's T, V^T V
  b += al((size_t)kQrGroup * n * 8);                     // group W
  return b + 256;
}

size_t qr_ws_bytes(int m, int n) {
  if (ts_fits(m, n)) return ts_ws_bytes(m, n);
  QrCfg c;
  if (!qr_config(m, &c)) c = {1024, 8, 4};
  const int np = (n + c.nb - 1) / c.nb;
  size_t b = al((size_t)m * n * 8);                 // Vall
  b += al((size_t)np * c.nb * c.nb * 8);            // T factors
  b += 2 * al((size_t)c.nb * n * 8);                // W, W2
  b += al((size_t)32 * c.nb * n * 8);               // split-K partials
  b += al((size_t)n * 8);                           // R diagonal signs
  return b + 256;
}

static int launch_panel(const QrCfg& c, double* A, int m, int n, int jb, int b, double* Vall,
                        double* Tp, hipStream_t s) {
#define TNX_PANEL(NT_, RPT_, NB_)                                                        \
  if (c.nt == NT_ && c.rpt == RPT_ && c.nb == NB_) {                                     \
    hipLaunchKernelGGL((qr_panel_kernel<NT_, RPT_, NB_>), dim3(1), dim3(NT_), 0, s, A, m, n, \
                       jb, b, Vall, Tp);                                                 \
    TNX_LAUNCH_CHECK();                                                                  \
    return TNX_OK;                                                                       \
  }
  TNX_PANEL(512, 1, 16)
  TNX_PANEL(512, 2, 16)
  TNX_PANEL(512, 4, 16)
  TNX_PANEL(1024, 4, 8)
  TNX_PANEL(1024, 8, 4)
#undef TNX_PANEL
  set_error("qr: no panel configuration");
  return TNX_ERR_UNSUPPORTED;
}

// Blocked QR with Cholesky-QR2 panels (three kernels per panel) and two
// GEMMs per panel for the trailing update; a panel that fails the
// Cholesky-QR checks sends the whole factorisation to the TSQR panel path
// (leaf, root, apply after every panel). Q formed backwards with two GEMMs
// per panel.
static int qr_tsqr(double* A, int m, int n, double* R, void* ws, hipStream_t s) {
  char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* Vall = (double*)p; p += al((size_t)m * n * 8);
  double* Y1 = (double*)p; p += al((size_t)m * n * 8);
  double* Y2 = (double*)p; p += al((size_t)m * n * 8);
  double* Qst = (double*)p; p += al((size_t)m * kTsNB * 8);
  double* Rst = (double*)p; p += al((size_t)kTsMaxLeaves * kTsNB * kTsNB * 8);
  double* Qtop = (double*)p; p += al((size_t)kTsMaxLeaves * kTsNB * kTsNB * 8);
  double* Tst = (double*)p; p += al((size_t)kTsNB * kTsNB * 8);
  double* Ust = (double*)p; p += al((size_t)kTsNB * kTsNB * 8);
  double* Wb = (double*)p; p += al((size_t)kTsNB * n * 8);
  double* part = (double*)p; p += al((size_t)32 * kTsNB * n * 8);
  double* sg = (double*)p; p += al((size_t)n * 8);
  CqBuf q;
  q.Qb = (double*)p; p += al((size_t)m * kTsNB * 8);
  q.Gp = (double*)p; p += al((size_t)((m + kCqRB - 1) / kCqRB) * kCqG * 8);
  double** smalls[7] = {&q.Li, &q.L1s, &q.Rs, &q.Rbuf, &q.L1f, &q.Ui, &q.T};
  for (double** sm : smalls) { *sm = (double*)p; p += al((size_t)kCqG * 8); }
  q.pnorm = (double*)p; p += al(8);
  q.cnt = (int*)p; p += al(sizeof(int));
  const int np = (n + kTsNB - 1) / kTsNB;
  int* flags = (int*)p; p += al((size_t)np * sizeof(int));
  double* Asave = (double*)p; p += al((size_t)m * n * 8);
  double* Tall = (double*)p; p += al((size_t)np * kCqG * 8);
  double* TG = (double*)p; p += al((size_t)(n / kQrGroup + 1) * kQrGroup * kQrGroup * 8);
  double* GG = (double*)p; p += al((size_t)(n / kQrGroup + 1) * kQrGroup * kQrGroup * 8);
  double* Wg = (double*)p;
  // Q is formed four panels at a time (one 128-wide block reflector) when
  // n is a whole number of such groups and no panel fell back to TSQR; the
  // group's V then needs zeros above each panel's diagonal block.
  const bool group_q = n % kQrGroup == 0;
  if (group_q) {
    const int64_t mn = (int64_t)m * n;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((int)std::min<int64_t>((mn + 255) / 256, 4096)),
                       dim3(256), 0, s, Vall, mn);
    TNX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(cq_init_kernel, dim3(1), dim3(1), 0, s, q.cnt);
  TNX_LAUNCH_CHECK();
  const size_t part_bytes = al((size_t)32 * kTsNB * n * 8);
  // One pass over the panels. deferred: Cholesky-QR panels only (the rows
  // kernel writes each accepted R into A); a panel that fails its checks
  // only sets its flag, and the caller refactors. Otherwise the TSQR kernels
  // follow every panel and factor the panels that were flagged.
  auto panels = [&](bool deferred) -> int {
    for (int pi = 0; pi < np; ++pi) {
      const int jb = pi * kTsNB;
      const int b = std::min(kTsNB, n - jb);
      const int mp = m - jb;
      TsPlan pl;
      if (!ts_plan(mp, &pl)) {
        set_error("qr: panel of %d rows has no TSQR plan", mp);
        return TNX_ERR_UNSUPPORTED;
      }
      const int nblk = (mp + kCqRB - 1) / kCqRB;
      int* fl = flags + pi;
      CqBuf qp = q;
      qp.T = Tall + (size_t)pi * kCqG;  // kept for the grouped Q formation
      hipLaunchKernelGGL(cq_pass_kernel<0>, dim3(nblk), dim3(256), 0, s, (const double*)A, n, jb,
                         mp, b, qp, nblk, fl);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(cq_pass_kernel<1>, dim3(nblk), dim3(256), 0, s, (const double*)A, n, jb,
                         mp, b, qp, nblk, fl);
      TNX_LAUNCH_CHECK();
      hipLaunchKernelGGL(cq_rows_kernel, dim3((mp + kCqRows - 1) / kCqRows), dim3(256), 0, s, A, n,
                         jb, mp, b, qp, Vall, Y1, Y2, fl, deferred ? 1 : 0);
      TNX_LAUNCH_CHECK();
      if (!deferred) {
        hipLaunchKernelGGL(tsqr_leaf_kernel, dim3(pl.nleaf), dim3(256), 0, s, A, n, jb, mp, b,
                           pl.h, pl.nleaf, Rst, Qst, fl);
        TNX_LAUNCH_CHECK();
        hipLaunchKernelGGL(tsqr_root_kernel, dim3(1), dim3(512), 0, s, Rst, pl.nleaf, b, Qst, A, n,
                           jb, Vall, Tst, Qtop, Ust, fl, (const double*)q.Rbuf);
        TNX_LAUNCH_CHECK();
        hipLaunchKernelGGL(tsqr_apply_kernel, dim3(pl.nleaf), dim3(256), 0, s, Qst, Qtop, Ust, Tst,
                           mp, b, pl.h, pl.nleaf, n, jb, Vall, Y1, Y2, fl);
        TNX_LAUNCH_CHECK();
      }
      const int nt = n - jb - b;
      if (nt <= 0) continue;
      const double* Vp = Vall + (size_t)jb * n + jb;
      const double* Y1p = Y1 + (size_t)jb * n + jb;
      double* At = A + (size_t)jb * n + jb + b;
      // W = V^T A_trail   (b x nt)
      GemmDesc g{};
      g.M = b; g.N = nt; g.K = mp; g.batch = 1;
      g.A = Vp; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = false;
      g.B = At; g.tb = DT::F64; g.ldb = n; g.sB = 0; g.b_ncontig = true;
      g.C = Wb; g.tc = DT::F64; g.ldc = nt; g.sC = 0;
      g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
      g.splitk_ws = part; g.splitk_ws_bytes = part_bytes;
      TNX_TRY(gemm(g, s));
      // A_trail -= (V T^T) W
      GemmDesc g3{};
      g3.M = mp; g3.N = nt; g3.K = b; g3.batch = 1;
      g3.A = Y1p; g3.ta = DT::F64; g3.lda = n; g3.sA = 0; g3.a_kcontig = true;
      g3.B = Wb; g3.tb = DT::F64; g3.ldb = nt; g3.sB = 0; g3.b_ncontig = true;
      g3.C = At; g3.tc = DT::F64; g3.ldc = n; g3.sC = 0;
      g3.alpha = -1.0; g3.beta = 1.0; g3.prec = Prec::F64;
      TNX_TRY(gemm(g3, s));
    }
    return TNX_OK;
  };
  // R, then Q formed backwards: 128-wide groups, or one panel at a time.
  auto finish = [&](bool grouped) -> int {
    {
      const int64_t nn = (int64_t)n * n;
      const int blocks = (int)std::min<int64_t>((nn + 255) / 256, 4096);
      hipLaunchKernelGGL(extract_r_kernel, dim3(blocks), dim3(256), 0, s, A, n, R);
      TNX_LAUNCH_CHECK();
      const int64_t mn = (int64_t)m * n;
      const int b2 = (int)std::min<int64_t>((mn + 255) / 256, 4096);
      hipLaunchKernelGGL(eye_kernel, dim3(b2), dim3(256), 0, s, A, m, n);
      TNX_LAUNCH_CHECK();
    }
    if (grouped) {
      // Q = H_G0 ... H_G(ng-1) [I; 0] backwards, one 128-wide group at a time:
      // Q_s -= (V_G T_G) (V_G^T Q_s) on rows and columns from the group's
      // first. What depends on the panels only comes first, batched over the
      // groups: every V_G^T V_G in one GEMM over all m rows (V is zero above
      // each panel's diagonal block, so the rows above the group add
      // nothing), every T_G (from its four panel T's and V_G^T V_G) in one
      // launch of a workgroup per group, and every Y_G = V_G T_G in one GEMM
      // (into Y2, which only the panel-by-panel form reads).
      const int ng = n / kQrGroup;
      {
        GemmDesc gg{};
        gg.M = kQrGroup; gg.N = kQrGroup; gg.K = m; gg.batch = ng;
        gg.A = Vall; gg.ta = DT::F64; gg.lda = n; gg.sA = kQrGroup; gg.a_kcontig = false;
        gg.B = Vall; gg.tb = DT::F64; gg.ldb = n; gg.sB = kQrGroup; gg.b_ncontig = true;
        gg.C = GG; gg.tc = DT::F64; gg.ldc = kQrGroup; gg.sC = (int64_t)kQrGroup * kQrGroup;
        gg.alpha = 1.0; gg.beta = 0.0; gg.prec = Prec::F64;
        gg.splitk_ws = part; gg.splitk_ws_bytes = part_bytes;
        TNX_TRY(gemm(gg, s));
        hipLaunchKernelGGL(qr_group_t_kernel, dim3(ng), dim3(256), 0, s, (const double*)GG,
                           (const double*)Tall, TG);
        TNX_LAUNCH_CHECK();
        GemmDesc gy{};
        gy.M = m; gy.N = kQrGroup; gy.K = kQrGroup; gy.batch = ng;
        gy.A = Vall; gy.ta = DT::F64; gy.lda = n; gy.sA = kQrGroup; gy.a_kcontig = true;
        gy.B = TG; gy.tb = DT::F64; gy.ldb = kQrGroup; gy.sB = (int64_t)kQrGroup * kQrGroup;
        gy.b_ncontig = true;
        gy.C = Y2; gy.tc = DT::F64; gy.ldc = n; gy.sC = kQrGroup;
        gy.alpha = 1.0; gy.beta = 0.0; gy.prec = Prec::F64;
        TNX_TRY(gemm(gy, s));
      }
      for (int gi = ng - 1; gi >= 0; --gi) {
        const int g0 = gi * kQrGroup;
        const int mp = m - g0, nq = n - g0;
        const double* Vp = Vall + (size_t)g0 * n + g0;
        double* Qs = A + (size_t)g0 * n + g0;
        const double* Yp = Y2 + (size_t)g0 * n + g0;
        GemmDesc g{};
        g.M = kQrGroup; g.N = nq; g.K = mp; g.batch = 1;
        g.A = Vp; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = false;
        g.B = Qs; g.tb = DT::F64; g.ldb = n; g.sB = 0; g.b_ncontig = true;
        g.C = Wg; g.tc = DT::F64; g.ldc = nq; g.sC = 0;
        g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
        g.splitk_ws = part; g.splitk_ws_bytes = part_bytes;
        TNX_TRY(gemm(g, s));
        GemmDesc g3{};
        g3.M = mp; g3.N = nq; g3.K = kQrGroup; g3.batch = 1;
        g3.A = Yp; g3.ta = DT::F64; g3.lda = n; g3.sA = 0; g3.a_kcontig = true;
        g3.B = Wg; g3.tb = DT::F64; g3.ldb = nq; g3.sB = 0; g3.b_ncontig = true;
        g3.C = Qs; g3.tc = DT::F64; g3.ldc = n; g3.sC = 0;
        g3.alpha = -1.0; g3.beta = 1.0; g3.prec = Prec::F64;
        TNX_TRY(gemm(g3, s));
      }
    } else {
      for (int pi = np - 1; pi >= 0; --pi) {
        const int jb = pi * kTsNB;
        const int b = std::min(kTsNB, n - jb);
        const int mp = m - jb;
        const int nq = n - jb;
        const double* Vp = Vall + (size_t)jb * n + jb;
        const double* Y2p = Y2 + (size_t)jb * n + jb;
        double* Qs = A + (size_t)jb * n + jb;
        GemmDesc g{};
        g.M = b; g.N = nq; g.K = mp; g.batch = 1;
        g.A = Vp; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = false;
        g.B = Qs; g.tb = DT::F64; g.ldb = n; g.sB = 0; g.b_ncontig = true;
        g.C = Wb; g.tc = DT::F64; g.ldc = nq; g.sC = 0;
        g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
        g.splitk_ws = part; g.splitk_ws_bytes = part_bytes;
        TNX_TRY(gemm(g, s));
        GemmDesc g3{};
        g3.M = mp; g3.N = nq; g3.K = b; g3.batch = 1;
        g3.A = Y2p; g3.ta = DT::F64; g3.lda = n; g3.sA = 0; g3.a_kcontig = true;
        g3.B = Wb; g3.tb = DT::F64; g3.ldb = nq; g3.sB = 0; g3.b_ncontig = true;
        g3.C = Qs; g3.tc = DT::F64; g3.ldc = n; g3.sC = 0;
        g3.alpha = -1.0; g3.beta = 1.0; g3.prec = Prec::F64;
        TNX_TRY(gemm(g3, s));
      }
    }
    return qr_fix_signs(A, m, R, n, sg, s);
  };
  // Cholesky-QR panels without the TSQR kernels (three launches per panel
  // that return at once when the panel is accepted), then one look at the
  // flags. The flags are copied into a pinned block and the host waits for
  // that copy only, while R and Q, formed as if every panel was accepted,
  // stay queued behind it; if a panel failed, A is restored, factored again
  // with the TSQR fallback after every panel, and R and Q formed again. A
  // thread whose last QR fell back waits for the flags before queueing R
  // and Q (fallbacks come in runs: the same kind of matrix again).
  static thread_local bool last_fell_back = false;
  const size_t abytes = (size_t)m * n * 8;
  TNX_HIP(hipMemcpyAsync(Asave, A, abytes, hipMemcpyDeviceToDevice, s));
  TNX_TRY(panels(true));
  PinnedLease pin(((size_t)np + 1) / 2);
  StepEvent fev;
  bool queued = false, failed = false;
  if (pin.get() && fev.e && !last_fell_back) {
    int* hf = reinterpret_cast<int*>(pin.get());
    TNX_HIP(hipMemcpyAsync(hf, flags, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, s));
    TNX_HIP(hipEventRecord(fev.e, s));
    const int st = finish(group_q);
    // the copy must land before the block goes back to the pool
    const hipError_t we = hipEventSynchronize(fev.e);
    TNX_TRY(st);
    TNX_HIP(we);
    queued = true;
    failed = std::any_of(hf, hf + np, [](int f) { return f != 0; });
  } else {
    std::vector<int> hf(np);
    TNX_HIP(hipMemcpyAsync(hf.data(), flags, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, s));
    TNX_HIP(hipStreamSynchronize(s));
    failed = std::any_of(hf.begin(), hf.end(), [](int f) { return f != 0; });
  }
  last_fell_back = failed;
  if (failed) {
    TNX_HIP(hipMemcpyAsync(A, Asave, abytes, hipMemcpyDeviceToDevice, s));
    TNX_TRY(panels(false));
    return finish(false);
  }
  return queued ? TNX_OK : finish(group_q);
}

int qr(double* A, int m, int n, double* R, void* ws, size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(A && R, "qr: null pointer");
  TNX_CHECK_ARG(m >= 1 && n >= 1 && m >= n, "qr: need m >= n >= 1 (m=%d n=%d)", m, n);
  if (ts_fits(m, n)) {
    TNX_CHECK_ARG(ws && ws_bytes >= qr_ws_bytes(m, n), "qr: workspace too small");
    return qr_tsqr(A, m, n, R, ws, s);
  }
  QrCfg c;
  if (!qr_config(m, &c)) {
    set_error("qr: m=%d rows exceeds the supported 8192", m);
    return TNX_ERR_UNSUPPORTED;
  }
  TNX_CHECK_ARG(ws && ws_bytes >= qr_ws_bytes(m, n), "qr: workspace too small");
  const int NB = c.nb;
  const int np = (n + NB - 1) / NB;
  char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* Vall = (double*)p; p += al((size_t)m * n * 8);
  double* Tall = (double*)p; p += al((size_t)np * NB * NB * 8);
  double* Wb = (double*)p; p += al((size_t)NB * n * 8);
  double* W2 = (double*)p; p += al((size_t)NB * n * 8);
  double* part = (double*)p; p += al((size_t)32 * NB * n * 8);
  double* sg = (double*)p;
  const size_t part_bytes = al((size_t)32 * NB * n * 8);

  // ---- factorisation
  for (int pi = 0; pi < np; ++pi) {
    const int jb = pi * NB;
    const int b = std::min(NB, n - jb);
    const int mp = m - jb;
    double* Tp = Tall + (size_t)pi * NB * NB;
    TNX_TRY(launch_panel(c, A, m, n, jb, b, Vall, Tp, s));
    const int nt = n - jb - b;
    if (nt <= 0) continue;
    const double* Vp = Vall + (size_t)jb * n + jb;
    double* At = A + (size_t)jb * n + jb + b;
    // W = V^T A_trail   (b x nt)
    GemmDesc g{};
    g.M = b; g.N = nt; g.K = mp; g.batch = 1;
    g.A = Vp; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = false;
    g.B = At; g.tb = DT::F64; g.ldb = n; g.sB = 0; g.b_ncontig = true;
    g.C = Wb; g.tc = DT::F64; g.ldc = nt; g.sC = 0;
    g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
    g.splitk_ws = part; g.splitk_ws_bytes = part_bytes;
    TNX_TRY(gemm(g, s));
    // W2 = T^T W
    GemmDesc g2{};
    g2.M = b; g2.N = nt; g2.K = b; g2.batch = 1;
    g2.A = Tp; g2.ta = DT::F64; g2.lda = NB; g2.sA = 0; g2.a_kcontig = false;
    g2.B = Wb; g2.tb = DT::F64; g2.ldb = nt; g2.sB = 0; g2.b_ncontig = true;
    g2.C = W2; g2.tc = DT::F64; g2.ldc = nt; g2.sC = 0;
    g2.alpha = 1.0; g2.beta = 0.0; g2.prec = Prec::F64;
    TNX_TRY(gemm(g2, s));
    // A_trail -= V W2
    GemmDesc g3{};
    g3.M = mp; g3.N = nt; g3.K = b; g3.batch = 1;
    g3.A = Vp; g3.ta = DT::F64; g3.lda = n; g3.sA = 0; g3.a_kcontig = true;
    g3.B = W2; g3.tb = DT::F64; g3.ldb = nt; g3.sB = 0; g3.b_ncontig = true;
    g3.C = At; g3.tc = DT::F64; g3.ldc = n; g3.sC = 0;
    g3.alpha = -1.0; g3.beta = 1.0; g3.prec = Prec::F64;
    TNX_TRY(gemm(g3, s));
  }
  // ---- R, then Q = H_0 ... H_{np-1} [I; 0] accumulated backwards into A
  {
    const int64_t nn = (int64_t)n * n;
    const int blocks = (int)std::min<int64_t>((nn + 255) / 256, 4096);
    hipLaunchKernelGGL(extract_r_kernel, dim3(blocks), dim3(256), 0, s, A, n, R);
    TNX_LAUNCH_CHECK();
    const int64_t mn = (int64_t)m * n;
    const int b2 = (int)std::min<int64_t>((mn + 255) / 256, 4096);
    hipLaunchKernelGGL(eye_kernel, dim3(b2), dim3(256), 0, s, A, m, n);
    TNX_LAUNCH_CHECK();
  }
  for (int pi = np - 1; pi >= 0; --pi) {
    const int jb = pi * NB;
    const int b = std::min(NB, n - jb);
    const int mp = m - jb;
    const int nq = n - jb;
    const double* Tp = Tall + (size_t)pi * NB * NB;
    const double* Vp = Vall + (size_t)jb * n + jb;
    double* Qs = A + (size_t)jb * n + jb;
    GemmDesc g{};
    g.M = b; g.N = nq; g.K = mp; g.batch = 1;
    g.A = Vp; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = false;
    g.B = Qs; g.tb = DT::F64; g.ldb = n; g.sB = 0; g.b_ncontig = true;
    g.C = Wb; g.tc = DT::F64; g.ldc = nq; g.sC = 0;
    g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
    g.splitk_ws = part; g.splitk_ws_bytes = part_bytes;
    TNX_TRY(gemm(g, s));
    GemmDesc g2{};
    g2.M = b; g2.N = nq; g2.K = b; g2.batch = 1;
    g2.A = Tp; g2.ta = DT::F64; g2.lda = NB; g2.sA = 0; g2.a_kcontig = true;
    g2.B = Wb; g2.tb = DT::F64; g2.ldb = nq; g2.sB = 0; g2.b_ncontig = true;
    g2.C = W2; g2.tc = DT::F64; g2.ldc = nq; g2.sC = 0;
    g2.alpha = 1.0; g2.beta = 0.0; g2.prec = Prec::F64;
    TNX_TRY(gemm(g2, s));
    GemmDesc g3{};
    g3.M = mp; g3.N = nq; g3.K = b; g3.batch = 1;
    g3.A = Vp; g3.ta = DT::F64; g3.lda = n; g3.sA = 0; g3.a_kcontig = true;
    g3.B = W2; g3.tb = DT::F64; g3.ldb = nq; g3.sB = 0; g3.b_ncontig = true;
    g3.C = Qs; g3.tc = DT::F64; g3.ldc = n; g3.sC = 0;
    g3.alpha = -1.0; g3.beta = 1.0; g3.prec = Prec::F64;
    TNX_TRY(gemm(g3, s));
  }
  return qr_fix_signs(A, m, R, n, sg, s);
}

// ---- gauge moves -------------------------------------------------------------
size_t gauge_ws_bytes(int chil, int d, int chir) {
  const int64_t m1 = (int64_t)chil * d, n1 = chir;          // rightward
  const int64_t m2 = (int64_t)d * chir, n2 = chil;          // leftward (transposed)
  size_t a = 0, b = 0;
  if (m1 >= n1) a = al((size_t)n1 * n1 * 8) + qr_ws_bytes((int)m1, (int)n1);
  if (m2 >= n2) b = al((size_t)m2 * n2 * 8) + al((size_t)n2 * n2 * 8) + qr_ws_bytes((int)m2, (int)n2);
  return std::max(a, b) + 256;
}

int gauge_right(double* A, const double* next, double* next_out, int chil, int d, int chir,
                int next_cols, void* ws, size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(A && next && next_out && next != next_out, "gauge_right: bad pointers");
  TNX_CHECK_ARG(chil > 0 && d > 0 && chir > 0 && next_cols > 0, "gauge_right: bad dims");
  const int m = chil * d, n = chir;
  TNX_CHECK_ARG(m >= n, "gauge_right: chil*d=%d < chir=%d (bond would shrink)", m, n);
  TNX_CHECK_ARG(ws && ws_bytes >= gauge_ws_bytes(chil, d, chir), "gauge_right: workspace too small");
  char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* R = (double*)p; p += al((size_t)n * n * 8);
  TNX_TRY(qr(A, m, n, R, p, qr_ws_bytes(m, n), s));
  // next_out = R * next
  GemmDesc g{};
  g.M = n; g.N = next_cols; g.K = n; g.batch = 1;
  g.A = R; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = true;
  g.B = next; g.tb = DT::F64; g.ldb = next_cols; g.sB = 0; g.b_ncontig = true;
  g.C = next_out; g.tc = DT::F64; g.ldc = next_cols; g.sC = 0;
  g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
  return gemm(g, s);
}

int gauge_left(double* A, const double* prev, double* prev_out, int chil, int d, int chir,
               int prev_rows, void* ws, size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(A && prev && prev_out && prev != prev_out, "gauge_left: bad pointers");
  TNX_CHECK_ARG(chil > 0 && d > 0 && chir > 0 && prev_rows > 0, "gauge_left: bad dims");
  const int m = d * chir, n = chil;
  TNX_CHECK_ARG(m >= n, "gauge_left: d*chir=%d < chil=%d (bond would shrink)", m, n);
  TNX_CHECK_ARG(ws && ws_bytes >= gauge_ws_bytes(chil, d, chir), "gauge_left: workspace too small");
  char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* At = (double*)p; p += al((size_t)m * n * 8);
  double* R = (double*)p; p += al((size_t)n * n * 8);
  // A (chil x d*chir) -> At (d*chir x chil); At = Q R; A = Q^T; prev <- prev R^T
  TNX_TRY(transpose_f64(A, At, n, m, s));
  TNX_TRY(qr(At, m, n, R, p, qr_ws_bytes(m, n), s));
  TNX_TRY(transpose_f64(At, A, m, n, s));
  GemmDesc g{};
  g.M = prev_rows; g.N = n; g.K = n; g.batch = 1;
  g.A = prev; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = true;
  g.B = R; g.tb = DT::F64; g.ldb = n; g.sB = 0; g.b_ncontig = false;
  g.C = prev_out; g.tc = DT::F64; g.ldc = n; g.sC = 0;
  g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
  return gemm(g, s);
}

// ---- batched small QR ----------------------------------------------------------
// Householder QR of `batch` small m x n matrices (m >= n), one workgroup per
// matrix held in LDS (row stride n + 1). trans = 0: matrix p is A_p, m x n
// row-major. trans = 1: matrix p is A_p^T, where A_p is n x m row-major (the
// leftward gauge move: A = (Q R)^T). On exit A_p holds the thin Q (Q^T when
// trans) and R_p = R (n x n, upper triangular). Same reflector convention as
// the blocked path (beta = -sign(alpha) |x|).
constexpr int kQrSmallBig = 16640;   // m * (n + 1) doubles: 130 KiB of LDS
constexpr int kQrSmallLittle = 4352; // 34 KiB

__device__ __forceinline__ double qs_block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wsum(v);
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

template <int MAXE>
__global__ __launch_bounds__(256) void qr_small_kernel(double* __restrict__ A, int m, int n,
                                                       int trans, double* __restrict__ R) {
  __shared__ double S[MAXE];
  __shared__ double tau_s[256];
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ld = n + 1;
  const int64_t mn = (int64_t)m * n;
  double* Ap = A + (int64_t)blockIdx.x * mn;
  for (int e = tid; e < mn; e += 256) {
    int i, c;
    if (trans) { c = e / m; i = e - c * m; } else { i = e / n; c = e - i * n; }
    S[i * ld + c] = Ap[e];
  }
  __syncthreads();
  // ---- factorisation
  for (int j = 0; j < n; ++j) {
    double part = 0.0;
    for (int i = j + 1 + tid; i < m; i += 256) part += S[i * ld + j] * S[i * ld + j];
    const double sigma = qs_block_sum(part, red);
    const double alpha = S[j * ld + j];
    double tj = 0.0, beta = alpha, scale = 0.0;
    if (sigma != 0.0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < m; i += 256) S[i * ld + j] *= scale;
    if (tid == 0) { S[j * ld + j] = beta; tau_s[j] = tj; }
    __syncthreads();
    if (tj != 0.0) {
      for (int c = j + 1 + wave; c < n; c += 4) {
        double acc = 0.0;
        for (int i = j + lane; i < m; i += 64) acc += (i == j ? 1.0 : S[i * ld + j]) * S[i * ld + c];
        const double w = tj * wsum(acc);
        for (int i = j + lane; i < m; i += 64) S[i * ld + c] -= (i == j ? 1.0 : S[i * ld + j]) * w;
      }
    }
    __syncthreads();
  }
  // ---- R, with a non-negative diagonal (same convention as the blocked QR)
  __shared__ double sgn[256];
  for (int j = tid; j < n; j += 256) sgn[j] = S[j * ld + j] < 0.0 ? -1.0 : 1.0;
  __syncthreads();
  double* Rp = R + (int64_t)blockIdx.x * n * n;
  for (int e = tid; e < n * n; e += 256) {
    const int i = e / n, c = e - i * n;
    Rp[e] = (c >= i) ? sgn[i] * S[i * ld + c] : 0.0;
  }
  __syncthreads();
  // ---- thin Q = H_0 ... H_{n-1} [I; 0], formed in place from the last column
  for (int j = n - 1; j >= 0; --j) {
    const double tj = tau_s[j];
    if (tj != 0.0) {
      for (int c = j + 1 + wave; c < n; c += 4) {
        double acc = 0.0;
        for (int i = j + lane; i < m; i += 64) acc += (i == j ? 1.0 : S[i * ld + j]) * S[i * ld + c];
        const double w = tj * wsum(acc);
        for (int i = j + lane; i < m; i += 64) S[i * ld + c] -= (i == j ? 1.0 : S[i * ld + j]) * w;
      }
    }
    __syncthreads();
    for (int i = tid; i < m; i += 256)
      S[i * ld + j] = (i < j) ? 0.0 : (i == j ? 1.0 - tj : -tj * S[i * ld + j]);
    __syncthreads();
  }
  for (int e = tid; e < mn; e += 256) {
    int i, c;
    if (trans) { c = e / m; i = e - c * m; } else { i = e / n; c = e - i * n; }
    Ap[e] = sgn[c] * S[i * ld + c];
  }
}

static bool qr_small_fits(int m, int n) {
  return m >= n && n >= 1 && n <= 256 && (int64_t)m * (n + 1) <= kQrSmallBig;
}

static int qr_small_b(double* A, int m, int n, int trans, double* R, int batch, hipStream_t s) {
  TNX_CHECK_ARG(qr_small_fits(m, n), "qr_batched: %d x %d exceeds the LDS-resident limit", m, n);
  TNX_CHECK_ARG(batch >= 1 && batch <= 2147483647, "qr_batched: bad batch %d", batch);
  if ((int64_t)m * (n + 1) <= kQrSmallLittle)
    hipLaunchKernelGGL(qr_small_kernel<kQrSmallLittle>, dim3(batch), dim3(256), 0, s, A, m, n,
                       trans, R);
  else
    hipLaunchKernelGGL(qr_small_kernel<kQrSmallBig>, dim3(batch), dim3(256), 0, s, A, m, n, trans,
                       R);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

size_t gauge_b_ws_bytes(int chil, int d, int chir, int batch) {
  const size_t n = (size_t)std::max(chil, chir);
  return al(n * n * 8 * (size_t)batch) + 256;
}

int gauge_right_b(double* A, const double* next, double* next_out, int chil, int d, int chir,
                  int next_cols, int batch, void* ws, size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(A && next && next_out && next != next_out, "gauge_right_batched: bad pointers");
  TNX_CHECK_ARG(chil > 0 && d > 0 && chir > 0 && next_cols > 0 && batch > 0,
                "gauge_right_batched: bad dims");
  const int m = chil * d, n = chir;
  TNX_CHECK_ARG(m >= n, "gauge_right_batched: chil*d=%d < chir=%d", m, n);
  TNX_CHECK_ARG(ws && ws_bytes >= gauge_b_ws_bytes(chil, d, chir, batch),
                "gauge_right_batched: workspace too small");
  double* R = (double*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  TNX_TRY(qr_small_b(A, m, n, 0, R, batch, s));
  GemmDesc g{};
  g.M = n; g.N = next_cols; g.K = n; g.batch = batch;
  g.A = R; g.ta = DT::F64; g.lda = n; g.sA = (int64_t)n * n; g.a_kcontig = true;
  g.B = next; g.tb = DT::F64; g.ldb = next_cols; g.sB = (int64_t)n * next_cols; g.b_ncontig = true;
  g.C = next_out; g.tc = DT::F64; g.ldc = next_cols; g.sC = (int64_t)n * next_cols;
  g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
  return gemm(g, s);
}

int gauge_left_b(double* A, const double* prev, double* prev_out, int chil, int d, int chir,
                 int prev_rows, int batch, void* ws, size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(A && prev && prev_out && prev != prev_out, "gauge_left_batched: bad pointers");
  TNX_CHECK_ARG(chil > 0 && d > 0 && chir > 0 && prev_rows > 0 && batch > 0,
                "gauge_left_batched: bad dims");
  const int m = d * chir, n = chil;
  TNX_CHECK_ARG(m >= n, "gauge_left_batched: d*chir=%d < chil=%d", m, n);
  TNX_CHECK_ARG(ws && ws_bytes >= gauge_b_ws_bytes(chil, d, chir, batch),
                "gauge_left_batched: workspace too small");
  double* R = (double*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  TNX_TRY(qr_small_b(A, m, n, 1, R, batch, s));
  // prev <- prev R^T
  GemmDesc g{};
  g.M = prev_rows; g.N = n; g.K = n; g.batch = batch;
  g.A = prev; g.ta = DT::F64; g.lda = n; g.sA = (int64_t)prev_rows * n; g.a_kcontig = true;
  g.B = R; g.tb = DT::F64; g.ldb = n; g.sB = (int64_t)n * n; g.b_ncontig = false;
  g.C = prev_out; g.tc = DT::F64; g.ldc = n; g.sC = (int64_t)prev_rows * n;
  g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
  return gemm(g, s);
}

}  // namespace tnx

extern "C" {

size_t tnx_qr_workspace_bytes(int m, int n) { return tnx::qr_ws_bytes(m, n); }

int tnx_qr(double* A, int m, int n, double* R, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::qr(A, m, n, R, ws, ws_bytes, (hipStream_t)stream);
}

size_t tnx_gauge_workspace_bytes(int chil, int d, int chir) {
  return tnx::gauge_ws_bytes(chil, d, chir);
}

int tnx_gauge_right(double* A, const double* next, double* next_out, int chil, int d, int chir,
                    int next_cols, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::gauge_right(A, next, next_out, chil, d, chir, next_cols, ws, ws_bytes,
                          (hipStream_t)stream);
}

int tnx_gauge_left(double* A, const double* prev, double* prev_out, int chil, int d, int chir,
                   int prev_rows, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::gauge_left(A, prev, prev_out, chil, d, chir, prev_rows, ws, ws_bytes,
                         (hipStream_t)stream);
}

size_t tnx_gauge_batched_workspace_bytes(int chil, int d, int chir, int batch) {
  return tnx::gauge_b_ws_bytes(chil, d, chir, batch);
}

int tnx_gauge_right_batched(double* A, const double* next, double* next_out,
                            int chil, int d, int chir, int next_cols, int batch,
                            void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::gauge_right_b(A, next, next_out, chil, d, chir, next_cols, batch, ws, ws_bytes,
                            (hipStream_t)stream);
}

int tnx_gauge_left_batched(double* A, const double* prev, double* prev_out,
                           int chil, int d, int chir, int prev_rows, int batch,
                           void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::gauge_left_b(A, prev, prev_out, chil, d, chir, prev_rows, batch, ws, ws_bytes,
                           (hipStream_t)stream);
}

}  // extern "C"
