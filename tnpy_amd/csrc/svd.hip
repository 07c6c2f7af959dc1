// Truncated SVD of a dense FP64 matrix: the top-k left singular vectors, the
// singular values, and U_k^T X, for the two-site DMRG truncation to chi and
// the Schmidt spectra. No Gram matrix of X is ever formed from X itself, so
// singular values keep absolute accuracy ~ eps * s_max (a density matrix
// X X^T would square the condition number).
//
// Algorithm (a >= b shown; a < b first reduces X to the square R0^T of
// X^T = Q0 R0):
//   1. X = Q1 R1 (Householder QR, b x b R1),
//   2. R1^T = Q2 R2 (a second QR step: R2 is close to diagonal in the graded
//      sense, which cuts the Jacobi sweeps from 15..35 to 6..8 on decaying
//      spectra),
//   3. one-sided block Jacobi (Hestenes) on the rows of R2: plane rotations
//      from the left until the rows are mutually orthogonal, R2 -> Sigma W^T
//      row by row. Blocks of kJw rows; the block pairs of a round follow the
//      round-robin tournament (block 0 fixed, the others rotating), so the
//      two blocks of a pair are adjacent after a fixed permutation of block
//      positions applied between rounds. Per round: the Gram matrices of
//      all pairs (one batched GEMM), one sweep of two-sided cyclic Jacobi
//      on each 2kJw x 2kJw Gram matrix in LDS (its rotations kept as Z),
//      the rows rotated (P <- Z^T P, one batched GEMM), then the positions
//      permuted. Rotation angles are the small ones (|theta| <= pi/4), so a
//      pair that is already orthogonal gets Z ~ I and rows keep their
//      blocks (sorting them would stall the tournament),
//   4. singular values = row norms, sorted; W_k = the top-k normalised
//      rows, made exactly orthonormal by a Householder QR of W_k^T;
//      U_k = Q1 W_k and U_k^T X = W_k^T R1.
// Convergence: every pair's off-diagonal |G_ij| <= tol * max(sqrt(G_ii
// G_jj), floor), checked once per sweep (one host read).
#include "tnx_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace tnx {

int qr(double* A, int m, int n, double* R, void* ws, size_t ws_bytes, hipStream_t s);
size_t qr_ws_bytes(int m, int n);

constexpr int kJw = 32;           // rows per Jacobi block
constexpr int kJp = 2 * kJw;      // rows of a block pair
constexpr int kJld = kJp + 1;     // LDS row stride
constexpr int kJacMaxSweeps = 30;
constexpr double kJacTol = 1e-14;   // relative orthogonality of a row pair
constexpr double kJacFloor = 1e-28; // times ||X||_F^2: rows below ~1e-14 s_max
                                    // are judged in absolute terms

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// One sweep of two-sided cyclic Jacobi (round-robin, kJp - 1 rounds of kJw
// disjoint rotations) on each pair's Gram matrix G (kJp x kJp, row-major);
// Z accumulates the rotations (G_final = Z^T G Z). A round's kJw rotations
// partition the kJp indices, so G <- J^T G J is one pass over 2 x 2 blocks
// (rows {p_k, q_k}, columns {p_l, q_l} of rotations k and l) with no two
// threads touching one entry, and Z <- Z J one pass over row pairs: two
// barriers per round. Also the pair's off-diagonal measure, max-reduced
// into *off_bits (a non-negative double read as an unsigned integer, so
// integer max = double max).
__global__ __launch_bounds__(256) void jacobi_pair_kernel(const double* __restrict__ Gall,
                                                         double* __restrict__ Zall,
                                                         const double* __restrict__ floor2,
                                                         unsigned long long* __restrict__ off_bits) {
  __shared__ double G[kJp * kJld];
  __shared__ double Z[kJp * kJld];
  __shared__ double cs[kJw], sn[kJw];
  __shared__ int pp[kJw], qq[kJw];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const double* Gp = Gall + (size_t)blockIdx.x * kJp * kJp;
  for (int e = tid; e < kJp * kJp; e += 256) {
    const int i = e >> 6, j = e & (kJp - 1);
    // symmetrise (the GEMM's two halves round independently)
    G[i * kJld + j] = 0.5 * (Gp[e] + Gp[j * kJp + i]);
    Z[i * kJld + j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  // off-diagonal measure of the incoming pair
  {
    const double fl = floor2[0];
    double mx = 0.0;
    for (int e = tid; e < kJp * kJp; e += 256) {
      const int i = e >> 6, j = e & (kJp - 1);
      if (i != j) {
        const double den = fmax(sqrt(fmax(G[i * kJld + i], 0.0) * fmax(G[j * kJld + j], 0.0)), fl);
        mx = fmax(mx, fabs(G[i * kJld + j]) / den);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
      const double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      atomicMax(off_bits, (unsigned long long)__double_as_longlong(m));
    }
  }
  for (int r = 0; r < kJp - 1; ++r) {
    if (tid < kJw) {
      const int i = tid;
      int p, q;
      if (i == 0) {
        p = 0;
        q = 1 + ((kJp - 2 - r) % (kJp - 1) + (kJp - 1)) % (kJp - 1);
      } else {
        p = 1 + ((i - 1 - r) % (kJp - 1) + (kJp - 1)) % (kJp - 1);
        q = 1 + ((kJp - 2 - i - r) % (kJp - 1) + (kJp - 1)) % (kJp - 1);
      }
      const double a = G[p * kJld + p], d = G[q * kJld + q], b = G[p * kJld + q];
      double c = 1.0, s = 0.0;
      if (b != 0.0) {
        // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = (d - a) / (2 b);
        // the refined reciprocals only see arguments >= 1
        const double zeta = (d - a) / (2.0 * b);
        double t;
        if (fabs(zeta) > 1e100) {
          t = 0.5 / zeta;
        } else {
          const double w = fma(zeta, zeta, 1.0);
          t = (zeta >= 0.0 ? 1.0 : -1.0) * rcp_nr(fabs(zeta) + w * rsqrt_nr(w));
        }
        c = rsqrt_nr(fma(t, t, 1.0));
        s = c * t;
      }
      cs[i] = c; sn[i] = s; pp[i] = p; qq[i] = q;
    }
    __syncthreads();
    // G <- J^T G J, one 2 x 2 block per (k, l)
    for (int e = tid; e < kJw * kJw; e += 256) {
      const int k = e >> 5, l = e & (kJw - 1);
      const int pk = pp[k], qk = qq[k], pl = pp[l], ql = qq[l];
      const double ck = cs[k], sk = sn[k], cl = cs[l], sl = sn[l];
      const double g00 = G[pk * kJld + pl], g01 = G[pk * kJld + ql];
      const double g10 = G[qk * kJld + pl], g11 = G[qk * kJld + ql];
      // rows: [r0; r1] = J_k^T [g0; g1]
      const double r00 = ck * g00 - sk * g10, r01 = ck * g01 - sk * g11;
      const double r10 = sk * g00 + ck * g10, r11 = sk * g01 + ck * g11;
      // columns: [c0 c1] = [r0 r1] J_l
      G[pk * kJld + pl] = cl * r00 - sl * r01;
      G[pk * kJld + ql] = sl * r00 + cl * r01;
      G[qk * kJld + pl] = cl * r10 - sl * r11;
      G[qk * kJld + ql] = sl * r10 + cl * r11;
    }
    // Z <- Z J
    for (int e = tid; e < kJp * kJw; e += 256) {
      const int j = e >> 5, l = e & (kJw - 1);
      const int pl = pp[l], ql = qq[l];
      const double cl = cs[l], sl = sn[l];
      const double zp = Z[j * kJld + pl], zq = Z[j * kJld + ql];
      Z[j * kJld + pl] = cl * zp - sl * zq;
      Z[j * kJld + ql] = sl * zp + cl * zq;
    }
    __syncthreads();
  }
  double* Zp = Zall + (size_t)blockIdx.x * kJp * kJp;
  for (int e = tid; e < kJp * kJp; e += 256) {
    const int i = e >> 6, j = e & (kJp - 1);
    Zp[e] = Z[i * kJld + j];
  }
}

__global__ void jacobi_permute_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                      const int* __restrict__ perm, int nblk, int64_t q) {
  const int64_t blk = (int64_t)kJw * q;
  const int64_t total = (int64_t)nblk * blk;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / blk);
    const int64_t o = e - (int64_t)b * blk;
    dst[(int64_t)perm[b] * blk + o] = src[e];
  }
}

// ||X||_F^2 scaled by the floor factor: per-block partial sums of squares,
// then their sum in a fixed order (deterministic).
constexpr int kFloorBlocks = 256;

__global__ __launch_bounds__(256) void jacobi_floor_part_kernel(const double* __restrict__ X,
                                                                int64_t n,
                                                                double* __restrict__ part) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    acc = fma(X[e], X[e], acc);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void jacobi_floor_kernel(const double* __restrict__ part, int np,
                                                           double factor, double* __restrict__ out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int e = threadIdx.x; e < np; e += 256) acc += part[e];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = factor * ((red[0] + red[1]) + (red[2] + red[3]));
}

__global__ void zero_u64_kernel(unsigned long long* p) { p[0] = 0ull; }

// norms[i] = |row i| of the p x q matrix X
__global__ void row_norms_kernel(const double* __restrict__ X, int p, int64_t q,
                                 double* __restrict__ norms) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= p) return;
  double acc = 0.0;
  for (int64_t j = lane; j < q; j += 64) acc += X[row * q + j] * X[row * q + j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) norms[row] = sqrt(acc);
}

// Wt[j][c] = X[order[c]][j] / norm[order[c]]   (q x k, row-major): the
// transposed, normalised top-k rows; a zero row gives a zero column (the QR
// that follows completes the basis).
__global__ void gather_rows_t_kernel(const double* __restrict__ X, int64_t q,
                                     const int* __restrict__ order, const double* __restrict__ norms,
                                     int k, double* __restrict__ Wt) {
  const int64_t total = (int64_t)q * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / k;
    const int c = (int)(e - j * k);
    const int row = order[c];
    const double nrm = norms[row];
    Wt[e] = nrm > 0.0 ? X[(int64_t)row * q + j] / nrm : 0.0;
  }
}

// Jacobi working size: p rows padded to a whole, even number of blocks.
static int jac_blocks(int p) {
  int B = (p + kJw - 1) / kJw;
  if (B < 2) B = 2;
  return B + (B & 1);
}

static size_t jacobi_ws_bytes(int p, int q) {
  const int B = jac_blocks(p);
  const int npair = B / 2;
  size_t b = 2 * al((size_t)B * kJw * q * 8);       // X, Y
  b += 2 * al((size_t)npair * kJp * kJp * 8);       // G, Z
  b += al((size_t)8 * npair * kJp * kJp * 8);       // split-K partials of the Gram GEMM
  b += al((size_t)B * 4) + al(64);                  // perm, off bits + floor
  return b + 256;
}

// Rows of the p x q matrix X0 (row-major) orthogonalised in place of a
// padded copy; returns the padded row count and leaves the result in *Xout
// (inside ws). sweeps_out receives the sweep count.
static int jacobi_rows(const double* X0, int p, int q, void* ws, double** Xout, int* sweeps_out,
                       hipStream_t s) {
  const int B = jac_blocks(p);
  const int npair = B / 2;
  const int64_t rows = (int64_t)B * kJw;
  char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* X = (double*)w; w += al((size_t)rows * q * 8);
  double* Y = (double*)w; w += al((size_t)rows * q * 8);
  double* G = (double*)w; w += al((size_t)npair * kJp * kJp * 8);
  double* Z = (double*)w; w += al((size_t)npair * kJp * kJp * 8);
  double* part = (double*)w; w += al((size_t)8 * npair * kJp * kJp * 8);
  int* perm = (int*)w; w += al((size_t)B * 4);
  unsigned long long* off_bits = (unsigned long long*)w;
  double* floor2 = (double*)(w + 8);
  // round-robin position map (block 0 fixed, the others rotate)
  {
    const int h = B / 2;
    std::vector<int> pm(B, 0);
    pm[0] = 0;
    pm[1] = h > 1 ? 2 : 1;
    for (int i = 1; i < h - 1; ++i) pm[2 * i] = 2 * i + 2;
    if (h > 1) pm[2 * (h - 1)] = 2 * h - 1;
    for (int i = 1; i < h; ++i) pm[2 * i + 1] = 2 * i - 1;
    TNX_HIP(hipMemcpyAsync(perm, pm.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  }
  TNX_TRY(fill_f64(X, rows * q, 0.0, s));
  TNX_HIP(hipMemcpyAsync(X, X0, sizeof(double) * (size_t)p * q, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(jacobi_floor_part_kernel, dim3(kFloorBlocks), dim3(256), 0, s, X,
                     (int64_t)p * q, part);
  TNX_LAUNCH_CHECK();
  hipLaunchKernelGGL(jacobi_floor_kernel, dim3(1), dim3(256), 0, s, part, kFloorBlocks, kJacFloor,
                     floor2);
  TNX_LAUNCH_CHECK();
  const int64_t pq = (int64_t)kJp * q;
  int sweep = 0;
  for (; sweep < kJacMaxSweeps; ++sweep) {
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, s, off_bits);
    TNX_LAUNCH_CHECK();
    for (int r = 0; r < B - 1; ++r) {
      // G_i = P_i P_i^T
      GemmDesc g{};
      g.M = kJp; g.N = kJp; g.K = q; g.batch = npair;
      g.A = X; g.ta = DT::F64; g.lda = q; g.sA = pq; g.a_kcontig = true;
      g.B = X; g.tb = DT::F64; g.ldb = q; g.sB = pq; g.b_ncontig = false;
      g.C = G; g.tc = DT::F64; g.ldc = kJp; g.sC = (int64_t)kJp * kJp;
      g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
      g.splitk_ws = part; g.splitk_ws_bytes = (size_t)8 * npair * kJp * kJp * 8;
      TNX_TRY(gemm(g, s));
      hipLaunchKernelGGL(jacobi_pair_kernel, dim3(npair), dim3(256), 0, s, G, Z, floor2, off_bits);
      TNX_LAUNCH_CHECK();
      // Y_i = Z_i^T P_i
      GemmDesc g2{};
      g2.M = kJp; g2.N = q; g2.K = kJp; g2.batch = npair;
      g2.A = Z; g2.ta = DT::F64; g2.lda = kJp; g2.sA = (int64_t)kJp * kJp; g2.a_kcontig = false;
      g2.B = X; g2.tb = DT::F64; g2.ldb = q; g2.sB = pq; g2.b_ncontig = true;
      g2.C = Y; g2.tc = DT::F64; g2.ldc = q; g2.sC = pq;
      g2.alpha = 1.0; g2.beta = 0.0; g2.prec = Prec::F64;
      TNX_TRY(gemm(g2, s));
      const int64_t total = rows * q;
      const int nb = (int)std::min<int64_t>((total + 255) / 256, 8192);
      hipLaunchKernelGGL(jacobi_permute_kernel, dim3(nb), dim3(256), 0, s, Y, X, perm, B, (int64_t)q);
      TNX_LAUNCH_CHECK();
    }
    unsigned long long bits = 0;
    TNX_HIP(hipMemcpyAsync(&bits, off_bits, sizeof(bits), hipMemcpyDeviceToHost, s));
    TNX_HIP(hipStreamSynchronize(s));
    double off;
    static_assert(sizeof(off) == sizeof(bits), "double size");
    memcpy(&off, &bits, sizeof(off));
    if (!(off > kJacTol)) {  // also stops on NaN-free zero input
      ++sweep;
      break;
    }
  }
  *Xout = X;
  *sweeps_out = sweep;
  return TNX_OK;
}

size_t svd_trunc_ws_bytes(int a, int b, int k) {
  // a < b: X^T = Q0 R0 first (Q0: b x a), then the square a x a problem
  const int m = std::max(a, b), n = std::min(a, b);
  size_t s = 0;
  s += al((size_t)m * n * 8);                      // Q of the first QR (X or X^T copy)
  s += al((size_t)n * n * 8);                      // R1
  s += al((size_t)n * n * 8);                      // R1^T, then its Q
  s += al((size_t)n * n * 8);                      // R2
  s += al((size_t)n * n * 8);                      // square core input (a < b) and Q1 of it
  s += al((size_t)n * n * 8);                      // R of the core's first QR
  s += al((size_t)n * k * 8);                      // W_k^T
  s += al((size_t)k * k * 8);                      // its R
  s += al((size_t)jac_blocks(n) * kJw * 8) + al((size_t)n * 4);  // norms, order
  s += al((size_t)n * k * 8);                      // U_k of the core (a < b)
  s += al((size_t)k * n * 8);                      // U_k^T core (a < b)
  size_t q = qr_ws_bytes(m, n);
  q = std::max(q, qr_ws_bytes(n, n));
  q = std::max(q, qr_ws_bytes(n, k));
  s += al(q);
  s += al(jacobi_ws_bytes(n, n));
  return s + 1024;
}

// Core on a square n x n matrix S (row-major): top-k left singular vectors
// U (n x k) and singular values sv (k, descending, host), and Y = U^T S
// (k x n). S is overwritten.
struct SvdScratch {
  double *Q1, *R1, *T1, *R2, *Wt, *Rw, *norms;
  int* order;
  void* qws;
  size_t qws_bytes;
  void* jws;
};

static int svd_core(double* S, int a, int n, int k, double* U, double* Y, double* sv_host,
                    int* sweeps, SvdScratch& sc, hipStream_t s) {
  // S (a x n, a >= n) = Q1 R1 ; Q1 -> S in place
  TNX_TRY(qr(S, a, n, sc.R1, sc.qws, sc.qws_bytes, s));
  // R1^T = Q2 R2
  TNX_TRY(transpose_f64(sc.R1, sc.T1, n, n, s));
  TNX_TRY(qr(sc.T1, n, n, sc.R2, sc.qws, sc.qws_bytes, s));
  double* Xj = nullptr;
  TNX_TRY(jacobi_rows(sc.R2, n, n, sc.jws, &Xj, sweeps, s));
  // rotations and the block permutation move rows through the padded
  // positions too: every padded row is a candidate
  const int prow = jac_blocks(n) * kJw;
  const int nb4 = (prow + 3) / 4;
  hipLaunchKernelGGL(row_norms_kernel, dim3(nb4), dim3(256), 0, s, Xj, prow, (int64_t)n, sc.norms);
  TNX_LAUNCH_CHECK();
  std::vector<double> nh(prow);
  TNX_HIP(hipMemcpyAsync(nh.data(), sc.norms, sizeof(double) * prow, hipMemcpyDeviceToHost, s));
  TNX_HIP(hipStreamSynchronize(s));
  std::vector<int> ord(prow);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nh[x] > nh[y]; });
  for (int c = 0; c < k; ++c) sv_host[c] = nh[ord[c]];
  TNX_HIP(hipMemcpyAsync(sc.order, ord.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
  // W_k^T (n x k), orthonormalised by QR
  {
    const int64_t total = (int64_t)n * k;
    const int nb = (int)std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(gather_rows_t_kernel, dim3(nb), dim3(256), 0, s, Xj, (int64_t)n, sc.order,
                       sc.norms, k, sc.Wt);
    TNX_LAUNCH_CHECK();
  }
  TNX_TRY(qr(sc.Wt, n, k, sc.Rw, sc.qws, sc.qws_bytes, s));
  // U = Q1 W_k  (a x k)
  GemmDesc g{};
  g.M = a; g.N = k; g.K = n; g.batch = 1;
  g.A = S; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = true;
  g.B = sc.Wt; g.tb = DT::F64; g.ldb = k; g.sB = 0; g.b_ncontig = true;
  g.C = U; g.tc = DT::F64; g.ldc = k; g.sC = 0;
  g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
  TNX_TRY(gemm(g, s));
  // Y = W_k^T R1  (k x n)
  GemmDesc g2{};
  g2.M = k; g2.N = n; g2.K = n; g2.batch = 1;
  g2.A = sc.Wt; g2.ta = DT::F64; g2.lda = k; g2.sA = 0; g2.a_kcontig = false;
  g2.B = sc.R1; g2.tb = DT::F64; g2.ldb = n; g2.sB = 0; g2.b_ncontig = true;
  g2.C = Y; g2.tc = DT::F64; g2.ldc = n; g2.sC = 0;
  g2.alpha = 1.0; g2.beta = 0.0; g2.prec = Prec::F64;
  return gemm(g2, s);
}

int svd_trunc(const double* X, int a, int b, int k, double* U, double* Y, double* sv_host,
              int* sweeps, void* ws, size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(X && U && Y && sv_host, "svd_trunc: null pointer");
  TNX_CHECK_ARG(a >= 1 && b >= 1 && k >= 1 && k <= std::min(a, b),
                "svd_trunc: need 1 <= k <= min(a, b) (a=%d b=%d k=%d)", a, b, k);
  TNX_CHECK_ARG(ws && ws_bytes >= svd_trunc_ws_bytes(a, b, k), "svd_trunc: workspace too small");
  const int m = std::max(a, b), n = std::min(a, b);
  char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* Qa = (double*)w; w += al((size_t)m * n * 8);
  SvdScratch sc;
  sc.R1 = (double*)w; w += al((size_t)n * n * 8);
  sc.T1 = (double*)w; w += al((size_t)n * n * 8);
  sc.R2 = (double*)w; w += al((size_t)n * n * 8);
  double* Sq = (double*)w; w += al((size_t)n * n * 8);
  double* R0 = (double*)w; w += al((size_t)n * n * 8);
  sc.Wt = (double*)w; w += al((size_t)n * k * 8);
  sc.Rw = (double*)w; w += al((size_t)k * k * 8);
  sc.norms = (double*)w; w += al((size_t)jac_blocks(n) * kJw * 8);
  sc.order = (int*)w; w += al((size_t)n * 4);
  double* Uc = (double*)w; w += al((size_t)n * k * 8);
  double* Yc = (double*)w; w += al((size_t)k * n * 8);
  size_t q = qr_ws_bytes(m, n);
  q = std::max(q, qr_ws_bytes(n, n));
  q = std::max(q, qr_ws_bytes(n, k));
  sc.qws = w; sc.qws_bytes = al(q); w += al(q);
  sc.jws = w;
  if (a >= b) {
    TNX_HIP(hipMemcpyAsync(Qa, X, sizeof(double) * (size_t)a * b, hipMemcpyDeviceToDevice, s));
    return svd_core(Qa, a, b, k, U, Y, sv_host, sweeps, sc, s);
  }
  // a < b: X^T (b x a) = Q0 R0, X = R0^T Q0^T; the left vectors of X are
  // those of the square R0^T, and U^T X = (U^T R0^T) Q0^T
  TNX_TRY(transpose_f64(X, Qa, a, b, s));
  TNX_TRY(qr(Qa, b, a, R0, sc.qws, sc.qws_bytes, s));
  TNX_TRY(transpose_f64(R0, Sq, a, a, s));
  TNX_TRY(svd_core(Sq, a, a, k, U, Yc, sv_host, sweeps, sc, s));
  GemmDesc g{};
  g.M = k; g.N = b; g.K = a; g.batch = 1;
  g.A = Yc; g.ta = DT::F64; g.lda = a; g.sA = 0; g.a_kcontig = true;
  g.B = Qa; g.tb = DT::F64; g.ldb = a; g.sB = 0; g.b_ncontig = false;  // B(k,n) = Q0[n][k]
  g.C = Y; g.tc = DT::F64; g.ldc = b; g.sC = 0;
  g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
  (void)Uc;
  return gemm(g, s);
}

// ---- dense symmetric eigensolver (all pairs) --------------------------------
// H + c I with c = the largest absolute row sum of H (Gershgorin) is
// positive definite, so its SVD is its eigendecomposition: singular values
// lambda + c, left singular vectors = eigenvectors. Eigenvalues come back
// ascending, eigenvectors as the columns of evecs (n x n, row-major).

// out[0] = max_i sum_j |A_ij| + 1 (one workgroup)
__global__ __launch_bounds__(256) void gershgorin_kernel(const double* __restrict__ A, int n,
                                                         double* __restrict__ out) {
  __shared__ double red[4];
  double mx = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += fabs(A[(int64_t)i * n + j]);
    mx = fmax(mx, acc);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3])) + 1.0;
}

// Hs = (A + A^T)/2 + c I
__global__ void shift_sym_kernel(const double* __restrict__ A, int n, const double* __restrict__ c,
                                 double* __restrict__ Hs) {
  const int64_t nn = (int64_t)n * n;
  const double cv = c[0];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e - i * n;
    Hs[e] = 0.5 * (A[e] + A[j * n + i]) + (i == j ? cv : 0.0);
  }
}

// E[:, j] = U[:, n-1-j]
__global__ void reverse_cols_kernel(const double* __restrict__ U, int n, double* __restrict__ E) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e - i * n;
    E[e] = U[i * n + (n - 1 - j)];
  }
}

size_t syev_all_ws_bytes(int n) { return 3 * al((size_t)n * n * 8) + al(64) + svd_trunc_ws_bytes(n, n, n); }

int syev_all(const double* A, int n, double* evals_host, double* evecs, int* sweeps, void* ws,
             size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(A && evals_host && evecs, "syev: null pointer");
  TNX_CHECK_ARG(n >= 1, "syev: n=%d", n);
  TNX_CHECK_ARG(ws && ws_bytes >= syev_all_ws_bytes(n), "syev: workspace too small");
  char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* Hs = (double*)w; w += al((size_t)n * n * 8);
  double* U = (double*)w; w += al((size_t)n * n * 8);
  double* Y = (double*)w; w += al((size_t)n * n * 8);
  double* c = (double*)w; w += al(64);
  hipLaunchKernelGGL(gershgorin_kernel, dim3(1), dim3(256), 0, s, A, n, c);
  TNX_LAUNCH_CHECK();
  const int64_t nn = (int64_t)n * n;
  const int nb = (int)std::min<int64_t>((nn + 255) / 256, 8192);
  hipLaunchKernelGGL(shift_sym_kernel, dim3(nb), dim3(256), 0, s, A, n, c, Hs);
  TNX_LAUNCH_CHECK();
  double ch = 0.0;
  TNX_HIP(hipMemcpyAsync(&ch, c, sizeof(double), hipMemcpyDeviceToHost, s));
  std::vector<double> sv(n);
  const size_t rest = ws_bytes - (size_t)(w - (char*)ws);
  TNX_TRY(svd_trunc(Hs, n, n, n, U, Y, sv.data(), sweeps, w, rest, s));
  for (int j = 0; j < n; ++j) evals_host[j] = sv[n - 1 - j] - ch;
  hipLaunchKernelGGL(reverse_cols_kernel, dim3(nb), dim3(256), 0, s, U, n, evecs);
  TNX_LAUNCH_CHECK();
  TNX_HIP(hipStreamSynchronize(s));
  return TNX_OK;
}

// ---- dense generalized symmetric-definite problem (lowest pair) -----------
// A x = lambda B x, B positive definite: B = W diag(l) W^T, S = W diag(l)^-1/2
// (so S^T B S = I), C = S^T A S, lowest pair (mu, y) of C, x = S y with
// x^T B x = 1.

// S[i][j] = W[i][j] / sqrt(l[j]), l ascending (host values uploaded)
__global__ void scale_cols_kernel(const double* __restrict__ W, int n, const double* __restrict__ l,
                                  double* __restrict__ S) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % n;
    S[e] = W[e] / sqrt(l[j]);
  }
}

size_t sygv_lowest_ws_bytes(int n) {
  return 5 * al((size_t)n * n * 8) + al((size_t)n * 8) + syev_all_ws_bytes(n);
}

int sygv_lowest(const double* A, const double* B, int n, double* x, double* eval_host, void* ws,
                size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(A && B && x && eval_host, "sygv: null pointer");
  TNX_CHECK_ARG(n >= 1, "sygv: n=%d", n);
  TNX_CHECK_ARG(ws && ws_bytes >= sygv_lowest_ws_bytes(n), "sygv: workspace too small");
  char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  double* Wb = (double*)w; w += al((size_t)n * n * 8);
  double* S = (double*)w; w += al((size_t)n * n * 8);
  double* T = (double*)w; w += al((size_t)n * n * 8);
  double* Cm = (double*)w; w += al((size_t)n * n * 8);
  double* Y = (double*)w; w += al((size_t)n * n * 8);
  double* ld = (double*)w; w += al((size_t)n * 8);
  const size_t rest = ws_bytes - (size_t)(w - (char*)ws);
  std::vector<double> lb(n), mu(n);
  int sw = 0;
  TNX_TRY(syev_all(B, n, lb.data(), Wb, &sw, w, rest, s));
  TNX_CHECK_ARG(lb[0] > 0.0, "sygv: B is not positive definite (lowest eigenvalue %g)", lb[0]);
  TNX_HIP(hipMemcpyAsync(ld, lb.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  const int64_t nn = (int64_t)n * n;
  const int nb = (int)std::min<int64_t>((nn + 255) / 256, 8192);
  hipLaunchKernelGGL(scale_cols_kernel, dim3(nb), dim3(256), 0, s, Wb, n, ld, S);
  TNX_LAUNCH_CHECK();
  // T = A S ; C = S^T T
  GemmDesc g{};
  g.M = n; g.N = n; g.K = n; g.batch = 1;
  g.A = A; g.ta = DT::F64; g.lda = n; g.sA = 0; g.a_kcontig = true;
  g.B = S; g.tb = DT::F64; g.ldb = n; g.sB = 0; g.b_ncontig = true;
  g.C = T; g.tc = DT::F64; g.ldc = n; g.sC = 0;
  g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
  TNX_TRY(gemm(g, s));
  GemmDesc g2 = g;
  g2.A = S; g2.a_kcontig = false;
  g2.B = T; g2.C = Cm;
  TNX_TRY(gemm(g2, s));
  TNX_TRY(syev_all(Cm, n, mu.data(), Y, &sw, w, rest, s));
  // x = S y0 (y0 = column 0 of Y)
  GemmDesc g3{};
  g3.M = n; g3.N = 1; g3.K = n; g3.batch = 1;
  g3.A = S; g3.ta = DT::F64; g3.lda = n; g3.sA = 0; g3.a_kcontig = true;
  g3.B = Y; g3.tb = DT::F64; g3.ldb = n; g3.sB = 0; g3.b_ncontig = true;
  g3.C = x; g3.tc = DT::F64; g3.ldc = 1; g3.sC = 0;
  g3.alpha = 1.0; g3.beta = 0.0; g3.prec = Prec::F64;
  TNX_TRY(gemm(g3, s));
  TNX_HIP(hipStreamSynchronize(s));
  *eval_host = mu[0];
  return TNX_OK;
}

// One site of the MPS of H|phi> (H an MPO): out[(l, wl)][k][(r, wr)] =
// sum_b W[wl][wr][k][b] A[l][b][r] (the MPO's bra index contracted with the
// state), bonds fused state-major. ShiftInvertDMRG._restore_mps
// (tnpy/finite_dmrg.py:316-345).
__global__ void mpo_apply_site_kernel(const double* __restrict__ W, const double* __restrict__ A,
                                      int Dl, int Dr, int d, int l, int r,
                                      double* __restrict__ out) {
  const int64_t total = (int64_t)l * Dl * d * r * Dr;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t t = e;
    const int wr = (int)(t % Dr); t /= Dr;
    const int rr = (int)(t % r); t /= r;
    const int k = (int)(t % d); t /= d;
    const int wl = (int)(t % Dl); t /= Dl;
    const int ll = (int)t;
    double acc = 0.0;
    for (int b = 0; b < d; ++b)
      acc += W[(((int64_t)wl * Dr + wr) * d + k) * d + b] * A[((int64_t)ll * d + b) * r + rr];
    out[e] = acc;
  }
}

// One site of an MPO absorbed into the dense operator built so far:
// Tout[v][K*d + k][B*d + b] = sum_w T[w][K][B] W[w][v][k][b]
// (T: Dt x dimK x dimK, W: Dt x Dr x d x d). Builds FullHamiltonian
// (tnpy/operators.py:119-154) site by site.
__global__ void mpo_absorb_kernel(const double* __restrict__ T, int Dt, int64_t dimK,
                                  const double* __restrict__ W, int Dr, int d,
                                  double* __restrict__ Tout) {
  const int64_t dimO = dimK * d;
  const int64_t total = (int64_t)Dr * dimO * dimO;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = e / (dimO * dimO);
    const int64_t rem = e - v * dimO * dimO;
    const int64_t ro = rem / dimO, co = rem - ro * dimO;
    const int64_t K = ro / d, B = co / d;
    const int k = (int)(ro - K * d), b = (int)(co - B * d);
    double acc = 0.0;
    for (int w = 0; w < Dt; ++w)
      acc += T[((int64_t)w * dimK + K) * dimK + B] * W[(((int64_t)w * Dr + v) * d + k) * d + b];
    Tout[e] = acc;
  }
}

}  // namespace tnx

extern "C" {

size_t tnx_svd_trunc_workspace_bytes(int a, int b, int k) { return tnx::svd_trunc_ws_bytes(a, b, k); }

int tnx_svd_trunc_f64(const double* X, int a, int b, int k, double* U, double* Y, double* sv_host,
                      int* sweeps, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  int sw = 0;
  const int st = tnx::svd_trunc(X, a, b, k, U, Y, sv_host, &sw, ws, ws_bytes, (hipStream_t)stream);
  if (sweeps) *sweeps = sw;
  return st;
}

size_t tnx_syev_workspace_bytes(int n) { return tnx::syev_all_ws_bytes(n); }

int tnx_syev_f64(const double* A, int n, double* evals_host, double* evecs, void* ws,
                 size_t ws_bytes, tnx_stream_t stream) {
  int sw = 0;
  return tnx::syev_all(A, n, evals_host, evecs, &sw, ws, ws_bytes, (hipStream_t)stream);
}

size_t tnx_sygv_lowest_workspace_bytes(int n) { return tnx::sygv_lowest_ws_bytes(n); }

int tnx_sygv_lowest(const double* A, const double* B, int n, double* x, double* eval_host, void* ws,
                    size_t ws_bytes, tnx_stream_t stream) {
  return tnx::sygv_lowest(A, B, n, x, eval_host, ws, ws_bytes, (hipStream_t)stream);
}

int tnx_mpo_apply_site(const double* W, const double* A, int Dl, int Dr, int d, int l, int r,
                       double* out, tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(W && A && out, "tnx_mpo_apply_site: null pointer");
  TNX_CHECK_ARG(Dl >= 1 && Dr >= 1 && d >= 1 && l >= 1 && r >= 1, "tnx_mpo_apply_site: bad dims");
  const int64_t total = (int64_t)l * Dl * d * r * Dr;
  const int nb = (int)std::min<int64_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(mpo_apply_site_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, W, A, Dl, Dr,
                     d, l, r, out);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

int tnx_mpo_absorb(const double* T, int Dt, int64_t dimK, const double* W, int Dr, int d,
                   double* Tout, tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(T && W && Tout, "tnx_mpo_absorb: null pointer");
  TNX_CHECK_ARG(Dt >= 1 && dimK >= 1 && Dr >= 1 && d >= 1, "tnx_mpo_absorb: bad dims");
  const int64_t total = (int64_t)Dr * dimK * d * dimK * d;
  const int nb = (int)std::min<int64_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(mpo_absorb_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, T, Dt, dimK, W,
                     Dr, d, Tout);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

}  // extern "C"
