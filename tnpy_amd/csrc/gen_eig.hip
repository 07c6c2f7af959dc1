// Lowest eigenpair of the generalized one-site problem of shift-invert DMRG,
//   H_eff x = lambda M_eff x,
// H_eff the one-site operator of an MPO (H - eps), M_eff that of its square
// (positive definite on the canonical-form local space). The reference asks
// primme for it (eigshmv with M=, tnpy/finite_dmrg.py:355-360 via
// tnpy/linalg.py:64-87). Here: a generalized Davidson iteration without
// preconditioner on a basis V kept M-orthonormal (V^T M V = I), with H V
// and M V stored, so the projected problem G = V^T H V is a standard one:
//   x = V s (lowest pair of G), r = H x - theta M x; stop when
//   |r| <= tol (|H x| + |theta| |M x|); else t = r, M-orthogonalised against
//   V twice (t -= V (MV)^T t), M-normalised, H t and M t appended.
// A full basis restarts with its `keep` lowest Ritz vectors (V S_k, HV S_k,
// MV S_k; G becomes diag(theta)). The returned x has x^T M x = 1. All vector
// arithmetic is GEMM/GEMV on the device; the projected J x J eigenproblem is
// solved by one workgroup (cyclic Jacobi in LDS).
#include "tnx_common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tnx {

int heff1_matvec(const double* L, const double* W, const double* R, const double* x, double* y,
                 int chil, int Dl, int Dr, int d, int chir, int prec, void* ws, size_t ws_bytes,
                 hipStream_t s);
size_t heff1_ws_bytes(int chil, int Dl, int Dr, int d, int chir);

constexpr int kGenMaxBasis = 64;
constexpr int kGenLd = kGenMaxBasis + 1;

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// Every eigenpair of the symmetric n x n (n <= 64) matrix G (row stride
// ldg): cyclic Jacobi sweeps (round-robin over 64 slots, rotations touching
// a slot >= n skipped) until the off-diagonal part is below 1e-30 of the
// squared norm; eigenvalues ascending to ev, eigenvectors as the columns
// of Z (row stride kGenMaxBasis).
__global__ __launch_bounds__(256) void small_syev_kernel(const double* __restrict__ G, int n, int ldg,
                                                        double* __restrict__ ev,
                                                        double* __restrict__ Zout) {
  __shared__ double A[kGenMaxBasis * kGenLd];
  __shared__ double V[kGenMaxBasis * kGenLd];
  __shared__ double cs[32], sn[32];
  __shared__ int pp[32], qq[32];
  __shared__ double red[8];
  __shared__ int stop;
  constexpr int N64 = kGenMaxBasis;
  const int tid = threadIdx.x;
  for (int e = tid; e < N64 * N64; e += 256) {
    const int i = e / N64, j = e - i * N64;
    A[i * kGenLd + j] = (i < n && j < n) ? 0.5 * (G[i * ldg + j] + G[j * ldg + i]) : 0.0;
    V[i * kGenLd + j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 30; ++sweep) {
    // off-diagonal and total squared norms
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < n * n; e += 256) {
      const int i = e / n, j = e - i * n;
      const double a = A[i * kGenLd + j];
      tot += a * a;
      if (i != j) off += a * a;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o, 64);
      tot += __shfl_xor(tot, o, 64);
    }
    if ((tid & 63) == 0) { red[tid >> 6] = off; red[4 + (tid >> 6)] = tot; }
    __syncthreads();
    if (tid == 0) {
      const double o2 = (red[0] + red[1]) + (red[2] + red[3]);
      const double t2 = (red[4] + red[5]) + (red[6] + red[7]);
      stop = (o2 <= 1e-30 * t2) ? 1 : 0;
    }
    __syncthreads();
    if (stop) break;
    for (int r = 0; r < N64 - 1; ++r) {
      if (tid < 32) {
        const int i = tid;
        int p, q;
        if (i == 0) {
          p = 0;
          q = 1 + ((N64 - 2 - r) % (N64 - 1) + (N64 - 1)) % (N64 - 1);
        } else {
          p = 1 + ((i - 1 - r) % (N64 - 1) + (N64 - 1)) % (N64 - 1);
          q = 1 + ((N64 - 2 - i - r) % (N64 - 1) + (N64 - 1)) % (N64 - 1);
        }
        double c = 1.0, s = 0.0;
        if (p < n && q < n) {
          const double a = A[p * kGenLd + p], d = A[q * kGenLd + q], b = A[p * kGenLd + q];
          if (b != 0.0) {
            const double zeta = (d - a) / (2.0 * b);
            double t;
            if (fabs(zeta) > 1e100) t = 0.5 / zeta;
            else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            s = c * t;
          }
        }
        cs[i] = c; sn[i] = s; pp[i] = p; qq[i] = q;
      }
      __syncthreads();
      for (int e = tid; e < 32 * N64; e += 256) {
        const int k = e / N64, j = e - k * N64;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], s = sn[k];
        const double gp = A[p * kGenLd + j], gq = A[q * kGenLd + j];
        A[p * kGenLd + j] = c * gp - s * gq;
        A[q * kGenLd + j] = s * gp + c * gq;
      }
      __syncthreads();
      for (int e = tid; e < 32 * N64; e += 256) {
        const int k = e / N64, j = e - k * N64;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], s = sn[k];
        const double gp = A[j * kGenLd + p], gq = A[j * kGenLd + q];
        A[j * kGenLd + p] = c * gp - s * gq;
        A[j * kGenLd + q] = s * gp + c * gq;
        const double vp = V[j * kGenLd + p], vq = V[j * kGenLd + q];
        V[j * kGenLd + p] = c * vp - s * vq;
        V[j * kGenLd + q] = s * vp + c * vq;
      }
      __syncthreads();
    }
  }
  // ascending order: slot i goes to its rank among the first n diagonals
  if (tid < n) {
    const double ai = A[tid * kGenLd + tid];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double aj = A[j * kGenLd + j];
      rank += (aj < ai || (aj == ai && j < tid)) ? 1 : 0;
    }
    ev[rank] = ai;
    for (int i = 0; i < n; ++i) Zout[i * kGenMaxBasis + rank] = V[i * kGenLd + tid];
  }
}

// G[i][J] = G[J][i] = h[i] for i <= J   (G row stride kGenMaxBasis)
__global__ void gen_gcol_kernel(const double* __restrict__ h, int J, double* __restrict__ G) {
  const int i = threadIdx.x;
  if (i <= J) {
    G[i * kGenMaxBasis + J] = h[i];
    G[J * kGenMaxBasis + i] = h[i];
  }
}

// G = diag(theta[0..k-1])
__global__ void gen_gdiag_kernel(const double* __restrict__ theta, int k, double* __restrict__ G) {
  for (int e = threadIdx.x; e < kGenMaxBasis * kGenMaxBasis; e += blockDim.x) {
    const int i = e / kGenMaxBasis, j = e - i * kGenMaxBasis;
    G[e] = (i == j && i < k) ? theta[i] : 0.0;
  }
}

// r = hx - theta[0] mx
__global__ void gen_resid_kernel(const double* __restrict__ hx, const double* __restrict__ mx,
                                 const double* __restrict__ theta, double* __restrict__ r, int64_t N) {
  const double th = theta[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    r[i] = hx[i] - th * mx[i];
}

// y = x * (1 / sqrt(max(nrm2[0], tiny)))   for two vectors at once
__global__ void gen_scale2_kernel(double* __restrict__ a, double* __restrict__ b,
                                  const double* __restrict__ nrm2, int64_t N) {
  const double f = 1.0 / sqrt(fmax(nrm2[0], 1e-300));
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    a[i] *= f;
    b[i] *= f;
  }
}

static int vblocks(int64_t N) { return (int)std::min<int64_t>((N + 255) / 256, 4096); }

// C (M x 1) = alpha * A b + beta * C  with A given as rows (a_rows: A = V, M
// = count, K = N, V row-major count x N) or as columns (A = V^T).
static int gemv(const double* V, int64_t N, int count, bool transpose, const double* b, double* c,
                double alpha, double beta, double* part, size_t part_bytes, hipStream_t s) {
  GemmDesc g{};
  g.batch = 1;
  if (!transpose) {  // c (count) = V b  (b length N)
    g.M = count; g.N = 1; g.K = (int)N;
    g.A = V; g.ta = DT::F64; g.lda = N; g.sA = 0; g.a_kcontig = true;
  } else {           // c (N) = V^T b  (b length count)
    g.M = (int)N; g.N = 1; g.K = count;
    g.A = V; g.ta = DT::F64; g.lda = N; g.sA = 0; g.a_kcontig = false;
  }
  g.B = b; g.tb = DT::F64; g.ldb = 1; g.sB = 0; g.b_ncontig = true;
  g.C = c; g.tc = DT::F64; g.ldc = 1; g.sC = 0;
  g.alpha = alpha; g.beta = beta; g.prec = Prec::F64;
  g.splitk_ws = part; g.splitk_ws_bytes = part_bytes;
  return gemm(g, s);
}

struct GenOp {
  const double *L, *W, *R;
  int Dl, Dr;
};

size_t eigsh_gen_ws_bytes(int chil, int Dl, int Dr, int D2l, int D2r, int d, int chir,
                          int max_basis) {
  const int m = std::max(3, std::min(max_basis, kGenMaxBasis));
  const size_t N = (size_t)chil * d * chir;
  size_t b = 6 * al((size_t)m * N * 8);            // V, HV, MV and their restart targets
  b += 4 * al(N * 8);                              // x, hx, mx, r
  b += 2 * al((size_t)kGenMaxBasis * kGenMaxBasis * 8);
  b += al((size_t)(3 * kGenMaxBasis + 8) * 8);
  b += al((size_t)32 * kGenMaxBasis * 8 * 8);      // split-K partials of the GEMVs
  b += al(std::max(heff1_ws_bytes(chil, Dl, Dr, d, chir), heff1_ws_bytes(chil, D2l, D2r, d, chir)));
  return b + 1024;
}

int eigsh_gen_heff1(const double* L, const double* W, const double* R, const double* L2,
                    const double* W2, const double* R2, const double* v0, double* vec,
                    double* hvec, int chil, int Dl, int Dr, int D2l, int D2r, int d, int chir,
                    const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                    size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(opts && info, "eigsh_gen: null opts/info");
  TNX_CHECK_ARG(L && W && R && L2 && W2 && R2 && v0 && vec, "eigsh_gen: null pointer");
  TNX_CHECK_ARG(chil > 0 && Dl > 0 && Dr > 0 && D2l > 0 && D2r > 0 && d > 0 && chir > 0,
                "eigsh_gen: bad dims");
  TNX_CHECK_ARG(opts->precision == TNX_F64, "eigsh_gen: only FP64 matvecs");
  TNX_CHECK_ARG(opts->tol >= 0.0, "eigsh_gen: negative tol");
  TNX_CHECK_ARG(ws && ws_bytes >= eigsh_gen_ws_bytes(chil, Dl, Dr, D2l, D2r, d, chir, opts->max_basis),
                "eigsh_gen: workspace too small");
  const int64_t N = (int64_t)chil * d * chir;
  const int m = (int)std::min<int64_t>(std::max(3, std::min(opts->max_basis, kGenMaxBasis)),
                                       std::max<int64_t>(N, 1));
  const int keep = std::max(1, std::min(m - 1, opts->keep > 0 ? opts->keep : m / 4 + 2));
  const int max_mv = opts->max_matvecs > 0 ? opts->max_matvecs : 10000;
  char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  const size_t blk = al((size_t)m * N * 8);
  double* V = (double*)w; w += blk;
  double* HV = (double*)w; w += blk;
  double* MV = (double*)w; w += blk;
  double* V2 = (double*)w; w += blk;
  double* HV2 = (double*)w; w += blk;
  double* MV2 = (double*)w; w += blk;
  double* x = (double*)w; w += al(N * 8);
  double* hx = (double*)w; w += al(N * 8);
  double* mx = (double*)w; w += al(N * 8);
  double* r = (double*)w; w += al(N * 8);
  double* G = (double*)w; w += al((size_t)kGenMaxBasis * kGenMaxBasis * 8);
  double* Z = (double*)w; w += al((size_t)kGenMaxBasis * kGenMaxBasis * 8);
  double* theta = (double*)w;                  // eigenvalues of G
  double* hb = theta + kGenMaxBasis;           // orthogonalisation coefficients
  double* hg = hb + kGenMaxBasis;              // new column of G
  double* s8 = hg + kGenMaxBasis;              // scalars
  w += al((size_t)(3 * kGenMaxBasis + 8) * 8);
  double* part = (double*)w; w += al((size_t)32 * kGenMaxBasis * 8 * 8);
  const size_t part_bytes = (size_t)32 * kGenMaxBasis * 8 * 8;
  void* mvws = w;
  const size_t mvb = std::max(heff1_ws_bytes(chil, Dl, Dr, d, chir), heff1_ws_bytes(chil, D2l, D2r, d, chir));
  const int nb = vblocks(N);
  int matvecs = 0, restarts = 0;
  auto applyH = [&](const double* in, double* out) -> int {
    ++matvecs;
    return heff1_matvec(L, W, R, in, out, chil, Dl, Dr, d, chir, TNX_F64, mvws, mvb, s);
  };
  auto applyM = [&](const double* in, double* out) -> int {
    ++matvecs;
    return heff1_matvec(L2, W2, R2, in, out, chil, D2l, D2r, d, chir, TNX_F64, mvws, mvb, s);
  };
  auto dot = [&](const double* a, const double* b, double* out) -> int {
    return gemv(a, N, 1, false, b, out, 1.0, 0.0, part, part_bytes, s);
  };
  auto to_host = [&](const double* dev, double* h, int n) -> int {
    TNX_HIP(hipMemcpyAsync(h, dev, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    TNX_HIP(hipStreamSynchronize(s));
    return TNX_OK;
  };
  // Ritz vector combinations: out = Src^T z0 (z0 = column 0 of Z, stride
  // kGenMaxBasis); restart: Dst = Z[:, :keep]^T Src
  auto comb0 = [&](const double* Src, int J, double* out) -> int {
    GemmDesc g{};
    g.M = (int)N; g.N = 1; g.K = J; g.batch = 1;
    g.A = Src; g.ta = DT::F64; g.lda = N; g.sA = 0; g.a_kcontig = false;
    g.B = Z; g.tb = DT::F64; g.ldb = kGenMaxBasis; g.sB = 0; g.b_ncontig = true;
    g.C = out; g.tc = DT::F64; g.ldc = 1; g.sC = 0;
    g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
    return gemm(g, s);
  };
  auto restart_block = [&](const double* Src, int J, double* Dst) -> int {  // Dst = Z[:, :keep]^T Src
    GemmDesc g{};
    g.M = keep; g.N = (int)N; g.K = J; g.batch = 1;
    g.A = Z; g.ta = DT::F64; g.lda = kGenMaxBasis; g.sA = 0; g.a_kcontig = false;
    g.B = Src; g.tb = DT::F64; g.ldb = N; g.sB = 0; g.b_ncontig = true;
    g.C = Dst; g.tc = DT::F64; g.ldc = N; g.sC = 0;
    g.alpha = 1.0; g.beta = 0.0; g.prec = Prec::F64;
    return gemm(g, s);
  };
  // V[0] = v0 / |v0|_M, its images, G[0][0]
  TNX_TRY(copy_f64(v0, V, N, s));
  TNX_TRY(applyM(V, MV));
  TNX_TRY(dot(V, MV, s8));
  {
    double q = 0.0;
    TNX_TRY(to_host(s8, &q, 1));
    TNX_CHECK_ARG(q > 0.0, "eigsh_gen: v0^T M v0 = %g is not positive", q);
  }
  hipLaunchKernelGGL(gen_scale2_kernel, dim3(nb), dim3(256), 0, s, V, MV, s8, N);
  TNX_LAUNCH_CHECK();
  TNX_TRY(applyH(V, HV));
  TNX_TRY(dot(V, HV, hg));
  hipLaunchKernelGGL(gen_gcol_kernel, dim3(1), dim3(kGenMaxBasis), 0, s, hg, 0, G);
  TNX_LAUNCH_CHECK();
  int J = 1;
  bool converged = false;
  double th = 0.0, rho = 0.0;
  while (true) {
    hipLaunchKernelGGL(small_syev_kernel, dim3(1), dim3(256), 0, s, G, J, kGenMaxBasis, theta, Z);
    TNX_LAUNCH_CHECK();
    TNX_TRY(comb0(V, J, x));
    TNX_TRY(comb0(HV, J, hx));
    TNX_TRY(comb0(MV, J, mx));
    hipLaunchKernelGGL(gen_resid_kernel, dim3(nb), dim3(256), 0, s, hx, mx, theta, r, N);
    TNX_LAUNCH_CHECK();
    TNX_TRY(dot(r, r, s8 + 1));
    TNX_TRY(dot(hx, hx, s8 + 2));
    TNX_TRY(dot(mx, mx, s8 + 3));
    TNX_HIP(hipMemcpyAsync(s8, theta, sizeof(double), hipMemcpyDeviceToDevice, s));
    double hs[4];
    TNX_TRY(to_host(s8, hs, 4));
    th = hs[0];
    rho = std::sqrt(std::max(hs[1], 0.0));
    const double scale = std::sqrt(std::max(hs[2], 0.0)) + std::fabs(th) * std::sqrt(std::max(hs[3], 0.0));
    if (rho <= opts->tol * scale || rho == 0.0 || J >= N) {
      converged = true;
      break;
    }
    if (matvecs + 2 > max_mv) break;
    if (J >= m) {
      // thick restart with the `keep` lowest Ritz vectors: M-orthonormal,
      // and G becomes diag(theta)
      TNX_TRY(restart_block(V, J, V2));
      TNX_TRY(restart_block(HV, J, HV2));
      TNX_TRY(restart_block(MV, J, MV2));
      TNX_HIP(hipMemcpyAsync(V, V2, sizeof(double) * keep * N, hipMemcpyDeviceToDevice, s));
      TNX_HIP(hipMemcpyAsync(HV, HV2, sizeof(double) * keep * N, hipMemcpyDeviceToDevice, s));
      TNX_HIP(hipMemcpyAsync(MV, MV2, sizeof(double) * keep * N, hipMemcpyDeviceToDevice, s));
      hipLaunchKernelGGL(gen_gdiag_kernel, dim3(1), dim3(256), 0, s, theta, keep, G);
      TNX_LAUNCH_CHECK();
      J = keep;
      ++restarts;
    }
    // expansion vector t = r, M-orthogonalised twice, M-normalised
    double* t = V + (int64_t)J * N;
    double* mt = MV + (int64_t)J * N;
    double* ht = HV + (int64_t)J * N;
    TNX_TRY(copy_f64(r, t, N, s));
    for (int pass = 0; pass < 2; ++pass) {
      TNX_TRY(gemv(MV, N, J, false, t, hb, 1.0, 0.0, part, part_bytes, s));   // hb = (MV) t
      TNX_TRY(gemv(V, N, J, true, hb, t, -1.0, 1.0, part, part_bytes, s));    // t -= V^T hb
    }
    TNX_TRY(applyM(t, mt));
    TNX_TRY(dot(t, mt, s8 + 4));
    double q = 0.0;
    TNX_TRY(to_host(s8 + 4, &q, 1));
    if (!(q > 0.0)) break;  // the residual lies in the M-span of V: no progress possible
    hipLaunchKernelGGL(gen_scale2_kernel, dim3(nb), dim3(256), 0, s, t, mt, s8 + 4, N);
    TNX_LAUNCH_CHECK();
    TNX_TRY(applyH(t, ht));
    TNX_TRY(gemv(V, N, J + 1, false, ht, hg, 1.0, 0.0, part, part_bytes, s));  // G[0..J][J]
    hipLaunchKernelGGL(gen_gcol_kernel, dim3(1), dim3(kGenMaxBasis), 0, s, hg, J, G);
    TNX_LAUNCH_CHECK();
    ++J;
  }
  TNX_TRY(copy_f64(x, vec, N, s));
  if (hvec) TNX_TRY(copy_f64(hx, hvec, N, s));
  TNX_HIP(hipStreamSynchronize(s));
  info->eval = th;
  info->resnorm = rho;
  info->anorm = 0.0;
  info->matvecs = matvecs;
  info->matvecs_f64 = matvecs;
  info->restarts = restarts;
  info->converged = converged ? 1 : 0;
  return converged ? TNX_OK : TNX_ERR_NOCONV;
}

}  // namespace tnx

extern "C" {

size_t tnx_eigsh_gen_heff1_workspace_bytes(int chil, int Dl, int Dr, int D2l, int D2r, int d,
                                           int chir, int max_basis) {
  return tnx::eigsh_gen_ws_bytes(chil, Dl, Dr, D2l, D2r, d, chir, max_basis);
}

int tnx_eigsh_gen_heff1(const double* L, const double* W, const double* R, const double* L2,
                        const double* W2, const double* R2, const double* v0, double* vec,
                        double* hvec, int chil, int Dl, int Dr, int D2l, int D2r, int d,
                        int chir, const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info, void* ws,
                        size_t ws_bytes, tnx_stream_t stream) {
  return tnx::eigsh_gen_heff1(L, W, R, L2, W2, R2, v0, vec, hvec, chil, Dl, Dr, D2l, D2r, d, chir,
                              opts, info, ws, ws_bytes, (hipStream_t)stream);
}

}  // extern "C"
