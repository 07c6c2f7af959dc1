// One-site effective-Hamiltonian matvec, environment updates and the dense
// effective matrix.
//
// Each contraction is GEMM -> W-apply -> GEMM:
//   matvec  (tnpy/matrix_product_state.py:411-440)
//     T1[a,k,v,s] = sum_r   x[a,k,r] R[r,v,s]            GEMM, K = chir
//     T2[a,w,c,s] = sum_k,v W[w,v,k,c] T1[a,k,v,s]        W-apply
//     y[b,c,s]    = sum_a,w L[a,w,b] T2[a,w,c,s]          GEMM, K = chil*Dl
//   left env  (tnpy/matrix_product_state.py:296-315)
//     T1[w,b,k,r] = sum_a   L[a,w,b] A[a,k,r]            GEMM, K = chil
//     T2[v,b,c,r] = sum_w,k W[w,v,k,c] T1[w,b,k,r]        W-apply
//     L'[r,v,s]   = sum_b,c T2[v,b,c,r] A[b,c,s]          Dr batched GEMMs
//   right env (tnpy/matrix_product_state.py:317-336)
//     T1[a,k,v,s] = sum_r   A[a,k,r] R[r,v,s]            GEMM
//     T2[a,w,c,s] = sum_k,v W[w,v,k,c] T1[a,k,v,s]        W-apply
//     R'[a,w,b]   = sum_c,s T2[a,w,c,s] A[b,c,s]          GEMM (B k-contiguous)
// The GEMMs carry 2*D*d*chi^3 flops each; the W-apply is a D*d-wide local
// map, bound by HBM (each (outer, inner) pair reads d*Dr values and writes
// Dl*d values, coalesced along the inner bond index).
#include "tnx_common.h"

namespace tnx {

// out[o][p][q][i] = sum_{x,y} C(p,q,x,y) in[o][x][y][i], C from W by a
// permutation selected by `mode`:
//   mode 0 (matvec / right env): p=w, q=c, x=k, y=v : C = W[w][v][k][c]
//   mode 1 (left env)          : p=v, q=c, x=w, y=k : C = W[w][v][k][c]
struct WApply {
  const void* in; void* out; const double* W;
  int in_f32, out_f32;
  int64_t n_outer, n_inner;
  int64_t in_so, in_sx, in_sy;     // strides of o, x, y in `in`
  int64_t out_so, out_sp, out_sq;  // strides of o, p, q in `out`
  int NX, NY, NP, NQ;
  int Dl, Dr, d, mode;
  // batch of independent problems (grid.y): strides of in, out and W
  int batch = 1;
  int64_t in_sb = 0, out_sb = 0, W_sb = 0;
  const int* bidx = nullptr;  // launched -> logical problem, or null
};

template <int MAXIN>
__global__ __launch_bounds__(256) void wapply_kernel(WApply a) {
  __shared__ double C[64 * 64];
  const int nin = a.NX * a.NY, nout = a.NP * a.NQ;
  {
    const int b = a.bidx ? a.bidx[blockIdx.y] : (int)blockIdx.y;
    a.in = (const char*)a.in + (size_t)b * a.in_sb * (a.in_f32 ? 4 : 8);
    a.out = (char*)a.out + (size_t)b * a.out_sb * (a.out_f32 ? 4 : 8);
    a.W += (size_t)b * a.W_sb;
  }
  for (int t = threadIdx.x; t < nout * nin; t += blockDim.x) {
    const int oi = t / nin, ii = t % nin;
    const int p = oi / a.NQ, q = oi % a.NQ;
    const int x = ii / a.NY, y = ii % a.NY;
    int w, v, k, c;
    if (a.mode == 0) { w = p; c = q; k = x; v = y; }
    else             { v = p; c = q; w = x; k = y; }
    C[t] = a.W[((w * a.Dr + v) * a.d + k) * a.d + c];
  }
  __syncthreads();
  const int64_t total = a.n_outer * a.n_inner;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = g / a.n_inner, i = g % a.n_inner;
    double xin[MAXIN];
#pragma unroll
    for (int t = 0; t < MAXIN; ++t) {
      if (t < nin) {
        const int x = t / a.NY, y = t % a.NY;
        const int64_t off = o * a.in_so + x * a.in_sx + y * a.in_sy + i;
        xin[t] = a.in_f32 ? (double)reinterpret_cast<const float*>(a.in)[off]
                          : reinterpret_cast<const double*>(a.in)[off];
      }
    }
    for (int oi = 0; oi < nout; ++oi) {
      const int p = oi / a.NQ, q = oi % a.NQ;
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < MAXIN; ++t)
        if (t < nin) s += C[oi * nin + t] * xin[t];
      const int64_t off = o * a.out_so + p * a.out_sp + q * a.out_sq + i;
      if (a.out_f32) reinterpret_cast<float*>(a.out)[off] = (float)s;
      else reinterpret_cast<double*>(a.out)[off] = s;
    }
  }
}

static int launch_wapply(const WApply& a, hipStream_t s) {
  const int nin = a.NX * a.NY;
  TNX_CHECK_ARG(a.NP * a.NQ <= 64 && nin <= 64,
                "W-apply: D*d = %d / %d exceeds 64", a.NP * a.NQ, nin);
  const int64_t total = a.n_outer * a.n_inner;
  if (total == 0) return TNX_OK;
  TNX_CHECK_ARG(a.batch >= 1 && a.batch <= 65535, "W-apply: bad batch %d", a.batch);
  int blocks = (int)std::min<int64_t>((total + 255) / 256, std::max<int64_t>(8192 / a.batch, 64));
  if (nin <= 16)
    hipLaunchKernelGGL(wapply_kernel<16>, dim3(blocks, a.batch), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(wapply_kernel<64>, dim3(blocks, a.batch), dim3(256), 0, s, a);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

__global__ void to_f32_cvt_kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

static bool valid_dims(int chil, int Dl, int Dr, int d, int chir) {
  return chil > 0 && Dl > 0 && Dr > 0 && d > 0 && chir > 0;
}

// Element size of the intermediates for a given precision.
static size_t inter_bytes(int prec) { return prec == TNX_F64 ? 8 : 4; }

// Workspace: T1, T2, and FP64 split-K partials for the last GEMM (its
// output, chil x d*chir, has few tiles at chil = chir and K = chil*Dl is
// long; the split rule in gemm.hip decides how many are used).
// partial sums of up to 4 K-splits of an output of out_elems doubles
static size_t splitk_bytes(int64_t out_elems) { return (size_t)out_elems * 4 * 8; }

size_t heff1_ws_bytes_b(int chil, int Dl, int Dr, int d, int chir, int batch) {
  const size_t t1 = (size_t)chil * d * Dr * chir;
  const size_t t2 = (size_t)chil * Dl * d * chir;
  return ((t1 + t2) * 8 + splitk_bytes((int64_t)chil * d * chir)) * (size_t)batch + 512;
}

size_t heff1_ws_bytes(int chil, int Dl, int Dr, int d, int chir) {
  return heff1_ws_bytes_b(chil, Dl, Dr, d, chir, 1);
}

// y_p = H_eff(L_p, W_p, R_p) x_p for the problems p of a stack of `batch`:
// all of them (bidx null) or the `count` listed in bidx. L, W, R are
// contiguous stacks; x_p = x + p*sx, y_p = y + p*sy (sx = sy = chil*d*chir
// for contiguous stacks). One problem: batch = count = 1. x is FP64, or
// FP32 when x_f32 (a low-precision Krylov basis); y is FP64.
int heff1_matvec_b(const void* L, const double* W, const void* R, const void* x, int x_f32,
                   double* y, int chil, int Dl, int Dr, int d, int chir, int prec, int batch,
                   int count, const int* bidx, int64_t sx, int64_t sy, void* ws, size_t ws_bytes,
                   hipStream_t s, int env_f32) {
  TNX_CHECK_ARG(valid_dims(chil, Dl, Dr, d, chir), "heff1_matvec: bad dims");
  TNX_CHECK_ARG(batch >= 1 && batch <= 65535, "heff1_matvec: bad batch %d", batch);
  TNX_CHECK_ARG(count >= 1 && count <= batch && (bidx || count == batch),
                "heff1_matvec: bad count %d", count);
  TNX_CHECK_ARG(L && W && R && x && y, "heff1_matvec: null pointer");
  TNX_CHECK_ARG(prec >= TNX_F64 && prec <= TNX_BF16X3, "heff1_matvec: bad precision");
  TNX_CHECK_ARG(!env_f32 || prec != TNX_F64, "heff1_matvec: FP32 environments need a low-precision arithmetic");
  TNX_CHECK_ARG(ws && ws_bytes >= heff1_ws_bytes_b(chil, Dl, Dr, d, chir, batch),
                "heff1_matvec: workspace too small");
  const size_t eb = inter_bytes(prec);
  const DT it = prec == TNX_F64 ? DT::F64 : DT::F32;
  const int64_t nx = (int64_t)chil * d * chir;
  const int64_t nt1 = (int64_t)chil * d * Dr * chir, nt2 = (int64_t)chil * Dl * d * chir;
  char* T1 = (char*)ws;
  char* T2 = T1 + (((size_t)nt1 * batch * eb + 255) & ~(size_t)255);
  char* P = T2 + (((size_t)nt2 * batch * eb + 255) & ~(size_t)255);
  // T1[(a,k),(v,s)] = X[(a,k), r] R[r, (v,s)]
  GemmDesc g1{};
  g1.M = chil * d; g1.N = Dr * chir; g1.K = chir; g1.batch = count; g1.bidx = bidx;
  g1.A = x; g1.ta = x_f32 ? DT::F32 : DT::F64; g1.lda = chir; g1.sA = sx; g1.a_kcontig = true;
  g1.B = R; g1.tb = env_f32 ? DT::F32 : DT::F64; g1.ldb = (int64_t)Dr * chir; g1.sB = (int64_t)chir * Dr * chir; g1.b_ncontig = true;
  g1.C = T1; g1.tc = it; g1.ldc = (int64_t)Dr * chir; g1.sC = nt1;
  g1.alpha = 1.0; g1.beta = 0.0; g1.prec = (Prec)prec;
  TNX_TRY(gemm(g1, s));
  WApply wa{};
  wa.in = T1; wa.out = T2; wa.W = W; wa.in_f32 = wa.out_f32 = (prec != TNX_F64);
  wa.n_outer = chil; wa.n_inner = chir;
  wa.in_so = (int64_t)d * Dr * chir; wa.in_sx = (int64_t)Dr * chir; wa.in_sy = chir;
  wa.out_so = (int64_t)Dl * d * chir; wa.out_sp = (int64_t)d * chir; wa.out_sq = chir;
  wa.NX = d; wa.NY = Dr; wa.NP = Dl; wa.NQ = d;
  wa.Dl = Dl; wa.Dr = Dr; wa.d = d; wa.mode = 0;
  wa.batch = count; wa.bidx = bidx;
  wa.in_sb = nt1; wa.out_sb = nt2; wa.W_sb = (int64_t)Dl * Dr * d * d;
  TNX_TRY(launch_wapply(wa, s));
  // y[b, (c,s)] = sum_{(a,w)} L[(a,w), b] T2[(a,w), (c,s)]
  GemmDesc g3{};
  g3.M = chil; g3.N = d * chir; g3.K = chil * Dl; g3.batch = count; g3.bidx = bidx;
  g3.A = L; g3.ta = env_f32 ? DT::F32 : DT::F64; g3.lda = chil; g3.sA = (int64_t)chil * Dl * chil; g3.a_kcontig = false;
  g3.B = T2; g3.tb = it; g3.ldb = (int64_t)d * chir; g3.sB = nt2; g3.b_ncontig = true;
  g3.C = y; g3.tc = DT::F64; g3.ldc = (int64_t)d * chir; g3.sC = sy;
  g3.alpha = 1.0; g3.beta = 0.0; g3.prec = (Prec)prec;
  g3.splitk_ws = reinterpret_cast<double*>(P);
  g3.splitk_ws_bytes = splitk_bytes(nx) * batch;
  return gemm(g3, s);
}

int heff1_matvec(const double* L, const double* W, const double* R,
                 const double* x, double* y, int chil, int Dl, int Dr, int d,
                 int chir, int prec, void* ws, size_t ws_bytes, hipStream_t s) {
  const int64_t nx = (int64_t)chil * d * chir;
  return heff1_matvec_b(L, W, R, x, 0, y, chil, Dl, Dr, d, chir, prec, 1, 1, nullptr, nx, nx, ws,
                        ws_bytes, s, 0);
}

// Two-site effective Hamiltonian (no reference function: the two-site
// solver is a stub at tnpy/finite_dmrg.py:113-114):
//   y[l',c1,c2,r'] = sum L[l,w,l'] W1[w,u,k1,c1] W2[u,v,k2,c2] R[r,v,r'] x[l,k1,k2,r]
//   T1[(l,k1,k2),(v,r')] = X[(l,k1,k2), r] R[r,(v,r')]          GEMM, K = chir
//   T2[l,k1,u,c2,r']     = sum_k2,v W2[u,v,k2,c2] T1[l,k1,k2,v,r']   W-apply
//   T3[l,w,c1,c2,r']     = sum_k1,u W1[w,u,k1,c1] T2[l,k1,u,c2,r']   W-apply
//   y[l',(c1,c2,r')]     = sum_(l,w) L[(l,w),l'] T3[(l,w),(c1,c2,r')]  GEMM, K = chil*Dl
// T3 reuses T1's memory.
size_t heff2_ws_bytes(int chil, int Dl, int Dm, int Dr, int d, int chir) {
  const size_t t1 = (size_t)chil * d * d * Dr * chir;
  const size_t t3 = (size_t)chil * Dl * d * d * chir;
  const size_t t2 = (size_t)chil * d * Dm * d * chir;
  return (std::max(t1, t3) + t2) * 8 + 256;
}

int heff2_matvec(const double* L, const double* W1, const double* W2, const double* R,
                 const void* x, int x_f32, double* y, int chil, int Dl, int Dm, int Dr, int d,
                 int chir, int prec, void* ws, size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(chil > 0 && Dl > 0 && Dm > 0 && Dr > 0 && d > 0 && chir > 0, "heff2_matvec: bad dims");
  TNX_CHECK_ARG(L && W1 && W2 && R && x && y, "heff2_matvec: null pointer");
  TNX_CHECK_ARG(prec >= TNX_F64 && prec <= TNX_BF16X3, "heff2_matvec: bad precision");
  TNX_CHECK_ARG(ws && ws_bytes >= heff2_ws_bytes(chil, Dl, Dm, Dr, d, chir),
                "heff2_matvec: workspace too small");
  const size_t eb = inter_bytes(prec);
  const DT it = prec == TNX_F64 ? DT::F64 : DT::F32;
  const size_t t1 = (size_t)chil * d * d * Dr * chir;
  const size_t t3 = (size_t)chil * Dl * d * d * chir;
  char* T1 = (char*)ws;
  char* T2 = T1 + ((std::max(t1, t3) * eb + 255) & ~(size_t)255);
  char* T3 = T1;
  GemmDesc g1{};
  g1.M = chil * d * d; g1.N = Dr * chir; g1.K = chir; g1.batch = 1;
  g1.A = x; g1.ta = x_f32 ? DT::F32 : DT::F64; g1.lda = chir; g1.sA = 0; g1.a_kcontig = true;
  g1.B = R; g1.tb = DT::F64; g1.ldb = (int64_t)Dr * chir; g1.sB = 0; g1.b_ncontig = true;
  g1.C = T1; g1.tc = it; g1.ldc = (int64_t)Dr * chir; g1.sC = 0;
  g1.alpha = 1.0; g1.beta = 0.0; g1.prec = (Prec)prec;
  TNX_TRY(gemm(g1, s));
  WApply a2{};  // W2 over (k2, v) -> (u, c2), outer (l,k1), inner r'
  a2.in = T1; a2.out = T2; a2.W = W2; a2.in_f32 = a2.out_f32 = (prec != TNX_F64);
  a2.n_outer = (int64_t)chil * d; a2.n_inner = chir;
  a2.in_so = (int64_t)d * Dr * chir; a2.in_sx = (int64_t)Dr * chir; a2.in_sy = chir;
  a2.out_so = (int64_t)Dm * d * chir; a2.out_sp = (int64_t)d * chir; a2.out_sq = chir;
  a2.NX = d; a2.NY = Dr; a2.NP = Dm; a2.NQ = d;
  a2.Dl = Dm; a2.Dr = Dr; a2.d = d; a2.mode = 0;
  TNX_TRY(launch_wapply(a2, s));
  WApply a1{};  // W1 over (k1, u) -> (w, c1), outer l, inner (c2, r')
  a1.in = T2; a1.out = T3; a1.W = W1; a1.in_f32 = a1.out_f32 = (prec != TNX_F64);
  a1.n_outer = chil; a1.n_inner = (int64_t)d * chir;
  a1.in_so = (int64_t)d * Dm * d * chir; a1.in_sx = (int64_t)Dm * d * chir; a1.in_sy = (int64_t)d * chir;
  a1.out_so = (int64_t)Dl * d * d * chir; a1.out_sp = (int64_t)d * d * chir; a1.out_sq = (int64_t)d * chir;
  a1.NX = d; a1.NY = Dm; a1.NP = Dl; a1.NQ = d;
  a1.Dl = Dl; a1.Dr = Dm; a1.d = d; a1.mode = 0;
  TNX_TRY(launch_wapply(a1, s));
  GemmDesc g3{};
  g3.M = chil; g3.N = d * d * chir; g3.K = chil * Dl; g3.batch = 1;
  g3.A = L; g3.ta = DT::F64; g3.lda = chil; g3.sA = 0; g3.a_kcontig = false;
  g3.B = T3; g3.tb = it; g3.ldb = (int64_t)d * d * chir; g3.sB = 0; g3.b_ncontig = true;
  g3.C = y; g3.tc = DT::F64; g3.ldc = (int64_t)d * d * chir; g3.sC = 0;
  g3.alpha = 1.0; g3.beta = 0.0; g3.prec = (Prec)prec;
  return gemm(g3, s);
}

size_t env_ws_bytes_b(int chil, int Dl, int Dr, int d, int chir, int batch) {
  return (size_t)chil * d * chir * (Dl + Dr) * 8 * (size_t)batch + 512;
}

size_t env_ws_bytes(int chil, int Dl, int Dr, int d, int chir) {
  return env_ws_bytes_b(chil, Dl, Dr, d, chir, 1);
}

// Left environment update for a contiguous stack of `batch` problems
// (L, A, W, Lout each a stack of per-problem blocks).
int env_left_b(const double* L, const double* A, const double* W, double* Lout,
               int chil, int Dl, int Dr, int d, int chir, int prec, int batch, void* ws,
               size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(valid_dims(chil, Dl, Dr, d, chir), "env_left: bad dims");
  TNX_CHECK_ARG(batch >= 1 && (int64_t)batch * Dr <= 65535, "env_left: bad batch %d", batch);
  TNX_CHECK_ARG(L && A && W && Lout, "env_left: null pointer");
  TNX_CHECK_ARG(prec >= TNX_F64 && prec <= TNX_BF16X3, "env_left: bad precision");
  TNX_CHECK_ARG(ws && ws_bytes >= env_ws_bytes_b(chil, Dl, Dr, d, chir, batch),
                "env_left: workspace too small");
  const size_t eb = inter_bytes(prec);
  const DT it = prec == TNX_F64 ? DT::F64 : DT::F32;
  const int64_t nA = (int64_t)chil * d * chir;
  const int64_t nT1 = (int64_t)Dl * nA, nT2 = (int64_t)Dr * nA;
  char* T1 = (char*)ws;
  char* T2 = T1 + (((size_t)nT1 * batch * eb + 255) & ~(size_t)255);
  // T1[(w,b),(k,r)] = sum_a L[a,(w,b)] A[a,(k,r)]
  GemmDesc g1{};
  g1.M = Dl * chil; g1.N = d * chir; g1.K = chil; g1.batch = batch;
  g1.A = L; g1.ta = DT::F64; g1.lda = (int64_t)Dl * chil; g1.sA = (int64_t)chil * Dl * chil; g1.a_kcontig = false;
  g1.B = A; g1.tb = DT::F64; g1.ldb = (int64_t)d * chir; g1.sB = nA; g1.b_ncontig = true;
  g1.C = T1; g1.tc = it; g1.ldc = (int64_t)d * chir; g1.sC = nT1;
  g1.alpha = 1.0; g1.beta = 0.0; g1.prec = (Prec)prec;
  TNX_TRY(gemm(g1, s));
  WApply wa{};
  wa.in = T1; wa.out = T2; wa.W = W; wa.in_f32 = wa.out_f32 = (prec != TNX_F64);
  wa.n_outer = chil; wa.n_inner = chir;
  wa.in_so = (int64_t)d * chir; wa.in_sx = nA; wa.in_sy = chir;
  wa.out_so = (int64_t)d * chir; wa.out_sp = nA; wa.out_sq = chir;
  wa.NX = Dl; wa.NY = d; wa.NP = Dr; wa.NQ = d;
  wa.Dl = Dl; wa.Dr = Dr; wa.d = d; wa.mode = 1;
  wa.batch = batch; wa.in_sb = nT1; wa.out_sb = nT2; wa.W_sb = (int64_t)Dl * Dr * d * d;
  TNX_TRY(launch_wapply(wa, s));
  // L'[r, v, s] = sum_{(b,c)} T2[v][(b,c)][r] A[(b,c)][s], batched over
  // (problem, v)
  GemmDesc g3{};
  g3.M = chir; g3.N = chir; g3.K = chil * d; g3.batch = batch * Dr; g3.nb2 = Dr;
  g3.A = T2; g3.ta = it; g3.lda = chir; g3.sA = nT2; g3.sA2 = nA; g3.a_kcontig = false;
  g3.B = A; g3.tb = DT::F64; g3.ldb = chir; g3.sB = nA; g3.sB2 = 0; g3.b_ncontig = true;
  g3.C = Lout; g3.tc = DT::F64; g3.ldc = (int64_t)Dr * chir; g3.sC = (int64_t)chir * Dr * chir;
  g3.sC2 = chir;
  g3.alpha = 1.0; g3.beta = 0.0; g3.prec = (Prec)prec;
  return gemm(g3, s);
}

int env_left(const double* L, const double* A, const double* W, double* Lout,
             int chil, int Dl, int Dr, int d, int chir, int prec, void* ws,
             size_t ws_bytes, hipStream_t s) {
  return env_left_b(L, A, W, Lout, chil, Dl, Dr, d, chir, prec, 1, ws, ws_bytes, s);
}

// Right environment update for a contiguous stack of `batch` problems.
int env_right_b(const double* R, const double* A, const double* W, double* Rout,
                int chil, int Dl, int Dr, int d, int chir, int prec, int batch, void* ws,
                size_t ws_bytes, hipStream_t s) {
  TNX_CHECK_ARG(valid_dims(chil, Dl, Dr, d, chir), "env_right: bad dims");
  TNX_CHECK_ARG(batch >= 1 && batch <= 65535, "env_right: bad batch %d", batch);
  TNX_CHECK_ARG(R && A && W && Rout, "env_right: null pointer");
  TNX_CHECK_ARG(prec >= TNX_F64 && prec <= TNX_BF16X3, "env_right: bad precision");
  TNX_CHECK_ARG(ws && ws_bytes >= env_ws_bytes_b(chil, Dl, Dr, d, chir, batch),
                "env_right: workspace too small");
  const size_t eb = inter_bytes(prec);
  const DT it = prec == TNX_F64 ? DT::F64 : DT::F32;
  const int64_t nA = (int64_t)chil * d * chir;
  const int64_t nT1 = (int64_t)Dr * nA, nT2 = (int64_t)Dl * nA;
  char* T1 = (char*)ws;
  char* T2 = T1 + (((size_t)nT1 * batch * eb + 255) & ~(size_t)255);
  GemmDesc g1{};
  g1.M = chil * d; g1.N = Dr * chir; g1.K = chir; g1.batch = batch;
  g1.A = A; g1.ta = DT::F64; g1.lda = chir; g1.sA = nA; g1.a_kcontig = true;
  g1.B = R; g1.tb = DT::F64; g1.ldb = (int64_t)Dr * chir; g1.sB = (int64_t)chir * Dr * chir; g1.b_ncontig = true;
  g1.C = T1; g1.tc = it; g1.ldc = (int64_t)Dr * chir; g1.sC = nT1;
  g1.alpha = 1.0; g1.beta = 0.0; g1.prec = (Prec)prec;
  TNX_TRY(gemm(g1, s));
  WApply wa{};
  wa.in = T1; wa.out = T2; wa.W = W; wa.in_f32 = wa.out_f32 = (prec != TNX_F64);
  wa.n_outer = chil; wa.n_inner = chir;
  wa.in_so = (int64_t)d * Dr * chir; wa.in_sx = (int64_t)Dr * chir; wa.in_sy = chir;
  wa.out_so = (int64_t)Dl * d * chir; wa.out_sp = (int64_t)d * chir; wa.out_sq = chir;
  wa.NX = d; wa.NY = Dr; wa.NP = Dl; wa.NQ = d;
  wa.Dl = Dl; wa.Dr = Dr; wa.d = d; wa.mode = 0;
  wa.batch = batch; wa.in_sb = nT1; wa.out_sb = nT2; wa.W_sb = (int64_t)Dl * Dr * d * d;
  TNX_TRY(launch_wapply(wa, s));
  // R'[(a,w), b] = sum_{(c,s)} T2[(a,w),(c,s)] A[b,(c,s)]
  GemmDesc g3{};
  g3.M = chil * Dl; g3.N = chil; g3.K = d * chir; g3.batch = batch;
  g3.A = T2; g3.ta = it; g3.lda = (int64_t)d * chir; g3.sA = nT2; g3.a_kcontig = true;
  g3.B = A; g3.tb = DT::F64; g3.ldb = (int64_t)d * chir; g3.sB = nA; g3.b_ncontig = false;
  g3.C = Rout; g3.tc = DT::F64; g3.ldc = chil; g3.sC = (int64_t)chil * Dl * chil;
  g3.alpha = 1.0; g3.beta = 0.0; g3.prec = (Prec)prec;
  return gemm(g3, s);
}

int env_right(const double* R, const double* A, const double* W, double* Rout,
              int chil, int Dl, int Dr, int d, int chir, int prec, void* ws,
              size_t ws_bytes, hipStream_t s) {
  return env_right_b(R, A, W, Rout, chil, Dl, Dr, d, chir, prec, 1, ws, ws_bytes, s);
}

// M[(a,k,r),(b,c,s)] = sum_{w,v} L[a,w,b] W[w,v,k,c] R[r,v,s]
// Stacks over blockIdx.y: problem p reads L, W, R, M + p * (their sizes).
__global__ void heff_dense_kernel(const double* L, const double* W, const double* R,
                                  double* M, int chil, int Dl, int Dr, int d, int chir) {
  const int64_t N = (int64_t)chil * d * chir;
  {
    const int64_t p = blockIdx.y;
    L += p * chil * Dl * chil;
    W += p * Dl * Dr * d * d;
    R += p * chir * Dr * chir;
    M += p * N * N;
  }
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < N * N;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = g / N, col = g % N;
    const int a = (int)(row / ((int64_t)d * chir)), k = (int)((row / chir) % d), r = (int)(row % chir);
    const int b = (int)(col / ((int64_t)d * chir)), c = (int)((col / chir) % d), s = (int)(col % chir);
    double acc = 0.0;
    for (int w = 0; w < Dl; ++w) {
      const double l = L[((int64_t)a * Dl + w) * chil + b];
      if (l == 0.0) continue;
      for (int v = 0; v < Dr; ++v)
        acc += l * W[((w * Dr + v) * d + k) * d + c] * R[((int64_t)r * Dr + v) * chir + s];
    }
    M[g] = acc;
  }
}

int heff1_dense_b(const double* L, const double* W, const double* R, double* M,
                  int chil, int Dl, int Dr, int d, int chir, int batch, hipStream_t s) {
  TNX_CHECK_ARG(valid_dims(chil, Dl, Dr, d, chir), "heff1_dense: bad dims");
  TNX_CHECK_ARG(batch >= 1 && batch <= 65535, "heff1_dense: bad batch %d", batch);
  TNX_CHECK_ARG(L && W && R && M, "heff1_dense: null pointer");
  const int64_t N = (int64_t)chil * d * chir;
  TNX_CHECK_ARG(N <= 8192, "heff1_dense: dimension %lld too large for a dense H_eff", (long long)N);
  const int blocks = (int)std::min<int64_t>((N * N + 255) / 256, std::max(4096 / batch, 16));
  hipLaunchKernelGGL(heff_dense_kernel, dim3(blocks, batch), dim3(256), 0, s, L, W, R, M, chil,
                     Dl, Dr, d, chir);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

int heff1_dense(const double* L, const double* W, const double* R, double* M,
                int chil, int Dl, int Dr, int d, int chir, hipStream_t s) {
  return heff1_dense_b(L, W, R, M, chil, Dl, Dr, d, chir, 1, s);
}

// ---- elementwise helpers ---------------------------------------------------
__global__ void transpose_kernel(const double* in, double* out, int rows, int cols) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int r = by + j, c = bx + threadIdx.x;
    if (r < rows && c < cols) tile[j][threadIdx.x] = in[(int64_t)r * cols + c];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int c = bx + j, r = by + threadIdx.x;
    if (r < rows && c < cols) out[(int64_t)c * rows + r] = tile[threadIdx.x][j];
  }
}

int transpose_f64(const double* in, double* out, int rows, int cols, hipStream_t s) {
  if (rows == 0 || cols == 0) return TNX_OK;
  TNX_CHECK_ARG(in && out && in != out, "transpose: bad pointers");
  TNX_CHECK_ARG(ceil_div(rows, 32) <= 65535, "transpose: too many rows");
  dim3 grid(ceil_div(cols, 32), ceil_div(rows, 32));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, s, in, out, rows, cols);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

__global__ void copy_kernel(const double* in, double* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}
__global__ void fill_kernel(double* out, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = v;
}

int copy_f64(const double* in, double* out, int64_t n, hipStream_t s) {
  if (n <= 0) return TNX_OK;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s, in, out, n);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

int fill_f64(double* out, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return TNX_OK;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, out, n, v);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

}  // namespace tnx

// ---- C-ABI -----------------------------------------------------------------
extern "C" {

size_t tnx_heff1_workspace_bytes(int chil, int Dl, int Dr, int d, int chir) {
  return tnx::heff1_ws_bytes(chil, Dl, Dr, d, chir);
}

int tnx_heff1_matvec(const double* L, const double* W, const double* R,
                     const double* x, double* y, int chil, int Dl, int Dr,
                     int d, int chir, int precision, void* ws, size_t ws_bytes,
                     tnx_stream_t stream) {
  return tnx::heff1_matvec(L, W, R, x, y, chil, Dl, Dr, d, chir, precision, ws,
                           ws_bytes, (hipStream_t)stream);
}

int tnx_heff1_dense(const double* L, const double* W, const double* R,
                    double* M, int chil, int Dl, int Dr, int d, int chir,
                    tnx_stream_t stream) {
  return tnx::heff1_dense(L, W, R, M, chil, Dl, Dr, d, chir, (hipStream_t)stream);
}

size_t tnx_env_workspace_bytes(int chil, int Dl, int Dr, int d, int chir) {
  return tnx::env_ws_bytes(chil, Dl, Dr, d, chir);
}

size_t tnx_heff2_workspace_bytes(int chil, int Dl, int Dm, int Dr, int d, int chir) {
  return tnx::heff2_ws_bytes(chil, Dl, Dm, Dr, d, chir);
}

int tnx_heff2_matvec(const double* L, const double* W1, const double* W2, const double* R,
                     const double* x, double* y, int chil, int Dl, int Dm, int Dr, int d,
                     int chir, int precision, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::heff2_matvec(L, W1, W2, R, x, 0, y, chil, Dl, Dm, Dr, d, chir, precision, ws,
                           ws_bytes, (hipStream_t)stream);
}

int tnx_env_left(const double* L, const double* A, const double* W,
                 double* Lout, int chil, int Dl, int Dr, int d, int chir,
                 int precision, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::env_left(L, A, W, Lout, chil, Dl, Dr, d, chir, precision, ws,
                       ws_bytes, (hipStream_t)stream);
}

int tnx_env_right(const double* R, const double* A, const double* W,
                  double* Rout, int chil, int Dl, int Dr, int d, int chir,
                  int precision, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  return tnx::env_right(R, A, W, Rout, chil, Dl, Dr, d, chir, precision, ws,
                        ws_bytes, (hipStream_t)stream);
}

size_t tnx_heff1_batched_workspace_bytes(int chil, int Dl, int Dr, int d, int chir, int batch) {
  return tnx::heff1_ws_bytes_b(chil, Dl, Dr, d, chir, batch);
}

int tnx_heff1_matvec_ex(const void* L, const double* W, const void* R, const void* x, double* y,
                        int chil, int Dl, int Dr, int d, int chir, int env_f32, int x_f32,
                        int precision, void* ws, size_t ws_bytes, tnx_stream_t stream) {
  const int64_t nx = (int64_t)chil * d * chir;
  return tnx::heff1_matvec_b(L, W, R, x, x_f32, y, chil, Dl, Dr, d, chir, precision, 1, 1, nullptr,
                             nx, nx, ws, ws_bytes, (hipStream_t)stream, env_f32);
}

int tnx_to_f32(const double* x, float* y, int64_t n, tnx_stream_t stream) {
  TNX_CHECK_ARG(x && y && n >= 0, "tnx_to_f32: bad arguments");
  if (n == 0) return TNX_OK;
  const int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(tnx::to_f32_cvt_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, x, y, n);
  TNX_LAUNCH_CHECK();
  return TNX_OK;
}

int tnx_heff1_matvec_batched(const double* L, const double* W, const double* R,
                             const double* x, double* y, int chil, int Dl, int Dr,
                             int d, int chir, int batch, int precision, void* ws,
                             size_t ws_bytes, tnx_stream_t stream) {
  const int64_t nx = (int64_t)chil * d * chir;
  return tnx::heff1_matvec_b(L, W, R, x, 0, y, chil, Dl, Dr, d, chir, precision, batch, batch,
                             nullptr, nx, nx, ws, ws_bytes, (hipStream_t)stream, 0);
}

int tnx_heff1_dense_batched(const double* L, const double* W, const double* R,
                            double* M, int chil, int Dl, int Dr, int d, int chir,
                            int batch, tnx_stream_t stream) {
  return tnx::heff1_dense_b(L, W, R, M, chil, Dl, Dr, d, chir, batch, (hipStream_t)stream);
}

size_t tnx_env_batched_workspace_bytes(int chil, int Dl, int Dr, int d, int chir, int batch) {
  return tnx::env_ws_bytes_b(chil, Dl, Dr, d, chir, batch);
}

int tnx_env_left_batched(const double* L, const double* A, const double* W,
                         double* Lout, int chil, int Dl, int Dr, int d, int chir,
                         int batch, int precision, void* ws, size_t ws_bytes,
                         tnx_stream_t stream) {
  return tnx::env_left_b(L, A, W, Lout, chil, Dl, Dr, d, chir, precision, batch, ws, ws_bytes,
                         (hipStream_t)stream);
}

int tnx_env_right_batched(const double* R, const double* A, const double* W,
                          double* Rout, int chil, int Dl, int Dr, int d, int chir,
                          int batch, int precision, void* ws, size_t ws_bytes,
                          tnx_stream_t stream) {
  return tnx::env_right_b(R, A, W, Rout, chil, Dl, Dr, d, chir, precision, batch, ws, ws_bytes,
                          (hipStream_t)stream);
}

}  // extern "C"
