/*
 * tnx — C-ABI of the MI355X finite-DMRG hot path (libtnx.so).
 *
 * The reference (tanlin2013/tnpy) is pure Python: its hot path is a set of
 * Python call sites into quimb / primme / LAPACK. Each entry point below
 * replaces one of those call sites; the comment on each names the
 * reference interface it stands in for (file:line in tnpy).
 *
 * Conventions
 *  - All tensor arguments are device pointers (HBM), caller-owned, dense,
 *    row-major, FP64 unless stated. The library never frees caller memory.
 *  - MPS site tensor A[l][p][r]  ('lpr', tnpy/matrix_product_state.py:45).
 *    Edge sites use a bond of size 1.
 *  - MPO tensor W[wl][wr][k][b]  (quimb 'lrud', k = ket, b = bra,
 *    tnpy/operators.py:97-98), edge tensors padded to a bond of size 1.
 *  - Environments L[l_ket][w][l_bra], R[r_ket][w][r_bra]
 *    (tnpy/matrix_product_state.py:386-407,427-437).
 *  - Scratch memory is a caller-supplied workspace; its size comes from the
 *    matching *_workspace_bytes query.
 *  - Every call is stream-ordered on `stream` (a hipStream_t, NULL = the
 *    default stream) and reentrant. Calls that return host scalars
 *    (tnx_eigsh_heff1, tnx_syev_lowest's info) synchronise that stream;
 *    tnx_qr and tnx_gauge_right/left wait for one read of the QR's panel
 *    flags, with the formation of Q and R already queued behind it (they
 *    may return before that work ends, like any stream-ordered call).
 *  - Status: 0 on success, negative on failure; tnx_last_error() returns a
 *    thread-local message. No C++ exception crosses this boundary.
 */
#ifndef TNX_H_
#define TNX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* tnx_stream_t; /* hipStream_t */

enum {
  TNX_OK = 0,
  TNX_ERR_INVALID = -1,     /* bad argument (null pointer, bad dims, small workspace) */
  TNX_ERR_HIP = -2,         /* HIP runtime error */
  TNX_ERR_NOCONV = -3,      /* iterative solver hit its iteration limit */
  TNX_ERR_UNSUPPORTED = -4  /* shape outside what the kernels support */
};

/* Arithmetic of a contraction. FP32 and BF16X3 read FP64 operands, convert
 * while staging tiles in LDS, accumulate in FP32 on MFMA, and write FP64. */
enum { TNX_F64 = 0, TNX_F32 = 1, TNX_BF16X3 = 2 };

const char* tnx_last_error(void);
int tnx_version(void);

/* ---- generic batched GEMM (building block, FP64 operands) -------------
 * C[b] = alpha * op(A[b]) op(B[b]) + beta * C[b]; A(m,k) = a_kcontig ?
 * A[m*lda+k] : A[k*lda+m]; B(k,n) = b_ncontig ? B[k*ldb+n] : B[n*ldb+k];
 * C(m,n) = C[m*ldc+n]. Replaces the numpy tensordot / BLAS dgemm calls that
 * quimb issues (tnpy/matrix_product_state.py:315,336,438). */
int tnx_gemm_f64(int M, int N, int K, int batch, double alpha,
                 const double* A, int64_t lda, int64_t strideA, int a_kcontig,
                 const double* B, int64_t ldb, int64_t strideB, int b_ncontig,
                 double beta, double* C, int64_t ldc, int64_t strideC,
                 int precision, tnx_stream_t stream);

/* The same GEMM with each operand's element type chosen (f32 flag: 1 =
 * float, 0 = double). With precision TNX_F32 or TNX_BF16X3 and float
 * operands the tiles are staged from FP32 directly (the values a double
 * operand would be rounded to, half the bytes). */
int tnx_gemm(int M, int N, int K, int batch, double alpha, const void* A, int a_f32,
             int64_t lda, int64_t strideA, int a_kcontig, const void* B, int b_f32,
             int64_t ldb, int64_t strideB, int b_ncontig, double beta, void* C, int c_f32,
             int64_t ldc, int64_t strideC, int precision, tnx_stream_t stream);

/* GEMM tile choice. Every GEMM takes its tile configuration from the tile
 * table (shape -> configuration) when the shape is there, else from a fixed
 * heuristic; neither allocates, synchronises or changes results (only
 * split-K changes the summation order, and it follows a fixed rule on the
 * shape). The library's only global state is this table, guarded by a
 * mutex.
 * tnx_set_autotune(1) turns on an inline tuning mode for tuning runs (off by
 * default; TNX_AUTOTUNE=1 in the environment also turns it on): the first
 * GEMM of a shape missing from the table then times every configuration on a
 * scratch output it allocates with hipMalloc, waits on the stream, and adds
 * the fastest to the table. Mode 1 tunes shapes of at least 2^28
 * multiply-adds; tnx_set_autotune(2) (TNX_AUTOTUNE=2) tunes every shape. Do not use it under graph capture or where the
 * caller cannot allow allocation. tnx_tune_clear empties the table.
 * No reference counterpart (quimb leaves this to BLAS). */
int tnx_set_autotune(int on);
int tnx_tune_clear(void);

/* Tile table files: tnx_tune_save writes the table to a text file;
 * tnx_tune_load adds a saved table (returns the number of entries, or a
 * negative status). TNX_TUNE_CACHE=<file> in the environment loads a table
 * before the first GEMM. tnpy_amd loads its committed MI355X table
 * (tnpy_amd/gemm_tiles_mi355x.txt) when it loads the library. No reference
 * counterpart. */
int tnx_tune_load(const char* path);
int tnx_tune_save(const char* path);

/* ---- one-site effective Hamiltonian matvec -----------------------------
 * y[l'][b][r'] = sum L[l][wl][l'] W[wl][wr][k][b] R[r][wr][r'] x[l][k][r]
 * Replaces Environment.one_site_matvec's closure
 * (tnpy/matrix_product_state.py:411-440), the LinearOperator primme calls. */
size_t tnx_heff1_workspace_bytes(int chil, int Dl, int Dr, int d, int chir);
int tnx_heff1_matvec(const double* L, const double* W, const double* R,
                     const double* x, double* y, int chil, int Dl, int Dr,
                     int d, int chir, int precision, void* ws, size_t ws_bytes,
                     tnx_stream_t stream);

/* Allocation-free tnx_heff1_matvec. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * kernels and results (bitwise) as tnx_heff1_matvec, with the scratch in a
 * fixed per-device array of tnx_noalloc_scratch_bytes() = 64 MiB instead of
 * a workspace; sizes whose tnx_heff1_workspace_bytes exceeds it return
 * TNX_ERR_INVALID. Calls no hipMalloc/hipFree. All calls on one GPU share
 * the array: calls that overlap on one GPU race and must be serialised by
 * the caller. */
size_t tnx_noalloc_scratch_bytes(void);
int tnx_heff1_matvec_noalloc(const double* L, const double* W, const double* R,
                             const double* x, double* y, int chil, int Dl, int Dr,
                             int d, int chir, int precision, tnx_stream_t stream);

/* The one-site matvec with FP32 environments (env_f32: L and R are float
 * copies, the form the low-precision cycles of tnx_eigsh_heff1 use; needs
 * precision TNX_F32 or TNX_BF16X3) and/or FP32 input (x_f32). tnx_to_f32
 * rounds n doubles to floats (device). */
int tnx_heff1_matvec_ex(const void* L, const double* W, const void* R, const void* x, double* y,
                        int chil, int Dl, int Dr, int d, int chir, int env_f32, int x_f32,
                        int precision, void* ws, size_t ws_bytes, tnx_stream_t stream);
int tnx_to_f32(const double* x, float* y, int64_t n, tnx_stream_t stream);

/* Dense H_eff, M[(l,k,r)][(l',b,r')]. Replaces
 * Environment.one_site_full_matrix (tnpy/matrix_product_state.py:372-409). */
int tnx_heff1_dense(const double* L, const double* W, const double* R,
                    double* M, int chil, int Dl, int Dr, int d, int chir,
                    tnx_stream_t stream);

/* ---- environment updates ----------------------------------------------
 * Lout[r][wr][r'] = sum L[l][wl][l'] A[l][k][r] W[wl][wr][k][b] A[l'][b][r']
 * Replaces Environment.update_left (tnpy/matrix_product_state.py:296-315). */
size_t tnx_env_workspace_bytes(int chil, int Dl, int Dr, int d, int chir);
int tnx_env_left(const double* L, const double* A, const double* W,
                 double* Lout, int chil, int Dl, int Dr, int d, int chir,
                 int precision, void* ws, size_t ws_bytes, tnx_stream_t stream);
/* Rout[l][wl][l'] = sum R[r][wr][r'] A[l][k][r] W[wl][wr][k][b] A[l'][b][r']
 * Replaces Environment.update_right (tnpy/matrix_product_state.py:317-336). */
int tnx_env_right(const double* R, const double* A, const double* W,
                  double* Rout, int chil, int Dl, int Dr, int d, int chir,
                  int precision, void* ws, size_t ws_bytes, tnx_stream_t stream);

/* ---- small dense symmetric eigensolver (lowest pair) -------------------
 * M: n x n symmetric (device), n <= 256. Writes the unit eigenvector of
 * the smallest eigenvalue to evec (device, n) and the eigenvalue to
 * *eval_host. Replaces linalg.eigh -> np.linalg.eigh (tnpy/linalg.py:42-61)
 * as used by FiniteDMRG.one_site_solver below exact_solver_dim
 * (tnpy/finite_dmrg.py:108-110). Synchronises the stream. */
size_t tnx_syev_lowest_workspace_bytes(int n);
int tnx_syev_lowest(const double* M, int n, double* evec, double* eval_host,
                    void* ws, size_t ws_bytes, tnx_stream_t stream);

/* ---- Krylov lowest eigenpair of the one-site H_eff ---------------------
 * Thick-restart Lanczos with full CGS2 reorthogonalisation, keeping the
 * `keep` lowest Ritz vectors at a restart (clustered spectra). Converged when
 * the residual norm ||H y - theta y|| <= tol * ||H||_est (||H||_est = largest
 * |Ritz value| seen, PRIMME's criterion). With precision != TNX_F64 the
 * Krylov matvecs run in that precision and the Ritz pair is corrected by
 * FP64 residuals until the FP64 criterion holds.
 * v0 (device, N=chil*d*chir) is the start vector; vec (device, N) receives
 * the unit eigenvector; hvec (device, N, may be NULL) receives H*vec.
 * Replaces linalg.eigshmv -> primme.eigsh(which='SA', k=1)
 * (tnpy/linalg.py:64-87) as called at tnpy/finite_dmrg.py:111. */
typedef struct {
  double tol;         /* residual tolerance (relative to ||H||_est) */
  int max_basis;      /* Lanczos basis size before a thick restart (>= 3) */
  int max_matvecs;    /* iteration limit (all precisions) */
  int precision;      /* TNX_F64 / TNX_F32 / TNX_BF16X3 for the Krylov matvecs */
  int keep;           /* Ritz vectors kept at a thick restart (0 = default) */
  int steps_per_check;  /* batched solver only: Lanczos steps queued between
                           host reads of the per-chain statuses, each chain's
                           convergence decided on the device (0 or 1 = read
                           after every step, with the next step queued
                           first); ignored by the single-problem solvers */
} tnx_eigsh_opts_t;

typedef struct {
  double eval;        /* Rayleigh quotient (FP64) of the returned vector */
  double resnorm;     /* residual norm of the returned pair */
  double anorm;       /* ||H|| estimate used by the criterion */
  int matvecs;        /* total H_eff applications */
  int matvecs_f64;    /* of which in FP64 */
  int restarts;
  int converged;      /* 1 if the criterion was met */
} tnx_eigsh_info_t;

size_t tnx_eigsh_heff1_workspace_bytes(int chil, int Dl, int Dr, int d,
                                       int chir, int max_basis);
int tnx_eigsh_heff1(const double* L, const double* W, const double* R,
                    const double* v0, double* vec, double* hvec, int chil,
                    int Dl, int Dr, int d, int chir,
                    const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info,
                    void* ws, size_t ws_bytes, tnx_stream_t stream);

/* ---- two-site effective Hamiltonian ------------------------------------
 * y[l'][c1][c2][r'] = sum L[l][w][l'] W1[w][u][k1][c1] W2[u][v][k2][c2]
 *                         R[r][v][r'] x[l][k1][k2][r]
 * No reference function: FiniteDMRG.two_site_solver is a stub returning
 * NotImplemented (tnpy/finite_dmrg.py:113-114); these entry points fill it. */
size_t tnx_heff2_workspace_bytes(int chil, int Dl, int Dm, int Dr, int d, int chir);
int tnx_heff2_matvec(const double* L, const double* W1, const double* W2,
                     const double* R, const double* x, double* y, int chil,
                     int Dl, int Dm, int Dr, int d, int chir, int precision,
                     void* ws, size_t ws_bytes, tnx_stream_t stream);
/* Lowest eigenpair of the two-site H_eff; same solver and options as
 * tnx_eigsh_heff1 (the two-site counterpart of tnpy/linalg.py:64-87). */
size_t tnx_eigsh_heff2_workspace_bytes(int chil, int Dl, int Dm, int Dr, int d,
                                       int chir, int max_basis);
int tnx_eigsh_heff2(const double* L, const double* W1, const double* W2,
                    const double* R, const double* v0, double* vec, double* hvec,
                    int chil, int Dl, int Dm, int Dr, int d, int chir,
                    const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info,
                    void* ws, size_t ws_bytes, tnx_stream_t stream);

/* ---- gauge moves (untruncated) ----------------------------------------
 * Rightward: A (chil*d x chir) = Q R; A <- Q (left-orthonormal);
 *   next_out (chir x next_cols) = R * next. Replaces
 *   MatrixProductState.split_tensor RIGHTWARD + linalg.svd
 *   (tnpy/matrix_product_state.py:200-211, tnpy/linalg.py:9-23): the
 *   untruncated SVD U*(sV^T) is replaced by Q*R, the same state in another
 *   gauge of the bond.
 * Leftward: A (chil x d*chir) = L Q; A <- Q (right-orthonormal);
 *   prev_out (prev_rows x chil) = prev * L. Replaces the LEFTWARD branch
 *   (tnpy/matrix_product_state.py:212-223). next/prev must not alias the
 *   outputs. */
size_t tnx_gauge_workspace_bytes(int chil, int d, int chir);
int tnx_gauge_right(double* A, const double* next, double* next_out,
                    int chil, int d, int chir, int next_cols, void* ws,
                    size_t ws_bytes, tnx_stream_t stream);
int tnx_gauge_left(double* A, const double* prev, double* prev_out,
                   int chil, int d, int chir, int prev_rows, void* ws,
                   size_t ws_bytes, tnx_stream_t stream);

/* ---- truncated SVD ------------------------------------------------------
 * X (a x b, row-major, device, not modified): the k leading left singular
 * vectors U (a x k, orthonormal columns, device), Y = U^T X (k x b, device)
 * and the k leading singular values, descending, to sv_host (host). The
 * decomposition works on X itself (two QR steps, then one-sided block
 * Jacobi on the rows of the triangular factor), never on X X^T, so the
 * singular values are accurate to ~ eps * s_max. sweeps (host, may be NULL)
 * receives the number of Jacobi sweeps. Synchronises the stream.
 * Replaces linalg.svd (tnpy/linalg.py:9-23, numpy.linalg.svd cut to
 * `cutoff` values) applied to the two-site block, the truncation to chi of
 * two-site DMRG (FiniteDMRG.two_site_solver, a stub in the reference at
 * tnpy/finite_dmrg.py:113-114), and the Schmidt spectra of
 * ExactDiagonalization.reduced_density_matrix
 * (tnpy/exact_diagonalization.py:62-90). */
size_t tnx_svd_trunc_workspace_bytes(int a, int b, int k);
int tnx_svd_trunc_f64(const double* X, int a, int b, int k, double* U, double* Y,
                      double* sv_host, int* sweeps, void* ws, size_t ws_bytes,
                      tnx_stream_t stream);

/* ---- dense symmetric eigensolver and the dense MPO -----------------------
 * tnx_syev_f64: every eigenpair of the symmetric n x n matrix A (device, not
 * modified): eigenvalues ascending to evals_host (host, n), eigenvectors as
 * the columns of evecs (device, n x n row-major, orthonormal). Computed as
 * the SVD (tnx_svd_trunc_f64's algorithm) of the positive-definite A + cI,
 * c = 1 + the largest absolute row sum. Synchronises the stream. Replaces
 * numpy.linalg.eigh in ExactDiagonalization._eigen_solver
 * (tnpy/exact_diagonalization.py:58-60).
 * tnx_mpo_absorb: one MPO site absorbed into the dense operator built so
 * far, Tout[v][K*d+k][B*d+b] = sum_w T[w][K][B] W[w][v][k][b] (T: Dt x dimK x
 * dimK, W: Dt x Dr x d x d, Tout: Dr x dimK*d x dimK*d, all device). Site by
 * site from T = [[[1]]] this is FullHamiltonian.matrix
 * (tnpy/operators.py:119-154, quimb's MPO contract + fuse). */
size_t tnx_syev_workspace_bytes(int n);
int tnx_syev_f64(const double* A, int n, double* evals_host, double* evecs, void* ws,
                 size_t ws_bytes, tnx_stream_t stream);
int tnx_mpo_absorb(const double* T, int Dt, int64_t dimK, const double* W, int Dr, int d,
                   double* Tout, tnx_stream_t stream);

/* ---- shift-invert DMRG: the generalized one-site problem -----------------
 * Lowest eigenpair of A x = lambda B x with B symmetric positive definite,
 * x normalised to x^T B x = 1. The reference's ShiftInvertDMRG.one_site_solver
 * (tnpy/finite_dmrg.py:347-360) passes the one-site operators of the MPO
 * (H - eps) as A and of its square as B (M=) to linalg.eigh / eigshmv
 * (tnpy/linalg.py:42-87: scipy.linalg.eigh(a, b) below exact_solver_dim,
 * primme.eigsh(A, M=B) above).
 * tnx_sygv_lowest: dense n x n A and B (device): B = W diag(l) W^T
 * (tnx_syev_f64), S = W diag(l)^-1/2, lowest pair (mu, y) of S^T A S,
 * x = S y. Synchronises the stream.
 * tnx_eigsh_gen_heff1: A = H_eff(L, W, R) (bond Dl, Dr), B = H_eff(L2, W2,
 * R2) (bond D2l, D2r), one-site operators of the same site; generalized
 * Davidson on a B-orthonormal basis (no preconditioner), stopping when
 * ||A x - lambda B x|| <= tol (||A x|| + |lambda| ||B x||); thick restart
 * keeps opts->keep Ritz vectors; FP64 only. hvec (may be NULL) receives A x.
 * Returns TNX_ERR_NOCONV at the matvec limit. Synchronises the stream. */
size_t tnx_sygv_lowest_workspace_bytes(int n);
int tnx_sygv_lowest(const double* A, const double* B, int n, double* x, double* eval_host,
                    void* ws, size_t ws_bytes, tnx_stream_t stream);
size_t tnx_eigsh_gen_heff1_workspace_bytes(int chil, int Dl, int Dr, int D2l, int D2r, int d,
                                           int chir, int max_basis);
int tnx_eigsh_gen_heff1(const double* L, const double* W, const double* R, const double* L2,
                        const double* W2, const double* R2, const double* v0, double* vec,
                        double* hvec, int chil, int Dl, int Dr, int D2l, int D2r, int d,
                        int chir, const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info,
                        void* ws, size_t ws_bytes, tnx_stream_t stream);
/* One site of H|phi> as an MPS: out[(l,wl)][k][(r,wr)] =
 * sum_b W[wl][wr][k][b] A[l][b][r] (device; out is l*Dl x d x r*Dr). The
 * restored state of ShiftInvertDMRG._restore_mps (tnpy/finite_dmrg.py:316-345,
 * quimb contract + fuse). */
int tnx_mpo_apply_site(const double* W, const double* A, int Dl, int Dr, int d, int l, int r,
                       double* out, tnx_stream_t stream);

/* Householder QR of a row-major m x n matrix, m >= n: A <- Q (m x n,
 * orthonormal columns), R (n x n, upper triangular, row-major).
 * Building block of the gauge moves (numpy.linalg.qr / LAPACK dgeqrf+dorgqr,
 * tnpy/linalg.py:26-39). */
size_t tnx_qr_workspace_bytes(int m, int n);
int tnx_qr(double* A, int m, int n, double* R, void* ws, size_t ws_bytes,
           tnx_stream_t stream);

/* ---- BLAS-1 (FP64, HBM bound) ------------------------------------------
 * Perturbation psi += alpha * H psi (tnpy/finite_dmrg.py:139-141) is
 * tnx_axpy with alpha = 1e-5. tnx_dot writes x.y to *out_host and
 * synchronises; its reduction order is fixed (deterministic). */
int tnx_axpy(int64_t n, double alpha, const double* x, double* y,
             tnx_stream_t stream);
int tnx_scal(int64_t n, double alpha, double* x, tnx_stream_t stream);
size_t tnx_dot_workspace_bytes(int64_t n);
int tnx_dot(int64_t n, const double* x, const double* y, double* out_host,
            void* ws, size_t ws_bytes, tnx_stream_t stream);

/* ---- batched forms: many independent chains of one shape -----------------
 * The reference runs one FiniteDMRG per chain (one disorder realisation,
 * tnpy/finite_dmrg.py:24-41 with an MPO from tnpy/model/); sweeping a set of
 * realisations is a Python loop over those objects. These entry points do
 * the same site step for `batch` chains of equal length at once: every
 * operand is a contiguous stack of `batch` per-chain blocks of the size the
 * single-chain call uses (e.g. L: batch x chil*Dl*chil), and one launch per
 * kernel covers the whole stack. Results per chain equal those of the
 * single-chain call on that chain up to rounding. */

/* Stacked tnx_heff1_matvec / tnx_heff1_dense
 * (tnpy/matrix_product_state.py:411-440 and :372-409, per chain). */
size_t tnx_heff1_batched_workspace_bytes(int chil, int Dl, int Dr, int d, int chir, int batch);
int tnx_heff1_matvec_batched(const double* L, const double* W, const double* R,
                             const double* x, double* y, int chil, int Dl, int Dr,
                             int d, int chir, int batch, int precision, void* ws,
                             size_t ws_bytes, tnx_stream_t stream);
int tnx_heff1_dense_batched(const double* L, const double* W, const double* R,
                            double* M, int chil, int Dl, int Dr, int d, int chir,
                            int batch, tnx_stream_t stream);

/* Stacked tnx_env_left / tnx_env_right
 * (tnpy/matrix_product_state.py:296-315 and :317-336, per chain). */
size_t tnx_env_batched_workspace_bytes(int chil, int Dl, int Dr, int d, int chir, int batch);
int tnx_env_left_batched(const double* L, const double* A, const double* W,
                         double* Lout, int chil, int Dl, int Dr, int d, int chir,
                         int batch, int precision, void* ws, size_t ws_bytes,
                         tnx_stream_t stream);
int tnx_env_right_batched(const double* R, const double* A, const double* W,
                          double* Rout, int chil, int Dl, int Dr, int d, int chir,
                          int batch, int precision, void* ws, size_t ws_bytes,
                          tnx_stream_t stream);

/* Stacked tnx_syev_lowest: eval_host receives `batch` eigenvalues
 * (np.linalg.eigh below exact_solver_dim, tnpy/finite_dmrg.py:108-110). */
size_t tnx_syev_lowest_batched_workspace_bytes(int n, int batch);
int tnx_syev_lowest_batched(const double* M, int n, int batch, double* evec,
                            double* eval_host, void* ws, size_t ws_bytes,
                            tnx_stream_t stream);

/* Stacked tnx_eigsh_heff1, FP64 only (precision must be TNX_F64, else
 * TNX_ERR_UNSUPPORTED). Lockstep thick-restart Lanczos: every chain keeps
 * its own Krylov scalars and leaves the launched set when its own residual
 * meets the criterion; info receives `batch` records; hvec is required.
 * (primme.eigsh via tnpy/linalg.py:64-87, called at
 * tnpy/finite_dmrg.py:111, per chain.) */
size_t tnx_eigsh_heff1_batched_workspace_bytes(int chil, int Dl, int Dr, int d, int chir,
                                               int max_basis, int batch);
int tnx_eigsh_heff1_batched(const double* L, const double* W, const double* R,
                            const double* v0, double* vec, double* hvec, int chil,
                            int Dl, int Dr, int d, int chir, int batch,
                            const tnx_eigsh_opts_t* opts, tnx_eigsh_info_t* info,
                            void* ws, size_t ws_bytes, tnx_stream_t stream);

/* Stacked tnx_gauge_right / tnx_gauge_left for small bonds: each QR is held
 * in LDS by one workgroup, so m * (n + 1) <= 16640 (m = chil*d, n = chir
 * rightward; m = d*chir, n = chil leftward), else TNX_ERR_INVALID.
 * (tnpy/matrix_product_state.py:200-223 + tnpy/linalg.py:9-23, per chain.) */
size_t tnx_gauge_batched_workspace_bytes(int chil, int d, int chir, int batch);
int tnx_gauge_right_batched(double* A, const double* next, double* next_out,
                            int chil, int d, int chir, int next_cols, int batch,
                            void* ws, size_t ws_bytes, tnx_stream_t stream);
int tnx_gauge_left_batched(double* A, const double* prev, double* prev_out,
                           int chil, int d, int chir, int prev_rows, int batch,
                           void* ws, size_t ws_bytes, tnx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TNX_H_ */
