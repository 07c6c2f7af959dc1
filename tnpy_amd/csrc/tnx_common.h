// Shared definitions of the tnx library: status codes, error reporting,
// launch checks and small device helpers. Internal header (not part of the
// C-ABI in include/tnx.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/tnx.h"

namespace tnx {

// Thread-local message of the last failing call on this host thread.
void set_error(const char* fmt, ...);
const char* last_error();

#define TNX_CHECK_ARG(cond, ...)                                   \
  do {                                                             \
    if (!(cond)) {                                                 \
      ::tnx::set_error(__VA_ARGS__);                               \
      return TNX_ERR_INVALID;                                      \
    }                                                              \
  } while (0)

#define TNX_HIP(call)                                              \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) {                                        \
      ::tnx::set_error("%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                        \
      return TNX_ERR_HIP;                                          \
    }                                                              \
  } while (0)

#define TNX_LAUNCH_CHECK()                                         \
  do {                                                             \
    hipError_t e_ = hipGetLastError();                             \
    if (e_ != hipSuccess) {                                        \
      ::tnx::set_error("kernel launch: %s (%s:%d)",                \
                       hipGetErrorString(e_), __FILE__, __LINE__); \
      return TNX_ERR_HIP;                                          \
    }                                                              \
  } while (0)

#define TNX_TRY(expr)                                              \
  do {                                                             \
    int s_ = (expr);                                               \
    if (s_ != TNX_OK) return s_;                                   \
  } while (0)

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Precision of a contraction (matches tnx_precision_t in include/tnx.h).
enum class Prec : int { F64 = TNX_F64, F32 = TNX_F32, BF16X3 = TNX_BF16X3 };

// ---------------------------------------------------------------------------
// Generic strided, batched GEMM on MFMA:
//   C[b](m, n) = alpha * sum_k A[b](m, k) * B[b](k, n) + beta * C[b](m, n)
// A(m, k) = a_kcontig ? A[m*lda + k] : A[k*lda + m]
// B(k, n) = b_ncontig ? B[k*ldb + n] : B[n*ldb + k]
// C(m, n) = C[m*ldc + n]                      (row-major output)
// Element types: A/B are double or float, C is double or float.
// ---------------------------------------------------------------------------
enum class DT : int { F64 = 0, F32 = 1 };

struct GemmDesc {
  int M, N, K, batch;
  const void* A; DT ta; int64_t lda; int64_t sA; bool a_kcontig;
  const void* B; DT tb; int64_t ldb; int64_t sB; bool b_ncontig;
  void* C;       DT tc; int64_t ldc; int64_t sC;
  double alpha, beta;
  Prec prec;
  // Optional FP64 scratch for split-K partial sums (used when the output
  // has few tiles and K is long); null disables splitting.
  double* splitk_ws = nullptr;
  size_t splitk_ws_bytes = 0;
  // Optional second batch level: batch index b = b1 * nb2 + b2 addresses
  // A + b1*sA + b2*sA2 (same for B, C). nb2 = 1 means a single level.
  int nb2 = 1;
  int64_t sA2 = 0, sB2 = 0, sC2 = 0;
  // Optional index list: launched batch i computes logical batch bidx[i]
  // (a subset of a larger stack); null means launched == logical.
  const int* bidx = nullptr;
};

int gemm(const GemmDesc& g, hipStream_t stream);
void set_autotune(int on);
void tune_clear();

// While one of these lives on a host thread, the GEMMs that thread issues
// never enter the inline tuning mode (which allocates and synchronises):
// they take the tile table entry or the heuristic.
struct NoTuneScope {
  NoTuneScope();
  ~NoTuneScope();
  NoTuneScope(const NoTuneScope&) = delete;
  NoTuneScope& operator=(const NoTuneScope&) = delete;
};

// Elementwise kernels used by several modules.
int transpose_f64(const double* in, double* out, int rows, int cols,
                  hipStream_t stream);  // out[c][r] = in[r][c]
int copy_f64(const double* in, double* out, int64_t n, hipStream_t stream);
int fill_f64(double* out, int64_t n, double v, hipStream_t stream);

// 1/sqrt(x) and 1/a on the device: the hardware estimates (v_rsq_f64,
// v_rcp_f64) refined by two Newton steps. Shorter dependent chains than the
// correctly rounded sequences; the results are within an ulp or two, which
// the factorisations that use them absorb.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

__device__ __forceinline__ double rcp_nr(double a) {
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  return fma(r, fma(-a, r, 1.0), r);
}

// Pinned host blocks for status reads (the Lanczos solvers' per-step
// statuses, the QR's panel flags): a copy into one can be waited for with an
// event while later work stays queued behind it. A call leases a block for
// its duration and returns it; blocks come from one process-wide pool and
// are never freed (they live until the process ends), so no HIP call runs
// when a host thread or the process exits, and the pool never holds more
// blocks than calls ever ran at once. get() is null when the allocation
// fails (the caller then reads its status the plain way).
class PinnedLease {
 public:
  explicit PinnedLease(size_t n) {
    {
      std::lock_guard<std::mutex> lk(mu());
      auto& fl = free_list();
      size_t best = fl.size();
      for (size_t i = 0; i < fl.size(); ++i)
        if (fl[i].second >= n && (best == fl.size() || fl[i].second < fl[best].second)) best = i;
      if (best < fl.size()) {
        p_ = fl[best].first;
        cap_ = fl[best].second;
        fl.erase(fl.begin() + (std::ptrdiff_t)best);
        return;
      }
    }
    const size_t want = std::max<size_t>(n, 64);
    if (hipHostMalloc((void**)&p_, want * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      p_ = nullptr;
      (void)hipGetLastError();
      return;
    }
    cap_ = want;
  }
  ~PinnedLease() {
    if (!p_) return;
    std::lock_guard<std::mutex> lk(mu());
    free_list().emplace_back(p_, cap_);
  }
  PinnedLease(const PinnedLease&) = delete;
  PinnedLease& operator=(const PinnedLease&) = delete;
  double* get() const { return p_; }

 private:
  static std::mutex& mu() {
    static std::mutex m;
    return m;
  }
  static std::vector<std::pair<double*, size_t>>& free_list() {
    // never destroyed: returning a block during process exit stays valid
    static auto* v = new std::vector<std::pair<double*, size_t>>();
    return *v;
  }
  double* p_ = nullptr;
  size_t cap_ = 0;
};

// An event for the duration of one call (a solve, a QR).
struct StepEvent {
  hipEvent_t e = nullptr;
  StepEvent() {
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      e = nullptr;
      (void)hipGetLastError();
    }
  }
  ~StepEvent() {
    if (e) (void)hipEventDestroy(e);
  }
  StepEvent(const StepEvent&) = delete;
  StepEvent& operator=(const StepEvent&) = delete;
};

}  // namespace tnx
