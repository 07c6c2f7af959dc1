// Allocation-free form of the one-site H_eff matvec, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). Its scratch (the T1 and T2 intermediates
// and the split-K partial sums) lives in a file-scope device array of fixed
// size instead of a caller workspace; the arithmetic is the one of
// tnx_heff1_matvec, so both give bitwise identical results.
//
// The array is per device: each GPU the library runs on holds its own copy
// until the process ends. Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address): the first
// call on a device, with that device current, looks up the array's device
// address with hipGetSymbolAddress and keeps it per device; every later use
// goes through that address.
//
// All calls on one GPU share that GPU's array, and it keeps the last call's
// contents. Nothing in it needs to start at zero: every intermediate is
// written in full before it is read. Calls that overlap on one GPU (two
// streams, two host threads) race on the array and must be serialised by
// the caller.
#include "tnx_common.h"

#include <mutex>

namespace tnx {

int heff1_matvec(const double* L, const double* W, const double* R, const double* x, double* y,
                 int chil, int Dl, int Dr, int d, int chir, int prec, void* ws, size_t ws_bytes,
                 hipStream_t s);
size_t heff1_ws_bytes(int chil, int Dl, int Dr, int d, int chir);

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array, looked up once per device.
static int scratch_address(void** out) {
  int dev = -1;
  TNX_HIP(hipGetDevice(&dev));
  TNX_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "noalloc: device %d beyond the supported %d", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("noalloc: hipGetSymbolAddress of the scratch array failed on device %d: %s", dev,
                hipGetErrorString(e));
      return TNX_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TNX_OK;
}

}  // namespace tnx

extern "C" {

size_t tnx_noalloc_scratch_bytes(void) { return tnx::kNoAllocScratchBytes; }

int tnx_heff1_matvec_noalloc(const double* L, const double* W, const double* R, const double* x,
                             double* y, int chil, int Dl, int Dr, int d, int chir, int precision,
                             tnx_stream_t stream) {
  using namespace tnx;
  TNX_CHECK_ARG(chil > 0 && Dl > 0 && Dr > 0 && d > 0 && chir > 0,
                "tnx_heff1_matvec_noalloc: bad dims");
  const size_t need = heff1_ws_bytes(chil, Dl, Dr, d, chir);
  TNX_CHECK_ARG(need <= kNoAllocScratchBytes,
                "tnx_heff1_matvec_noalloc: chil=%d Dl=%d Dr=%d d=%d chir=%d needs %zu bytes of "
                "scratch, above the fixed %zu",
                chil, Dl, Dr, d, chir, need, kNoAllocScratchBytes);
  void* ws = nullptr;
  TNX_TRY(scratch_address(&ws));
  NoTuneScope no_tune;  // its GEMMs take table or heuristic tiles: no allocation
  return heff1_matvec(L, W, R, x, y, chil, Dl, Dr, d, chir, precision, ws, kNoAllocScratchBytes,
                      (hipStream_t)stream);
}

}  // extern "C"
