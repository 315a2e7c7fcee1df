// tc_direct_soa_noalloc: the FP64 all-pairs sum of tc_direct_soa with no
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, serving loops).
//
// The bodies are gathered into a fixed file-scope device array, `scratch`,
// instead of the context's growable buffers.  Host code never passes the
// array's name to a kernel: its device address is looked up once per device
// with hipGetSymbolAddress (at tc_create, or on the first call on a device)
// and that address is used everywhere.  Every call on one device shares the
// array, so calls that overlap on one device (two streams, two threads) race.
// Each call rewrites everything it reads from the array (the gathered bodies
// and the identity index list) before reading it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tc_engine.h"

namespace tcfmm {

// Largest n the call accepts, and the array's layout for it: eight FP64
// columns (x y z vx vy vz q acc_old), two int columns (original index, the
// identity gather list) and one byte column (flags); ~19 MiB.
constexpr int kNoAllocMaxBodies = 1 << 18;
constexpr size_t kNaCol = sizeof(double) * (size_t)kNoAllocMaxBodies;
constexpr size_t kNaInt = sizeof(int) * (size_t)kNoAllocMaxBodies;
constexpr size_t kNaBytes = 8 * kNaCol + 2 * kNaInt + (size_t)kNoAllocMaxBodies;

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static unsigned char *g_na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once).
hipError_t noalloc_scratch(unsigned char **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kNaMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_na_ptr[dev];
  return hipSuccess;
}

int noalloc_max_bodies() { return kNoAllocMaxBodies; }

}  // namespace tcfmm

using namespace tcfmm;

extern "C" int tc_direct_soa_noalloc(tc_context *ctx, int n, const double *x, const double *v,
                                     const double *m, const unsigned char *flags, int nt,
                                     const int *targets, double *pot, double *acc,
                                     double *jerk) {
  if (!ctx) return -1;
  if (n <= 0 || nt < 0 || !x || !v || !m || (nt > 0 && (!targets || !pot || !acc || !jerk))) {
    ctx->err = "bad arguments";
    return -1;
  }
  if (n > kNoAllocMaxBodies) {
    ctx->err = "tc_direct_soa_noalloc: n exceeds the fixed scratch size (262144 bodies)";
    return -1;
  }
  unsigned char *base = nullptr;
  hipError_t e = noalloc_scratch(&base);
  if (e != hipSuccess) {
    ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return -2;
  }
  double *col = reinterpret_cast<double *>(base);
  int *ints = reinterpret_cast<int *>(base + 8 * kNaCol);
  const size_t N = kNoAllocMaxBodies;
  BodiesDev b = {col,         col + N,     col + 2 * N,
                 col + 3 * N, col + 4 * N, col + 5 * N,
                 col + 6 * N, base + 8 * kNaCol + 2 * kNaInt,
                 ints,        col + 7 * N};
  int *ident = ints + N;
  hipStream_t s = ctx->stream;
  launch_iota(ident, n, s);
  launch_gather(x, v, m, flags, ident, n, b, s);
  if (nt > 0) {
    launch_zero_f64(pot, (size_t)nt, s);
    launch_zero_f64(acc, 3 * (size_t)nt, s);
    launch_zero_f64(jerk, 3 * (size_t)nt, s);
    launch_direct_sum(b, n, targets, nt, pot, acc, jerk, s);
  }
  e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = hipGetErrorString(e);
    return -2;
  }
  return 0;
}
