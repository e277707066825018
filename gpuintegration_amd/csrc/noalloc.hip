// noalloc.hip -- the refinement with its sums from a fixed device array, for
// callers that cannot allocate device memory (code under stream capture,
// serving loops that forbid hipMalloc).
//
// pagani_refine_dev_noalloc runs the kernels of pagani_refine_dev (k_refine,
// k_fold, k_finalize) with their chunk records, fold records and result
// record carved from `scratch`, a file-scope device array of kScratchBytes
// (1.5 MB; every GPU the library runs on holds one copy until the process
// ends).  Its device address is looked up once per device, with that device
// current, on the entry point's first call there; host code uses that
// address, never the array's name.  Nothing in scratch needs clearing:
// every record a launch reads was written earlier in the same call.  All
// calls on one GPU share its array, so calls that overlap on one GPU (two
// streams, two threads) race and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "pagani.h"
#include "phases.hpp"
#include "ws_internal.hpp"

namespace pagani {
namespace internal {
namespace {

namespace ph = pagani::phases;

// one chunk more than a flat list of the maximum size: a list with parents
// is cut at its half (ph::num_chunks(n, true))
constexpr std::size_t kMaxChunks = PAGANI_NOALLOC_MAX_REGIONS / ph::kChunk + 1;
constexpr std::size_t kMaxFolds = (kMaxChunks + ph::kFold - 1) / ph::kFold;
static_assert(PAGANI_NOALLOC_MAX_REGIONS % ph::kChunk == 0, "whole chunks");

constexpr std::size_t align256(std::size_t b) { return (b + 255) / 256 * 256; }

// byte offsets of the sub-buffers inside scratch
constexpr std::size_t kOffPart = 0;
constexpr std::size_t kOffFpart = kOffPart + align256(kMaxChunks * ph::kPartDoubles * 8);
constexpr std::size_t kOffOff = kOffFpart + align256(kMaxFolds * ph::kPartDoubles * 8);
constexpr std::size_t kOffFoff = kOffOff + align256(kMaxChunks * 8);
constexpr std::size_t kOffCnt = kOffFoff + align256(kMaxFolds * 8);
constexpr std::size_t kOffFcnt = kOffCnt + align256(kMaxChunks * 4);
constexpr std::size_t kOffScal = kOffFcnt + align256(kMaxFolds * 4);
constexpr std::size_t kScratchBytes = kOffScal + align256(sizeof(ph::Scalars));

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up on first use with
// that device current (each GPU holds its own copy at its own address).
int scratch_address(pagani_ws* ws, int device, char** out)
{
  if (device < 0 || device >= kMaxDevices)
    return fail(ws, PAGANI_EUNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[device]) {
    int prev = 0;
    HIP_TRY(ws, hipGetDevice(&prev));
    if (prev != device) HIP_TRY(ws, hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return hip_fail(ws, e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(ws, PAGANI_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch[device] = p;
  }
  *out = static_cast<char*>(g_scratch[device]);
  return PAGANI_OK;
}

}  // namespace
}  // namespace internal
}  // namespace pagani

using namespace pagani::internal;

extern "C" int pagani_refine_dev_noalloc(pagani_ws* ws, size_t n, const double* est,
                                         const double* err_in, double* err_out,
                                         const double* parent_est, double epsrel,
                                         int relerr_classification, uint8_t* active,
                                         double* sums_dev, uint64_t* num_active_dev)
{
  if (!ws || !est || !err_in || !err_out || !active)
    return fail(ws, PAGANI_EBADARG, "null argument");
  if (parent_est && (n % 2)) return fail(ws, PAGANI_EBADARG, "refinement needs an even count");
  if (parent_est && err_in == err_out)
    return fail(ws, PAGANI_EBADARG, "refinement cannot run in place");
  if (n > PAGANI_NOALLOC_MAX_REGIONS)
    return fail(ws, PAGANI_EBADARG,
                "pagani_refine_dev_noalloc takes at most PAGANI_NOALLOC_MAX_REGIONS (" +
                    std::to_string(PAGANI_NOALLOC_MAX_REGIONS) + ") regions, got " +
                    std::to_string(n));
  if (n == 0) return PAGANI_OK;
  char* base = nullptr;
  int rc = scratch_address(ws, ws->device, &base);
  if (rc) return rc;
  double* part = reinterpret_cast<double*>(base + kOffPart);
  double* fpart = reinterpret_cast<double*>(base + kOffFpart);
  std::uint64_t* off = reinterpret_cast<std::uint64_t*>(base + kOffOff);
  std::uint64_t* foff = reinterpret_cast<std::uint64_t*>(base + kOffFoff);
  std::uint32_t* cnt = reinterpret_cast<std::uint32_t*>(base + kOffCnt);
  std::uint32_t* fcnt = reinterpret_cast<std::uint32_t*>(base + kOffFcnt);
  ph::Scalars* scal = reinterpret_cast<ph::Scalars*>(base + kOffScal);
  HIP_TRY(ws, ph::refine(ws->stream, n, est, err_in, err_out, parent_est, epsrel,
                         relerr_classification, active, part, cnt));
  HIP_TRY(ws, ph::finalize(ws->stream, ph::num_chunks(n, parent_est != nullptr), part, cnt, off,
                           fpart, fcnt, foff, scal, 0));
  if (sums_dev)
    HIP_TRY(ws, hipMemcpyAsync(sums_dev, scal->sum, 4 * sizeof(double), hipMemcpyDeviceToDevice,
                               ws->stream));
  if (num_active_dev)
    HIP_TRY(ws, hipMemcpyAsync(num_active_dev, &scal->count, sizeof(std::uint64_t),
                               hipMemcpyDeviceToDevice, ws->stream));
  return PAGANI_OK;
}
