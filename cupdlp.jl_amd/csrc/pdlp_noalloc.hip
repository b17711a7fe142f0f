// pdlp_noalloc.hip — the device region behind pdlp_solve_noalloc.
//
// One file-scope array per device holds every device buffer of a
// no-allocation solve (the LP's arena, the slice and block tables, the
// partial sums, the preparation scratch), so that call makes no hipMalloc,
// hipFree or hipHostMalloc.  Each device the library runs on holds its own
// copy of the array, at its own address, until the process ends.  In host
// code the array's name only stands in for the symbol: its device address is
// taken once per device with hipGetSymbolAddress (scratch_setup, called by
// pdlp_ctx_create with that device current) and only that address is used.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pdlp_noalloc.h"

namespace pdlp {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
}  // namespace

int scratch_setup(int device, std::string *err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = "device index beyond the no-allocation scratch table";
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return 0;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    *err = std::string("hipGetSymbolAddress of the scratch region: ") + hipGetErrorString(e);
    return -(int)(e == hipSuccess ? hipErrorInvalidSymbol : e);
  }
  g_ptr[device] = p;
  return 0;
}

void *scratch_region(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace pdlp
