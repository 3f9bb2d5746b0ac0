// noalloc.hip -- imsrg_hip_commutator_ss_dev_noalloc: the device commutator
// with every per-call workspace in one static device array, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).
//
// The array is file-scope device memory; each GPU holds its own copy from
// module load until the process ends.  Host code never uses the array's name
// as an address: noalloc_setup looks up each device's copy once
// (hipGetSymbolAddress, with that device current) and every later use goes
// through that pointer.  All calls on one GPU share its copy, so calls that
// overlap on one GPU (other host threads, other contexts' streams) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "noalloc.hpp"

namespace {
// 56 MiB: the workspaces of bases up to emax = 6 (O16 reference: 24.7 MiB of
// Pandya matrices plus small lists, 7.1 MiB for comm122's W, and 15.1 MiB of
// bf16x3 operand panels in mixed precision).  Larger bases are rejected.
constexpr long kScratchDoubles = 7L << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

namespace imsrg {

int noalloc_setup(int device, std::string &err) {
  if (device < 0 || device >= kMaxDevices) {
    err = "device index beyond the static workspace table";
    return IMSRG_E_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return IMSRG_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    err = std::string("hipGetSymbolAddress of the static workspace: ") + hipGetErrorString(e);
    return IMSRG_E_HIP;
  }
  g_scratch_ptr[device] = static_cast<double *>(p);
  return IMSRG_OK;
}

}  // namespace imsrg

extern "C" int imsrg_hip_commutator_ss_dev_noalloc(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y,
                                                   imsrg_dop *Z, const imsrg_comm_opts *opts) {
  if (!ctx) return IMSRG_E_INVALID;
  const int device = imsrg::ctx_device(ctx);
  double *ws = nullptr;
  if (device >= 0 && device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ws = g_scratch_ptr[device];
  }
  if (!ws) return imsrg::ctx_fail(ctx, IMSRG_E_HIP, "the static workspace of this device is not set up");
  return imsrg::commutator_in_workspace(ctx, X, Y, Z, opts, ws, kScratchDoubles);
}
