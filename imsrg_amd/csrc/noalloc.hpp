// noalloc.hpp -- the static per-device workspace of the allocation-free
// commutator (noalloc.hip) and the commutator that runs in it (api.cpp).
#pragma once
#include <string>

#include "../../include/imsrg_hip.h"

namespace imsrg {

// Looks up this device's copy of the static workspace (once per device, with
// the device current).  Called when a context is created on the device.
int noalloc_setup(int device, std::string &err);

// One commutator whose per-call workspaces (Pandya matrices, partial sums,
// comm122 coefficients) are carved from `ws` (`cap` doubles) instead of the
// context's own buffers.  Allocates nothing.
// the device a context lives on; sets the context's error message and returns code
int ctx_device(const imsrg_ctx *ctx);
int ctx_fail(imsrg_ctx *ctx, int code, const char *msg);

int commutator_in_workspace(imsrg_ctx *ctx, const imsrg_dop *X, const imsrg_dop *Y, imsrg_dop *Z,
                            const imsrg_comm_opts *opts, double *ws, long cap);

}  // namespace imsrg
