// hd_noalloc.hip -- hd_norm_and_error_noalloc: the L2 norm and error of
// hd_norm_and_error for callers that cannot allocate (code under HIP graph
// capture, tight serving loops).  Its scratch -- the norm kernel's block
// partial sums -- is a file-scope device array of fixed size instead of a
// buffer the library allocates, and the result goes to a caller-owned device
// array instead of host memory, so a call allocates nothing and does not
// synchronise.
//
// The array's name is a host-side stand-in in host code: its device address
// comes from hipGetSymbolAddress, once per device (each GPU holds its own
// copy), and only that address is handed to kernels.
#include "../../include/hyperdeal_amd.h"
#include "hd_device.h"
#include "hd_handles.h"

#include <algorithm>
#include <mutex>

namespace
{
// the norm kernel's partial sums: two per block, at most kNoallocBlocks
// blocks striding over the cells (so any mesh fits)
constexpr int kNoallocBlocks = HD_NOALLOC_NORM_BLOCKS;
__device__ __align__(256) double scratch[2 * kNoallocBlocks];

constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// the device address of this device's scratch (looked up on the first call
// on the device, with that device current)
hipError_t scratch_of_current_device(double **out)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return e;
  if (dev < 0 || dev >= kMaxDevices)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev])
    {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess)
        return e;
      g_scratch[dev] = static_cast<double *>(p);
    }
  *out = g_scratch[dev];
  return hipSuccess;
}
} // namespace

using hd_abi::fail;
using hd_abi::hip_fail;
using hd_abi::ok;

extern "C" {

int hd_norm_and_error_noalloc(hd_matrix_free *mf, const double *src,
                              int function, const double *params, double time,
                              double *out_dev, void *stream)
{
  if (!mf || !src || !out_dev)
    return fail(HD_ERR_INVALID, "hd_norm_and_error_noalloc: null argument");
  const int vol = hd::ipow(std::max(mf->nd, mf->nq), mf->D);
  const size_t lds = (2 * (size_t)vol + 512) * sizeof(double);
  if (lds > 160 * 1024)
    return fail(HD_ERR_UNSUPPORTED,
                "cell too large for hd_norm_and_error_noalloc");
  double *partial = nullptr;
  hipError_t e = scratch_of_current_device(&partial);
  if (e != hipSuccess)
    return hip_fail(e, "hd_norm_and_error_noalloc: scratch address");
  hd::NormDev nt{};
  std::copy(mf->tables.S.begin(), mf->tables.S.end(), nt.S);
  std::copy(mf->tables.qpoints.begin(), mf->tables.qpoints.end(), nt.qp);
  std::copy(mf->tables.qweights.begin(), mf->tables.qweights.end(), nt.w);
  nt.nd = mf->nd;
  nt.nq = mf->nq;
  hd::FuncDev f{};
  f.id = function;
  for (int d = 0; d < mf->D; ++d)
    f.prm[d] = params ? params[d] : 0.0;
  f.nd = mf->nd;
  f.dim_x = mf->dim_x;
  // every block writes its two partials, and the finish kernel reads those
  // of this launch only: nothing to clear between calls
  const int blocks =
    (int)std::min<long long>(mf->mesh.n_cells, (long long)kNoallocBlocks);
  hipStream_t s = hd_abi::as_stream(stream);
  e = hd::launch_norm(src, mf->mesh, nt, f, time, vol, partial, mf->geo, s,
                      blocks);
  if (e == hipSuccess)
    e = hd::launch_norm_finish(partial, blocks, out_dev, s);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_norm_and_error_noalloc");
}

} // extern "C"
