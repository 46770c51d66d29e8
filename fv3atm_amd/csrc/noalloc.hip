// noalloc.hip -- fv3_step_dev_noalloc: the FP32 advection + remap step of
// fv3_step_dev with its device scratch in a file-scope array instead of the
// context's allocated arenas, for callers that cannot allocate (code under
// hipGraph capture, serving loops).
//
// The array's name is a host-side stand-in in host code: its device address
// is looked up once per device with hipGetSymbolAddress (at the first
// context created on that device) and only that address is handed to
// kernels.  Every call on a device shares its array and finds the previous
// call's contents there; the step initialises everything it reads (the
// Courant maxima are zeroed by a fill kernel, every other scratch value is
// written before it is read), so no clearing is needed between calls.  Calls
// that overlap on one device race on the array.
#include <mutex>

#include "../../include/fv3atm_amd.h"
#include "fused.h"
#include "fv3_internal.h"
#include "mapz_column.h"

namespace fv3 {

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = FV3_NOALLOC_SCRATCH_BYTES;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void *g_scratch[NOALLOC_MAX_DEVICES] = {};

void *scratch_of(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  return dev >= 0 && dev < NOALLOC_MAX_DEVICES ? g_scratch[dev] : nullptr;
}

}  // namespace

int noalloc_setup(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return set_error("device %d out of the no-allocation scratch's 0..%d", dev, NOALLOC_MAX_DEVICES - 1);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[dev]) return 0;
  int cur = -1;
  FV3_HIP(hipGetDevice(&cur));
  if (cur != dev) FV3_HIP(hipSetDevice(dev));
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || p == nullptr)
    return set_error("hipGetSymbolAddress of the no-allocation scratch failed on device %d: %s", dev,
                     hipGetErrorString(e));
  g_scratch[dev] = p;
  return 0;
}

}  // namespace fv3

using namespace fv3;

extern "C" {

int fv3_step_dev_noalloc(fv3_ctx *h, int form, int nq, const void *q_in, void *q_out, void *pt,
                         const double *dp1, double *delp_out, const void *cx, const void *cy,
                         const void *mfx, const void *mfy, const double *ak, const double *bk,
                         double ptop, int hord, int kord_tr, int kord_tm, double t_min,
                         int nsplit_cap, void *stream) {
  if (!h) return set_error("NULL context");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.grid_set) return set_error("grid metrics not uploaded (fv3_ctx_set_grid)");
  if (form != FV3_FORM_FUSED32)
    return set_error("fv3_step_dev_noalloc runs the FP32 form only (FV3_FORM_FUSED32, got %d)", form);
  if (!q_in || !q_out || !dp1 || !delp_out || !cx || !cy || !mfx || !mfy || !ak || !bk)
    return set_error("fv3_step_dev_noalloc: NULL argument");
  if (q_in == q_out) return set_error("fv3_step_dev_noalloc: q_in and q_out must be distinct buffers");
  if (nq < 1 || nq > REMAP_MAX_NQ) return set_error("nq=%d out of 1..%d", nq, REMAP_MAX_NQ);
  void *p = scratch_of(c.dev);
  if (!p) {
    if (noalloc_setup(c.dev)) return 1;
    p = scratch_of(c.dev);
  }
  hipStream_t saved = c.stream;
  if (stream) c.stream = (hipStream_t)stream;
  const int rc = fused_step(c, true, nq, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk, ptop,
                            hord, kord_tr, kord_tm, t_min, nsplit_cap,
                            StepArena{(char *)p, NOALLOC_SCRATCH_BYTES});
  c.stream = saved;
  return rc;
}

}  // extern "C"
