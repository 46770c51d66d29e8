// gsum.hip -- global tracer mass on the device: for each tracer, the sum over
// every tile, level and compute-domain cell of q * delp * area, in FP64 with
// Neumaier compensation and a fixed reduction order, so the result is the
// same on every run and within a couple of units in the last place of the
// exact sum of the products.  It plays the part of the dycore's global mass
// sums (g_sum over mpp_global_sum, whose bit-reproducible mode the reference
// imports as BITWISE_EXACT_SUM, moving_nest/fv_moving_nest_utils.F90:35,41),
// which the FP32 step keeps in FP64.
#include <hip/hip_runtime.h>

#include "../../include/fv3atm_amd.h"
#include "fv3_internal.h"

namespace fv3 {

namespace {

struct Ks {
  double s, c;
};

// Neumaier's update: the rounding error of s + x is recovered exactly and
// kept in c
__device__ __forceinline__ void ks_add(Ks &k, double x) {
  const double t = k.s + x;
  k.c += fabs(k.s) >= fabs(x) ? (k.s - t) + x : (x - t) + k.s;
  k.s = t;
}
__device__ __forceinline__ void ks_merge(Ks &a, const Ks &b) {
  ks_add(a, b.s);
  a.c += b.c;
}

constexpr int GS_BLOCK = 256, GS_GRID = 512;

struct MassDims {
  int ni, nj, nid, ioff, joff, npz, nt;
  long long s2, gs2;
};

// block partials of one tracer (blockIdx.y): thread r takes cells r,
// r + GS_BLOCK * GS_GRID, ... of the tracer's (tile, level, j, i) cells
template <typename T>
__global__ void __launch_bounds__(GS_BLOCK) k_mass_part(const T *q, const double *delp, const double *area,
                                                        MassDims m, Ks *part) {
  __shared__ Ks sh[GS_BLOCK];
  const int iq = blockIdx.y;
  const long long per = (long long)m.ni * m.nj, n = per * m.npz * m.nt;
  Ks k{0.0, 0.0};
  for (long long c = (long long)blockIdx.x * GS_BLOCK + threadIdx.x; c < n; c += (long long)GS_BLOCK * GS_GRID) {
    const int i = (int)(c % m.ni);
    const long long r = c / m.ni;
    const int j = (int)(r % m.nj);
    const long long z = r / m.nj;  // t * npz + k
    const int kk = (int)(z % m.npz), t = (int)(z / m.npz);
    const long long cell = (long long)(j + m.joff) * m.nid + (i + m.ioff);
    const double qv = (double)q[(((long long)t * gridDim.y + iq) * m.npz + kk) * m.s2 + cell];
    ks_add(k, qv * delp[z * m.s2 + cell] * area[t * m.gs2 + cell]);
  }
  sh[threadIdx.x] = k;
  __syncthreads();
  for (int s = GS_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) ks_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)iq * GS_GRID + blockIdx.x] = sh[0];
}

__global__ void k_mass_final(const Ks *part, double *out) {
  if (threadIdx.x != 0) return;
  Ks k{0.0, 0.0};
  for (int b = 0; b < GS_GRID; b++) ks_merge(k, part[(long long)blockIdx.x * GS_GRID + b]);
  out[blockIdx.x] = k.s + k.c;
}

}  // namespace

static int global_mass(Ctx &c, const void *q, int nq, const double *delp, int elem, double *out, hipStream_t s) {
  const Dims &d = c.d;
  MassDims m;
  m.ni = d.ie - d.is + 1;
  m.nj = d.je - d.js + 1;
  m.nid = d.nid;
  m.ioff = d.is - d.isd;
  m.joff = d.js - d.jsd;
  m.npz = c.npz;
  m.nt = c.ntiles;
  m.s2 = d.s2;
  m.gs2 = c.grid64.s2;
  const size_t pb = ((sizeof(Ks) * GS_GRID * nq + 255) & ~(size_t)255);
  char *base = (char *)ctx_stage(c, pb + sizeof(double) * nq);
  if (!base) return 1;
  Ks *part = (Ks *)base;
  double *dout = (double *)(base + pb);
  const dim3 g(GS_GRID, (unsigned)nq);
  if (elem == 8)
    hipLaunchKernelGGL(k_mass_part<double>, g, dim3(GS_BLOCK), 0, s, (const double *)q, delp, c.grid64.area, m, part);
  else
    hipLaunchKernelGGL(k_mass_part<float>, g, dim3(GS_BLOCK), 0, s, (const float *)q, delp, c.grid64.area, m, part);
  FV3_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mass_final, dim3((unsigned)nq), dim3(64), 0, s, part, dout);
  FV3_HIP(hipGetLastError());
  FV3_HIP(hipMemcpyAsync(out, dout, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace fv3

using namespace fv3;

extern "C" int fv3_global_mass(fv3_ctx *h, const void *q, int nq, const double *delp, int elem_bytes, double *out,
                               void *stream) {
  if (!h || !q || !delp || !out) return set_error("fv3_global_mass: NULL argument");
  if (nq < 1 || nq > 65535) return set_error("nq must be in 1..65535 (got %d)", nq);
  if (elem_bytes != 4 && elem_bytes != 8) return set_error("element size must be 4 or 8 bytes (got %d)", elem_bytes);
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.grid_set) return set_error("grid metrics not set (fv3_ctx_set_grid)");
  return global_mass(c, q, nq, delp, elem_bytes, out, stream ? (hipStream_t)stream : c.stream);
}
