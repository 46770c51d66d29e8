// fv3_internal.h -- host-side internal interface of libfv3atm_amd (not part
// of the C-ABI): domain dimensions, device-resident grid metrics, context.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "fv3_device.h"

namespace fv3 {

// Fortran bounds of one tile and the derived extents of every array shape on
// the path (fv_grid_bounds_type, moving_nest/fv_moving_nest.F90:177-185).
struct Dims {
  int is, ie, js, je, isd, ied, jsd, jed, npx, npy;
  int nid, njd;  // (isd:ied, jsd:jed)
  int nix;       // is:ie+1
  int niy;       // is:ie
  int nj;        // js:je
  int njy;       // js:je+1
  long long s2;  // q / dp1 / metric slab
  long long sx;  // cx, xfx   (is:ie+1, jsd:jed)
  long long sy;  // cy, yfx   (isd:ied, js:je+1)
  long long smx; // mfx       (is:ie+1, js:je)
  long long smy; // mfy       (is:ie, js:je+1)
  long long srx; // ra_x      (is:ie, jsd:jed)
  long long sry; // ra_y      (isd:ied, js:je)
  long long sdp; // dp2       (is:ie, js:je)
  unsigned corners;  // bit0 sw, bit1 se, bit2 ne, bit3 nw
  bool edge;         // !bounded_domain && grid_type < 3
  bool bounded;
};

Dims make_dims(int is, int ie, int js, int je, int isd, int ied, int jsd, int jed, int npx,
               int npy, unsigned corners, bool bounded, int grid_type);

// Grid metrics of all tiles, stacked tile-outermost, in FP64 and FP32.
template <typename T>
struct GridDev {
  const T *area, *rarea, *dx, *dy, *dxa, *dya, *sin_sg, *del6_u, *del6_v;
  long long s2, sdx, sdy;  // tile strides of (isd:ied,jsd:jed), dx, dy
};

// FP32 fused launch configurations (waves per workgroup x tracers per wave
// per kernel kind); the default: 8 x 2 for the interior kernel (two
// workgroups, 16 waves per CU) and 4 x 1 at three workgroups per CU for the
// edge kernels
constexpr int FUSED32_NCFG = 15;
constexpr int FUSED32_DEFAULT_CFG = 12;
// objects the FP32 configurations are compiled into (cfg % FUSED32_NPART)
constexpr int FUSED32_NPART = 6;

struct Ctx {
  Dims d;
  int npz = 0, ntiles = 1;
  double da_min = 0.0;
  int dev = 0;
  hipStream_t stream = nullptr;
  // device metrics
  double *g64 = nullptr;
  float *g32 = nullptr;
  GridDev<double> grid64{};
  GridDev<float> grid32{};
  // FP32 upwind area-flux factors of tile 1: gxp, gxm (isd:ied+1, jsd:jed),
  // gyp, gym (isd:ied, jsd:jed+1)
  float *gfac32 = nullptr;
  bool grid_set = false;
  // six-tile edge-halo table (dst, src) flat offsets within a tile-stacked slab set
  long long *halo_dst = nullptr, *halo_src = nullptr;
  int halo_n = 0;
  // scratch arena (grown on demand, reused)
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  void *stage = nullptr;  // staging area for host-pointer entry points
  size_t stage_bytes = 0;
  std::mutex mu;
  // optional phase timing of fv3_step_dev (hipEvents on the step's stream)
  bool timing = false;
  // fused-kernel launch shape: configuration id and interior rows per block
  int fused_cfg = FUSED32_DEFAULT_CFG, fused_h = 0;
  // FP32 step remap: 0 column-shared fast kernel, 1 one thread per column
  int remap_mode = 0;
  // fused step's sub-cycling: per-level nsplit(k) (tracer_2d_1L) or one
  // nsplit for every level (tracer_2d), forced to q_split when > 0
  bool step_per_level = true;
  int step_q_split = 0;
  // graph-safe steps: no host synchronisation, no memset, no allocation; the
  // nsplit_cap check lands in step_flag for fv3_step_status
  bool graph_safe = false;
  int *step_flag = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // the fused transport's edge kinds run beside the lean kind on three
  // auxiliary streams (default priority), forked from and joined to the
  // step's stream with events every pass; a fourth runs the remap's column
  // prep beside the transport
  bool concurrent_kinds = true;
  hipStream_t aux[4] = {nullptr, nullptr, nullptr, nullptr};  // [3]: the remap's column prep
  hipEvent_t fork_ev = nullptr, join_ev[3] = {nullptr, nullptr, nullptr}, prep_ev = nullptr;
  // FP32 step pipelining: the tracers are transported in chunks of this many
  // (-1: one round of the launch shape, 0: all at once, the default), and
  // each chunk's remap runs on aux[3] beside the next chunk's transport
  int pipe_chunk = 0;
  hipEvent_t chunk_ev = nullptr;
  int last_passes = 0;
  bool ev_valid = false;
  // six-face halo tables of the halo service (scalar, A, C, D grids), built
  // on first use
  void *halo_tab[4] = {nullptr, nullptr, nullptr, nullptr};
  int halo_tab_n[4] = {0, 0, 0, 0};
  // the edge-halo table grouped by source cell (the adjoint's transposed
  // refresh), built on first use: distinct sources, offsets, halo cells
  void *halo_adj = nullptr;
  int halo_adj_nsrc = 0;
  // the fused step without its remap (the host-array transport entry points)
  // and the sub-cycle counts of the last step that derived them on the host
  bool transport_only = false;
  std::vector<int> last_nsplit;
  // Lagrangian_to_Eulerian buffers that outlive the column-kernel calls
  void *l2e_buf = nullptr;
  size_t l2e_bytes = 0;
};

// error reporting (thread-local last error string)
int set_error(const char *fmt, ...);
#define FV3_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::fv3::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
  } while (0)

void *ctx_scratch(Ctx &c, size_t bytes);
void *ctx_stage(Ctx &c, size_t bytes);

// the no-allocation step's device scratch: looks up its address on device
// dev once (the first context created on that device does it)
int noalloc_setup(int dev);

template <typename T>
const GridDev<T> &grid_of(const Ctx &c);

// ---- plain (staged) transport, FP64 or FP32 --------------------------------
// All pointers are device pointers with the Fortran layouts of tracer_2d:
// q (isd:ied,jsd:jed,npz,nq), dp1 (isd:ied,jsd:jed,npz), cx (is:ie+1,jsd:jed,npz),
// cy (isd:ied,js:je+1,npz), mfx (is:ie+1,js:je,npz), mfy (is:ie,js:je+1,npz),
// each repeated per tile (tile outermost).  In-place semantics as upstream.
template <typename T>
int tracer_2d_plain(Ctx &c, bool per_level, T *q, T *dp1, T *mfx, T *mfy, T *cx, T *cy, int nq,
                    int hord, int q_split, T dt, int id_divg, int nord_tr, T trdm, T lim_fac,
                    int *nsplit_host);

template <typename T>
int fv_tp_2d_plain(Ctx &c, T *q, const T *crx, const T *cry, int hord, T *fx, T *fy,
                   const T *xfx, const T *yfx, const T *ra_x, const T *ra_y, T lim_fac,
                   const T *mfx, const T *mfy, const T *mass, int nord, T damp_c, bool damp);

// six-tile edge-halo refresh of nslab stacked slabs per tile
template <typename T>
int halo_fill(Ctx &c, T *base, long long tile_stride, const int *slab_ids, int nslab_list,
              int nslab_dense);

// ---- remap ------------------------------------------------------------------
template <typename T>
int remap_tracers_plain(Ctx &c, int km, int nq, const double *ak, const double *bk, double ptop,
                        double *delp, T *q, T *pt, const int *kord_tr, int kord_tm, double t_min,
                        int fill);

// cs_profile over im independent columns: a1 (km, im) in, out (4, km, im)
int cs_profile_rows(Ctx &c, const double *a1, const double *dp, const double *qs, double *out,
                    int im, int km, int iv, int kord);

}  // namespace fv3

// the opaque handle of the C-ABI (include/fv3atm_amd.h)
struct fv3_ctx {
  fv3::Ctx c;
};
