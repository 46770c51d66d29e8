// fused.h -- the fused transport kernel launchers and the step driver of the
// device-resident forms.
#pragma once
#include "fused_tp.h"
#include "fv3_internal.h"

namespace fv3 {

// streams of the kernel kinds (lean, edge rows, edge strips / left strips,
// right strips)
struct KindStreams {
  hipStream_t k[4];
};
int launch_fused_f64(const FusedArgs<double> &a, const KindStreams &s);
int launch_fused_f32(const FusedArgs<float> &a, int cfg, const KindStreams &s);
// tracers per round (waves x tracers per wave) of the lean kind of cfg
int fused32_round(int cfg);

// per-level sub-cycle counts on the device (cmax, nsplit; no scaling)
template <typename T>
int device_nsplit(Ctx &c, const T *cx, const T *cy, bool per_level, int q_split, double *cmax,
                  int *nsplit, T *frac);

// Device scratch handed to a step by its caller instead of the context's
// arenas (the no-allocation entry point); base == nullptr: the context's.
struct StepArena {
  char *base;
  size_t bytes;
};

int fused_step(Ctx &c, bool fp32, int nq, const void *q_in, void *q_out, void *pt,
               const double *dp1, double *delp_out, const void *cx, const void *cy,
               const void *mfx, const void *mfy, const double *ak, const double *bk, double ptop,
               int hord, int kord_tr, int kord_tm, double t_min, int nsplit_cap,
               const StepArena &arena = StepArena{nullptr, 0});

// the FP32 column-shared remap alone: q_src -> q_dst (distinct buffers), pt
// in place, delp_out the Eulerian thickness from the Lagrangian dp1
int remap_only_f32(Ctx &c, int nq, const float *q_src, float *q_dst, float *pt, const double *dp1,
                   double *delp_out, const double *ak, const double *bk, double ptop, int kord_tr,
                   int kord_tm, double t_min);

// tangent-linear / adjoint of the FP64 advection + tracer remap step
int step_tl(Ctx &c, bool per_level, int q_split, int nq, const double *q0, const double *dq0,
            double *q1, double *dq1, const double *dp1, double *delp_out, const double *cx,
            const double *cy, const double *mfx, const double *mfy, const double *ak,
            const double *bk, double ptop, int hord, int kord_tr);
int step_ad(Ctx &c, bool per_level, int q_split, int nq, const double *q0, const double *qbar1,
            double *qbar0, const double *dp1, const double *cx, const double *cy,
            const double *mfx, const double *mfy, const double *ak, const double *bk,
            double ptop, int hord, int kord_tr);

}  // namespace fv3
