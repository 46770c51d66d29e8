#include "fused.h"
namespace fv3 {
int fused_step(Ctx &c, bool fp32, int nq, const void *q_in, void *q_out, void *pt,
               const double *dp1, double *delp_out, const void *cx, const void *cy,
               const void *mfx, const void *mfy, const double *ak, const double *bk, double ptop,
               int hord, int kord_tr, int kord_tm, double t_min, int nsplit_cap) {
  return set_error("fused forms not built");
}
}  // namespace fv3
