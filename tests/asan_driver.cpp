// Host AddressSanitizer / UBSan driver (built and run by tests/test_host_asan.py):
// exercises the host C code the GPU tests rely on -- the synthetic-case
// generator (fv3atm_amd/csrc/test_cases.c), the FP64 oracle (oracle/*.c) and
// the netCDF classic reader/writer (fv3atm_amd/csrc/ncfile.h) -- on small
// exactly-sized heap arrays, so that any access outside them aborts the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../fv3atm_amd/csrc/ncfile.h"
#include "../oracle/fv3_oracle.h"

extern "C" {
int fv3c_halo_map(int n, int tile, long long *map);
int fv3c_grid(int n, int tile, double radius, double *area, double *rarea, double *dx, double *dy, double *dxa,
              double *dya, double *sin_sg, double *del6_u, double *del6_v, double *xyz_c, double *xyz_n);
int fv3c_winds(int n, int tile, int npz, int kind, const double *par, double radius, const double *speed_k,
               double dt, const double *dp_ref, const double *dxa, const double *dy, const double *dya,
               const double *dx, const double *sin_sg, double *cx, double *cy, double *mfx, double *mfy);
int fv3c_tracer(int n, int tile, int kind, const double *par, double *q2d);
int fv3c_delp(int n, int tile, int npz, int kind, const double *amp, uint64_t seed, const double *dp_ref,
              double ps_minus_ptop, double *delp);
int fv3c_cell_centers(int n, int tile, double *xyz);
}

#define CHECK(x)                                             \
  do {                                                       \
    if (!(x)) {                                              \
      fprintf(stderr, "check failed: %s (line %d)\n", #x, __LINE__); \
      return 1;                                              \
    }                                                        \
  } while (0)

using V = std::vector<double>;

struct Tile {
  V area, rarea, dx, dy, dxa, dya, sin_sg, d6u, d6v, xyz_c, xyz_n;
  fv3o_grid g;
};

static int one_case(int n, int npz, int nq) {
  const int NG = 3, nd = n + 2 * NG;
  const size_t s2 = (size_t)nd * nd;
  std::vector<Tile> tiles(6);
  fv3o_bounds bd{1, n, 1, n, 1 - NG, n + NG, 1 - NG, n + NG};
  V dp_ref(npz), speed(npz);
  for (int k = 0; k < npz; k++) {
    dp_ref[k] = 1000.0 + 50.0 * k;
    speed[k] = 20.0 + k;
  }
  const double par[2] = {0.3, 1.0};
  const size_t sx = (size_t)(n + 1) * nd, sy = (size_t)nd * (n + 1), smx = (size_t)(n + 1) * n;
  V cx(sx * npz * 6), cy(sy * npz * 6), mfx(smx * npz * 6), mfy(smx * npz * 6);
  V q(s2 * npz * nq * 6), dp1(s2 * npz * 6);
  for (int t = 0; t < 6; t++) {
    Tile &T = tiles[t];
    T.area.resize(s2); T.rarea.resize(s2); T.dx.resize((size_t)nd * (nd + 1)); T.dy.resize((size_t)(nd + 1) * nd);
    T.dxa.resize(s2); T.dya.resize(s2); T.sin_sg.resize(9 * s2); T.d6u.resize((size_t)nd * (nd + 1));
    T.d6v.resize((size_t)(nd + 1) * nd); T.xyz_c.resize(3 * s2); T.xyz_n.resize(3 * (size_t)(nd + 1) * (nd + 1));
    CHECK(fv3c_grid(n, t + 1, 6.371e6, T.area.data(), T.rarea.data(), T.dx.data(), T.dy.data(), T.dxa.data(),
                    T.dya.data(), T.sin_sg.data(), T.d6u.data(), T.d6v.data(), T.xyz_c.data(), T.xyz_n.data()) == 0);
    std::vector<long long> map(s2);
    CHECK(fv3c_halo_map(n, t + 1, map.data()) == 0);
    V cc(3 * s2);
    CHECK(fv3c_cell_centers(n, t + 1, cc.data()) == 0);
    CHECK(fv3c_winds(n, t + 1, npz, 0, par, 6.371e6, speed.data(), 600.0, dp_ref.data(), T.dxa.data(), T.dy.data(),
                     T.dya.data(), T.dx.data(), T.sin_sg.data(), cx.data() + t * sx * npz, cy.data() + t * sy * npz,
                     mfx.data() + t * smx * npz, mfy.data() + t * smx * npz) == 0);
    const double amp[2] = {0.05, 0.2};
    CHECK(fv3c_delp(n, t + 1, npz, 1, amp, 7, dp_ref.data(), 1.0e5 - 300.0, dp1.data() + t * s2 * npz) == 0);
    for (int iq = 0; iq < nq; iq++) {
      const double tp[4] = {0.5, 0.2, 0.8, 0.1};
      CHECK(fv3c_tracer(n, t + 1, iq % 3, tp, q.data() + ((size_t)t * nq + iq) * npz * s2) == 0);
      for (int k = 1; k < npz; k++)
        for (size_t m = 0; m < s2; m++)
          q[(((size_t)t * nq + iq) * npz + k) * s2 + m] = q[((size_t)t * nq + iq) * npz * s2 + m];
    }
    T.g = fv3o_grid{T.area.data(), T.rarea.data(), T.dx.data(), T.dy.data(), T.dxa.data(), T.dya.data(),
                    T.sin_sg.data(), T.d6u.data(), T.d6v.data(), 1.0e8, 1, 1, 1, 1, 0, 0};
  }
  // tracer_2d_1L and tracer_2d (with deln_flux damping) on tile 1
  int nsplit[64];
  {
    V a = q, b = dp1, c1 = cx, c2 = cy, m1 = mfx, m2 = mfy;
    CHECK(fv3o_tracer_2d_1L(a.data(), b.data(), m1.data(), m2.data(), c1.data(), c2.data(), &tiles[0].g, &bd, 1,
                            n + 1, n + 1, npz, nq, 8, 0, 600.0, 0, 0, 0.0, 1.0, nullptr, nullptr, nsplit) == 0);
  }
  for (int nord = 0; nord <= 2; nord++) {
    V a = q, b = dp1, c1 = cx, c2 = cy, m1 = mfx, m2 = mfy;
    CHECK(fv3o_tracer_2d(a.data(), b.data(), m1.data(), m2.data(), c1.data(), c2.data(), &tiles[0].g, &bd, 1,
                         n + 1, n + 1, npz, nq, 13, 2, 600.0, 1, nord, 0.1, 1.0, nullptr, nullptr, nsplit) == 0);
  }
  // tracer/pt remap on tile 1
  {
    V ak(npz + 1), bk(npz + 1), delp(dp1.begin(), dp1.begin() + s2 * npz), a(q.begin(), q.begin() + s2 * npz * nq);
    V pt(s2 * npz, 250.0);
    for (int k = 0; k <= npz; k++) {
      bk[k] = (double)k / npz;
      ak[k] = 300.0 * (1.0 - bk[k]);
    }
    std::vector<int> kord(nq, 9);
    CHECK(fv3o_lagrangian_to_eulerian_tracers(npz, nq, &bd, ak.data(), bk.data(), 300.0, delp.data(), a.data(),
                                              pt.data(), kord.data(), 10, 184.0, 1) == 0);
  }
  // dyn_core inputs on tile 1, two sub-steps
  {
    V uc(2 * sx * npz, 12.0), vc(2 * sy * npz, -7.0), delp(dp1.begin(), dp1.begin() + s2 * npz), d1(s2 * npz);
    V ox(sx * npz), oy(sy * npz), omx(smx * npz), omy(smx * npz);
    CHECK(fv3o_dyn_core_fluxes(2, 300.0, 8, uc.data(), vc.data(), delp.data(), d1.data(), ox.data(), oy.data(),
                               omx.data(), omy.data(), &tiles[0].g, &bd, 1, n + 1, n + 1, npz, 1.0, nullptr,
                               nullptr) == 0);
  }
  return 0;
}

static int nc_round_trip() {
  using namespace fv3::nc;
  NcFile f;
  f.numrecs = 1;
  const int dx = f.dim("x", 5), dt = f.dim("T", 0);
  Var &a = f.var("a", NC_DOUBLE, {dt, dx});
  a.atts.push_back(att_text("units", "m"));
  f.var("b", NC_FLOAT, {dt, dx}).atts.push_back(att_int("n", 3));
  f.gatts.push_back(att_doubles("ak", std::vector<double>{1.0, 2.0}.data(), 2));
  f.layout();
  std::vector<unsigned char> h = f.header();
  NcFile g;
  Reader r;
  CHECK(r.parse(h, g));
  CHECK(g.vars.size() == 2 && g.vars[0].begin == f.vars[0].begin && g.vars[1].begin == f.vars[1].begin);
  CHECK(g.recsize == f.recsize && g.gatts.size() == 1);
  // every truncation of the header must be refused without reading past it
  for (size_t cut = 0; cut < h.size(); cut++) {
    std::vector<unsigned char> t(h.begin(), h.begin() + cut);
    NcFile z;
    Reader rz;
    CHECK(!rz.parse(t, z));
  }
  return 0;
}

int main() {
  if (one_case(8, 4, 2) || one_case(12, 6, 7)) return 1;
  if (nc_round_trip()) return 1;
  printf("asan driver: all checks passed\n");
  return 0;
}
