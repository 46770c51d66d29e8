/*
 * test_cases.c -- deterministic synthetic inputs for the tracer transport and
 * remap path: an analytic gnomonic equal-distance cubed-sphere grid with
 * 3-cell halos filled from the neighbouring faces, stream-function winds
 * (discretely non-divergent face fluxes), tracer patterns and perturbed layer
 * thicknesses.  Plays the role of the dycore's grid generator and
 * test_cases module, which supply these arrays to tracer_2d and
 * Lagrangian_to_Eulerian upstream.  Host C, FP64.
 *
 * Geometry.  Face f has a unit normal n_f and tangent axes (ex_f, ey_f); the
 * cube surface is an integer lattice with faces at |coordinate| = N, cell
 * centres at odd and corners at even offsets.  A lattice point is mapped to
 * the sphere by x -> G(x) per coordinate, G(v) = sqrt(2) tan(alpha v/N),
 * alpha = asin(1/sqrt 3), which places corners equally spaced along the cube
 * edges (gnomonic equal-distance).  Halo indices that leave a face are folded
 * across the cube edge onto the neighbouring face, so a halo cell IS the
 * neighbour's cell, bit for bit.  The 3x3 corner blocks, which exist on no
 * face, take the face's own gnomonic extension.
 *
 * Face frames follow the FMS cubic mosaic orientation (tile 1 east -> tile 2
 * west, tile 1 north -> tile 3 west rotated, ...; tiles 3 and 6 at the poles).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NG 3

static const int FACE_N[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
static const int FACE_X[6][3] = {{0, 1, 0}, {-1, 0, 0}, {-1, 0, 0}, {0, 0, -1}, {0, 0, -1}, {0, 1, 0}};
static const int FACE_Y[6][3] = {{0, 0, 1}, {0, 0, 1}, {0, -1, 0}, {0, -1, 0}, {1, 0, 0}, {1, 0, 0}};

static inline int dot3i(const int *a, const int *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Lattice point of face-index (u, v) offsets: L = N n + u ex + v ey. */
static void lattice(int n, int tile, int u, int v, int L[3]) {
  const int f = tile - 1;
  for (int c = 0; c < 3; c++) L[c] = n * FACE_N[f][c] + u * FACE_X[f][c] + v * FACE_Y[f][c];
}

/* Fold a lattice point that overshoots one face edge onto the neighbour face.
 * Returns the number of overshooting coordinates (0, 1, or 2 for the corner
 * blocks, which are left unfolded).  On return for 0/1, *face is the tile
 * (1..6) the point lies on. */
static int fold(int n, int L[3], int tile, int *face) {
  const int f = tile - 1;
  int over = 0, oc = -1;
  for (int c = 0; c < 3; c++)
    if (abs(L[c]) > n) {
      over++;
      oc = c;
    }
  if (over == 1) {
    int d = abs(L[oc]) - n;
    int nc = FACE_N[f][0] != 0 ? 0 : (FACE_N[f][1] != 0 ? 1 : 2);
    int s = L[oc] > 0 ? 1 : -1;
    L[oc] = s * n;
    L[nc] = (L[nc] > 0 ? 1 : -1) * (n - d);
    int nrm[3] = {0, 0, 0};
    nrm[oc] = s;
    for (int t = 0; t < 6; t++)
      if (dot3i(FACE_N[t], nrm) == 1) *face = t + 1;
  } else if (over == 0) {
    *face = tile;
  } else {
    *face = tile;
  }
  return over;
}

/* G(v) = sqrt(2) tan(alpha v / N), exact +-1 on the cube edges, odd in v. */
static double gmap(int n, int v) {
  int a = abs(v);
  double g = a == n ? 1.0 : sqrt(2.0) * tan(asin(1.0 / sqrt(3.0)) * (double)a / (double)n);
  return v < 0 ? -g : g;
}

static void lattice_to_unit(int n, const int L[3], double p[3]) {
  double x = gmap(n, L[0]), y = gmap(n, L[1]), z = gmap(n, L[2]);
  double r = sqrt(x * x + y * y + z * z);
  p[0] = x / r;
  p[1] = y / r;
  p[2] = z / r;
}

/* Unit vector of corner (node) (i, j), 1-based, of a tile, halo-folded. */
static void node_xyz(int n, int tile, int i, int j, double p[3]) {
  int L[3], face;
  lattice(n, tile, 2 * i - 2 - n, 2 * j - 2 - n, L);
  fold(n, L, tile, &face);
  lattice_to_unit(n, L, p);
}

/* Unit vector of the centre of cell (i, j), halo-folded. */
static void cell_xyz(int n, int tile, int i, int j, double p[3]) {
  int L[3], face;
  lattice(n, tile, 2 * i - 1 - n, 2 * j - 1 - n, L);
  fold(n, L, tile, &face);
  lattice_to_unit(n, L, p);
}

/* Source (tile, i, j) of halo cell (i, j): returns 1 when it lies on a
 * neighbouring face, 0 for interior cells (source = itself), -1 for the
 * corner blocks. */
int fv3c_halo_source(int n, int tile, int i, int j, int *st, int *si, int *sj) {
  int L[3], face;
  lattice(n, tile, 2 * i - 1 - n, 2 * j - 1 - n, L);
  int over = fold(n, L, tile, &face);
  if (over == 2) return -1;
  const int f = face - 1;
  *st = face;
  *si = (dot3i(L, FACE_X[f]) + n + 1) / 2;
  *sj = (dot3i(L, FACE_Y[f]) + n + 1) / 2;
  return over;
}

/* map(isd:ied, jsd:jed) of a tile: for every edge-halo cell the flat index
 * (t-1)*S + (sj-jsd)*ni + (si-isd) of its source cell in a stack of six
 * (isd:ied, jsd:jed) slabs of size S; -1 for interior and corner cells. */
int fv3c_halo_map(int n, int tile, long long *map) {
  const int isd = 1 - NG, ied = n + NG, ni = ied - isd + 1;
  const long long S = (long long)ni * ni;
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied; i++) {
      int st, si, sj;
      int r = fv3c_halo_source(n, tile, i, j, &st, &si, &sj);
      map[(size_t)(j - isd) * ni + (i - isd)] =
          r == 1 ? (st - 1) * S + (long long)(sj - isd) * ni + (si - isd) : -1;
    }
  return 0;
}

static void normalize3(double p[3]) {
  double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  p[0] /= r;
  p[1] /= r;
  p[2] /= r;
}

static void mid3(const double a[3], const double b[3], double m[3]) {
  m[0] = a[0] + b[0];
  m[1] = a[1] + b[1];
  m[2] = a[2] + b[2];
  normalize3(m);
}

/* great-circle distance on the unit sphere, symmetric in (a, b) */
static double gdist(const double a[3], const double b[3]) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return 2.0 * asin(0.5 * sqrt(dx * dx + dy * dy + dz * dz));
}

/* unit tangent at m pointing toward p */
static void tangent(const double m[3], const double p[3], double t[3]) {
  double d = m[0] * p[0] + m[1] * p[1] + m[2] * p[2];
  t[0] = p[0] - d * m[0];
  t[1] = p[1] - d * m[1];
  t[2] = p[2] - d * m[2];
  normalize3(t);
}

static double sin_between(const double a[3], const double b[3]) {
  double c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  double s = sqrt(fmax(0.0, 1.0 - c * c));
  return s > 1.0 ? 1.0 : s;
}

static double tri_area(const double a[3], const double b[3], const double c[3]) {
  double bxc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  double num = fabs(a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2]);
  double den = 1.0 + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) + (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) +
               (c[0] * a[0] + c[1] * a[1] + c[2] * a[2]);
  return 2.0 * atan2(num, den);
}

/* Grid metrics of one tile, all arrays with a 3-cell halo (Fortran layout):
 * area, rarea, dxa, dya (isd:ied, jsd:jed); dx, del6_u (isd:ied, jsd:jed+1);
 * dy, del6_v (isd:ied+1, jsd:jed); sin_sg (isd:ied, jsd:jed, 9);
 * optional xyz_c (3, isd:ied, jsd:jed) cell centres and xyz_n (3, isd:ied+1,
 * jsd:jed+1) corners. */
int fv3c_grid(int n, int tile, double radius, double *area, double *rarea, double *dx,
              double *dy, double *dxa, double *dya, double *sin_sg, double *del6_u,
              double *del6_v, double *xyz_c, double *xyz_n) {
  if (n < 8 || tile < 1 || tile > 6) return 1;
  const int isd = 1 - NG, ied = n + NG, ni = ied - isd + 1, nn = ni + 1;
  const size_t s2 = (size_t)ni * ni;
  double *pn = (double *)malloc(sizeof(double) * 3 * nn * nn);
  double *pc = (double *)malloc(sizeof(double) * 3 * s2);
#define PN(i, j) (pn + 3 * ((size_t)((j) - isd) * nn + ((i) - isd)))
#define PC(i, j) (pc + 3 * ((size_t)((j) - isd) * ni + ((i) - isd)))
#define C2(a, i, j) (a)[(size_t)((j) - isd) * ni + ((i) - isd)]
  for (int j = isd; j <= ied + 1; j++)
    for (int i = isd; i <= ied + 1; i++) node_xyz(n, tile, i, j, PN(i, j));
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied; i++) cell_xyz(n, tile, i, j, PC(i, j));
  const double R2 = radius * radius;
  for (int j = isd; j <= ied + 1; j++)
    for (int i = isd; i <= ied; i++)
      dx[(size_t)(j - isd) * ni + (i - isd)] = radius * gdist(PN(i, j), PN(i + 1, j));
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied + 1; i++)
      dy[(size_t)(j - isd) * nn + (i - isd)] = radius * gdist(PN(i, j), PN(i, j + 1));
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied; i++) {
      const double *p1 = PN(i, j), *p2 = PN(i + 1, j), *p3 = PN(i + 1, j + 1), *p4 = PN(i, j + 1);
      double w[3], e[3], s[3], no[3], t1[3], t2[3];
      mid3(p1, p4, w);
      mid3(p2, p3, e);
      mid3(p1, p2, s);
      mid3(p4, p3, no);
      C2(dxa, i, j) = radius * gdist(w, e);
      C2(dya, i, j) = radius * gdist(s, no);
      double a = R2 * (tri_area(p1, p2, p3) + tri_area(p1, p3, p4));
      C2(area, i, j) = a;
      C2(rarea, i, j) = 1.0 / a;
      double *sg = sin_sg + (size_t)(j - isd) * ni + (i - isd);
      tangent(w, e, t1);
      tangent(w, p4, t2);
      sg[0 * s2] = sin_between(t1, t2);
      tangent(s, no, t1);
      tangent(s, p2, t2);
      sg[1 * s2] = sin_between(t1, t2);
      tangent(e, w, t1);
      tangent(e, p3, t2);
      sg[2 * s2] = sin_between(t1, t2);
      tangent(no, s, t1);
      tangent(no, p3, t2);
      sg[3 * s2] = sin_between(t1, t2);
      const double *c = PC(i, j);
      double we[3] = {e[0] - w[0], e[1] - w[1], e[2] - w[2]};
      double sn[3] = {no[0] - s[0], no[1] - s[1], no[2] - s[2]};
      double d1 = we[0] * c[0] + we[1] * c[1] + we[2] * c[2];
      double d2 = sn[0] * c[0] + sn[1] * c[1] + sn[2] * c[2];
      for (int t = 0; t < 3; t++) {
        we[t] -= d1 * c[t];
        sn[t] -= d2 * c[t];
      }
      normalize3(we);
      normalize3(sn);
      sg[4 * s2] = sin_between(we, sn);
      tangent(p1, p2, t1);
      tangent(p1, p4, t2);
      sg[5 * s2] = sin_between(t1, t2);
      tangent(p2, p3, t1);
      tangent(p2, p1, t2);
      sg[6 * s2] = sin_between(t1, t2);
      tangent(p3, p4, t1);
      tangent(p3, p2, t2);
      sg[7 * s2] = sin_between(t1, t2);
      tangent(p4, p1, t1);
      tangent(p4, p3, t2);
      sg[8 * s2] = sin_between(t1, t2);
    }
  /* The 3x3 corner blocks exist on no face.  Like the q values copy_corners
   * puts there, their cell metrics are the rotated copies of the adjacent
   * halo strip: x-sweep quantities (dxa, sides 1/3, area) from the cell the
   * dir=1 rotation reads, y-sweep quantities (dya, sides 2/4) from the dir=2
   * one. */
  {
    const int npx = n + 1, npy = n + 1;
    double *cp_dxa = (double *)malloc(sizeof(double) * s2);
    double *cp_dya = (double *)malloc(sizeof(double) * s2);
    double *cp_area = (double *)malloc(sizeof(double) * s2);
    double *cp_sg = (double *)malloc(sizeof(double) * s2 * 9);
    memcpy(cp_dxa, dxa, sizeof(double) * s2);
    memcpy(cp_dya, dya, sizeof(double) * s2);
    memcpy(cp_area, area, sizeof(double) * s2);
    memcpy(cp_sg, sin_sg, sizeof(double) * s2 * 9);
    for (int c = 0; c < 4; c++) { /* 0 SW, 1 SE, 2 NE, 3 NW */
      const int ilo = (c == 0 || c == 3) ? 1 - NG : npx;
      const int jlo = (c == 0 || c == 1) ? 1 - NG : npy;
      for (int j = jlo; j < jlo + NG; j++)
        for (int i = ilo; i < ilo + NG; i++) {
          int a1, b1, a2, b2; /* dir=1 and dir=2 sources */
          int e1, w1, n2, s2side; /* sin_sg side of the source (0-based) */
          if (c == 0) { a1 = j; b1 = 1 - i; a2 = 1 - j; b2 = i; }
          else if (c == 1) { a1 = npy - j; b1 = i - npx + 1; a2 = npy + j - 1; b2 = npx - i; }
          else if (c == 2) { a1 = j; b1 = 2 * npx - 1 - i; a2 = 2 * npy - 1 - j; b2 = i; }
          else { a1 = npy - j; b1 = i - 1 + npx; a2 = j + 1 - npx; b2 = npy - i; }
          if (c == 0 || c == 2) { e1 = 1; w1 = 3; } else { e1 = 3; w1 = 1; }
          if (c == 0 || c == 2) { n2 = 0; s2side = 2; } else { n2 = 2; s2side = 0; }
          const size_t d = (size_t)(j - isd) * ni + (i - isd);
          const size_t o1 = (size_t)(b1 - isd) * ni + (a1 - isd);
          const size_t o2 = (size_t)(b2 - isd) * ni + (a2 - isd);
          dxa[d] = cp_dya[o1];
          dya[d] = cp_dxa[o2];
          area[d] = cp_area[o1];
          rarea[d] = 1.0 / cp_area[o1];
          sin_sg[0 * s2 + d] = cp_sg[w1 * s2 + o1];
          sin_sg[2 * s2 + d] = cp_sg[e1 * s2 + o1];
          sin_sg[1 * s2 + d] = cp_sg[s2side * s2 + o2];
          sin_sg[3 * s2 + d] = cp_sg[n2 * s2 + o2];
          sin_sg[4 * s2 + d] = cp_sg[4 * s2 + o1];
        }
    }
    free(cp_dxa);
    free(cp_dya);
    free(cp_area);
    free(cp_sg);
  }
  /* del6_v(i,j) = 0.5 (sin_sg(i-1,j,3) + sin_sg(i,j,1)) dy(i,j) / dxc(i,j);
   * del6_u(i,j) = 0.5 (sin_sg(i,j-1,4) + sin_sg(i,j,2)) dx(i,j) / dyc(i,j). */
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied + 1; i++) {
      double v = 0.0;
      if (i > isd && i <= ied) {
        double dxc = radius * gdist(PC(i - 1, j), PC(i, j));
        v = 0.5 * (sin_sg[2 * s2 + (size_t)(j - isd) * ni + (i - 1 - isd)] +
                   sin_sg[0 * s2 + (size_t)(j - isd) * ni + (i - isd)]) *
            dy[(size_t)(j - isd) * nn + (i - isd)] / dxc;
      }
      del6_v[(size_t)(j - isd) * nn + (i - isd)] = v;
    }
  for (int j = isd; j <= ied + 1; j++)
    for (int i = isd; i <= ied; i++) {
      double v = 0.0;
      if (j > isd && j <= ied) {
        double dyc = radius * gdist(PC(i, j - 1), PC(i, j));
        v = 0.5 * (sin_sg[3 * s2 + (size_t)(j - 1 - isd) * ni + (i - isd)] +
                   sin_sg[1 * s2 + (size_t)(j - isd) * ni + (i - isd)]) *
            dx[(size_t)(j - isd) * ni + (i - isd)] / dyc;
      }
      del6_u[(size_t)(j - isd) * ni + (i - isd)] = v;
    }
  if (xyz_c) memcpy(xyz_c, pc, sizeof(double) * 3 * s2);
  if (xyz_n) memcpy(xyz_n, pn, sizeof(double) * 3 * nn * nn);
#undef PN
#undef PC
#undef C2
  free(pn);
  free(pc);
  return 0;
}

/* ---- stream functions at corners ---------------------------------------- */
/* kind 0: solid-body rotation about the axis at colatitude-tilt par[0]
 *         (radians from the pole toward lon 0): psi = -R (w . p), |u| = 1 m/s
 *         on the rotation equator.
 * kind 1: Nair-Lauritzen non-divergent deformational flow at time
 *         t/T = par[0], period par[1] seconds:
 *         psi = (R^2/T) [10 sin^2(lon - 2 pi t/T) cos^2(lat) cos(pi t/T) - 2 pi sin(lat)] / R
 *         scaled so its peak speed is O(1) m/s per unit amplitude. */
static double stream(int kind, const double *par, const double p[3], double radius) {
  if (kind == 0) {
    double a = par[0];
    double w[3] = {sin(a), 0.0, cos(a)};
    return -radius * (w[0] * p[0] + w[1] * p[1] + w[2] * p[2]);
  } else {
    const double pi = 3.14159265358979323846;
    double tt = par[0];
    double lon = atan2(p[1], p[0]), s = p[2];
    double c2 = 1.0 - s * s;
    double sl = sin(lon - 2.0 * pi * tt);
    return radius / (2.0 * pi) * (10.0 * sl * sl * c2 * cos(pi * tt) - 2.0 * pi * s);
  }
}

/* Face Courant numbers and mass fluxes of one tile for npz levels from a
 * corner stream function.  Area flux through x-face (i,j):
 *   xfx = dt speed_k (psi(i,j) - psi(i,j+1)),  y-face: dt speed_k (psi(i+1,j) - psi(i,j))
 * (discretely divergence-free); Courant numbers invert the upwind area-flux
 * formula of tracer_2d; mass fluxes are xfx * dp_ref(k) on interior faces.
 * cx (is:ie+1, jsd:jed, npz), cy (isd:ied, js:je+1, npz),
 * mfx (is:ie+1, js:je, npz), mfy (is:ie, js:je+1, npz). */
int fv3c_winds(int n, int tile, int npz, int kind, const double *par, double radius,
               const double *speed_k, double dt, const double *dp_ref, const double *dxa,
               const double *dy, const double *dya, const double *dx, const double *sin_sg,
               double *cx, double *cy, double *mfx, double *mfy) {
  const int is = 1, ie = n, js = 1, je = n, isd = 1 - NG, ied = n + NG, jsd = isd, jed = ied;
  const int ni = ied - isd + 1, nn = ni + 1, nix = ie - is + 2, niy = ie - is + 1;
  const size_t s2 = (size_t)ni * ni;
  double *psi = (double *)malloc(sizeof(double) * nn * nn);
  for (int j = isd; j <= jed + 1; j++)
    for (int i = isd; i <= ied + 1; i++) {
      double p[3];
      node_xyz(n, tile, i, j, p);
      psi[(size_t)(j - isd) * nn + (i - isd)] = stream(kind, par, p, radius);
    }
#define PSI(i, j) psi[(size_t)((j) - isd) * nn + ((i) - isd)]
#define M2(a, i, j) (a)[(size_t)((j) - jsd) * ni + ((i) - isd)]
  for (int k = 0; k < npz; k++) {
    const double sc = dt * speed_k[k];
    double *cxk = cx + (size_t)k * nix * (jed - jsd + 1);
    double *cyk = cy + (size_t)k * ni * (je - js + 2);
    double *mxk = mfx + (size_t)k * nix * (je - js + 1);
    double *myk = mfy + (size_t)k * niy * (je - js + 2);
    for (int j = jsd; j <= jed; j++)
      for (int i = is; i <= ie + 1; i++) {
        double f = sc * (PSI(i, j) - PSI(i, j + 1));
        double dyf = dy[(size_t)(j - jsd) * nn + (i - isd)];
        double den = f > 0.0 ? M2(dxa, i - 1, j) * dyf * sin_sg[2 * s2 + (size_t)(j - jsd) * ni + (i - 1 - isd)]
                             : M2(dxa, i, j) * dyf * sin_sg[0 * s2 + (size_t)(j - jsd) * ni + (i - isd)];
        cxk[(size_t)(j - jsd) * nix + (i - is)] = f / den;
        if (j >= js && j <= je) mxk[(size_t)(j - js) * nix + (i - is)] = f * dp_ref[k];
      }
    for (int j = js; j <= je + 1; j++)
      for (int i = isd; i <= ied; i++) {
        double f = sc * (PSI(i + 1, j) - PSI(i, j));
        double dxf = dx[(size_t)(j - jsd) * ni + (i - isd)];
        double den = f > 0.0 ? M2(dya, i, j - 1) * dxf * sin_sg[3 * s2 + (size_t)(j - 1 - jsd) * ni + (i - isd)]
                             : M2(dya, i, j) * dxf * sin_sg[1 * s2 + (size_t)(j - jsd) * ni + (i - isd)];
        cyk[(size_t)(j - js) * ni + (i - isd)] = f / den;
        if (i >= is && i <= ie) myk[(size_t)(j - js) * niy + (i - is)] = f * dp_ref[k];
      }
  }
#undef PSI
#undef M2
  free(psi);
  return 0;
}

/* ---- counter-based RNG (splitmix64) -------------------------------------- */
static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* U(-1, 1) keyed by (seed, tile, i, j, k, field) */
double fv3c_uniform(uint64_t seed, int tile, int i, int j, int k, int field) {
  uint64_t h = splitmix64(seed);
  h = splitmix64(h ^ (uint64_t)(uint32_t)tile);
  h = splitmix64(h ^ (uint64_t)(uint32_t)(i + 1024));
  h = splitmix64(h ^ (uint64_t)(uint32_t)(j + 1024));
  h = splitmix64(h ^ (uint64_t)(uint32_t)k);
  h = splitmix64(h ^ (uint64_t)(uint32_t)field);
  return ((double)(h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}

/* ---- tracer patterns ------------------------------------------------------ */
/* kind 0: cosine bell centred at (lon par[0], lat par[1]), radius par[2]
 *         (radians), background par[3]: q = b + (1-b) * 0.5 (1 + cos(pi r/r0)).
 * kind 1: slotted cylinder, same centre/radius, 1 inside (slot of half-width
 *         r0/6 cut from the south edge up to 5 r0/12 north of the centre), 0 out.
 * kind 2: sin(lon) cos(lat).
 * Evaluated at halo-folded cell centres; q2d is (isd:ied, jsd:jed). */
int fv3c_tracer(int n, int tile, int kind, const double *par, double *q2d) {
  const int isd = 1 - NG, ied = n + NG, ni = ied - isd + 1;
  const double pi = 3.14159265358979323846;
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied; i++) {
      double p[3];
      cell_xyz(n, tile, i, j, p);
      double v;
      if (kind == 2) {
        double lon = atan2(p[1], p[0]);
        double lat = asin(fmax(-1.0, fmin(1.0, p[2])));
        v = sin(lon) * cos(lat);
      } else {
        double c0[3] = {cos(par[1]) * cos(par[0]), cos(par[1]) * sin(par[0]), sin(par[1])};
        double r = gdist(p, c0), r0 = par[2];
        if (kind == 0) {
          double h = r < r0 ? 0.5 * (1.0 + cos(pi * r / r0)) : 0.0;
          v = par[3] + (1.0 - par[3]) * h;
        } else {
          /* local east/north offsets around the centre */
          double e[3] = {-sin(par[0]), cos(par[0]), 0.0};
          double nv[3] = {-sin(par[1]) * cos(par[0]), -sin(par[1]) * sin(par[0]), cos(par[1])};
          double xe = p[0] * e[0] + p[1] * e[1] + p[2] * e[2];
          double xn = p[0] * nv[0] + p[1] * nv[1] + p[2] * nv[2];
          v = 0.0;
          if (r <= r0 && (fabs(xe) >= r0 / 6.0 || xn >= 5.0 * r0 / 12.0)) v = 1.0;
        }
      }
      q2d[(size_t)(j - isd) * ni + (i - isd)] = v;
    }
  return 0;
}

/* Smooth field on the sphere: sum of four seeded plane waves, range ~[-1, 1]. */
static double smooth_noise(uint64_t seed, int k, const double p[3]) {
  double s = 0.0;
  for (int m = 0; m < 4; m++) {
    double a = fv3c_uniform(seed, 0, m, 0, k, 101);
    double b = fv3c_uniform(seed, 0, m, 1, k, 101);
    double c = fv3c_uniform(seed, 0, m, 2, k, 101);
    double ph = 3.14159265358979323846 * fv3c_uniform(seed, 0, m, 3, k, 101);
    double w = 2.0 + 2.0 * (m + 1);
    s += 0.25 * sin(w * (a * p[0] + b * p[1] + c * p[2]) + ph);
  }
  return s;
}

/* Perturbed layer thicknesses delp (isd:ied, jsd:jed, npz) of one tile:
 * kind 0: dp_ref(k) (1 + amp U(-1,1)), U keyed by the cell's source
 *         (tile, i, j) so halos equal the neighbours' interior values;
 * kind 1: dp_ref(k) (1 + amp_k * smooth(p)), amp_k from amp[0]..amp[1]
 *         linearly over the levels.
 * Every column is then rescaled so its sum is ps - ptop. */
int fv3c_delp(int n, int tile, int npz, int kind, const double *amp, uint64_t seed,
              const double *dp_ref, double ps_minus_ptop, double *delp) {
  const int isd = 1 - NG, ied = n + NG, ni = ied - isd + 1;
  const size_t s2 = (size_t)ni * ni;
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied; i++) {
      int st = tile, si = i, sj = j;
      int r = fv3c_halo_source(n, tile, i, j, &st, &si, &sj);
      if (r < 0) {
        st = tile;
        si = i;
        sj = j;
      }
      double p[3];
      cell_xyz(n, tile, i, j, p);
      double sum = 0.0;
      for (int k = 0; k < npz; k++) {
        double f;
        if (kind == 0) {
          f = 1.0 + amp[0] * fv3c_uniform(seed, st, si, sj, k, 7);
        } else {
          double ak = npz > 1 ? amp[0] + (amp[1] - amp[0]) * (double)k / (double)(npz - 1) : amp[0];
          f = 1.0 + ak * smooth_noise(seed, k, p);
        }
        double v = dp_ref[k] * f;
        delp[(size_t)k * s2 + (size_t)(j - isd) * ni + (i - isd)] = v;
        sum += v;
      }
      double sc = ps_minus_ptop / sum;
      for (int k = 0; k < npz; k++) {
        double *d = &delp[(size_t)k * s2 + (size_t)(j - isd) * ni + (i - isd)];
        *d = *d * sc;
      }
    }
  return 0;
}

/* Cell-centre unit vectors (3, isd:ied, jsd:jed) of a tile (halo-folded). */
int fv3c_cell_centers(int n, int tile, double *xyz) {
  const int isd = 1 - NG, ied = n + NG, ni = ied - isd + 1;
  for (int j = isd; j <= ied; j++)
    for (int i = isd; i <= ied; i++) cell_xyz(n, tile, i, j, xyz + 3 * ((size_t)(j - isd) * ni + (i - isd)));
  return 0;
}
