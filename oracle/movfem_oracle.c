/*
 * movfem_oracle.c -- CPU restatement of MoVFEM_3DMT's element-integration,
 * assembly and sort path (reference: MoVFEM_3DMT/src, GPL-3.0; nothing here is
 * copied, every routine is re-expressed from the reference's behaviour).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product path
 * is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never calls it.
 *
 * It reproduces the reference's effective arithmetic, including:
 *   - float32 Gauss points / 3-point weights      (integration.f90:297-299,400-402)
 *   - float32 omega in the element matrix         (integration.f90:84)
 *   - real-part truncation of f1/f2               (integration.f90:154,211)
 *   - real-part storage of the GPML stretch h     (integration.f90:15,126)
 *   - GPML zone flags taken from the previously processed element, because
 *     get_pml runs after int_elem_params          (MoVFEM_3DMT.f90:203-206)
 *   - 8-node d/dzeta typo                         (n_fem.f90:193)
 *   - J^-1 divided by |det J|                     (n_fem.f90:378-386)
 *   - float32 constants in primary field / source (problem.f90:112,250,352-356)
 *   - the update_sigma linear-index range         (geometry.f90:144-153)
 *   - complex64 values and the column-only value permutation of ga_sort_sparse
 *                                                 (global_assembly.f90:152-181)
 *
 * Pinning status of each function is stated in the header of oracle/oracle.py.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

#define ORC_PI 3.1415926535897932384626433
#define ORC_EPS 8.854187817e-12
#define ORC_B0 1.0e-9

/* ------------------------------------------------------------------------- */
/* Reference-element tables.  Data tables that must agree with the reference:
 * node coordinates (n_fem.f90:113-133), edge (basis node, direction) pairs
 * (v_fem.f90:491-504), neighbour copy maps (global_assembly.f90:242-265,
 * 310-351,396-443), which double as the Dirichlet face lists
 * (boundary_conds.f90:276-388).  All indices 1-based as in the reference.   */
static const signed char NR27[27][3] = {
    {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},  {-1, -1, 1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1}, {0, -1, -1},
    {1, 0, 1},   {0, 1, 1},  {-1, 0, 1}, {0, -1, 1},   {1, -1, 0}, {1, 1, 0},
    {-1, 1, 0},  {-1, -1, 0}, {1, 0, 0}, {0, 1, 0},    {-1, 0, 0}, {0, -1, 0},
    {0, 0, -1},  {0, 0, 1},  {0, 0, 0}};

static const unsigned char ED12[12][2] = {{4, 3}, {4, 2}, {8, 2}, {3, 3}, {4, 1}, {8, 1},
                                          {3, 1}, {7, 1}, {1, 3}, {1, 2}, {5, 2}, {2, 3}};
static const unsigned char ED36[36][2] = {
    {4, 3},  {8, 3},  {4, 2},  {3, 2},  {8, 2},  {7, 2},  {3, 3},  {7, 3},  {4, 1},
    {1, 1},  {8, 1},  {5, 1},  {3, 1},  {2, 1},  {7, 1},  {6, 1},  {1, 3},  {5, 3},
    {1, 2},  {2, 2},  {5, 2},  {6, 2},  {2, 3},  {6, 3},  {12, 3}, {20, 2}, {12, 2},
    {16, 2}, {10, 3}, {17, 2}, {11, 3}, {20, 1}, {11, 1}, {15, 1}, {19, 1}, {9, 3}};
static const unsigned char ED54[54][2] = {
    {4, 3},  {8, 3},  {4, 2},  {20, 2}, {8, 2},  {11, 3}, {15, 3}, {3, 2},  {19, 2},
    {7, 2},  {3, 3},  {7, 3},  {4, 1},  {20, 1}, {8, 1},  {11, 1}, {15, 1}, {3, 1},
    {19, 1}, {7, 1},  {12, 3}, {16, 3}, {12, 2}, {10, 2}, {16, 2}, {14, 2}, {10, 3},
    {14, 3}, {1, 1},  {17, 1}, {5, 1},  {9, 1},  {13, 1}, {2, 1},  {18, 1}, {6, 1},
    {1, 3},  {5, 3},  {1, 2},  {17, 2}, {5, 2},  {9, 3},  {13, 3}, {2, 2},  {18, 2},
    {6, 2},  {2, 3},  {6, 3},  {25, 3}, {26, 3}, {24, 2}, {22, 2}, {23, 1}, {21, 1}};

/* copy maps: [face z,y,x][k] = {dst (this element), src (neighbour at -face)} */
static const unsigned char CP12[3][4][2] = {
    {{2, 3}, {5, 6}, {7, 8}, {10, 11}},
    {{1, 4}, {5, 7}, {6, 8}, {9, 12}},
    {{1, 9}, {2, 10}, {3, 11}, {4, 12}}};
static const unsigned char CP36[3][10][2] = {
    {{3, 5}, {4, 6}, {9, 11}, {10, 12}, {13, 15}, {14, 16}, {19, 21}, {20, 22}, {27, 28}, {33, 34}},
    {{1, 7}, {2, 8}, {9, 13}, {10, 14}, {11, 15}, {12, 16}, {17, 23}, {18, 24}, {25, 29}, {32, 35}},
    {{1, 17}, {2, 18}, {3, 19}, {4, 20}, {5, 21}, {6, 22}, {7, 23}, {8, 24}, {26, 30}, {31, 36}}};
static const unsigned char CP54[3][12][2] = {
    {{3, 5}, {8, 10}, {13, 15}, {16, 17}, {18, 20}, {23, 25}, {24, 26}, {29, 31}, {32, 33}, {34, 36},
     {39, 41}, {44, 46}},
    {{1, 11}, {2, 12}, {13, 18}, {14, 19}, {15, 20}, {21, 27}, {22, 28}, {29, 34}, {30, 35}, {31, 36},
     {37, 47}, {38, 48}},
    {{1, 37}, {2, 38}, {3, 39}, {4, 40}, {5, 41}, {6, 42}, {7, 43}, {8, 44}, {9, 45}, {10, 46},
     {11, 47}, {12, 48}}};

/* ------------------------------------------------------------------------- */
typedef struct orc_state {
  int mn, me, nord, ngd, gpt, nf; /* nodes, edges, lattice order, GP/dir, GP, per-face */
  int nx, ny, nz;                 /* elements along x,y,z (reference g_nx-1 ...) */
  int nnx, nny, nnz;
  int64_t nyz, npt, ne;
  double *xp, *yp, *zp;
  double *sig; /* complex (6,npt), interleaved re/im, mutable (update_sigma) */
  double *mu;  /* real (6,npt) */
  int dirichlet, inimod, gpml_sch;
  double a0, b0, nn;
  int nextd, nzl_top;
  /* boundary 1-D model of bd_setmodel (boundary_conds.f90:396-430) */
  double ztop, h_sigma;
  int nl;
  double *lsig, *ldz, *zl; /* (nl), (nl-1), (nl) */
  zc *cz, *ez;             /* Wait recursion state (nl) */
  double omega;
  int el_xa[2], el_xb[2], el_ya[2], el_yb[2], el_za[2], el_zb[2];
  double xa[2], xb[2], ya[2], yb[2], za[2], zb[2];
  int in_pml[3]; /* persists across calls like the reference's module variable */
  double rw[27][4];
  int32_t *gne; /* (ne, me) element-major: gne[e*me + im] */
  int32_t nne;
  int32_t *nzindx; /* (ne*me*me) */
  int64_t nnze;
  const unsigned char (*ed)[2];
  const unsigned char *cp; /* [3][nf][2] */
} orc_state;

/* ---------------------------- shape functions ---------------------------- */
/* nodal shape function N_i (n_fem.f90:138-178), i 1-based */
static double nf_ln(const orc_state *s, int i, const double *x) {
  const signed char *r = NR27[i - 1];
  double v = 1.0;
  if (s->mn == 8) {
    for (int c = 0; c < 3; c++) v *= (1 + r[c] * x[c]);
    return v / 8.0;
  }
  if (s->mn == 20) {
    if (i <= 8) {
      for (int c = 0; c < 3; c++) v *= (1 + r[c] * x[c]);
      return v * (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] - 2) / 8.0;
    }
    for (int c = 0; c < 3; c++) v *= (r[c] == 0) ? (1 - x[c] * x[c]) : (1 + r[c] * x[c]);
    return v / 4.0;
  }
  /* 27: tensor product of 1-D quadratic Lagrange factors */
  int nz0 = 0;
  for (int c = 0; c < 3; c++) {
    if (r[c] == 0) {
      v *= (1 - x[c] * x[c]);
      nz0++;
    } else {
      v *= (1 + r[c] * x[c]) * (r[c] * x[c]);
    }
  }
  static const double den[4] = {8.0, 4.0, 2.0, 1.0};
  return v / den[nz0];
}

/* d N_i / d xi_d (n_fem.f90:181-317); reproduces the 8-node d=3 form (:193) */
static double nf_dln(const orc_state *s, int d, int i, const double *x) {
  const signed char *r = NR27[i - 1];
  int a = (d + 1) % 3, b = (d + 2) % 3; /* the two other coordinates */
  if (s->mn == 8) {
    if (d == 2) /* reference uses (1+r1) instead of (1+r1*xi) */
      return ((1 + r[0]) * (1 + r[1] * x[1]) * r[2]) / 8.0;
    return r[d] * (1 + r[a] * x[a]) * (1 + r[b] * x[b]) / 8.0;
  }
  if (s->mn == 20) {
    if (i <= 8)
      return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) *
             (r[d] * (2 * r[d] * x[d] + r[a] * x[a] + r[b] * x[b] - 1)) / 8.0;
    if (r[d] == 0) return -x[d] * (1 + r[a] * x[a]) * (1 + r[b] * x[b]) / 2.0;
    /* derivative along a non-zero coordinate: the zero coordinate gives (1-x^2) */
    int z = (r[a] == 0) ? a : b, o = (r[a] == 0) ? b : a;
    return r[d] * (1 - x[z] * x[z]) * (1 + r[o] * x[o]) / 4.0;
  }
  /* 27-node: product of 1-D factors, derivative on coordinate d */
  double v = 1.0;
  int nz0 = 0;
  for (int c = 0; c < 3; c++) {
    if (r[c] == 0) {
      nz0++;
      v *= (c == d) ? (-2.0 * x[c]) : (1 - x[c] * x[c]);
    } else {
      v *= (c == d) ? (r[c] * (1 + 2 * r[c] * x[c])) : ((1 + r[c] * x[c]) * (r[c] * x[c]));
    }
  }
  static const double den[4] = {8.0, 4.0, 2.0, 1.0};
  return v / den[nz0];
}

/* mixed-order edge polynomial phi(node i, direction dir) (v_fem.f90:78-179) */
static double mix_ln(const orc_state *s, int i, int dir, const double *x) {
  const signed char *r = NR27[i - 1];
  int a = (dir + 1) % 3, b = (dir + 2) % 3; /* dir is 0-based; a,b transverse */
  if (s->mn == 8) return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) / 4.0;
  if (s->mn == 20) {
    if (i <= 8)
      return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) *
             (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] - 1) / 8.0;
    int z = (r[a] == 0) ? a : b, o = (r[a] == 0) ? b : a;
    return (1 - x[z] * x[z]) * (1 + r[o] * x[o]) / 2.0;
  }
  /* 27: linear along dir, quadratic Lagrange across */
  double v = (1 + r[dir] * x[dir]);
  int nz0 = 0;
  for (int k = 0; k < 2; k++) {
    int c = k ? b : a;
    if (r[c] == 0) {
      v *= (1 - x[c] * x[c]);
      nz0++;
    } else {
      v *= (1 + r[c] * x[c]) * (r[c] * x[c]);
    }
  }
  static const double den[3] = {8.0, 4.0, 2.0};
  return v / den[nz0];
}

/* d phi / d xi_d (v_fem.f90:184-466) */
static double mix_dln(const orc_state *s, int dir, int d, int i, const double *x) {
  const signed char *r = NR27[i - 1];
  int a = (dir + 1) % 3, b = (dir + 2) % 3;
  if (s->mn == 8) {
    if (d == dir) return 0.0;
    int o = (d == a) ? b : a;
    return r[d] * (1 + r[o] * x[o]) / 4.0;
  }
  if (s->mn == 20) {
    if (i <= 8) {
      if (d == dir) return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) * r[d] / 8.0;
      int o = (d == a) ? b : a;
      return r[d] * (1 + r[o] * x[o]) * (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] + r[d] * x[d]) / 8.0;
    }
    if (d == dir) return 0.0;
    int z = (r[a] == 0) ? a : b, o = (r[a] == 0) ? b : a;
    if (d == z) return -x[z] * (1 + r[o] * x[o]);
    return (1 - x[z] * x[z]) * r[o] / 2.0;
  }
  double v = (d == dir) ? (double)r[dir] : (1 + r[dir] * x[dir]);
  int nz0 = 0;
  for (int k = 0; k < 2; k++) {
    int c = k ? b : a;
    if (r[c] == 0) {
      nz0++;
      v *= (c == d) ? (-2.0 * x[c]) : (1 - x[c] * x[c]);
    } else {
      v *= (c == d) ? (r[c] * (1 + 2 * r[c] * x[c])) : ((1 + r[c] * x[c]) * (r[c] * x[c]));
    }
  }
  static const double den[3] = {8.0, 4.0, 2.0};
  return v / den[nz0];
}

/* ------------------------------ Jacobian --------------------------------- */
/* J(m,n) = sum_l dN_l/dxi_m x_l,n ; inverse by cofactors / |det| (n_fem.f90:355-395) */
static double orc_det(const double J[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
         J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}
static int orc_jac(const orc_state *s, const double re[][3], const double *x, double J[3][3],
                   double Ji[3][3], double *det) {
  for (int m = 0; m < 3; m++)
    for (int n = 0; n < 3; n++) {
      double acc = 0.0;
      for (int l = 0; l < s->mn; l++) acc += nf_dln(s, m, l + 1, x) * re[l][n];
      J[m][n] = acc;
    }
  double dj = orc_det(J);
  *det = dj;
  if (dj == 0.0) return -1;
  double ad = fabs(dj);
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / ad;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / ad;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / ad;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / ad;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / ad;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / ad;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / ad;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / ad;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / ad;
  return 0;
}

/* ------------------------------ quadrature ------------------------------- */
static void orc_quadrature(orc_state *s) {
  double p[3], w[3];
  int n = s->ngd;
  if (n == 2) {
    p[0] = (double)-0.577350269189625764509148780502f;
    p[1] = (double)0.577350269189625764509148780502f;
    w[0] = 1.0;
    w[1] = 1.0;
  } else {
    p[0] = (double)-0.774596669241483377035853079956f;
    p[1] = 0.0;
    p[2] = (double)0.774596669241483377035853079956f;
    w[0] = (double)0.555555555555555555555555555556f;
    w[1] = (double)0.888888888888888888888888888889f;
    w[2] = (double)0.555555555555555555555555555556f;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        int id = (i * n + j) * n + k;
        s->rw[id][0] = p[i];
        s->rw[id][1] = p[j];
        s->rw[id][2] = p[k];
        s->rw[id][3] = w[i] * w[j] * w[k];
      }
}

/* ------------------------------ state ------------------------------------ */
static inline double sig_re(const orc_state *s, int64_t node, int c) { return s->sig[2 * (node * 6 + c)]; }
static inline zc sig_z(const orc_state *s, int64_t node, int c) {
  return CMPLX(s->sig[2 * (node * 6 + c)], s->sig[2 * (node * 6 + c) + 1]);
}

int orc_set_bdmodel(orc_state *s, double ztop, double h_sigma, int nl, const double *l_sigma, const double *l_dz);

orc_state *orc_create(int nx, int ny, int nz, int mn, const double *xp, const double *yp, const double *zp,
                      const double *sigma, const double *mu, int dirichlet, int inimod, int gpml_sch,
                      double a0, double b0, double nn, int nextd, int nzl_top) {
  if (mn != 8 && mn != 20 && mn != 27) return NULL;
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->mn = mn;
  s->me = mn == 8 ? 12 : (mn == 20 ? 36 : 54);
  s->nf = mn == 8 ? 4 : (mn == 20 ? 10 : 12);
  s->nord = mn == 8 ? 2 : 3;
  s->ngd = mn == 8 ? 2 : 3;
  s->gpt = s->ngd * s->ngd * s->ngd;
  s->ed = mn == 8 ? ED12 : (mn == 20 ? ED36 : ED54);
  s->cp = mn == 8 ? &CP12[0][0][0] : (mn == 20 ? &CP36[0][0][0] : &CP54[0][0][0]);
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->nnx = nx * (s->nord - 1) + 1;
  s->nny = ny * (s->nord - 1) + 1;
  s->nnz = nz * (s->nord - 1) + 1;
  s->nyz = (int64_t)s->nny * s->nnz;
  s->npt = (int64_t)s->nnx * s->nyz;
  s->ne = (int64_t)nx * ny * nz;
  s->xp = (double *)malloc(sizeof(double) * s->nnx);
  s->yp = (double *)malloc(sizeof(double) * s->nny);
  s->zp = (double *)malloc(sizeof(double) * s->npt);
  s->sig = (double *)malloc(sizeof(double) * 12 * s->npt);
  s->mu = (double *)malloc(sizeof(double) * 6 * s->npt);
  memcpy(s->xp, xp, sizeof(double) * s->nnx);
  memcpy(s->yp, yp, sizeof(double) * s->nny);
  memcpy(s->zp, zp, sizeof(double) * s->npt);
  memcpy(s->sig, sigma, sizeof(double) * 12 * s->npt);
  memcpy(s->mu, mu, sizeof(double) * 6 * s->npt);
  s->dirichlet = dirichlet;
  s->inimod = inimod;
  s->gpml_sch = gpml_sch;
  s->a0 = a0;
  s->b0 = b0;
  s->nn = nn;
  s->nextd = nextd;
  s->nzl_top = nzl_top;
  orc_quadrature(s);
  /* init_gpml (boundary_conds.f90:51-70); g_nx = nx+1 etc. */
  int gnx = nx + 1, gny = ny + 1, gnz = nz + 1, o = s->nord - 1;
  s->el_xa[0] = nextd; s->el_xa[1] = gnx - nextd; s->el_xb[0] = 1; s->el_xb[1] = gnx - 1;
  s->el_ya[0] = nextd; s->el_ya[1] = gny - nextd; s->el_yb[0] = 1; s->el_yb[1] = gny - 1;
  s->el_za[0] = nextd; s->el_za[1] = gnz - nextd; s->el_zb[0] = 1; s->el_zb[1] = gnz - 1;
  if (nextd >= 1 && nextd * o + 1 <= s->nnx && nextd * o + 1 <= s->nny && nextd * o + 1 <= s->nnz) {
    s->xa[0] = s->xp[nextd * o]; s->xa[1] = s->xp[s->nnx - nextd * o - 1];
    s->ya[0] = s->yp[nextd * o]; s->ya[1] = s->yp[s->nny - nextd * o - 1];
    int64_t id1 = nextd * o; /* 0-based node id nextd*o+1 */
    int64_t id2 = (int64_t)(s->nnx - 1) * s->nyz + (int64_t)(s->nny - 1) * s->nnz + s->nnz - nzl_top * o - 1;
    s->za[0] = s->zp[id1];
    s->za[1] = (id2 >= 0 && id2 < s->npt) ? s->zp[id2] : 0.0;
  }
  s->xb[0] = s->xp[0]; s->xb[1] = s->xp[s->nnx - 1];
  s->yb[0] = s->yp[0]; s->yb[1] = s->yp[s->nny - 1];
  s->zb[0] = s->zp[0]; s->zb[1] = s->zp[s->npt - 1];
  s->in_pml[0] = s->in_pml[1] = s->in_pml[2] = 0;
  const double one_layer = 0.01;
  orc_set_bdmodel(s, 0.0, 0.01, 1, &one_layer, NULL);
  return s;
}

void orc_free(orc_state *s) {
  if (!s) return;
  free(s->xp); free(s->yp); free(s->zp); free(s->sig); free(s->mu);
  free(s->gne); free(s->nzindx);
  free(s->lsig); free(s->ldz); free(s->zl); free(s->cz); free(s->ez);
  free(s);
}

/* update_omega (geometry.f90:137-140) */
void orc_set_omega(orc_state *s, double omega) { s->omega = omega; }

/* bd_setmodel (boundary_conds.f90:396-430) with g_ztop (geometry.f90:389-394):
 * half-space conductivity h_sigma (inimod 2) or nl layers l_sigma with nl-1
 * thicknesses l_dz in km, interfaces at zl(1) = 0, zl(i) = zl(i-1) - dl(i-1)
 * (inimod 3).  Returns 0, or -1 for nl < 1. */
int orc_set_bdmodel(orc_state *s, double ztop, double h_sigma, int nl, const double *l_sigma, const double *l_dz) {
  if (nl < 1) return -1;
  free(s->lsig); free(s->ldz); free(s->zl); free(s->cz); free(s->ez);
  s->ztop = ztop;
  s->h_sigma = h_sigma;
  s->nl = nl;
  s->lsig = (double *)malloc(sizeof(double) * nl);
  s->ldz = (double *)malloc(sizeof(double) * nl);
  s->zl = (double *)malloc(sizeof(double) * nl);
  s->cz = (zc *)malloc(sizeof(zc) * nl);
  s->ez = (zc *)malloc(sizeof(zc) * nl);
  memcpy(s->lsig, l_sigma, sizeof(double) * nl);
  if (nl > 1) memcpy(s->ldz, l_dz, sizeof(double) * (nl - 1));
  s->zl[0] = 0.0;
  for (int i = 1; i < nl; i++) s->zl[i] = s->zl[i - 1] - s->ldz[i - 1];
  return 0;
}

/* update_sigma (geometry.f90:144-153): the loop runs i<=npt, j<=6 over a
 * (6,npt) array, so it touches linear entries 0..npt+29 only.  Each touched
 * entry with a non-zero imaginary part gets imag = float32(eps*omega). */
void orc_update_sigma(orc_state *s) {
  int64_t lim = s->npt + 30;
  if (lim > 6 * s->npt) lim = 6 * s->npt;
  double im = (double)(float)(ORC_EPS * s->omega);
  for (int64_t L = 0; L < lim; L++)
    if (s->sig[2 * L + 1] != 0.0) s->sig[2 * L + 1] = im;
}

/* read back the (mutated) sigma array */
void orc_get_sigma(const orc_state *s, double *out) { memcpy(out, s->sig, sizeof(double) * 12 * s->npt); }

/* ------------------------------ numbering -------------------------------- */
static int edge_face(const orc_state *s, int ie, int je, int ke, int im) {
  /* edge_bdary (boundary_conds.f90:262-390): faces tested in order -x,-y,-z,+x,+y,+z */
  const unsigned char *cp = s->cp;
  int nf = s->nf;
  int conds[6] = {ie == 1, je == 1, ke == 1, ie == s->nx, je == s->ny, ke == s->nz};
  for (int f = 0; f < 6; f++) {
    if (!conds[f]) continue;
    int axis = (f % 3 == 0) ? 2 : (f % 3 == 1 ? 1 : 0); /* copy table order z,y,x */
    int col = (f < 3) ? 0 : 1;                            /* -face: dst list, +face: src list */
    for (int k = 0; k < nf; k++)
      if (cp[(axis * nf + k) * 2 + col] == im) return f + 1;
  }
  return 0;
}

/* c_gne12/36/54 (global_assembly.f90:231-475) */
int32_t orc_build_gne(orc_state *s) {
  int me = s->me, nf = s->nf;
  free(s->gne);
  s->gne = (int32_t *)calloc((size_t)s->ne * me, sizeof(int32_t));
  int32_t indx = 0;
  int copied[54];
  for (int ie = 1; ie <= s->nx; ie++)
    for (int je = 1; je <= s->ny; je++)
      for (int ke = 1; ke <= s->nz; ke++) {
        int64_t ide = ((int64_t)(ie - 1) * s->ny + (je - 1)) * s->nz + (ke - 1);
        int32_t *g = s->gne + ide * me;
        memset(copied, 0, sizeof(copied));
        int cond[3] = {ke > 1, je > 1, ie > 1};
        int64_t nb[3] = {ide - 1, ide - s->nz, ide - (int64_t)s->ny * s->nz};
        for (int ax = 0; ax < 3; ax++) {
          if (!cond[ax]) continue;
          const int32_t *gn = s->gne + nb[ax] * me;
          for (int k = 0; k < nf; k++) {
            int dst = s->cp[(ax * nf + k) * 2 + 0], src = s->cp[(ax * nf + k) * 2 + 1];
            g[dst - 1] = gn[src - 1];
            copied[dst - 1] = 1;
          }
        }
        int bel = (ie == 1 || ie == s->nx || je == 1 || je == s->ny || ke == 1 || ke == s->nz);
        for (int im = 1; im <= me; im++) {
          if (s->dirichlet && bel) {
            int f = edge_face(s, ie, je, ke, im);
            if (f) {
              g[im - 1] = -f;
              continue;
            }
          }
          if (!copied[im - 1]) g[im - 1] = ++indx;
        }
      }
  int32_t mx = 0;
  for (int64_t i = 0; i < s->ne * me; i++)
    if (s->gne[i] > mx) mx = s->gne[i];
  s->nne = mx;
  return mx;
}

/* ga_nzindx + shr_nzindx12/36/54 (global_assembly.f90:197-229,477-1906).
 * The reference's case tables are instances of one rule, restated here:
 *   a DOF on this element's +z/+y/+x face(s) is shared forward.  Processing its
 *   row assigns slots for the own row, then (in order z,y,x) for the neighbour
 *   across each such face -- first copying the slots of DOFs shared across that
 *   face, then numbering the rest -- and finally, for a DOF on two + faces,
 *   the diagonal neighbour, whose face-shared slots are copied first from the
 *   lower-order face neighbour through the higher-order face map, then from the
 *   higher-order face neighbour through the lower-order face map.  DOFs on no +
 *   face ("case default") number their own row's unassigned slots. */
static int face_pos(const orc_state *s, int ax, int col, int im) {
  for (int k = 0; k < s->nf; k++)
    if (s->cp[(ax * s->nf + k) * 2 + col] == im) return k;
  return -1;
}
static int32_t *slot(orc_state *s, int64_t ide, int im, int jm) {
  return s->nzindx + (ide * s->me + (im - 1)) * s->me + (jm - 1);
}
static void own_row(orc_state *s, int64_t ide, int im, int64_t *indx) {
  const int32_t *g = s->gne + ide * s->me;
  for (int jm = 1; jm <= s->me; jm++) {
    if (g[jm - 1] < 0) continue;
    int32_t *p = slot(s, ide, im, jm);
    if (*p == -1) *p = (int32_t)(++(*indx));
  }
}
/* copy slots (src elem, row im_s) -> (dst elem, row im_d) through face map ax */
static void copy_face(orc_state *s, int64_t src, int im_s, int64_t dst, int im_d, int ax) {
  for (int k = 0; k < s->nf; k++) {
    int js = s->cp[(ax * s->nf + k) * 2 + 1], jd = s->cp[(ax * s->nf + k) * 2 + 0];
    int32_t *pd = slot(s, dst, im_d, jd);
    if (*pd != -1) continue;
    *pd = *slot(s, src, im_s, js);
  }
}

int64_t orc_build_nzindx(orc_state *s) {
  if (!s->gne) orc_build_gne(s);
  int me = s->me;
  free(s->nzindx);
  size_t n = (size_t)s->ne * me * me;
  s->nzindx = (int32_t *)malloc(n * sizeof(int32_t));
  for (size_t i = 0; i < n; i++) s->nzindx[i] = -1;
  int64_t indx = 0;
  int64_t stride[3] = {1, s->nz, (int64_t)s->ny * s->nz};
  for (int ie = 1; ie <= s->nx; ie++)
    for (int je = 1; je <= s->ny; je++)
      for (int ke = 1; ke <= s->nz; ke++) {
        int64_t ide = ((int64_t)(ie - 1) * s->ny + (je - 1)) * s->nz + (ke - 1);
        int lt[3] = {ke < s->nz, je < s->ny, ie < s->nx};
        for (int im = 1; im <= me; im++) {
          if (s->gne[ide * me + im - 1] < 0) continue;
          int pos[3], on = 0;
          for (int ax = 0; ax < 3; ax++) {
            pos[ax] = face_pos(s, ax, 1, im); /* on + face of axis ax ? */
            if (pos[ax] >= 0) on++;
          }
          own_row(s, ide, im, &indx);
          if (on == 0) continue; /* case default: own row only */
          int im1[3];
          for (int ax = 0; ax < 3; ax++) {
            if (pos[ax] < 0) continue;
            im1[ax] = s->cp[(ax * s->nf + pos[ax]) * 2 + 0];
            if (!lt[ax]) continue;
            int64_t nb = ide + stride[ax];
            copy_face(s, ide, im, nb, im1[ax], ax);
            own_row(s, nb, im1[ax], &indx);
          }
          if (on == 2) {
            int lo = -1, hi = -1;
            for (int ax = 0; ax < 3; ax++)
              if (pos[ax] >= 0) {
                if (lo < 0) lo = ax;
                else hi = ax;
              }
            if (lt[lo] && lt[hi]) {
              int64_t nlo = ide + stride[lo], nhi = ide + stride[hi], nd = ide + stride[lo] + stride[hi];
              /* row index of the DOF in the diagonal element */
              int pd = face_pos(s, hi, 1, im1[lo]);
              int imd = s->cp[(hi * s->nf + pd) * 2 + 0];
              copy_face(s, nlo, im1[lo], nd, imd, hi);
              copy_face(s, nhi, im1[hi], nd, imd, lo);
              own_row(s, nd, imd, &indx);
            }
          }
        }
      }
  int64_t mx = 0;
  for (size_t i = 0; i < n; i++)
    if (s->nzindx[i] > mx) mx = s->nzindx[i];
  s->nnze = mx;
  return mx;
}

int32_t orc_nne(const orc_state *s) { return s->nne; }
int64_t orc_nnze(const orc_state *s) { return s->nnze; }
void orc_get_gne(const orc_state *s, int32_t *out) { memcpy(out, s->gne, sizeof(int32_t) * s->ne * s->me); }
void orc_get_nzindx(const orc_state *s, int32_t *out) {
  memcpy(out, s->nzindx, sizeof(int32_t) * s->ne * s->me * s->me);
}

/* ------------------------------ element ---------------------------------- */
typedef struct {
  double re[27][3];
  int64_t idx[27];
  double inmu[27][6], dmu[27][6];
  zc dsig[27][6];
  zc ep[2][27][3], hp[2][27][3];
  double wgt[27];
  double mf1[27][6], mf2[27][6]; /* real parts; imaginary parts never reach f1/f2 */
  double h[27][3];
  zc src[2][27][3];
  double cv1[27][54][3], cv2[27][54][3], ve[27][54][3];
} orc_ws;

static void inv_tensor(const double *a, double *b) {
  /* problem.f90:279-306 */
  double d = a[0] * (a[3] * a[5] - a[4] * a[4]) + a[1] * (a[2] * a[4] - a[1] * a[5]) +
             a[2] * (a[1] * a[4] - a[3] * a[2]);
  b[0] = (a[3] * a[5] - a[4] * a[4]) / d;
  b[1] = (a[2] * a[4] - a[1] * a[5]) / d;
  b[2] = (a[1] * a[4] - a[2] * a[3]) / d;
  b[3] = (a[0] * a[5] - a[2] * a[2]) / d;
  b[4] = (a[2] * a[1] - a[0] * a[4]) / d;
  b[5] = (a[3] * a[0] - a[1] * a[1]) / d;
}

/* gpml_h real part (boundary_conds.f90:84-186); Im parts are discarded by the
 * real-valued gpml storage, scheme 1 therefore has Re h = 1 exactly. */
static double gpml_re(const orc_state *s, int axis, double r) {
  int f = s->in_pml[axis];
  if (f == 0) return 1.0;
  const double *a = axis == 0 ? s->xa : (axis == 1 ? s->ya : s->za);
  const double *b = axis == 0 ? s->xb : (axis == 1 ? s->yb : s->zb);
  int k = (f == -1) ? 0 : 1;
  double rp = sqrt((b[k] - a[k]) * (b[k] - a[k]));
  double rr = sqrt((r - a[k]) * (r - a[k]));
  if (s->gpml_sch == 0) return 1.0 + s->a0 * pow(rr / rp, s->nn);
  return 1.0;
}

/* get_pml (boundary_conds.f90:72-82) */
static void orc_get_pml(orc_state *s, int i, int j, int k) {
  s->in_pml[0] = s->in_pml[1] = s->in_pml[2] = 0;
  if (i >= s->el_xa[0] && i <= s->el_xb[0]) s->in_pml[0] = -1;
  if (i >= s->el_xa[1] && i <= s->el_xb[1]) s->in_pml[0] = 1;
  if (j >= s->el_ya[0] && j <= s->el_yb[0]) s->in_pml[1] = -1;
  if (j >= s->el_ya[1] && j <= s->el_yb[1]) s->in_pml[1] = 1;
  if (k >= s->el_za[0] && k <= s->el_zb[0]) s->in_pml[2] = -1;
  if (k >= s->el_za[1] && k <= s->el_zb[1]) s->in_pml[2] = 1;
}

/* p_source (problem.f90:91-128) at reference point x of the element in w
 * (Ji: inverse Jacobian there), both polarizations, into src[d][g][c]:
 * (sum_n grad N_n x (mu_n^-1 dmu_n) Hp_n + sum_n N_n dsigma_n Ep_n) (-i fl32(w)) */
static void orc_psource(const orc_state *s, const orc_ws *w, const double *x, double Ji[3][3], double w32,
                        zc src[][27][3], int g) {
  int mn = s->mn;
  for (int d = 0; d < 2; d++) {
    zc pc[3] = {0, 0, 0}, dm[3] = {0, 0, 0};
    for (int n = 0; n < mn; n++) {
      /* p_pcurl: grad N x (mu^-1 dmu) Hp */
      double dn[3];
      for (int a = 0; a < 3; a++) {
        double acc = 0.0;
        for (int b = 0; b < 3; b++) acc += Ji[a][b] * nf_dln(s, b, n + 1, x);
        dn[a] = acc;
      }
      const double *im = w->inmu[n], *dmu = w->dmu[n];
      double M[3][3] = {
          {im[0] * dmu[0] + im[1] * dmu[1] + im[2] * dmu[2], im[0] * dmu[1] + im[1] * dmu[3] + im[2] * dmu[4],
           im[0] * dmu[2] + im[1] * dmu[4] + im[2] * dmu[5]},
          {im[1] * dmu[0] + im[3] * dmu[1] + im[4] * dmu[2], im[1] * dmu[1] + im[3] * dmu[3] + im[4] * dmu[4],
           im[1] * dmu[2] + im[3] * dmu[4] + im[4] * dmu[5]},
          {im[2] * dmu[0] + im[4] * dmu[1] + im[5] * dmu[2], im[2] * dmu[1] + im[4] * dmu[3] + im[5] * dmu[4],
           im[2] * dmu[2] + im[4] * dmu[4] + im[5] * dmu[5]}};
      zc v[3];
      for (int a = 0; a < 3; a++) v[a] = M[a][0] * w->hp[d][n][0] + M[a][1] * w->hp[d][n][1] + M[a][2] * w->hp[d][n][2];
      pc[0] += v[2] * dn[1] - v[1] * dn[2];
      pc[1] += v[0] * dn[2] - v[2] * dn[0];
      pc[2] += v[1] * dn[0] - v[0] * dn[1];
      /* p_dmpf: N * (dsigma . Ep) */
      const zc *ds = w->dsig[n];
      const zc *e = w->ep[d][n];
      zc u0 = ds[0] * e[0] + ds[1] * e[1] + ds[2] * e[2];
      zc u1 = ds[1] * e[0] + ds[3] * e[1] + ds[4] * e[2];
      zc u2 = ds[2] * e[0] + ds[4] * e[1] + ds[5] * e[2];
      double N = nf_ln(s, n + 1, x);
      dm[0] += N * u0;
      dm[1] += N * u1;
      dm[2] += N * u2;
    }
    zc mw = CMPLX(0.0, -w32);
    for (int c = 0; c < 3; c++) src[d][g][c] = (dm[c] + pc[c]) * mw;
  }
}

/* nf_get_r + p_elem_fields + int_elem_params for element (ie,je,ke), 1-based */
static int orc_elem_params(orc_state *s, orc_ws *w, int ie, int je, int ke) {
  int o = s->nord - 1, mn = s->mn, G = s->gpt, me = s->me;
  int64_t eno = (int64_t)(ie - 1) * s->nyz * o + (int64_t)(je - 1) * s->nnz * o + (int64_t)(ke - 1) * o;
  for (int n = 0; n < mn; n++) {
    const signed char *r = NR27[n];
    int i1 = (r[0] + 1) * o / 2, j1 = (r[1] + 1) * o / 2, k1 = (r[2] + 1) * o / 2;
    int64_t id = eno + (int64_t)i1 * s->nyz + (int64_t)j1 * s->nnz + k1;
    w->idx[n] = id;
    w->re[n][0] = s->xp[(ie - 1) * o + i1];
    w->re[n][1] = s->yp[(je - 1) * o + j1];
    w->re[n][2] = s->zp[id];
  }
  /* p_elem_fields (problem.f90:70-87) */
  double w32 = (double)(float)(s->omega);
  zc psig = CMPLX(0.0, (double)(float)(s->omega * ORC_EPS)); /* pset_pmodel */
  double pmu = 4 * ORC_PI * 1.e-7;
  for (int n = 0; n < mn; n++) {
    int64_t id = w->idx[n];
    zc sg[6];
    for (int c = 0; c < 6; c++) sg[c] = sig_z(s, id, c);
    zc cd = sg[0] * (sg[3] * sg[5] - sg[4] * sg[4]) + sg[1] * (sg[2] * sg[4] - sg[1] * sg[5]) +
            sg[2] * (sg[1] * sg[4] - sg[3] * sg[2]);
    const double *m = s->mu + 6 * id;
    double md = m[0] * (m[3] * m[5] - m[4] * m[4]) + m[1] * (m[2] * m[4] - m[1] * m[5]) +
                m[2] * (m[1] * m[4] - m[3] * m[2]);
    if (cd == 0.0 || md == 0.0) return -2; /* pinv_geomodel stops (problem.f90:260-271) */
    inv_tensor(m, w->inmu[n]);
    for (int c = 0; c < 6; c++) {
      w->dsig[n][c] = sg[c];
      w->dmu[n][c] = m[c];
    }
    w->dsig[n][0] -= psig; w->dsig[n][3] -= psig; w->dsig[n][5] -= psig;
    w->dmu[n][0] -= pmu; w->dmu[n][3] -= pmu; w->dmu[n][5] -= pmu;
    /* p_pfields (problem.f90:340-358) */
    double zz = (double)(float)(s->omega * ORC_B0 * s->zp[id]);
    double hh = (double)(float)ORC_B0 / (4 * ORC_PI * 1.e-7);
    for (int d = 0; d < 2; d++)
      for (int c = 0; c < 3; c++) w->ep[d][n][c] = w->hp[d][n][c] = 0.0;
    w->ep[0][n][0] = CMPLX(0.0, -zz);
    w->hp[0][n][1] = CMPLX(hh, 0.0);
    w->ep[1][n][1] = CMPLX(0.0, zz);
    w->hp[1][n][0] = CMPLX(hh, 0.0);
  }
  /* int_elem_params (integration.f90:60-74) */
  for (int g = 0; g < G; g++) {
    const double *x = s->rw[g];
    double J[3][3], Ji[3][3], det;
    if (orc_jac(s, w->re, x, J, Ji, &det)) return -1;
    w->wgt[g] = det * x[3];
    /* p_intmodels, pe_sch = 1 */
    for (int c = 0; c < 6; c++) {
      double a1 = 0.0, a2 = 0.0;
      for (int n = 0; n < mn; n++) {
        double N = nf_ln(s, n + 1, x);
        a1 += N * w->inmu[n][c];
        a2 += N * sig_re(s, w->idx[n], c);
      }
      w->mf1[g][c] = a1;
      w->mf2[g][c] = a2;
    }
    /* GPML stretch at the mapped point, using the zone flags currently held */
    if (!s->dirichlet) {
      double gr[3] = {0, 0, 0};
      for (int n = 0; n < mn; n++) {
        double N = nf_ln(s, n + 1, x);
        gr[0] += N * w->re[n][0];
        gr[1] += N * w->re[n][1];
        gr[2] += N * w->re[n][2];
      }
      for (int c = 0; c < 3; c++) w->h[g][c] = gpml_re(s, c, gr[c]);
    } else {
      w->h[g][0] = w->h[g][1] = w->h[g][2] = 1.0;
    }
    /* p_source (problem.f90:91-128) */
    orc_psource(s, w, x, Ji, w32, w->src, g);
    /* int_edges (integration.f90:139-152) */
    for (int j = 0; j < me; j++) {
      int node = s->ed[j][0], dir = s->ed[j][1] - 1;
      double vij[3] = {Ji[0][dir], Ji[1][dir], Ji[2][dir]};
      double dni[3];
      for (int a = 0; a < 3; a++) {
        double acc = 0.0;
        for (int b = 0; b < 3; b++) acc += Ji[a][b] * mix_dln(s, dir, b, node, x);
        dni[a] = acc;
      }
      w->cv1[g][j][0] = dni[1] * vij[2];
      w->cv2[g][j][0] = dni[2] * vij[1];
      w->cv1[g][j][1] = dni[2] * vij[0];
      w->cv2[g][j][1] = dni[0] * vij[2];
      w->cv1[g][j][2] = dni[0] * vij[1];
      w->cv2[g][j][2] = dni[1] * vij[0];
      double phi = mix_ln(s, node, dir, x);
      for (int c = 0; c < 3; c++) w->ve[g][j][c] = phi * vij[c];
    }
  }
  (void)w32;
  return 0;
}

/* tensor index of (a,b), 6-vector order [11,12,13,22,23,33] */
static const int T6[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

/* f1 (integration.f90:154-209): terms in the reference's order a, b, then
 * (k,l) = (1,1),(2,1),(1,2),(2,2); GPML weights are the products of h. */
static double orc_f1(const orc_state *s, const orc_ws *w, int g, int im, int jm) {
  const double *m = w->mf1[g];
  double c1[3][2], c2[3][2];
  for (int a = 0; a < 3; a++) {
    c1[a][0] = w->cv1[g][im][a]; c1[a][1] = w->cv2[g][im][a];
    c2[a][0] = w->cv1[g][jm][a]; c2[a][1] = w->cv2[g][jm][a];
  }
  static const int KL[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
  double acc = 0.0;
  if (s->dirichlet) {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        for (int t = 0; t < 4; t++) {
          int k = KL[t][0], l = KL[t][1];
          double v = (m[T6[a][b]] * c1[a][k]) * c2[b][l];
          acc = (k == l) ? acc + v : acc - v;
        }
    return acc;
  }
  const double *h = w->h[g];
  /* weight W(a,k;b,l) = g(a,k) g(b,l) written as the products of h the
   * reference uses; g^2 table: (1,1)=h1h3/h2 (1,2)=h1h2/h3 (2,1)=h1h2/h3
   * (2,2)=h2h3/h1 (3,1)=h2h3/h1 (3,2)=h1h3/h2 */
  double q[3] = {h[0] * h[2] / h[1], h[0] * h[1] / h[2], h[1] * h[2] / h[0]};
  static const int GI[3][2] = {{0, 1}, {1, 2}, {2, 0}}; /* index into q for g^2(a,k) */
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int t = 0; t < 4; t++) {
        int k = KL[t][0], l = KL[t][1];
        int ia = GI[a][k], ib = GI[b][l];
        double W;
        if (ia == ib) W = q[ia];
        else {
          /* sqrt(q_i q_j) is the h component shared by both products */
          /* q0*q1 = h1^2, q0*q2 = h3^2, q1*q2 = h2^2 */
          int other = 3 - ia - ib;
          W = (other == 2) ? h[0] : (other == 1 ? h[2] : h[1]);
        }
        double v = ((W * m[T6[a][b]]) * c1[a][k]) * c2[b][l];
        acc = (k == l) ? acc + v : acc - v;
      }
  return acc;
}

/* f2 (integration.f90:211-238) */
static double orc_f2(const orc_state *s, const orc_ws *w, int g, int im, int jm) {
  const double *m = w->mf2[g];
  const double *u = w->ve[g][im], *v = w->ve[g][jm];
  if (s->dirichlet)
    return (v[0] * m[0] * u[0] + v[1] * m[1] * u[0] + v[2] * m[2] * u[0]) +
           (v[0] * m[1] * u[1] + v[1] * m[3] * u[1] + v[2] * m[4] * u[1]) +
           (v[0] * m[2] * u[2] + v[1] * m[4] * u[2] + v[2] * m[5] * u[2]);
  double h3 = w->h[g][0] * w->h[g][1] * w->h[g][2];
  return h3 * v[0] * m[0] * u[0] + h3 * v[1] * m[1] * u[0] + h3 * v[2] * m[2] * u[0] +
         h3 * v[0] * m[1] * u[1] + h3 * v[1] * m[3] * u[1] + h3 * v[2] * m[4] * u[1] +
         h3 * v[0] * m[2] * u[2] + h3 * v[1] * m[4] * u[2] + h3 * v[2] * m[5] * u[2];
}

/* alocal (integration.f90:76-86) */
static zc orc_alocal(const orc_state *s, const orc_ws *w, int im, int jm) {
  double w32 = (double)(float)s->omega;
  double ar = 0.0, ai = 0.0;
  for (int g = 0; g < s->gpt; g++) {
    ar += w->wgt[g] * orc_f1(s, w, g, im, jm);
    ai += w->wgt[g] * (w32 * orc_f2(s, w, g, im, jm));
  }
  return CMPLX(ar, ai);
}

/* blocal (integration.f90:88-106) with f3 (:240-265) */
static void orc_blocal(const orc_state *s, const orc_ws *w, int im, zc out[2]) {
  for (int d = 0; d < 2; d++) {
    zc acc = 0.0;
    for (int g = 0; g < s->gpt; g++) {
      const double *v = w->ve[g][im];
      const zc *sr = w->src[d][g];
      zc f3;
      if (s->dirichlet) f3 = v[0] * sr[0] + v[1] * sr[1] + v[2] * sr[2];
      else {
        double h3 = w->h[g][0] * w->h[g][1] * w->h[g][2];
        f3 = (h3 * v[0]) * sr[0] + (h3 * v[1]) * sr[1] + (h3 * v[2]) * sr[2];
      }
      acc += w->wgt[g] * f3;
    }
    out[d] = acc;
  }
}

/* Element matrices for one element in reference loop position (ie,je,ke).
 * Emulates the call sequence of global_vfem: int_elem_params uses the GPML
 * flags left by the previous get_pml call; get_pml(ie,je,ke) then runs.
 * aloc: complex (me,me) row-major [im][jm]; bloc: complex (me,2) [im][d]. */
static orc_ws *g_ws = NULL;
int orc_element(orc_state *s, int ie, int je, int ke, double *aloc, double *bloc) {
  if (!g_ws) g_ws = (orc_ws *)malloc(sizeof(orc_ws));
  orc_ws *w = g_ws;
  int rc = orc_elem_params(s, w, ie, je, ke);
  if (!s->dirichlet) orc_get_pml(s, ie, je, ke);
  if (rc) return rc;
  int me = s->me;
  if (aloc)
    for (int im = 0; im < me; im++)
      for (int jm = 0; jm < me; jm++) {
        zc a = orc_alocal(s, w, im, jm);
        aloc[2 * (im * me + jm)] = creal(a);
        aloc[2 * (im * me + jm) + 1] = cimag(a);
      }
  if (bloc)
    for (int im = 0; im < me; im++) {
      zc b[2];
      orc_blocal(s, w, im, b);
      for (int d = 0; d < 2; d++) {
        bloc[2 * (im * 2 + d)] = creal(b[d]);
        bloc[2 * (im * 2 + d) + 1] = cimag(b[d]);
      }
    }
  return 0;
}

/* force the GPML flag state (the module variable) for tests */
void orc_set_in_pml(orc_state *s, int a, int b, int c) {
  s->in_pml[0] = a; s->in_pml[1] = b; s->in_pml[2] = c;
}

/* ----------------------- Dirichlet boundary values ----------------------- */
/* Complex arithmetic in the form the reference's Fortran evaluates it: a real
 * operand is promoted to (r, 0) and multiplied in full; the quotient is the
 * scaled formula of C99 Annex G.5.1 (divisor scaled by 2^-logb(max(|c|,|d|)),
 * an exact rescaling), which is what the reference build's compiler emits. */
static inline zc zr(double r) { return CMPLX(r, 0.0); }
static inline zc zmul(zc a, zc b) {
  return CMPLX(creal(a) * creal(b) - cimag(a) * cimag(b), creal(a) * cimag(b) + cimag(a) * creal(b));
}
static inline zc zadd(zc a, zc b) { return CMPLX(creal(a) + creal(b), cimag(a) + cimag(b)); }
static inline zc zsub(zc a, zc b) { return CMPLX(creal(a) - creal(b), cimag(a) - cimag(b)); }
static inline zc zneg(zc a) { return CMPLX(-creal(a), -cimag(a)); }
static zc zdiv(zc x, zc y) {
  double a = creal(x), b = cimag(x), c = creal(y), d = cimag(y);
  int k = 0;
  double lg = logb(fmax(fabs(c), fabs(d)));
  if (isfinite(lg)) {
    k = (int)lg;
    c = scalbn(c, -k);
    d = scalbn(d, -k);
  }
  double den = c * c + d * d;
  return CMPLX(scalbn((a * c + b * d) / den, -k), scalbn((b * c - a * d) / den, -k));
}

/* gamma = sqrt(i alpha), alpha = omega mu0 psigma with psigma = float32(sigma)
 * + i float32(eps omega): bd_setmodel's default-kind CMPLX keeps the real part,
 * bd_updatemodel replaces the imaginary one (boundary_conds.f90:47-49,411-428,
 * 452-454, 532-533) */
static zc bd_gamma(const orc_state *s, double sigma) {
  zc psig = CMPLX((double)(float)sigma, (double)(float)(ORC_EPS * s->omega));
  zc alpha = zmul(zr(s->omega * (4 * ORC_PI * 1.e-7)), psig);
  return csqrt(zmul(CMPLX(0.0, 1.0), alpha));
}

/* wait_recursion (boundary_conds.f90:521-553): surface transfer functions cz
 * from the bottom layer up, then the interface amplitudes ez downwards */
static void orc_wait(orc_state *s) {
  int nl = s->nl;
  zc g = bd_gamma(s, s->lsig[nl - 1]);
  s->cz[nl - 1] = zdiv(zr(1.0), g);
  for (int l = nl - 2; l >= 0; l--) {
    g = bd_gamma(s, s->lsig[l]);
    zc t = zmul(g, s->cz[l + 1]);
    zc r = zdiv(zsub(zr(1.0), t), zadd(zr(1.0), t));
    zc e = cexp(zneg(zmul(zmul(zr(2.0), g), zr(s->ldz[l]))));
    zc re = zmul(r, e);
    s->cz[l] = zdiv(zsub(zr(1.0), re), zmul(g, zadd(zr(1.0), re)));
  }
  s->ez[0] = zr(1.0);
  for (int l = 0; l < nl - 1; l++) {
    g = bd_gamma(s, s->lsig[l]);
    zc e = cexp(zneg(zmul(g, zr(s->ldz[l]))));
    zc p = zmul(zmul(s->ez[l], s->cz[l + 1]), zadd(zr(1.0), zmul(s->cz[l], g)));
    zc q = zmul(s->cz[l], zadd(zr(1.0), zmul(s->cz[l + 1], g)));
    s->ez[l + 1] = zdiv(zmul(e, p), q);
  }
}

/* ezl(0, z) (boundary_conds.f90:558-578): the layer test compares the depth
 * z >= 0 with the interface heights zl <= 0, so below the surface only z = 0
 * (layer 1, or the half-space when nl = 1) yields a non-zero value */
static zc orc_ezl0(const orc_state *s, double z) {
  int nl = s->nl;
  zc v = 0.0;
  for (int l = 0; l < nl - 1; l++) {
    if (z <= s->zl[l] && z > s->zl[l + 1]) {
      zc g = bd_gamma(s, s->lsig[l]);
      zc t = zmul(g, s->cz[l + 1]);
      zc r = zdiv(zsub(zr(1.0), t), zadd(zr(1.0), t));
      zc a1 = zadd(zr(1.0), zdiv(zr(1.0), zmul(s->cz[l], g)));
      zc a2 = zsub(zr(1.0), zmul(r, cexp(zneg(zmul(zmul(zr(2.0), g), zr(z - s->zl[l + 1]))))));
      zc a3 = cexp(zneg(zmul(g, zr(s->zl[l] - z))));
      v = zmul(zmul(zmul(zdiv(s->ez[l], zr(2.0)), a1), a2), a3);
    }
  }
  if (z <= s->zl[nl - 1]) {
    zc g = bd_gamma(s, s->lsig[nl - 1]);
    v = zmul(s->ez[nl - 1], cexp(zneg(zmul(g, zr(s->zl[nl - 1] - z)))));
  }
  return v;
}

/* p_pfields of the boundary module (boundary_conds.f90:436-516), E part for
 * pe_sch = 1: ep[0] = Ex of polarization 1, ep[1] = Ey of polarization 2 at a
 * node of height zn (every other component is zero).  Requires orc_wait for
 * inimod 3. */
static void orc_bd_pfields(const orc_state *s, double zn, zc ep[2]) {
  double z = (s->ztop - zn) / 1000.0;
  int air = zn > s->ztop;
  double cm = (double)(float)(-2.0 * s->omega * ORC_B0), cp = (double)(float)(2.0 * s->omega * ORC_B0);
  ep[0] = ep[1] = 0.0;
  if (s->inimod == 2) {
    zc sq = bd_gamma(s, air ? 0.0 : s->h_sigma);
    zc q = zdiv(zr(1.0), sq);
    zc fp = air ? zmul(q, zsub(zr(1.0), zmul(zr(z), sq))) : zmul(q, cexp(zneg(zmul(zr(z), sq))));
    ep[0] = zmul(CMPLX(0.0, cm), fp);
    ep[1] = zmul(CMPLX(0.0, cp), fp);
  } else if (s->inimod == 3) {
    if (air) {
      zc t = zsub(s->cz[0], zr(z));
      ep[0] = zmul(CMPLX(0.0, cp), t);
      ep[1] = zmul(CMPLX(0.0, cm), t);
    } else {
      zc e = orc_ezl0(s, z);
      ep[0] = zmul(zmul(CMPLX(0.0, cp), s->cz[0]), e);
      ep[1] = zmul(zmul(CMPLX(0.0, cm), s->cz[0]), e);
    }
  }
}

/* f_boundary (boundary_conds.f90:188-240) for every Dirichlet edge of the
 * element whose nodes w holds: faces 1,2,4,5 (x and y) take the tangential
 * component E_p . dr/dxi_d at the edge's basis node, faces 3,6 and inimod 1
 * are zero.  fb: [me][2]; entries of non-Dirichlet edges are left untouched. */
static void orc_f_boundary(const orc_state *s, const orc_ws *w, const int32_t *g, zc (*fb)[2]) {
  for (int jm = 0; jm < s->me; jm++) {
    if (g[jm] >= 0) continue;
    int code = -g[jm];
    fb[jm][0] = fb[jm][1] = 0.0;
    if (!s->dirichlet || (s->inimod != 2 && s->inimod != 3)) continue;
    if (code == 3 || code == 6) continue;
    int i = s->ed[jm][0], d = s->ed[jm][1] - 1;
    double x[3] = {NR27[i - 1][0], NR27[i - 1][1], NR27[i - 1][2]};
    double J[3][3], Ji[3][3], det;
    orc_jac(s, w->re, x, J, Ji, &det); /* only J is used (nf_jacobian, nf_dr_dxi) */
    zc ep[2];
    orc_bd_pfields(s, w->re[i - 1][2], ep);
    zc e1[3] = {ep[0], 0.0, 0.0}, e2[3] = {0.0, ep[1], 0.0};
    zc f1 = 0.0, f2 = 0.0;
    for (int c = 0; c < 3; c++) {
      f1 = c ? zadd(f1, zmul(e1[c], zr(J[d][c]))) : zmul(e1[c], zr(J[d][c]));
      f2 = c ? zadd(f2, zmul(e2[c], zr(J[d][c]))) : zmul(e2[c], zr(J[d][c]));
    }
    fb[jm][0] = f1;
    fb[jm][1] = f2;
  }
}

/* Boundary primary field at a node of height zn for the current omega:
 * out complex (2) = {Ex of polarization 1, Ey of polarization 2}. */
void orc_bd_field(orc_state *s, double zn, double *out) {
  if (s->inimod == 3) orc_wait(s);
  zc ep[2];
  orc_bd_pfields(s, zn, ep);
  for (int d = 0; d < 2; d++) {
    out[2 * d] = creal(ep[d]);
    out[2 * d + 1] = cimag(ep[d]);
  }
}

/* Dirichlet values of every boundary edge of element (ie,je,ke), 1-based,
 * after set_frequency: out complex (me, 2), zero for interior edges. */
int orc_dirichlet_values(orc_state *s, int ie, int je, int ke, double *out) {
  if (!s->gne) orc_build_gne(s);
  if (!g_ws) g_ws = (orc_ws *)malloc(sizeof(orc_ws));
  int rc = orc_elem_params(s, g_ws, ie, je, ke);
  if (rc) return rc;
  if (s->dirichlet && s->inimod == 3) orc_wait(s);
  int64_t e = ((int64_t)(ie - 1) * s->ny + (je - 1)) * s->nz + (ke - 1);
  zc fb[54][2];
  for (int jm = 0; jm < s->me; jm++) fb[jm][0] = fb[jm][1] = 0.0;
  orc_f_boundary(s, g_ws, s->gne + e * s->me, fb);
  for (int jm = 0; jm < s->me; jm++)
    for (int d = 0; d < 2; d++) {
      out[2 * (jm * 2 + d)] = creal(fb[jm][d]);
      out[2 * (jm * 2 + d) + 1] = cimag(fb[jm][d]);
    }
  return 0;
}

/* ------------------------------ assembly --------------------------------- */
/* stable argsort of key[perm[*]] (merge_sort/ga_merge, global_assembly.f90:1908-1962) */
static void stable_argsort(const int32_t *key, int64_t *perm, int64_t n) {
  int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  for (int64_t w = 1; w < n; w *= 2) {
    for (int64_t lo = 0; lo < n; lo += 2 * w) {
      int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      int64_t i = lo, j = mid, k = lo;
      while (i < mid && j < hi) tmp[k++] = (key[perm[i]] <= key[perm[j]]) ? perm[i++] : perm[j++];
      while (i < mid) tmp[k++] = perm[i++];
      while (j < hi) tmp[k++] = perm[j++];
    }
    memcpy(perm, tmp, sizeof(int64_t) * n);
  }
  free(tmp);
}

/* global_vfem + local_vfem (MoVFEM_3DMT.f90:167-263) with ga_assemble_nze,
 * assign_aij/assign_bi and ga_sort_sparse.  ia/ja are 1-based; a, b complex
 * interleaved; b has length 2*nne (polarization-major).  Returns 0 or <0. */
int orc_global_vfem(orc_state *s, int32_t *ia, int32_t *ja, double *a, double *b) {
  if (!s->gne) orc_build_gne(s);
  if (!s->nzindx) orc_build_nzindx(s);
  if (s->dirichlet && (s->inimod < 1 || s->inimod > 3)) return -4;
  if (s->dirichlet && s->inimod == 3) orc_wait(s);
  int me = s->me;
  int64_t nnze = s->nnze;
  int32_t nne = s->nne;
  /* ga_assemble_nze */
  for (int64_t i = 0; i < nnze; i++) ia[i] = ja[i] = -1;
  for (int64_t e = 0; e < s->ne; e++) {
    const int32_t *g = s->gne + e * me;
    for (int im = 0; im < me; im++) {
      if (g[im] < 0) continue;
      for (int jm = 0; jm < me; jm++) {
        if (g[jm] < 0) continue;
        int32_t p = s->nzindx[(e * me + im) * me + jm];
        if (p == -1) continue;
        if (ia[p - 1] == -1 && ja[p - 1] == -1) {
          ia[p - 1] = g[im];
          ja[p - 1] = g[jm];
        } else if (ia[p - 1] != g[im] || ja[p - 1] != g[jm]) {
          return -3; /* 'ia, ja index are not same' */
        }
      }
    }
  }
  zc *A = (zc *)calloc((size_t)nnze, sizeof(zc));
  zc *B = (zc *)calloc((size_t)2 * nne, sizeof(zc));
  double *aloc = (double *)malloc(sizeof(double) * 2 * me * me);
  double *bloc = (double *)malloc(sizeof(double) * 2 * me * 2);
  for (int ie = 1; ie <= s->nx; ie++)
    for (int je = 1; je <= s->ny; je++)
      for (int ke = 1; ke <= s->nz; ke++) {
        int64_t e = ((int64_t)(ie - 1) * s->ny + (je - 1)) * s->nz + (ke - 1);
        int rc = orc_element(s, ie, je, ke, aloc, bloc);
        if (rc) {
          free(A); free(B); free(aloc); free(bloc);
          return rc;
        }
        const int32_t *g = s->gne + e * me;
        for (int im = 0; im < me; im++) {
          if (g[im] < 0) continue;
          for (int jm = 0; jm < me; jm++) {
            if (g[jm] < 0) continue;
            if (g[im] < g[jm]) continue; /* sym: keep row >= col */
            int32_t p = s->nzindx[(e * me + im) * me + jm];
            if (p < 1 || ia[p - 1] != g[im] || ja[p - 1] != g[jm]) {
              free(A); free(B); free(aloc); free(bloc);
              return -3;
            }
            A[p - 1] += CMPLX(aloc[2 * (im * me + jm)], aloc[2 * (im * me + jm) + 1]);
          }
        }
        /* Dirichlet lift: b(im) = blocal(im) - sum_jm f_boundary(jm) alocal(im,jm)
         * over the element's Dirichlet edges jm (MoVFEM_3DMT.f90:251-261) */
        zc fb[54][2];
        orc_f_boundary(s, g_ws, g, fb);
        for (int im = 0; im < me; im++) {
          if (g[im] < 0) continue;
          zc bda[2] = {0.0, 0.0};
          for (int jm = 0; jm < me; jm++) {
            if (g[jm] >= 0) continue;
            zc al = CMPLX(aloc[2 * (im * me + jm)], aloc[2 * (im * me + jm) + 1]);
            for (int d = 0; d < 2; d++) bda[d] = zadd(bda[d], zmul(fb[jm][d], al));
          }
          for (int d = 0; d < 2; d++) {
            zc bi = CMPLX(bloc[2 * (im * 2 + d)], bloc[2 * (im * 2 + d) + 1]) - bda[d];
            B[(int64_t)d * nne + g[im] - 1] += bi;
          }
        }
      }
  free(aloc); free(bloc);
  /* ga_sort_sparse: stable by column (values permuted, complex64), then
   * stable by row (indices permuted, values left in column order) */
  int64_t *q = (int64_t *)malloc(sizeof(int64_t) * nnze);
  int32_t *ti = (int32_t *)malloc(sizeof(int32_t) * nnze), *tj = (int32_t *)malloc(sizeof(int32_t) * nnze);
  for (int64_t i = 0; i < nnze; i++) q[i] = i;
  stable_argsort(ja, q, nnze);
  zc *ta = (zc *)malloc(sizeof(zc) * nnze);
  for (int64_t i = 0; i < nnze; i++) {
    ti[i] = ia[q[i]];
    tj[i] = ja[q[i]];
    float complex c = (float complex)A[q[i]];
    ta[i] = (zc)c;
  }
  memcpy(ia, ti, sizeof(int32_t) * nnze);
  memcpy(ja, tj, sizeof(int32_t) * nnze);
  for (int64_t i = 0; i < nnze; i++) q[i] = i;
  stable_argsort(ia, q, nnze);
  for (int64_t i = 0; i < nnze; i++) {
    ti[i] = ia[q[i]];
    tj[i] = ja[q[i]];
  }
  memcpy(ia, ti, sizeof(int32_t) * nnze);
  memcpy(ja, tj, sizeof(int32_t) * nnze);
  for (int64_t i = 0; i < nnze; i++) {
    float complex c = (float complex)ta[i];
    a[2 * i] = crealf(c);
    a[2 * i + 1] = cimagf(c);
  }
  for (int64_t i = 0; i < 2 * (int64_t)nne; i++) {
    b[2 * i] = creal(B[i]);
    b[2 * i + 1] = cimag(B[i]);
  }
  free(q); free(ti); free(tj); free(ta); free(A); free(B);
  return 0;
}

/* ------------------------- PROBLEM.OUT fields ----------------------------- */
/* set_write_problem + write_problem_elem (problem.f90:459-507): the element
 * loop visits every element in reference order and writes, at each of its
 * nord^3 lattice points r, the interpolated primary fields get_ep/get_hp
 * (problem.f90:155-186) and the source p_source(r) (:91-128) of polarization
 * d, so the last element containing a node sets its values.  psigma is
 * pe_psigma(1) = (0, omega*eps) (pset_pmodel, :247-252), whose real part
 * PROBLEM.OUT prints.  ep/hp/src: complex (2*npt, 3), polarization-major. */
int orc_problem_fields(orc_state *s, double *ep, double *hp, double *src) {
  int o = s->nord - 1, mn = s->mn, nord = s->nord;
  int64_t npt = s->npt;
  double as[3] = {-1.0, 0.0, 1.0};
  if (nord == 2) as[1] = 1.0;
  orc_ws *w = (orc_ws *)malloc(sizeof(orc_ws));
  const double w32 = (double)(float)(s->omega);
  for (int ie = 1; ie <= s->nx; ie++)
    for (int je = 1; je <= s->ny; je++)
      for (int ke = 1; ke <= s->nz; ke++) {
        int64_t eno = (int64_t)(ie - 1) * s->nyz * o + (int64_t)(je - 1) * s->nnz * o + (int64_t)(ke - 1) * o;
        int rc = orc_elem_params(s, w, ie, je, ke);
        if (rc) {
          free(w);
          return rc;
        }
        for (int i1 = 0; i1 < nord; i1++)
          for (int j1 = 0; j1 < nord; j1++)
            for (int k1 = 0; k1 < nord; k1++) {
              int64_t id = eno + (int64_t)i1 * s->nyz + (int64_t)j1 * s->nnz + k1;
              double r[3] = {as[i1], as[j1], as[k1]};
              double J[3][3], Ji[3][3], det;
              orc_jac(s, w->re, r, J, Ji, &det);
              zc sr[2][27][3];
              orc_psource(s, w, r, Ji, w32, sr, 0);
              for (int d = 0; d < 2; d++) {
                zc e[3] = {0, 0, 0}, h[3] = {0, 0, 0};
                for (int n = 0; n < mn; n++) {
                  double N = nf_ln(s, n + 1, r);
                  for (int c = 0; c < 3; c++) {
                    e[c] += w->ep[d][n][c] * N;
                    h[c] += w->hp[d][n][c] * N;
                  }
                }
                int64_t idd = id + d * npt;
                for (int c = 0; c < 3; c++) {
                  ep[2 * (idd * 3 + c)] = creal(e[c]);
                  ep[2 * (idd * 3 + c) + 1] = cimag(e[c]);
                  hp[2 * (idd * 3 + c)] = creal(h[c]);
                  hp[2 * (idd * 3 + c) + 1] = cimag(h[c]);
                  src[2 * (idd * 3 + c)] = creal(sr[d][0][c]);
                  src[2 * (idd * 3 + c) + 1] = cimag(sr[d][0][c]);
                }
              }
            }
      }
  free(w);
  return 0;
}

/* ------------------------------ readout ---------------------------------- */
/* node_solution (solution.f90:18-49) with get_elem_sol/get_er/get_hr
 * (:207-438) for pe_sch = 1: total E at every lattice node (secondary from the
 * edge solution x plus the interpolated primary field), then H =
 * -i fl32(1/omega) mu^-1 curl E from the element's nodal E; the first element
 * (reference order) visiting a node sets its value.  x: complex 2*nne,
 * polarization-major.  esol/hsol: complex (2*npt, 3), polarization-major. */
int orc_node_solution(orc_state *s, const double *x, double *esol, double *hsol) {
  if (!s->gne) orc_build_gne(s);
  if (s->dirichlet && (s->inimod < 1 || s->inimod > 3)) return -4;
  if (s->dirichlet && s->inimod == 3) orc_wait(s);
  int o = s->nord - 1, mn = s->mn, me = s->me, nord = s->nord;
  int64_t npt = s->npt;
  char *ve = (char *)calloc((size_t)2 * npt, 1), *vh = (char *)calloc((size_t)2 * npt, 1);
  memset(esol, 0, sizeof(double) * 2 * 3 * 2 * npt);
  memset(hsol, 0, sizeof(double) * 2 * 3 * 2 * npt);
  double as[3] = {-1.0, 0.0, 1.0};
  if (nord == 2) as[1] = 1.0;
  orc_ws *w = (orc_ws *)malloc(sizeof(orc_ws));
  const zc mi = CMPLX(0.0, (double)(float)(-1.0 / s->omega));
  for (int ie = 1; ie <= s->nx; ie++)
    for (int je = 1; je <= s->ny; je++)
      for (int ke = 1; ke <= s->nz; ke++) {
        int64_t ide = ((int64_t)(ie - 1) * s->ny + (je - 1)) * s->nz + (ke - 1);
        int64_t eno = (int64_t)(ie - 1) * s->nyz * o + (int64_t)(je - 1) * s->nnz * o + (int64_t)(ke - 1) * o;
        /* nf_get_r + p_elem_fields: node coordinates and primary fields */
        int rc = orc_elem_params(s, w, ie, je, ke);
        if (rc) {
          free(ve); free(vh); free(w);
          return rc;
        }
        const int32_t *g = s->gne + ide * me;
        zc fb[54][2];
        orc_f_boundary(s, w, g, fb);
        for (int d = 0; d < 2; d++) {
          for (int i1 = 0; i1 < nord; i1++)
            for (int j1 = 0; j1 < nord; j1++)
              for (int k1 = 0; k1 < nord; k1++) {
                int64_t id = eno + (int64_t)i1 * s->nyz + (int64_t)j1 * s->nnz + k1;
                int64_t idd = id + d * npt;
                if (ve[idd]) continue;
                double r[3] = {as[i1], as[j1], as[k1]};
                double J[3][3], Ji[3][3], det;
                orc_jac(s, w->re, r, J, Ji, &det);
                zc ef[3] = {0, 0, 0}, ep[3] = {0, 0, 0};
                for (int im = 0; im < me; im++) {
                  zc f = fb[im][d]; /* Dirichlet edge: f_boundary (solution.f90:327-329) */
                  if (g[im] >= 0) f = CMPLX(x[2 * ((int64_t)g[im] - 1 + d * (int64_t)s->nne)],
                                            x[2 * ((int64_t)g[im] - 1 + d * (int64_t)s->nne) + 1]);
                  int node = s->ed[im][0], dir = s->ed[im][1] - 1;
                  double phi = mix_ln(s, node, dir, r);
                  for (int c = 0; c < 3; c++) ef[c] += f * (phi * Ji[c][dir]);
                }
                for (int n = 0; n < mn; n++) {
                  double N = nf_ln(s, n + 1, r);
                  for (int c = 0; c < 3; c++) ep[c] += w->ep[d][n][c] * N;
                }
                for (int c = 0; c < 3; c++) {
                  zc v = ef[c] + ep[c];
                  esol[2 * (idd * 3 + c)] = creal(v);
                  esol[2 * (idd * 3 + c) + 1] = cimag(v);
                }
                ve[idd] = 1;
              }
          for (int i1 = 0; i1 < nord; i1++)
            for (int j1 = 0; j1 < nord; j1++)
              for (int k1 = 0; k1 < nord; k1++) {
                int64_t id = eno + (int64_t)i1 * s->nyz + (int64_t)j1 * s->nnz + k1;
                int64_t idd = id + d * npt;
                if (vh[idd]) continue;
                double r[3] = {as[i1], as[j1], as[k1]};
                double J[3][3], Ji[3][3], det;
                orc_jac(s, w->re, r, J, Ji, &det);
                zc ef[3] = {0, 0, 0};
                for (int n = 0; n < mn; n++) {
                  int64_t nd = w->idx[n] + d * npt;
                  zc f2[3];
                  for (int c = 0; c < 3; c++) f2[c] = CMPLX(esol[2 * (nd * 3 + c)], esol[2 * (nd * 3 + c) + 1]);
                  double dn[3];
                  for (int a = 0; a < 3; a++) {
                    double acc = 0.0;
                    for (int b = 0; b < 3; b++) acc += Ji[a][b] * nf_dln(s, b, n + 1, r);
                    dn[a] = acc;
                  }
                  ef[0] += (f2[2] * dn[1] - f2[1] * dn[2]);
                  ef[1] += (f2[0] * dn[2] - f2[2] * dn[0]);
                  ef[2] += (f2[1] * dn[0] - f2[0] * dn[1]);
                }
                double m1[6];
                for (int c = 0; c < 6; c++) {
                  double a1 = 0.0;
                  for (int n = 0; n < mn; n++) a1 += nf_ln(s, n + 1, r) * w->inmu[n][c];
                  m1[c] = a1;
                }
                zc md[6];
                for (int c = 0; c < 6; c++) md[c] = mi * m1[c];
                zc hv[3] = {md[0] * ef[0] + md[1] * ef[1] + md[2] * ef[2], md[1] * ef[0] + md[3] * ef[1] + md[4] * ef[2],
                            md[2] * ef[0] + md[4] * ef[1] + md[5] * ef[2]};
                for (int c = 0; c < 3; c++) {
                  hsol[2 * (idd * 3 + c)] = creal(hv[c]);
                  hsol[2 * (idd * 3 + c) + 1] = cimag(hv[c]);
                }
                vh[idd] = 1;
              }
        }
      }
  free(ve); free(vh); free(w);
  return 0;
}

/* get_impedance + inv_matrix + get_res_phase (solution.f90:456-505):
 * e, h = (x pol1, x pol2, y pol1, y pol2) */
void orc_impedance(const double *e_, const double *h_, double omega, double *z_, double *rho, double *phi) {
  zc e[4], h[4], ih[4], z[4];
  for (int i = 0; i < 4; i++) {
    e[i] = CMPLX(e_[2 * i], e_[2 * i + 1]);
    h[i] = CMPLX(h_[2 * i], h_[2 * i + 1]);
  }
  /* inv_matrix, get_impedance (solution.f90:456-469,493-505) in the
   * reference build's complex arithmetic */
  zc det = zsub(zmul(h[0], h[3]), zmul(h[1], h[2]));
  ih[0] = zdiv(h[3], det);
  ih[1] = zneg(zdiv(h[1], det));
  ih[2] = zneg(zdiv(h[2], det));
  ih[3] = zdiv(h[0], det);
  z[0] = zadd(zmul(ih[0], e[0]), zmul(ih[2], e[1]));
  z[1] = zadd(zmul(ih[1], e[0]), zmul(ih[3], e[1]));
  z[2] = zadd(zmul(ih[0], e[2]), zmul(ih[2], e[3]));
  z[3] = zadd(zmul(ih[1], e[2]), zmul(ih[3], e[3]));
  double mu0 = 4 * ORC_PI * 1.e-7;
  for (int i = 0; i < 4; i++) {
    z_[2 * i] = creal(z[i]);
    z_[2 * i + 1] = cimag(z[i]);
    rho[i] = (1.0 / (omega * mu0)) * (creal(z[i]) * creal(z[i]) + cimag(z[i]) * cimag(z[i]));
    if (rho[i] < 1e-2) {
      rho[i] = 0.0;
      phi[i] = 0.0;
    } else {
      phi[i] = atan(cimag(z[i]) / creal(z[i])) * (180.0 / ORC_PI);
    }
  }
}

/* rx_data with get_rx_element and interpolate_rx (receiver_data.f90:111-253):
 * bilinear interpolation of the nodal E and H on the top face of element
 * (ie, je, ke_top) containing (x, y), then impedance, rho_a and phase.
 * ie runs over nextd+1 .. nx-nextd (the reference's inner region).
 * out_e/out_h: complex (nrx, 2 pol, 3); out_z complex (nrx, 4); rho/phi (nrx, 4). */
int orc_receivers(orc_state *s, const double *esol, const double *hsol, int ke_top, int nrx, const double *rx,
                  const double *ry, double *out_e, double *out_h, double *out_z, double *rho, double *phi) {
  int o = s->nord - 1;
  /* get_rx_element (receiver_data.f90:186-209): y is searched only inside the
   * x cell that holds the site; an index not found keeps the previous
   * receiver's value (rx_data's loop variables, :121-124), and the site is
   * then extrapolated from that element. */
  int ie = -1, je = -1;
  for (int ir = 0; ir < nrx; ir++) {
    for (int a = s->nextd + 1; a <= s->nx - s->nextd; a++)
      if (rx[ir] >= s->xp[(a - 1) * o] && rx[ir] < s->xp[a * o]) {
        ie = a;
        for (int b = s->nextd + 1; b <= s->ny - s->nextd; b++)
          if (ry[ir] >= s->yp[(b - 1) * o] && ry[ir] < s->yp[b * o]) {
            je = b;
            break;
          }
      }
    if (ie < 0 || je < 0) return -5; /* no earlier receiver to inherit from */
    int ke = ke_top;
    int64_t eno = (int64_t)(ie - 1) * o * s->nyz + (int64_t)(je - 1) * o * s->nnz + (int64_t)(ke - 1) * o;
    double x1 = s->xp[(ie - 1) * o], x2 = s->xp[(ie - 1) * o + o];
    double y1 = s->yp[(je - 1) * o], y2 = s->yp[(je - 1) * o + o];
    int64_t id11 = eno + o, id12 = eno + (int64_t)o * s->nnz + o, id21 = eno + (int64_t)o * s->nyz + o,
            id22 = eno + (int64_t)o * s->nyz + (int64_t)o * s->nnz + o;
    double x = rx[ir], y = ry[ir];
    double ax = (x2 - x) / (x2 - x1), bx = (x - x1) / (x2 - x1);
    double ay = (y2 - y) / (y2 - y1), by = (y - y1) / (y2 - y1);
    double e4[8], h4[8];
    for (int d = 0; d < 2; d++) {
      for (int c = 0; c < 3; c++) {
        for (int q = 0; q < 2; q++) {
          const double *S[2] = {esol, hsol};
          for (int t = 0; t < 2; t++) {
            const double *F = S[t];
            double f11 = F[2 * ((id11 + d * s->npt) * 3 + c) + q], f12 = F[2 * ((id12 + d * s->npt) * 3 + c) + q];
            double f21 = F[2 * ((id21 + d * s->npt) * 3 + c) + q], f22 = F[2 * ((id22 + d * s->npt) * 3 + c) + q];
            double fx1 = ax * f11 + bx * f21, fx2 = ax * f12 + bx * f22;
            double v = ay * fx1 + by * fx2;
            double *O = t == 0 ? out_e : out_h;
            O[2 * ((ir * 2 + d) * 3 + c) + q] = v;
          }
        }
      }
    }
    /* assign_rxmatrix: (x pol1, x pol2, y pol1, y pol2) */
    for (int q = 0; q < 2; q++) {
      e4[0 + q] = out_e[2 * ((ir * 2 + 0) * 3 + 0) + q];
      e4[2 + q] = out_e[2 * ((ir * 2 + 1) * 3 + 0) + q];
      e4[4 + q] = out_e[2 * ((ir * 2 + 0) * 3 + 1) + q];
      e4[6 + q] = out_e[2 * ((ir * 2 + 1) * 3 + 1) + q];
      h4[0 + q] = out_h[2 * ((ir * 2 + 0) * 3 + 0) + q];
      h4[2 + q] = out_h[2 * ((ir * 2 + 1) * 3 + 0) + q];
      h4[4 + q] = out_h[2 * ((ir * 2 + 0) * 3 + 1) + q];
      h4[6 + q] = out_h[2 * ((ir * 2 + 1) * 3 + 1) + q];
    }
    orc_impedance(e4, h4, s->omega, out_z + 8 * ir, rho + 4 * ir, phi + 4 * ir);
  }
  return 0;
}

/* find_zeros + rem_zeros (global_assembly.f90:123-150): compacts in place,
 * returns the new length */
int64_t orc_remove_zeros(int64_t n, int32_t *ia, int32_t *ja, double *a) {
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++) {
    if (a[2 * i] == 0.0 && a[2 * i + 1] == 0.0) continue;
    ia[k] = ia[i];
    ja[k] = ja[i];
    a[2 * k] = a[2 * i];
    a[2 * k + 1] = a[2 * i + 1];
    k++;
  }
  return k;
}

/* ---------------------- timed element loop (CPU baseline) ---------------------
 * The element loop of global_vfem (MoVFEM_3DMT.f90:193-211) for elements
 * [e0, e1) in reference loop order: int_elem_params, then alocal for every
 * pair im <= jm and blocal for every edge (integration.f90:60-106), on
 * nthreads OpenMP threads.  Each element takes the GPML zone flags of the
 * element before it, as get_pml after int_elem_params leaves them
 * (MoVFEM_3DMT.f90:203-206; element 0 takes none).  Every thread has its own
 * workspace and a copy of the state's scalars, so the results do not depend
 * on the thread count.  out[0] = sum of Re alocal, out[1] = sum of Im alocal,
 * out[2] = sum of |blocal| over the range (a checksum that keeps the work
 * observable), out[3] = the thread count used.  Returns 0 or the first
 * element's error code. */
int orc_element_loop_timed(orc_state *s, int64_t e0, int64_t e1, int nthreads, double *out) {
  if (nthreads < 1) nthreads = 1;
  int me = s->me, rc_all = 0;
  double sr = 0.0, si = 0.0, sb = 0.0;
  int used = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : sr, si, sb)
  {
#pragma omp single
    used = omp_get_num_threads();
    orc_state st = *s;
    orc_ws *w = (orc_ws *)malloc(sizeof(orc_ws));
#pragma omp for schedule(dynamic, 16)
    for (int64_t e = e0; e < e1; e++) {
      int ie = (int)(e / ((int64_t)s->ny * s->nz)) + 1;
      int je = (int)((e / s->nz) % s->ny) + 1;
      int ke = (int)(e % s->nz) + 1;
      st.in_pml[0] = st.in_pml[1] = st.in_pml[2] = 0;
      if (e > 0 && !s->dirichlet) {
        int64_t p = e - 1;
        orc_get_pml(&st, (int)(p / ((int64_t)s->ny * s->nz)) + 1, (int)((p / s->nz) % s->ny) + 1,
                    (int)(p % s->nz) + 1);
      }
      int rc = orc_elem_params(&st, w, ie, je, ke);
      if (rc) {
#pragma omp critical
        if (!rc_all) rc_all = rc;
        continue;
      }
      for (int im = 0; im < me; im++) {
        for (int jm = im; jm < me; jm++) {
          zc a = orc_alocal(&st, w, im, jm);
          sr += creal(a);
          si += cimag(a);
        }
        zc b[2];
        orc_blocal(&st, w, im, b);
        sb += cabs(b[0]) + cabs(b[1]);
      }
    }
    free(w);
  }
  if (out) {
    out[0] = sr;
    out[1] = si;
    out[2] = sb;
    out[3] = used;
  }
  return rc_all;
}
