// Reference-element data for the 8-, 20- and 27-node hexahedra.
//
// Tables that must agree with MoVFEM_3DMT (kept alike on purpose, 1-based
// entries as in the reference):
//   node reference coordinates      n_fem.f90:113-133
//   edge DOF (basis node, direction) v_fem.f90:491-504
//   neighbour copy maps (-z,-y,-x)   global_assembly.f90:242-265, 310-351, 396-443;
//   the same lists are the Dirichlet face lists of boundary_conds.f90:276-388
// Gauss-Legendre points/weights are the float32-rounded literals of
// integration.f90:297-299,400-402.
#pragma once
#include <cmath>
#include <cstdint>

namespace mvf {

static const signed char kNodeRef[27][3] = {
    {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, 1}, {1, 1, 1},  {-1, 1, 1},
    {-1, -1, 1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1},  {0, -1, -1}, {1, 0, 1}, {0, 1, 1},
    {-1, 0, 1},  {0, -1, 1}, {1, -1, 0},  {1, 1, 0},    {-1, 1, 0}, {-1, -1, 0}, {1, 0, 0},
    {0, 1, 0},   {-1, 0, 0}, {0, -1, 0},  {0, 0, -1},   {0, 0, 1},  {0, 0, 0}};

static const unsigned char kEdge12[12][2] = {{4, 3}, {4, 2}, {8, 2}, {3, 3}, {4, 1}, {8, 1},
                                             {3, 1}, {7, 1}, {1, 3}, {1, 2}, {5, 2}, {2, 3}};
static const unsigned char kEdge36[36][2] = {
    {4, 3},  {8, 3},  {4, 2},  {3, 2},  {8, 2}, {7, 2},  {3, 3},  {7, 3},  {4, 1},  {1, 1},  {8, 1},  {5, 1},
    {3, 1},  {2, 1},  {7, 1},  {6, 1},  {1, 3}, {5, 3},  {1, 2},  {2, 2},  {5, 2},  {6, 2},  {2, 3},  {6, 3},
    {12, 3}, {20, 2}, {12, 2}, {16, 2}, {10, 3}, {17, 2}, {11, 3}, {20, 1}, {11, 1}, {15, 1}, {19, 1}, {9, 3}};
static const unsigned char kEdge54[54][2] = {
    {4, 3},  {8, 3},  {4, 2},  {20, 2}, {8, 2},  {11, 3}, {15, 3}, {3, 2},  {19, 2}, {7, 2},  {3, 3},
    {7, 3},  {4, 1},  {20, 1}, {8, 1},  {11, 1}, {15, 1}, {3, 1},  {19, 1}, {7, 1},  {12, 3}, {16, 3},
    {12, 2}, {10, 2}, {16, 2}, {14, 2}, {10, 3}, {14, 3}, {1, 1},  {17, 1}, {5, 1},  {9, 1},  {13, 1},
    {2, 1},  {18, 1}, {6, 1},  {1, 3},  {5, 3},  {1, 2},  {17, 2}, {5, 2},  {9, 3},  {13, 3}, {2, 2},
    {18, 2}, {6, 2},  {2, 3},  {6, 3},  {25, 3}, {26, 3}, {24, 2}, {22, 2}, {23, 1}, {21, 1}};

// copy maps, axis order z, y, x: {dst in this element, src in the -axis neighbour}
static const unsigned char kCopy12[3][4][2] = {{{2, 3}, {5, 6}, {7, 8}, {10, 11}},
                                               {{1, 4}, {5, 7}, {6, 8}, {9, 12}},
                                               {{1, 9}, {2, 10}, {3, 11}, {4, 12}}};
static const unsigned char kCopy36[3][10][2] = {
    {{3, 5}, {4, 6}, {9, 11}, {10, 12}, {13, 15}, {14, 16}, {19, 21}, {20, 22}, {27, 28}, {33, 34}},
    {{1, 7}, {2, 8}, {9, 13}, {10, 14}, {11, 15}, {12, 16}, {17, 23}, {18, 24}, {25, 29}, {32, 35}},
    {{1, 17}, {2, 18}, {3, 19}, {4, 20}, {5, 21}, {6, 22}, {7, 23}, {8, 24}, {26, 30}, {31, 36}}};
static const unsigned char kCopy54[3][12][2] = {
    {{3, 5}, {8, 10}, {13, 15}, {16, 17}, {18, 20}, {23, 25}, {24, 26}, {29, 31}, {32, 33}, {34, 36}, {39, 41},
     {44, 46}},
    {{1, 11}, {2, 12}, {13, 18}, {14, 19}, {15, 20}, {21, 27}, {22, 28}, {29, 34}, {30, 35}, {31, 36}, {37, 47},
     {38, 48}},
    {{1, 37}, {2, 38}, {3, 39}, {4, 40}, {5, 41}, {6, 42}, {7, 43}, {8, 44}, {9, 45}, {10, 46}, {11, 47},
     {12, 48}}};

struct ElemType {
  int nn, ne, ng1, ng, nf, nord;  // nodes, edge DOFs, GP per axis, GP, DOFs per face, lattice order
  const unsigned char (*edge)[2];
  const unsigned char *copy;  // [3][nf][2]
};

inline ElemType elem_type(int nodes) {
  ElemType t{};
  t.nn = nodes;
  if (nodes == 8) {
    t.ne = 12; t.ng1 = 2; t.nf = 4; t.nord = 2; t.edge = kEdge12; t.copy = &kCopy12[0][0][0];
  } else if (nodes == 20) {
    t.ne = 36; t.ng1 = 3; t.nf = 10; t.nord = 3; t.edge = kEdge36; t.copy = &kCopy36[0][0][0];
  } else {
    t.ne = 54; t.ng1 = 3; t.nf = 12; t.nord = 3; t.edge = kEdge54; t.copy = &kCopy54[0][0][0];
  }
  t.ng = t.ng1 * t.ng1 * t.ng1;
  return t;
}

// float32-rounded Gauss-Legendre rule, point index with the last axis fastest
inline void gauss_rule(int n1, double pts[27][3], double w[27]) {
  double p[3], q[3];
  if (n1 == 2) {
    p[0] = (double)-0.5773502691896257645f; p[1] = (double)0.5773502691896257645f;
    q[0] = 1.0; q[1] = 1.0;
  } else {
    p[0] = (double)-0.7745966692414833770f; p[1] = 0.0; p[2] = (double)0.7745966692414833770f;
    q[0] = (double)0.5555555555555555556f; q[1] = (double)0.8888888888888888889f;
    q[2] = (double)0.5555555555555555556f;
  }
  int id = 0;
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n1; k++, id++) {
        pts[id][0] = p[i]; pts[id][1] = p[j]; pts[id][2] = p[k];
        w[id] = q[i] * q[j] * q[k];
      }
}

// ---------------------------------------------------------------------------
// Basis functions on the reference cube.  One-dimensional factor form:
//   8-node  nodal:  prod_c (1 + r_c x_c)/2
//   20-node nodal:  serendipity (corner: prod(1+r x) (sum r x - 2)/8,
//                   mid-edge: (1-x_0^2) prod_other(1+r x)/4)
//   27-node nodal:  prod_c L_c(x_c), L = (1+r x) r x / 2 for r=+-1, 1-x^2 for r=0
//   edge DOF (node n, direction d): 8-node  prod_{c!=d} (1+r_c x_c)/2
//                   20-node corner  prod_{c!=d}(1+r_c x_c) (sum_c r_c x_c - 1)/8
//                   20-node mid     (1-x_z^2)(1+r_o x_o)/2 (z the zero coordinate)
//                   27-node         (1+r_d x_d)/2 * prod_{c!=d} L_c(x_c)
// The 8-node d/dzeta of the nodal functions keeps the reference's form
// (1+r_1)(1+r_2 eta) r_3/8 (n_fem.f90:193), which only matters when an
// element's z varies with x.
// ---------------------------------------------------------------------------
inline double lag1(int r, double x) { return r == 0 ? 1.0 - x * x : (1.0 + r * x) * (r * x) * 0.5; }
inline double dlag1(int r, double x) { return r == 0 ? -2.0 * x : r * (1.0 + 2.0 * r * x) * 0.5; }

inline double nodal(int nodes, int n, const double *x) {
  const signed char *r = kNodeRef[n];
  if (nodes == 8) return (1 + r[0] * x[0]) * (1 + r[1] * x[1]) * (1 + r[2] * x[2]) / 8.0;
  if (nodes == 20) {
    if (n < 8)
      return (1 + r[0] * x[0]) * (1 + r[1] * x[1]) * (1 + r[2] * x[2]) * (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] - 2) /
             8.0;
    double v = 1.0;
    for (int c = 0; c < 3; c++) v *= (r[c] == 0) ? (1 - x[c] * x[c]) : (1 + r[c] * x[c]);
    return v / 4.0;
  }
  return lag1(r[0], x[0]) * lag1(r[1], x[1]) * lag1(r[2], x[2]);
}

inline double dnodal(int nodes, int n, int d, const double *x) {
  const signed char *r = kNodeRef[n];
  int a = (d + 1) % 3, b = (d + 2) % 3;
  if (nodes == 8) {
    if (d == 2) return (1 + r[0]) * (1 + r[1] * x[1]) * r[2] / 8.0;
    return r[d] * (1 + r[a] * x[a]) * (1 + r[b] * x[b]) / 8.0;
  }
  if (nodes == 20) {
    if (n < 8)
      return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) * r[d] * (2 * r[d] * x[d] + r[a] * x[a] + r[b] * x[b] - 1) / 8.0;
    if (r[d] == 0) return -x[d] * (1 + r[a] * x[a]) * (1 + r[b] * x[b]) / 2.0;
    int z = (r[a] == 0) ? a : b, o = (r[a] == 0) ? b : a;
    return r[d] * (1 - x[z] * x[z]) * (1 + r[o] * x[o]) / 4.0;
  }
  double v = 1.0;
  for (int c = 0; c < 3; c++) v *= (c == d) ? dlag1(r[c], x[c]) : lag1(r[c], x[c]);
  return v;
}

// edge polynomial for DOF (node n 0-based, direction d 0-based)
inline double edgepoly(int nodes, int n, int d, const double *x) {
  const signed char *r = kNodeRef[n];
  int a = (d + 1) % 3, b = (d + 2) % 3;
  if (nodes == 8) return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) / 4.0;
  if (nodes == 20) {
    if (n < 8) return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) * (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] - 1) / 8.0;
    int z = (r[a] == 0) ? a : b, o = (r[a] == 0) ? b : a;
    return (1 - x[z] * x[z]) * (1 + r[o] * x[o]) / 2.0;
  }
  return (1 + r[d] * x[d]) * 0.5 * lag1(r[a], x[a]) * lag1(r[b], x[b]);
}

inline double dedgepoly(int nodes, int n, int d, int k, const double *x) {
  const signed char *r = kNodeRef[n];
  int a = (d + 1) % 3, b = (d + 2) % 3;
  if (nodes == 8) {
    if (k == d) return 0.0;
    int o = (k == a) ? b : a;
    return r[k] * (1 + r[o] * x[o]) / 4.0;
  }
  if (nodes == 20) {
    if (n < 8) {
      if (k == d) return (1 + r[a] * x[a]) * (1 + r[b] * x[b]) * r[d] / 8.0;
      int o = (k == a) ? b : a;
      return r[k] * (1 + r[o] * x[o]) * (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] + r[k] * x[k]) / 8.0;
    }
    if (k == d) return 0.0;
    int z = (r[a] == 0) ? a : b, o = (r[a] == 0) ? b : a;
    if (k == z) return -x[z] * (1 + r[o] * x[o]);
    return (1 - x[z] * x[z]) * r[o] / 2.0;
  }
  double fd = (k == d) ? r[d] * 0.5 : (1 + r[d] * x[d]) * 0.5;
  double fa = (k == a) ? dlag1(r[a], x[a]) : lag1(r[a], x[a]);
  double fb = (k == b) ? dlag1(r[b], x[b]) : lag1(r[b], x[b]);
  return fd * fa * fb;
}

// Tabulated basis at the Gauss points; uploaded once per handle.
struct BasisTab {
  // sizes fixed for the largest element (27 nodes, 54 DOFs, 27 points)
  double w[27];            // Gauss weights
  double N[27][27];        // [g][n] nodal values
  double dN[27][3][27];    // [g][d][n]
  double phi[27][54];      // [g][e] edge polynomial
  double dphi[27][54][3];  // [g][e][k]
  double pts[27][3];
  int32_t dir[54];         // 0-based direction of each DOF
  int32_t node[54];        // 0-based basis node of each DOF
  signed char nref[27][3]; // node reference coordinates
  unsigned char pi[1485], pj[1485];  // upper-packed pair p -> (i, j), row-major i <= j
};

// Basis tabulated at npts given points (w = 0); fill_basis below uses the
// Gauss rule, the readout uses the element's lattice nodes.
inline void fill_basis_points(int nodes, const double (*pts)[3], int npts, const double *w, BasisTab *t);

inline void fill_basis(int nodes, BasisTab *t) {
  ElemType et = elem_type(nodes);
  double w[27], pts[27][3];
  gauss_rule(et.ng1, pts, w);
  fill_basis_points(nodes, pts, et.ng, w, t);
}

inline void fill_basis_points(int nodes, const double (*pts)[3], int npts, const double *w, BasisTab *t) {
  ElemType et = elem_type(nodes);
  for (int g = 0; g < npts; g++)
    for (int c = 0; c < 3; c++) t->pts[g][c] = pts[g][c];
  for (int g = 0; g < npts; g++) {
    t->w[g] = w ? w[g] : 0.0;
    for (int n = 0; n < et.nn; n++) {
      t->N[g][n] = nodal(nodes, n, t->pts[g]);
      for (int d = 0; d < 3; d++) t->dN[g][d][n] = dnodal(nodes, n, d, t->pts[g]);
    }
    for (int e = 0; e < et.ne; e++) {
      int n = et.edge[e][0] - 1, d = et.edge[e][1] - 1;
      t->phi[g][e] = edgepoly(nodes, n, d, t->pts[g]);
      for (int k = 0; k < 3; k++) t->dphi[g][e][k] = dedgepoly(nodes, n, d, k, t->pts[g]);
    }
  }
  for (int e = 0; e < et.ne; e++) {
    t->node[e] = et.edge[e][0] - 1;
    t->dir[e] = et.edge[e][1] - 1;
  }
  for (int n = 0; n < 27; n++)
    for (int c = 0; c < 3; c++) t->nref[n][c] = kNodeRef[n][c];
  int p = 0;
  for (int i = 0; i < et.ne; i++)
    for (int j = i; j < et.ne; j++, p++) {
      t->pi[p] = (unsigned char)i;
      t->pj[p] = (unsigned char)j;
    }
}

}  // namespace mvf
