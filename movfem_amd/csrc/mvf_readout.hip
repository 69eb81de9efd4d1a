// Readout on the device: total E and H at every lattice node (node_solution
// with get_elem_sol / get_er / get_hr, solution.f90:18-49,207-438, pe_sch = 1)
// and the receivers' fields, impedance, apparent resistivity and phase
// (rx_data / interpolate_rx, receiver_data.f90:111-253; get_impedance,
// get_res_phase, solution.f90:456-505).
//
// The reference visits the elements in loop order and the first element that
// contains a lattice node sets the node's E, then its H (curl of the nodal E
// of that element).  The first element containing lattice node a along an
// axis of order o is max(1, ceil(a / o)), independently per axis, so both
// passes are one thread per (node, polarization): pass 1 evaluates E in the
// owner element, pass 2 evaluates H in the owner element from pass 1's E.
#include "mvf_internal.h"

namespace mvf {

__device__ inline int owner_axis(int a, int o, int n) {
  int e = a == 0 ? 1 : (a + o - 1) / o;
  return e > n ? n : e;
}

// node coordinates of element (ie, je, ke) and the lattice point index of
// node id inside it
template <int NN>
__device__ inline int owner_setup(const BasisTab *tab, const Geom &g, const double *xp, const double *yp,
                                  const double *zp, int64_t id, int &ie, int &je, int &ke, double (*xs)[3],
                                  int64_t *nid) {
  const int o = g.nord - 1;
  const int a = (int)(id / g.nyz), b = (int)((id / g.nnz) % g.nny), c = (int)(id % g.nnz);
  ie = owner_axis(a, o, g.nx);
  je = owner_axis(b, o, g.ny);
  ke = owner_axis(c, o, g.nz);
  const int i1 = a - (ie - 1) * o, j1 = b - (je - 1) * o, k1 = c - (ke - 1) * o;
  for (int n = 0; n < NN; n++) {
    int64_t q;
    int ix, jy;
    load_node(tab, g, ie, je, ke, n, q, ix, jy);
    xs[n][0] = xp[ix];
    xs[n][1] = yp[jy];
    xs[n][2] = zp[q];
    nid[n] = q;
  }
  return (i1 * g.nord + j1) * g.nord + k1;
}

// pass 1: esol[(d*npt + id)*3 + c] = secondary E from the edge solution plus
// the interpolated primary field (p_pfields, problem.f90:340-358)
template <int NN, int NE>
__global__ __launch_bounds__(256) void k_node_e(const BasisTab *ntab, Geom g, const double *xp, const double *yp,
                                                const double *zp, const int32_t *gne, const double2 *x, int ncol,
                                                int col0, double omega, const int32_t *bidx, const double2 *dirf,
                                                double2 *esol) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 2 * g.npt) return;
  const int d = (int)(t / g.npt);
  const int64_t id = t % g.npt;
  int ie, je, ke;
  double xs[NN][3];
  int64_t nid[NN];
  const int p = owner_setup<NN>(ntab, g, xp, yp, zp, id, ie, je, ke, xs, nid);
  double Ji[9];
  jacobian<NN>(ntab, p, xs, Ji);
  const int64_t e = ((int64_t)(ie - 1) * g.ny + (je - 1)) * g.nz + (ke - 1);
  double er[3] = {0, 0, 0}, ei[3] = {0, 0, 0};
  // Dirichlet edges carry f_boundary (solution.f90:327-329): zero for
  // inimod 1, the boundary values of the element's row of dirf otherwise
  const int32_t be = bidx ? bidx[e] : -1;
  for (int im = 0; im < NE; im++) {
    const int32_t gi = gne[e * NE + im];
    if (gi == 0 || (gi < 0 && be < 0)) continue;
    const double2 f = gi > 0 ? x[(int64_t)(gi - 1) * ncol + col0 + d] : dirf[((int64_t)be * NE + im) * 2 + d];
    const double ph = ntab->phi[p][im];
    const int dir = ntab->dir[im];
    for (int c = 0; c < 3; c++) {
      const double v = ph * Ji[3 * c + dir];
      er[c] += f.x * v;
      ei[c] += f.y * v;
    }
  }
  // air primary field at the element's nodes: pol 1 Ex = -i fl32(w b0 z), pol 2 Ey = +i fl32(w b0 z)
  double pz = 0.0;
  for (int n = 0; n < NN; n++) pz += ntab->N[p][n] * (double)(float)(omega * MVF_B0 * xs[n][2]);
  if (d == 0) ei[0] -= pz;
  else ei[1] += pz;
  for (int c = 0; c < 3; c++) esol[(d * g.npt + id) * 3 + c] = make_double2(er[c], ei[c]);
}

// pass 2: H = -i fl32(1/omega) mu^-1 curl E from the owner element's nodal E
template <int NN>
__global__ __launch_bounds__(256) void k_node_h(const BasisTab *ntab, Geom g, const double *xp, const double *yp,
                                                const double *zp, const double *inmu, double omega,
                                                const double2 *esol, double2 *hsol) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 2 * g.npt) return;
  const int d = (int)(t / g.npt);
  const int64_t id = t % g.npt;
  int ie, je, ke;
  double xs[NN][3];
  int64_t nid[NN];
  const int p = owner_setup<NN>(ntab, g, xp, yp, zp, id, ie, je, ke, xs, nid);
  double Ji[9];
  jacobian<NN>(ntab, p, xs, Ji);
  double cr[3] = {0, 0, 0}, ci[3] = {0, 0, 0}, m1[6] = {0, 0, 0, 0, 0, 0};
  for (int n = 0; n < NN; n++) {
    double dn[3];
    for (int a = 0; a < 3; a++) dn[a] = Ji[3 * a] * ntab->dN[p][0][n] + Ji[3 * a + 1] * ntab->dN[p][1][n] +
                                        Ji[3 * a + 2] * ntab->dN[p][2][n];
    const double2 *f = esol + (d * g.npt + nid[n]) * 3;
    const double2 f0 = f[0], f1 = f[1], f2 = f[2];
    cr[0] += f2.x * dn[1] - f1.x * dn[2];
    ci[0] += f2.y * dn[1] - f1.y * dn[2];
    cr[1] += f0.x * dn[2] - f2.x * dn[0];
    ci[1] += f0.y * dn[2] - f2.y * dn[0];
    cr[2] += f1.x * dn[0] - f0.x * dn[1];
    ci[2] += f1.y * dn[0] - f0.y * dn[1];
    const double Nn = ntab->N[p][n];
    for (int c = 0; c < 6; c++) m1[c] += Nn * inmu[nid[n] * 6 + c];
  }
  // md = (0, mi) m1 with mi = fl32(-1/omega); H = md . curl
  const double mi = (double)(float)(-1.0 / omega);
  double tr[3], ti[3];
  sym6_mul(m1, cr, tr);
  sym6_mul(m1, ci, ti);
  // the full complex product (0, mi)(tr, ti), as the reference forms it (signed zeros included)
  for (int c = 0; c < 3; c++)
    hsol[(d * g.npt + id) * 3 + c] = make_double2(0.0 * tr[c] - mi * ti[c], 0.0 * ti[c] + mi * tr[c]);
}

// PROBLEM.OUT's fields (set_write_problem / write_problem_elem,
// problem.f90:459-507): the reference's element loop writes every lattice
// point of each element in turn, so the last element containing a node sets
// its values: max over the elements along each axis, min(n, a/o + 1).  At
// that point: the interpolated primary fields (get_ep / get_hp, :155-186) and
// the source p_source (:91-128) of polarization d.
template <int NN>
__global__ __launch_bounds__(256) void k_problem(const BasisTab *ntab, Geom g, const double *xp, const double *yp,
                                                 const double *zp, const double *sig, const double *inmu,
                                                 const double *dmu, double omega, double2 *ep, double2 *hp,
                                                 double2 *src) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 2 * g.npt) return;
  const int d = (int)(t / g.npt);
  const int64_t id = t % g.npt;
  const int o = g.nord - 1;
  const int a = (int)(id / g.nyz), b = (int)((id / g.nnz) % g.nny), c = (int)(id % g.nnz);
  const int ie = min(g.nx, a / o + 1), je = min(g.ny, b / o + 1), ke = min(g.nz, c / o + 1);
  const int p = ((a - (ie - 1) * o) * g.nord + (b - (je - 1) * o)) * g.nord + (c - (ke - 1) * o);
  double xs[NN][3];
  int64_t nid[NN];
  for (int n = 0; n < NN; n++) {
    int ix, jy;
    load_node(ntab, g, ie, je, ke, n, nid[n], ix, jy);
    xs[n][0] = xp[ix];
    xs[n][1] = yp[jy];
    xs[n][2] = zp[nid[n]];
  }
  double Ji[9];
  jacobian<NN>(ntab, p, xs, Ji);
  const double w32 = (double)(float)omega;
  const double psig = (double)(float)(omega * MVF_EPS0);
  const double hh = (double)(float)MVF_B0 / (4 * MVF_PI * 1.e-7);
  // Ep has one imaginary component per node (x: -i zz for d = 0, y: +i zz
  // for d = 1), Hp one real component (y for d = 0, x for d = 1)
  const int ce = d == 0 ? 0 : 1, chh = d == 0 ? 1 : 0;
  static const int T6[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  double epi = 0.0, hpr = 0.0, sr[3] = {0, 0, 0}, si[3] = {0, 0, 0};
  for (int n = 0; n < NN; n++) {
    const double N = ntab->N[p][n];
    const double zz = (double)(float)(omega * MVF_B0 * xs[n][2]);
    const double e = d == 0 ? -zz : zz;
    epi += e * N;
    hpr += hh * N;
    // p_dmpf: N (dsigma . Ep)
    for (int q = 0; q < 3; q++) {
      const int k6 = T6[q][ce];
      double dr = sig[2 * (nid[n] * 6 + k6)], di = sig[2 * (nid[n] * 6 + k6) + 1];
      if (k6 == 0 || k6 == 3 || k6 == 5) di -= psig;
      sr[q] += N * (-di * e);
      si[q] += N * (dr * e);
    }
    // p_pcurl: grad N x (mu^-1 dmu Hp)
    const double *im = inmu + nid[n] * 6, *dm = dmu + nid[n] * 6;
    double v[3];
    for (int q = 0; q < 3; q++) {
      double m = 0.0;
      for (int r = 0; r < 3; r++) m += im[T6[q][r]] * dm[T6[r][chh]];
      v[q] = m * hh;
    }
    double dn[3];
    for (int q = 0; q < 3; q++)
      dn[q] = Ji[3 * q] * ntab->dN[p][0][n] + Ji[3 * q + 1] * ntab->dN[p][1][n] + Ji[3 * q + 2] * ntab->dN[p][2][n];
    sr[0] += v[2] * dn[1] - v[1] * dn[2];
    sr[1] += v[0] * dn[2] - v[2] * dn[0];
    sr[2] += v[1] * dn[0] - v[0] * dn[1];
  }
  const int64_t base = (d * g.npt + id) * 3;
  for (int q = 0; q < 3; q++) {
    ep[base + q] = make_double2(0.0, q == ce ? epi : 0.0);
    hp[base + q] = make_double2(q == chh ? hpr : 0.0, 0.0);
    // times -i w32: (a + i b)(-i w) = b w - i a w
    src[base + q] = make_double2(si[q] * w32, -sr[q] * w32);
  }
}

__device__ inline double2 zmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 zdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// Z = E H^-1 over the two polarizations (get_impedance, inv_matrix,
// solution.f90:456-505) from e, h = (x pol1, x pol2, y pol1, y pol2), and
// rho_a = |Z|^2 / (omega mu0), phase atan(Im/Re) in degrees, both 0 when
// rho_a < 1e-2 (get_res_phase)
__device__ inline void mt_impedance(const double2 *e, const double2 *hh, double omega, double2 *z, double *rho,
                                    double *phi) {
  const double2 m01 = zmul(hh[0], hh[3]), m12 = zmul(hh[1], hh[2]);
  const double2 det = make_double2(m01.x - m12.x, m01.y - m12.y);
  const double2 i0 = zdiv(hh[3], det), i1 = zdiv(make_double2(-hh[1].x, -hh[1].y), det);
  const double2 i2 = zdiv(make_double2(-hh[2].x, -hh[2].y), det), i3 = zdiv(hh[0], det);
  const double2 a0 = zmul(i0, e[0]), a1 = zmul(i2, e[1]), b0 = zmul(i1, e[0]), b1 = zmul(i3, e[1]);
  const double2 c0 = zmul(i0, e[2]), c1 = zmul(i2, e[3]), q0 = zmul(i1, e[2]), q1 = zmul(i3, e[3]);
  z[0] = make_double2(a0.x + a1.x, a0.y + a1.y);
  z[1] = make_double2(b0.x + b1.x, b0.y + b1.y);
  z[2] = make_double2(c0.x + c1.x, c0.y + c1.y);
  z[3] = make_double2(q0.x + q1.x, q0.y + q1.y);
  const double mu0 = 4 * MVF_PI * 1.e-7;
  for (int i = 0; i < 4; i++) {
    double r = (1.0 / (omega * mu0)) * (z[i].x * z[i].x + z[i].y * z[i].y);
    double ph = 0.0;
    if (r < 1e-2) r = 0.0;
    else ph = atan(z[i].y / z[i].x) * (180.0 / MVF_PI);
    rho[i] = r;
    phi[i] = ph;
  }
}

// Z, rho_a, phase at every lattice node from its nodal fields (z_rho_phi,
// hor_fields, solution.f90:51-69,440-454)
__global__ void k_node_z(int64_t npt, const double2 *esol, const double2 *hsol, double omega, double2 *oz,
                         double *rho, double *phi) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= npt) return;
  const double2 e[4] = {esol[id * 3 + 0], esol[(npt + id) * 3 + 0], esol[id * 3 + 1], esol[(npt + id) * 3 + 1]};
  const double2 hh[4] = {hsol[id * 3 + 0], hsol[(npt + id) * 3 + 0], hsol[id * 3 + 1], hsol[(npt + id) * 3 + 1]};
  mt_impedance(e, hh, omega, oz + id * 4, rho + id * 4, phi + id * 4);
}

// receivers: bilinear interpolation of the nodal fields on the top face of
// element (ie, je, ke_top) (interpolate_rx), Z = E H^-1 over the two
// polarizations (get_impedance), rho_a = |Z|^2 / (omega mu0) and phase
// atan(Im/Re) in degrees, both 0 when rho_a < 1e-2 (get_res_phase)
__global__ void k_receivers(Geom g, int nrx, const int32_t *rie, const int32_t *rje, int ke, const double *rx,
                            const double *ry, const double *xp, const double *yp, const double *zp,
                            const double2 *esol, const double2 *hsol, double omega, double2 *oe, double2 *oh,
                            double2 *oz, double *rho, double *phi, double *elev) {
  const int ir = blockIdx.x * blockDim.x + threadIdx.x;
  if (ir >= nrx) return;
  const int o = g.nord - 1, ie = rie[ir], je = rje[ir];
  const int64_t eno = (int64_t)(ie - 1) * o * g.nyz + (int64_t)(je - 1) * o * g.nnz + (int64_t)(ke - 1) * o;
  const double x1 = xp[(ie - 1) * o], x2 = xp[(ie - 1) * o + o];
  const double y1 = yp[(je - 1) * o], y2 = yp[(je - 1) * o + o];
  const int64_t id11 = eno + o, id12 = eno + (int64_t)o * g.nnz + o, id21 = eno + (int64_t)o * g.nyz + o,
                id22 = eno + (int64_t)o * g.nyz + (int64_t)o * g.nnz + o;
  const double x = rx[ir], y = ry[ir];
  const double ax = (x2 - x) / (x2 - x1), bx = (x - x1) / (x2 - x1);
  const double ay = (y2 - y) / (y2 - y1), by = (y - y1) / (y2 - y1);
  // the receiver's height on the element's top face (interpolate_rx, receiver_data.f90:222-225)
  const double zx1 = ax * zp[id11] + bx * zp[id21], zx2 = ax * zp[id12] + bx * zp[id22];
  elev[ir] = ay * zx1 + by * zx2;
  for (int t = 0; t < 2; t++) {
    const double2 *F = t == 0 ? esol : hsol;
    double2 *O = t == 0 ? oe : oh;
    for (int d = 0; d < 2; d++)
      for (int c = 0; c < 3; c++) {
        const double2 f11 = F[(d * g.npt + id11) * 3 + c], f12 = F[(d * g.npt + id12) * 3 + c];
        const double2 f21 = F[(d * g.npt + id21) * 3 + c], f22 = F[(d * g.npt + id22) * 3 + c];
        const double2 fx1 = make_double2(ax * f11.x + bx * f21.x, ax * f11.y + bx * f21.y);
        const double2 fx2 = make_double2(ax * f12.x + bx * f22.x, ax * f12.y + bx * f22.y);
        O[(ir * 2 + d) * 3 + c] = make_double2(ay * fx1.x + by * fx2.x, ay * fx1.y + by * fx2.y);
      }
  }
  // assign_rxmatrix order: (x pol1, x pol2, y pol1, y pol2)
  const double2 e[4] = {oe[(ir * 2 + 0) * 3 + 0], oe[(ir * 2 + 1) * 3 + 0], oe[(ir * 2 + 0) * 3 + 1],
                        oe[(ir * 2 + 1) * 3 + 1]};
  const double2 hh[4] = {oh[(ir * 2 + 0) * 3 + 0], oh[(ir * 2 + 1) * 3 + 0], oh[(ir * 2 + 0) * 3 + 1],
                         oh[(ir * 2 + 1) * 3 + 1]};
  mt_impedance(e, hh, omega, oz + ir * 4, rho + ir * 4, phi + ir * 4);
}

template <int NN, int NE>
static void launch_node_fields(mvf_handle *h, const double2 *x, int ncol, int col0, double omega, double2 *es,
                               double2 *hs) {
  const int64_t n = 2 * h->geo.npt;
  hipLaunchKernelGGL((k_node_e<NN, NE>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, h->d_ntab, h->geo, h->d_xp,
                     h->d_yp, h->d_zp, h->d_gne, x, ncol, col0, omega, h->bd_on ? h->d_bidx : nullptr,
                     h->d_dirf, es);
  hipLaunchKernelGGL((k_node_h<NN>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, h->d_ntab, h->geo, h->d_xp,
                     h->d_yp, h->d_zp, h->d_inmu, omega, es, hs);
}

// basis functions and derivatives at the nord^3 lattice points of an element
// (uploaded once)
int node_table(mvf_handle *h) {
  if (!h->d_ntab) {
    const int nord = h->et.nord;
    double pts[27][3];
    const double as3[3] = {-1.0, 0.0, 1.0}, as2[2] = {-1.0, 1.0};
    int np = 0;
    for (int i1 = 0; i1 < nord; i1++)
      for (int j1 = 0; j1 < nord; j1++)
        for (int k1 = 0; k1 < nord; k1++, np++) {
          const double *as = nord == 2 ? as2 : as3;
          pts[np][0] = as[i1];
          pts[np][1] = as[j1];
          pts[np][2] = as[k1];
        }
    BasisTab *t = new BasisTab();
    fill_basis_points(h->nodes, pts, np, nullptr, t);
    int rc = dalloc(h, &h->d_ntab, 1);
    if (!rc) {
      hipError_t e = hipMemcpy(h->d_ntab, t, sizeof(BasisTab), hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = set_err(h, MVF_E_DEVICE, "hipMemcpy(node table): %s", hipGetErrorString(e));
    }
    delete t;
    if (rc) return rc;
  }
  return 0;
}

// nodal E and H (device arrays of 2*npt*3 complex) for solution columns
// (col0, col0+1) of x ([nne][ncol] complex, device) at angular frequency omega
int node_fields(mvf_handle *h, const double *x, int ncol, int col0, double omega, double *es, double *hs) {
  MVF_TRY(node_table(h));
  MVF_TRY(bd_values(h, omega));
  const double2 *xv = (const double2 *)x;
  double2 *e2 = (double2 *)es, *h2 = (double2 *)hs;
  switch (h->nodes) {
    case 8: launch_node_fields<8, 12>(h, xv, ncol, col0, omega, e2, h2); break;
    case 20: launch_node_fields<20, 36>(h, xv, ncol, col0, omega, e2, h2); break;
    default: launch_node_fields<27, 54>(h, xv, ncol, col0, omega, e2, h2); break;
  }
  MVF_HIP(h, hipGetLastError());
  return 0;
}

int receivers(mvf_handle *h, const double *es, const double *hs, int ke_top, int nrx, const int32_t *rie,
              const int32_t *rje, const double *rx, const double *ry, double omega, double *oe, double *oh, double *oz,
              double *rho, double *phi, double *elev) {
  hipLaunchKernelGGL(k_receivers, dim3(grid1d(nrx, 64)), dim3(64), 0, h->stream, h->geo, nrx, rie, rje, ke_top, rx, ry,
                     h->d_xp, h->d_yp, h->d_zp, (const double2 *)es, (const double2 *)hs, omega, (double2 *)oe,
                     (double2 *)oh, (double2 *)oz, rho, phi, elev);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

// PROBLEM.OUT's fields at the current frequency into device arrays of
// 2*npt*3 complex each
int problem_fields(mvf_handle *h, double omega, double *ep, double *hp, double *src) {
  MVF_TRY(node_table(h));
  const int64_t n = 2 * h->geo.npt;
  double2 *e2 = (double2 *)ep, *h2 = (double2 *)hp, *s2 = (double2 *)src;
#define MVF_PROBLEM(NN)                                                                                            \
  hipLaunchKernelGGL((k_problem<NN>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, h->d_ntab, h->geo, h->d_xp,   \
                     h->d_yp, h->d_zp, h->d_sig, h->d_inmu, h->d_dmu, omega, e2, h2, s2)
  switch (h->nodes) {
    case 8: MVF_PROBLEM(8); break;
    case 20: MVF_PROBLEM(20); break;
    default: MVF_PROBLEM(27); break;
  }
#undef MVF_PROBLEM
  MVF_HIP(h, hipGetLastError());
  return 0;
}

// nodal Z, rho_a, phase from device nodal fields (2*npt*3 complex each)
int node_impedance(mvf_handle *h, const double *es, const double *hs, double omega, double *oz, double *rho,
                   double *phi) {
  const int64_t npt = h->geo.npt;
  hipLaunchKernelGGL(k_node_z, dim3(grid1d(npt, 256)), dim3(256), 0, h->stream, npt, (const double2 *)es,
                     (const double2 *)hs, omega, (double2 *)oz, rho, phi);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

}  // namespace mvf
