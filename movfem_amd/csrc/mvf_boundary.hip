// Dirichlet boundary values for inimod 2 (homogeneous half-space) and 3
// (layered earth) and their lift into the element RHS.
//
// The reference removes the Dirichlet edges from the system and moves their
// prescribed values to the right-hand side element by element:
//   b(im) = blocal(im) - sum_{jm Dirichlet} f_boundary(jm) alocal(im, jm)
// for every free edge im (local_vfem, MoVFEM_3DMT.f90:251-261).  f_boundary
// (boundary_conds.f90:188-240) is the tangential component E_p . dr/dxi_d of
// a 1-D primary field at the edge's basis node on the x and y faces (codes
// 1, 2, 4, 5) and zero on the z faces and for inimod 1.  The field depends on
// the node height only (p_pfields, :436-516):
//   inimod 2: fp = (1/g)(1 - z g) in the air, exp(-z g)/g below the surface;
//   inimod 3: i c (cz1 - z) in the air and i c cz1 ezl(0, z) below, with cz
//             from the Wait recursion (:521-553) and ezl (:558-578);
// z = (ztop - z_node)/1000 and c = float32(2 omega b0).  ezl compares the
// depth z >= 0 with interface heights <= 0, so below the surface only z = 0
// gives a non-zero value.
//
// Per frequency the host forms the scalars of the 1-D model (square roots,
// the recursion, ezl(0, 0)) in the reference's complex arithmetic; the device
// evaluates the per-node fields, the Jacobian rows at the basis nodes and the
// lift.  Only elements on the x and y faces carry such edges; their
// frequency-independent K_e and M_e are kept from setup, so that
// alocal = K_e + i float32(omega) M_e (integration.f90:76-86) is formed on the
// fly for them.
#include <complex>

#include "mvf_internal.h"

namespace mvf {

// ---------------------------------------------------------------- host side
// Complex arithmetic as the reference's Fortran evaluates it: real operands
// are promoted to (r, 0) and multiplied in full; quotients use the scaled
// formula of C99 Annex G.5.1 (the divisor is scaled by 2^-logb(max(|c|,|d|)),
// an exact rescaling); sqrt and exp are the C library's csqrt and cexp.
namespace {

struct hz {
  double re, im;
};

hz hz_mul(hz a, hz b) {
#pragma clang fp contract(off)
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
hz hz_add(hz a, hz b) { return {a.re + b.re, a.im + b.im}; }
hz hz_sub(hz a, hz b) { return {a.re - b.re, a.im - b.im}; }
hz hz_neg(hz a) { return {-a.re, -a.im}; }
hz hz_r(double r) { return {r, 0.0}; }
hz hz_div(hz x, hz y) {
#pragma clang fp contract(off)
  double c = y.re, d = y.im;
  int k = 0;
  double lg = std::logb(std::fmax(std::fabs(c), std::fabs(d)));
  if (std::isfinite(lg)) {
    k = (int)lg;
    c = std::scalbn(c, -k);
    d = std::scalbn(d, -k);
  }
  double den = c * c + d * d;
  return {std::scalbn((x.re * c + x.im * d) / den, -k), std::scalbn((x.im * c - x.re * d) / den, -k)};
}
hz hz_sqrt(hz a) {
  std::complex<double> r = std::sqrt(std::complex<double>(a.re, a.im));
  return {r.real(), r.imag()};
}
hz hz_exp(hz a) {
  std::complex<double> r = std::exp(std::complex<double>(a.re, a.im));
  return {r.real(), r.imag()};
}

// sqrt(i omega mu0 psigma), psigma = float32(sigma) + i float32(eps omega)
// (bd_setmodel's default-kind CMPLX, bd_updatemodel, boundary_conds.f90:47-49,
// 411-428)
hz bd_gamma(double omega, double sigma) {
  hz psig = {(double)(float)sigma, (double)(float)(MVF_EPS0 * omega)};
  hz alpha = hz_mul(hz_r(omega * (4 * MVF_PI * 1.e-7)), psig);
  return hz_sqrt(hz_mul({0.0, 1.0}, alpha));
}

}  // namespace

BdParams bd_params(const mvf_handle *h, double omega) {
  BdParams P{};
  P.inimod = h->bc.inimod;
  P.ztop = h->bd_ztop;
  P.cm = (double)(float)(-2.0 * omega * MVF_B0);
  P.cp = (double)(float)(2.0 * omega * MVF_B0);
  if (P.inimod == 2) {
    hz sa = bd_gamma(omega, 0.0), se = bd_gamma(omega, h->bd_hsigma);
    hz qa = hz_div(hz_r(1.0), sa), qe = hz_div(hz_r(1.0), se);
    P.sa = make_double2(sa.re, sa.im);
    P.se = make_double2(se.re, se.im);
    P.qa = make_double2(qa.re, qa.im);
    P.qe = make_double2(qe.re, qe.im);
  } else if (P.inimod == 3) {
    const int nl = (int)h->bd_lsig.size();
    const std::vector<double> &sg = h->bd_lsig, &dl = h->bd_ldz;
    std::vector<hz> cz(nl), ez(nl), gm(nl);
    std::vector<double> zl(nl);
    for (int l = 0; l < nl; l++) gm[l] = bd_gamma(omega, sg[l]);
    zl[0] = 0.0;
    for (int l = 1; l < nl; l++) zl[l] = zl[l - 1] - dl[l - 1];
    // wait_recursion (boundary_conds.f90:521-553)
    cz[nl - 1] = hz_div(hz_r(1.0), gm[nl - 1]);
    for (int l = nl - 2; l >= 0; l--) {
      hz t = hz_mul(gm[l], cz[l + 1]);
      hz r = hz_div(hz_sub(hz_r(1.0), t), hz_add(hz_r(1.0), t));
      hz re = hz_mul(r, hz_exp(hz_neg(hz_mul(hz_mul(hz_r(2.0), gm[l]), hz_r(dl[l])))));
      cz[l] = hz_div(hz_sub(hz_r(1.0), re), hz_mul(gm[l], hz_add(hz_r(1.0), re)));
    }
    ez[0] = hz_r(1.0);
    for (int l = 0; l < nl - 1; l++) {
      hz e = hz_exp(hz_neg(hz_mul(gm[l], hz_r(dl[l]))));
      hz p = hz_mul(hz_mul(ez[l], cz[l + 1]), hz_add(hz_r(1.0), hz_mul(cz[l], gm[l])));
      hz q = hz_mul(cz[l], hz_add(hz_r(1.0), hz_mul(cz[l + 1], gm[l])));
      ez[l + 1] = hz_div(hz_mul(e, p), q);
    }
    // ezl(0, 0) (boundary_conds.f90:558-578): the surface node's factor
    const double z = 0.0;
    hz v = {0.0, 0.0};
    for (int l = 0; l < nl - 1; l++) {
      if (z <= zl[l] && z > zl[l + 1]) {
        hz t = hz_mul(gm[l], cz[l + 1]);
        hz r = hz_div(hz_sub(hz_r(1.0), t), hz_add(hz_r(1.0), t));
        hz a1 = hz_add(hz_r(1.0), hz_div(hz_r(1.0), hz_mul(cz[l], gm[l])));
        hz a2 = hz_sub(hz_r(1.0), hz_mul(r, hz_exp(hz_neg(hz_mul(hz_mul(hz_r(2.0), gm[l]), hz_r(z - zl[l + 1]))))));
        hz a3 = hz_exp(hz_neg(hz_mul(gm[l], hz_r(zl[l] - z))));
        v = hz_mul(hz_mul(hz_mul(hz_div(ez[l], hz_r(2.0)), a1), a2), a3);
      }
    }
    if (z <= zl[nl - 1]) v = hz_mul(ez[nl - 1], hz_exp(hz_neg(hz_mul(gm[nl - 1], hz_r(zl[nl - 1] - z)))));
    P.cz1 = make_double2(cz[0].re, cz[0].im);
    P.e0 = make_double2(v.re, v.im);
  }
  return P;
}

// -------------------------------------------------------------- device side
__device__ inline double2 bz_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Ex of polarization 1 and Ey of polarization 2 at a node of height zn
__device__ inline void bd_field(const BdParams &P, double zn, double2 &e1, double2 &e2) {
  const double z = (P.ztop - zn) / 1000.0;
  const bool air = zn > P.ztop;
  const double2 im = make_double2(0.0, P.cm), ip = make_double2(0.0, P.cp);
  e1 = e2 = make_double2(0.0, 0.0);
  if (P.inimod == 2) {
    double2 fp;
    if (air) {
      fp = bz_mul(P.qa, make_double2(1.0 - z * P.sa.x, -(z * P.sa.y)));
    } else {
      double s, c;
      sincos(-(z * P.se.y), &s, &c);
      const double m = exp(-(z * P.se.x));
      fp = bz_mul(P.qe, make_double2(m * c, m * s));
    }
    e1 = bz_mul(im, fp);
    e2 = bz_mul(ip, fp);
  } else if (P.inimod == 3) {
    if (air) {
      const double2 t = make_double2(P.cz1.x - z, P.cz1.y);
      e1 = bz_mul(ip, t);
      e2 = bz_mul(im, t);
    } else if (z == 0.0) {
      e1 = bz_mul(bz_mul(ip, P.cz1), P.e0);
      e2 = bz_mul(bz_mul(im, P.cz1), P.e0);
    }
  }
}

// f_boundary of every edge of the boundary elements:
// dirf[(b*NE + jm)*2 + pol], zero for free edges and z-face codes.
// btab holds the basis at the element's own basis nodes.
template <int NN, int NE>
__global__ __launch_bounds__(64) void k_bd_values(const BasisTab *__restrict__ btab, Geom g,
                                                  const double *__restrict__ xp, const double *__restrict__ yp,
                                                  const double *__restrict__ zp, const int32_t *__restrict__ gne,
                                                  const int64_t *__restrict__ bel, BdParams P,
                                                  double2 *__restrict__ dirf) {
  const int64_t b = blockIdx.x;
  const int t = threadIdx.x;
  __shared__ double xs[NN][3];
  const int64_t e = bel[b];
  int i, j, k;
  elem_index(g, e, i, j, k);
  if (t < NN) {
    int64_t id;
    int ix, jy;
    load_node(btab, g, i, j, k, t, id, ix, jy);
    xs[t][0] = xp[ix];
    xs[t][1] = yp[jy];
    xs[t][2] = zp[id];
  }
  __syncthreads();
  if (t < NE) {
    double2 f1 = make_double2(0.0, 0.0), f2 = f1;
    const int32_t gi = gne[e * NE + t];
    if (gi < 0 && gi != -3 && gi != -6) {
      const int n = btab->node[t], d = btab->dir[t];
      // row d of the Jacobian at basis node n (nf_jacobian, nf_dr_dxi)
      double jr[3];
      for (int c = 0; c < 3; c++) {
        double acc = 0.0;
        for (int l = 0; l < NN; l++) acc += btab->dN[n][d][l] * xs[l][c];
        jr[c] = acc;
      }
      double2 e1, e2;
      bd_field(P, xs[n][2], e1, e2);
      f1 = make_double2(e1.x * jr[0], e1.y * jr[0]);
      f2 = make_double2(e2.x * jr[1], e2.y * jr[1]);
    }
    dirf[(b * NE + t) * 2 + 0] = f1;
    dirf[(b * NE + t) * 2 + 1] = f2;
  }
}

// be[e][im][pol] -= sum_{jm Dirichlet} f(jm, pol) (K_e + i w32 M_e)(im, jm)
template <int NE>
__global__ __launch_bounds__(64) void k_bd_lift(const int32_t *__restrict__ gne, const int64_t *__restrict__ bel,
                                                const double *__restrict__ bK, const double *__restrict__ bM,
                                                const double2 *__restrict__ dirf, double w32,
                                                double *__restrict__ be) {
  constexpr int64_t NP = NE * (NE + 1) / 2;
  const int64_t b = blockIdx.x;
  const int t = threadIdx.x;
  __shared__ double2 f[NE][2];
  __shared__ int dir[NE];
  const int64_t e = bel[b];
  if (t < NE) {
    f[t][0] = dirf[(b * NE + t) * 2 + 0];
    f[t][1] = dirf[(b * NE + t) * 2 + 1];
    dir[t] = gne[e * NE + t] < 0;
  }
  __syncthreads();
  if (t >= NE || dir[t]) return;
  const double *K = bK + b * NP, *M = bM + b * NP;
  double2 s0 = make_double2(0.0, 0.0), s1 = s0;
  for (int jm = 0; jm < NE; jm++) {
    if (!dir[jm]) continue;
    const int lo = t < jm ? t : jm, hi = t < jm ? jm : t;
    const int64_t p = (int64_t)lo * NE - (int64_t)lo * (lo - 1) / 2 + (hi - lo);
    const double2 a = make_double2(K[p], w32 * M[p]);
    const double2 p0 = bz_mul(f[jm][0], a), p1 = bz_mul(f[jm][1], a);
    s0.x += p0.x; s0.y += p0.y;
    s1.x += p1.x; s1.y += p1.y;
  }
  double *o = be + (e * NE + t) * 4;
  o[0] -= s0.x; o[1] -= s0.y;
  o[2] -= s1.x; o[3] -= s1.y;
}

// ----------------------------------------------------------------- drivers
// Boundary element list, element -> list index, K_e/M_e of those elements
// (from the setup integration dKe/dMe of all elements) and the basis at the
// basis nodes.
int bd_setup(mvf_handle *h, const double *dKe, const double *dMe) {
  const Geom &g = h->geo;
  std::vector<int64_t> bel;
  std::vector<int32_t> bidx(g.ne, -1);
  for (int i = 1; i <= g.nx; i++)
    for (int j = 1; j <= g.ny; j++)
      for (int k = 1; k <= g.nz; k++) {
        if (i != 1 && i != g.nx && j != 1 && j != g.ny) continue;
        int64_t e = ((int64_t)(i - 1) * g.ny + (j - 1)) * g.nz + (k - 1);
        bidx[e] = (int32_t)bel.size();
        bel.push_back(e);
      }
  h->nbel = (int64_t)bel.size();
  const int64_t np = (int64_t)h->et.ne * (h->et.ne + 1) / 2;
  MVF_TRY(dalloc(h, &h->d_bel, h->nbel));
  MVF_TRY(dalloc(h, &h->d_bidx, g.ne));
  MVF_TRY(dalloc(h, &h->d_bK, np * h->nbel));
  MVF_TRY(dalloc(h, &h->d_bM, np * h->nbel));
  MVF_TRY(dalloc(h, &h->d_dirf, (size_t)h->nbel * h->et.ne * 2));
  MVF_HIP(h, hipMemcpy(h->d_bel, bel.data(), 8 * h->nbel, hipMemcpyHostToDevice));
  MVF_HIP(h, hipMemcpy(h->d_bidx, bidx.data(), 4 * g.ne, hipMemcpyHostToDevice));
  // runs of consecutive element ids copy as one block
  for (int64_t s = 0; s < h->nbel;) {
    int64_t r = s + 1;
    while (r < h->nbel && bel[r] == bel[r - 1] + 1) r++;
    const size_t bytes = 8 * (size_t)np * (r - s);
    MVF_HIP(h, hipMemcpyAsync(h->d_bK + s * np, dKe + bel[s] * np, bytes, hipMemcpyDeviceToDevice, h->stream));
    MVF_HIP(h, hipMemcpyAsync(h->d_bM + s * np, dMe + bel[s] * np, bytes, hipMemcpyDeviceToDevice, h->stream));
    s = r;
  }
  BasisTab *tab = new BasisTab();
  memset(tab, 0, sizeof(BasisTab));
  double pts[27][3];
  for (int n = 0; n < h->nodes; n++)
    for (int c = 0; c < 3; c++) pts[n][c] = kNodeRef[n][c];
  fill_basis_points(h->nodes, pts, h->nodes, nullptr, tab);
  int rc = dalloc(h, &h->d_btab, 1);
  if (!rc && hipMemcpy(h->d_btab, tab, sizeof(BasisTab), hipMemcpyHostToDevice) != hipSuccess)
    rc = set_err(h, MVF_E_DEVICE, "basis-node table upload failed");
  delete tab;
  MVF_TRY(rc);
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  h->bd_omega = -1.0;
  return 0;
}

// f_boundary of every boundary element at omega into h->d_dirf
int bd_values(mvf_handle *h, double omega) {
  if (!h->bd_on || h->nbel == 0) return 0;
  if (h->bd_omega == omega) return 0;
  const BdParams P = bd_params(h, omega);
  const unsigned nb = (unsigned)h->nbel;
  if (h->nodes == 8)
    hipLaunchKernelGGL((k_bd_values<8, 12>), dim3(nb), dim3(64), 0, h->stream, h->d_btab, h->geo, h->d_xp, h->d_yp,
                       h->d_zp, h->d_gne, h->d_bel, P, h->d_dirf);
  else if (h->nodes == 20)
    hipLaunchKernelGGL((k_bd_values<20, 36>), dim3(nb), dim3(64), 0, h->stream, h->d_btab, h->geo, h->d_xp, h->d_yp,
                       h->d_zp, h->d_gne, h->d_bel, P, h->d_dirf);
  else
    hipLaunchKernelGGL((k_bd_values<27, 54>), dim3(nb), dim3(64), 0, h->stream, h->d_btab, h->geo, h->d_xp, h->d_yp,
                       h->d_zp, h->d_gne, h->d_bel, P, h->d_dirf);
  MVF_HIP(h, hipGetLastError());
  h->bd_omega = omega;
  return 0;
}

// lift of the Dirichlet values into the element RHS be of all elements
int bd_lift(mvf_handle *h, double omega, double *d_be) {
  if (!h->bd_on || h->nbel == 0) return 0;
  MVF_TRY(bd_values(h, omega));
  const double w32 = (double)(float)omega;
  const unsigned nb = (unsigned)h->nbel;
  if (h->nodes == 8)
    hipLaunchKernelGGL((k_bd_lift<12>), dim3(nb), dim3(64), 0, h->stream, h->d_gne, h->d_bel, h->d_bK, h->d_bM,
                       h->d_dirf, w32, d_be);
  else if (h->nodes == 20)
    hipLaunchKernelGGL((k_bd_lift<36>), dim3(nb), dim3(64), 0, h->stream, h->d_gne, h->d_bel, h->d_bK, h->d_bM,
                       h->d_dirf, w32, d_be);
  else
    hipLaunchKernelGGL((k_bd_lift<54>), dim3(nb), dim3(64), 0, h->stream, h->d_gne, h->d_bel, h->d_bK, h->d_bM,
                       h->d_dirf, w32, d_be);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

}  // namespace mvf
