// mvf_shepard.h -- modified quadratic Shepard interpolation of scattered
// surface points (Renka, ACM TOMS algorithm 660, QSHEP2D), the interpolant the
// reference uses for its topography and interface heights
// (toms660.f90; called from geometry.f90:456-458,562-563).
//
// Host code.  Lattice heights must match the reference bit for bit, so every
// routine keeps the reference's operation order and control flow, including
// two of its deviations from the published algorithm:
//   * the nearest-point search stops when the search box covers the
//     candidate circle on the +x, -y and +y sides only: its exit test compares
//     i1 with itself (toms660.f90:246) where the published code tests the
//     -x side, so a closer point left of the box can be missed;
//   * the routines do not stop on an error condition they cannot report
//     (getnp2 with nr < 1, toms660.f90:118-121).
// Cell lists and node numbers are 1-based as in the reference (0 = empty cell,
// lnext(k) = k ends a list).
//
// Provenance: the algorithm and its routine structure (store2, getnp2,
// givens, rotate, setup2, qshep2, qs2val) are Robert J. Renka's, published as
// "Algorithm 660: QSHEP2D, Quadratic Shepard method for bivariate
// interpolation of scattered data", ACM Transactions on Mathematical Software
// 14 (1988) 149-150, and vendored by the reference as toms660.f90 (its header
// names that author and reference, toms660.f90:35-45; the file states no
// licence of its own).  ACM publishes its collected algorithms under the ACM
// Software License Agreement, which allows use for non-commercial purposes.
// This file restates the routines in C++ for bit-identical lattices and adds
// nothing to the method.
#pragma once

#include <cmath>
#include <cstdlib>
#include <vector>

namespace shep {

// One fitted interpolant: the cell grid of store2 and the nodal-function
// coefficients of qshep2 (toms660.f90:679-1207).  x, y, f are copies of the
// data points.
struct Fit {
  int n = 0, nr = 0, ier = 0;
  std::vector<double> x, y, f, rsq, a;  // a(5, n) column-major
  std::vector<int> lcell, lnext;        // lcell(nr, nr) column-major
  double xmin = 0, ymin = 0, dx = 0, dy = 0, rmax = 0;
};

// store2 (toms660.f90:1579-1745): cell grid of nr x nr cells over the
// bounding box, node lists threaded through lnext with the lowest node first.
static int store2(int n, const double *x, const double *y, int nr, int *lcell, int *lnext, double &xmin,
                  double &ymin, double &dx, double &dy) {
  if (n < 2 || nr < 1) return 1;
  double xmn = x[0], xmx = x[0], ymn = y[0], ymx = y[0];
  for (int k = 1; k < n; ++k) {
    if (x[k] < xmn) xmn = x[k];
    if (x[k] > xmx) xmx = x[k];
    if (y[k] < ymn) ymn = y[k];
    if (y[k] > ymx) ymx = y[k];
  }
  xmin = xmn;
  ymin = ymn;
  dx = (xmx - xmn) / (double)nr;
  dy = (ymx - ymn) / (double)nr;
  if (dx == 0.0 || dy == 0.0) return 2;
  for (int c = 0; c < nr * nr; ++c) lcell[c] = 0;
  for (int k = n; k >= 1; --k) {
    int i = (int)((x[k - 1] - xmin) / dx) + 1;
    if (i > nr) i = nr;
    int j = (int)((y[k - 1] - ymin) / dy) + 1;
    if (j > nr) j = nr;
    int &cell = lcell[(j - 1) * nr + (i - 1)];
    lnext[k - 1] = cell != 0 ? cell : k;
    cell = k;
  }
  return 0;
}

// getnp2 (toms660.f90:4-287): the closest unmarked node (lnext >= 0) to
// (px, py); marks it by negating its lnext entry.  np = 0 if none.
static void getnp2(double px, double py, const double *x, const double *y, int nr, const int *lcell, int *lnext,
                   double xmin, double ymin, double dx, double dy, int &np, double &dsq) {
  const double xp = px, yp = py;
  if (nr < 1 || dx <= 0.0 || dy <= 0.0) {  // the reference does not return here
    np = 0;
    dsq = 0.0;
  }
  bool first = true;
  int imin = 1, imax = nr, jmin = 1, jmax = nr;
  const double delx = xp - xmin, dely = yp - ymin;
  int i0 = (int)(delx / dx) + 1;
  if (i0 < 1) i0 = 1;
  if (i0 > nr) i0 = nr;
  int j0 = (int)(dely / dy) + 1;
  if (j0 < 1) j0 = 1;
  if (j0 > nr) j0 = nr;
  int i1 = i0, i2 = i0, j1 = j0, j2 = j0;
  int lmin = 0;
  double rsmin = 0.0;
  for (;;) {
    for (int j = j1; j <= j2; ++j) {
      if (jmax < j) break;
      if (j < jmin) continue;
      for (int i = i1; i <= i2; ++i) {
        if (imax < i) break;
        if (i < imin) continue;
        if (j != j1 && j != j2 && i != i1 && i != i2) continue;
        int l = lcell[(j - 1) * nr + (i - 1)];
        if (l > 0) {
          for (;;) {
            const int ln = lnext[l - 1];
            if (ln >= 0) {
              const double ddx = x[l - 1] - xp, ddy = y[l - 1] - yp;
              const double rsq = ddx * ddx + ddy * ddy;
              if (first) {
                lmin = l;
                rsmin = rsq;
                const double r = std::sqrt(rsmin);
                imin = (int)((delx - r) / dx) + 1;
                if (imin < 1) imin = 1;
                imax = (int)((delx + r) / dx) + 1;
                if (imax > nr) imax = nr;
                jmin = (int)((dely - r) / dy) + 1;
                if (jmin < 1) jmin = 1;
                jmax = (int)((dely + r) / dy) + 1;
                if (jmax > nr) jmax = nr;
                first = false;
              } else if (rsq < rsmin) {
                lmin = l;
                rsmin = rsq;
              }
            }
            if (std::abs(ln) != l) {
              l = std::abs(ln);
              continue;
            }
            break;
          }
        }
      }
    }
    // the reference's exit test: its first clause compares i1 with itself
    if (imax <= i2 && j1 <= jmin && jmax <= j2) break;
    --i1;
    ++i2;
    --j1;
    ++j2;
  }
  if (first) {
    np = 0;
    dsq = 0.0;
  } else {
    np = lmin;
    dsq = rsmin;
    lnext[lmin - 1] = -lnext[lmin - 1];
  }
}

// givens (toms660.f90:289-394): plane rotation zeroing b; a <- r, b <- the
// value from which c and s can be recovered.
static void givens(double &a, double &b, double &c, double &s) {
  if (std::fabs(b) < std::fabs(a)) {
    const double u = 2.0 * a;
    const double v = b / u;
    const double r = std::sqrt(0.25 + v * v) * u;
    c = a / r;
    s = 2.0 * v * c;
    b = s;
    a = r;
  } else if (b != 0.0) {
    const double u = 2.0 * b;
    const double v = a / u;
    a = std::sqrt(0.25 + v * v) * u;
    s = b / a;
    c = 2.0 * v * s;
    b = c != 0.0 ? 1.0 / c : 1.0;
  } else {
    c = 1.0;
    s = 0.0;
  }
}

// rotate (toms660.f90:1426-1482): apply the rotation to the vector pair.
static void rotate(int n, double c, double s, double *x, double *y) {
  if (n <= 0) return;
  if (c == 1.0 && s == 0.0) return;
  for (int i = 0; i < n; ++i) {
    const double xi = x[i], yi = y[i];
    x[i] = c * xi + s * yi;
    y[i] = -(s * xi) + c * yi;
  }
}

// setup2 (toms660.f90:1484-1577): one weighted least-squares equation for the
// quadratic nodal function at node k.
static void setup2(double xk, double yk, double fk, double xi, double yi, double fi, double s1, double s2, double r,
                   double *row) {
  const double dx = xi - xk, dy = yi - yk;
  const double d = std::sqrt(dx * dx + dy * dy);
  if (d <= 0.0 || r <= d) {
    for (int i = 0; i < 6; ++i) row[i] = 0.0;
  } else {
    const double w = (r - d) / r / d;
    row[0] = dx * dx * w / s2;
    row[1] = dx * dy * w / s2;
    row[2] = dy * dy * w / s2;
    row[3] = dx * w / s1;
    row[4] = dy * w / s1;
    row[5] = (fi - fk) * w;
  }
}

// qshep2 (toms660.f90:679-1207): the nodal functions of every data point.
// Returns the reference's ier (0 ok, 1 bad sizes, 2 duplicate nodes,
// 3 collinear or ill-conditioned).
static int qshep2(int n, const double *x, const double *y, const double *f, int nq, int nw, int nr, int *lcell,
                  int *lnext, double &xmin, double &ymin, double &dx, double &dy, double &rmax, double *rsq,
                  double *a) {
  const double dtol = 0.01, rtol = 1.0e-5, sf = 1.0;
  const int nn = n, nnr = nr;
  const int nqwmax = nq > nw ? nq : nw;
  const int lmax = n - 1 < 40 ? n - 1 : 40;
  if (nq < 5 || nw < 1 || lmax < nqwmax || nr < 1) return 1;
  double xmn, ymn, ddx, ddy;
  if (store2(nn, x, y, nnr, lcell, lnext, xmn, ymn, ddx, ddy) != 0) {
    xmin = xmn;
    ymin = ymn;
    dx = ddx;
    dy = ddy;
    return 3;
  }
  double b[36];  // b(6,6) column-major: setup2 fills column irow
#define B(i, j) b[((j)-1) * 6 + ((i)-1)]
  int npts[40];
  double rsmx = 0.0;
  for (int k = 1; k <= nn; ++k) {
    const double xk = x[k - 1], yk = y[k - 1], fk = f[k - 1];
    lnext[k - 1] = -lnext[k - 1];
    double rs = 0.0, sum2 = 0.0, rws = 0.0, rq = 0.0, rsold = 0.0, avsq = 0.0, av, c, s;
    int lnp = 0, neq = 0, np = 0, i, irow;
  L1:
    sum2 = sum2 + rs;
    if (lnp == lmax) goto L3;
    lnp = lnp + 1;
    rsold = rs;
    getnp2(xk, yk, x, y, nnr, lcell, lnext, xmn, ymn, ddx, ddy, np, rs);
    if (rs == 0.0) return 2;
    npts[lnp - 1] = np;
    if ((rs - rsold) / rs < rtol) goto L1;
    if (rws == 0.0 && nw < lnp) rws = rs;
    if (rq == 0.0 && nq < lnp) {
      neq = lnp - 1;
      rq = std::sqrt(rs);
      avsq = sum2 / (double)neq;
    }
    if (nqwmax < lnp) goto L4;
    goto L1;
  L3:
    if (rws == 0.0) rws = 1.1 * rs;
    if (rq == 0.0) {
      neq = lmax;
      rq = std::sqrt(1.1 * rs);
      avsq = sum2 / (double)neq;
    }
  L4:
    rsq[k - 1] = rws;
    if (rws > rsmx) rsmx = rws;
    av = std::sqrt(avsq);
    i = 0;
  L5:
    i = i + 1;
    np = npts[i - 1];
    irow = i < 6 ? i : 6;
    setup2(xk, yk, fk, x[np - 1], y[np - 1], f[np - 1], av, avsq, rq, &B(1, irow));
    if (i == 1) goto L5;
    for (int j = 1; j <= irow - 1; ++j) {
      givens(B(j, j), B(j, irow), c, s);
      rotate(6 - j, c, s, &B(j + 1, j), &B(j + 1, irow));
    }
    if (i < neq) goto L5;
    {
      double dmin = std::fabs(B(1, 1));
      for (int t = 2; t <= 5; ++t)
        if (std::fabs(B(t, t)) < dmin) dmin = std::fabs(B(t, t));
      if (dtol <= dmin * rq) goto L13;
    }
    if (neq == lmax) goto L10;
  L7:
    rsold = rs;
    neq = neq + 1;
    if (neq == lmax) goto L9;
    if (neq != lnp) {
      np = npts[neq];  // npts(neq+1)
      const double ex = x[np - 1] - xk, ey = y[np - 1] - yk;
      rs = ex * ex + ey * ey;
      if ((rs - rsold) / rs < rtol) goto L7;
      rq = std::sqrt(rs);
      goto L5;
    }
    lnp = lnp + 1;
    getnp2(xk, yk, x, y, nnr, lcell, lnext, xmn, ymn, ddx, ddy, np, rs);
    if (np == 0) return 2;
    npts[lnp - 1] = np;
    if ((rs - rsold) / rs < rtol) goto L7;
    rq = std::sqrt(rs);
    goto L5;
  L9:
    rq = std::sqrt(1.1 * rs);
    goto L5;
  L10:
    for (int ii = 1; ii <= 3; ++ii) {
      B(ii, 6) = sf;
      for (int t = ii + 1; t <= 6; ++t) B(t, 6) = 0.0;
      for (int j = ii; j <= 5; ++j) {
        givens(B(j, j), B(j, 6), c, s);
        rotate(6 - j, c, s, &B(j + 1, j), &B(j + 1, 6));
      }
    }
    {
      double dmin = std::fabs(B(1, 1));
      for (int t = 2; t <= 5; ++t)
        if (std::fabs(B(t, t)) < dmin) dmin = std::fabs(B(t, t));
      if (dmin * rq < dtol) {
        xmin = xmn;
        ymin = ymn;
        dx = ddx;
        dy = ddy;
        return 3;
      }
    }
  L13:
    for (int ii = 5; ii >= 1; --ii) {
      double t = 0.0;
      for (int j = ii + 1; j <= 5; ++j) t = t + B(j, ii) * a[(k - 1) * 5 + (j - 1)];
      a[(k - 1) * 5 + (ii - 1)] = (B(6, ii) - t) / B(ii, ii);
    }
    for (int t = 0; t < 3; ++t) a[(k - 1) * 5 + t] = a[(k - 1) * 5 + t] / avsq;
    a[(k - 1) * 5 + 3] = a[(k - 1) * 5 + 3] / av;
    a[(k - 1) * 5 + 4] = a[(k - 1) * 5 + 4] / av;
    lnext[k - 1] = -lnext[k - 1];
    for (int t = 1; t <= lnp; ++t) {
      const int p = npts[t - 1];
      lnext[p - 1] = -lnext[p - 1];
    }
  }
#undef B
  xmin = xmn;
  ymin = ymn;
  dx = ddx;
  dy = ddy;
  rmax = std::sqrt(rsmx);
  return 0;
}

// qs2val (toms660.f90:1209-1424): the interpolant at (px, py).
static double qs2val(const Fit &q, double px, double py) {
  const double xp = px, yp = py;
  const int n = q.n, nr = q.nr;
  if (n < 6 || nr < 1 || q.dx <= 0.0 || q.dy <= 0.0 || q.rmax < 0.0) return 0.0;
  int imin = (int)((xp - q.xmin - q.rmax) / q.dx) + 1;
  if (imin < 1) imin = 1;
  int imax = (int)((xp - q.xmin + q.rmax) / q.dx) + 1;
  if (imax > nr) imax = nr;
  int jmin = (int)((yp - q.ymin - q.rmax) / q.dy) + 1;
  if (jmin < 1) jmin = 1;
  int jmax = (int)((yp - q.ymin + q.rmax) / q.dy) + 1;
  if (jmax > nr) jmax = nr;
  if (imax < imin || jmax < jmin) return 0.0;
  double sw = 0.0, swq = 0.0;
  for (int j = jmin; j <= jmax; ++j) {
    for (int i = imin; i <= imax; ++i) {
      int k = q.lcell[(j - 1) * nr + (i - 1)];
      if (k == 0) continue;
      for (;;) {
        const double delx = xp - q.x[k - 1], dely = yp - q.y[k - 1];
        const double ds = delx * delx + dely * dely;
        const double rs = q.rsq[k - 1];
        if (ds < rs) {
          if (ds == 0.0) return q.f[k - 1];
          const double rds = rs * ds;
          const double rd = std::sqrt(rds);
          const double w = (rs + ds - rd - rd) / rds;
          const double *ak = &q.a[(k - 1) * 5];
          sw = sw + w;
          swq = swq + w * (q.f[k - 1] + ak[0] * delx * delx + ak[1] * delx * dely + ak[2] * dely * dely +
                           ak[3] * delx + ak[4] * dely);
        }
        const int kp = k;
        k = q.lnext[kp - 1];
        if (k == kp) break;
      }
    }
  }
  return sw == 0.0 ? 0.0 : swq / sw;
}

// The reference's call: nr = int(sqrt(float(n)/3.)) in single precision
// (geometry.f90:454-455), then qshep2 on copies of the points.
static int fit(Fit &q, int n, const double *x, const double *y, const double *f, int nq, int nw) {
  q.n = n;
  q.nr = (int)std::sqrt((float)n / 3.0f);
  q.x.assign(x, x + n);
  q.y.assign(y, y + n);
  q.f.assign(f, f + n);
  q.rsq.assign(n, 0.0);
  q.a.assign(5 * (size_t)n, 0.0);
  q.lnext.assign(n, 0);
  q.lcell.assign((size_t)(q.nr > 0 ? q.nr * q.nr : 0), 0);
  q.ier = qshep2(n, q.x.data(), q.y.data(), q.f.data(), nq, nw, q.nr, q.lcell.data(), q.lnext.data(), q.xmin,
                 q.ymin, q.dx, q.dy, q.rmax, q.rsq.data(), q.a.data());
  return q.ier;
}

}  // namespace shep
