// Input files and lattice generation (host only; no device use):
//   * 3DEM_FEM.INP (input(), read_input.f90:37-201) and PARAM.INP
//     (MoVFEM_3DMT.f90:302-334), read with Fortran list-directed rules;
//   * grid_3d (geometry.f90:48-122): extension zones sized by the skin
//     depth, interface layering and the lattice heights from the Shepard
//     interpolant of each interface (zblocks, gauss_grid, complete_grid), and
//     the inner conductivity/permeability model copied onto the lattice by
//     nearest model point (innermodel_gqg, min_dd_inner, assign_model).
//
// The lattice must equal the reference's bit for bit, so each step keeps the
// reference's arithmetic (operation order, float32 where it rounds) and its
// write order where later writes overwrite earlier ones.  Two things are
// done once instead of per lattice column: the Shepard fits, which the
// reference recomputes for every (x, y) cell from the same points, and the
// nearest-model search, which depends on the node's coordinates only and is
// run once per node (threads over nodes).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "mvf_internal.h"
#include "mvf_shepard.h"

#pragma float_control(push)
#pragma clang fp contract(off)

struct mvf_mesh {
  int64_t dims[12];
  double vals[6];
  std::vector<double> xp, yp, zp, sigma, mu, xsr, ysr, zsr;
  std::vector<int32_t> nzl;
};

struct mvf_inputs {
  bool has_inp = false, has_param = false;
  // 3DEM_FEM.INP
  double dx = 0, dy = 0, dz = 0, high = 0, depth = 0, bcksigma = 0;
  int32_t mn = 0, nsr = 0, nsf = 0, imu = 0, isigma = 0, mx = 0, my = 0, mz = 0;
  std::vector<double> xsr, ysr, zsr, xto, yto, zto, xm, ym, zm, sigma, mu;
  std::vector<int32_t> nsp, ijsigma, ijmu;
  // PARAM.INP
  int32_t nf = 0, boundary = 0, inimod = 0, nl = 0, gpml_sch = 0;
  double h_sigma = 0, a0 = 0, b0 = 0, nn = 0;
  std::vector<double> freqs, l_sigma, l_dz;
};

namespace mvf {
namespace {

int host_fail(const std::string &m) {
  g_host_err = m;
  return MVF_E_ARG;
}

constexpr double kPi = MVF_PI;
constexpr double kEps0 = MVF_EPS0;
constexpr double kMu0 = 4.0 * kPi * 1.0e-7;

// ----------------------------------------------------------------------------
// Fortran list-directed input (the subset the input files use): values
// separated by blanks, tabs or commas, r*c repeats, each READ statement
// starting at the next record and leaving the rest of its last record unread;
// a READ with an empty list consumes one record.
// ----------------------------------------------------------------------------
struct LdFile {
  std::vector<std::string> rec;
  size_t pos = 0;
  std::string name, err;

  bool load(const char *path) {
    name = path;
    std::ifstream f(path);
    if (!f) {
      err = std::string("cannot open ") + path;
      return false;
    }
    std::string s;
    while (std::getline(f, s)) {
      if (!s.empty() && s.back() == '\r') s.pop_back();
      rec.push_back(s);
    }
    return true;
  }
  bool fail(const std::string &m) {
    err = name + ":" + std::to_string(pos) + ": " + m;
    return false;
  }
  // read(u,'(a)') / read(u,*) with an empty list
  bool skip() {
    if (pos >= rec.size()) return fail("unexpected end of file");
    ++pos;
    return true;
  }
  bool line(std::string &s) {
    if (pos >= rec.size()) return fail("unexpected end of file");
    s = rec[pos++];
    return true;
  }
  // n value strings of one list-directed READ statement
  bool values(size_t n, std::vector<std::string> &out) {
    out.clear();
    if (n == 0) return skip();
    while (out.size() < n) {
      if (pos >= rec.size()) return fail("unexpected end of file");
      const std::string &r = rec[pos++];
      size_t i = 0;
      while (i < r.size() && out.size() < n) {
        const char c = r[i];
        if (c == ' ' || c == '\t' || c == ',') {
          // a comma after a value is a separator; two commas would be a null value
          if (c == ',') {
            size_t k = i + 1;
            while (k < r.size() && (r[k] == ' ' || r[k] == '\t')) ++k;
            if (k < r.size() && r[k] == ',') return fail("null values are not supported");
          }
          ++i;
          continue;
        }
        if (c == '/') return fail("a slash ends the list before all values were read");
        size_t j = i;
        while (j < r.size() && r[j] != ' ' && r[j] != '\t' && r[j] != ',' && r[j] != '/') ++j;
        std::string tok = r.substr(i, j - i);
        i = j;
        const size_t star = tok.find('*');
        if (star != std::string::npos) {
          const std::string cnt = tok.substr(0, star), val = tok.substr(star + 1);
          char *e = nullptr;
          const long rep = std::strtol(cnt.c_str(), &e, 10);
          if (cnt.empty() || *e || rep < 1) return fail("bad repeat count '" + tok + "'");
          if (val.empty()) return fail("null values are not supported");
          for (long t = 0; t < rep && out.size() < n; ++t) out.push_back(val);
        } else {
          out.push_back(tok);
        }
      }
    }
    return true;
  }
  bool to_int(const std::string &t, int32_t &v) {
    size_t i = 0;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
    if (i == t.size()) return fail("expected an integer, found '" + t + "'");
    for (size_t k = i; k < t.size(); ++k)
      if (t[k] < '0' || t[k] > '9') return fail("expected an integer, found '" + t + "'");
    const long long x = std::strtoll(t.c_str(), nullptr, 10);
    if (x < INT32_MIN || x > INT32_MAX) return fail("integer out of range '" + t + "'");
    v = (int32_t)x;
    return true;
  }
  // Fortran real: [sign] digits [. digits] [(E|D|Q) [sign] digits | sign digits]
  bool to_real(const std::string &t, double &v) {
    std::string s;
    size_t i = 0;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) s += t[i++];
    size_t nd = 0;
    while (i < t.size() && t[i] >= '0' && t[i] <= '9') s += t[i++], ++nd;
    if (i < t.size() && t[i] == '.') {
      s += t[i++];
      while (i < t.size() && t[i] >= '0' && t[i] <= '9') s += t[i++], ++nd;
    }
    if (nd == 0) return fail("expected a real, found '" + t + "'");
    if (i < t.size()) {
      const char c = t[i];
      if (c == 'e' || c == 'E' || c == 'd' || c == 'D' || c == 'q' || c == 'Q') ++i;
      else if (c != '+' && c != '-') return fail("expected a real, found '" + t + "'");
      s += 'e';
      if (i < t.size() && (t[i] == '+' || t[i] == '-')) s += t[i++];
      size_t ne = 0;
      while (i < t.size() && t[i] >= '0' && t[i] <= '9') s += t[i++], ++ne;
      if (ne == 0 || i != t.size()) return fail("expected a real, found '" + t + "'");
    }
    v = std::strtod(s.c_str(), nullptr);
    return true;
  }
  bool ints(std::vector<int32_t *> vs) {
    std::vector<std::string> tk;
    if (!values(vs.size(), tk)) return false;
    for (size_t k = 0; k < vs.size(); ++k)
      if (!to_int(tk[k], *vs[k])) return false;
    return true;
  }
  bool reals(double *v, size_t n) {
    std::vector<std::string> tk;
    if (!values(n, tk)) return false;
    for (size_t k = 0; k < n; ++k)
      if (!to_real(tk[k], v[k])) return false;
    return true;
  }
  // Iw field of a formatted record, blanks ignored (an all-blank field is 0)
  bool ifield(const std::string &r, size_t at, int32_t &v) {
    const char c = at < r.size() ? r[at] : ' ';
    if (c == ' ') v = 0;
    else if (c >= '0' && c <= '9') v = c - '0';
    else return fail(std::string("expected a digit in column ") + std::to_string(at + 1));
    return true;
  }
};

// input() (read_input.f90:37-201)
bool read_3dem(LdFile &f, mvf_inputs &in) {
  double p[5];
  std::vector<std::string> tk;
  if (!f.skip() || !f.values(7, tk)) return false;
  for (int k = 0; k < 5; ++k)
    if (!f.to_real(tk[k], p[k])) return false;
  if (!f.to_int(tk[5], in.mn) || !f.to_real(tk[6], in.bcksigma)) return false;
  in.dx = p[0], in.dy = p[1], in.dz = p[2], in.high = p[3], in.depth = p[4];
  // sources/receivers (read_input.f90:56-68)
  if (!f.skip() || !f.ints({&in.nsr})) return false;
  if (in.nsr < 0) return f.fail("negative source/receiver count");
  in.xsr.resize(in.nsr), in.ysr.resize(in.nsr), in.zsr.resize(in.nsr);
  for (int i = 0; i < in.nsr; ++i) {
    if (!f.values(4, tk)) return false;
    int32_t n;
    if (!f.to_int(tk[0], n) || !f.to_real(tk[1], in.xsr[i]) || !f.to_real(tk[2], in.ysr[i]) ||
        !f.to_real(tk[3], in.zsr[i]))
      return false;
  }
  // interfaces (read_input.f90:72-101): slots 2..nsf0+1 of nsf = nsf0 + 3
  int32_t nsf0;
  if (!f.skip() || !f.ints({&nsf0})) return false;
  if (nsf0 < 1) return f.fail("at least one interface is needed");
  in.nsf = nsf0 + 3;
  in.nsp.assign(in.nsf, 0);
  {
    std::vector<int32_t *> ps;
    for (int l = 1; l <= nsf0; ++l) ps.push_back(&in.nsp[l]);
    if (!f.ints(ps)) return false;
  }
  int nmax = 0;
  for (int l = 1; l <= nsf0; ++l) {
    if (in.nsp[l] < 0) return f.fail("negative interface point count");
    nmax = std::max(nmax, (int)in.nsp[l]);
  }
  in.xto.assign((size_t)in.nsf * nmax, 0.0);
  in.yto = in.xto, in.zto = in.xto;
  for (int l = 1; l <= nsf0; ++l) {
    if (!f.skip()) return false;
    for (int j = 0; j < in.nsp[l]; ++j) {
      if (!f.values(4, tk)) return false;
      int32_t n;
      const size_t at = (size_t)j * in.nsf + l;
      if (!f.to_int(tk[0], n) || !f.to_real(tk[1], in.xto[at]) || !f.to_real(tk[2], in.yto[at]) ||
          !f.to_real(tk[3], in.zto[at]))
        return false;
    }
  }
  // inner model (read_input.f90:105-162)
  if (!f.skip() || !f.ints({&in.imu, &in.isigma, &in.mx, &in.my, &in.mz})) return false;
  if (in.isigma < 1 || in.imu < 0 || in.mx < 1 || in.my < 1 || in.mz < 1)
    return f.fail("bad inner-model sizes");
  in.ijmu.assign(2 * (size_t)in.imu, 0);
  in.ijsigma.assign(2 * (size_t)in.isigma, 0);
  std::string r;
  for (int i = 0; i < in.imu; ++i)  // (a2,i1,i1)
    if (!f.line(r) || !f.ifield(r, 2, in.ijmu[i]) || !f.ifield(r, 3, in.ijmu[in.imu + i])) return false;
  for (int i = 0; i < in.isigma; ++i)  // (a5,i1,i1)
    if (!f.line(r) || !f.ifield(r, 5, in.ijsigma[i]) || !f.ifield(r, 6, in.ijsigma[in.isigma + i])) return false;
  in.xm.resize(in.mx), in.ym.resize(in.my);
  if (!f.skip() || !f.reals(in.xm.data(), in.mx) || !f.reals(in.ym.data(), in.my)) return false;
  const size_t n = (size_t)in.mx * in.my * in.mz;
  in.zm.resize(n);
  if (!f.skip()) return false;
  for (size_t c = 0; c < (size_t)in.mx * in.my; ++c)
    if (!f.reals(&in.zm[c * in.mz], in.mz)) return false;
  // conductivity and permeability blocks (read_input.f90:166-201), (ncomp, n)
  std::vector<double> col(in.mz);
  auto block = [&](int ncomp, std::vector<double> &dst) {
    dst.assign((size_t)ncomp * n, 0.0);
    for (int ic = 0; ic < ncomp; ++ic) {
      if (!f.skip()) return false;
      for (size_t c = 0; c < (size_t)in.mx * in.my; ++c) {
        if (!f.reals(col.data(), in.mz)) return false;
        for (int k = 0; k < in.mz; ++k) dst[(c * in.mz + k) * ncomp + ic] = col[k];
      }
    }
    return true;
  };
  if (!block(in.isigma, in.sigma)) return false;
  if (in.imu != 0) {
    if (!block(in.imu, in.mu)) return false;
  } else {
    in.mu.assign(n, 1.0);
  }
  in.has_inp = true;
  return true;
}

// PARAM.INP (MoVFEM_3DMT.f90:302-334)
bool read_param(LdFile &f, mvf_inputs &in) {
  if (!f.skip() || !f.ints({&in.nf})) return false;
  if (in.nf < 1) return f.fail("at least one frequency is needed");
  in.freqs.resize(in.nf);
  if (!f.reals(in.freqs.data(), in.nf)) return false;
  if (!f.skip() || !f.ints({&in.boundary}) || !f.skip() || !f.ints({&in.inimod})) return false;
  if (!f.skip() || !f.reals(&in.h_sigma, 1) || !f.skip() || !f.ints({&in.nl})) return false;
  if (in.nl < 1) return f.fail("at least one layer is needed");
  in.l_sigma.resize(in.nl);
  in.l_dz.assign(in.nl, 0.0);
  if (!f.skip() || !f.reals(in.l_sigma.data(), in.nl)) return false;
  if (!f.skip() || !f.reals(in.l_dz.data(), in.nl - 1)) return false;
  in.l_dz.resize(in.nl - 1);
  if (in.boundary == 0) {
    if (!f.skip() || !f.ints({&in.gpml_sch})) return false;
    double v[3];
    if (!f.skip() || !f.reals(v, 3)) return false;
    in.a0 = v[0], in.b0 = v[1], in.nn = v[2];
  }
  in.has_param = true;
  return true;
}

// ----------------------------------------------------------------------------
// grid_3d (geometry.f90:48-122)
// ----------------------------------------------------------------------------
struct Slot {
  std::vector<double> x, y, z;
};

struct Grid3d {
  const mvf_mesh_input &in;
  mvf_mesh &out;
  std::string err;
  int nord = 2, nsf = 0, nextd = 2, gnx = 0, gny = 0, gnz = 0;
  int64_t nnx = 0, nny = 0, nnz = 0, npt = 0, nyz = 0;
  double omega = 0, xmin = 0, ymin = 0, zmin = 0, xmax = 0, ymax = 0, zmax = 0, zbot = 0, ztop = 0;
  double x0 = 0, y0 = 0, z0 = 0, gztop = 0, skin = 0;
  double fact = 1.3;
  std::vector<double> x, y, dmz, xsr, ysr, zsr, xm, ym, zm, asx, asy, asz;
  std::vector<Slot> slot;  // slot[l-1] = interface l
  std::vector<int> nzl;    // nzl[l-1]
  std::map<std::tuple<int, int, int, int>, shep::Fit> fits;

  Grid3d(const mvf_mesh_input &i, mvf_mesh &o) : in(i), out(o) {}

  bool fail(const std::string &m) {
    err = m;
    return false;
  }

  // min_max_input (geometry.f90:159-219); the y range takes xsr where ysr
  // is compared (:178-179)
  void min_max_input() {
    xmin = 1.e20, xmax = -1.e20, ymin = 1.e20, ymax = -1.e20, zmin = 1.e20, zmax = -1.e20;
    double x1 = 1.e20, x2 = -1.e20, y1 = 1.e20, y2 = -1.e20, zl = 1.e20, zu = -1.e20;
    for (int i = 0; i < in.nsr; ++i) {
      if (xsr[i] < xmin) xmin = xsr[i];
      if (xsr[i] > xmax) xmax = xsr[i];
      if (ysr[i] < ymin) ymin = xsr[i];
      if (ysr[i] > ymax) ymax = xsr[i];
      if (zsr[i] < zmin) zmin = zsr[i];
      if (zsr[i] > zmax) zmax = zsr[i];
    }
    for (int l = 2; l <= nsf - 2; ++l) {
      const Slot &s = slot[l - 1];
      for (size_t j = 0; j < s.x.size(); ++j) {
        if (s.x[j] < x1) x1 = s.x[j];
        if (s.x[j] > x2) x2 = s.x[j];
        if (s.y[j] < y1) y1 = s.y[j];
        if (s.y[j] > y2) y2 = s.y[j];
        if (s.z[j] < zl) zl = s.z[j];
        if (s.z[j] > zu) zu = s.z[j];
      }
    }
    xmin = std::min(xmin, x1), xmax = std::max(xmax, x2);
    ymin = std::min(ymin, y1), ymax = std::max(ymax, y2);
    zu = zu + in.high, zl = zl - in.depth;
    zmin = std::min(zmin, zl), zmax = std::max(zmax, zu);
    for (double v : xm) {
      if (v < xmin) xmin = v;
      if (v > xmax) xmax = v;
    }
    for (double v : ym) {
      if (v < ymin) ymin = v;
      if (v > ymax) ymax = v;
    }
    for (double v : zm) {
      if (v < zmin) zmin = v;
      if (v > zmax) zmax = v;
    }
  }

  // extension (geometry.f90:223-342)
  bool extension() {
    double minf = 1.e20;
    nextd = 2;
    double ddx = std::min(std::min(xmax - xmin, ymax - ymin), zmax - zmin);
    for (int i = 0; i < in.nf; ++i)
      if (in.freqs[i] < minf) minf = in.freqs[i];
    skin = std::sqrt(2.0 / (kMu0 * in.sigma_bg * 2.0 * kPi * minf));
    // depth factor; for 1.5 <= ratio < 2.5 the reference assigns none
    // (geometry.f90:245-251) and its build takes the last branch's 3
    double nd = 3;
    if (ddx / skin >= 2.5) nd = 1.5;
    else if (ddx / skin < 2.5 && skin / ddx >= 1.5) nd = 2;
    else if (ddx / skin < 1.5) nd = 3;
    ddx = std::min(std::min(in.dx, in.dy), in.dz);
    for (;;) {
      double sum_ex = 0.0;
      for (int i = 1; i <= nextd; ++i) sum_ex = sum_ex + fact * ddx * (double)i;
      if (sum_ex > nd * skin) break;
      ++nextd;
      if (nextd > 100000) return fail("extension zone does not converge");
    }
    std::vector<double> dmx(nextd + 1), dmy(nextd + 1);
    double x1 = xmin, x2 = xmax, y1 = ymin, y2 = ymax;
    for (int i = 1; i <= nextd; ++i) {
      dmx[i] = fact * (double)i * in.dx;
      x1 = x1 - dmx[i];
      x2 = x2 + dmx[i];
    }
    for (int i = 1; i <= nextd; ++i) {
      dmy[i] = fact * (double)i * in.dy;
      y1 = y1 - dmy[i];
      y2 = y2 + dmy[i];
    }
    x0 = 0.5 * (x1 + x2);
    y0 = 0.5 * (y1 + y2);
    x1 = x1 - x0, x2 = x2 - x0;
    y1 = y1 - y0, y2 = y2 - y0;
    xmin = xmin - x0, xmax = xmax - x0;
    ymin = ymin - y0, ymax = ymax - y0;
    const double fx = (xmax - xmin) / in.dx, fy = (ymax - ymin) / in.dy;
    if (!(fx >= 0 && fx < 1e8 && fy >= 0 && fy < 1e8)) return fail("inner extent / cell size out of range");
    const int nx0 = (int)fx + 1, ny0 = (int)fy + 1;
    gnx = nx0 + 2 * nextd;
    gny = ny0 + 2 * nextd;
    x.assign(gnx + 1, 0.0);  // 1-based
    y.assign(gny + 1, 0.0);
    x[1] = x1;
    y[1] = y1;
    for (int i = 1; i <= nextd; ++i) x[i + 1] = x[i] + dmx[nextd + 1 - i];
    for (int i = 1; i <= nextd; ++i) y[i + 1] = y[i] + dmy[nextd + 1 - i];
    for (int i = 1; i <= nx0; ++i) x[i + nextd] = x[nextd + 1] + (double)(i - 1) * in.dx;
    for (int i = 1; i <= ny0; ++i) y[i + nextd] = y[nextd + 1] + (double)(i - 1) * in.dy;
    int i0 = nextd + nx0;
    for (int i = 1; i <= nextd; ++i) x[i0 + i] = x[i0 + i - 1] + dmx[i];
    i0 = nextd + ny0;
    for (int i = 1; i <= nextd; ++i) y[i0 + i] = y[i0 + i - 1] + dmy[i];
    dmz.assign(nextd + 1, 0.0);
    zbot = zmin, ztop = zmax;
    for (int i = 1; i <= nextd; ++i) {
      dmz[i] = fact * (double)i * in.dz;
      zbot = zbot - dmz[i];
      ztop = ztop + dmz[i];
    }
    // bottom, air and top interfaces (geometry.f90:326-337)
    slot[0] = slot[1];
    std::fill(slot[0].z.begin(), slot[0].z.end(), zmin);
    slot[nsf - 2] = slot[nsf - 3];
    std::fill(slot[nsf - 2].z.begin(), slot[nsf - 2].z.end(), zmax);
    slot[nsf - 1] = slot[nsf - 2];
    std::fill(slot[nsf - 1].z.begin(), slot[nsf - 1].z.end(), ztop);
    z0 = zbot;
    zmin = zmin - z0, zmax = zmax - z0;
    return true;
  }

  // coord_transform (geometry.f90:364-405)
  void coord_transform() {
    for (int i = 0; i < in.nsr; ++i) xsr[i] = xsr[i] - x0, ysr[i] = ysr[i] - y0, zsr[i] = zsr[i] - z0;
    for (Slot &s : slot)
      for (size_t j = 0; j < s.x.size(); ++j) s.x[j] = s.x[j] - x0, s.y[j] = s.y[j] - y0, s.z[j] = s.z[j] - z0;
    double m = 1.e20;
    for (double v : slot[nsf - 3].z)
      if (v < m) m = v;
    gztop = m;
    for (double &v : xm) v = v - x0;
    for (double &v : ym) v = v - y0;
    for (double &v : zm) v = v - z0;
  }

  // Shepard fit of interface l (zero heights when zero) with the pair's nq, nw
  const shep::Fit *fit(int l, bool zero, int nq, int nw) {
    const auto key = std::make_tuple(l, (int)zero, nq, nw);
    auto it = fits.find(key);
    if (it != fits.end()) return &it->second;
    const Slot &s = slot[l - 1];
    const int n = (int)s.x.size();
    std::vector<double> f = zero ? std::vector<double>(n, 0.0) : s.z;
    shep::Fit &q = fits[key];
    const int ier = shep::fit(q, n, s.x.data(), s.y.data(), f.data(), nq, nw);
    if (ier != 0) {
      static const char *why[] = {"", "too few points (at least 20 are needed)", "duplicate points",
                                  "collinear or ill-conditioned points"};
      err = "interface " + std::to_string(l) + ": the Shepard fit fails (qshep2 ier=" + std::to_string(ier) +
            ", " + why[ier] + "); the reference continues with undefined coefficients there";
      return nullptr;
    }
    return &q;
  }
  bool pair(int l, const shep::Fit *&f1, const shep::Fit *&f2) {
    const int np1 = (int)slot[(l == 1 ? 1 : l - 1) - 1].x.size(), np2 = (int)slot[l - 1].x.size();
    const int nq = std::min(13, std::min(np1, np2)), nw = std::min(19, std::min(np1, np2));
    f2 = fit(l, false, nq, nw);
    f1 = l == 1 ? fit(1, true, nq, nw) : fit(l - 1, false, nq, nw);
    return f1 && f2;
  }

  // zblocks (geometry.f90:411-470)
  bool zblocks() {
    nzl.assign(nsf, 0);
    nzl[0] = nextd;
    nzl[nsf - 1] = nextd;
    for (int l = 2; l <= nsf - 1; ++l) {
      const shep::Fit *f1, *f2;
      if (!pair(l, f1, f2)) return false;
      double h = 0.0;
      for (int i = nextd + 1; i <= gnx - nextd - 1; ++i) {
        const double xi = x[i];
        for (int j = nextd + 1; j <= gny - nextd - 1; ++j) {
          const double yj = y[j];
          const double h2 = shep::qs2val(*f2, xi, yj);
          const double h1 = shep::qs2val(*f1, xi, yj);
          const double dh = h2 - h1;
          if (dh > h) h = dh;
        }
      }
      const double q = h / in.dz;
      if (q >= 1e7) return fail("layer " + std::to_string(l) + " needs too many cells");
      nzl[l - 1] = std::max((int)q, 1);
    }
    gnz = 1;
    for (int l = 0; l < nsf; ++l) gnz += nzl[l];
    return true;
  }

  // gauss_grid (geometry.f90:474-577) and local_gqg (:743-766)
  bool gauss_grid() {
    nnx = (int64_t)(gnx - 1) * (nord - 1) + 1;
    nny = (int64_t)(gny - 1) * (nord - 1) + 1;
    nnz = (int64_t)(gnz - 1) * (nord - 1) + 1;
    npt = nnx * nny * nnz;
    nyz = nny * nnz;
    if (npt > ((int64_t)1 << 31) - 1) return fail("lattice exceeds 2^31-1 nodes");
    out.xp.assign(nnx, 0.0);
    out.yp.assign(nny, 0.0);
    out.zp.assign(npt, 0.0);
    // gqg_nodes (geometry.f90:714-739)
    const double fn = 2 / (double)(nord - 1);
    asx.assign(nord + 1, 0.0);
    for (int j = 1; j <= nord; ++j) asx[j] = -1 + (double)(j - 1) * fn;
    asy = asx, asz = asx;
    std::vector<const shep::Fit *> f1(nsf + 1), f2(nsf + 1);
    for (int l = 1; l <= nsf; ++l)
      if (!pair(l, f1[l], f2[l])) return false;
    const int no2 = nord * nord;
    std::vector<double> z1(no2), z2(no2), dm(no2);
    double *xp = out.xp.data(), *yp = out.yp.data(), *zp = out.zp.data();
    for (int i = nextd + 1; i <= gnx - nextd - 1; ++i) {
      const double x1 = x[i], x2 = x[i + 1];
      for (int j = nextd + 1; j <= gny - nextd + 1; ++j) {  // the reference's y range (geometry.f90:516)
        const double y1 = y[j], y2 = y[j + 1];
        std::fill(z2.begin(), z2.end(), 0.0);
        int k = 0;
        for (int l = 1; l <= nsf; ++l) {
          // per-layer increment at each (x, y) point: the same for every sub-layer
          for (int i1 = 1; i1 <= nord; ++i1) {
            const double xi = 0.5 * (x2 - x1) * asx[i1] + 0.5 * (x1 + x2);
            for (int j1 = 1; j1 <= nord; ++j1) {
              const double yj = 0.5 * (y2 - y1) * asy[j1] + 0.5 * (y1 + y2);
              const double h2 = shep::qs2val(*f2[l], xi, yj);
              const double h1 = shep::qs2val(*f1[l], xi, yj);
              dm[(i1 - 1) * nord + (j1 - 1)] = (h2 - h1) / (double)nzl[l - 1];
            }
          }
          for (int l1 = 1; l1 <= nzl[l - 1]; ++l1) {
            ++k;
            const int64_t no = (int64_t)(i - 1) * (nord - 1) * nyz + (int64_t)(j - 1) * (nord - 1) * nnz +
                               (int64_t)(k - 1) * (nord - 1) + 1;
            for (int t = 0; t < no2; ++t) {
              z1[t] = z2[t];
              z2[t] = z1[t] + dm[t];
            }
            for (int i1 = 1; i1 <= nord; ++i1) {
              const double xi = 0.5 * (x2 - x1) * asx[i1] + 0.5 * (x1 + x2);
              xp[(i - 1) * (nord - 1) + i1 - 1] = xi;
              for (int j1 = 1; j1 <= nord; ++j1) {
                const double yj = 0.5 * (y2 - y1) * asy[j1] + 0.5 * (y1 + y2);
                yp[(j - 1) * (nord - 1) + j1 - 1] = yj;
                const double a = z1[(i1 - 1) * nord + j1 - 1], b = z2[(i1 - 1) * nord + j1 - 1];
                for (int k1 = 1; k1 <= nord; ++k1) {
                  const double zk = 0.5 * (b - a) * asz[k1] + 0.5 * (a + b);
                  zp[no + (int64_t)(i1 - 1) * nyz + (int64_t)(j1 - 1) * nnz + (k1 - 1) - 1] = zk;
                }
              }
            }
          }
        }
      }
    }
    complete_grid();
    return true;
  }

  // complete_grid (geometry.f90:579-710): the extension columns copy the
  // nearest inner column's heights
  void complete_grid() {
    double *xp = out.xp.data(), *yp = out.yp.data(), *zp = out.zp.data();
    const int s = nord - 1;
    auto nodeno = [&](int i, int j, int k) {
      return (int64_t)(i - 1) * s * nyz + (int64_t)(j - 1) * s * nnz + (int64_t)(k - 1) * s + 1;
    };
    auto copy_col = [&](int64_t no, int64_t no1, int l, int l1, int m, int m1) {
      for (int n = 1; n <= nord; ++n) {
        const int64_t id = no + (l - 1) * nyz + (m - 1) * nnz + (n - 1);
        const int64_t id1 = no1 + (l1 - 1) * nyz + (m1 - 1) * nnz + (n - 1);
        zp[id - 1] = zp[id1 - 1];
      }
    };
    // (y,z) planes
    for (int side = 0; side < 2; ++side) {
      const int ia = side == 0 ? 1 : gnx - nextd, ib = side == 0 ? nextd : gnx - 1;
      const int i1 = side == 0 ? nextd + 1 : gnx - nextd - 1, l1 = side == 0 ? 1 : nord;
      for (int i = ia; i <= ib; ++i) {
        const double x1 = x[i], x2 = x[i + 1];
        for (int j = nextd + 1; j <= gny - nextd - 1; ++j)
          for (int k = 1; k <= gnz - 1; ++k) {
            const int64_t no = nodeno(i, j, k), no1 = nodeno(i1, j, k);
            for (int l = 1; l <= nord; ++l) {
              xp[(i - 1) * s + l - 1] = 0.5 * (x2 - x1) * asx[l] + 0.5 * (x1 + x2);
              for (int m = 1; m <= nord; ++m) {
                const int64_t jj = (int64_t)(j - 1) * s + m;
                yp[jj - 1] = yp[(int64_t)(j - 1) * s + m - 1];
                copy_col(no, no1, l, l1, m, m);
              }
            }
          }
      }
    }
    // (x,z) planes; the y offsets use asx (geometry.f90:668,697)
    for (int side = 0; side < 2; ++side) {
      const int ja = side == 0 ? 1 : gny - nextd, jb = side == 0 ? nextd : gny - 1;
      const int j1 = side == 0 ? nextd + 1 : gny - nextd - 1, m1 = side == 0 ? 1 : nord;
      for (int i = 1; i <= gnx - 1; ++i)
        for (int j = ja; j <= jb; ++j) {
          const double y1 = y[j], y2 = y[j + 1];
          for (int k = 1; k <= gnz - 1; ++k) {
            const int64_t no = nodeno(i, j, k), no1 = nodeno(i, j1, k);
            for (int l = 1; l <= nord; ++l) {
              const int64_t ii = (int64_t)(i - 1) * s + l;
              xp[ii - 1] = xp[(int64_t)(i - 1) * s + l - 1];
              for (int m = 1; m <= nord; ++m) {
                yp[(j - 1) * s + m - 1] = 0.5 * (y2 - y1) * asx[m] + 0.5 * (y1 + y2);
                copy_col(no, no1, l, l, m, m1);
              }
            }
          }
        }
    }
  }

  // assign_model (geometry.f90:1033-1079) at 0-based node id from model point p
  void assign_model(int64_t id, int64_t p) {
    double *sg = &out.sigma[12 * id];
    double *mu = &out.mu[6 * id];
    const double ie = (double)(float)(kEps0 * omega);  // cmplx(0.d0, eps*omega): complex64
    for (int c = 0; c < 12; ++c) sg[c] = 0.0;
    if (in.isigma == 1) {
      const double s = in.sigma[p];
      for (int c : {0, 3, 5}) sg[2 * c] = s, sg[2 * c + 1] = ie;
    } else {
      for (int i = 0; i < in.isigma; ++i) {
        const int a = in.ijsigma[i], b = in.ijsigma[in.isigma + i];
        const int t = a == 1 ? b : (a == 2 || a == 3) ? a + b : 0;
        if (t) sg[2 * (t - 1)] = in.sigma[p * in.isigma + i], sg[2 * (t - 1) + 1] = 0.0;
      }
      for (int c : {0, 3, 5}) sg[2 * c] = sg[2 * c] + 0.0, sg[2 * c + 1] = sg[2 * c + 1] + ie;
    }
    for (int c = 0; c < 6; ++c) mu[c] = 0.0;
    if (in.imu <= 1) {
      const double m = in.imu == 0 ? 1.0 : in.mu[p];
      for (int c : {0, 3, 5}) mu[c] = kMu0 * m;
    } else {
      for (int i = 0; i < in.imu; ++i) {
        const int a = in.ijmu[i], b = in.ijmu[in.imu + i];
        const int t = a == 1 ? b : (a == 2 || a == 3) ? a + b : 0;
        if (t) mu[t - 1] = kMu0 * in.mu[p * in.imu + i];
      }
    }
  }

  // innermodel_gqg (geometry.f90:772-966) with min_dd_inner (:971-1027)
  bool innermodel() {
    out.sigma.assign(12 * (size_t)npt, 0.0);  // cmplx(-1.d0, 0.d0)
    for (size_t c = 0; c < out.sigma.size(); c += 2) out.sigma[c] = -1.0;
    out.mu.assign(6 * (size_t)npt, -1.0);
    std::vector<uint8_t> valued(npt, 0);
    const int s = nord - 1;
    const int ktopz = gnz - nzl[nsf - 1] - nzl[nsf - 2] - 1;  // last inner element layer
    // the nodes of the inner elements: a box of the lattice
    const int64_t ia = (int64_t)nextd * s + 1, ib = (int64_t)(gnx - nextd - 1) * s + 1;
    const int64_t ja = (int64_t)nextd * s + 1, jb = (int64_t)(gny - nextd - 1) * s + 1;
    const int64_t ka = (int64_t)(nextd - 1) * s + 1, kb = (int64_t)ktopz * s + 1;
    const bool any = gnx - nextd - 1 >= nextd + 1 && gny - nextd - 1 >= nextd + 1 && ktopz >= nextd;
    if (any) {
      const int64_t cx = ib - ia + 1, cy = jb - ja + 1, cz = kb - ka + 1, cnt = cx * cy * cz;
      std::vector<int64_t> pick(cnt, -1);
      const int64_t nm = (int64_t)in.mx * in.my * in.mz;
      const double *xmv = xm.data(), *ymv = ym.data(), *zmv = zm.data();
      const double *xp = out.xp.data(), *yp = out.yp.data(), *zp = out.zp.data();
      auto work = [&](int64_t t0, int64_t t1) {
        for (int64_t t = t0; t < t1; ++t) {
          const int64_t ii = ia + t / (cy * cz), jj = ja + (t / cz) % cy, kk = ka + t % cz;
          const int64_t id = (ii - 1) * nyz + (jj - 1) * nnz + kk;
          const double px = xp[ii - 1], py = yp[jj - 1], pz = zp[id - 1];
          double ddmin = 1.e20;
          int64_t best = -1;
          for (int64_t q = 0; q < nm; ++q) {
            const int64_t ij = q / in.mz;
            const double ex = px - xmv[ij / in.my], ey = py - ymv[ij % in.my], ez = pz - zmv[q];
            const double dd = std::sqrt(ex * ex + ey * ey + ez * ez);
            if (dd <= 1.e-5) {  // first coincident model point wins
              best = q;
              break;
            }
            if (dd < ddmin) ddmin = dd, best = q;
          }
          pick[t] = best;
        }
      };
      unsigned nth = std::thread::hardware_concurrency();
      nth = std::max(1u, std::min(nth, 16u));
      if (cnt * nm < (int64_t)4000000) nth = 1;
      std::vector<std::thread> th;
      const int64_t chunk = (cnt + nth - 1) / nth;
      for (unsigned w = 0; w < nth; ++w) {
        const int64_t t0 = w * chunk, t1 = std::min(cnt, t0 + chunk);
        if (t0 < t1) th.emplace_back(work, t0, t1);
      }
      for (auto &t : th) t.join();
      if (std::find(pick.begin(), pick.end(), (int64_t)-1) != pick.end())
        return fail("a lattice node lies 1e20 or more from every model point");
      for (int64_t t = 0; t < cnt; ++t) {
        const int64_t ii = ia + t / (cy * cz), jj = ja + (t / cz) % cy, kk = ka + t % cz;
        const int64_t id = (ii - 1) * nyz + (jj - 1) * nnz + kk - 1;
        assign_model(id, pick[t]);
        valued[id] = 1;
      }
    }
    auto copy_if = [&](int64_t id, int64_t ip) {  // 1-based
      if (valued[id - 1]) return;
      std::copy(&out.sigma[12 * (ip - 1)], &out.sigma[12 * ip], &out.sigma[12 * (id - 1)]);
      std::copy(&out.mu[6 * (ip - 1)], &out.mu[6 * ip], &out.mu[6 * (id - 1)]);
      valued[id - 1] = 1;
    };
    const int64_t ex = (int64_t)nextd * s;
    const int64_t kt = nnz - (int64_t)nzl[nsf - 1] * s - (int64_t)nzl[nsf - 2] * s;
    // 1. (y,z) planes
    for (int64_t im = 1; im <= ex; ++im)
      for (int64_t jm = ex + 1; jm <= nny - ex; ++jm)
        for (int64_t km = ex; km <= kt; ++km) copy_if((im - 1) * nyz + (jm - 1) * nnz + km, ex * nyz + (jm - 1) * nnz + km);
    for (int64_t im = nnx - ex + 1; im <= nnx; ++im)
      for (int64_t jm = ex + 1; jm <= nny - ex; ++jm)
        for (int64_t km = ex; km <= kt; ++km)
          copy_if((im - 1) * nyz + (jm - 1) * nnz + km, (nnx - ex - 1) * nyz + (jm - 1) * nnz + km);
    // 2. (x,z) planes
    for (int64_t im = 1; im <= nnx; ++im)
      for (int64_t jm = 1; jm <= ex; ++jm)
        for (int64_t km = ex; km <= kt; ++km)
          copy_if((im - 1) * nyz + (jm - 1) * nnz + km, (im - 1) * nyz + ex * nnz + km);
    for (int64_t im = 1; im <= nnx; ++im)
      for (int64_t jm = nny - ex + 1; jm <= nny; ++jm)
        for (int64_t km = ex; km <= kt; ++km)
          copy_if((im - 1) * nyz + (jm - 1) * nnz + km, (im - 1) * nyz + (nny - ex - 1) * nnz + km);
    // 3. (x,y) planes below the inner model
    for (int64_t im = 1; im <= nnx; ++im)
      for (int64_t jm = 1; jm <= nny; ++jm)
        for (int64_t km = 1; km <= ex - 1; ++km)
          copy_if((im - 1) * nyz + (jm - 1) * nnz + km, (im - 1) * nyz + (jm - 1) * nnz + ex);
    // 4. air
    const double ie = (double)(float)(kEps0 * omega);
    for (int64_t im = 1; im <= nnx; ++im)
      for (int64_t jm = 1; jm <= nny; ++jm)
        for (int64_t km = kt + 1; km <= nnz; ++km) {
          const int64_t id = (im - 1) * nyz + (jm - 1) * nnz + km - 1;
          double *sg = &out.sigma[12 * id], *mu = &out.mu[6 * id];
          for (int c = 0; c < 12; ++c) sg[c] = 0.0;
          for (int c = 0; c < 6; ++c) mu[c] = 0.0;
          for (int c : {0, 3, 5}) sg[2 * c + 1] = ie, mu[c] = kMu0;
        }
    // 5. components left negative take the value of the node below
    for (int64_t id = 0; id < npt; ++id)
      for (int c = 0; c < 6; ++c) {
        const bool ns = out.sigma[12 * id + 2 * c] < 0, nm = out.mu[6 * id + c] < 0;
        if ((ns || nm) && id == 0) return fail("the first lattice node was left without a model value");
        if (ns) {
          out.sigma[12 * id + 2 * c] = out.sigma[12 * (id - 1) + 2 * c];
          out.sigma[12 * id + 2 * c + 1] = out.sigma[12 * (id - 1) + 2 * c + 1];
        }
        if (nm) out.mu[6 * id + c] = out.mu[6 * (id - 1) + c];
      }
    return true;
  }

  bool validate() {
    const mvf_mesh_input &q = in;
    if (q.mn != 8 && q.mn != 20 && q.mn != 27) return fail("mn must be 8, 20 or 27");
    if (!(q.dx > 0 && q.dy > 0 && q.dz > 0)) return fail("dx, dy, dz must be positive");
    if (!(q.sigma_bg > 0)) return fail("sigma_bg must be positive");
    if (q.nsf < 4) return fail("nsf must be at least 4 (one interface plus the three added slots)");
    if (q.nsr < 0 || (q.nsr > 0 && (!q.xsr || !q.ysr || !q.zsr))) return fail("bad source/receiver arrays");
    if (q.mx < 1 || q.my < 1 || q.mz < 1 || !q.xm || !q.ym || !q.zm) return fail("bad inner-model lattice");
    if (q.isigma < 1 || !q.ijsigma || !q.sigma) return fail("bad conductivity arrays");
    if (q.imu < 0 || (q.imu > 0 && (!q.ijmu || !q.mu))) return fail("bad permeability arrays");
    if (q.nf < 1 || !q.freqs) return fail("at least one frequency is needed");
    if (!q.nsp || !q.xto || !q.yto || !q.zto) return fail("bad interface arrays");
    for (int i = 0; i < q.nf; ++i)
      if (!(q.freqs[i] > 0)) return fail("frequencies must be positive");
    for (int l = 2; l <= q.nsf - 2; ++l)
      if (q.nsp[l - 1] < 1) return fail("interface " + std::to_string(l) + " has no points");
    for (int i = 0; i < q.isigma; ++i) {
      const int a = q.ijsigma[i], b = q.ijsigma[q.isigma + i];
      if (q.isigma > 1 && (a == 1 || a == 2 || a == 3) && (a == 1 ? b : a + b) > 6)
        return fail("conductivity component index out of range");
      if (q.isigma > 1 && (a == 1 ? b : a + b) < 1 && (a >= 1 && a <= 3))
        return fail("conductivity component index out of range");
    }
    for (int i = 0; i < q.imu; ++i) {
      const int a = q.ijmu[i], b = q.ijmu[q.imu + i];
      if (q.imu > 1 && (a >= 1 && a <= 3) && ((a == 1 ? b : a + b) > 6 || (a == 1 ? b : a + b) < 1))
        return fail("permeability component index out of range");
    }
    return true;
  }

  bool run() {
    if (!validate()) return false;
    nsf = in.nsf;
    nord = in.mn == 8 ? 2 : 3;
    omega = 2.0 * kPi * in.freqs[0];
    xsr.assign(in.xsr, in.xsr + in.nsr);
    ysr.assign(in.ysr, in.ysr + in.nsr);
    zsr.assign(in.zsr, in.zsr + in.nsr);
    xm.assign(in.xm, in.xm + in.mx);
    ym.assign(in.ym, in.ym + in.my);
    zm.assign(in.zm, in.zm + (size_t)in.mx * in.my * in.mz);
    slot.assign(nsf, Slot());
    for (int l = 2; l <= nsf - 2; ++l) {
      Slot &s = slot[l - 1];
      const int n = in.nsp[l - 1];
      for (int j = 0; j < n; ++j) {
        const size_t at = (size_t)j * nsf + (l - 1);
        s.x.push_back(in.xto[at]), s.y.push_back(in.yto[at]), s.z.push_back(in.zto[at]);
      }
    }
    min_max_input();
    if (!extension()) return false;
    coord_transform();
    if (!zblocks() || !gauss_grid() || !innermodel()) return false;
    int64_t *d = out.dims;
    d[0] = gnx - 1, d[1] = gny - 1, d[2] = gnz - 1, d[3] = in.mn, d[4] = nnx, d[5] = nny, d[6] = nnz, d[7] = npt;
    d[8] = nextd, d[9] = nsf, d[10] = in.nsr, d[11] = nord;
    double *v = out.vals;
    v[0] = gztop, v[1] = x0, v[2] = y0, v[3] = z0, v[4] = omega, v[5] = skin;
    out.nzl.assign(nzl.begin(), nzl.end());
    out.xsr = xsr, out.ysr = ysr, out.zsr = zsr;
    return true;
  }
};

}  // namespace
}  // namespace mvf

#pragma float_control(pop)

extern "C" {

int mvf_grid3d(const mvf_mesh_input *in, mvf_mesh **out) {
  if (!in || !out) return mvf::host_fail("mvf_grid3d: null argument");
  *out = nullptr;
  mvf_mesh *m = new (std::nothrow) mvf_mesh();
  if (!m) return mvf::host_fail("mvf_grid3d: out of host memory");
  try {
    mvf::Grid3d g(*in, *m);
    if (!g.run()) {
      delete m;
      return mvf::host_fail("mvf_grid3d: " + g.err);
    }
  } catch (const std::bad_alloc &) {
    delete m;
    return mvf::host_fail("mvf_grid3d: out of host memory");
  }
  *out = m;
  return MVF_OK;
}

int mvf_mesh_sizes(const mvf_mesh *m, int64_t *dims, double *vals) {
  if (!m) return mvf::host_fail("mvf_mesh_sizes: null mesh");
  if (dims) std::copy(m->dims, m->dims + 12, dims);
  if (vals) std::copy(m->vals, m->vals + 6, vals);
  return MVF_OK;
}

int mvf_mesh_arrays(const mvf_mesh *m, double *xp, double *yp, double *zp, double *sigma, double *mu, int32_t *nzl,
                    double *xsr, double *ysr, double *zsr) {
  if (!m) return mvf::host_fail("mvf_mesh_arrays: null mesh");
  auto put = [](const auto &v, auto *dst) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  put(m->xp, xp), put(m->yp, yp), put(m->zp, zp), put(m->sigma, sigma), put(m->mu, mu), put(m->nzl, nzl);
  put(m->xsr, xsr), put(m->ysr, ysr), put(m->zsr, zsr);
  return MVF_OK;
}

void mvf_mesh_free(mvf_mesh *m) { delete m; }

int mvf_read_inputs(const char *inp_path, const char *param_path, mvf_inputs **out) {
  if (!out || (!inp_path && !param_path)) return mvf::host_fail("mvf_read_inputs: nothing to read");
  *out = nullptr;
  mvf_inputs *h = new (std::nothrow) mvf_inputs();
  if (!h) return mvf::host_fail("mvf_read_inputs: out of host memory");
  try {
    if (inp_path) {
      mvf::LdFile f;
      if (!f.load(inp_path) || !mvf::read_3dem(f, *h)) {
        delete h;
        return mvf::host_fail("mvf_read_inputs: " + f.err);
      }
    }
    if (param_path) {
      mvf::LdFile f;
      if (!f.load(param_path) || !mvf::read_param(f, *h)) {
        delete h;
        return mvf::host_fail("mvf_read_inputs: " + f.err);
      }
    }
  } catch (const std::bad_alloc &) {
    delete h;
    return mvf::host_fail("mvf_read_inputs: out of host memory");
  }
  *out = h;
  return MVF_OK;
}

int mvf_inputs_view(const mvf_inputs *h, mvf_mesh_input *mesh, mvf_params *params) {
  if (!h) return mvf::host_fail("mvf_inputs_view: null handle");
  if (mesh) {
    *mesh = mvf_mesh_input{};
    if (h->has_inp) {
      mesh->dx = h->dx, mesh->dy = h->dy, mesh->dz = h->dz, mesh->high = h->high, mesh->depth = h->depth;
      mesh->sigma_bg = h->bcksigma;
      mesh->mn = h->mn, mesh->nsr = h->nsr, mesh->nsf = h->nsf;
      mesh->mx = h->mx, mesh->my = h->my, mesh->mz = h->mz, mesh->isigma = h->isigma, mesh->imu = h->imu;
      mesh->xsr = h->xsr.data(), mesh->ysr = h->ysr.data(), mesh->zsr = h->zsr.data();
      mesh->nsp = h->nsp.data(), mesh->xto = h->xto.data(), mesh->yto = h->yto.data(), mesh->zto = h->zto.data();
      mesh->xm = h->xm.data(), mesh->ym = h->ym.data(), mesh->zm = h->zm.data();
      mesh->ijsigma = h->ijsigma.data(), mesh->sigma = h->sigma.data();
      mesh->ijmu = h->imu ? h->ijmu.data() : nullptr, mesh->mu = h->imu ? h->mu.data() : nullptr;
    }
    if (h->has_param) mesh->nf = h->nf, mesh->freqs = h->freqs.data();
  }
  if (params) {
    *params = mvf_params{};
    if (h->has_param) {
      params->nf = h->nf, params->boundary = h->boundary, params->inimod = h->inimod, params->nl = h->nl;
      params->gpml_sch = h->gpml_sch;
      params->h_sigma = h->h_sigma, params->a0 = h->a0, params->b0 = h->b0, params->nn = h->nn;
      params->freqs = h->freqs.data(), params->l_sigma = h->l_sigma.data(), params->l_dz = h->l_dz.data();
    }
  }
  return MVF_OK;
}

void mvf_inputs_free(mvf_inputs *h) { delete h; }

}  // extern "C"
