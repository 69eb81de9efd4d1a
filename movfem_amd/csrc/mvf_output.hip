// Text writers for the reference's output files (host only): SOLUTION.OUT
// (write_solution, solution.f90:73-205), RECEIVER.OUT (write_receiver_data,
// receiver_data.f90:25-67), SURFACE.OUT (write_surface_data, :69-109) and
// PROBLEM.OUT (write_problem_ugrid, problem.f90:188-240).
//
// Column formats are the reference's edit descriptors.  Its list-directed
// records (write(u,*)) take the reference build's list-directed form: a
// leading blank, one blank before every item, a new record before the line
// would pass 79 columns, integers in minimal width and reals as the shortest
// decimal that reads back to the same double (the middle one when several
// do), printed positionally (".1", "1000.") when the value rounded to one
// significant digit lies in [0.1, 1e15) and as d.dddE+xx otherwise.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mvf_internal.h"

namespace mvf {
namespace {

std::string rjust(const std::string &s, int w) {
  if ((int)s.size() >= w) return s;
  return std::string(w - s.size(), ' ') + s;
}

// ESw.d: one digit before the point, d after, exponent E+dd (+ddd without
// the letter past two digits); w asterisks when the field does not fit
std::string fmt_es(double v, int w, int d) {
  if (std::isnan(v)) return rjust("NaN", w);
  if (std::isinf(v)) return rjust(v < 0 ? "-Inf" : "Inf", w);
  char buf[64];
  snprintf(buf, sizeof(buf), "%.*E", d, v);
  char *e = strchr(buf, 'E');
  const int ex = atoi(e + 1);
  std::string r(buf, e - buf);
  char eb[16];
  if (std::abs(ex) <= 99) snprintf(eb, sizeof(eb), "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
  else snprintf(eb, sizeof(eb), "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
  r += eb;
  if ((int)r.size() > w) return std::string(w, '*');
  return rjust(r, w);
}

// Fw.d
std::string fmt_f(double v, int w, int d) {
  if (std::isnan(v)) return rjust("NaN", w);
  if (std::isinf(v)) return rjust(v < 0 ? "-Inf" : "Inf", w);
  char buf[512];
  snprintf(buf, sizeof(buf), "%.*f", d, v);
  if ((int)strlen(buf) > w) return std::string(w, '*');
  return rjust(buf, w);
}

// k * 10^ss == m (a long double) exactly?  k * 5^ss * 2^ss against
// M * 2^E with m = M * 2^E, M a 64-bit integer, in base-2^32 integers.
bool decimal_equals(unsigned long long k, int ss, long double m) {
  int em;
  const long double fr = frexpl(m, &em);
  const unsigned long long M = (unsigned long long)ldexpl(fr, 64);
  em -= 64;
  typedef std::vector<uint32_t> Big;
  auto from = [](unsigned long long v) { return Big{(uint32_t)v, (uint32_t)(v >> 32)}; };
  auto mul = [](Big &b, uint32_t f) {
    unsigned long long c = 0;
    for (auto &w : b) {
      c += (unsigned long long)w * f;
      w = (uint32_t)c;
      c >>= 32;
    }
    if (c) b.push_back((uint32_t)c);
  };
  auto shl = [&](Big &b, int n) {
    for (; n >= 31; n -= 31) mul(b, 1u << 31);
    if (n > 0) mul(b, 1u << n);
  };
  auto trim = [](Big &b) {
    while (!b.empty() && b.back() == 0) b.pop_back();
  };
  Big l = from(k), r = from(M);
  for (int i = 0; i < std::abs(ss); i++) mul(ss >= 0 ? l : r, 5);
  const int sh = ss - em;
  shl(sh >= 0 ? l : r, std::abs(sh));
  trim(l);
  trim(r);
  return l == r;
}

// list-directed REAL(8)
std::string ld_real(double x) {
  if (x == 0.0) return std::signbit(x) ? "-0." : "0.";
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x < 0 ? "-Inf" : "Inf";
  const double a = std::fabs(x);
  // Shortest digit strings k * 10^sc (p digits) strictly inside the interval
  // of reals that round to a.  The reference build excludes the interval's
  // ends (where a decimal reaches a only by the ties-to-even rule) and, when
  // several p-digit strings qualify, prints the middle one,
  // floor((kmin + kmax) / 2), not the nearest: measured on its list-directed
  // output, every case agreeing (tests/golden/ld_reals.npz).  The ends are
  // exact in 80-bit long double; a decimal that reads back as one is compared
  // with it exactly.
  const long double lo_end = ((long double)a + (long double)std::nextafter(a, 0.0)) / 2;
  const long double hi_end = ((long double)a + (long double)std::nextafter(a, (double)INFINITY)) / 2;
  char buf[64];
  long long k = 0;
  int p = 1, sc = 0;
  auto reads_back = [&](long long kk, int ss) {
    char b[48];
    snprintf(b, sizeof(b), "%llde%d", kk, ss);
    if (strtod(b, nullptr) != a) return false;
    const long double c = strtold(b, nullptr);
    if (c == lo_end) return !decimal_equals((unsigned long long)kk, ss, lo_end);
    if (c == hi_end) return !decimal_equals((unsigned long long)kk, ss, hi_end);
    return true;
  };
  for (;; p++) {
    snprintf(buf, sizeof(buf), "%.*e", p - 1, a);
    k = 0;
    for (const char *c = buf; *c && *c != 'e'; c++)
      if (*c >= '0' && *c <= '9') k = 10 * k + (*c - '0');
    sc = atoi(strchr(buf, 'e') + 1) - (p - 1);
    if (p >= 17 || reads_back(k, sc)) break;
    if (reads_back(k + 1, sc)) {  // the nearest lies outside a one-sided interval
      k = k + 1;
      break;
    }
    if (k > 1 && reads_back(k - 1, sc)) {
      k = k - 1;
      break;
    }
  }
  long long lo = k, hi = k;
  while (lo > 1 && reads_back(lo - 1, sc)) lo--;
  while (reads_back(hi + 1, sc)) hi++;
  std::string D = std::to_string((lo + hi) / 2);
  const int E = sc + (int)D.size();  // a = 0.D x 10^E
  while (D.size() > 1 && D.back() == '0') D.pop_back();
  // positional form when the value rounded to one digit lies in [0.1, 1e15)
  char r1[32];
  snprintf(r1, sizeof(r1), "%.0e", a);
  const int E1 = atoi(strchr(r1, 'e') + 1) + 1;
  std::string s = x < 0 ? "-" : "";
  const int n = (int)D.size();
  if (E1 >= 0 && E1 <= 15) {
    if (E <= 0) {
      s += "." + std::string(-E, '0') + D;
    } else if (n <= E) {
      s += D + std::string(E - n, '0') + ".";
    } else {
      s += D.substr(0, E) + "." + D.substr(E);
    }
  } else {
    char eb[16];
    const int ex = E - 1;
    snprintf(eb, sizeof(eb), "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    s += D.substr(0, 1) + "." + D.substr(1) + eb;
  }
  return s;
}

// one list-directed write statement
struct ListOut {
  FILE *f;
  std::string line;
  void item(const std::string &s) {
    const std::string t = " " + s;
    if (!line.empty() && line.size() + t.size() > 79) {
      fprintf(f, "%s\n", line.c_str());
      line.clear();
    }
    line += t;
  }
  void real(double v) { item(ld_real(v)); }
  // complex: " (re," and "im)" are placed separately, so a record may end
  // after the comma and the next one then starts with the imaginary part
  void cplx(double re, double im) {
    item("(" + ld_real(re) + ",");
    const std::string t = ld_real(im) + ")";
    if (line.size() + t.size() > 79) {
      fprintf(f, "%s\n", line.c_str());
      line.clear();
    }
    line += t;
  }
  void integer(long long v) { item(std::to_string(v)); }
  void end() {
    fprintf(f, "%s\n", line.c_str());
    line.clear();
  }
};

const char *kFieldLabels[24] = {"Re[Ex_(1)]", "Im[Ex_(1)]", "Re[Ey_(1)]", "Im[Ey_(1)]", "Re[Ez_(1)]", "Im[Ez_(1)]",
                                "Re[Bx_(1)]", "Im[Bx_(1)]", "Re[By_(1)]", "Im[By_(1)]", "Re[Bz_(1)]", "Im[Bz_(1)]",
                                "Re[Ex_(2)]", "Im[Ex_(2)]", "Re[Ey_(2)]", "Im[Ey_(2)]", "Re[Ez_(2)]", "Im[Ez_(2)]",
                                "Re[Bx_(2)]", "Im[Bx_(2)]", "Re[By_(2)]", "Im[By_(2)]", "Re[Bz_(2)]", "Im[Bz_(2)]"};
const char *kZLabels[8] = {"Re(Zxx)", "Im(Zxx)", "Re(Zxy)", "Im(Zxy)", "Re(Zyx)", "Im(Zyx)", "Re(Zyy)", "Im(Zyy)"};
const char *kRhoLabels[8] = {"App_RHO_xx", "App_RHO_xy", "App_RHO_yx", "App_RHO_yy",
                             "App_PHI_xx", "App_PHI_xy", "App_PHI_yx", "App_PHI_yy"};

// the 12 field values of node id and polarization d: E then H, (Re, Im) per component
void write_node_row(FILE *f, const double *esol, const double *hsol, int64_t npt, int64_t id) {
  std::string s;
  for (int d = 0; d < 2; d++) {
    const double *e = esol + 2 * ((d * npt + id) * 3), *h = hsol + 2 * ((d * npt + id) * 3);
    for (int c = 0; c < 6; c++) s += fmt_es(e[c], 20, 5);
    for (int c = 0; c < 6; c++) s += fmt_es(h[c], 20, 5);
  }
  fprintf(f, "%s\n", s.c_str());
}

void labels(FILE *f, const char *const *lab, int n, int nf, bool lead, bool xy) {
  std::string s;
  if (lead) {
    char b[16];
    snprintf(b, sizeof(b), "%5d ", nf);
    s = b;
    if (xy) s += rjust("x", 15) + rjust("y", 15);
  }
  for (int i = 0; i < n; i++) s += rjust(lab[i], 15);
  fprintf(f, "%s\n", s.c_str());
}

FILE *open_out(const char *path, int append) { return fopen(path, append ? "a" : "w"); }

}  // namespace
}  // namespace mvf

using namespace mvf;

extern "C" {

int mvf_write_solution(const char *path, int32_t nf, int32_t append, int64_t npt, const double *esol,
                       const double *hsol, const double *node_z, const double *node_rho, const double *node_phi,
                       int64_t nsel, const int32_t *sel) {
  if (!path || npt < 1 || !esol || !hsol || !node_z || !node_rho || !node_phi || nsel < 0 || (nsel > 0 && !sel))
    return MVF_E_ARG;
  for (int64_t i = 0; i < nsel; i++)
    if (sel[i] < 1 || sel[i] > npt) return MVF_E_ARG;
  FILE *f = open_out(path, append);
  if (!f) return MVF_E_ARG;
  ListOut lo{f, ""};
  lo.item("vector-fem solution: ");
  lo.integer(nf);
  lo.end();
  const int64_t n = nsel > 0 ? nsel : npt;
  auto node = [&](int64_t i) -> int64_t { return nsel > 0 ? (int64_t)sel[i] - 1 : i; };
  labels(f, kFieldLabels, 24, nf, false, false);
  for (int64_t i = 0; i < n; i++) write_node_row(f, esol, hsol, npt, node(i));
  labels(f, kZLabels, 8, nf, false, false);
  for (int64_t i = 0; i < n; i++) {
    std::string s;
    for (int c = 0; c < 8; c++) s += fmt_es(node_z[8 * node(i) + c], 20, 5);
    fprintf(f, "%s\n", s.c_str());
  }
  labels(f, kRhoLabels, 8, nf, false, false);
  for (int64_t i = 0; i < n; i++) {
    std::string s;
    const int64_t id = node(i);
    for (int c = 0; c < 4; c++) s += nsel > 0 ? fmt_f(node_rho[4 * id + c], 25, 6) : fmt_f(node_rho[4 * id + c], 20, 5);
    for (int c = 0; c < 4; c++) s += nsel > 0 ? fmt_f(node_phi[4 * id + c], 25, 6) : fmt_f(node_phi[4 * id + c], 20, 5);
    fprintf(f, "%s\n", s.c_str());
  }
  const int bad = ferror(f);
  return (fclose(f) != 0 || bad) ? MVF_E_ARG : 0;
}

int mvf_write_receivers(const char *path, int32_t nf, int32_t append, int32_t nfreq, int32_t nrx, const double *rx,
                        const double *ry, const double *elev, const double *rx_e, const double *rx_h,
                        const double *rx_z, const double *rho, const double *phi) {
  if (!path || nrx < 0 || (nrx > 0 && (!rx || !ry || !elev || !rx_e || !rx_h || !rx_z || !rho || !phi)))
    return MVF_E_ARG;
  FILE *f = open_out(path, append);
  if (!f) return MVF_E_ARG;
  ListOut lo{f, ""};
  if (!append) {
    lo.item("----------------rx coords-----------------------");
    lo.end();
    lo.integer(nrx);
    lo.integer(2);
    lo.integer(nfreq);
    lo.end();
    for (int i = 0; i < nrx; i++) {
      lo.real(rx[i]);
      lo.real(ry[i]);
      lo.real(elev[i]);
      lo.end();
    }
  }
  const char *xy_fields[24];
  for (int i = 0; i < 24; i++) xy_fields[i] = kFieldLabels[i];
  labels(f, xy_fields, 24, nf, true, true);
  for (int i = 0; i < nrx; i++) {
    std::string s = fmt_es(rx[i], 20, 5) + fmt_es(ry[i], 20, 5);
    for (int d = 0; d < 2; d++) {
      const double *e = rx_e + 2 * ((i * 2 + d) * 3), *h = rx_h + 2 * ((i * 2 + d) * 3);
      for (int c = 0; c < 6; c++) s += fmt_es(e[c], 20, 5);
      for (int c = 0; c < 6; c++) s += fmt_es(h[c], 20, 5);
    }
    fprintf(f, "%s\n", s.c_str());
  }
  labels(f, kZLabels, 8, nf, true, true);
  for (int i = 0; i < nrx; i++) {
    std::string s = fmt_es(rx[i], 20, 5) + fmt_es(ry[i], 20, 5);
    for (int c = 0; c < 8; c++) s += fmt_es(rx_z[8 * i + c], 20, 5);
    fprintf(f, "%s\n", s.c_str());
  }
  labels(f, kRhoLabels, 8, nf, true, true);
  for (int i = 0; i < nrx; i++) {
    std::string s = fmt_f(rx[i], 20, 5) + fmt_f(ry[i], 20, 5);
    for (int c = 0; c < 4; c++) s += fmt_f(rho[4 * i + c], 20, 5);
    for (int c = 0; c < 4; c++) s += fmt_f(phi[4 * i + c], 20, 5);
    fprintf(f, "%s\n", s.c_str());
  }
  const int bad = ferror(f);
  return (fclose(f) != 0 || bad) ? MVF_E_ARG : 0;
}

int mvf_write_surface(const char *path, int32_t nf, int32_t append, int32_t nfreq, int32_t nxb, int32_t nyb,
                      const double *xs, const double *ys, const double *zs, const int64_t *ids, int64_t npt,
                      const double *esol, const double *hsol, const double *node_z, const double *node_rho,
                      const double *node_phi) {
  const int64_t nsp = (int64_t)nxb * nyb;
  if (!path || nxb < 1 || nyb < 1 || !xs || !ys || !zs || !ids || npt < 1 || !esol || !hsol || !node_z ||
      !node_rho || !node_phi)
    return MVF_E_ARG;
  for (int64_t i = 0; i < nsp; i++)
    if (ids[i] < 0 || ids[i] >= npt) return MVF_E_ARG;
  FILE *f = open_out(path, append);
  if (!f) return MVF_E_ARG;
  ListOut lo{f, ""};
  if (!append) {
    lo.item("-----------------surface coords----------------");
    lo.end();
    lo.integer(nxb);
    lo.integer(nyb);
    lo.integer(nsp);
    lo.integer(2);
    lo.integer(nfreq);
    lo.end();
    for (int i = 0; i < nxb; i++) lo.real(xs[i]);
    lo.end();
    for (int i = 0; i < nyb; i++) lo.real(ys[i]);
    lo.end();
    for (int64_t i = 0; i < nsp; i++) lo.real(zs[i]);
    lo.end();
  }
  labels(f, kFieldLabels, 24, nf, true, false);
  for (int64_t i = 0; i < nsp; i++) write_node_row(f, esol, hsol, npt, ids[i]);
  labels(f, kZLabels, 8, nf, true, false);
  for (int64_t i = 0; i < nsp; i++) {
    std::string s;
    for (int c = 0; c < 8; c++) s += fmt_es(node_z[8 * ids[i] + c], 20, 5);
    fprintf(f, "%s\n", s.c_str());
  }
  labels(f, kRhoLabels, 8, nf, true, false);
  for (int64_t i = 0; i < nsp; i++) {
    std::string s;
    for (int c = 0; c < 4; c++) s += fmt_f(node_rho[4 * ids[i] + c], 20, 5);
    for (int c = 0; c < 4; c++) s += fmt_f(node_phi[4 * ids[i] + c], 20, 5);
    fprintf(f, "%s\n", s.c_str());
  }
  const int bad = ferror(f);
  return (fclose(f) != 0 || bad) ? MVF_E_ARG : 0;
}

// write_problem_ugrid (problem.f90:188-240), ndir = 2: per polarization the
// node count and frequency index, the column line, and per node the
// list-directed record dble(psigma), ep(1:3), hp(1:3), src(1:3).  psigma is
// pe_psigma(1) = (0, omega*eps) at every node (pset_pmodel, :247-252), so its
// real part, 0, is printed.
int mvf_write_problem(const char *path, int32_t nf, int32_t append, int64_t npt, const double *ep, const double *hp,
                      const double *src, int64_t nsel, const int32_t *sel) {
  if (!path || npt <= 0 || !ep || !hp || !src || nsel < 0 || (nsel > 0 && !sel)) return MVF_E_ARG;
  for (int64_t i = 0; i < nsel; i++)
    if (sel[i] < 1 || sel[i] > npt) return MVF_E_ARG;
  FILE *f = open_out(path, append);
  if (!f) return MVF_E_ARG;
  const int64_t nrow = sel ? nsel : npt;
  for (int d = 0; d < 2; d++) {
    ListOut lo{f, ""};
    lo.item("problem main functions in structured grid");
    lo.integer(nrow);
    lo.item("f= ");
    lo.integer(nf);
    lo.end();
    lo.item("--psigma--epx--epy--epz--hpx--hpy--hpz--srcx--srcy--srcz");
    lo.end();
    for (int64_t r = 0; r < nrow; r++) {
      const int64_t id = sel ? sel[r] - 1 : r;
      lo.real(0.0);
      for (const double *a : {ep, hp, src}) {
        const double *v = a + 2 * ((d * npt + id) * 3);
        for (int c = 0; c < 3; c++) lo.cplx(v[2 * c], v[2 * c + 1]);
      }
      lo.end();
    }
  }
  const int bad = ferror(f);
  return (fclose(f) != 0 || bad) ? MVF_E_ARG : 0;
}

}  // extern "C"
