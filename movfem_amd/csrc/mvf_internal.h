// Internal state of an mvf_handle and shared helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/movfem_gpu.h"
#include "mvf_tables.h"

#define MVF_PI 3.1415926535897932384626433
#define MVF_EPS0 8.854187817e-12
#define MVF_B0 1.0e-9

namespace mvf {

// message of the last failing handle-less call (file readers, mesh
// generation) on this thread; mvf_last_error(NULL) returns it
inline thread_local std::string g_host_err;

struct Gpml {
  // zone limits in element indices (1-based) and coordinates (init_gpml,
  // boundary_conds.f90:51-70); side 0 = "-", side 1 = "+"
  int el_a[3][2], el_b[3][2];
  double a[3][2], b[3][2];
  double a0, nn;
  int sch;
  int on;  // 0 for Dirichlet runs
};

// GPML flags of one element (get_pml, boundary_conds.f90:72-82), packed as
// three 2-bit fields: 0 none, 1 "+", 2 "-"
__host__ __device__ inline int zone_flags(const Gpml &g, int i, int j, int k) {
  int idx[3] = {i, j, k}, f = 0;
  for (int c = 0; c < 3; c++) {
    int v = 0;
    if (idx[c] >= g.el_a[c][0] && idx[c] <= g.el_b[c][0]) v = 2;
    if (idx[c] >= g.el_a[c][1] && idx[c] <= g.el_b[c][1]) v = 1;
    f |= v << (2 * c);
  }
  return f;
}

// Re of the GPML stretch along axis c at coordinate r (gpml_h,
// boundary_conds.f90:84-186; the reference stores only the real part, so the
// Fang scheme keeps 1 + a0 (r/R)^n and the Zhou scheme reduces to 1)
__host__ __device__ inline double gpml_stretch(const Gpml &g, int flags, int c, double r) {
  int v = (flags >> (2 * c)) & 3;
  if (v == 0 || g.sch != 0) return 1.0;
  int s = (v == 2) ? 0 : 1;
  double rp = sqrt((g.b[c][s] - g.a[c][s]) * (g.b[c][s] - g.a[c][s]));
  double rr = sqrt((r - g.a[c][s]) * (r - g.a[c][s]));
  return 1.0 + g.a0 * pow(rr / rp, g.nn);
}

struct Geom {
  int nx, ny, nz, nnx, nny, nnz, nord;
  int64_t nyz, npt, ne;
};

// per-frequency scalars of the Dirichlet boundary 1-D model (inimod 2/3)
struct BdParams {
  int inimod;
  double ztop;
  double cm, cp;           // float32(-2 omega b0), float32(2 omega b0)
  double2 sa, qa, se, qe;  // inimod 2: gamma and 1/gamma of the air and the half-space
  double2 cz1, e0;         // inimod 3: surface transfer function, ezl(0, 0)
};

// Reference-element tables in the lane order of the matrix-core element
// kernel: for Gauss-point group u, DOF tile t and lane l (point 4u + l/16, DOF
// 16t + l%16) the DOF's derivative (d0, d1, d2) and value (p) of the edge
// polynomial, [u][t][l], zero for padding; the DOF's direction [t][l%16]; and
// the nodal functions with the point fastest, dN [3][node][32], N [node][32],
// weights w [32].
template <typename T>
struct LaneTab {
  const T *d0, *d1, *d2, *p;
  const int *dir;
  const T *dN, *N, *w;
};

// Workgroups are dealt round-robin to the 8 XCDs.  xcd_block renumbers block
// b of nb so that each XCD takes a contiguous range of logical blocks (a
// bijection of [0, nb)): neighbouring work then meets in that XCD's L2.
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t x = b & 7, i = b >> 3, q = nb >> 3, rm = nb & 7;
  return x * q + (x < rm ? x : rm) + i;
}

}  // namespace mvf

struct mvf_handle {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  int nodes = 8;
  mvf::ElemType et{};
  mvf::Geom geo{};
  mvf::Gpml gp{};
  int emulate = MVF_EMU_ALL;
  int verbose = 0;
  mvf_bc bc{};
  // lattice and materials (device)
  double *d_xp = nullptr, *d_yp = nullptr, *d_zp = nullptr;
  double *d_sig = nullptr;   // complex (npt, 6) interleaved, mutable by update_sigma
  double *d_inmu = nullptr;  // (npt, 6) inverse permeability tensor
  double *d_dmu = nullptr;   // (npt, 6) mu - mu0 on the diagonal
  mvf::BasisTab *d_tab = nullptr;
  // numbering
  int32_t *d_gne = nullptr;
  int32_t nne = 0;
  // full CSR pattern
  int64_t nnz = 0, nnz_triu = 0;
  int64_t *d_rowptr = nullptr;
  int32_t *d_col = nullptr;
  int64_t *d_diagpos = nullptr;
  // element -> entry gather lists (entry n sums contributions seg[n]..seg[n+1])
  int64_t npairs = 0;
  int64_t *d_seg = nullptr;
  int64_t *d_contrib = nullptr;  // e*me*me + im*me + jm
  // DOF -> (element, local DOF) gather lists for the RHS
  int64_t *d_rseg = nullptr;
  int64_t *d_rcontrib = nullptr;  // e*me + im
  // frequency-independent operators
  double *d_K = nullptr, *d_M = nullptr;
  float *d_K32 = nullptr, *d_M32 = nullptr;
  int ops_fp32 = 0;  // d_K/d_M were last assembled from FP32 element integration
  int elem_legacy = 0;  // element matrices from the one-workgroup-per-element VALU kernel (comparison)
  mvf::LaneTab<double> lt64{};
  int autotune = 1;                  // SpMV lanes per row timed on first use (mvf_solver.hip)
  int setup_rows = 0;                // pattern and assembly built row by row (large models, mvf_setup.hip)
  int spmv_lpr[4] = {0, 0, 0, 0};    // tuned lanes per row for 2, 4, 8, 16 columns
  mvf::LaneTab<float> lt32{};
  double *d_Ke = nullptr, *d_Me = nullptr;  // element matrices (upper packed), kept for mvf_integrate
  int64_t ke_elems = 0;
  // element-0 correction for the carried GPML flags (see MVF_EMU_GPML_PREV_FLAGS)
  double *d_dK0 = nullptr, *d_dM0 = nullptr;  // packed upper, me*(me+1)/2
  int carry_nonzero = 0;
  int carry_applied = 0;  // element 0 currently holds the carried-flag contribution
  int flags_first = 0, flags_late = 0;
  int64_t calls = 0;  // global_vfem-equivalent calls performed (frequencies processed)
  // pinned host staging of mvf_global_vfem's chunks (two buffers of gv_pin_cap entries, 24 B each)
  void *gv_pin[2] = {nullptr, nullptr};
  int64_t gv_pin_cap = 0;
  // current frequency
  double freq = 0.0, omega = 0.0;
  int have_freq = 0;
  // last batched solve
  int ncols_last = 0;
  double *d_x = nullptr, *d_b = nullptr;
  size_t xb_cap = 0;
  double t_setup[4] = {0, 0, 0, 0};
  // sparse direct factorization state (mvf_direct.hip)
  void *direct = nullptr;
  // Dirichlet inimod 2/3 values (mvf_boundary.hip): elements on the x and y
  // faces, their K_e/M_e, f_boundary per (element, edge, polarization)
  int bd_on = 0;
  double bd_ztop = 0.0, bd_hsigma = 0.0;
  std::vector<double> bd_lsig, bd_ldz;
  int64_t nbel = 0;
  int64_t *d_bel = nullptr;
  int32_t *d_bidx = nullptr;  // element -> index in d_bel, or -1
  double *d_bK = nullptr, *d_bM = nullptr;
  double2 *d_dirf = nullptr;
  double bd_omega = -1.0;     // frequency d_dirf holds
  mvf::BasisTab *d_btab = nullptr;  // basis at the element's basis nodes
  // readout: basis at the element's lattice nodes; frequencies of the last batch
  mvf::BasisTab *d_ntab = nullptr;
  std::vector<double> last_freqs;
  // scratch
  std::vector<void *> owned;
};

namespace mvf {

inline int set_err(mvf_handle *h, int code, const char *fmt, ...) {
  if (h) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    h->err = buf;
  }
  return code;
}

#define MVF_HIP(h, call)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return mvf::set_err((h), (e_ == hipErrorOutOfMemory) ? MVF_E_MEMORY : MVF_E_DEVICE,       \
                          "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define MVF_TRY(expr)      \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)

template <typename T>
inline int dalloc(mvf_handle *h, T **p, size_t n) {
  void *q = nullptr;
  size_t bytes = (n ? n : 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess)
    return set_err(h, MVF_E_MEMORY, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  *p = (T *)q;
  return 0;
}

inline void dfree(void *p) {
  if (p) (void)hipFree(p);
}

// Temporary device buffers of one call: whatever is allocated through it is
// freed when the call returns, on every return path (errors included).
struct DevScope {
  std::vector<void *> bufs;
  DevScope() = default;
  DevScope(const DevScope &) = delete;
  DevScope &operator=(const DevScope &) = delete;
  ~DevScope() {
    for (void *p : bufs)
      if (p) (void)hipFree(p);
  }
  template <typename T>
  int alloc(mvf_handle *h, T **p, size_t n) {
    int rc = dalloc(h, p, n);
    if (!rc) bufs.push_back((void *)*p);
    return rc;
  }
  // free one buffer before the call ends (lowers the peak)
  void drop(void *p) {
    if (!p) return;
    for (auto &q : bufs)
      if (q == p) q = nullptr;
    (void)hipFree(p);
  }
};

inline int grid1d(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// internal entry points shared by the translation units
int setup_numbering(mvf_handle *h);
int setup_pattern(mvf_handle *h);
int setup_operators(mvf_handle *h);
int integrate_elements(mvf_handle *h, int64_t e0, int64_t e1, int first_flags, double *dK, double *dM);
int integrate_elements32(mvf_handle *h, int64_t e0, int64_t e1, int first_flags, float *dK, float *dM);
int assemble_operators(mvf_handle *h);
int assemble_operators32(mvf_handle *h, const float *Ke, const float *Me);
int element_rhs_range(mvf_handle *h, double omega, int first_flags, int64_t e0, int64_t e1, double *d_be);
int assemble_rhs(mvf_handle *h, const double *d_be, double *d_b, int ncol, int col0);
int update_sigma(mvf_handle *h, double omega);
void direct_free(mvf_handle *h);
int node_table(mvf_handle *h);
int node_fields(mvf_handle *h, const double *x, int ncol, int col0, double omega, double *es, double *hs);
int problem_fields(mvf_handle *h, double omega, double *ep, double *hp, double *src);
int receivers(mvf_handle *h, const double *es, const double *hs, int ke_top, int nrx, const int32_t *rie,
              const int32_t *rje, const double *rx, const double *ry, double omega, double *oe, double *oh, double *oz,
              double *rho, double *phi, double *elev);
int node_impedance(mvf_handle *h, const double *es, const double *hs, double omega, double *oz, double *rho,
                   double *phi);
void direct_invalidate(mvf_handle *h);
int bd_setup(mvf_handle *h, const double *dKe, const double *dMe);
int bd_values(mvf_handle *h, double omega);
int bd_lift(mvf_handle *h, double omega, double *d_be);
int noalloc_setup(mvf_handle *h);
int lane_tables(mvf_handle *h, const BasisTab *tab);

}  // namespace mvf
