// C ABI of the MI355X element-integration, assembly and solve path
// (declared in include/movfem_gpu.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <new>
#include <thread>

#include "mvf_internal.h"

namespace mvf {
int solve_device(mvf_handle *h, const double *omegas, int ncol, int ncol_pad, const mvf_solver_opts *o,
                 const double *b, double *x, mvf_solve_info *info);
int solve_device_guess(mvf_handle *h, const double *omegas, const mvf_solver_opts *o, const double *b, double *x,
                       mvf_solve_info *info, double *alpha_out);

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the flags element 0 uses at assembly call index c (reference: in_pml is 0
// before the first get_pml, and holds the last element's flags afterwards)
static int first_flags_for(const mvf_handle *h, int64_t call) {
  if (!h->gp.on) return 0;
  return call >= 1 ? h->flags_late : h->flags_first;
}

// element 0's contribution switches from the first-call flags to the carried
// ones once the first assembly is done: K += dK0, M += dM0 on its pairs
__global__ void k_apply_carry(const int32_t *gne0, int me, const int64_t *rowptr, const int32_t *col,
                              const double *dK, const double *dM, double *K, double *M, float *K32, float *M32,
                              double sign) {
  int np = me * (me + 1) / 2;
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    int i = 0, rem = p;
    while (rem >= me - i) {
      rem -= me - i;
      i++;
    }
    int j = i + rem;
    int gi = gne0[i], gj = gne0[j];
    if (gi < 0 || gj < 0) continue;
    for (int s = 0; s < (i == j ? 1 : 2); s++) {
      int r = (s == 0 ? gi : gj) - 1, c = (s == 0 ? gj : gi) - 1;
      int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1, pos = -1;
      while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        if (col[mid] == c) {
          pos = mid;
          break;
        }
        if (col[mid] < c) lo = mid + 1;
        else hi = mid - 1;
      }
      if (pos < 0) continue;
      K[pos] += sign * dK[p];
      M[pos] += sign * dM[p];
      K32[pos] = (float)K[pos];
      M32[pos] = (float)M[pos];
    }
  }
}

// rows [r0, r1) of the reference's coordinate form: (ia, ja) 1-based,
// a = fl32(K) + i fl32(fl32(omega) M) on and above the diagonal and 0 below
// (ga_sort_sparse leaves the lower slots unassigned).  One wave per row, lanes
// over the row's entries.  DROP = false: every pattern entry, written at
// p - rowptr[r0].  DROP = true: only the nonzero ones (find_zeros/rem_zeros),
// in pattern order, row r's first at off[r] - off[r0].
template <bool DROP>
__global__ void k_gv_format(int32_t r0, int32_t r1, const int64_t *rowptr, const int32_t *col, const double *K,
                            const double *M, double w32, const int64_t *off, int32_t *ia, int32_t *ja, double *a) {
  const int lane = threadIdx.x & 63;
  const int64_t q0 = DROP ? off[r0] : rowptr[r0];
  const int64_t rstep = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t r = r0 + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; r < r1; r += rstep) {
    const int64_t pb = rowptr[r], pe = rowptr[r + 1];
    int64_t q = DROP ? off[r] - q0 : pb - q0;
    for (int64_t p1 = pb; p1 < pe; p1 += 64) {  // wave-uniform trip count
      const int64_t p = p1 + lane;
      const bool in = p < pe;
      const int32_t c = in ? col[p] : 0;
      const bool up = in && c >= r;
      const double re = up ? (double)(float)K[p] : 0.0, im = up ? (double)(float)(w32 * M[p]) : 0.0;
      bool keep = in;
      int64_t qi = q + lane;
      if (DROP) {
        keep = in && (re != 0.0 || im != 0.0);
        const uint64_t mask = __ballot(keep);
        qi = q + __popcll(mask & ((uint64_t(1) << lane) - 1));
        q += __popcll(mask);
      } else {
        q += 64;
      }
      if (keep && (!DROP || qi < off[r + 1] - q0)) {  // the row's counted slots bound its writes
        ia[qi] = (int32_t)r + 1;
        ja[qi] = c + 1;
        a[2 * qi] = re;
        a[2 * qi + 1] = im;
      }
    }
  }
}

// cnt[r - r0] = nonzero entries of row r in the reference's coordinate form
__global__ void k_gv_count(int32_t r0, int32_t r1, const int64_t *rowptr, const int32_t *col, const double *K,
                           const double *M, double w32, int32_t *cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t rstep = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t r = r0 + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; r < r1; r += rstep) {
    const int64_t pe = rowptr[r + 1];
    int n = 0;
    for (int64_t p1 = rowptr[r]; p1 < pe; p1 += 64) {
      const int64_t p = p1 + lane;
      bool keep = false;
      if (p < pe && col[p] >= r) keep = (double)(float)K[p] != 0.0 || (double)(float)(w32 * M[p]) != 0.0;
      n += __popcll(__ballot(keep));
    }
    if (lane == 0) cnt[r - r0] = n;
  }
}

static int ensure_call_state(mvf_handle *h, int64_t call) {
  if (!h->carry_nonzero) return 0;
  int want = call >= 1 ? 1 : 0;
  if (want == h->carry_applied) return 0;
  double sign = want ? 1.0 : -1.0;
  hipLaunchKernelGGL(k_apply_carry, dim3(1), dim3(256), 0, h->stream, h->d_gne, h->et.ne, h->d_rowptr, h->d_col,
                     h->d_dK0, h->d_dM0, h->d_K, h->d_M, h->d_K32, h->d_M32, sign);
  MVF_HIP(h, hipGetLastError());
  direct_invalidate(h);
  h->carry_applied = want;
  return 0;
}

static void inv6(const double *a, double *b, double &d) {
  d = a[0] * (a[3] * a[5] - a[4] * a[4]) + a[1] * (a[2] * a[4] - a[1] * a[5]) + a[2] * (a[1] * a[4] - a[3] * a[2]);
  b[0] = (a[3] * a[5] - a[4] * a[4]) / d;
  b[1] = (a[2] * a[4] - a[1] * a[5]) / d;
  b[2] = (a[1] * a[4] - a[2] * a[3]) / d;
  b[3] = (a[0] * a[5] - a[2] * a[2]) / d;
  b[4] = (a[2] * a[1] - a[0] * a[4]) / d;
  b[5] = (a[3] * a[0] - a[1] * a[1]) / d;
}

}  // namespace mvf

using namespace mvf;

extern "C" {

void mvf_default_solver_opts(mvf_solver_opts *o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->method = MVF_KRYLOV_COCG;
  o->precision = MVF_PREC_FP64;
  o->matrix = MVF_MATRIX_C64REF;
  o->precond = MVF_PC_JACOBI;
  o->tol = 1e-8;
  o->maxit = 100000;
  o->check_every = 25;
  o->inner_tol = 1e-2;
  o->use_graph = 0;
}

const char *mvf_last_error(const mvf_handle *h) { return h ? h->err.c_str() : mvf::g_host_err.c_str(); }

void mvf_free(mvf_handle *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  void *ptrs[] = {h->d_xp,  h->d_yp,  h->d_zp,   h->d_sig,   h->d_inmu,     h->d_dmu,      h->d_tab,
                  h->d_gne, h->d_rowptr, h->d_col, h->d_diagpos, h->d_seg, h->d_contrib, h->d_rseg,
                  h->d_rcontrib, h->d_K, h->d_M, h->d_K32, h->d_M32, h->d_Ke, h->d_Me, h->d_dK0, h->d_dM0,
                  h->d_x,  h->d_b, h->d_ntab, h->d_bel, h->d_bidx, h->d_bK, h->d_bM, h->d_dirf, h->d_btab};
  for (void *p : ptrs) dfree(p);
  for (void *p : h->owned) dfree(p);
  for (void *p : h->gv_pin)
    if (p) (void)hipHostFree(p);
  direct_free(h);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int mvf_init(const mvf_grid *grid, const mvf_model *model, const mvf_bc *bc, const mvf_options *opts,
             mvf_handle **out) {
  if (!out) return MVF_E_ARG;
  *out = nullptr;
  if (!grid || !model || !bc || !grid->xp || !grid->yp || !grid->zp || !model->sigma || !model->mu) return MVF_E_ARG;
  if (grid->nodes_per_elem != 8 && grid->nodes_per_elem != 20 && grid->nodes_per_elem != 27) return MVF_E_ARG;
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return MVF_E_ARG;
  mvf_handle *h = new (std::nothrow) mvf_handle();
  if (!h) return MVF_E_MEMORY;
  *out = h;
  h->device = opts ? opts->device : 0;
  h->emulate = opts ? opts->emulate : MVF_EMU_ALL;
  h->verbose = opts ? opts->verbose : 0;
  h->bc = *bc;
  h->bc.l_sigma = h->bc.l_dz = nullptr;  // the caller's arrays are copied below
  if (bc->dirichlet) {
    if (bc->inimod < 1 || bc->inimod > 3)
      return set_err(h, MVF_E_ARG, "Dirichlet boundary model inimod=%d (must be 1, 2 or 3)", bc->inimod);
    if (bc->inimod == 3) {
      if (bc->nl < 1 || !bc->l_sigma || (bc->nl > 1 && !bc->l_dz))
        return set_err(h, MVF_E_ARG, "inimod 3 needs nl >= 1 layer conductivities and nl-1 thicknesses");
      h->bd_lsig.assign(bc->l_sigma, bc->l_sigma + bc->nl);
      h->bd_ldz.assign(bc->nl - 1, 0.0);
      for (int l = 0; l + 1 < bc->nl; l++) h->bd_ldz[l] = bc->l_dz[l];
    }
    h->bd_on = bc->inimod != 1;
    h->bd_ztop = bc->ztop;
    h->bd_hsigma = bc->h_sigma;
  }
  MVF_HIP(h, hipSetDevice(h->device));
  {
    // MVF_ELEM_KERNEL=valu selects the one-workgroup-per-element VALU element
    // kernel instead of the matrix-core one (same results to rounding)
    const char *ek = getenv("MVF_ELEM_KERNEL");
    h->elem_legacy = (ek && strcmp(ek, "valu") == 0) ? 1 : 0;
  }
  // device address of this GPU's no-allocation scratch array (mvf_solve_noalloc)
  MVF_TRY(noalloc_setup(h));
  // a blocking stream: synchronous runtime calls on the null stream stay ordered after its work
  MVF_HIP(h, hipStreamCreate(&h->stream));
  h->nodes = grid->nodes_per_elem;
  h->et = elem_type(h->nodes);
  Geom &g = h->geo;
  g.nx = grid->nx; g.ny = grid->ny; g.nz = grid->nz;
  g.nord = h->et.nord;
  int o = g.nord - 1;
  g.nnx = g.nx * o + 1; g.nny = g.ny * o + 1; g.nnz = g.nz * o + 1;
  g.nyz = (int64_t)g.nny * g.nnz;
  g.npt = (int64_t)g.nnx * g.nyz;
  g.ne = (int64_t)g.nx * g.ny * g.nz;
  // element indices are decomposed into lattice positions in 32-bit arithmetic
  if (g.ne > INT32_MAX) return set_err(h, MVF_E_ARG, "grid too large: %lld elements, at most 2^31 - 1", (long long)g.ne);
  if ((int64_t)h->et.ne * h->et.ne * g.ne / h->et.ne > INT64_MAX / 64) return set_err(h, MVF_E_ARG, "grid too large");
  // GPML zone state (init_gpml, boundary_conds.f90:51-70)
  Gpml &gp = h->gp;
  gp.on = bc->dirichlet ? 0 : 1;
  gp.sch = bc->gpml_sch;
  gp.a0 = bc->a0;
  gp.nn = bc->nn;
  int nel[3] = {g.nx, g.ny, g.nz};
  for (int c = 0; c < 3; c++) {
    gp.el_a[c][0] = bc->nextd;
    gp.el_a[c][1] = nel[c] + 1 - bc->nextd;
    gp.el_b[c][0] = 1;
    gp.el_b[c][1] = nel[c];
  }
  int nd = bc->nextd * o;
  if (gp.on && (bc->nextd < 1 || nd + 1 > g.nnx || nd + 1 > g.nny || nd + 1 > g.nnz))
    return set_err(h, MVF_E_ARG, "nextd=%d does not fit the grid", bc->nextd);
  if (gp.on) {
    gp.a[0][0] = grid->xp[nd]; gp.a[0][1] = grid->xp[g.nnx - nd - 1];
    gp.b[0][0] = grid->xp[0]; gp.b[0][1] = grid->xp[g.nnx - 1];
    gp.a[1][0] = grid->yp[nd]; gp.a[1][1] = grid->yp[g.nny - nd - 1];
    gp.b[1][0] = grid->yp[0]; gp.b[1][1] = grid->yp[g.nny - 1];
    int64_t id2 = (int64_t)(g.nnx - 1) * g.nyz + (int64_t)(g.nny - 1) * g.nnz + g.nnz - (int64_t)bc->nzl_top * o - 1;
    if (id2 < 0 || id2 >= g.npt) return set_err(h, MVF_E_ARG, "nzl_top=%d does not fit the grid", bc->nzl_top);
    gp.a[2][0] = grid->zp[nd]; gp.a[2][1] = grid->zp[id2];
    gp.b[2][0] = grid->zp[0]; gp.b[2][1] = grid->zp[g.npt - 1];
    h->flags_first = 0;
    h->flags_late = zone_flags(gp, g.nx, g.ny, g.nz);
  }
  double t0 = now();
  // lattice and materials
  MVF_TRY(dalloc(h, &h->d_xp, g.nnx));
  MVF_TRY(dalloc(h, &h->d_yp, g.nny));
  MVF_TRY(dalloc(h, &h->d_zp, g.npt));
  MVF_TRY(dalloc(h, &h->d_sig, 12 * g.npt));
  MVF_TRY(dalloc(h, &h->d_inmu, 6 * g.npt));
  MVF_TRY(dalloc(h, &h->d_dmu, 6 * g.npt));
  MVF_HIP(h, hipMemcpy(h->d_xp, grid->xp, 8 * g.nnx, hipMemcpyHostToDevice));
  MVF_HIP(h, hipMemcpy(h->d_yp, grid->yp, 8 * g.nny, hipMemcpyHostToDevice));
  MVF_HIP(h, hipMemcpy(h->d_zp, grid->zp, 8 * g.npt, hipMemcpyHostToDevice));
  MVF_HIP(h, hipMemcpy(h->d_sig, model->sigma, 96 * g.npt, hipMemcpyHostToDevice));
  {
    // per-node inverse permeability and contrast (pinv_geomodel, pdelta_model,
    // problem.f90:257-335); a singular tensor stops the reference (:260-271)
    std::vector<double> inmu(6 * g.npt), dmu(6 * g.npt);
    const double mu0 = 4 * MVF_PI * 1.e-7;
    for (int64_t n = 0; n < g.npt; n++) {
      const double *m = model->mu + 6 * n;
      double d;
      inv6(m, &inmu[6 * n], d);
      if (d == 0.0) return set_err(h, MVF_E_SINGULAR, "singular mu tensor at node %lld", (long long)n + 1);
      const double *s = model->sigma + 12 * n;
      // complex determinant of the sigma tensor
      auto cm = [](double ar, double ai, double br, double bi, double &r, double &i) {
        r = ar * br - ai * bi;
        i = ar * bi + ai * br;
      };
      double p1r, p1i, p2r, p2i, tr, ti, sr = 0, si = 0;
      // a1*(a4*a6 - a5*a5)
      cm(s[6], s[7], s[10], s[11], p1r, p1i); cm(s[8], s[9], s[8], s[9], p2r, p2i);
      cm(s[0], s[1], p1r - p2r, p1i - p2i, tr, ti); sr += tr; si += ti;
      // a2*(a3*a5 - a2*a6)
      cm(s[4], s[5], s[8], s[9], p1r, p1i); cm(s[2], s[3], s[10], s[11], p2r, p2i);
      cm(s[2], s[3], p1r - p2r, p1i - p2i, tr, ti); sr += tr; si += ti;
      // a3*(a2*a5 - a4*a3)
      cm(s[2], s[3], s[8], s[9], p1r, p1i); cm(s[6], s[7], s[4], s[5], p2r, p2i);
      cm(s[4], s[5], p1r - p2r, p1i - p2i, tr, ti); sr += tr; si += ti;
      if (sr == 0.0 && si == 0.0)
        return set_err(h, MVF_E_SINGULAR, "singular sigma tensor at node %lld", (long long)n + 1);
      for (int c = 0; c < 6; c++) dmu[6 * n + c] = m[c];
      dmu[6 * n + 0] -= mu0; dmu[6 * n + 3] -= mu0; dmu[6 * n + 5] -= mu0;
    }
    MVF_HIP(h, hipMemcpy(h->d_inmu, inmu.data(), 48 * g.npt, hipMemcpyHostToDevice));
    MVF_HIP(h, hipMemcpy(h->d_dmu, dmu.data(), 48 * g.npt, hipMemcpyHostToDevice));
  }
  {
    BasisTab *tab = new BasisTab();
    memset(tab, 0, sizeof(BasisTab));
    fill_basis(h->nodes, tab);
    int rc = dalloc(h, &h->d_tab, 1);
    if (!rc && hipMemcpy(h->d_tab, tab, sizeof(BasisTab), hipMemcpyHostToDevice) != hipSuccess)
      rc = set_err(h, MVF_E_DEVICE, "basis table upload failed");
    if (!rc) rc = lane_tables(h, tab);
    delete tab;
    if (rc) return rc;
  }
  double t1 = now();
  MVF_TRY(setup_numbering(h));
  double t2 = now();
  MVF_TRY(setup_pattern(h));
  double t3 = now();
  int64_t np = (int64_t)h->et.ne * (h->et.ne + 1) / 2;
  MVF_TRY(dalloc(h, &h->d_Ke, np * g.ne));
  MVF_TRY(dalloc(h, &h->d_Me, np * g.ne));
  int ff = first_flags_for(h, 0);
  MVF_TRY(integrate_elements(h, 0, g.ne, ff, h->d_Ke, h->d_Me));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  double t4 = now();
  MVF_TRY(assemble_operators(h));
  if (h->bd_on) MVF_TRY(bd_setup(h, h->d_Ke, h->d_Me));
  // element 0 under the carried flags (see MVF_EMU_GPML_PREV_FLAGS)
  if (gp.on && (h->emulate & MVF_EMU_GPML_PREV_FLAGS) && h->flags_late != h->flags_first) {
    MVF_TRY(dalloc(h, &h->d_dK0, np));
    MVF_TRY(dalloc(h, &h->d_dM0, np));
    MVF_TRY(integrate_elements(h, 0, 1, h->flags_late, h->d_dK0, h->d_dM0));
    std::vector<double> a(np), b(np), c(np), d(np);
    MVF_HIP(h, hipMemcpy(a.data(), h->d_dK0, 8 * np, hipMemcpyDeviceToHost));
    MVF_HIP(h, hipMemcpy(b.data(), h->d_dM0, 8 * np, hipMemcpyDeviceToHost));
    MVF_HIP(h, hipMemcpy(c.data(), h->d_Ke, 8 * np, hipMemcpyDeviceToHost));
    MVF_HIP(h, hipMemcpy(d.data(), h->d_Me, 8 * np, hipMemcpyDeviceToHost));
    bool nz = false;
    for (int64_t p = 0; p < np; p++) {
      a[p] -= c[p];
      b[p] -= d[p];
      nz |= (a[p] != 0.0 || b[p] != 0.0);
    }
    MVF_HIP(h, hipMemcpy(h->d_dK0, a.data(), 8 * np, hipMemcpyHostToDevice));
    MVF_HIP(h, hipMemcpy(h->d_dM0, b.data(), 8 * np, hipMemcpyHostToDevice));
    h->carry_nonzero = nz ? 1 : 0;
  }
  h->carry_applied = 0;
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  double t5 = now();
  dfree(h->d_Ke); h->d_Ke = nullptr;
  dfree(h->d_Me); h->d_Me = nullptr;
  h->t_setup[0] = t2 - t1;
  h->t_setup[1] = t3 - t2;
  h->t_setup[2] = t4 - t3;
  h->t_setup[3] = t5 - t4;
  (void)t0;
  return 0;
}

int mvf_assemble(mvf_handle *h, int32_t precision, double *integrate_ms) {
  DevScope ds;
  if (!h || (precision != 0 && precision != 1)) return h ? set_err(h, MVF_E_ARG, "precision must be 0 or 1") : MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  const int64_t ne = h->geo.ne, np = (int64_t)h->et.ne * (h->et.ne + 1) / 2;
  // element 0 takes the flags of the current call state directly
  const int ff = first_flags_for(h, h->calls);
  hipEvent_t ev[2];
  MVF_HIP(h, hipEventCreate(&ev[0]));
  MVF_HIP(h, hipEventCreate(&ev[1]));
  int rc = 0;
  if (precision == 0) {
    MVF_TRY(dalloc(h, &h->d_Ke, np * ne));
    MVF_TRY(dalloc(h, &h->d_Me, np * ne));
    MVF_HIP(h, hipEventRecord(ev[0], h->stream));
    rc = integrate_elements(h, 0, ne, ff, h->d_Ke, h->d_Me);
    (void)hipEventRecord(ev[1], h->stream);
    if (!rc) rc = assemble_operators(h);
    dfree(h->d_Ke); dfree(h->d_Me);
    h->d_Ke = h->d_Me = nullptr;
  } else {
    float *k32, *m32;
    MVF_TRY(ds.alloc(h, &k32, np * ne));
    MVF_TRY(ds.alloc(h, &m32, np * ne));
    MVF_HIP(h, hipEventRecord(ev[0], h->stream));
    rc = integrate_elements32(h, 0, ne, ff, k32, m32);
    (void)hipEventRecord(ev[1], h->stream);
    if (!rc) rc = assemble_operators32(h, k32, m32);
    ds.drop(k32); ds.drop(m32);
  }
  direct_invalidate(h);
  if (!rc) {
    MVF_HIP(h, hipStreamSynchronize(h->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (integrate_ms) *integrate_ms = ms;
    h->carry_applied = (h->carry_nonzero && h->calls >= 1) ? 1 : 0;
    h->ops_fp32 = precision;
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return rc;
}

int mvf_sizes(const mvf_handle *h, int32_t *nne, int64_t *nnze, int64_t *nnz_triu) {
  if (!h) return MVF_E_ARG;
  if (nne) *nne = h->nne;
  if (nnze) *nnze = h->nnz;
  if (nnz_triu) *nnz_triu = h->nnz_triu;
  return 0;
}

int mvf_setup_times(const mvf_handle *h, double *t4) {
  if (!h || !t4) return MVF_E_ARG;
  for (int i = 0; i < 4; i++) t4[i] = h->t_setup[i];
  return 0;
}

int mvf_get_operators(mvf_handle *h, int64_t *rowptr, int32_t *col, double *K, double *M) {
  if (!h) return MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  if (rowptr) MVF_HIP(h, hipMemcpy(rowptr, h->d_rowptr, 8 * ((size_t)h->nne + 1), hipMemcpyDeviceToHost));
  if (col) MVF_HIP(h, hipMemcpy(col, h->d_col, 4 * (size_t)h->nnz, hipMemcpyDeviceToHost));
  if (K) MVF_HIP(h, hipMemcpy(K, h->d_K, 8 * (size_t)h->nnz, hipMemcpyDeviceToHost));
  if (M) MVF_HIP(h, hipMemcpy(M, h->d_M, 8 * (size_t)h->nnz, hipMemcpyDeviceToHost));
  return 0;
}

int mvf_get_rows(mvf_handle *h, int64_t r0, int64_t r1, int64_t *rowptr, int32_t *col, double *K, double *M) {
  if (!h || r0 < 0 || r1 > h->nne || r1 < r0 || !rowptr) return h ? set_err(h, MVF_E_ARG, "bad row range") : MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  MVF_HIP(h, hipMemcpy(rowptr, h->d_rowptr + r0, 8 * (size_t)(r1 - r0 + 1), hipMemcpyDeviceToHost));
  const int64_t p0 = rowptr[0], n = rowptr[r1 - r0] - p0;
  for (int64_t i = 0; i <= r1 - r0; i++) rowptr[i] -= p0;
  if (col) MVF_HIP(h, hipMemcpy(col, h->d_col + p0, 4 * (size_t)n, hipMemcpyDeviceToHost));
  if (K) MVF_HIP(h, hipMemcpy(K, h->d_K + p0, 8 * (size_t)n, hipMemcpyDeviceToHost));
  if (M) MVF_HIP(h, hipMemcpy(M, h->d_M + p0, 8 * (size_t)n, hipMemcpyDeviceToHost));
  return 0;
}

int mvf_get_gne(mvf_handle *h, int32_t *gne_out) {
  if (!h || !gne_out) return MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  MVF_HIP(h, hipMemcpy(gne_out, h->d_gne, sizeof(int32_t) * h->geo.ne * h->et.ne, hipMemcpyDeviceToHost));
  return 0;
}

int mvf_set_frequency(mvf_handle *h, double freq_hz) {
  if (!h || !(freq_hz > 0)) return MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  h->freq = freq_hz;
  h->omega = 2.0 * MVF_PI * freq_hz;
  h->have_freq = 1;
  MVF_TRY(update_sigma(h, h->omega));
  return 0;
}

int mvf_integrate(mvf_handle *h, int64_t e0, int64_t e1, double *Ke, double *Me, double *be) {
  DevScope ds;
  if (!h || e0 < 0 || e1 > h->geo.ne || e1 < e0) return MVF_E_ARG;
  if (be && !h->have_freq) return set_err(h, MVF_E_ARG, "mvf_set_frequency must precede element RHS");
  MVF_HIP(h, hipSetDevice(h->device));
  int64_t n = e1 - e0;
  if (n == 0) return 0;
  int64_t np = (int64_t)h->et.ne * (h->et.ne + 1) / 2;
  int ff = first_flags_for(h, h->calls);
  if (Ke || Me) {
    double *dK, *dM;
    MVF_TRY(ds.alloc(h, &dK, np * n));
    MVF_TRY(ds.alloc(h, &dM, np * n));
    int rc = integrate_elements(h, e0, e1, ff, dK, dM);
    if (!rc && Ke)
      rc = hipMemcpyAsync(Ke, dK, 8 * np * n, hipMemcpyDeviceToHost, h->stream) == hipSuccess ? 0 : MVF_E_DEVICE;
    if (!rc && Me)
      rc = hipMemcpyAsync(Me, dM, 8 * np * n, hipMemcpyDeviceToHost, h->stream) == hipSuccess ? 0 : MVF_E_DEVICE;
    if (!rc) rc = hipStreamSynchronize(h->stream) == hipSuccess ? 0 : MVF_E_DEVICE;
    ds.drop(dK);
    ds.drop(dM);
    if (rc) return rc;
  }
  if (be) {
    double *d;
    MVF_TRY(ds.alloc(h, &d, (size_t)n * h->et.ne * 4));
    int rc = element_rhs_range(h, h->omega, ff, e0, e1, d);
    if (!rc)
      rc = hipMemcpyAsync(be, d, 8 * (size_t)n * h->et.ne * 4, hipMemcpyDeviceToHost, h->stream) == hipSuccess
               ? 0
               : MVF_E_DEVICE;
    if (!rc) rc = hipStreamSynchronize(h->stream) == hipSuccess ? 0 : MVF_E_DEVICE;
    ds.drop(d);
    if (rc) return rc;
  }
  return 0;
}

// RHS of the current frequency into column pair col0 of an [nne][ncol] device array
static int build_rhs(mvf_handle *h, double omega, int64_t call, double *d_b, int ncol, int col0) {
  DevScope ds;
  double *be;
  MVF_TRY(ds.alloc(h, &be, (size_t)h->geo.ne * h->et.ne * 4));
  int rc = element_rhs_range(h, omega, first_flags_for(h, call), 0, h->geo.ne, be);
  if (!rc) rc = bd_lift(h, omega, be);
  if (!rc) rc = assemble_rhs(h, be, d_b, ncol, col0);
  ds.drop(be);
  return rc;
}

// dst[i] = src[i] for the three arrays of a staged chunk, split over host
// threads when the chunk is large (the caller's arrays are pageable memory,
// written at the host's copy rate; one thread does not reach the DMA rate)
static void par_copy(void *const dst[3], const void *const src[3], const size_t bytes[3]) {
  const size_t total = bytes[0] + bytes[1] + bytes[2];
  const unsigned hw = std::thread::hardware_concurrency();
  const int nt = total < (size_t(8) << 20) ? 1 : (int)std::min<unsigned>(8, hw ? hw : 1);
  auto part = [&](int t) {
    for (int k = 0; k < 3; k++) {
      const size_t b0 = bytes[k] * t / nt, b1 = bytes[k] * (t + 1) / nt;
      if (b1 > b0) memcpy((char *)dst[k] + b0, (const char *)src[k] + b0, b1 - b0);
    }
  };
  if (nt == 1) {
    part(0);
    return;
  }
  // a thread that cannot be started leaves its part (and the later ones) to
  // this thread: nothing is thrown across the C interface
  std::vector<std::thread> th;
  int started = 1;
  try {
    th.reserve(nt - 1);
    for (int t = 1; t < nt; t++) {
      th.emplace_back(part, t);
      started = t + 1;
    }
  } catch (...) {
  }
  for (int t = started; t < nt; t++) part(t);
  part(0);
  for (auto &x : th) x.join();
}

struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~EventPair() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// global_vfem's (ia, ja, a, b) into the caller's host arrays; with drop, the
// triplet after find_zeros/rem_zeros, at most cap entries, its length in *n_out.
// The coordinate form is written on the device in row chunks of at most
// `chunk` entries (a row is never split; MVF_GV_CHUNK overrides the size),
// copied by DMA into one of two pinned host buffers, and from there into the
// caller's arrays by host threads while the next chunk is formatted and
// copied into the other buffer.
static int global_coo(mvf_handle *h, int32_t *ia, int32_t *ja, double *a, double *b, bool drop, int64_t cap,
                      int64_t *n_out) {
  DevScope ds;
  const int32_t nne = h->nne;
  if (!h->have_freq) return set_err(h, MVF_E_ARG, "mvf_set_frequency must precede mvf_global_vfem");
  if (h->ops_fp32)
    return set_err(h, MVF_E_ARG, "K and M were integrated in FP32: mvf_assemble(h, 0, ...) restores the FP64 matrix");
  MVF_HIP(h, hipSetDevice(h->device));
  MVF_TRY(ensure_call_state(h, h->calls));
  double *db;
  MVF_TRY(ds.alloc(h, &db, (size_t)nne * 4));
  MVF_TRY(build_rhs(h, h->omega, h->calls, db, 2, 0));
  const double w32 = (double)(float)h->omega;
  auto waves_grid = [](int64_t rows) { return dim3((unsigned)std::min<int64_t>((rows + 3) / 4, 65535 * 16)); };
  // rp: row starts of what is written (the pattern's, or the nonzero counts' prefix sums)
  std::vector<int64_t> rp(nne + 1);
  std::vector<double> bb((size_t)nne * 4);
  int64_t *doff = nullptr;
  if (drop) {
    int32_t *dcnt;
    MVF_TRY(ds.alloc(h, &dcnt, (size_t)nne));
    MVF_TRY(ds.alloc(h, &doff, (size_t)nne + 1));
    hipLaunchKernelGGL(k_gv_count, waves_grid(nne), dim3(256), 0, h->stream, 0, nne, h->d_rowptr, h->d_col, h->d_K,
                       h->d_M, w32, dcnt);
    MVF_HIP(h, hipGetLastError());
    std::vector<int32_t> cnt(nne);
    MVF_HIP(h, hipMemcpyAsync(cnt.data(), dcnt, 4 * (size_t)nne, hipMemcpyDeviceToHost, h->stream));
    MVF_HIP(h, hipStreamSynchronize(h->stream));
    rp[0] = 0;
    for (int32_t r = 0; r < nne; r++) rp[r + 1] = rp[r] + cnt[r];
    if (rp[nne] > cap)
      return set_err(h, MVF_E_ARG, "%lld nonzero entries do not fit the arrays' %lld", (long long)rp[nne],
                     (long long)cap);
    MVF_HIP(h, hipMemcpyAsync(doff, rp.data(), 8 * ((size_t)nne + 1), hipMemcpyHostToDevice, h->stream));
  } else {
    MVF_HIP(h, hipMemcpyAsync(rp.data(), h->d_rowptr, 8 * ((size_t)nne + 1), hipMemcpyDeviceToHost, h->stream));
  }
  MVF_HIP(h, hipMemcpyAsync(bb.data(), db, 32 * (size_t)nne, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  ds.drop(db);
  int64_t chunk = int64_t(1) << 22;
  if (const char *e = getenv("MVF_GV_CHUNK")) chunk = std::max<int64_t>(1, atoll(e));
  int64_t stage = 0;
  for (int32_t r = 0, r1; r < nne; r = r1) {
    for (r1 = r + 1; r1 < nne && rp[r1 + 1] - rp[r] <= chunk; r1++) {
    }
    stage = std::max(stage, rp[r1] - rp[r]);
  }
  int32_t *sia, *sja;
  double *sa;
  MVF_TRY(ds.alloc(h, &sia, (size_t)std::max<int64_t>(stage, 1)));
  MVF_TRY(ds.alloc(h, &sja, (size_t)std::max<int64_t>(stage, 1)));
  MVF_TRY(ds.alloc(h, &sa, (size_t)std::max<int64_t>(2 * stage, 1)));
  if (h->gv_pin_cap < stage) {
    for (void *&p : h->gv_pin)
      if (p) {
        (void)hipHostFree(p);
        p = nullptr;
      }
    h->gv_pin_cap = 0;
    for (void *&p : h->gv_pin)
      if (hipHostMalloc(&p, 24 * (size_t)stage, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        return set_err(h, MVF_E_MEMORY, "pinned host staging of %lld bytes failed", (long long)(48 * stage));
      }
    h->gv_pin_cap = stage;
  }
  // pinned buffer layout: ia (4 B per entry), ja (4 B), a (16 B)
  auto pin = [&](int k, int part) {
    char *b = (char *)h->gv_pin[k];
    return part == 0 ? (void *)b : part == 1 ? (void *)(b + 4 * h->gv_pin_cap) : (void *)(b + 8 * h->gv_pin_cap);
  };
  EventPair ev;
  for (hipEvent_t &e : ev.e) MVF_HIP(h, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct Staged {
    int64_t p0 = -1, np = 0;
    int k = 0;
  } prev;
  // waits for a staged chunk's DMA, then copies it into the caller's arrays
  auto drain = [&](const Staged &c) -> int {
    MVF_HIP(h, hipEventSynchronize(ev.e[c.k]));
    void *dst[3] = {ia + c.p0, ja + c.p0, a + 2 * c.p0};
    const void *src[3] = {pin(c.k, 0), pin(c.k, 1), pin(c.k, 2)};
    const size_t bytes[3] = {4 * (size_t)c.np, 4 * (size_t)c.np, 16 * (size_t)c.np};
    par_copy(dst, src, bytes);
    return 0;
  };
  int k = 0;
  for (int32_t r = 0, r1; r < nne; r = r1) {
    for (r1 = r + 1; r1 < nne && rp[r1 + 1] - rp[r] <= chunk; r1++) {
    }
    const int64_t p0 = rp[r], np = rp[r1] - rp[r];
    if (np == 0) continue;
    if (drop)
      hipLaunchKernelGGL(k_gv_format<true>, waves_grid(r1 - r), dim3(256), 0, h->stream, r, r1, h->d_rowptr, h->d_col,
                         h->d_K, h->d_M, w32, doff, sia, sja, sa);
    else
      hipLaunchKernelGGL(k_gv_format<false>, waves_grid(r1 - r), dim3(256), 0, h->stream, r, r1, h->d_rowptr,
                         h->d_col, h->d_K, h->d_M, w32, doff, sia, sja, sa);
    MVF_HIP(h, hipGetLastError());
    // buffer k was drained (synchronously, below) before this chunk is queued into it
    MVF_HIP(h, hipMemcpyAsync(pin(k, 0), sia, 4 * np, hipMemcpyDeviceToHost, h->stream));
    MVF_HIP(h, hipMemcpyAsync(pin(k, 1), sja, 4 * np, hipMemcpyDeviceToHost, h->stream));
    MVF_HIP(h, hipMemcpyAsync(pin(k, 2), sa, 16 * np, hipMemcpyDeviceToHost, h->stream));
    MVF_HIP(h, hipEventRecord(ev.e[k], h->stream));
    if (prev.p0 >= 0) MVF_TRY(drain(prev));
    prev = {p0, np, k};
    k ^= 1;
  }
  if (prev.p0 >= 0) MVF_TRY(drain(prev));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  for (int32_t r = 0; r < nne; r++)
    for (int d = 0; d < 2; d++) {
      b[2 * ((size_t)d * nne + r)] = bb[4 * (size_t)r + 2 * d];
      b[2 * ((size_t)d * nne + r) + 1] = bb[4 * (size_t)r + 2 * d + 1];
    }
  if (n_out) *n_out = rp[nne];
  h->calls++;
  return 0;
}

int mvf_global_vfem(mvf_handle *h, int32_t *ia, int32_t *ja, double *a, double *b, int64_t nnze, int32_t nne,
                    int32_t ndir) {
  if (!h || !ia || !ja || !a || !b) return MVF_E_ARG;
  if (nnze != h->nnz || nne != h->nne || ndir != 2)
    return set_err(h, MVF_E_ARG, "size mismatch: nnze %lld/%lld nne %d/%d ndir %d", (long long)nnze,
                   (long long)h->nnz, nne, h->nne, ndir);
  return global_coo(h, ia, ja, a, b, false, nnze, nullptr);
}

int mvf_global_vfem_nz(mvf_handle *h, int32_t *ia, int32_t *ja, double *a, double *b, int64_t cap, int32_t nne,
                       int32_t ndir, int64_t *nnz_out) {
  if (!h || !ia || !ja || !a || !b || !nnz_out) return MVF_E_ARG;
  if (cap < 0 || nne != h->nne || ndir != 2)
    return set_err(h, MVF_E_ARG, "size mismatch: cap %lld nne %d/%d ndir %d", (long long)cap, nne, h->nne, ndir);
  return global_coo(h, ia, ja, a, b, true, cap, nnz_out);
}

int mvf_remove_zeros(int64_t n, int32_t *ia, int32_t *ja, double *a, int64_t *n_out) {
  if (n < 0 || (n > 0 && (!ia || !ja || !a)) || !n_out) return MVF_E_ARG;
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++) {
    if (a[2 * i] == 0.0 && a[2 * i + 1] == 0.0) continue;
    ia[k] = ia[i];
    ja[k] = ja[i];
    a[2 * k] = a[2 * i];
    a[2 * k + 1] = a[2 * i + 1];
    k++;
  }
  *n_out = k;
  return 0;
}

static int pad_cols(int n) { return n <= 2 ? 2 : (n <= 4 ? 4 : (n <= 8 ? 8 : (n <= 16 ? 16 : -1))); }

int mvf_solve(mvf_handle *h, double *rhs_inout, const mvf_solver_opts *o, mvf_solve_info *info) {
  DevScope ds;
  if (!h || !rhs_inout) return MVF_E_ARG;
  if (!h->have_freq) return set_err(h, MVF_E_ARG, "mvf_set_frequency must precede mvf_solve");
  mvf_solver_opts od;
  mvf_default_solver_opts(&od);
  if (!o) o = &od;
  mvf_solve_info dummy;
  if (!info) info = &dummy;
  memset(info, 0, sizeof(*info));
  MVF_HIP(h, hipSetDevice(h->device));
  int32_t nne = h->nne;
  std::vector<double> hb((size_t)nne * 4);
  for (int32_t r = 0; r < nne; r++)
    for (int d = 0; d < 2; d++) {
      hb[4 * (size_t)r + 2 * d] = rhs_inout[2 * ((size_t)d * nne + r)];
      hb[4 * (size_t)r + 2 * d + 1] = rhs_inout[2 * ((size_t)d * nne + r) + 1];
    }
  double *db, *dx;
  MVF_TRY(ds.alloc(h, &db, (size_t)nne * 4));
  MVF_TRY(ds.alloc(h, &dx, (size_t)nne * 4));
  MVF_HIP(h, hipMemcpyAsync(db, hb.data(), 32 * (size_t)nne, hipMemcpyHostToDevice, h->stream));
  MVF_HIP(h, hipMemsetAsync(dx, 0, 32 * (size_t)nne, h->stream));
  double om[2] = {h->omega, h->omega};
  double t0 = now();
  int rc = solve_device(h, om, 2, 2, o, db, dx, info);
  (void)hipStreamSynchronize(h->stream);
  info->seconds = now() - t0;
  info->operator_fp32 = h->ops_fp32;
  if (hipMemcpy(hb.data(), dx, 32 * (size_t)nne, hipMemcpyDeviceToHost) != hipSuccess && !rc) rc = MVF_E_DEVICE;
  ds.drop(db);
  ds.drop(dx);
  for (int32_t r = 0; r < nne; r++)
    for (int d = 0; d < 2; d++) {
      rhs_inout[2 * ((size_t)d * nne + r)] = hb[4 * (size_t)r + 2 * d];
      rhs_inout[2 * ((size_t)d * nne + r) + 1] = hb[4 * (size_t)r + 2 * d + 1];
    }
  return rc;
}

int mvf_solve_guess(mvf_handle *h, const double *x0, double *rhs_inout, const mvf_solver_opts *o,
                    mvf_solve_info *info, double *alpha_out) {
  DevScope ds;
  if (!h || !rhs_inout || !x0) return MVF_E_ARG;
  if (!h->have_freq) return set_err(h, MVF_E_ARG, "mvf_set_frequency must precede mvf_solve_guess");
  mvf_solver_opts od;
  mvf_default_solver_opts(&od);
  if (!o) o = &od;
  mvf_solve_info dummy;
  if (!info) info = &dummy;
  memset(info, 0, sizeof(*info));
  MVF_HIP(h, hipSetDevice(h->device));
  int32_t nne = h->nne;
  std::vector<double> hb((size_t)nne * 4), hx((size_t)nne * 4);
  for (int32_t r = 0; r < nne; r++)
    for (int d = 0; d < 2; d++)
      for (int p = 0; p < 2; p++) {
        hb[4 * (size_t)r + 2 * d + p] = rhs_inout[2 * ((size_t)d * nne + r) + p];
        hx[4 * (size_t)r + 2 * d + p] = x0[2 * ((size_t)d * nne + r) + p];
      }
  double *db, *dx;
  MVF_TRY(ds.alloc(h, &db, (size_t)nne * 4));
  MVF_TRY(ds.alloc(h, &dx, (size_t)nne * 4));
  MVF_HIP(h, hipMemcpyAsync(db, hb.data(), 32 * (size_t)nne, hipMemcpyHostToDevice, h->stream));
  MVF_HIP(h, hipMemcpyAsync(dx, hx.data(), 32 * (size_t)nne, hipMemcpyHostToDevice, h->stream));
  double om[2] = {h->omega, h->omega};
  double t0 = now();
  int rc = solve_device_guess(h, om, o, db, dx, info, alpha_out);
  (void)hipStreamSynchronize(h->stream);
  info->seconds = now() - t0;
  info->operator_fp32 = h->ops_fp32;
  if (hipMemcpy(hb.data(), dx, 32 * (size_t)nne, hipMemcpyDeviceToHost) != hipSuccess && !rc) rc = MVF_E_DEVICE;
  for (int32_t r = 0; r < nne; r++)
    for (int d = 0; d < 2; d++) {
      rhs_inout[2 * ((size_t)d * nne + r)] = hb[4 * (size_t)r + 2 * d];
      rhs_inout[2 * ((size_t)d * nne + r) + 1] = hb[4 * (size_t)r + 2 * d + 1];
    }
  return rc;
}

int mvf_solve_frequencies(mvf_handle *h, const double *freqs_hz, int32_t nfreq, const mvf_solver_opts *o,
                          mvf_solve_info *info) {
  DevScope ds;
  if (!h || !freqs_hz || nfreq < 1 || nfreq > 32) return h ? set_err(h, MVF_E_ARG, "nfreq must be 1..32") : MVF_E_ARG;
  mvf_solver_opts od;
  mvf_default_solver_opts(&od);
  if (!o) o = &od;
  mvf_solve_info dummy;
  if (!info) info = &dummy;
  memset(info, 0, sizeof(*info));
  MVF_HIP(h, hipSetDevice(h->device));
  int32_t nne = h->nne;
  int ncol = 2 * nfreq;
  size_t need = (size_t)nne * ncol * 2;
  if (h->xb_cap < need) {
    dfree(h->d_x);
    dfree(h->d_b);
    h->d_x = h->d_b = nullptr;
    MVF_TRY(dalloc(h, &h->d_x, need));
    MVF_TRY(dalloc(h, &h->d_b, need));
    h->xb_cap = need;
  }
  h->ncols_last = ncol;
  h->last_freqs.assign(freqs_hz, freqs_hz + nfreq);
  double t0 = now();
  // batches: the first assembly call (if still pending and the carried GPML
  // flags change element 0) is solved alone, the rest together
  int f = 0;
  int rc = 0;
  while (f < nfreq && !rc) {
    // at most 8 frequencies (16 columns) share one batched sweep
    int nb = std::min(nfreq - f, 8);
    if (h->carry_nonzero && h->calls == 0) nb = 1;
    // the direct preconditioner serves freqs_per_factor consecutive frequencies
    if (o->precond == MVF_PC_DIRECT) nb = std::min(nb, std::max(1, std::min(o->freqs_per_factor, 8)));
    MVF_TRY(ensure_call_state(h, h->calls));
    int pc = pad_cols(2 * nb);
    double *bb, *xb;
    MVF_TRY(ds.alloc(h, &bb, (size_t)nne * pc * 2));
    MVF_TRY(ds.alloc(h, &xb, (size_t)nne * pc * 2));
    MVF_HIP(h, hipMemsetAsync(bb, 0, 16 * (size_t)nne * pc, h->stream));
    MVF_HIP(h, hipMemsetAsync(xb, 0, 16 * (size_t)nne * pc, h->stream));
    double om[16] = {0};
    for (int q = 0; q < nb; q++) {
      MVF_TRY(mvf_set_frequency(h, freqs_hz[f + q]));
      MVF_TRY(build_rhs(h, h->omega, h->calls + q, bb, pc, 2 * q));
      om[2 * q] = om[2 * q + 1] = h->omega;
    }
    mvf_solve_info bi;
    memset(&bi, 0, sizeof(bi));
    rc = solve_device(h, om, 2 * nb, pc, o, bb, xb, &bi);
    info->factor_seconds += bi.factor_seconds;
    info->factor_flops = bi.factor_flops;
    info->factor_bytes = bi.factor_bytes;
    info->factorizations = bi.factorizations;
    info->static_pivots = std::max(info->static_pivots, bi.static_pivots);
    info->precond_applies += bi.precond_applies;
    if (bi.gemm_launches) {
      info->gemm_ms = bi.gemm_ms;
      info->gemm_flops = bi.gemm_flops;
      info->gemm_launches = bi.gemm_launches;
    }
    info->iterations += bi.iterations;
    info->graph_launches += bi.graph_launches;
    info->refinements += bi.refinements;
    info->fallbacks += bi.fallbacks;
    // SpMV timing is reported for the widest batch (bytes are per launch and
    // depend on the column count, so batches of different widths do not mix)
    if (bi.spmv_ncol >= info->spmv_ncol) {
      for (int p = 0; p < 2; p++) {
        info->spmv_launches[p] = bi.spmv_launches[p];
        info->spmv_ms[p] = bi.spmv_ms[p];
        info->spmv_bytes[p] = bi.spmv_bytes[p];
      }
      info->spmv_ncol = bi.spmv_ncol;
      info->spmv_lpr = bi.spmv_lpr;
    }
    for (int c = 0; c < 2 * nb && 2 * f + c < 64; c++) info->true_resid[2 * f + c] = bi.true_resid[c];
    MVF_HIP(h, hipMemcpy2DAsync(h->d_x + 2 * (size_t)(2 * f), 16 * (size_t)ncol, xb, 16 * (size_t)pc, 16 * (size_t)(2 * nb),
                                nne, hipMemcpyDeviceToDevice, h->stream));
    MVF_HIP(h, hipMemcpy2DAsync(h->d_b + 2 * (size_t)(2 * f), 16 * (size_t)ncol, bb, 16 * (size_t)pc, 16 * (size_t)(2 * nb),
                                nne, hipMemcpyDeviceToDevice, h->stream));
    MVF_HIP(h, hipStreamSynchronize(h->stream));
    ds.drop(bb);
    ds.drop(xb);
    h->calls += nb;
    f += nb;
  }
  info->converged = rc == 0;
  info->seconds = now() - t0;
  info->operator_fp32 = h->ops_fp32;
  return rc;
}

int mvf_solve_batch(mvf_handle *const *hs, int32_t nh, const double *freqs_hz, int32_t nfreq,
                    const mvf_solver_opts *o, mvf_solve_info *infos, int32_t *status, int32_t nthreads) {
  if (!hs || nh < 1 || !freqs_hz || nfreq < 1 || nfreq > 32) return MVF_E_ARG;
  for (int32_t i = 0; i < nh; i++)
    if (!hs[i]) return MVF_E_ARG;
  const int nt = std::max(1, std::min<int>(nthreads > 0 ? nthreads : 8, nh));
  // several host threads: no stream capture, since another thread's copy on
  // the null stream synchronizes with every blocking stream and would break a
  // capture in progress on one of them
  mvf_solver_opts od;
  if (o) {
    od = *o;
  } else {
    mvf_default_solver_opts(&od);
  }
  if (nt > 1) od.use_graph = 0;
  o = &od;
  std::atomic<int> next{0};
  std::vector<int> rcs(nh, 0);
  auto worker = [&]() {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= nh) break;
      rcs[i] = mvf_solve_frequencies(hs[i], freqs_hz + (size_t)i * nfreq, nfreq, o, infos ? &infos[i] : nullptr);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) pool.emplace_back(worker);
  worker();
  for (auto &t : pool) t.join();
  int rc = 0;
  for (int32_t i = 0; i < nh; i++) {
    if (status) status[i] = rcs[i];
    if (!rc && rcs[i]) rc = rcs[i];
  }
  return rc;
}

static int fetch_cols(mvf_handle *h, const double *d, double *out, int32_t ncols) {
  if (!h || !out || !d || ncols != h->ncols_last) return h ? set_err(h, MVF_E_ARG, "column count mismatch") : MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  int32_t nne = h->nne;
  std::vector<double> t((size_t)nne * ncols * 2);
  MVF_HIP(h, hipMemcpy(t.data(), d, t.size() * 8, hipMemcpyDeviceToHost));
  for (int32_t r = 0; r < nne; r++)
    for (int c = 0; c < ncols; c++) {
      out[2 * ((size_t)c * nne + r)] = t[2 * ((size_t)r * ncols + c)];
      out[2 * ((size_t)c * nne + r) + 1] = t[2 * ((size_t)r * ncols + c) + 1];
    }
  return 0;
}

int mvf_get_solution(mvf_handle *h, double *x_out, int32_t ncols) { return fetch_cols(h, h ? h->d_x : nullptr, x_out, ncols); }
int mvf_get_rhs(mvf_handle *h, double *b_out, int32_t ncols) { return fetch_cols(h, h ? h->d_b : nullptr, b_out, ncols); }

int mvf_readout(mvf_handle *h, const double *x, int32_t f, int32_t ke_top, int32_t nrx, const double *rx,
                const double *ry, double *esol, double *hsol, double *rx_e, double *rx_h, double *rx_z, double *rho,
                double *phi, double *rx_elev, double *node_z, double *node_rho, double *node_phi) {
  DevScope ds;
  if (!h) return MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  const Geom &g = h->geo;
  const int o = g.nord - 1;
  double omega;
  int ncol, col0;
  double *dx = nullptr;
  if (x) {
    if (!h->have_freq) return set_err(h, MVF_E_ARG, "mvf_readout: no current frequency");
    omega = h->omega;
    ncol = 2;
    col0 = 0;
    std::vector<double> t((size_t)h->nne * 4);
    for (int64_t r = 0; r < h->nne; r++)
      for (int d = 0; d < 2; d++) {
        t[4 * r + 2 * d] = x[2 * ((size_t)d * h->nne + r)];
        t[4 * r + 2 * d + 1] = x[2 * ((size_t)d * h->nne + r) + 1];
      }
    MVF_TRY(ds.alloc(h, &dx, t.size()));
    MVF_HIP(h, hipMemcpy(dx, t.data(), t.size() * 8, hipMemcpyHostToDevice));
  } else {
    if (!h->d_x || f < 0 || f >= (int)h->last_freqs.size())
      return set_err(h, MVF_E_ARG, "mvf_readout: frequency %d not in the last batch", f);
    omega = 2.0 * MVF_PI * h->last_freqs[f];
    ncol = h->ncols_last;
    col0 = 2 * f;
  }
  if (nrx < 0 || (nrx > 0 && (!rx || !ry))) return set_err(h, MVF_E_ARG, "mvf_readout: bad receiver arguments");
  if (nrx > 0 && (ke_top < 1 || ke_top > g.nz)) return set_err(h, MVF_E_ARG, "mvf_readout: ke_top=%d", ke_top);
  // receiver elements: the reference's search over the inner region (get_rx_element)
  std::vector<double> hxp(g.nnx), hyp(g.nny);
  std::vector<int32_t> rie(nrx > 0 ? nrx : 1), rje(nrx > 0 ? nrx : 1);
  if (nrx > 0) {
    MVF_HIP(h, hipMemcpy(hxp.data(), h->d_xp, 8 * g.nnx, hipMemcpyDeviceToHost));
    MVF_HIP(h, hipMemcpy(hyp.data(), h->d_yp, 8 * g.nny, hipMemcpyDeviceToHost));
    const int nd = h->bc.nextd;
    // y is searched only inside the x cell holding the site; an index not
    // found keeps the previous receiver's (rx_data's loop variables,
    // receiver_data.f90:121-124), and that element extrapolates the site
    int ie = -1, je = -1;
    for (int ir = 0; ir < nrx; ir++) {
      for (int a = nd + 1; a <= g.nx - nd; a++)
        if (rx[ir] >= hxp[(a - 1) * o] && rx[ir] < hxp[a * o]) {
          ie = a;
          for (int b = nd + 1; b <= g.ny - nd; b++)
            if (ry[ir] >= hyp[(b - 1) * o] && ry[ir] < hyp[b * o]) {
              je = b;
              break;
            }
        }
      if (ie < 0 || je < 0) {
        ds.drop(dx);
        return set_err(h, MVF_E_ARG,
                       "mvf_readout: receiver %d (%g, %g) lies outside the inner cells and no earlier receiver's "
                       "element is there to inherit (the reference's indices are undefined then)",
                       ir, rx[ir], ry[ir]);
      }
      rie[ir] = ie;
      rje[ir] = je;
    }
  }
  double *es = nullptr, *hs = nullptr;
  const size_t nf = (size_t)2 * g.npt * 3 * 2;
  int rc = dalloc(h, &es, nf);
  if (!rc) rc = dalloc(h, &hs, nf);
  if (!rc) rc = node_fields(h, x ? dx : h->d_x, ncol, col0, omega, es, hs);
  if (!rc && nrx > 0) {
    double *drx = nullptr, *dry = nullptr, *doe = nullptr, *doh = nullptr, *doz = nullptr, *drho = nullptr,
           *dphi = nullptr, *delev = nullptr;
    int32_t *die = nullptr, *dje = nullptr;
    rc = dalloc(h, &drx, nrx);
    if (!rc) rc = dalloc(h, &dry, nrx);
    if (!rc) rc = dalloc(h, &die, nrx);
    if (!rc) rc = dalloc(h, &dje, nrx);
    if (!rc) rc = dalloc(h, &doe, (size_t)nrx * 12);
    if (!rc) rc = dalloc(h, &doh, (size_t)nrx * 12);
    if (!rc) rc = dalloc(h, &doz, (size_t)nrx * 8);
    if (!rc) rc = dalloc(h, &drho, (size_t)nrx * 4);
    if (!rc) rc = dalloc(h, &dphi, (size_t)nrx * 4);
    if (!rc) rc = dalloc(h, &delev, (size_t)nrx);
    if (!rc) {
      hipError_t e = hipMemcpy(drx, rx, 8 * nrx, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(dry, ry, 8 * nrx, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(die, rie.data(), 4 * nrx, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(dje, rje.data(), 4 * nrx, hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = set_err(h, MVF_E_DEVICE, "mvf_readout: %s", hipGetErrorString(e));
    }
    if (!rc) rc = receivers(h, es, hs, ke_top, nrx, die, dje, drx, dry, omega, doe, doh, doz, drho, dphi, delev);
    if (!rc) {
      hipError_t e = hipStreamSynchronize(h->stream);
      if (e == hipSuccess && rx_e) e = hipMemcpy(rx_e, doe, 8 * (size_t)nrx * 12, hipMemcpyDeviceToHost);
      if (e == hipSuccess && rx_h) e = hipMemcpy(rx_h, doh, 8 * (size_t)nrx * 12, hipMemcpyDeviceToHost);
      if (e == hipSuccess && rx_z) e = hipMemcpy(rx_z, doz, 8 * (size_t)nrx * 8, hipMemcpyDeviceToHost);
      if (e == hipSuccess && rho) e = hipMemcpy(rho, drho, 8 * (size_t)nrx * 4, hipMemcpyDeviceToHost);
      if (e == hipSuccess && phi) e = hipMemcpy(phi, dphi, 8 * (size_t)nrx * 4, hipMemcpyDeviceToHost);
      if (e == hipSuccess && rx_elev) e = hipMemcpy(rx_elev, delev, 8 * (size_t)nrx, hipMemcpyDeviceToHost);
      if (e != hipSuccess) rc = set_err(h, MVF_E_DEVICE, "mvf_readout: %s", hipGetErrorString(e));
    }
    ds.drop(drx); ds.drop(dry); ds.drop(die); ds.drop(dje); ds.drop(doe); ds.drop(doh); ds.drop(doz); ds.drop(drho); ds.drop(dphi);
    ds.drop(delev);
  }
  if (!rc && (node_z || node_rho || node_phi)) {
    double *dz = nullptr, *drh = nullptr, *dph = nullptr;
    rc = dalloc(h, &dz, (size_t)g.npt * 8);
    if (!rc) rc = dalloc(h, &drh, (size_t)g.npt * 4);
    if (!rc) rc = dalloc(h, &dph, (size_t)g.npt * 4);
    if (!rc) rc = node_impedance(h, es, hs, omega, dz, drh, dph);
    if (!rc) {
      hipError_t e = hipStreamSynchronize(h->stream);
      if (e == hipSuccess && node_z) e = hipMemcpy(node_z, dz, 8 * (size_t)g.npt * 8, hipMemcpyDeviceToHost);
      if (e == hipSuccess && node_rho) e = hipMemcpy(node_rho, drh, 8 * (size_t)g.npt * 4, hipMemcpyDeviceToHost);
      if (e == hipSuccess && node_phi) e = hipMemcpy(node_phi, dph, 8 * (size_t)g.npt * 4, hipMemcpyDeviceToHost);
      if (e != hipSuccess) rc = set_err(h, MVF_E_DEVICE, "mvf_readout: %s", hipGetErrorString(e));
    }
    ds.drop(dz); ds.drop(drh); ds.drop(dph);
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && esol) e = hipMemcpy(esol, es, nf * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && hsol) e = hipMemcpy(hsol, hs, nf * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_err(h, MVF_E_DEVICE, "mvf_readout: %s", hipGetErrorString(e));
  }
  ds.drop(es);
  ds.drop(hs);
  ds.drop(dx);
  return rc;
}

int mvf_problem_fields(mvf_handle *h, double *ep, double *hp, double *src) {
  if (!h) return MVF_E_ARG;
  MVF_HIP(h, hipSetDevice(h->device));
  if (!h->have_freq) return set_err(h, MVF_E_ARG, "mvf_problem_fields: no current frequency");
  const size_t nf = (size_t)2 * h->geo.npt * 3 * 2;
  double *de = nullptr, *dh = nullptr, *ds = nullptr;
  int rc = dalloc(h, &de, nf);
  if (!rc) rc = dalloc(h, &dh, nf);
  if (!rc) rc = dalloc(h, &ds, nf);
  if (!rc) rc = problem_fields(h, h->omega, de, dh, ds);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && ep) e = hipMemcpy(ep, de, nf * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && hp) e = hipMemcpy(hp, dh, nf * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && src) e = hipMemcpy(src, ds, nf * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_err(h, MVF_E_DEVICE, "mvf_problem_fields: %s", hipGetErrorString(e));
  }
  dfree(de);
  dfree(dh);
  dfree(ds);
  return rc;
}

}  // extern "C"
