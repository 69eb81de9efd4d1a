/*
 * movfem_gpu.h -- C ABI of the MI355X element-integration, assembly and solve
 * path for MoVFEM_3DMT.  Plain pointers and sizes only; callable from Fortran
 * through ISO_C_BINDING (movfem_amd/fortran/movfem_gpu_mod.f90), from C/C++,
 * and from Python (ctypes).
 *
 * Layouts (all kept from the reference):
 *   node id       (i-1)*nny*nnz + (j-1)*nnz + k, z fastest      (n_fem.f90:81)
 *   element id    (ie-1)*ny*nz + (je-1)*nz + ke, z fastest     (MoVFEM_3DMT.f90:199)
 *   6-tensors     [11,12,13,22,23,33]                           (geometry.f90:1055-1060)
 *   complex       interleaved (re, im) doubles, Fortran complex(8) compatible
 *   COO indices   1-based (MUMPS convention)                    (MoVFEM_3DMT.f90:72-78)
 *   RHS/solution  column-major (nne, ndir), polarization d at offset (d-1)*nne
 *                                                               (global_assembly.f90:70-74)
 * Errors: every entry point returns 0 or a negative MVF_E_* code and never
 * aborts the process; mvf_last_error(h) holds the message (NULL: the calling
 * thread's last handle-less call, i.e. the file readers and mvf_grid3d; the reference
 * prints and stops: global_assembly.f90:55-57, n_fem.f90:374-377,
 * problem.f90:260-271).
 * Threading: one handle per host thread; calls on one handle are serialized
 * and stream-ordered on the handle's HIP stream.
 */
#ifndef MOVFEM_GPU_H
#define MOVFEM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MVF_OK = 0,
  MVF_E_ARG = -1,        /* invalid argument / size mismatch                       */
  MVF_E_DEVICE = -2,     /* HIP runtime error                                      */
  MVF_E_SINGULAR = -3,   /* zero Jacobian or singular sigma/mu tensor at a node   */
  MVF_E_INDEX = -4,      /* slot/index consistency failure                         */
  MVF_E_UNSUPPORTED = -5,/* option or combination the library does not provide     */
  MVF_E_NOCONV = -6,     /* Krylov did not reach the tolerance                     */
  MVF_E_BREAKDOWN = -7,  /* Krylov breakdown (rho or omega vanished)               */
  MVF_E_MEMORY = -8      /* device allocation failed                               */
};

/* Grid: the reference's node lattice (geometry.f90:17-26).  nx,ny,nz count
 * elements (the reference's g_nx-1, g_ny-1, g_nz-1). */
typedef struct mvf_grid {
  int32_t nx, ny, nz;
  int32_t nodes_per_elem; /* 8, 20 or 27 (nf_mn, n_fem.f90:30) */
  const double *xp;       /* g_xp(nnx) */
  const double *yp;       /* g_yp(nny) */
  const double *zp;       /* g_zp(npt), z measured from the grid bottom */
} mvf_grid;

/* Model: g_sigma complex(6,npt) and g_mu real(6,npt) as the reference holds
 * them after grid_3d (geometry.f90:1037-1085), i.e. with +i*float32(omega1*eps)
 * on sigma's diagonal for the first frequency. */
typedef struct mvf_model {
  const double *sigma; /* complex (6, npt) interleaved */
  const double *mu;    /* real    (6, npt) */
} mvf_model;

/* Boundary conditions (ga_init / init_gpml / bd_setmodel state,
 * global_assembly.f90:26-41, boundary_conds.f90:51-70,396-430). */
typedef struct mvf_bc {
  int32_t dirichlet; /* 1: Dirichlet edges removed; 0: GPML                     */
  int32_t inimod;    /* Dirichlet values: 1 zero, 2 half-space, 3 layered earth */
  int32_t gpml_sch;  /* 0 Fang, 1 Zhou (Re h = 1 under the reference semantics) */
  int32_t nextd;     /* extension cells; GPML zone = last nextd elements        */
  int32_t nzl_top;   /* g_nzl(g_nsf): top-zone layers, locates za(2)            */
  int32_t pad_;
  double a0, b0, nn; /* GPML constants                                          */
  /* boundary 1-D model of bd_setmodel, read for dirichlet with inimod 2/3 */
  double ztop;             /* g_ztop: surface height (geometry.f90:389-394)     */
  double h_sigma;          /* inimod 2: half-space conductivity                 */
  int32_t nl;              /* inimod 3: number of layers (>= 1)                 */
  int32_t pad2_;
  const double *l_sigma;   /* inimod 3: layer conductivities, top first (nl)    */
  const double *l_dz;      /* inimod 3: layer thicknesses in km (nl-1)          */
} mvf_bc;

/* Emulation switches (default: all on = the reference's effective semantics). */
enum {
  MVF_EMU_GPML_PREV_FLAGS = 1, /* GPML zone of element e taken from element e-1 (get_pml after
                                  int_elem_params, MoVFEM_3DMT.f90:203-206), carried between calls */
  MVF_EMU_UPDATE_SIGMA = 2,    /* update_sigma touches linear entries < npt+30 only (geometry.f90:146-152) */
  MVF_EMU_ALL = 3
};

typedef struct mvf_options {
  int32_t emulate;    /* MVF_EMU_* bit set */
  int32_t device;     /* HIP device ordinal */
  int32_t verbose;    /* 0 quiet */
  int32_t pad_;
} mvf_options;

typedef struct mvf_handle mvf_handle;

/* ---- setup (replaces ga_init + init_* of initialize_vfem,
 *      MoVFEM_3DMT.f90:354-388; global_assembly.f90:26-41) -------------------
 * Uploads lattice and model, numbers the DOFs (c_gne*, global_assembly.f90:
 * 183-475), builds the sparsity pattern and the element->entry gather lists
 * (ga_nzindx, :197-229), integrates the frequency-independent element
 * curl-curl (K) and mass (M) matrices and assembles them on the device. */
int mvf_init(const mvf_grid *grid, const mvf_model *model, const mvf_bc *bc,
             const mvf_options *opts, mvf_handle **out);
void mvf_free(mvf_handle *h);
const char *mvf_last_error(const mvf_handle *h);

/* sizes: nne (global DOFs, ga_cgne :194), nnze (full pattern entries, the
 * reference's slot count :228), nnz_triu (entries with row <= col) */
int mvf_sizes(const mvf_handle *h, int32_t *nne, int64_t *nnze, int64_t *nnz_triu);

/* DOF numbering gne(ne, me), element-major, negative = Dirichlet face code
 * (global_assembly.f90:17,266-285). */
int mvf_get_gne(mvf_handle *h, int32_t *gne_out);

/* ---- per frequency (update_omega, update_sigma, pset_pmodel, bd_updatemodel;
 *      MoVFEM_3DMT.f90:66-69, geometry.f90:137-153, problem.f90:247-252) ---- */
int mvf_set_frequency(mvf_handle *h, double freq_hz);

/* ---- element level (int_elem_params / alocal / blocal,
 *      integration.f90:60-106) ----------------------------------------------
 * For elements [e0, e1) (0-based, reference loop order): K and M upper-packed
 * (row-major i<=j, E(E+1)/2 doubles per element; A_ij = K_ij + i*fl32(omega)*M_ij
 * is alocal) and b (E x 2 complex per element, blocal) at the current
 * frequency.  Any output pointer may be NULL.  Host buffers. */
int mvf_integrate(mvf_handle *h, int64_t e0, int64_t e1, double *Ke, double *Me, double *be);

/* ---- drop-in for global_vfem(ia, ja, a, b) (MoVFEM_3DMT.f90:167-216) -------
 * Fills exactly what the reference's global_vfem leaves after ga_sort_sparse
 * and before zero removal: ia/ja 1-based row-major over the full pattern,
 * a the complex64-rounded values in the reference's column-sorted order, b the
 * 2*nne RHS.  nnze/nne/ndir must match mvf_sizes; ndir must be 2.  The
 * handle keeps two pinned host staging buffers (up to 2^22 entries, 96 MiB,
 * each) from the first call of either global_vfem form until mvf_free. */
int mvf_global_vfem(mvf_handle *h, int32_t *ia, int32_t *ja, double *a, double *b, int64_t nnze,
                    int32_t nne, int32_t ndir);

/* global_vfem followed by find_zeros/rem_zeros (MoVFEM_3DMT.f90:82-97,
 * global_assembly.f90:123-150) in one call: the zero-removed triplet the
 * reference hands MUMPS (ia, ja, a in mvf_global_vfem's order with every
 * entry equal to (0, 0) dropped), formatted on the device, so that only the
 * nonzero entries cross PCIe.  cap = capacity of ia/ja/a in entries (nnz_triu
 * of mvf_sizes always suffices: every entry below the diagonal is zero);
 * *nnz_out = entries written; b as mvf_global_vfem.  Counts as one
 * global_vfem call.  Fewer than the nonzero entries in cap: MVF_E_ARG. */
int mvf_global_vfem_nz(mvf_handle *h, int32_t *ia, int32_t *ja, double *a, double *b, int64_t cap, int32_t nne,
                       int32_t ndir, int64_t *nnz_out);

/* find_zeros + rem_zeros (global_assembly.f90:123-150): in-place compaction of
 * a COO triplet, returns the new length in *n_out. Host only. */
int mvf_remove_zeros(int64_t n, int32_t *ia, int32_t *ja, double *a, int64_t *n_out);

/* ---- solve (replaces zmumps(job=6), MoVFEM_3DMT.f90:113-115) --------------- */
enum { MVF_KRYLOV_COCG = 0, MVF_KRYLOV_BICGSTAB = 1, MVF_KRYLOV_GMRES = 2 };
/* MIXED: FP32 inner Krylov, FP64 refinement; with MVF_PC_DIRECT: the factor's
 * entries rounded to complex64 after every update (an emulated complex64
 * factorization, FP64 arithmetic between the roundings), FP64 GMRES outside */
enum { MVF_PREC_FP64 = 0, MVF_PREC_MIXED = 1 };
enum { MVF_MATRIX_C64REF = 0, MVF_MATRIX_FP64 = 1 };/* which operator the residual is judged on */
/* MVF_PC_DIRECT: multifrontal LU (FP64, nested dissection of the element box)
 * of the operator at the geometric mean of the batch's frequencies, with a small imaginary
 * shift on the diagonal of massless (all-air) DOFs; the outer method is then
 * right-preconditioned GMRES on the unshifted operator whatever `method` says.
 * It stands where zmumps(job=6) stands (MoVFEM_3DMT.f90:113-115). */
enum { MVF_PC_NONE = 0, MVF_PC_JACOBI = 1, MVF_PC_DIRECT = 2 };

typedef struct mvf_solver_opts {
  int32_t method;       /* MVF_KRYLOV_* */
  int32_t precision;    /* MVF_PREC_* */
  int32_t matrix;       /* MVF_MATRIX_*: C64REF = the complex64-rounded matrix the reference hands MUMPS */
  int32_t precond;      /* MVF_PC_* */
  double tol;           /* FP64 true relative residual ||b-Ax||/||b|| per polarization */
  int32_t maxit;        /* total Krylov iterations */
  int32_t check_every;  /* iterations between convergence checks (host sync) */
  double inner_tol;     /* MIXED: FP32 residual drop before an FP64 refinement step */
  int32_t use_graph;    /* COCG/BiCGStab: 1 = capture check_every iterations once as a HIP graph and replay
                           it between convergence checks; COCG: 2 = run each check_every iterations in one
                           cooperative launch of a persistent kernel (grid barriers between the steps);
                           MVF_PC_DIRECT: nonzero = replay each application of the factor
                           (its triangular sweeps) from a graph captured once per column count.  Ignored by
                           mvf_solve_noalloc, and by mvf_solve_batch with more than one thread */
  int32_t gmres_m;      /* GMRES restart length (MVF_PC_DIRECT), 0 = 30 */
  int32_t freqs_per_factor; /* MVF_PC_DIRECT in mvf_solve_frequencies: consecutive frequencies sharing one
                               factorization (at their geometric-mean frequency), 0 or 1 = one each */
  int32_t direct_tile;  /* MVF_PC_DIRECT: GEMM tiles of the factorization, 0 = by front size, 64 = 64 x 64
                           (four real products per complex one), 128 = 128 x 64 (three) forced */
} mvf_solver_opts;

typedef struct mvf_solve_info {
  int32_t iterations;      /* Krylov iterations performed (max over columns) */
  int32_t refinements;     /* FP64 refinement steps (MIXED) */
  int32_t converged;       /* 1 if every column met tol */
  int32_t fallbacks;       /* restarts after breakdown */
  double true_resid[64];   /* final FP64 ||b-Ax||/||b|| per column (column = freq*2 + pol) */
  double seconds;          /* wall time of the solve on the device stream */
  /* SpMV kernels timed with HIP events on the handle's stream
   * [0] FP64 operator, [1] FP32 operator (MIXED inner iterations) */
  int64_t spmv_launches[2];
  double spmv_ms[2];       /* summed kernel time */
  double spmv_bytes[2];    /* algorithmic bytes of one launch: nnz*(K+M+col) + nne*(rowptr + ncol*(x+y)) */
  int32_t spmv_ncol;       /* columns per launch (last batch) */
  int32_t spmv_lpr;        /* lanes per matrix row of the FP64 SpMV, chosen by the autotuner */
  /* MVF_PC_DIRECT */
  double factor_seconds;   /* numeric factorizations performed by this call, wall time */
  double factor_flops;     /* algorithmic real flops of one factorization */
  double factor_bytes;     /* bytes of the stored factor */
  int32_t factorizations;  /* factorizations performed by the handle so far */
  int32_t static_pivots;   /* pivots replaced because the owned block was singular */
  int32_t precond_applies; /* triangular-solve sweeps performed by this call */
  int32_t operator_fp32;   /* 1: K and M in use were integrated in FP32 (mvf_assemble precision 1);
                              residuals are then judged on that operator */
  /* batched complex GEMM launches of the last factorization (its dominant
   * kernel): summed HIP-event time, algorithmic real flops, launch count */
  double gemm_ms;
  double gemm_flops;
  int64_t gemm_launches;
  int64_t graph_launches;  /* use_graph: replays of the captured block of check_every Krylov iterations */
} mvf_solve_info;

void mvf_default_solver_opts(mvf_solver_opts *o);

/* zmumps(job=6)-shaped: rhs_inout is the host (nne, 2) complex RHS as
 * global_vfem filled it; it is overwritten by the solution, as MUMPS does. */
int mvf_solve(mvf_handle *h, double *rhs_inout, const mvf_solver_opts *o, mvf_solve_info *info);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  mvf_solve for method COCG or BiCGStab in FP64 with the
 * Jacobi or no preconditioner, the same kernels, with every buffer (Krylov
 * vectors, reduction partials, the staging of rhs_inout) in one static device
 * array of 64 MiB per GPU instead of allocations: COCG up to about 299,000
 * DOFs, BiCGStab up to about 174,000; larger systems are MVF_E_ARG, the
 * direct preconditioner, GMRES and MIXED are MVF_E_UNSUPPORTED.  It calls no
 * allocating routine.  All calls on one GPU share that array, so calls that
 * overlap on one GPU race: the caller serializes them.  rhs_inout as in
 * mvf_solve (overwritten by the solution). */
int mvf_solve_noalloc(mvf_handle *h, double *rhs_inout, const mvf_solver_opts *o, mvf_solve_info *info);

/* Device-resident path: RHS integrated on the device for each of nfreq
 * frequencies (processed in that order, as the reference's frequency loop
 * would), solved in batches of up to 8 frequencies (16 columns) per batched
 * Krylov sweep sharing K and M.  Solutions stay on the device; fetch with
 * mvf_get_solution.  1 <= nfreq <= 32 (true_resid holds 64 columns). */
int mvf_solve_frequencies(mvf_handle *h, const double *freqs_hz, int32_t nfreq, const mvf_solver_opts *o,
                          mvf_solve_info *info);
/* Many models (handles, of any size and element type) x nfreq frequencies
 * each: mvf_solve_frequencies on every handle, nthreads host workers (0: 8)
 * keeping that many models' work in flight at once, each on its handle's own
 * stream, so small models' kernels overlap on the device.  freqs_hz is
 * (nh, nfreq) row-major; infos and status (each nh entries, may be NULL)
 * receive each model's info and return code; the result is the first
 * non-zero code.  No counterpart in the reference, which runs one model
 * through its frequency loop (MoVFEM_3DMT.f90:62-143). */
int mvf_solve_batch(mvf_handle *const *hs, int32_t nh, const double *freqs_hz, int32_t nfreq,
                    const mvf_solver_opts *o, mvf_solve_info *infos, int32_t *status, int32_t nthreads);
/* column-major (nne, 2*nfreq) complex solution of the last solve_frequencies */
int mvf_get_solution(mvf_handle *h, double *x_out, int32_t ncols);
/* device RHS (nne, 2*nfreq) of the last solve_frequencies, for checks */
int mvf_get_rhs(mvf_handle *h, double *b_out, int32_t ncols);

/* Timing of the last setup phases, seconds: [0] numbering, [1] pattern,
 * [2] element integration (K, M), [3] assembly. */
int mvf_setup_times(const mvf_handle *h, double *t4);

/* The assembled operators as the device holds them: the full symmetric CSR
 * pattern (0-based rowptr[nne+1] and col[nnze], rows in DOF order, columns
 * ascending) and the FP64 values of K (curl-curl) and M (mass) at the current
 * call state, A(omega) = K + i*fl32(omega)*M before the complex64 rounding of
 * ga_sort_sparse (global_assembly.f90:152-181).  Any pointer may be NULL.
 * No counterpart in the reference (whose matrix exists only as the COO
 * global_vfem returns): for host-side residual checks. */
int mvf_get_operators(mvf_handle *h, int64_t *rowptr, int32_t *col, double *K, double *M);
/* Rows [r0, r1) of the same: rowptr[r1 - r0 + 1] relative to the first row's
 * start, then that many col, K, M entries (any of col, K, M may be NULL:
 * a call with all three NULL returns the row lengths to size the buffers). */
int mvf_get_rows(mvf_handle *h, int64_t r0, int64_t r1, int64_t *rowptr, int32_t *col, double *K, double *M);

/* Re-integrates every element's K and M from the resident lattice and model
 * and re-assembles them in place (the element loop of global_vfem,
 * MoVFEM_3DMT.f90:193-211, without any host transfer).  precision 0: FP64
 * integration; 1: FP32 integration (K and M rounded from FP32 element sums).
 * The handle records which: solves report it in mvf_solve_info.operator_fp32,
 * and mvf_global_vfem (the reference's FP64 matrix) refuses an FP32-integrated
 * operator with MVF_E_ARG until an FP64 re-assembly.
 * integrate_ms (optional) receives the element kernel's time from HIP events. */
int mvf_assemble(mvf_handle *h, int32_t precision, double *integrate_ms);

/* ---- readout (node_solution, solution.f90:18-49; rx_data,
 *      receiver_data.f90:111-133; get_impedance/get_res_phase, :456-505) -------
 * x: host solution (nne, 2) complex column-major at the current frequency, as
 * mvf_solve returns it; or NULL for frequency f (0-based) of the last
 * mvf_solve_frequencies, read from the device.
 * esol/hsol (optional): (2, npt, 3) complex, total E and H at every lattice
 * node per polarization in node-id order; the first element in the
 * reference's loop order that contains a node sets it (solution.f90:230-238).
 * Receivers (nrx >= 0 at rx/ry): fields interpolated on the top face of
 * element layer ke_top (1-based: the reference's top earth layer) inside the
 * region ie, je in [nextd+1, n-nextd] (get_rx_element); rx_e/rx_h (nrx, 2, 3)
 * complex, rx_z (nrx, 4) complex [Zxx Zxy Zyx Zyy], rho/phi (nrx, 4), degrees.
 * A receiver outside those cells keeps the previous receiver's element index
 * along each axis not found (y is searched only when x is found; the
 * reference's loop variables, receiver_data.f90:121-124,186-209) and is
 * extrapolated from that element; the first receiver outside them is
 * MVF_E_ARG (the reference's indices are undefined there).  rx_elev (nrx): the receiver's
 * height on that face (interpolate_rx overwrites g_zsr with it,
 * receiver_data.f90:222-225).  node_z (npt, 4) complex, node_rho/node_phi
 * (npt, 4): Z, rho_a and phase at every lattice node (z_rho_phi,
 * solution.f90:51-69).  Outputs may be NULL. */
int mvf_readout(mvf_handle *h, const double *x, int32_t f, int32_t ke_top, int32_t nrx, const double *rx,
                const double *ry, double *esol, double *hsol, double *rx_e, double *rx_h, double *rx_z, double *rho,
                double *phi, double *rx_elev, double *node_z, double *node_rho, double *node_phi);

/* ---- text writers (host only; no device use) ------------------------------
 * The reference's output files, one frequency block per call: nf is the
 * 1-based frequency index printed in the block headers; append = 0 creates
 * the file and writes the file header the reference writes for nf = 1,
 * append = 1 adds a block.  Column formats are the reference's edit
 * descriptors (es20.5, f20.5, f25.6, A15, i5); its list-directed lines
 * (write(u,*)) follow the reference build's list-directed form: a leading
 * blank, one blank before every item, records wrapped before 80 columns,
 * integers in minimal width and reals as the shortest decimal that reads
 * back exactly.  Field arrays use mvf_readout's layouts.  Returns 0 or
 * MVF_E_ARG (bad arguments or an I/O error). */

/* SOLUTION.OUT (write_solution, solution.f90:73-205): nodal E, H (24es20.5
 * per node), Z (8es20.5), rho_a/phase (8f20.5; 8f25.6 when a node selection
 * is given, the reference's inner-region form).  sel: nsel 1-based node ids
 * (the reference's g_nb), or NULL for every node in id order. */
int mvf_write_solution(const char *path, int32_t nf, int32_t append, int64_t npt, const double *esol,
                       const double *hsol, const double *node_z, const double *node_rho, const double *node_phi,
                       int64_t nsel, const int32_t *sel);

/* RECEIVER.OUT (write_receiver_data, receiver_data.f90:25-67): on create the
 * receiver count, polarizations (2), frequency count nfreq and the receiver
 * coordinates (x, y, elev); per block the fields (26es20.5), Z (10es20.5) and
 * rho_a/phase (10f20.5) with the receiver coordinates in front. */
int mvf_write_receivers(const char *path, int32_t nf, int32_t append, int32_t nfreq, int32_t nrx, const double *rx,
                        const double *ry, const double *elev, const double *rx_e, const double *rx_h,
                        const double *rx_z, const double *rho, const double *phi);

/* SURFACE.OUT (write_surface_data, receiver_data.f90:69-109): nsp = nxb*nyb
 * surface nodes ids (0-based lattice ids, x fastest within y as sur_data
 * walks them) with coordinates xs (nxb), ys (nyb), zs (nsp); on create the
 * counts and coordinates, per block their E, H (24es20.5), Z (8es20.5) and
 * rho_a/phase (8f20.5) taken from the nodal arrays (npt nodes). */
int mvf_write_surface(const char *path, int32_t nf, int32_t append, int32_t nfreq, int32_t nxb, int32_t nyb,
                      const double *xs, const double *ys, const double *zs, const int64_t *ids, int64_t npt,
                      const double *esol, const double *hsol, const double *node_z, const double *node_rho,
                      const double *node_phi);

/* PROBLEM.OUT (write_problem_ugrid, problem.f90:188-240), one frequency:
 * for polarizations 1 and 2 the header (node count, nf), the column line and
 * per node the list-directed record psigma, ep(3), hp(3), src(3), complex
 * values as (re,im).  psigma is the real part of pe_psigma(1) = (0, omega*eps)
 * (pset_pmodel, problem.f90:247-252): 0 at every node.  ep, hp, src:
 * mvf_problem_fields' arrays (npt nodes); sel: nsel 1-based node ids (the
 * reference's g_nb for inner = 1) or NULL for every node. */
int mvf_write_problem(const char *path, int32_t nf, int32_t append, int64_t npt, const double *ep, const double *hp,
                      const double *src, int64_t nsel, const int32_t *sel);

/* The fields PROBLEM.OUT prints (set_write_problem / write_problem_elem,
 * problem.f90:459-507) at the current frequency (mvf_set_frequency): at every
 * lattice node the interpolated primary E and H (get_ep / get_hp, :155-186)
 * and the source term p_source (:91-128), each complex (2, npt, 3)
 * polarization-major, taken from the last element (reference order) holding
 * the node.  Computed on the device; host output buffers, any may be NULL. */
int mvf_problem_fields(mvf_handle *h, double *ep, double *hp, double *src);

/* ---- input files and lattice generation (host only; no device use) --------
 * Failures return MVF_E_ARG; mvf_last_error(NULL) holds the message of the
 * calling thread's last failing call of this group.  Inputs on which the
 * reference's own routines run into undefined values are rejected here: a
 * Shepard fit that fails (fewer than 20 points on an interface, duplicate or
 * collinear points; the reference ignores qshep2's ier, geometry.f90:458-459).
 * For a domain-length / skin-depth ratio in [1.5, 2.5) the reference leaves
 * its extension depth factor unset (geometry.f90:245-251); its build takes 3
 * there, and so does this library. */

/* grid_3d's arguments (geometry.f90:48-62) in the reference's array layouts,
 * so that the arrays read by input() (read_input.f90:14-20) pass unchanged.
 * Interfaces occupy slots 2..nsf-2 of nsf = (input interfaces) + 3; slots 1,
 * nsf-1 and nsf (bottom, air and top) are generated. */
typedef struct mvf_mesh_input {
  double dx, dy, dz;      /* inner cell sizes (in_dx, in_dy, in_dz)                    */
  double high, depth;     /* air above / earth below the interface points (in_high...) */
  double sigma_bg;        /* conductivity for the skin depth (in_bcksigma)              */
  int32_t mn;             /* nodes per element: 8, 20 or 27                             */
  int32_t nsr;            /* source/receiver sites                                      */
  int32_t nsf;            /* interface slots (in_nsf)                                   */
  int32_t mx, my, mz;     /* inner model lattice                                        */
  int32_t isigma, imu;    /* conductivity / permeability components; imu = 0: mu_r = 1 */
  int32_t nf;             /* frequencies                                                */
  int32_t pad_;
  const double *xsr, *ysr, *zsr;   /* (nsr)                                             */
  const int32_t *nsp;              /* (nsf) points per slot; slots 2..nsf-2 are read     */
  const double *xto, *yto, *zto;   /* (nsf, *) column-major: point j of slot l at        */
                                   /* (j-1)*nsf + (l-1)                                  */
  const double *xm, *ym;           /* (mx), (my)                                        */
  const double *zm;                /* (mx*my*mz): point (i,j,k) at (i-1)*my*mz+(j-1)*mz+k-1 */
  const int32_t *ijsigma;          /* (isigma, 2) column-major tensor index pairs       */
  const double *sigma;             /* (isigma, mx*my*mz) column-major                   */
  const int32_t *ijmu;             /* (imu, 2) column-major, NULL when imu = 0          */
  const double *mu;                /* (imu, mx*my*mz) column-major, NULL when imu = 0   */
  const double *freqs;             /* (nf) Hz; freqs[0] gives sigma's i*omega*eps term  */
} mvf_mesh_input;

/* PARAM.INP (MoVFEM_3DMT.f90:302-334).  gpml_sch, a0, b0, nn are read only
 * for boundary = 0 (GPML) and are 0 otherwise. */
typedef struct mvf_params {
  int32_t nf, boundary, inimod, nl, gpml_sch, pad_;
  double h_sigma, a0, b0, nn;
  const double *freqs;   /* (nf) Hz                    */
  const double *l_sigma; /* (nl)                       */
  const double *l_dz;    /* (nl-1)                     */
} mvf_params;

typedef struct mvf_mesh mvf_mesh;
typedef struct mvf_inputs mvf_inputs;

/* Replaces input() (read_input.f90:37-43) and the PARAM.INP reads of
 * initialize_vfem (MoVFEM_3DMT.f90:302-334): either path may be NULL.
 * Fortran list-directed rules: blanks, tabs or commas separate values, r*c
 * repeats, a READ continues over records until its list is filled and
 * leaves the rest of its last record unread, an empty list consumes one
 * record; the conductivity/permeability names are read as (a5,i1,i1) and
 * (a2,i1,i1). */
int mvf_read_inputs(const char *inp_path, const char *param_path, mvf_inputs **out);
/* Pointers into the handle (valid until mvf_inputs_free); nf and freqs of
 * mesh come from PARAM.INP when it was read. */
int mvf_inputs_view(const mvf_inputs *h, mvf_mesh_input *mesh, mvf_params *params);
void mvf_inputs_free(mvf_inputs *h);

/* Replaces grid_3d (geometry.f90:48-122; iout = 0): the node lattice, the
 * model on it and the layer counts, equal to the reference's bit for bit. */
int mvf_grid3d(const mvf_mesh_input *in, mvf_mesh **out);
/* dims[12] = nx, ny, nz (elements: g_nx-1, g_ny-1, g_nz-1), nodes_per_elem,
 * nnx, nny, nnz, npt, nextd, nsf, nsr, nodes per element edge (2 or 3);
 * vals[6] = g_ztop, x0, y0, z0 (the shift applied to every coordinate),
 * omega of freqs[0], skin depth at the lowest frequency.  Either may be NULL. */
int mvf_mesh_sizes(const mvf_mesh *m, int64_t *dims, double *vals);
/* Copies g_xp (nnx), g_yp (nny), g_zp (npt), g_sigma (complex (6, npt)),
 * g_mu ((6, npt)), g_nzl (nsf) and the shifted site coordinates g_xsr,
 * g_ysr, g_zsr (nsr); any pointer may be NULL.  The lattice, sigma and mu
 * feed mvf_grid / mvf_model directly. */
int mvf_mesh_arrays(const mvf_mesh *m, double *xp, double *yp, double *zp, double *sigma, double *mu, int32_t *nzl,
                    double *xsr, double *ysr, double *zsr);
void mvf_mesh_free(mvf_mesh *m);

/* ---- learned initial guess (no counterpart in the reference) --------------
 * A 3-D U-Net (four levels, 16/32/64/96 channels, 3x3x3 convolutions with
 * leaky ReLU, 2x2x2 average pooling and nearest upsampling, skip
 * connections, a 1x1x1 head; 1.16M parameters) trained and evaluated on the
 * BF16 matrix cores with FP32 accumulation and FP32 master weights (Adam).
 * Inputs and targets are channels-last FP32 host arrays (batch, nx, ny, nz, c).
 * It predicts the solution of a frequency-model system from per-cell
 * features; mvf_solve_guess starts the Krylov solve from that prediction,
 * the FP64 true residual judging the result, and plain zero-start Krylov is
 * what remains when the guess does not help (the least-squares scaling never
 * leaves a larger initial residual than x0 = 0). */
typedef struct mvf_nn mvf_nn;
typedef struct mvf_nn_config {
  int32_t nx, ny, nz;   /* voxel grid, each a multiple of 8                      */
  int32_t cin, cout;    /* real input / output channels, 1..16                    */
  int32_t width;        /* channels of the finest level (16)                      */
  int32_t batch;        /* samples per forward/backward                          */
  int32_t device;
  uint64_t seed;        /* weight initialisation                                  */
  float lr, beta1, beta2, eps; /* Adam                                           */
  float leaky;          /* negative slope of the activation                       */
  int32_t pad_;
} mvf_nn_config;
void mvf_nn_default_config(mvf_nn_config *c);
int mvf_nn_create(const mvf_nn_config *cfg, mvf_nn **out);
void mvf_nn_free(mvf_nn *n);
const char *mvf_nn_last_error(const mvf_nn *n);
/* parameters and gradients: per layer (encoder levels 0..3, convolutions 1
 * and 2; decoder levels 0..2, convolutions 1 and 2; head) the weights
 * [tap][cout][cin] (tap = 9 dx + 3 dy + dz over offsets -1..1) then the bias
 * [cout]; the first layer's cin and the head's cout are padded to 16 */
int64_t mvf_nn_num_params(const mvf_nn *n);
int mvf_nn_get_params(mvf_nn *n, float *out);
int mvf_nn_set_params(mvf_nn *n, const float *in);
int mvf_nn_get_grads(mvf_nn *n, float *out);
/* loss = sum m (y - t)^2 / sum m over the batch (m = 1 when NULL) and the
 * gradients of every parameter; nb must equal cfg.batch */
int mvf_nn_forward_backward(mvf_nn *n, int32_t nb, const float *x, const float *t, const float *m, float *loss);
int mvf_nn_adam(mvf_nn *n, float lr);
int mvf_nn_predict(mvf_nn *n, int32_t nb, const float *x, float *y);
/* HIP-event timing of the matrix-core kernels of the next calls: s3 = summed
 * kernel ms, algorithmic flops (2 per multiply-add) and launches of the last
 * forward(-backward) */
int mvf_nn_set_timing(mvf_nn *n, int32_t on);
int mvf_nn_stats(const mvf_nn *n, double *s3);

/* mvf_solve starting from x0 (host, (nne, 2) complex like rhs_inout) instead
 * of zero: each column of x0 is first scaled by the complex factor alpha that
 * minimises ||b - alpha A x0|| (returned in alpha_out[4] as re, im per
 * polarization; NULL allowed), so the initial residual is never above ||b||. */
int mvf_solve_guess(mvf_handle *h, const double *x0, double *rhs_inout, const mvf_solver_opts *o,
                    mvf_solve_info *info, double *alpha_out);

/* Host-only symbolic analysis of the sparse direct factorization for a
 * numbering (gne, ne x me, 1-based, <= 0 removed) and the full CSR pattern
 * (0-based rowptr[nne+1], col): no device use.  stats[10] = fronts, tree
 * depth, max owned DOFs, max update DOFs, max front, factor entries, real
 * flops of one factorization, sum of owned DOFs, seconds, update entries.
 * Diagnostic; no counterpart in the reference (whose ordering is MUMPS's). */
int mvf_direct_plan_stats(int32_t nx, int32_t ny, int32_t nz, int32_t me, const int32_t *gne, int64_t nne,
                          const int64_t *rowptr, const int32_t *col, double *stats);

#ifdef __cplusplus
}
#endif
#endif /* MOVFEM_GPU_H */
