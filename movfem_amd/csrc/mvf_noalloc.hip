// mvf_solve_noalloc: mvf_solve for the Krylov methods with no device
// allocation.  The Krylov buffers, the staging of the RHS and the solution
// live in one file-scope device array, `scratch`, of kScratchBytes per GPU.
//
// `scratch` is a device symbol: in host code its name is only a stand-in, so
// host code never passes it to a kernel or a copy.  noalloc_setup() looks up
// its device address once per GPU (with that GPU current) and keeps it in
// g_scratch; everything below uses that address.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents: the solution buffer is cleared by a kernel before each solve, and
// calls that overlap on one GPU (two host threads, two handles on one device)
// race on it.  The caller serializes them.
#include <mutex>

namespace mvf {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static char *g_scratch[kScratchDevices] = {};

// once per device, with h->device current (called from mvf_init)
int noalloc_setup(mvf_handle *h) {
  if (h->device < 0 || h->device >= kScratchDevices)
    return set_err(h, MVF_E_ARG, "device %d beyond the %d devices the scratch table holds", h->device,
                   kScratchDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[h->device]) return 0;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_err(h, MVF_E_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  g_scratch[h->device] = (char *)p;
  return 0;
}

static char *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kScratchDevices) ? g_scratch[device] : nullptr;
}

// column-major host layout (nne, 2) -> [nne][2] interleaved columns, and back
__global__ void k_cols_to_rows(int64_t n, const double2 *cm, double2 *rows) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    rows[2 * i] = cm[i];
    rows[2 * i + 1] = cm[n + i];
  }
}

__global__ void k_rows_to_cols(int64_t n, const double2 *rows, double2 *cm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    cm[i] = rows[2 * i];
    cm[n + i] = rows[2 * i + 1];
  }
}

__global__ void k_zero2(int64_t n, double2 *x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = make_double2(0.0, 0.0);
}

}  // namespace mvf

extern "C" int mvf_solve_noalloc(mvf_handle *h, double *rhs_inout, const mvf_solver_opts *o, mvf_solve_info *info) {
  using namespace mvf;
  if (!h || !rhs_inout) return MVF_E_ARG;
  if (!h->have_freq) return set_err(h, MVF_E_ARG, "mvf_set_frequency must precede mvf_solve_noalloc");
  mvf_solver_opts od;
  mvf_default_solver_opts(&od);
  if (!o) o = &od;
  if (o->precond == MVF_PC_DIRECT || o->precision != MVF_PREC_FP64 || o->method == MVF_KRYLOV_GMRES)
    return set_err(h, MVF_E_UNSUPPORTED,
                   "mvf_solve_noalloc runs FP64 COCG or BiCGStab with the Jacobi or no preconditioner");
  mvf_solve_info dummy;
  if (!info) info = &dummy;
  memset(info, 0, sizeof(*info));
  MVF_HIP(h, hipSetDevice(h->device));
  char *base = scratch_of(h->device);
  if (!base) return set_err(h, MVF_E_DEVICE, "no scratch address for device %d", h->device);
  const int64_t n = h->nne;
  // staging: host-layout RHS/solution, then the interleaved b and x
  const size_t vec = (((size_t)n * 2 * sizeof(double2)) + 255) & ~(size_t)255;
  // the lanes per row the handle's autotuner chose for two columns, if it ran
  Solver S{h, 2, {}, o->matrix == MVF_MATRIX_C64REF, h->spmv_lpr[0] ? h->spmv_lpr[0] : 16, n, nullptr};
  S.w.w[0] = S.w.w[1] = (double)(float)h->omega;
  Workspace W;
  W.sizing = true;
  if (o->method == MVF_KRYLOV_BICGSTAB) {
    BicgBufs<2> B;
    MVF_TRY(bicg_bufs<2>(W, h, n, B));
  } else {
    CocgBufs<2> B;
    MVF_TRY(cocg_bufs<2>(W, h, n, B));
  }
  const size_t need = 3 * vec + W.used;
  if (need > kScratchBytes)
    return set_err(h, MVF_E_ARG,
                   "mvf_solve_noalloc: %lld DOFs need %zu bytes of scratch, beyond its %zu (COCG fits about %lld "
                   "DOFs, BiCGStab about %lld)",
                   (long long)n, need, kScratchBytes, (long long)(kScratchBytes / 224),
                   (long long)(kScratchBytes / 384));
  double2 *stage = (double2 *)base, *db = (double2 *)(base + vec), *dx = (double2 *)(base + 2 * vec);
  S.ws = base + 3 * vec;
  S.ws_cap = kScratchBytes - 3 * vec;
  const unsigned gb = (unsigned)grid1d(n, 256);
  MVF_HIP(h, hipMemcpyAsync(stage, rhs_inout, sizeof(double2) * 2 * (size_t)n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_cols_to_rows, dim3(gb), dim3(256), 0, h->stream, n, (const double2 *)stage, db);
  hipLaunchKernelGGL(k_zero2, dim3(grid1d(2 * n, 256)), dim3(256), 0, h->stream, 2 * n, dx);
  MVF_HIP(h, hipGetLastError());
  double t0 = now();
  int rc = o->method == MVF_KRYLOV_BICGSTAB ? run_bicgstab<2>(S, o, (const double *)db, (double *)dx, info)
                                            : run_cocg64<2>(S, o, (const double *)db, (double *)dx, info);
  hipLaunchKernelGGL(k_rows_to_cols, dim3(gb), dim3(256), 0, h->stream, n, (const double2 *)dx, stage);
  if (hipMemcpyAsync(rhs_inout, stage, sizeof(double2) * 2 * (size_t)n, hipMemcpyDeviceToHost, h->stream) !=
          hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess) {
    if (!rc) rc = set_err(h, MVF_E_DEVICE, "copying the solution back failed");
  }
  info->seconds = now() - t0;
  info->operator_fp32 = h->ops_fp32;
  info->spmv_lpr = S.lpr;
  return rc;
}
