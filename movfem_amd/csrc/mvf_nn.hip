// Learned initial guess: a 3-D U-Net on the BF16 matrix cores.
//
// No counterpart in the reference, which factors every system from scratch
// with MUMPS (MoVFEM_3DMT.f90:113-115).  The network maps per-cell inputs
// (log10 sigma, air mask, log10 f) on a voxel grid to the edge unknowns of the
// 8-node system, and its prediction, scaled by the least-squares factor that
// minimises ||b - A x0||, starts the Krylov solve (mvf_solve_guess).
//
// Tensors are channels-last [batch][x][y][z][c] with c innermost, activations
// and weights in BF16, every product accumulated in FP32 on
// v_mfma_f32_16x16x16bf16_1k.  Lane layouts (16 x 16 x 16): A lane l ->
// A[l%16][4(l/16)+j], B lane l -> B[4(l/16)+j][l%16] (j = 0..3, four BF16 in
// one 64-bit register pair), D register v of lane l -> D[4(l/16)+v][l%16].
//
//   k_conv   implicit-GEMM convolution (3x3x3 or 1x1x1, stride 1, zero
//            padding): Y[v][n] = sum_tap sum_k X[v+off(tap)][k] W[tap][n][k];
//            the data gradient is the same product with the weights flipped
//            and transposed (Wt[tap'][ci][co] = W[26-tap'][co][ci]).
//   k_wgrad  weight gradient dW[tap][co][ci] = sum_v dY[v][co] X[v+off][ci],
//            the voxel sum split over workgroups and added with FP32 atomics.
//   element-wise kernels: input packing, 2x2x2 average pooling and nearest
//   upsampling with their adjoints, the masked MSE loss and Adam.
#include <cstring>
#include <random>

namespace mvf {
namespace nn {

typedef uint16_t bf16;
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__host__ __device__ inline float bf2f(bf16 b) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = (uint32_t)b << 16;
  return c.f;
}

// round to nearest even (finite inputs)
__host__ __device__ inline bf16 f2bf(float f) {
  union {
    uint32_t u;
    float f;
  } c;
  c.f = f;
  uint32_t u = c.u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

struct Dims {
  int nb, nx, ny, nz;  // batch, voxel grid
  __host__ __device__ int64_t vox() const { return (int64_t)nb * nx * ny * nz; }
};

// voxel v -> (b, x, y, z), z fastest
__device__ inline void vsplit(const Dims &d, int64_t v, int &b, int &x, int &y, int &z) {
  z = (int)(v % d.nz);
  int64_t t = v / d.nz;
  y = (int)(t % d.ny);
  t /= d.ny;
  x = (int)(t % d.nx);
  b = (int)(t / d.nx);
}

__device__ inline int64_t vjoin(const Dims &d, int b, int x, int y, int z) {
  return (((int64_t)b * d.nx + x) * d.ny + y) * d.nz + z;
}

__device__ inline s4v load4(const bf16 *p) { return *(const s4v *)p; }

enum { EPI_ACT = 0, EPI_F32 = 1, EPI_MASK = 2, EPI_RAW = 3 };

// Y = conv(X, W) over K = taps * cin; each wave computes 32 voxels (two row
// tiles) x NT*16 output channels, four waves per workgroup.
//   EPI_ACT:  Y bf16 = leaky(acc + bias)
//   EPI_F32:  Y fp32 = acc + bias
//   EPI_MASK: Y bf16 = acc * leaky'(mask tensor)   (data gradient inside a level)
//   EPI_RAW:  Y bf16 = acc                          (data gradient at a level boundary)
// X: channel stride cx (input may be a slice of a wider tensor); Y: channel
// stride cy; mask: channel stride cm.
template <int KS, int NT, int EPI>
__global__ __launch_bounds__(256) void k_conv(Dims d, const bf16 *__restrict__ X, int cx, int cin,
                                              const bf16 *__restrict__ W, const float *__restrict__ bias, void *Yv,
                                              int cy, const bf16 *__restrict__ mask, int cm, float alpha) {
  constexpr int NTAP = KS * KS * KS, R = KS / 2;
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  const int64_t V = d.vox();
  const int64_t v0 = ((int64_t)blockIdx.x * 4 + wv) * 32;  // first voxel of the wave
  f4v acc[2][NT];
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[mt][nt] = f4v{0.f, 0.f, 0.f, 0.f};
  int bb[2], xx[2], yy[2], zz[2];
  bool rv[2];
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
    const int64_t v = v0 + 16 * mt + r;
    rv[mt] = v < V;
    vsplit(d, rv[mt] ? v : 0, bb[mt], xx[mt], yy[mt], zz[mt]);
  }
  const s4v zero4 = s4v{0, 0, 0, 0};
#pragma unroll 1
  for (int tap = 0; tap < NTAP; tap++) {
    const int dx = tap / (KS * KS) - R, dy = (tap / KS) % KS - R, dz = tap % KS - R;
    const bf16 *xp[2];
    bool ok[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
      const int x = xx[mt] + dx, y = yy[mt] + dy, z = zz[mt] + dz;
      ok[mt] = rv[mt] && x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz;
      xp[mt] = X + (ok[mt] ? vjoin(d, bb[mt], x, y, z) : 0) * cx + 4 * q;
    }
    const bf16 *wp = W + ((int64_t)tap * NT * 16 + r) * cin + 4 * q;
#pragma unroll 1
    for (int c0 = 0; c0 < cin; c0 += 16) {
      s4v a[2];
#pragma unroll
      for (int mt = 0; mt < 2; mt++) {
        s4v t = load4(xp[mt] + c0);
        a[mt] = ok[mt] ? t : zero4;
      }
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        const s4v b = load4(wp + (int64_t)nt * 16 * cin + c0);
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mt], b, acc[mt][nt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int64_t vox = v0 + 16 * mt + 4 * q + v;
      if (vox >= V) continue;
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        const int co = 16 * nt + r;
        float val = acc[mt][nt][v];
        if (EPI == EPI_ACT) {
          val += bias[co];
          val = val > 0.f ? val : alpha * val;
          ((bf16 *)Yv)[vox * cy + co] = f2bf(val);
        } else if (EPI == EPI_F32) {
          ((float *)Yv)[vox * cy + co] = val + bias[co];
        } else if (EPI == EPI_MASK) {
          const float m = bf2f(mask[vox * cm + co]);
          ((bf16 *)Yv)[vox * cy + co] = f2bf(m > 0.f ? val : alpha * val);
        } else {
          ((bf16 *)Yv)[vox * cy + co] = f2bf(val);
        }
      }
    }
}

// dW[tap][co][ci] += sum over a chunk of voxels of dY[v][co] X[v+off(tap)][ci].
// Grid (chunks, taps).  With at least four (co tile, ci tile) pairs each wave
// takes the pairs p with p % 4 == wave over the whole chunk; with fewer, every
// wave takes all pairs over every fourth group of 16 voxels, so that no wave
// idles on the finest level (16 x 16 channels, one pair).  A operand (rows co,
// depth voxels) and B operand (depth voxels, columns ci) gathered four voxels
// per lane.
template <int KS>
__global__ __launch_bounds__(256) void k_wgrad(Dims d, const bf16 *__restrict__ X, int cx, int cin,
                                               const bf16 *__restrict__ dY, int cdy, int cout, int64_t chunk,
                                               float *__restrict__ dW) {
  constexpr int R = KS / 2;
  constexpr int MAXP = 16;  // pairs per wave (cout/16 * cin/16 <= 64)
  const int tap = blockIdx.y;
  const int dx = tap / (KS * KS) - R, dy = (tap / KS) % KS - R, dz = tap % KS - R;
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  const int mtn = cout / 16, ntn = cin / 16, np = mtn * ntn;
  const int pw = np >= 4 ? 4 : 1;            // pair stride across waves
  const int pofs = wv % pw, vsl = wv / pw;   // this wave's first pair and voxel slice
  const int64_t vstep = 16 * (4 / pw);
  const int64_t V = d.vox();
  const int64_t vb = (int64_t)blockIdx.x * chunk, ve = vb + chunk < V ? vb + chunk : V;
  f4v acc[MAXP];
#pragma unroll
  for (int p = 0; p < MAXP; p++) acc[p] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int64_t k0 = vb + 16 * vsl; k0 < ve; k0 += vstep) {
    // the four voxels of this lane's depth slice, and their shifted partners
    int64_t vv[4], vs[4];
    bool okv[4], oks[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      vv[j] = k0 + 4 * q + j;
      okv[j] = vv[j] < ve;
      int b, x, y, z;
      vsplit(d, okv[j] ? vv[j] : 0, b, x, y, z);
      x += dx;
      y += dy;
      z += dz;
      oks[j] = okv[j] && x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz;
      vs[j] = oks[j] ? vjoin(d, b, x, y, z) : 0;
      if (!okv[j]) vv[j] = 0;
    }
#pragma unroll
    for (int pp = 0; pp < MAXP; pp++) {
      const int p = pofs + pw * pp;
      if (p >= np) continue;
      const int mt = p / ntn, nt = p % ntn;
      s4v a, b;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const bf16 av = dY[vv[j] * cdy + 16 * mt + r];
        const bf16 bv = X[vs[j] * cx + 16 * nt + r];
        a[j] = (short)(okv[j] ? av : 0);
        b[j] = (short)(oks[j] ? bv : 0);
      }
      acc[pp] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[pp], 0, 0, 0);
    }
  }
#pragma unroll
  for (int pp = 0; pp < MAXP; pp++) {
    const int p = pofs + pw * pp;
    if (p >= np) continue;
    const int mt = p / ntn, nt = p % ntn;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int co = 16 * mt + 4 * q + v, ci = 16 * nt + r;
      atomicAdd(&dW[((int64_t)tap * cout + co) * cin + ci], acc[pp][v]);
    }
  }
}

// db[c] += sum_v dY[v][c]
__global__ void k_bgrad(int64_t V, const bf16 *__restrict__ dY, int cdy, int c, int64_t chunk, float *db) {
  const int t = threadIdx.x;
  const int ch = t % c, lane0 = t / c, lanes = blockDim.x / c;
  if (lane0 >= lanes) return;
  const int64_t vb = (int64_t)blockIdx.x * chunk, ve = vb + chunk < V ? vb + chunk : V;
  float s = 0.f;
  for (int64_t v = vb + lane0; v < ve; v += lanes) s += bf2f(dY[v * cdy + ch]);
  atomicAdd(&db[ch], s);
}

// fp32 host layout [v][creal] -> bf16 [v][cpad], zero padded
__global__ void k_pack(int64_t V, const float *__restrict__ in, int creal, int cpad, bf16 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * cpad) return;
  const int64_t v = i / cpad;
  const int c = (int)(i % cpad);
  out[i] = f2bf(c < creal ? in[v * creal + c] : 0.f);
}

// out (coarse d/2) [v'][c] = mean of the 8 fine children of in (fine d) [v][c]
__global__ void k_pool(Dims fine, const bf16 *__restrict__ in, int cin_stride, int c, bf16 *__restrict__ out) {
  Dims co{fine.nb, fine.nx / 2, fine.ny / 2, fine.nz / 2};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= co.vox() * c) return;
  const int64_t v = i / c;
  const int ch = (int)(i % c);
  int b, x, y, z;
  vsplit(co, v, b, x, y, z);
  float s = 0.f;
  for (int t = 0; t < 8; t++)
    s += bf2f(in[vjoin(fine, b, 2 * x + (t >> 2), 2 * y + ((t >> 1) & 1), 2 * z + (t & 1)) * cin_stride + ch]);
  out[v * c + ch] = f2bf(0.125f * s);
}

// out (fine) [v][c] (channel stride cout) = in (coarse) [v/2][c]
__global__ void k_up(Dims fine, const bf16 *__restrict__ in, int c, bf16 *__restrict__ out, int cout_stride) {
  Dims co{fine.nb, fine.nx / 2, fine.ny / 2, fine.nz / 2};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fine.vox() * c) return;
  const int64_t v = i / c;
  const int ch = (int)(i % c);
  int b, x, y, z;
  vsplit(fine, v, b, x, y, z);
  out[v * cout_stride + ch] = in[vjoin(co, b, x / 2, y / 2, z / 2) * c + ch];
}

// adjoint of k_up followed by leaky' of the coarse activation h:
// dh[v'][c] = leaky'(h[v'][c]) * sum of the 8 children dX[v][c]
__global__ void k_up_bwd(Dims fine, const bf16 *__restrict__ dX, int cdx, const bf16 *__restrict__ h, int c,
                         float alpha, bf16 *__restrict__ dh) {
  Dims co{fine.nb, fine.nx / 2, fine.ny / 2, fine.nz / 2};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= co.vox() * c) return;
  const int64_t v = i / c;
  const int ch = (int)(i % c);
  int b, x, y, z;
  vsplit(co, v, b, x, y, z);
  float s = 0.f;
  for (int t = 0; t < 8; t++)
    s += bf2f(dX[vjoin(fine, b, 2 * x + (t >> 2), 2 * y + ((t >> 1) & 1), 2 * z + (t & 1)) * cdx + ch]);
  dh[v * c + ch] = f2bf(bf2f(h[v * c + ch]) > 0.f ? s : alpha * s);
}

// gradient of an encoder output h (fine) that feeds both the skip connection
// (dskip, a slice of a wider gradient tensor) and the pooled path (dpool,
// coarse): dh = leaky'(h) * (dskip + dpool[v/2] / 8).  h has channel stride ch_stride.
__global__ void k_pool_bwd(Dims fine, const bf16 *__restrict__ dskip, int cds, const bf16 *__restrict__ dpool,
                           const bf16 *__restrict__ h, int ch_stride, int c, float alpha, bf16 *__restrict__ dh) {
  Dims co{fine.nb, fine.nx / 2, fine.ny / 2, fine.nz / 2};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fine.vox() * c) return;
  const int64_t v = i / c;
  const int ch = (int)(i % c);
  int b, x, y, z;
  vsplit(fine, v, b, x, y, z);
  float s = 0.125f * bf2f(dpool[vjoin(co, b, x / 2, y / 2, z / 2) * c + ch]);
  if (dskip) s += bf2f(dskip[v * cds + ch]);
  dh[v * c + ch] = f2bf(bf2f(h[v * ch_stride + ch]) > 0.f ? s : alpha * s);
}

// masked MSE: loss += sum m (y - t)^2 / count; dY = 2 m (y - t) / count (bf16)
__global__ void k_loss(int64_t V, const float *__restrict__ y, int cy, const float *__restrict__ t,
                       const float *__restrict__ m, int creal, float inv_count, bf16 *__restrict__ dY, int cdy,
                       float *loss) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.f;
  if (i < V * cdy) {
    const int64_t v = i / cdy;
    const int c = (int)(i % cdy);
    float g = 0.f;
    if (c < creal) {
      const float mm = m ? m[v * creal + c] : 1.f;
      const float e = y[v * cy + c] - t[v * creal + c];
      s = mm * e * e * inv_count;
      g = 2.f * mm * e * inv_count;
    }
    dY[v * cdy + c] = f2bf(g);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// fp32 [v][cpad] -> host layout [v][creal]
__global__ void k_unpack(int64_t V, const float *__restrict__ in, int cpad, int creal, float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * creal) return;
  out[i] = in[(i / creal) * cpad + (i % creal)];
}

// Adam on one layer's fp32 master weights w[tap][co][ci] and bias, then the
// bf16 copies in both layouts: Wb[tap][co][ci] and Wt[flip(tap)][ci][co]
__global__ void k_adam(int64_t nw, int cout, int cin, int ntap, float *w, float *mw, float *vw, const float *gw,
                       float lr, float b1, float b2, float eps, float c1, float c2, bf16 *Wb, bf16 *Wt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const float g = gw[i];
  const float m = b1 * mw[i] + (1.f - b1) * g;
  const float v = b2 * vw[i] + (1.f - b2) * g * g;
  mw[i] = m;
  vw[i] = v;
  const float x = w[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
  w[i] = x;
  const int64_t per = (int64_t)cout * cin;
  const int tap = (int)(i / per);
  const int co = (int)((i % per) / cin), ci = (int)(i % cin);
  const bf16 xb = f2bf(x);
  Wb[i] = xb;
  Wt[((int64_t)(ntap - 1 - tap) * cin + ci) * cout + co] = xb;
}

__global__ void k_adam_bias(int n, float *b, float *mb, float *vb, const float *gb, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = gb[i];
  const float m = b1 * mb[i] + (1.f - b1) * g;
  const float v = b2 * vb[i] + (1.f - b2) * g * g;
  mb[i] = m;
  vb[i] = v;
  b[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

// bf16 copies of fp32 master weights (after set_params)
__global__ void k_wcopy(int64_t nw, int cout, int cin, int ntap, const float *w, bf16 *Wb, bf16 *Wt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const int64_t per = (int64_t)cout * cin;
  const int tap = (int)(i / per);
  const int co = (int)((i % per) / cin), ci = (int)(i % cin);
  const bf16 xb = f2bf(w[i]);
  Wb[i] = xb;
  Wt[((int64_t)(ntap - 1 - tap) * cin + ci) * cout + co] = xb;
}

__global__ void k_fill_f32(int64_t n, float *p, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace nn
}  // namespace mvf

// ------------------------------------------------------------------ network
struct mvf_nn {
  std::string err;
  mvf_nn_config cfg{};
  hipStream_t stream = nullptr;
  int nb = 0;  // batch the buffers hold
  static constexpr int NL = 4;  // levels
  int ch[NL] = {0, 0, 0, 0};
  struct Layer {
    int ks = 3, cin = 0, cout = 0;
    int64_t nw = 0;
    float *w = nullptr, *b = nullptr, *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    float *gw = nullptr, *gb = nullptr;
    mvf::nn::bf16 *Wb = nullptr, *Wt = nullptr;
  };
  // encoder e[k][0..1], decoder d[k][0..1] (k = 0..2), head
  Layer enc[NL][2], dec[NL - 1][2], head;
  // activations (bf16): input, per level: e1, cat (decoder input; its upper
  // channels hold the encoder output = skip), pooled input of the next
  // level, decoder d1, d2; bottleneck e2 at the last level
  mvf::nn::bf16 *x0 = nullptr;
  mvf::nn::bf16 *e1[NL] = {}, *cat[NL] = {}, *pin[NL] = {}, *d1[NL] = {}, *d2[NL] = {}, *bott = nullptr;
  // gradients (bf16), same shapes
  mvf::nn::bf16 *ge1[NL] = {}, *gcat[NL] = {}, *gpin[NL] = {}, *gd1[NL] = {}, *gd2[NL] = {}, *gbott = nullptr;
  mvf::nn::bf16 *gout = nullptr;
  float *out = nullptr, *xin = nullptr, *tin = nullptr, *min_ = nullptr, *dloss = nullptr, *hout = nullptr;
  std::vector<void *> owned;
  int64_t step = 0;
  // last forward+backward: matrix-core kernel time (HIP events) and flops
  double conv_ms = 0.0, conv_flops = 0.0;
  int64_t conv_launches = 0;
  std::vector<hipEvent_t> evs;
  bool timing = false;
};

namespace mvf {
namespace nn {

static int nn_err(mvf_nn *n, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  n->err = buf;
  return code;
}

#define NN_HIP(n, call)                                                                            \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return nn_err((n), (e_ == hipErrorOutOfMemory) ? MVF_E_MEMORY : MVF_E_DEVICE, "%s: %s", #call, \
                    hipGetErrorString(e_));                                                        \
  } while (0)

template <typename T>
static int nn_alloc(mvf_nn *n, T **p, size_t count) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
  if (e != hipSuccess) return nn_err(n, MVF_E_MEMORY, "hipMalloc(%zu) failed", count * sizeof(T));
  n->owned.push_back(q);
  *p = (T *)q;
  return 0;
}

static Dims level_dims(const mvf_nn *n, int k) {
  return Dims{n->nb, n->cfg.nx >> k, n->cfg.ny >> k, n->cfg.nz >> k};
}

static int make_layer(mvf_nn *n, mvf_nn::Layer &L, int ks, int cin, int cout, std::mt19937_64 &rng) {
  L.ks = ks;
  L.cin = cin;
  L.cout = cout;
  const int ntap = ks * ks * ks;
  L.nw = (int64_t)ntap * cin * cout;
  for (float **p : {&L.w, &L.mw, &L.vw, &L.gw}) MVF_TRY(nn_alloc(n, p, L.nw));
  for (float **p : {&L.b, &L.mb, &L.vb, &L.gb}) MVF_TRY(nn_alloc(n, p, cout));
  MVF_TRY(nn_alloc(n, &L.Wb, L.nw));
  MVF_TRY(nn_alloc(n, &L.Wt, L.nw));
  // He-uniform initialisation from the seeded generator
  std::vector<float> w(L.nw);
  const float bound = std::sqrt(6.0f / (float)(ntap * cin));
  std::uniform_real_distribution<float> U(-bound, bound);
  for (auto &x : w) x = U(rng);
  NN_HIP(n, hipMemcpy(L.w, w.data(), 4 * L.nw, hipMemcpyHostToDevice));
  std::vector<float> z(std::max<int64_t>(L.nw, cout), 0.f);
  for (float *p : {L.mw, L.vw, L.gw}) NN_HIP(n, hipMemcpy(p, z.data(), 4 * L.nw, hipMemcpyHostToDevice));
  for (float *p : {L.b, L.mb, L.vb, L.gb}) NN_HIP(n, hipMemcpy(p, z.data(), 4 * cout, hipMemcpyHostToDevice));
  return 0;
}

static void refresh_bf16(mvf_nn *n, mvf_nn::Layer &L) {
  hipLaunchKernelGGL(k_wcopy, dim3((unsigned)((L.nw + 255) / 256)), dim3(256), 0, n->stream, L.nw, L.cout, L.cin,
                     L.ks * L.ks * L.ks, L.w, L.Wb, L.Wt);
}

template <typename F>
static void for_layers(mvf_nn *n, F f) {
  for (int k = 0; k < mvf_nn::NL; k++)
    for (int s = 0; s < 2; s++) f(n->enc[k][s]);
  for (int k = 0; k < mvf_nn::NL - 1; k++)
    for (int s = 0; s < 2; s++) f(n->dec[k][s]);
  f(n->head);
}

static int alloc_acts(mvf_nn *n, int nb) {
  // activation buffers for nb samples (allocated once, for the largest batch)
  n->nb = nb;
  const int64_t V0 = level_dims(n, 0).vox();
  MVF_TRY(nn_alloc(n, &n->x0, V0 * 16));
  for (int k = 0; k < mvf_nn::NL; k++) {
    const int64_t V = level_dims(n, k).vox();
    const int c = n->ch[k], cu = k + 1 < mvf_nn::NL ? n->ch[k + 1] : 0;
    MVF_TRY(nn_alloc(n, &n->e1[k], V * c));
    MVF_TRY(nn_alloc(n, &n->ge1[k], V * c));
    if (k + 1 < mvf_nn::NL) {
      MVF_TRY(nn_alloc(n, &n->cat[k], V * (cu + c)));
      MVF_TRY(nn_alloc(n, &n->gcat[k], V * (cu + c)));
      MVF_TRY(nn_alloc(n, &n->d1[k], V * c));
      MVF_TRY(nn_alloc(n, &n->gd1[k], V * c));
      MVF_TRY(nn_alloc(n, &n->d2[k], V * c));
      MVF_TRY(nn_alloc(n, &n->gd2[k], V * c));
    }
    if (k > 0) {
      MVF_TRY(nn_alloc(n, &n->pin[k], V * n->ch[k - 1]));
      MVF_TRY(nn_alloc(n, &n->gpin[k], V * n->ch[k - 1]));
    }
  }
  const int64_t V3 = level_dims(n, mvf_nn::NL - 1).vox();
  MVF_TRY(nn_alloc(n, &n->bott, V3 * n->ch[mvf_nn::NL - 1]));
  MVF_TRY(nn_alloc(n, &n->gbott, V3 * n->ch[mvf_nn::NL - 1]));
  MVF_TRY(nn_alloc(n, &n->out, V0 * 16));
  MVF_TRY(nn_alloc(n, &n->gout, V0 * 16));
  MVF_TRY(nn_alloc(n, &n->xin, V0 * n->cfg.cin));
  MVF_TRY(nn_alloc(n, &n->tin, V0 * n->cfg.cout));
  MVF_TRY(nn_alloc(n, &n->min_, V0 * n->cfg.cout));
  MVF_TRY(nn_alloc(n, &n->hout, V0 * n->cfg.cout));
  MVF_TRY(nn_alloc(n, &n->dloss, 1));
  return 0;
}

// conv launches: y = conv(x) with the layer's forward weights, or the data
// gradient with its flipped transposed weights
template <int KS, int EPI>
static void conv_nt(mvf_nn *n, int nt, Dims d, const bf16 *X, int cx, int cin, const bf16 *W, const float *bias,
                    void *Y, int cy, const bf16 *mask, int cm) {
  const unsigned grid = (unsigned)((d.vox() + 127) / 128);
  const float a = n->cfg.leaky;
#define NN_CONV(NT_)                                                                                          \
  hipLaunchKernelGGL((k_conv<KS, NT_, EPI>), dim3(grid), dim3(256), 0, n->stream, d, X, cx, cin, W, bias, Y, cy, \
                     mask, cm, a)
  switch (nt) {
    case 1: NN_CONV(1); break;
    case 2: NN_CONV(2); break;
    case 3: NN_CONV(3); break;
    case 4: NN_CONV(4); break;
    case 6: NN_CONV(6); break;
    case 10: NN_CONV(10); break;
    default: break;
  }
#undef NN_CONV
}

static void tick(mvf_nn *n) {
  if (!n->timing) return;
  hipEvent_t e;
  if (hipEventCreate(&e) == hipSuccess) {
    (void)hipEventRecord(e, n->stream);
    n->evs.push_back(e);
  }
}

static void conv_fwd(mvf_nn *n, const mvf_nn::Layer &L, Dims d, const bf16 *X, int cx, void *Y, int cy, bool f32) {
  tick(n);
  if (L.ks == 3) {
    if (f32) conv_nt<3, EPI_F32>(n, L.cout / 16, d, X, cx, L.cin, L.Wb, L.b, Y, cy, nullptr, 0);
    else conv_nt<3, EPI_ACT>(n, L.cout / 16, d, X, cx, L.cin, L.Wb, L.b, Y, cy, nullptr, 0);
  } else {
    if (f32) conv_nt<1, EPI_F32>(n, L.cout / 16, d, X, cx, L.cin, L.Wb, L.b, Y, cy, nullptr, 0);
    else conv_nt<1, EPI_ACT>(n, L.cout / 16, d, X, cx, L.cin, L.Wb, L.b, Y, cy, nullptr, 0);
  }
  tick(n);
  n->conv_flops += 2.0 * (double)d.vox() * L.ks * L.ks * L.ks * L.cin * L.cout;
  n->conv_launches++;
}

// data gradient dX (cin channels, stride cdx) from dY (cout channels);
// mask: the layer input whose leaky' multiplies the result (nullptr: raw)
static void conv_dgrad(mvf_nn *n, const mvf_nn::Layer &L, Dims d, const bf16 *dY, bf16 *dX, int cdx, const bf16 *mask,
                       int cm) {
  tick(n);
  if (L.ks == 3) {
    if (mask) conv_nt<3, EPI_MASK>(n, L.cin / 16, d, dY, L.cout, L.cout, L.Wt, nullptr, dX, cdx, mask, cm);
    else conv_nt<3, EPI_RAW>(n, L.cin / 16, d, dY, L.cout, L.cout, L.Wt, nullptr, dX, cdx, nullptr, 0);
  } else {
    if (mask) conv_nt<1, EPI_MASK>(n, L.cin / 16, d, dY, L.cout, L.cout, L.Wt, nullptr, dX, cdx, mask, cm);
    else conv_nt<1, EPI_RAW>(n, L.cin / 16, d, dY, L.cout, L.cout, L.Wt, nullptr, dX, cdx, nullptr, 0);
  }
  tick(n);
  n->conv_flops += 2.0 * (double)d.vox() * L.ks * L.ks * L.ks * L.cin * L.cout;
  n->conv_launches++;
}

static void conv_wgrad(mvf_nn *n, const mvf_nn::Layer &L, Dims d, const bf16 *X, int cx, const bf16 *dY) {
  const int64_t V = d.vox();
  // about 16 voxel chunks of at least 256 voxels per tap and chunk, rounded to 16
  int64_t chunk = (V + 63) / 64;
  chunk = ((chunk + 15) / 16) * 16;
  if (chunk < 256) chunk = 256;
  const unsigned nch = (unsigned)((V + chunk - 1) / chunk);
  tick(n);
  if (L.ks == 3)
    hipLaunchKernelGGL((k_wgrad<3>), dim3(nch, 27), dim3(256), 0, n->stream, d, X, cx, L.cin, dY, L.cout, L.cout,
                       chunk, L.gw);
  else
    hipLaunchKernelGGL((k_wgrad<1>), dim3(nch, 1), dim3(256), 0, n->stream, d, X, cx, L.cin, dY, L.cout, L.cout, chunk,
                       L.gw);
  tick(n);
  n->conv_flops += 2.0 * (double)V * L.ks * L.ks * L.ks * L.cin * L.cout;
  n->conv_launches++;
  const int64_t bch = 4096;
  hipLaunchKernelGGL(k_bgrad, dim3((unsigned)((V + bch - 1) / bch)), dim3(256), 0, n->stream, V, dY, L.cout, L.cout,
                     bch, L.gb);
}

static unsigned g256(int64_t n) { return (unsigned)((n + 255) / 256); }

// forward pass on the packed input x0; writes out (fp32, 16 channels)
static void forward(mvf_nn *n) {
  const int NL = mvf_nn::NL;
  const bf16 *in = n->x0;
  int cin_stride = 16;
  for (int k = 0; k < NL; k++) {
    Dims d = level_dims(n, k);
    const int c = n->ch[k];
    conv_fwd(n, n->enc[k][0], d, in, cin_stride, n->e1[k], c, false);
    if (k + 1 < NL) {
      // the encoder output lands in the upper channels of the decoder's input
      const int cu = n->ch[k + 1], cc = cu + c;
      bf16 *skip = n->cat[k] + cu;
      conv_fwd(n, n->enc[k][1], d, n->e1[k], c, skip, cc, false);
      hipLaunchKernelGGL(k_pool, dim3(g256(level_dims(n, k + 1).vox() * c)), dim3(256), 0, n->stream, d, skip, cc, c,
                         n->pin[k + 1]);
      in = n->pin[k + 1];
      cin_stride = c;
    } else {
      conv_fwd(n, n->enc[k][1], d, n->e1[k], c, n->bott, c, false);
    }
  }
  const bf16 *low = n->bott;
  for (int k = NL - 2; k >= 0; k--) {
    Dims d = level_dims(n, k);
    const int c = n->ch[k], cu = n->ch[k + 1], cc = cu + c;
    hipLaunchKernelGGL(k_up, dim3(g256(d.vox() * cu)), dim3(256), 0, n->stream, d, low, cu, n->cat[k], cc);
    conv_fwd(n, n->dec[k][0], d, n->cat[k], cc, n->d1[k], c, false);
    conv_fwd(n, n->dec[k][1], d, n->d1[k], c, n->d2[k], c, false);
    low = n->d2[k];
  }
  conv_fwd(n, n->head, level_dims(n, 0), n->d2[0], n->ch[0], n->out, 16, true);
}

// backward pass from gout (bf16 gradient of the fp32 output)
static void backward(mvf_nn *n) {
  const int NL = mvf_nn::NL;
  const float a = n->cfg.leaky;
  Dims d0 = level_dims(n, 0);
  conv_wgrad(n, n->head, d0, n->d2[0], n->ch[0], n->gout);
  conv_dgrad(n, n->head, d0, n->gout, n->gd2[0], n->ch[0], n->d2[0], n->ch[0]);
  // decoder, finest first; gd2[k] holds the gradient of the pre-activation of d2[k]
  for (int k = 0; k < NL - 1; k++) {
    Dims d = level_dims(n, k);
    const int c = n->ch[k], cu = n->ch[k + 1], cc = cu + c;
    conv_wgrad(n, n->dec[k][1], d, n->d1[k], c, n->gd2[k]);
    conv_dgrad(n, n->dec[k][1], d, n->gd2[k], n->gd1[k], c, n->d1[k], c);
    conv_wgrad(n, n->dec[k][0], d, n->cat[k], cc, n->gd1[k]);
    conv_dgrad(n, n->dec[k][0], d, n->gd1[k], n->gcat[k], cc, nullptr, 0);
    // the upsampled part goes to the coarser decoder output (or the bottleneck)
    Dims dc = level_dims(n, k + 1);
    bf16 *gh = (k + 1 < NL - 1) ? n->gd2[k + 1] : n->gbott;
    const bf16 *h = (k + 1 < NL - 1) ? n->d2[k + 1] : n->bott;
    hipLaunchKernelGGL(k_up_bwd, dim3(g256(dc.vox() * cu)), dim3(256), 0, n->stream, d, n->gcat[k], cc, h, cu, a,
                       gh);
  }
  // encoder, coarsest first; ge holds gradients of pre-activations
  for (int k = NL - 1; k >= 0; k--) {
    Dims d = level_dims(n, k);
    const int c = n->ch[k];
    bf16 *g2;  // gradient of the pre-activation of the level's second conv
    if (k == NL - 1) {
      g2 = n->gbott;
    } else {
      const int cu = n->ch[k + 1], cc = cu + c;
      // skip gradient (upper channels of gcat) + pooled path, times leaky'
      g2 = n->gcat[k] + cu;  // written in place over the skip slice
      hipLaunchKernelGGL(k_pool_bwd, dim3(g256(d.vox() * c)), dim3(256), 0, n->stream, d, n->gcat[k] + cu, cc,
                         n->gpin[k + 1], n->cat[k] + cu, cc, c, a, n->ge1[k]);
      // k_pool_bwd wrote the combined gradient (stride c) into ge1[k]; the
      // second conv's gradient input must be stride c: use ge1 as g2 and the
      // first conv's gradient goes to the pooled buffer of this level
      g2 = n->ge1[k];
    }
    const bf16 *xin = k == 0 ? n->x0 : n->pin[k];
    const int cin_stride = k == 0 ? 16 : n->ch[k - 1];
    conv_wgrad(n, n->enc[k][1], d, n->e1[k], c, g2);
    // gradient of e1's pre-activation: into a buffer of stride c that is free
    // now (d1[k] for k < NL-1; at the bottleneck level the pooled gradient
    // buffer of the level is not needed any more)
    bf16 *g1 = (k < NL - 1) ? n->gd1[k] : n->gpin[k];
    if (k == NL - 1) g1 = n->gd1[0];  // any free buffer with >= V*c entries
    conv_dgrad(n, n->enc[k][1], d, g2, g1, c, n->e1[k], c);
    conv_wgrad(n, n->enc[k][0], d, xin, cin_stride, g1);
    if (k > 0) conv_dgrad(n, n->enc[k][0], d, g1, n->gpin[k], n->ch[k - 1], nullptr, 0);
  }
}

static int load_inputs(mvf_nn *n, int nb, const float *x, const float *t, const float *m) {
  const int64_t V = level_dims(n, 0).vox();
  (void)nb;
  NN_HIP(n, hipMemcpyAsync(n->xin, x, 4 * V * n->cfg.cin, hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(k_pack, dim3(g256(V * 16)), dim3(256), 0, n->stream, V, n->xin, n->cfg.cin, 16, n->x0);
  if (t) NN_HIP(n, hipMemcpyAsync(n->tin, t, 4 * V * n->cfg.cout, hipMemcpyHostToDevice, n->stream));
  if (m) NN_HIP(n, hipMemcpyAsync(n->min_, m, 4 * V * n->cfg.cout, hipMemcpyHostToDevice, n->stream));
  return 0;
}

static void zero_grads(mvf_nn *n) {
  for_layers(n, [&](mvf_nn::Layer &L) {
    hipLaunchKernelGGL(k_fill_f32, dim3(g256(L.nw)), dim3(256), 0, n->stream, L.nw, L.gw, 0.f);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, n->stream, (int64_t)L.cout, L.gb, 0.f);
  });
}

static void timing_begin(mvf_nn *n) {
  for (auto e : n->evs) (void)hipEventDestroy(e);
  n->evs.clear();
  n->conv_ms = n->conv_flops = 0.0;
  n->conv_launches = 0;
}

static void timing_end(mvf_nn *n) {
  if (!n->timing) return;
  (void)hipStreamSynchronize(n->stream);
  for (size_t i = 0; i + 1 < n->evs.size(); i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, n->evs[i], n->evs[i + 1]) == hipSuccess) n->conv_ms += ms;
  }
  for (auto e : n->evs) (void)hipEventDestroy(e);
  n->evs.clear();
}

}  // namespace nn
}  // namespace mvf

using namespace mvf::nn;

extern "C" {

void mvf_nn_default_config(mvf_nn_config *c) {
  if (!c) return;
  memset(c, 0, sizeof(*c));
  c->nx = 64;
  c->ny = 64;
  c->nz = 48;
  c->cin = 3;
  c->cout = 12;
  c->width = 16;
  c->batch = 2;
  c->device = 0;
  c->seed = 4;
  c->lr = 1e-3f;
  c->beta1 = 0.9f;
  c->beta2 = 0.999f;
  c->eps = 1e-8f;
  c->leaky = 0.1f;
}

int mvf_nn_create(const mvf_nn_config *cfg, mvf_nn **out) {
  if (!out || !cfg) return MVF_E_ARG;
  *out = nullptr;
  if (cfg->nx % 8 || cfg->ny % 8 || cfg->nz % 8 || cfg->nx < 8 || cfg->ny < 8 || cfg->nz < 8) return MVF_E_ARG;
  if (cfg->cin < 1 || cfg->cin > 16 || cfg->cout < 1 || cfg->cout > 16 || cfg->batch < 1) return MVF_E_ARG;
  if (cfg->width != 16) return MVF_E_UNSUPPORTED;
  mvf_nn *n = new (std::nothrow) mvf_nn();
  if (!n) return MVF_E_MEMORY;
  *out = n;
  n->cfg = *cfg;
  NN_HIP(n, hipSetDevice(cfg->device));
  NN_HIP(n, hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  const int w = cfg->width;
  n->ch[0] = w;
  n->ch[1] = 2 * w;
  n->ch[2] = 4 * w;
  n->ch[3] = 6 * w;
  std::mt19937_64 rng(cfg->seed);
  for (int k = 0; k < mvf_nn::NL; k++) {
    const int cin = k == 0 ? 16 : n->ch[k - 1];
    MVF_TRY(make_layer(n, n->enc[k][0], 3, cin, n->ch[k], rng));
    MVF_TRY(make_layer(n, n->enc[k][1], 3, n->ch[k], n->ch[k], rng));
  }
  for (int k = 0; k < mvf_nn::NL - 1; k++) {
    MVF_TRY(make_layer(n, n->dec[k][0], 3, n->ch[k + 1] + n->ch[k], n->ch[k], rng));
    MVF_TRY(make_layer(n, n->dec[k][1], 3, n->ch[k], n->ch[k], rng));
  }
  MVF_TRY(make_layer(n, n->head, 1, n->ch[0], 16, rng));
  for_layers(n, [&](mvf_nn::Layer &L) { refresh_bf16(n, L); });
  MVF_TRY(alloc_acts(n, cfg->batch));
  NN_HIP(n, hipStreamSynchronize(n->stream));
  return 0;
}

void mvf_nn_free(mvf_nn *n) {
  if (!n) return;
  (void)hipSetDevice(n->cfg.device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (auto e : n->evs) (void)hipEventDestroy(e);
  for (void *p : n->owned) (void)hipFree(p);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

const char *mvf_nn_last_error(const mvf_nn *n) { return n ? n->err.c_str() : "null network"; }

int64_t mvf_nn_num_params(const mvf_nn *n) {
  if (!n) return -1;
  int64_t s = 0;
  for_layers(const_cast<mvf_nn *>(n), [&](mvf_nn::Layer &L) { s += L.nw + L.cout; });
  return s;
}

// params / grads: per layer (encoder levels 0..3 conv 1, 2; decoder levels
// 0..2 conv 1, 2; head) w [tap][cout][cin] then b [cout]
static int copy_params(mvf_nn *n, float *host, bool grads, bool to_device) {
  if (!n || !host) return MVF_E_ARG;
  NN_HIP(n, hipSetDevice(n->cfg.device));
  NN_HIP(n, hipStreamSynchronize(n->stream));
  int64_t off = 0;
  int rc = 0;
  for_layers(n, [&](mvf_nn::Layer &L) {
    if (rc) return;
    float *w = grads ? L.gw : L.w, *b = grads ? L.gb : L.b;
    hipError_t e1, e2;
    if (to_device) {
      e1 = hipMemcpy(w, host + off, 4 * L.nw, hipMemcpyHostToDevice);
      e2 = hipMemcpy(b, host + off + L.nw, 4 * L.cout, hipMemcpyHostToDevice);
    } else {
      e1 = hipMemcpy(host + off, w, 4 * L.nw, hipMemcpyDeviceToHost);
      e2 = hipMemcpy(host + off + L.nw, b, 4 * L.cout, hipMemcpyDeviceToHost);
    }
    if (e1 != hipSuccess || e2 != hipSuccess) rc = nn_err(n, MVF_E_DEVICE, "parameter copy failed");
    off += L.nw + L.cout;
  });
  if (!rc && to_device && !grads) {
    for_layers(n, [&](mvf_nn::Layer &L) { refresh_bf16(n, L); });
    NN_HIP(n, hipStreamSynchronize(n->stream));
  }
  return rc;
}

int mvf_nn_get_params(mvf_nn *n, float *out) { return copy_params(n, out, false, false); }
int mvf_nn_set_params(mvf_nn *n, const float *in) { return copy_params(n, const_cast<float *>(in), false, true); }
int mvf_nn_get_grads(mvf_nn *n, float *out) { return copy_params(n, out, true, false); }

int mvf_nn_forward_backward(mvf_nn *n, int32_t nb, const float *x, const float *t, const float *m, float *loss) {
  if (!n || !x || !t || nb != n->nb) return n ? nn_err(n, MVF_E_ARG, "batch must equal the configured %d", n->nb)
                                              : MVF_E_ARG;
  NN_HIP(n, hipSetDevice(n->cfg.device));
  MVF_TRY(load_inputs(n, nb, x, t, m));
  timing_begin(n);
  forward(n);
  const int64_t V = level_dims(n, 0).vox();
  // normalisation: the masked entry count (all entries without a mask)
  double cnt = 0.0;
  if (m) {
    std::vector<float> hm(V * n->cfg.cout);
    memcpy(hm.data(), m, 4 * hm.size());
    for (float v : hm) cnt += v;
  } else {
    cnt = (double)V * n->cfg.cout;
  }
  const float inv = cnt > 0 ? (float)(1.0 / cnt) : 0.f;
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, n->stream, (int64_t)1, n->dloss, 0.f);
  hipLaunchKernelGGL(k_loss, dim3(g256(V * 16)), dim3(256), 0, n->stream, V, n->out, 16, n->tin,
                     m ? n->min_ : nullptr, n->cfg.cout, inv, n->gout, 16, n->dloss);
  zero_grads(n);
  backward(n);
  float hl = 0.f;
  NN_HIP(n, hipMemcpyAsync(&hl, n->dloss, 4, hipMemcpyDeviceToHost, n->stream));
  NN_HIP(n, hipStreamSynchronize(n->stream));
  NN_HIP(n, hipGetLastError());
  timing_end(n);
  if (loss) *loss = hl;
  return 0;
}

int mvf_nn_adam(mvf_nn *n, float lr) {
  if (!n) return MVF_E_ARG;
  NN_HIP(n, hipSetDevice(n->cfg.device));
  n->step++;
  const float b1 = n->cfg.beta1, b2 = n->cfg.beta2;
  const float c1 = 1.f / (1.f - std::pow(b1, (float)n->step)), c2 = 1.f / (1.f - std::pow(b2, (float)n->step));
  const float eps = n->cfg.eps;
  for_layers(n, [&](mvf_nn::Layer &L) {
    hipLaunchKernelGGL(k_adam, dim3(g256(L.nw)), dim3(256), 0, n->stream, L.nw, L.cout, L.cin, L.ks * L.ks * L.ks,
                       L.w, L.mw, L.vw, L.gw, lr, b1, b2, eps, c1, c2, L.Wb, L.Wt);
    hipLaunchKernelGGL(k_adam_bias, dim3(1), dim3(256), 0, n->stream, L.cout, L.b, L.mb, L.vb, L.gb, lr, b1, b2, eps,
                       c1, c2);
  });
  NN_HIP(n, hipStreamSynchronize(n->stream));
  NN_HIP(n, hipGetLastError());
  return 0;
}

int mvf_nn_predict(mvf_nn *n, int32_t nb, const float *x, float *y) {
  if (!n || !x || !y || nb != n->nb) return n ? nn_err(n, MVF_E_ARG, "batch must equal the configured %d", n->nb)
                                              : MVF_E_ARG;
  NN_HIP(n, hipSetDevice(n->cfg.device));
  MVF_TRY(load_inputs(n, nb, x, nullptr, nullptr));
  timing_begin(n);
  forward(n);
  const int64_t V = level_dims(n, 0).vox();
  hipLaunchKernelGGL(k_unpack, dim3(g256(V * n->cfg.cout)), dim3(256), 0, n->stream, V, n->out, 16, n->cfg.cout,
                     n->hout);
  NN_HIP(n, hipMemcpyAsync(y, n->hout, 4 * V * n->cfg.cout, hipMemcpyDeviceToHost, n->stream));
  NN_HIP(n, hipStreamSynchronize(n->stream));
  NN_HIP(n, hipGetLastError());
  timing_end(n);
  return 0;
}

int mvf_nn_set_timing(mvf_nn *n, int32_t on) {
  if (!n) return MVF_E_ARG;
  n->timing = on != 0;
  return 0;
}

int mvf_nn_stats(const mvf_nn *n, double *s3) {
  if (!n || !s3) return MVF_E_ARG;
  s3[0] = n->conv_ms;
  s3[1] = n->conv_flops;
  s3[2] = (double)n->conv_launches;
  return 0;
}

}  // extern "C"
