// Krylov solve on the device (replaces zmumps(job=6), MoVFEM_3DMT.f90:113-115).
//
// Operator: A_c = K + i w_c M for column c (w_c = float32(omega) of the
// column's frequency, as the reference's element matrix uses, integration.f90:84).
// MVF_MATRIX_C64REF applies the reference's complex64 rounding of every
// assembled value (ga_sort_sparse, global_assembly.f90:157-176) on the fly:
// A_c = fl32(K) + i fl32(w_c M).  K and M are the full symmetric CSR pattern.
//
// Vectors hold ncol complex columns interleaved per DOF ([nne][ncol]), so one
// pass over K and M serves every polarization and frequency of a batch.
// Methods (A is complex symmetric, A = A^T):
//   COCG      conjugate orthogonal CG: one SpMV per iteration, bilinear dots
//   BiCGStab  two SpMVs per iteration, Hermitian dots
// Jacobi preconditioning uses the per-column diagonal.  MVF_PREC_MIXED runs
// COCG with float32 matrix and vectors, and replaces the recursive residual by
// the FP64 true residual b - A x (FP64 operator, FP64 x) whenever it has
// dropped by inner_tol since the last replacement ("reliable updates"); all
// dot products accumulate in FP64.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>

#include "mvf_internal.h"

namespace mvf {

constexpr int kMaxCol = 16;
constexpr int kRedBlocks = 1024;
constexpr int kBS = 256;

struct ColW {
  double w[kMaxCol];
};

// per-column scalar state on the device
struct ColState {
  double rho_r, rho_i;      // COCG: r^T z ; BiCGStab: rhat^H r
  double alpha_r, alpha_i;
  double beta_r, beta_i;
  double om_r, om_i;        // BiCGStab omega
  double bnorm;             // ||b||
  double rnorm;             // ||r|| (recursive)
  double rref;              // MIXED: ||r|| at the last reliable update
  int active;
  int breakdown;
};

template <typename T>
__device__ inline double ld(const T *p, int64_t i) {
  return (double)p[i];
}

// ---------------------------------------------------------------- SpMV -----
// y = A x (columns 0..NC-1); LPR lanes share a row.  With DOT, block partials
// of p^T y (bilinear) are written to part[block][c].
template <typename VT, typename MT, int NC, int LPR, bool C64, bool DOT>
__device__ __forceinline__ void spmv_body(int64_t nne, const int64_t *__restrict__ rowptr,
                                          const int32_t *__restrict__ col, const MT *__restrict__ K,
                                          const MT *__restrict__ M, const ColW &w, const VT *__restrict__ x,
                                          VT *__restrict__ y, const VT *__restrict__ p, double *__restrict__ part) {
  const int lane = threadIdx.x % LPR;
  const int64_t rstride = (int64_t)gridDim.x * (kBS / LPR);
  double dr[NC], di[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) dr[c] = di[c] = 0.0;
  const int64_t rfirst = ((int64_t)blockIdx.x * kBS + threadIdx.x) / LPR;
  const int64_t rlast = rfirst + ((nne - rfirst + rstride - 1) / rstride) * rstride;  // uniform trip count per group
  for (int64_t row = rfirst; row < (rlast > rfirst ? rlast : rfirst + rstride); row += rstride) {
  double ar[NC], ai[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) ar[c] = ai[c] = 0.0;
  if (row < nne) {
    const int64_t k1 = rowptr[row + 1];
    // one entry of the lane: A_k x_col(k) accumulated into (ar, ai)
    auto entry = [&](int64_t cc, double kv, double mv) {
      const VT *xp = x + 2 * cc * NC;
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double vr = kv, vi = w.w[c] * mv;
        if (C64) {
          vr = (double)(float)vr;
          vi = (double)(float)vi;
        }
        const double xr = (double)xp[2 * c], xi = (double)xp[2 * c + 1];
        ar[c] += vr * xr - vi * xi;
        ai[c] += vr * xi + vi * xr;
      }
    };
    int64_t k = rowptr[row] + lane;
    // two entries per lane in flight, accumulated in the same order
    for (; k + LPR < k1; k += 2 * LPR) {
      const int64_t c0 = col[k], c1 = col[k + LPR];
      const double k0v = (double)K[k], m0v = (double)M[k], k1v = (double)K[k + LPR], m1v = (double)M[k + LPR];
      entry(c0, k0v, m0v);
      entry(c1, k1v, m1v);
    }
    if (k < k1) entry(col[k], (double)K[k], (double)M[k]);
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < NC; c++) {
      ar[c] += __shfl_xor(ar[c], off, LPR);
      ai[c] += __shfl_xor(ai[c], off, LPR);
    }
  if (row < nne) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      if (c % LPR == lane) {
        y[2 * (row * NC + c)] = (VT)ar[c];
        y[2 * (row * NC + c) + 1] = (VT)ai[c];
        if (DOT) {
          const double pr = (double)p[2 * (row * NC + c)], pi = (double)p[2 * (row * NC + c) + 1];
          dr[c] += pr * ar[c] - pi * ai[c];
          di[c] += pr * ai[c] + pi * ar[c];
        }
      }
    }
  }
  }
  if (DOT) {
    __shared__ double red[kBS / 64][NC][2];
#pragma unroll
    for (int c = 0; c < NC; c++)
      for (int off = 32; off > 0; off >>= 1) {
        dr[c] += __shfl_xor(dr[c], off, 64);
        di[c] += __shfl_xor(di[c], off, 64);
      }
    const int wv = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0)
      for (int c = 0; c < NC; c++) {
        red[wv][c][0] = dr[c];
        red[wv][c][1] = di[c];
      }
    __syncthreads();
    if (threadIdx.x < 2 * NC) {
      int c = threadIdx.x / 2, q = threadIdx.x % 2;
      double s = 0.0;
      for (int v = 0; v < kBS / 64; v++) s += red[v][c][q];
      part[((int64_t)blockIdx.x * NC + c) * 2 + q] = s;
    }
  }
}

template <typename VT, typename MT, int NC, int LPR, bool C64, bool DOT>
__global__ __launch_bounds__(kBS) void k_spmv(int64_t nne, const int64_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ col, const MT *__restrict__ K,
                                              const MT *__restrict__ M, ColW w, const VT *__restrict__ x,
                                              VT *__restrict__ y, const VT *__restrict__ p, double *__restrict__ part) {
  spmv_body<VT, MT, NC, LPR, C64, DOT>(nne, rowptr, col, K, M, w, x, y, p, part);
}

// ------------------------------------------------------- block reduction ----
// NQ complex quantities per column accumulated per thread, written as block
// partials part[block][q][c][2]
template <int NC, int NQ>
__device__ inline void block_partials(double (&v)[NQ][NC][2], double *part) {
  __shared__ double red[kBS / 64][NQ][NC][2];
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int s = 0; s < 2; s++)
        for (int off = 32; off > 0; off >>= 1) v[q][c][s] += __shfl_xor(v[q][c][s], off, 64);
  const int wv = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0)
    for (int q = 0; q < NQ; q++)
      for (int c = 0; c < NC; c++) {
        red[wv][q][c][0] = v[q][c][0];
        red[wv][q][c][1] = v[q][c][1];
      }
  __syncthreads();
  for (int t = threadIdx.x; t < NQ * NC * 2; t += kBS) {
    int q = t / (NC * 2), c = (t / 2) % NC, s = t % 2;
    double a = 0.0;
    for (int w = 0; w < kBS / 64; w++) a += red[w][q][c][s];
    part[(((int64_t)blockIdx.x * NQ + q) * NC + c) * 2 + s] = a;
  }
}

// sum of block partials -> out[q][c][2]; one block of 1024 threads
template <int NC, int NQ>
__device__ inline void reduce_partials(const double *part, int nblocks, double *out) {
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64, nw = blockDim.x / 64;
  for (int pr = wv; pr < NQ * NC * 2; pr += nw) {
    double a = 0.0;
    for (int b = lane; b < nblocks; b += 64) a += part[(int64_t)b * NQ * NC * 2 + pr];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) out[pr] = a;
  }
}

__device__ inline void cdiv(double ar, double ai, double br, double bi, double &qr, double &qi) {
  double d = br * br + bi * bi;
  qr = (ar * br + ai * bi) / d;
  qi = (ai * br - ar * bi) / d;
}

// ---------------------------------------------------------------- Jacobi ----
template <int NC, bool C64>
__global__ void k_invdiag(int64_t nne, const int64_t *diagpos, const double *K, const double *M, ColW w, double *invd) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nne) return;
  int64_t p = diagpos[r];
  double kv = p >= 0 ? K[p] : 0.0, mv = p >= 0 ? M[p] : 0.0;
  for (int c = 0; c < NC; c++) {
    double dr = kv, di = w.w[c] * mv;
    if (C64) {
      dr = (double)(float)dr;
      di = (double)(float)di;
    }
    double qr = 1.0, qi = 0.0;
    if (dr != 0.0 || di != 0.0) cdiv(1.0, 0.0, dr, di, qr, qi);
    invd[2 * (r * NC + c)] = qr;
    invd[2 * (r * NC + c) + 1] = qi;
  }
}

// ------------------------------------------------------------------ COCG ----
// init: r = b - A x (Ax given in q), z = D^-1 r, p = z; partials rho = r^T z, |r|^2, |b|^2
template <typename VT, int NC>
__global__ __launch_bounds__(kBS) void k_cocg_init(int64_t nne, const double *b, const VT *q, const double *invd,
                                                   VT *r, VT *p, double *part) {
  double v[3][NC][2];
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < NC; c++) v[a][c][0] = v[a][c][1] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double br = b[2 * i], bi = b[2 * i + 1];
    double rr = br - (double)q[2 * i], ri = bi - (double)q[2 * i + 1];
    double dr = invd[2 * i], di = invd[2 * i + 1];
    double zr = dr * rr - di * ri, zi = dr * ri + di * rr;
    r[2 * i] = (VT)rr; r[2 * i + 1] = (VT)ri;
    p[2 * i] = (VT)zr; p[2 * i + 1] = (VT)zi;
#pragma unroll
    for (int cc = 0; cc < NC; cc++)
      if (cc == c) {
        v[0][cc][0] += rr * zr - ri * zi;
        v[0][cc][1] += rr * zi + ri * zr;
        v[1][cc][0] += rr * rr + ri * ri;
        v[2][cc][0] += br * br + bi * bi;
      }
  }
  block_partials<NC, 3>(v, part);
}

template <int NC>
__global__ void k_cocg_fin_init(const double *part, int nb, ColState *st) {
  __shared__ double out[3 * NC * 2];
  reduce_partials<NC, 3>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) {
    int c = threadIdx.x;
    ColState &s = st[c];
    s.rho_r = out[(0 * NC + c) * 2];
    s.rho_i = out[(0 * NC + c) * 2 + 1];
    s.rnorm = sqrt(out[(1 * NC + c) * 2]);
    s.bnorm = sqrt(out[(2 * NC + c) * 2]);
    s.rref = s.rnorm;
    s.active = s.bnorm > 0.0 ? 1 : 0;
    s.breakdown = 0;
    s.alpha_r = s.alpha_i = s.beta_r = s.beta_i = 0.0;
  }
}

// alpha = rho / (p^T q); a zero denominator is a breakdown of that column
__device__ inline void cocg_alpha_step(ColState &s, double pr, double pi) {
  if (!s.active) {
    s.alpha_r = s.alpha_i = 0.0;
  } else if (pr == 0.0 && pi == 0.0) {
    s.alpha_r = s.alpha_i = 0.0;
    s.breakdown = 1;
    s.active = 0;
  } else {
    cdiv(s.rho_r, s.rho_i, pr, pi, s.alpha_r, s.alpha_i);
  }
}

template <int NC>
__global__ void k_cocg_alpha(const double *part, int nb, ColState *st) {
  __shared__ double out[NC * 2];
  reduce_partials<NC, 1>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) cocg_alpha_step(st[threadIdx.x], out[2 * threadIdx.x], out[2 * threadIdx.x + 1]);
}

// x += alpha p (x in XT precision), r -= alpha q; partials rho' = r^T D^-1 r, |r|^2
template <typename VT, typename XT, int NC>
__global__ __launch_bounds__(kBS) void k_cocg_upd1(int64_t nne, const ColState *st, XT *x, VT *r, const VT *p,
                                                   const VT *q, const double *invd, double *part) {
  double v[2][NC][2];
  for (int a = 0; a < 2; a++)
    for (int c = 0; c < NC; c++) v[a][c][0] = v[a][c][1] = 0.0;
  __shared__ double sa[NC][2];
  if (threadIdx.x < NC) {
    sa[threadIdx.x][0] = st[threadIdx.x].alpha_r;
    sa[threadIdx.x][1] = st[threadIdx.x].alpha_i;
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double alr = sa[c][0], ali = sa[c][1];
    double pr = (double)p[2 * i], pi = (double)p[2 * i + 1];
    double qr = (double)q[2 * i], qi = (double)q[2 * i + 1];
    x[2 * i] = (XT)((double)x[2 * i] + alr * pr - ali * pi);
    x[2 * i + 1] = (XT)((double)x[2 * i + 1] + alr * pi + ali * pr);
    double rr = (double)r[2 * i] - (alr * qr - ali * qi);
    double ri = (double)r[2 * i + 1] - (alr * qi + ali * qr);
    r[2 * i] = (VT)rr;
    r[2 * i + 1] = (VT)ri;
    rr = (double)r[2 * i];
    ri = (double)r[2 * i + 1];
    double dr = invd[2 * i], di = invd[2 * i + 1];
    double zr = dr * rr - di * ri, zi = dr * ri + di * rr;
#pragma unroll
    for (int cc = 0; cc < NC; cc++)
      if (cc == c) {
        v[0][cc][0] += rr * zr - ri * zi;
        v[0][cc][1] += rr * zi + ri * zr;
        v[1][cc][0] += rr * rr + ri * ri;
      }
  }
  block_partials<NC, 2>(v, part);
}

// beta = rho'/rho, rho = rho'; convergence of the recursive residual freezes a column
__device__ inline void cocg_beta_step(ColState &s, double nr, double ni, double rn2, double tol_abs_scale,
                                      int freeze) {
  s.rnorm = sqrt(rn2);
  if (!s.active) {
    s.beta_r = s.beta_i = 0.0;
    return;
  }
  if (s.rho_r == 0.0 && s.rho_i == 0.0) {
    s.breakdown = 1;
    s.active = 0;
    s.beta_r = s.beta_i = 0.0;
    return;
  }
  cdiv(nr, ni, s.rho_r, s.rho_i, s.beta_r, s.beta_i);
  s.rho_r = nr;
  s.rho_i = ni;
  if (freeze && s.rnorm <= tol_abs_scale * s.bnorm) s.active = 0;
}

template <int NC>
__global__ void k_cocg_beta(const double *part, int nb, ColState *st, double tol_abs_scale, int freeze) {
  __shared__ double out[2 * NC * 2];
  reduce_partials<NC, 2>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) {
    const int c = threadIdx.x;
    cocg_beta_step(st[c], out[(0 * NC + c) * 2], out[(0 * NC + c) * 2 + 1], out[(1 * NC + c) * 2], tol_abs_scale,
                   freeze);
  }
}

// p = D^-1 r + beta p
template <typename VT, int NC>
__global__ __launch_bounds__(kBS) void k_cocg_upd2(int64_t nne, const ColState *st, const VT *r, VT *p,
                                                   const double *invd) {
  __shared__ double sb[NC][2];
  if (threadIdx.x < NC) {
    sb[threadIdx.x][0] = st[threadIdx.x].beta_r;
    sb[threadIdx.x][1] = st[threadIdx.x].beta_i;
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double rr = (double)r[2 * i], ri = (double)r[2 * i + 1];
    double dr = invd[2 * i], di = invd[2 * i + 1];
    double zr = dr * rr - di * ri, zi = dr * ri + di * rr;
    double pr = (double)p[2 * i], pi = (double)p[2 * i + 1];
    double br = sb[c][0], bi = sb[c][1];
    p[2 * i] = (VT)(zr + br * pr - bi * pi);
    p[2 * i + 1] = (VT)(zi + br * pi + bi * pr);
  }
}

// ------------------------------------------------------- persistent COCG ----
// `iters` COCG iterations in one cooperative launch: the SpMV with its fused
// p^T A p partials, alpha, the x/r update with its partials, beta and the p
// update, separated by three grid barriers per iteration.  Every block reduces
// the partials itself (the same order as the one-block reduction kernels) and
// keeps its own copy of the column state, so no further barrier is needed; block 0
// writes the state back at the end.  Every block runs every iteration: inactive
// columns only carry zero alpha and beta.
template <int NC, int LPR, bool C64>
__global__ __launch_bounds__(kBS) void k_cocg_persist(int iters, int64_t nne, const int64_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col, const double *__restrict__ K,
                                                      const double *__restrict__ M, ColW w, double *__restrict__ x,
                                                      double *__restrict__ r, double *__restrict__ p,
                                                      double *__restrict__ q, const double *__restrict__ invd,
                                                      double *__restrict__ spart, double *__restrict__ part,
                                                      ColState *__restrict__ st, double tol) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  __shared__ ColState cs[NC];
  __shared__ double out1[NC * 2], out2[2 * NC * 2];
  if (threadIdx.x < NC) cs[threadIdx.x] = st[threadIdx.x];
  __syncthreads();
  const int nb = gridDim.x;
  for (int k = 0; k < iters; k++) {
    spmv_body<double, double, NC, LPR, C64, true>(nne, rowptr, col, K, M, w, p, q, p, spart);
    grid.sync();
    reduce_partials<NC, 1>(spart, nb, out1);
    __syncthreads();
    if (threadIdx.x < NC) cocg_alpha_step(cs[threadIdx.x], out1[2 * threadIdx.x], out1[2 * threadIdx.x + 1]);
    __syncthreads();
    {
      double v[2][NC][2];
      for (int a = 0; a < 2; a++)
        for (int c = 0; c < NC; c++) v[a][c][0] = v[a][c][1] = 0.0;
      for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
        const int c = (int)(i % NC);
        const double alr = cs[c].alpha_r, ali = cs[c].alpha_i;
        const double pr = p[2 * i], pi = p[2 * i + 1], qr = q[2 * i], qi = q[2 * i + 1];
        x[2 * i] = x[2 * i] + alr * pr - ali * pi;
        x[2 * i + 1] = x[2 * i + 1] + alr * pi + ali * pr;
        const double rr = r[2 * i] - (alr * qr - ali * qi), ri = r[2 * i + 1] - (alr * qi + ali * qr);
        r[2 * i] = rr;
        r[2 * i + 1] = ri;
        const double dr = invd[2 * i], di = invd[2 * i + 1];
        const double zr = dr * rr - di * ri, zi = dr * ri + di * rr;
#pragma unroll
        for (int cc = 0; cc < NC; cc++)
          if (cc == c) {
            v[0][cc][0] += rr * zr - ri * zi;
            v[0][cc][1] += rr * zi + ri * zr;
            v[1][cc][0] += rr * rr + ri * ri;
          }
      }
      block_partials<NC, 2>(v, part);
    }
    grid.sync();
    reduce_partials<NC, 2>(part, nb, out2);
    __syncthreads();
    if (threadIdx.x < NC) {
      const int c = threadIdx.x;
      cocg_beta_step(cs[c], out2[(0 * NC + c) * 2], out2[(0 * NC + c) * 2 + 1], out2[(1 * NC + c) * 2], tol, 1);
    }
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
      const int c = (int)(i % NC);
      const double rr = r[2 * i], ri = r[2 * i + 1], dr = invd[2 * i], di = invd[2 * i + 1];
      const double zr = dr * rr - di * ri, zi = dr * ri + di * rr;
      const double pr = p[2 * i], pi = p[2 * i + 1], br = cs[c].beta_r, bi = cs[c].beta_i;
      p[2 * i] = zr + br * pr - bi * pi;
      p[2 * i + 1] = zi + br * pi + bi * pr;
    }
    grid.sync();
  }
  if (blockIdx.x == 0 && threadIdx.x < NC) st[threadIdx.x] = cs[threadIdx.x];
}

// --------------------------------------------------------------- helpers ----
// out = b - q (fp64), partial |out|^2 and |b|^2
template <int NC>
__global__ __launch_bounds__(kBS) void k_resid(int64_t nne, const double *b, const double *q, double *out,
                                               double *part) {
  double v[2][NC][2];
  for (int a = 0; a < 2; a++)
    for (int c = 0; c < NC; c++) v[a][c][0] = v[a][c][1] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double rr = b[2 * i] - q[2 * i], ri = b[2 * i + 1] - q[2 * i + 1];
    if (out) {
      out[2 * i] = rr;
      out[2 * i + 1] = ri;
    }
#pragma unroll
    for (int cc = 0; cc < NC; cc++)
      if (cc == c) {
        v[0][cc][0] += rr * rr + ri * ri;
        v[1][cc][0] += b[2 * i] * b[2 * i] + b[2 * i + 1] * b[2 * i + 1];
      }
  }
  block_partials<NC, 2>(v, part);
}

template <int NC>
__global__ void k_norms(const double *part, int nb, double *out2) {
  __shared__ double out[2 * NC * 2];
  reduce_partials<NC, 2>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) {
    out2[2 * threadIdx.x] = sqrt(out[(0 * NC + threadIdx.x) * 2]);
    out2[2 * threadIdx.x + 1] = sqrt(out[(1 * NC + threadIdx.x) * 2]);
  }
}

template <typename A, typename B>
__global__ void k_convert(int64_t n, const A *a, B *b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (B)a[i];
}

// x64 += x32, x32 = 0
__global__ void k_accum(int64_t n, double *x, float *x32) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] += (double)x32[i];
    x32[i] = 0.0f;
  }
}

// ------------------------------------------------------------ BiCGStab -----
// state kept per column: rho (rhat^H r), alpha, omega.  Vectors: x, r, rh, p, v, s(aliases q), t
template <int NC>
__global__ __launch_bounds__(kBS) void k_bicg_init(int64_t nne, const double *b, const double *q, double *r,
                                                   double *rh, double *p, double *v, double *part) {
  double acc[3][NC][2];
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < NC; c++) acc[a][c][0] = acc[a][c][1] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double rr = b[2 * i] - q[2 * i], ri = b[2 * i + 1] - q[2 * i + 1];
    r[2 * i] = rr; r[2 * i + 1] = ri;
    rh[2 * i] = rr; rh[2 * i + 1] = ri;
    p[2 * i] = p[2 * i + 1] = 0.0;
    v[2 * i] = v[2 * i + 1] = 0.0;
#pragma unroll
    for (int cc = 0; cc < NC; cc++)
      if (cc == c) {
        acc[0][cc][0] += rr * rr + ri * ri;  // rhat^H r
        acc[1][cc][0] += rr * rr + ri * ri;
        acc[2][cc][0] += b[2 * i] * b[2 * i] + b[2 * i + 1] * b[2 * i + 1];
      }
  }
  block_partials<NC, 3>(acc, part);
}

template <int NC>
__global__ void k_bicg_fin_init(const double *part, int nb, ColState *st) {
  __shared__ double out[3 * NC * 2];
  reduce_partials<NC, 3>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) {
    int c = threadIdx.x;
    ColState &s = st[c];
    s.rho_r = 1.0; s.rho_i = 0.0;
    s.alpha_r = 1.0; s.alpha_i = 0.0;
    s.om_r = 1.0; s.om_i = 0.0;
    s.beta_r = 0.0; s.beta_i = 0.0;
    s.rnorm = sqrt(out[(1 * NC + c) * 2]);
    s.bnorm = sqrt(out[(2 * NC + c) * 2]);
    s.active = s.bnorm > 0.0 ? 1 : 0;
    s.breakdown = 0;
  }
}

// rho' = rhat^H r -> beta = (rho'/rho)(alpha/omega); p = r + beta (p - omega v); phat = D^-1 p
template <int NC>
__global__ __launch_bounds__(kBS) void k_bicg_p(int64_t nne, const ColState *st, const double *r, double *p,
                                                const double *v, const double *invd, double *phat) {
  __shared__ double sb[NC][4];
  if (threadIdx.x < NC) {
    sb[threadIdx.x][0] = st[threadIdx.x].beta_r;
    sb[threadIdx.x][1] = st[threadIdx.x].beta_i;
    sb[threadIdx.x][2] = st[threadIdx.x].om_r;
    sb[threadIdx.x][3] = st[threadIdx.x].om_i;
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double br = sb[c][0], bi = sb[c][1], wr = sb[c][2], wi = sb[c][3];
    double pr = p[2 * i] - (wr * v[2 * i] - wi * v[2 * i + 1]);
    double pi = p[2 * i + 1] - (wr * v[2 * i + 1] + wi * v[2 * i]);
    double nr = r[2 * i] + br * pr - bi * pi, ni = r[2 * i + 1] + br * pi + bi * pr;
    p[2 * i] = nr; p[2 * i + 1] = ni;
    double dr = invd[2 * i], di = invd[2 * i + 1];
    phat[2 * i] = dr * nr - di * ni;
    phat[2 * i + 1] = dr * ni + di * nr;
  }
}

// partial rhat^H v
template <int NC>
__global__ __launch_bounds__(kBS) void k_hdot(int64_t nne, const double *a, const double *b, double *part) {
  double acc[1][NC][2];
  for (int c = 0; c < NC; c++) acc[0][c][0] = acc[0][c][1] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
#pragma unroll
    for (int cc = 0; cc < NC; cc++)
      if (cc == c) {
        acc[0][cc][0] += ar * br + ai * bi;
        acc[0][cc][1] += ar * bi - ai * br;
      }
  }
  block_partials<NC, 1>(acc, part);
}

// alpha = rho / (rhat^H v)
template <int NC>
__global__ void k_bicg_alpha(const double *part, int nb, ColState *st) {
  __shared__ double out[NC * 2];
  reduce_partials<NC, 1>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) {
    ColState &s = st[threadIdx.x];
    double dr = out[2 * threadIdx.x], di = out[2 * threadIdx.x + 1];
    if (!s.active) {
      s.alpha_r = s.alpha_i = 0.0;
    } else if (dr == 0.0 && di == 0.0) {
      s.breakdown = 1;
      s.active = 0;
      s.alpha_r = s.alpha_i = 0.0;
    } else {
      cdiv(s.rho_r, s.rho_i, dr, di, s.alpha_r, s.alpha_i);
    }
  }
}

// s = r - alpha v ; shat = D^-1 s
template <int NC>
__global__ __launch_bounds__(kBS) void k_bicg_s(int64_t nne, const ColState *st, const double *r, const double *v,
                                                double *s, const double *invd, double *shat) {
  __shared__ double sa[NC][2];
  if (threadIdx.x < NC) {
    sa[threadIdx.x][0] = st[threadIdx.x].alpha_r;
    sa[threadIdx.x][1] = st[threadIdx.x].alpha_i;
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double ar = sa[c][0], ai = sa[c][1];
    double sr = r[2 * i] - (ar * v[2 * i] - ai * v[2 * i + 1]);
    double si = r[2 * i + 1] - (ar * v[2 * i + 1] + ai * v[2 * i]);
    s[2 * i] = sr; s[2 * i + 1] = si;
    double dr = invd[2 * i], di = invd[2 * i + 1];
    shat[2 * i] = dr * sr - di * si;
    shat[2 * i + 1] = dr * si + di * sr;
  }
}

// partials t^H s, t^H t
template <int NC>
__global__ __launch_bounds__(kBS) void k_bicg_ts(int64_t nne, const double *t, const double *s, double *part) {
  double acc[2][NC][2];
  for (int a = 0; a < 2; a++)
    for (int c = 0; c < NC; c++) acc[a][c][0] = acc[a][c][1] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double tr = t[2 * i], ti = t[2 * i + 1], sr = s[2 * i], si = s[2 * i + 1];
#pragma unroll
    for (int cc = 0; cc < NC; cc++)
      if (cc == c) {
        acc[0][cc][0] += tr * sr + ti * si;
        acc[0][cc][1] += tr * si - ti * sr;
        acc[1][cc][0] += tr * tr + ti * ti;
      }
  }
  block_partials<NC, 2>(acc, part);
}

template <int NC>
__global__ void k_bicg_omega(const double *part, int nb, ColState *st) {
  __shared__ double out[2 * NC * 2];
  reduce_partials<NC, 2>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) {
    int c = threadIdx.x;
    ColState &s = st[c];
    double nr = out[(0 * NC + c) * 2], ni = out[(0 * NC + c) * 2 + 1], tt = out[(1 * NC + c) * 2];
    if (!s.active || tt == 0.0) {
      s.om_r = s.om_i = 0.0;
      if (s.active && tt == 0.0) {
        s.breakdown = 1;
        s.active = 0;
      }
      return;
    }
    s.om_r = nr / tt;
    s.om_i = ni / tt;
  }
}

// x += alpha phat + omega shat ; r = s - omega t ; partials rhat^H r, |r|^2
template <int NC>
__global__ __launch_bounds__(kBS) void k_bicg_x(int64_t nne, const ColState *st, double *x, double *r,
                                                const double *s, const double *t, const double *phat,
                                                const double *shat, const double *rh, double *part) {
  __shared__ double sc[NC][4];
  if (threadIdx.x < NC) {
    sc[threadIdx.x][0] = st[threadIdx.x].alpha_r;
    sc[threadIdx.x][1] = st[threadIdx.x].alpha_i;
    sc[threadIdx.x][2] = st[threadIdx.x].om_r;
    sc[threadIdx.x][3] = st[threadIdx.x].om_i;
  }
  __syncthreads();
  double acc[2][NC][2];
  for (int a = 0; a < 2; a++)
    for (int c = 0; c < NC; c++) acc[a][c][0] = acc[a][c][1] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < nne * NC; i += (int64_t)gridDim.x * kBS) {
    int c = (int)(i % NC);
    double ar = sc[c][0], ai = sc[c][1], wr = sc[c][2], wi = sc[c][3];
    x[2 * i] += ar * phat[2 * i] - ai * phat[2 * i + 1] + wr * shat[2 * i] - wi * shat[2 * i + 1];
    x[2 * i + 1] += ar * phat[2 * i + 1] + ai * phat[2 * i] + wr * shat[2 * i + 1] + wi * shat[2 * i];
    double rr = s[2 * i] - (wr * t[2 * i] - wi * t[2 * i + 1]);
    double ri = s[2 * i + 1] - (wr * t[2 * i + 1] + wi * t[2 * i]);
    r[2 * i] = rr; r[2 * i + 1] = ri;
    double hr = rh[2 * i], hi = rh[2 * i + 1];
#pragma unroll
    for (int cc = 0; cc < NC; cc++)
      if (cc == c) {
        acc[0][cc][0] += hr * rr + hi * ri;
        acc[0][cc][1] += hr * ri - hi * rr;
        acc[1][cc][0] += rr * rr + ri * ri;
      }
  }
  block_partials<NC, 2>(acc, part);
}

// beta = (rho'/rho)(alpha/omega); rho = rho'
template <int NC>
__global__ void k_bicg_beta(const double *part, int nb, ColState *st, double tol, int freeze) {
  __shared__ double out[2 * NC * 2];
  reduce_partials<NC, 2>(part, nb, out);
  __syncthreads();
  if (threadIdx.x < NC) {
    int c = threadIdx.x;
    ColState &s = st[c];
    double nr = out[(0 * NC + c) * 2], ni = out[(0 * NC + c) * 2 + 1];
    s.rnorm = sqrt(out[(1 * NC + c) * 2]);
    if (!s.active) {
      s.beta_r = s.beta_i = 0.0;
      return;
    }
    if ((s.rho_r == 0.0 && s.rho_i == 0.0) || (s.om_r == 0.0 && s.om_i == 0.0)) {
      s.breakdown = 1;
      s.active = 0;
      return;
    }
    double q1r, q1i, q2r, q2i;
    cdiv(nr, ni, s.rho_r, s.rho_i, q1r, q1i);
    cdiv(s.alpha_r, s.alpha_i, s.om_r, s.om_i, q2r, q2i);
    s.beta_r = q1r * q2r - q1i * q2i;
    s.beta_i = q1r * q2i + q1i * q2r;
    s.rho_r = nr;
    s.rho_i = ni;
    if (freeze && s.rnorm <= tol * s.bnorm) s.active = 0;
  }
}

// ---------------------------------------------------------- orchestration --
// SpMV launches are bracketed by HIP events from a ring; elapsed times are
// collected at the solver's existing host synchronisation points.
struct SpmvTimer {
  static constexpr int kRing = 256;
  hipEvent_t ev[kRing][2];
  int kind[kRing];
  int used = 0;
  bool on = false;
  int64_t launches[2] = {0, 0};
  double ms[2] = {0, 0};
  void init() {
    for (int i = 0; i < kRing; i++) {
      (void)hipEventCreate(&ev[i][0]);
      (void)hipEventCreate(&ev[i][1]);
    }
    on = true;
  }
  void drain() {
    for (int i = 0; i < used; i++) {
      float t = 0.f;
      if (hipEventSynchronize(ev[i][1]) == hipSuccess && hipEventElapsedTime(&t, ev[i][0], ev[i][1]) == hipSuccess) {
        ms[kind[i]] += t;
        launches[kind[i]]++;
      }
    }
    used = 0;
  }
  void destroy() {
    if (!on) return;
    drain();
    for (int i = 0; i < kRing; i++) {
      (void)hipEventDestroy(ev[i][0]);
      (void)hipEventDestroy(ev[i][1]);
    }
    on = false;
  }
};

// A block of Krylov iterations captured once as a HIP graph and replayed:
// the kernels of check_every iterations take the same arguments every time
// (all state lives in device buffers), so one graph launch replaces
// 5 * check_every kernel launches between two convergence checks.
struct GraphBlock {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  int64_t launches = 0;
  GraphBlock() = default;
  GraphBlock(const GraphBlock &) = delete;
  GraphBlock &operator=(const GraphBlock &) = delete;
  ~GraphBlock() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
  }
};

struct Solver {
  mvf_handle *h;
  int ncol;
  ColW w{};
  bool c64;
  int lpr;
  int64_t n;  // nne
  SpmvTimer *tm;
  // device block the Krylov buffers are carved from; nullptr: allocated per call
  char *ws = nullptr;
  size_t ws_cap = 0;
};

// Krylov buffers of one solve: allocated (freed on return), carved from a
// caller's device block, or only counted (sizing).
struct Workspace {
  DevScope ds;
  char *base = nullptr;
  size_t cap = 0, used = 0;
  bool sizing = false;
  template <typename T>
  int get(mvf_handle *h, T **p, size_t n) {
    size_t bytes = (((n ? n : 1) * sizeof(T)) + 255) & ~(size_t)255;
    if (sizing) {
      used += bytes;
      *p = nullptr;
      return 0;
    }
    if (!base) return ds.alloc(h, p, n);
    if (used + bytes > cap)
      return set_err(h, MVF_E_ARG, "Krylov workspace needs more than the %zu bytes provided", cap);
    *p = (T *)(base + used);
    used += bytes;
    return 0;
  }
};

__global__ void k_fill_d(int64_t n, double *p, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// invd = 1 in every column (no preconditioner)
__global__ void k_unit_diag(int64_t nv, double *invd) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv / 2) {
    invd[2 * i] = 1.0;
    invd[2 * i + 1] = 0.0;
  }
}

// Replays the captured block of `count` iterations (capturing it on first use);
// the SpMV timer is off while capturing (its events are not part of the graph)
template <typename F>
static int graph_block(Solver &S, GraphBlock &gb, int count, F &&iter) {
  mvf_handle *h = S.h;
  if (!gb.x) {
    bool tm_on = S.tm && S.tm->on;
    if (S.tm) S.tm->on = false;
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int k = 0; k < count; k++) iter();
      e = hipStreamEndCapture(h->stream, &gb.g);
    }
    if (S.tm) S.tm->on = tm_on;
    if (e != hipSuccess) return set_err(h, MVF_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    MVF_HIP(h, hipGraphInstantiate(&gb.x, gb.g, nullptr, nullptr, 0));
  }
  MVF_HIP(h, hipGraphLaunch(gb.x, h->stream));
  gb.launches++;
  return 0;
}

constexpr int kSpmvBlocks = 4096;
static int spmv_blocks(int64_t n) {
  int64_t b = (n + kBS / 16 - 1) / (kBS / 16);
  return (int)(b < kSpmvBlocks ? b : kSpmvBlocks);
}

template <typename VT, typename MT, int NC, bool DOT>
static void spmv_launch(const Solver &S, const MT *K, const MT *M, const VT *x, VT *y, const VT *p, double *part) {
  mvf_handle *h = S.h;
  unsigned grid = (unsigned)spmv_blocks(S.n);
  SpmvTimer *tm = S.tm;
  if (tm && tm->on) {
    if (tm->used == SpmvTimer::kRing) tm->drain();
    tm->kind[tm->used] = sizeof(MT) == 4 ? 1 : 0;
    (void)hipEventRecord(tm->ev[tm->used][0], h->stream);
  }
  // lanes per row: the autotuned choice for the FP64 operator, 16 otherwise
  const int lpr = sizeof(MT) == 8 ? S.lpr : 16;
#define MVF_SPMV(L_)                                                                                             \
  do {                                                                                                           \
    if (S.c64)                                                                                                   \
      hipLaunchKernelGGL((k_spmv<VT, MT, NC, L_, true, DOT>), dim3(grid), dim3(kBS), 0, h->stream, S.n,          \
                         h->d_rowptr, h->d_col, K, M, S.w, x, y, p, part);                                       \
    else                                                                                                         \
      hipLaunchKernelGGL((k_spmv<VT, MT, NC, L_, false, DOT>), dim3(grid), dim3(kBS), 0, h->stream, S.n,         \
                         h->d_rowptr, h->d_col, K, M, S.w, x, y, p, part);                                       \
  } while (0)
  if (sizeof(MT) == 8 && lpr == 4) MVF_SPMV(4);
  else if (sizeof(MT) == 8 && lpr == 8) MVF_SPMV(8);
  else if (sizeof(MT) == 8 && lpr == 32) MVF_SPMV(32);
  else MVF_SPMV(16);
#undef MVF_SPMV
  if (tm && tm->on) {
    (void)hipEventRecord(tm->ev[tm->used][1], h->stream);
    tm->used++;
  }
}

static int vec_blocks(int64_t n, int nc) {
  int64_t b = (n * nc + kBS - 1) / kBS;
  return (int)(b < kRedBlocks ? (b > 0 ? b : 1) : kRedBlocks);
}

template <int NC>
static int host_state(mvf_handle *h, ColState *dst, const ColState *src) {
  MVF_HIP(h, hipMemcpyAsync(dst, src, sizeof(ColState) * NC, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  return 0;
}

// true FP64 residual norms: out[c] = ||b - A x|| / ||b||
template <int NC>
static int true_resid(const Solver &S, const double *b, const double *x, double *q, double *part, double *dnorm,
                      double *res, double *rvec_out) {
  mvf_handle *h = S.h;
  spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, x, q, nullptr, nullptr);
  int nb = vec_blocks(S.n, NC);
  hipLaunchKernelGGL(k_resid<NC>, dim3(nb), dim3(kBS), 0, h->stream, S.n, b, q, rvec_out, part);
  hipLaunchKernelGGL(k_norms<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, dnorm);
  double hn[2 * NC];
  MVF_HIP(h, hipMemcpyAsync(hn, dnorm, sizeof(hn), hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  MVF_HIP(h, hipGetLastError());
  for (int c = 0; c < NC; c++) res[c] = hn[2 * c + 1] > 0 ? hn[2 * c] / hn[2 * c + 1] : 0.0;
  return 0;
}

template <int NC>
struct CocgBufs {
  double *r, *p, *q, *invd, *part, *dnorm;
  ColState *st;
};

template <int NC>
static int cocg_bufs(Workspace &W, mvf_handle *h, int64_t n, CocgBufs<NC> &B) {
  const int64_t nv = n * NC * 2;
  MVF_TRY(W.get(h, &B.r, nv));
  MVF_TRY(W.get(h, &B.p, nv));
  MVF_TRY(W.get(h, &B.q, nv));
  MVF_TRY(W.get(h, &B.invd, nv));
  MVF_TRY(W.get(h, &B.part, (size_t)kRedBlocks * 3 * NC * 2 + (size_t)spmv_blocks(n) * NC * 2));
  MVF_TRY(W.get(h, &B.dnorm, 4 * NC));
  MVF_TRY(W.get(h, &B.st, NC));
  return 0;
}

// the persistent COCG kernel for this solver's lane width and matrix: grid of
// co-resident blocks (cooperative launch requires it), at most as many as the
// partial-sum buffers hold
template <int NC, int LPR, bool C64>
static int cocg_persist_grid(mvf_handle *h, int64_t n, int &grid) {
  int per = 0, ncu = 0;
  MVF_HIP(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_cocg_persist<NC, LPR, C64>, kBS, 0));
  MVF_HIP(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
  int64_t g = (int64_t)per * ncu;
  // MVF_PERSIST_BLOCKS caps the grid: fewer blocks make the barriers and the
  // per-block reductions cheaper and leave fewer waves for the SpMV
  if (const char *e = getenv("MVF_PERSIST_BLOCKS")) {
    const long v = strtol(e, nullptr, 10);
    if (v > 0) g = std::min<int64_t>(g, v);
  }
  g = std::min<int64_t>(g, kRedBlocks);
  g = std::min<int64_t>(g, spmv_blocks(n));
  if (g < 1) return set_err(h, MVF_E_DEVICE, "persistent COCG: no co-resident blocks");
  grid = (int)g;
  return 0;
}

template <int NC, int LPR, bool C64>
static int cocg_persist_run(Solver &S, int grid, int iters, double *x, double *r, double *p, double *q,
                            const double *invd, double *spart, double *part, ColState *st, double tol) {
  mvf_handle *h = S.h;
  int64_t n = S.n;
  const int64_t *rowptr = h->d_rowptr;
  const int32_t *col = h->d_col;
  const double *K = h->d_K, *M = h->d_M;
  ColW w = S.w;
  void *args[] = {&iters, &n, &rowptr, &col, &K, &M, &w, &x, &r, &p, &q, &invd, &spart, &part, &st, &tol};
  MVF_HIP(h, hipLaunchCooperativeKernel((const void *)k_cocg_persist<NC, LPR, C64>, dim3(grid), dim3(kBS), args, 0,
                                        h->stream));
  return 0;
}

// dispatch over the autotuned lane width and the matrix rounding
template <int NC>
static int cocg_persist(Solver &S, bool grid_only, int &grid, int iters, double *x, double *r, double *p, double *q,
                        const double *invd, double *spart, double *part, ColState *st, double tol) {
#define MVF_PERSIST(LPR_, C64_)                                                                                    \
  return grid_only ? cocg_persist_grid<NC, LPR_, C64_>(S.h, S.n, grid)                                            \
                   : cocg_persist_run<NC, LPR_, C64_>(S, grid, iters, x, r, p, q, invd, spart, part, st, tol)
  if (S.c64) {
    switch (S.lpr) {
      case 4: MVF_PERSIST(4, true);
      case 8: MVF_PERSIST(8, true);
      case 32: MVF_PERSIST(32, true);
      default: MVF_PERSIST(16, true);
    }
  } else {
    switch (S.lpr) {
      case 4: MVF_PERSIST(4, false);
      case 8: MVF_PERSIST(8, false);
      case 32: MVF_PERSIST(32, false);
      default: MVF_PERSIST(16, false);
    }
  }
#undef MVF_PERSIST
}

template <int NC>
static int run_cocg64(Solver &S, const mvf_solver_opts *o, const double *b, double *x, mvf_solve_info *info) {
  mvf_handle *h = S.h;
  int64_t n = S.n, nv = n * NC * 2;
  Workspace W;
  W.base = S.ws;
  W.cap = S.ws_cap;
  CocgBufs<NC> B;
  MVF_TRY(cocg_bufs<NC>(W, h, n, B));
  // use_graph 1: replay check_every iterations from a HIP graph; 2: run them in
  // one cooperative launch of the persistent kernel
  const bool persist = o->use_graph == 2 && !S.ws && o->check_every > 0;
  const bool use_graph = o->use_graph == 1 && !S.ws && o->check_every > 0;
  GraphBlock gb;
  double *r = B.r, *p = B.p, *q = B.q, *invd = B.invd, *part = B.part, *dnorm = B.dnorm;
  ColState *st = B.st;
  double *spart = part + (size_t)kRedBlocks * 3 * NC * 2;
  int pgrid = 0;
  if (persist) MVF_TRY(cocg_persist<NC>(S, true, pgrid, 0, x, r, p, q, invd, spart, part, st, o->tol));
  if (o->precond == MVF_PC_JACOBI) {
    if (S.c64)
      hipLaunchKernelGGL((k_invdiag<NC, true>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, n, h->d_diagpos, h->d_K,
                         h->d_M, S.w, invd);
    else
      hipLaunchKernelGGL((k_invdiag<NC, false>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, n, h->d_diagpos,
                         h->d_K, h->d_M, S.w, invd);
  } else {
    hipLaunchKernelGGL(k_unit_diag, dim3(grid1d(nv / 2, 256)), dim3(256), 0, h->stream, nv, invd);
  }
  int nb = vec_blocks(n, NC), sb = spmv_blocks(n);
  int it = 0, restarts = 0;
  ColState hs[NC];
  double res[NC];
  bool done = false;
  int rc = 0;
  while (!done) {
    // (re)start from the current x: r = b - A x
    spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, x, q, nullptr, nullptr);
    hipLaunchKernelGGL((k_cocg_init<double, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, b, q, invd, r, p, part);
    hipLaunchKernelGGL(k_cocg_fin_init<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st);
    int since = 0;
    auto iter = [&]() {
      spmv_launch<double, double, NC, true>(S, h->d_K, h->d_M, p, q, p, spart);
      hipLaunchKernelGGL(k_cocg_alpha<NC>, dim3(1), dim3(1024), 0, h->stream, spart, sb, st);
      hipLaunchKernelGGL((k_cocg_upd1<double, double, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, st, x, r, p, q,
                         invd, part);
      hipLaunchKernelGGL(k_cocg_beta<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st, o->tol, 1);
      hipLaunchKernelGGL((k_cocg_upd2<double, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, st, r, p, invd);
    };
    while (it < o->maxit) {
      if (persist && since == 0) {
        const int k = std::min(o->check_every, o->maxit - it);
        MVF_TRY(cocg_persist<NC>(S, false, pgrid, k, x, r, p, q, invd, spart, part, st, o->tol));
        gb.launches++;
        it += k;
        since = o->check_every;
      } else if (use_graph && since == 0 && it + o->check_every <= o->maxit) {
        MVF_TRY(graph_block(S, gb, o->check_every, iter));
        it += o->check_every;
        since = o->check_every;
      } else {
        iter();
        it++;
        since++;
      }
      if (since >= o->check_every) {
        since = 0;
        MVF_TRY(host_state<NC>(h, hs, st));
        bool any = false;
        for (int c = 0; c < S.ncol; c++) any |= hs[c].active != 0;
        if (!any) break;
      }
    }
    MVF_TRY(true_resid<NC>(S, b, x, q, part, dnorm, res, nullptr));
    MVF_TRY(host_state<NC>(h, hs, st));
    bool ok = true, brk = false;
    for (int c = 0; c < S.ncol; c++) {
      ok &= res[c] <= o->tol;
      brk |= hs[c].breakdown != 0;
    }
    if (ok) {
      done = true;
    } else if (it >= o->maxit || restarts >= 20) {
      rc = brk ? set_err(h, MVF_E_BREAKDOWN, "COCG breakdown") : set_err(h, MVF_E_NOCONV, "COCG did not converge");
      done = true;
    } else {
      restarts++;  // recursive residual met tol but the true one did not, or breakdown: restart
    }
  }
  info->iterations = it;
  info->fallbacks = restarts;
  info->converged = rc == 0;
  info->graph_launches = gb.launches;
  for (int c = 0; c < S.ncol && c < 64; c++) info->true_resid[c] = res[c];
  return rc;
}

// mixed precision: float32 COCG recurrences, FP64 x and FP64 true residual replacements
template <int NC>
static int run_cocg_mixed(Solver &S, const mvf_solver_opts *o, const double *b, double *x, mvf_solve_info *info) {
  DevScope ds;
  mvf_handle *h = S.h;
  int64_t n = S.n, nv = n * NC * 2;
  float *r, *p, *q, *x32;
  double *r64, *q64, *invd, *part, *dnorm;
  ColState *st;
  MVF_TRY(ds.alloc(h, &r, nv));
  MVF_TRY(ds.alloc(h, &p, nv));
  MVF_TRY(ds.alloc(h, &q, nv));
  MVF_TRY(ds.alloc(h, &x32, nv));
  MVF_TRY(ds.alloc(h, &r64, nv));
  MVF_TRY(ds.alloc(h, &q64, nv));
  MVF_TRY(ds.alloc(h, &invd, nv));
  MVF_TRY(ds.alloc(h, &part, (size_t)kRedBlocks * 3 * NC * 2 + (size_t)spmv_blocks(n) * NC * 2));
  MVF_TRY(ds.alloc(h, &dnorm, 4 * NC));
  MVF_TRY(ds.alloc(h, &st, NC));
  double *spart = part + (size_t)kRedBlocks * 3 * NC * 2;
  if (S.c64)
    hipLaunchKernelGGL((k_invdiag<NC, true>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, n, h->d_diagpos, h->d_K,
                       h->d_M, S.w, invd);
  else
    hipLaunchKernelGGL((k_invdiag<NC, false>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, n, h->d_diagpos, h->d_K,
                       h->d_M, S.w, invd);
  MVF_HIP(h, hipMemsetAsync(x32, 0, nv * sizeof(float), h->stream));
  int nb = vec_blocks(n, NC), sb = spmv_blocks(n);
  int it = 0, refinements = 0;
  ColState hs[NC];
  double res[NC];
  int rc = 0;
  double drop = o->inner_tol > 0 ? o->inner_tol : 1e-2;
  // r64 = b - A x64 ; r = fl32(r64) ; start the recurrences
  MVF_TRY(true_resid<NC>(S, b, x, q64, part, dnorm, res, r64));
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid1d(nv, 256)), dim3(256), 0, h->stream, nv, r64, r);
  // q = 0 so that init computes r = b' - 0 with b' = r64
  MVF_HIP(h, hipMemsetAsync(q, 0, nv * sizeof(float), h->stream));
  hipLaunchKernelGGL((k_cocg_init<float, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, r64, q, invd, r, p, part);
  hipLaunchKernelGGL(k_cocg_fin_init<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st);
  MVF_TRY(host_state<NC>(h, hs, st));
  double rref[NC];
  for (int c = 0; c < NC; c++) rref[c] = hs[c].rnorm;
  double bnorm[NC];
  {
    // ||b|| for the tolerance (the init above saw r64 as "b")
    double r0[NC];
    MVF_TRY(true_resid<NC>(S, b, x, q64, part, dnorm, r0, nullptr));
    double hn[2 * NC];
    MVF_HIP(h, hipMemcpy(hn, dnorm, sizeof(hn), hipMemcpyDeviceToHost));
    for (int c = 0; c < NC; c++) bnorm[c] = hn[2 * c + 1];
  }
  int since = 0;
  bool all_ok = false;
  while (it < o->maxit) {
    spmv_launch<float, float, NC, true>(S, h->d_K32, h->d_M32, p, q, p, spart);
    hipLaunchKernelGGL(k_cocg_alpha<NC>, dim3(1), dim3(1024), 0, h->stream, spart, sb, st);
    hipLaunchKernelGGL((k_cocg_upd1<float, float, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, st, x32, r, p, q, invd,
                       part);
    hipLaunchKernelGGL(k_cocg_beta<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st, 0.0, 0);
    hipLaunchKernelGGL((k_cocg_upd2<float, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, st, r, p, invd);
    it++;
    bool replace = false;
    if (++since >= o->check_every) {
      since = 0;
      MVF_TRY(host_state<NC>(h, hs, st));
      for (int c = 0; c < S.ncol; c++)
        if (hs[c].rnorm <= drop * rref[c] || hs[c].breakdown) replace = true;
    }
    if (replace) {
      // reliable update: x += x32, r = fl32(b - A x), keep p, rho = r^T D^-1 r
      hipLaunchKernelGGL(k_accum, dim3(grid1d(nv, 256)), dim3(256), 0, h->stream, nv, x, x32);
      MVF_TRY(true_resid<NC>(S, b, x, q64, part, dnorm, res, r64));
      refinements++;
      all_ok = true;
      for (int c = 0; c < S.ncol; c++) all_ok &= res[c] <= o->tol;
      if (all_ok) break;
      hipLaunchKernelGGL((k_convert<double, float>), dim3(grid1d(nv, 256)), dim3(256), 0, h->stream, nv, r64, r);
      bool brk = false;
      for (int c = 0; c < S.ncol; c++) brk |= hs[c].breakdown != 0;
      if (brk) {
        // restart the recurrences from the true residual
        MVF_HIP(h, hipMemsetAsync(q, 0, nv * sizeof(float), h->stream));
        hipLaunchKernelGGL((k_cocg_init<float, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, r64, q, invd, r, p, part);
        hipLaunchKernelGGL(k_cocg_fin_init<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st);
        info->fallbacks++;
      } else {
        // recompute rho from the replaced residual; keep the search direction
        MVF_HIP(h, hipMemsetAsync(q, 0, nv * sizeof(float), h->stream));
        hipLaunchKernelGGL((k_cocg_init<float, NC>), dim3(nb), dim3(kBS), 0, h->stream, n, r64, q, invd, r, q, part);
        hipLaunchKernelGGL(k_cocg_fin_init<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st);
      }
      MVF_TRY(host_state<NC>(h, hs, st));
      for (int c = 0; c < NC; c++) rref[c] = hs[c].rnorm;
      // freeze converged columns
      (void)bnorm;
    }
  }
  if (!all_ok) {
    hipLaunchKernelGGL(k_accum, dim3(grid1d(nv, 256)), dim3(256), 0, h->stream, nv, x, x32);
    MVF_TRY(true_resid<NC>(S, b, x, q64, part, dnorm, res, nullptr));
    all_ok = true;
    for (int c = 0; c < S.ncol; c++) all_ok &= res[c] <= o->tol;
    if (!all_ok) rc = set_err(h, MVF_E_NOCONV, "mixed-precision COCG did not converge");
  }
  info->iterations = it;
  info->refinements = refinements;
  info->converged = rc == 0;
  for (int c = 0; c < S.ncol && c < 64; c++) info->true_resid[c] = res[c];
  ds.drop(r); ds.drop(p); ds.drop(q); ds.drop(x32); ds.drop(r64); ds.drop(q64); ds.drop(invd); ds.drop(part); ds.drop(dnorm); ds.drop(st);
  return rc;
}

template <int NC>
struct BicgBufs {
  double *r, *rh, *p, *v, *s, *t, *ph, *sh, *invd, *part, *dnorm;
  ColState *st;
};

template <int NC>
static int bicg_bufs(Workspace &W, mvf_handle *h, int64_t n, BicgBufs<NC> &B) {
  const int64_t nv = n * NC * 2;
  for (double **v : {&B.r, &B.rh, &B.p, &B.v, &B.s, &B.t, &B.ph, &B.sh, &B.invd}) MVF_TRY(W.get(h, v, nv));
  MVF_TRY(W.get(h, &B.part, (size_t)kRedBlocks * 3 * NC * 2));
  MVF_TRY(W.get(h, &B.dnorm, 4 * NC));
  MVF_TRY(W.get(h, &B.st, NC));
  return 0;
}

template <int NC>
static int run_bicgstab(Solver &S, const mvf_solver_opts *o, const double *b, double *x, mvf_solve_info *info) {
  mvf_handle *h = S.h;
  int64_t n = S.n, nv = n * NC * 2;
  Workspace W;
  W.base = S.ws;
  W.cap = S.ws_cap;
  BicgBufs<NC> B;
  MVF_TRY(bicg_bufs<NC>(W, h, n, B));
  const bool use_graph = o->use_graph && !S.ws && o->check_every > 0;
  GraphBlock gb;
  double *r = B.r, *rh = B.rh, *p = B.p, *v = B.v, *s = B.s, *t = B.t, *ph = B.ph, *sh = B.sh, *invd = B.invd;
  double *part = B.part, *dnorm = B.dnorm;
  ColState *st = B.st;
  if (o->precond == MVF_PC_JACOBI) {
    if (S.c64)
      hipLaunchKernelGGL((k_invdiag<NC, true>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, n, h->d_diagpos, h->d_K,
                         h->d_M, S.w, invd);
    else
      hipLaunchKernelGGL((k_invdiag<NC, false>), dim3(grid1d(n, 256)), dim3(256), 0, h->stream, n, h->d_diagpos,
                         h->d_K, h->d_M, S.w, invd);
  } else {
    hipLaunchKernelGGL(k_unit_diag, dim3(grid1d(nv / 2, 256)), dim3(256), 0, h->stream, nv, invd);
  }
  int nb = vec_blocks(n, NC);
  int it = 0, restarts = 0, rc = 0;
  ColState hs[NC];
  double res[NC];
  bool done = false;
  while (!done) {
    spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, x, s, nullptr, nullptr);
    hipLaunchKernelGGL(k_bicg_init<NC>, dim3(nb), dim3(kBS), 0, h->stream, n, b, s, r, rh, p, v, part);
    hipLaunchKernelGGL(k_bicg_fin_init<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st);
    // first beta with rho = alpha = omega = 1: k_bicg_ts(rhat, r) yields
    // [rhat^H r, rhat^H rhat] = [rho', |r|^2] in the layout k_bicg_beta reads
    hipLaunchKernelGGL(k_bicg_ts<NC>, dim3(nb), dim3(kBS), 0, h->stream, n, rh, r, part);
    hipLaunchKernelGGL(k_bicg_beta<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st, o->tol, 0);
    int since = 0;
    auto iter = [&]() {
      hipLaunchKernelGGL(k_bicg_p<NC>, dim3(nb), dim3(kBS), 0, h->stream, n, st, r, p, v, invd, ph);
      spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, ph, v, nullptr, nullptr);
      hipLaunchKernelGGL(k_hdot<NC>, dim3(nb), dim3(kBS), 0, h->stream, n, rh, v, part);
      hipLaunchKernelGGL(k_bicg_alpha<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st);
      hipLaunchKernelGGL(k_bicg_s<NC>, dim3(nb), dim3(kBS), 0, h->stream, n, st, r, v, s, invd, sh);
      spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, sh, t, nullptr, nullptr);
      hipLaunchKernelGGL(k_bicg_ts<NC>, dim3(nb), dim3(kBS), 0, h->stream, n, t, s, part);
      hipLaunchKernelGGL(k_bicg_omega<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st);
      hipLaunchKernelGGL(k_bicg_x<NC>, dim3(nb), dim3(kBS), 0, h->stream, n, st, x, r, s, t, ph, sh, rh, part);
      hipLaunchKernelGGL(k_bicg_beta<NC>, dim3(1), dim3(1024), 0, h->stream, part, nb, st, o->tol, 1);
    };
    while (it < o->maxit) {
      if (use_graph && since == 0 && it + o->check_every <= o->maxit) {
        MVF_TRY(graph_block(S, gb, o->check_every, iter));
        it += o->check_every;
        since = o->check_every;
      } else {
        iter();
        it++;
        since++;
      }
      if (since >= o->check_every) {
        since = 0;
        MVF_TRY(host_state<NC>(h, hs, st));
        bool any = false;
        for (int c = 0; c < S.ncol; c++) any |= hs[c].active != 0;
        if (!any) break;
      }
    }
    MVF_TRY(true_resid<NC>(S, b, x, s, part, dnorm, res, nullptr));
    MVF_TRY(host_state<NC>(h, hs, st));
    bool ok = true, brk = false;
    for (int c = 0; c < S.ncol; c++) {
      ok &= res[c] <= o->tol;
      brk |= hs[c].breakdown != 0;
    }
    if (ok) done = true;
    else if (it >= o->maxit || restarts >= 20) {
      rc = brk ? set_err(h, MVF_E_BREAKDOWN, "BiCGStab breakdown") : set_err(h, MVF_E_NOCONV, "BiCGStab did not converge");
      done = true;
    } else restarts++;
  }
  info->iterations = it;
  info->fallbacks = restarts;
  info->converged = rc == 0;
  info->graph_launches = gb.launches;
  for (int c = 0; c < S.ncol && c < 64; c++) info->true_resid[c] = res[c];
  return rc;
}

template <int NC>
static int run_gmres_direct(Solver &S, const mvf_solver_opts *o, const double *b, double *x, mvf_solve_info *info);

// Per column: acc[4c..4c+3] += (Re, Im) of y^H b, |y|^2, |b|^2
template <int NC>
__global__ void k_guess_dots(int64_t n, const double2 *__restrict__ y, const double2 *__restrict__ b, double *acc) {
  __shared__ double red[4 * NC][256];
  double s[4 * NC];
#pragma unroll
  for (int k = 0; k < 4 * NC; k++) s[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double2 yy = y[i * NC + c], bb = b[i * NC + c];
      s[4 * c] += yy.x * bb.x + yy.y * bb.y;
      s[4 * c + 1] += yy.x * bb.y - yy.y * bb.x;
      s[4 * c + 2] += yy.x * yy.x + yy.y * yy.y;
      s[4 * c + 3] += bb.x * bb.x + bb.y * bb.y;
    }
#pragma unroll
  for (int k = 0; k < 4 * NC; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
#pragma unroll
      for (int k = 0; k < 4 * NC; k++) red[k][threadIdx.x] += red[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 4 * NC) atomicAdd(&acc[threadIdx.x], red[threadIdx.x][0]);
}

template <int NC>
__global__ void k_scale_cols(int64_t n, double2 *x, double ar0, double ai0, double ar1, double ai1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * NC) return;
  const int c = (int)(i % NC);
  // columns beyond the first pair keep alpha of their polarization (column % 2)
  const double ar = (c & 1) ? ar1 : ar0, ai = (c & 1) ? ai1 : ai0;
  const double2 v = x[i];
  x[i] = make_double2(ar * v.x - ai * v.y, ar * v.y + ai * v.x);
}

// x <- alpha x per polarization with alpha = (Ax)^H b / |Ax|^2 on the operator
// the solve is judged on; writes alpha (re, im per polarization) to alpha_out
template <int NC>
static int guess_scale(Solver &S, const double *b, double *x, double *alpha_out) {
  mvf_handle *h = S.h;
  const int64_t n = S.n;
  DevScope ds;
  double *q, *acc;
  MVF_TRY(ds.alloc(h, &q, (size_t)n * NC * 2));
  MVF_TRY(ds.alloc(h, &acc, 4 * NC));
  spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, x, q, nullptr, nullptr);
  hipLaunchKernelGGL(k_fill_d, dim3(1), dim3(64), 0, h->stream, (int64_t)(4 * NC), acc, 0.0);
  hipLaunchKernelGGL(k_guess_dots<NC>, dim3(512), dim3(256), 0, h->stream, n, (const double2 *)q,
                     (const double2 *)b, acc);
  double ha[4 * NC];
  MVF_HIP(h, hipMemcpyAsync(ha, acc, sizeof(ha), hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  double al[2][2] = {{0, 0}, {0, 0}};
  for (int c = 0; c < 2 && c < NC; c++) {
    const double yy = ha[4 * c + 2];
    if (yy > 0) {
      al[c][0] = ha[4 * c] / yy;
      al[c][1] = ha[4 * c + 1] / yy;
    }
    if (alpha_out) {
      alpha_out[2 * c] = al[c][0];
      alpha_out[2 * c + 1] = al[c][1];
    }
  }
  hipLaunchKernelGGL(k_scale_cols<NC>, dim3(grid1d(n * NC, 256)), dim3(256), 0, h->stream, n, (double2 *)x,
                     al[0][0], al[0][1], al[1][0], al[1][1]);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

constexpr int nc_slot(int nc) { return nc == 2 ? 0 : (nc == 4 ? 1 : (nc == 8 ? 2 : 3)); }

// SpMV autotuner: the FP64 operator's lanes per row (4, 8, 16 or 32) timed
// once per handle and column count with HIP events (one warm-up and three
// timed launches each), the fastest kept in the handle.  Rows of 8-node
// systems hold <= 33 entries, 20-node ~84, 27-node ~104: the best width
// follows the row length and the column count.
template <int NC>
static int tune_lpr(Solver &S) {
  mvf_handle *h = S.h;
  int &slot = h->spmv_lpr[nc_slot(NC)];
  if (slot) {
    S.lpr = slot;
    return 0;
  }
  DevScope ds;
  double *xa, *ya;
  const int64_t nv = S.n * NC * 2;
  MVF_TRY(ds.alloc(h, &xa, nv));
  MVF_TRY(ds.alloc(h, &ya, nv));
  hipLaunchKernelGGL(k_fill_d, dim3(grid1d(nv, 256)), dim3(256), 0, h->stream, nv, xa, 1.0);
  hipEvent_t e0, e1;
  MVF_HIP(h, hipEventCreate(&e0));
  MVF_HIP(h, hipEventCreate(&e1));
  SpmvTimer *tm = S.tm;
  S.tm = nullptr;
  int best = 16;
  float best_ms = 1e30f;
  for (int L : {4, 8, 16, 32}) {
    S.lpr = L;
    spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, xa, ya, nullptr, nullptr);
    (void)hipEventRecord(e0, h->stream);
    for (int k = 0; k < 3; k++) spmv_launch<double, double, NC, false>(S, h->d_K, h->d_M, xa, ya, nullptr, nullptr);
    (void)hipEventRecord(e1, h->stream);
    float ms = 0.f;
    if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best_ms) {
      best_ms = ms;
      best = L;
    }
  }
  S.tm = tm;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MVF_HIP(h, hipGetLastError());
  slot = best;
  S.lpr = best;
  return 0;
}

template <int NC>
static int solve_nc(mvf_handle *h, const double *omegas, int ncol, const mvf_solver_opts *o, const double *b,
                    double *x, mvf_solve_info *info, double *alpha_out = nullptr) {
  SpmvTimer *tm = new SpmvTimer();
  tm->init();
  Solver S{h, ncol, {}, o->matrix == MVF_MATRIX_C64REF, 16, (int64_t)h->nne, tm};
  for (int c = 0; c < NC; c++) S.w.w[c] = c < ncol ? (double)(float)omegas[c] : 0.0;
  int rc = h->autotune ? tune_lpr<NC>(S) : 0;
  if (!rc && alpha_out) rc = guess_scale<NC>(S, b, x, alpha_out);
  if (rc) {
  } else if (o->precond == MVF_PC_DIRECT) rc = run_gmres_direct<NC>(S, o, b, x, info);
  else if (o->precision == MVF_PREC_MIXED) rc = run_cocg_mixed<NC>(S, o, b, x, info);
  else if (o->method == MVF_KRYLOV_BICGSTAB) rc = run_bicgstab<NC>(S, o, b, x, info);
  else rc = run_cocg64<NC>(S, o, b, x, info);
  tm->destroy();
  const double n = (double)h->nne, nnz = (double)h->nnz;
  for (int k = 0; k < 2; k++) {
    info->spmv_launches[k] = tm->launches[k];
    info->spmv_ms[k] = tm->ms[k];
  }
  // algorithmic bytes: K, M and column index per entry, row pointer per row,
  // x read once and y written once per column
  info->spmv_bytes[0] = nnz * (8 + 8 + 4) + n * (8 + NC * (16 + 16));
  info->spmv_bytes[1] = nnz * (4 + 4 + 4) + n * (8 + NC * (8 + 8));
  info->spmv_ncol = NC;
  info->spmv_lpr = S.lpr;
  delete tm;
  return rc;
}

// b, x: device [nne][ncol_pad] with ncol_pad in {2,4,8,16}
int solve_device(mvf_handle *h, const double *omegas, int ncol, int ncol_pad, const mvf_solver_opts *o,
                 const double *b, double *x, mvf_solve_info *info) {
  switch (ncol_pad) {
    case 2: return solve_nc<2>(h, omegas, ncol, o, b, x, info);
    case 4: return solve_nc<4>(h, omegas, ncol, o, b, x, info);
    case 8: return solve_nc<8>(h, omegas, ncol, o, b, x, info);
    case 16: return solve_nc<16>(h, omegas, ncol, o, b, x, info);
  }
  return set_err(h, MVF_E_ARG, "unsupported column count %d", ncol_pad);
}

// the same, starting from the caller's x scaled per polarization by the
// least-squares factor (two columns)
int solve_device_guess(mvf_handle *h, const double *omegas, const mvf_solver_opts *o, const double *b, double *x,
                       mvf_solve_info *info, double *alpha_out) {
  double a4[4];
  return solve_nc<2>(h, omegas, 2, o, b, x, info, alpha_out ? alpha_out : a4);
}

}  // namespace mvf
