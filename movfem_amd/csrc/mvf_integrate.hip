// Element integration on the device.
//
// k_elem_km: frequency-independent element matrices.  The reference forms
//   alocal(i,j) = sum_g w_g (f1 + i fl32(omega) f2)        (integration.f90:76-86)
// with f1 the real part of the curl-curl integrand and f2 the real part of the
// mass integrand (:154-238).  Both parts are real and frequency independent,
// so one pass yields K_ij = sum_g w_g f1 and M_ij = sum_g w_g f2 and
// alocal = K + i fl32(omega) M for every frequency.
//
// Under GPML the reference weights each of the 36 curl products by a product
// of the (real) stretch factors h (:171-188).  That weight table factors as
// W = g g^T with g(a,1), g(a,2) = sqrt of h1h3/h2, h1h2/h3, h2h3/h1 cyclically,
// so f1 = chat_i^T m chat_j with chat(a) = g(a,1) curl1(a) - g(a,2) curl2(a);
// f2 gains the factor h1 h2 h3 (:228-232).
//
// One workgroup per element: node data, Jacobians and interpolated tensors at
// the Gauss points are computed once (the reference recomputes the Jacobian
// for every basis call, v_fem.f90:475-476,516-517), the per-(point, DOF)
// vectors chat, ve, m1*chat, m2*ve are staged in LDS for a chunk of points, and
// each thread accumulates a 3 x 3 tile of the upper (i <= j) pairs over the
// points in registers.
//
// k_elem_rhs: the per-frequency element RHS blocal (integration.f90:88-106)
// with the secondary-field source of p_source (problem.f90:91-128), the air
// primary field of p_pfields (:340-358) and pset_pmodel's float32 constants.
#include "mvf_internal.h"

namespace mvf {

// lattice position (1-based) of element e = ((i-1) ny + j-1) nz + k-1; mvf_init
// bounds the element count by 2^31 - 1, so two 32-bit divisions suffice
__device__ inline void elem_index(const Geom &g, int64_t e, int &i, int &j, int &k) {
  const uint32_t ue = (uint32_t)e, nz = (uint32_t)g.nz, ny = (uint32_t)g.ny;
  const uint32_t col = ue / nz, a = col / ny;
  k = (int)(ue - col * nz) + 1;
  j = (int)(col - a * ny) + 1;
  i = (int)a + 1;
}

// GPML flags used for element e at lattice position (i, j, k)
// (MVF_EMU_GPML_PREV_FLAGS reproduces that the reference's int_elem_params
// reads the flags get_pml left for the previous element,
// MoVFEM_3DMT.f90:203-206; that element is one step back along k, wrapping
// to j and then i)
__device__ inline int elem_flags(const Geom &g, const Gpml &gp, int prev, int first_flags, int64_t e, int i,
                                 int j, int k) {
  if (!gp.on) return 0;
  if (prev) {
    if (e == 0) return first_flags;
    if (k > 1) {
      k--;
    } else {
      k = g.nz;
      if (j > 1) {
        j--;
      } else {
        j = g.ny;
        i--;
      }
    }
  }
  return zone_flags(gp, i, j, k);
}

__device__ inline void load_node(const BasisTab *tab, const Geom &g, int i, int j, int k, int n, int64_t &id,
                                 int &ix, int &jy) {
  int o = g.nord - 1;
  int i1 = (tab->nref[n][0] + 1) * o / 2, j1 = (tab->nref[n][1] + 1) * o / 2, k1 = (tab->nref[n][2] + 1) * o / 2;
  int64_t eno = (int64_t)(i - 1) * g.nyz * o + (int64_t)(j - 1) * g.nnz * o + (int64_t)(k - 1) * o;
  id = eno + (int64_t)i1 * g.nyz + (int64_t)j1 * g.nnz + k1;
  ix = (i - 1) * o + i1;
  jy = (j - 1) * o + j1;
}

// Jacobian at point g from node coordinates; returns det, writes J^-1 / |det|
template <int NN>
__device__ inline double jacobian(const BasisTab *tab, int gpt, const double (*xs)[3], double *Ji) {
  double J[3][3];
  for (int m = 0; m < 3; m++)
    for (int n = 0; n < 3; n++) {
      double acc = 0.0;
      for (int l = 0; l < NN; l++) acc += tab->dN[gpt][m][l] * xs[l][n];
      J[m][n] = acc;
    }
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double ad = fabs(det);
  Ji[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / ad;
  Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / ad;
  Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / ad;
  Ji[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / ad;
  Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / ad;
  Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / ad;
  Ji[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / ad;
  Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / ad;
  Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / ad;
  return det;
}

__device__ inline void sym6_mul(const double *m, const double *v, double *o) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[1] * v[0] + m[3] * v[1] + m[4] * v[2];
  o[2] = m[2] * v[0] + m[4] * v[1] + m[5] * v[2];
}

// jacobian in compute type T (J^-1 / |det|); returns det
template <typename T, int NN>
__device__ inline T jacobian_t(const BasisTab *tab, int gpt, const T (*xs)[3], T *Ji) {
  T J[3][3];
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int n = 0; n < 3; n++) {
      T acc = 0;
#pragma unroll 4
      for (int l = 0; l < NN; l++) acc += (T)tab->dN[gpt][m][l] * xs[l][n];
      J[m][n] = acc;
    }
  T det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  T ad = det < 0 ? -det : det;
  Ji[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / ad;
  Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / ad;
  Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / ad;
  Ji[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / ad;
  Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / ad;
  Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / ad;
  Ji[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / ad;
  Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / ad;
  Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / ad;
  return det;
}

template <typename T>
__device__ inline void sym6_mul_t(const T *m, const T *v, T *o) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[1] * v[0] + m[3] * v[1] + m[4] * v[2];
  o[2] = m[2] * v[0] + m[4] * v[1] + m[5] * v[2];
}

// T = double: the FP64 path; T = float: FP32 element integration (the GPML
// stretch, its square-root weights and the element position stay FP64)
template <typename T, int NN, int NE, int NG, int GC>
__global__ __launch_bounds__(256) void k_elem_km(const BasisTab *__restrict__ tab, const double *__restrict__ xp,
                                                 const double *__restrict__ yp, const double *__restrict__ zp,
                                                 const double *__restrict__ inmu, const double *__restrict__ sig,
                                                 Geom g, Gpml gp, int prev, int first_flags, int64_t e0,
                                                 T *__restrict__ Kout, T *__restrict__ Mout, int *bad) {
  constexpr int NP = NE * (NE + 1) / 2;
  // pairs in 3 x 3 tiles (i in 3bi..3bi+2, j in 3bj..3bj+2, bi <= bj): a tile
  // reads six values of each of its three row DOFs and three column DOFs per
  // point for 54 products, a third of the LDS traffic of one pair per thread
  constexpr int TE = NE / 3, NT = TE * (TE + 1) / 2;
  static_assert(NE % 3 == 0 && NT <= 256, "one 3 x 3 tile per thread");
  const int64_t e = e0 + blockIdx.x;
  const int t = threadIdx.x;
  __shared__ double xs64[NN][3];
  __shared__ T xs[NN][3];
  __shared__ T nm[NN][12];
  __shared__ T sJi[NG][9];
  __shared__ T sm1[NG][6], sm2[NG][6], sgf[NG][6];
  __shared__ T sv[GC][NE][12];
  int i, j, k;
  elem_index(g, e, i, j, k);
  if (t < NN) {
    int64_t id;
    int ix, jy;
    load_node(tab, g, i, j, k, t, id, ix, jy);
    xs64[t][0] = xp[ix];
    xs64[t][1] = yp[jy];
    xs64[t][2] = zp[id];
    for (int c = 0; c < 3; c++) xs[t][c] = (T)xs64[t][c];
    for (int c = 0; c < 6; c++) {
      nm[t][c] = (T)inmu[id * 6 + c];
      nm[t][6 + c] = (T)sig[2 * (id * 6 + c)];
    }
  }
  __syncthreads();
  if (t < NG) {
    T Ji[9];
    T det = jacobian_t<T, NN>(tab, t, xs, Ji);
    if (det == (T)0) atomicExch(bad, 1);
    T wgt = det * (T)tab->w[t];
    T m1[6], m2[6];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      T a1 = 0, a2 = 0;
#pragma unroll 4
      for (int n = 0; n < NN; n++) {
        a1 += (T)tab->N[t][n] * nm[n][c];
        a2 += (T)tab->N[t][n] * nm[n][6 + c];
      }
      m1[c] = a1;
      m2[c] = a2;
    }
    double h[3] = {1.0, 1.0, 1.0};
    if (gp.on) {
      int fl = elem_flags(g, gp, prev, first_flags, e, i, j, k);
      double r[3] = {0, 0, 0};
      for (int n = 0; n < NN; n++)
        for (int c = 0; c < 3; c++) r[c] += tab->N[t][n] * xs64[n][c];
      for (int c = 0; c < 3; c++) h[c] = gpml_stretch(gp, fl, c, r[c]);
    }
    double h3 = h[0] * h[1] * h[2];
    double q0 = sqrt(h[0] * h[2] / h[1]), q1 = sqrt(h[0] * h[1] / h[2]), q2 = sqrt(h[1] * h[2] / h[0]);
    sgf[t][0] = (T)q0; sgf[t][1] = (T)q1;
    sgf[t][2] = (T)q1; sgf[t][3] = (T)q2;
    sgf[t][4] = (T)q2; sgf[t][5] = (T)q0;
    for (int c = 0; c < 6; c++) {
      sm1[t][c] = wgt * m1[c];
      sm2[t][c] = (T)((double)wgt * h3) * m2[c];
    }
    for (int c = 0; c < 9; c++) sJi[t][c] = Ji[c];
  }
  __syncthreads();
  int bi = 0, bj = 0;
  {
    int q = t < NT ? t : 0;  // row-major over the upper tiles
    while (q >= TE - bi) {
      q -= TE - bi;
      bi++;
    }
    bj = bi + q;
  }
  T kacc[3][3], macc[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) kacc[r][c] = macc[r][c] = 0;
  for (int g0 = 0; g0 < NG; g0 += GC) {
    const int gcount = (NG - g0) < GC ? (NG - g0) : GC;
    for (int q = t; q < gcount * NE; q += 256) {
      int gl = q / NE, ed = q - gl * NE, gg = g0 + gl;
      const T *Ji = sJi[gg];
      int d = tab->dir[ed];
      T v[3] = {Ji[0 * 3 + d], Ji[1 * 3 + d], Ji[2 * 3 + d]};
      T dp0 = (T)tab->dphi[gg][ed][0], dp1 = (T)tab->dphi[gg][ed][1], dp2 = (T)tab->dphi[gg][ed][2];
      T dn[3];
      for (int a = 0; a < 3; a++) dn[a] = Ji[a * 3 + 0] * dp0 + Ji[a * 3 + 1] * dp1 + Ji[a * 3 + 2] * dp2;
      const T *gf = sgf[gg];
      T ch[3] = {gf[0] * (dn[1] * v[2]) - gf[1] * (dn[2] * v[1]), gf[2] * (dn[2] * v[0]) - gf[3] * (dn[0] * v[2]),
                 gf[4] * (dn[0] * v[1]) - gf[5] * (dn[1] * v[0])};
      T ph = (T)tab->phi[gg][ed];
      T ve[3] = {ph * v[0], ph * v[1], ph * v[2]};
      T tk[3], tm[3];
      sym6_mul_t<T>(sm1[gg], ch, tk);
      sym6_mul_t<T>(sm2[gg], ve, tm);
      T *o = sv[gl][ed];
      o[0] = ch[0]; o[1] = ch[1]; o[2] = ch[2];
      o[3] = ve[0]; o[4] = ve[1]; o[5] = ve[2];
      o[6] = tk[0]; o[7] = tk[1]; o[8] = tk[2];
      o[9] = tm[0]; o[10] = tm[1]; o[11] = tm[2];
    }
    __syncthreads();
    if (t < NT)
      for (int gl = 0; gl < gcount; gl++) {
        T A[3][6], B[3][6];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 6; c++) {
            A[r][c] = sv[gl][3 * bi + r][c];
            B[r][c] = sv[gl][3 * bj + r][6 + c];
          }
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) {
            kacc[r][c] += A[r][0] * B[c][0] + A[r][1] * B[c][1] + A[r][2] * B[c][2];
            macc[r][c] += A[r][3] * B[c][3] + A[r][4] * B[c][4] + A[r][5] * B[c][5];
          }
      }
    __syncthreads();
  }
  const int64_t base = (e - e0) * (int64_t)NP;
  if (t < NT)
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const int ii = 3 * bi + r, jj = 3 * bj + c;
        if (ii <= jj) {
          const int p = ii * NE - ii * (ii - 1) / 2 + (jj - ii);  // row-major upper packing
          Kout[base + p] = kacc[r][c];
          Mout[base + p] = macc[r][c];
        }
      }
}

template <typename T>
static int lane_table(mvf_handle *h, const BasisTab *tab, LaneTab<T> &lt) {
  const int NN = h->et.nn, NE = h->et.ne, NG = h->et.ng;
  const int NT = (NE + 15) / 16, NU = (NG + 3) / 4;
  const size_t n4 = (size_t)NU * NT * 64, ndN = (size_t)3 * NN * 32, nN = (size_t)NN * 32;
  std::vector<T> v(4 * n4 + ndN + nN + 32, (T)0);
  std::vector<int> dir((size_t)NT * 16, 0);
  T *d0 = v.data(), *d1 = d0 + n4, *d2 = d1 + n4, *p = d2 + n4, *dN = p + n4, *N = dN + ndN, *w = N + nN;
  for (int u = 0; u < NU; u++)
    for (int t = 0; t < NT; t++)
      for (int l = 0; l < 64; l++) {
        const int g = 4 * u + (l >> 4), e = 16 * t + (l & 15);
        if (g >= NG || e >= NE) continue;
        const size_t i = ((size_t)u * NT + t) * 64 + l;
        d0[i] = (T)tab->dphi[g][e][0];
        d1[i] = (T)tab->dphi[g][e][1];
        d2[i] = (T)tab->dphi[g][e][2];
        p[i] = (T)tab->phi[g][e];
      }
  for (int t = 0; t < NT; t++)
    for (int r = 0; r < 16; r++)
      if (16 * t + r < NE) dir[(size_t)t * 16 + r] = tab->dir[16 * t + r];
  for (int g = 0; g < NG; g++) {
    for (int n = 0; n < NN; n++) {
      for (int m = 0; m < 3; m++) dN[((size_t)m * NN + n) * 32 + g] = (T)tab->dN[g][m][n];
      N[(size_t)n * 32 + g] = (T)tab->N[g][n];
    }
    w[g] = (T)tab->w[g];
  }
  char *dev;
  const size_t bytes = v.size() * sizeof(T), dbytes = dir.size() * sizeof(int);
  MVF_TRY(dalloc(h, &dev, bytes + dbytes));
  h->owned.push_back(dev);
  MVF_HIP(h, hipMemcpy(dev, v.data(), bytes, hipMemcpyHostToDevice));
  MVF_HIP(h, hipMemcpy(dev + bytes, dir.data(), dbytes, hipMemcpyHostToDevice));
  const T *b = (const T *)dev;
  lt.d0 = b;
  lt.d1 = b + n4;
  lt.d2 = b + 2 * n4;
  lt.p = b + 3 * n4;
  lt.dN = b + 4 * n4;
  lt.N = lt.dN + ndN;
  lt.w = lt.N + nN;
  lt.dir = (const int *)(dev + bytes);
  return 0;
}

int lane_tables(mvf_handle *h, const BasisTab *tab) {
  MVF_TRY(lane_table<double>(h, tab, h->lt64));
  MVF_TRY(lane_table<float>(h, tab, h->lt32));
  return 0;
}

// ---------------------------------------------------------------------------
// k_elem_km_mfma: the same K_e and M_e on the matrix cores, one wavefront per
// element (four elements per 256-thread workgroup).
//
// With chat_i(g, a) the modified curl and v_i(g, a) the value vector of DOF i
// at Gauss point g (component a), the element matrices are the products
//   K = C^T Q,  C[(g,a)][i] = chat_i(g, a),  Q[(g,a)][j] = (w_g m1_g chat_j(g))_a
//   M = V^T P,  V[(g,a)][i] = v_i(g, a),     P[(g,a)][j] = (w_g h1h2h3 m2_g v_j(g))_a
// of depth 3G, computed as 16 x 16 tiles with v_mfma_{f64,f32}_16x16x4.  The
// depth index is ordered (u, a, q) with g = 4u + q: in an MFMA step the lanes
// l with l / 16 = q supply depth entry q, so lane l holds, for every step of
// one u and all three a, the entries of Gauss point 4u + l/16 and DOFs
// 16 t + l%16.  Only tiles on or above the diagonal are formed; K first, then M.
//
// Operands per lane.  DOF i is N_i = phi_i grad(xi_d) with d = d_i its edge
// direction, so with u = J^-1 grad_ref(phi_i) and v = J^-1 e_d (column d of
// J^-1) the GPML-weighted curl is chat_a = q_a1 u_b v_c - q_a2 u_c v_b
// ((a, b, c) cyclic): a linear map of grad_ref(phi_i),
//   chat = Md(g) grad_ref(phi_i),  Md[a][m] = q_a1 v_c Ji[b][m] - q_a2 v_b Ji[c][m],
// whose 3 x 3 matrix depends only on the point and the direction.  The setup
// forms Md for the three directions at every point into LDS; a lane then
// reads the matrix of its DOF's direction (a per-lane LDS address, no
// selects), and its operands cost one 3 x 3 product each for C and for Q.
// The mass operands are V = phi v_d and P = phi u_d with u_d = (w h1h2h3 m2)
// v_d formed per point in the setup (the 20-node FP32 kernel applies the 3 x 3
// w h1h2h3 m2 per DOF instead, see PtRec).
// The reference-element tables are zero outside the element's points and
// DOFs, so padded rows and columns vanish without masking.
// MFMA lane layouts (16 x 16 x 4): A lane l -> A[l%16][l/16], B lane l ->
// B[l/16][l%16]; D register v of lane l -> D[row][l%16] with row = l/16 + 4v
// (f64) or 4 (l/16) + v (f32).
template <typename T>
__device__ inline int mfma_drow(int l, int v);
template <>
__device__ inline int mfma_drow<double>(int l, int v) {
  return (l >> 4) + 4 * v;
}
template <>
__device__ inline int mfma_drow<float>(int l, int v) {
  return 4 * (l >> 4) + v;
}

typedef double mvf_e4d __attribute__((ext_vector_type(4)));
typedef float mvf_e4f __attribute__((ext_vector_type(4)));
template <typename T>
struct Acc4;
template <>
struct Acc4<double> {
  typedef mvf_e4d type;
  static __device__ inline type mma(double a, double b, type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Acc4<float> {
  typedef mvf_e4f type;
  static __device__ inline type mma(float a, float b, type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// per-point record in LDS: Md for d = 0, 1, 2 (27), J^-1 columns v_d (9),
// w m1 (6), then u_d = (w h1h2h3 m2) v_d for d = 0, 1, 2 (9), so that the
// mass operand of a DOF is phi u_d (three products instead of a 3 x 3 one).
// The 20-node FP32 kernel keeps w h1h2h3 m2 itself (6) instead (measured: u_d
// makes the FP64 kernels 6-12 % faster and the 20-node FP32 one 11 % slower,
// with four or with three waves per workgroup under its 80-register budget).
template <typename T, int NN>
struct PtRec {
  static constexpr bool ud = !(sizeof(T) == 4 && NN == 20);
  static constexpr int n = ud ? 51 : 48;
};
constexpr int kPtV = 27, kPtM1 = 36, kPtM2 = 42;

// Stores the i <= j entries of the 16 x 16 tiles acc (tile pairs t <= t2 in
// order) into one element's packed upper triangle, row-major:
// (ii, jj) -> ii NE - ii (ii - 1) / 2 + jj - ii.  ob is the element's
// (wave-uniform) base.  Lane l holds rows ii = c + s, c = 16 t + drow(0, v)
// known at compile time and s = drow(l, 0), so the offset splits into a
// constant, one per-lane term and -c s.
template <typename T, int NE, int NT, typename A4>
__device__ inline void store_upper(T *__restrict__ ob, const A4 *acc, int l) {
  const int r = l & 15, s = mfma_drow<T>(l, 0);
  const int lo = s * (NE - 1) - s * (s - 1) / 2 + r, dq = r - s;
  int p = 0;
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int t2 = t; t2 < NT; t2++, p++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = 16 * t + mfma_drow<T>(0, v);
        const bool upper = t2 > t || dq >= c - 16 * t2;           // ii <= jj
        const bool inside = 16 * t2 + 16 <= NE || r < NE - 16 * t2;  // jj < NE
        if (upper && inside) ob[(uint32_t)(c * (NE - 1) - c * (c - 1) / 2 + 16 * t2 + lo - c * s)] = acc[p][v];
      }
}

template <typename T, int NN, int NE, int NG>
__global__ __launch_bounds__(256, sizeof(T) == 4 && NN == 20 ? 6 : 2) void k_elem_km_mfma(
    const BasisTab *__restrict__ tab, LaneTab<T> lt, const double *__restrict__ xp, const double *__restrict__ yp,
    const double *__restrict__ zp, const double *__restrict__ inmu, const double *__restrict__ sig, Geom g, Gpml gp,
    int prev, int first_flags, int64_t e0, int64_t n, T *__restrict__ Kout, T *__restrict__ Mout, int *bad) {
  constexpr int WPB = 4;                 // elements (waves) per workgroup
  constexpr int NT = (NE + 15) / 16;     // DOF tiles
  constexpr int NU = (NG + 3) / 4;       // Gauss-point groups of four
  constexpr int NPAIR = NT * (NT + 1) / 2;
  constexpr int NP = NE * (NE + 1) / 2;
  typedef typename Acc4<T>::type A4;
  // the wave index is uniform: taking it from lane 0 keeps the element's
  // index arithmetic scalar
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  const int64_t le = xcd_block(blockIdx.x, gridDim.x) * WPB + wv;  // element index within the launch
  const bool live = le < n;
  const int64_t e = e0 + (live ? le : 0);
  __shared__ double xs64[WPB][NN][3];
  __shared__ T nm[WPB][NN][12];
  __shared__ T gpd[WPB][NG][PtRec<T, NN>::n];
  int i, j, k;
  elem_index(g, e, i, j, k);
  if (l < NN) {
    int64_t id;
    int ix, jy;
    load_node(tab, g, i, j, k, l, id, ix, jy);
    xs64[wv][l][0] = xp[ix];
    xs64[wv][l][1] = yp[jy];
    xs64[wv][l][2] = zp[id];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      nm[wv][l][c] = (T)inmu[id * 6 + c];
      nm[wv][l][6 + c] = (T)sig[2 * (id * 6 + c)];
    }
  }
  // every array above is private to its wave: a wave barrier orders the
  // stores before the reads of other lanes
  __builtin_amdgcn_wave_barrier();
  {
    // Point setup on the whole wave: lane l takes point pt = l & 31, in half
    // hf = l >> 5.  The halves sum over complementary node sets in one
    // instruction stream (the Jacobian and the point coordinates over nodes
    // [hf NH, hf NH + NH), mu^-1 and Re sigma over the other half) and then
    // add the partial sums of lanes l and l ^ 32; a + b = b + a, so both
    // halves hold the same totals and write the same values.  Lanes with
    // pt >= NG read the tables' zero padding and write nothing.
    constexpr int NH = (NN + 1) / 2;
    static_assert(NG <= 32, "one point per lane of each half wave");
    const int hf = l >> 5, pt = l & 31;
    const bool on = pt < NG;
    const int pc = on ? pt : 0;
    const int nj0 = hf * NH, nm0 = NH - hf * NH;
    // GPML stretch of the point first, while little else is live
    double h[3] = {1.0, 1.0, 1.0};
    if (gp.on) {
      int fl = elem_flags(g, gp, prev, first_flags, e, i, j, k);
      double rr[3] = {0, 0, 0};
      for (int s = 0; s < NH; s++) {
        const int m = nj0 + s;
        if (NN % 2 == 0 || m < NN)
          for (int c = 0; c < 3; c++) rr[c] += tab->N[pc][m] * xs64[wv][m][c];
      }
      for (int c = 0; c < 3; c++) rr[c] += __shfl_xor(rr[c], 32);
      for (int c = 0; c < 3; c++) h[c] = gpml_stretch(gp, fl, c, rr[c]);
    }
    const double h3 = h[0] * h[1] * h[2];
    // GPML curl weights: component a takes q_a1 on u_b v_c and q_a2 on u_c v_b
    const T q0 = (T)sqrt(h[0] * h[2] / h[1]), q1 = (T)sqrt(h[0] * h[1] / h[2]), q2 = (T)sqrt(h[1] * h[2] / h[0]);
    // Jacobian from the point-fastest derivative table (coalesced across lanes)
    T J[3][3];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int c = 0; c < 3; c++) J[m][c] = 0;
#pragma unroll 2
    for (int s = 0; s < NH; s++) {
      const int m = nj0 + s;
      if (NN % 2 == 0 || m < NN) {
        const T x0 = (T)xs64[wv][m][0], x1 = (T)xs64[wv][m][1], x2 = (T)xs64[wv][m][2];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const T dn = lt.dN[(a * NN + m) * 32 + pt];
          J[a][0] += dn * x0;
          J[a][1] += dn * x1;
          J[a][2] += dn * x2;
        }
      }
    }
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int c = 0; c < 3; c++) J[m][c] += __shfl_xor(J[m][c], 32);
    T det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const T ad = det < 0 ? -det : det;
    T Ji[9];
    Ji[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / ad;
    Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / ad;
    Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / ad;
    Ji[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / ad;
    Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / ad;
    Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / ad;
    Ji[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / ad;
    Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / ad;
    Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / ad;
    if (live && on && det == (T)0) atomicExch(bad, 1);
    T wgt = det * lt.w[pt];
    T m1[6], m2[6];
#pragma unroll
    for (int c = 0; c < 6; c++) m1[c] = m2[c] = 0;
#pragma unroll 2
    for (int s = 0; s < NH; s++) {
      const int m = nm0 + s;
      if (NN % 2 == 0 || m < NN) {
        const T Nm = lt.N[m * 32 + pt];
#pragma unroll
        for (int c = 0; c < 6; c++) {
          m1[c] += Nm * nm[wv][m][c];
          m2[c] += Nm * nm[wv][m][6 + c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
      m1[c] += __shfl_xor(m1[c], 32);
      m2[c] += __shfl_xor(m2[c], 32);
    }
    if (on) {
      const T qa1[3] = {q0, q1, q2}, qa2[3] = {q1, q2, q0};
      T *o = gpd[wv][pt];
      const T wh = (T)((double)wgt * h3);
      T w2[6];
#pragma unroll
      for (int c = 0; c < 6; c++) w2[c] = wh * m2[c];
      // direction maps: half hf writes direction hf, then both halves direction 2
#pragma unroll
      for (int it = 0; it < 2; it++) {
        const int d = it == 0 ? hf : 2;
        T v[3];
#pragma unroll
        for (int a = 0; a < 3; a++) v[a] = it == 0 ? (hf ? Ji[a * 3 + 1] : Ji[a * 3]) : Ji[a * 3 + 2];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const int b = (a + 1) % 3, c = (a + 2) % 3;
#pragma unroll
          for (int m = 0; m < 3; m++)
            o[d * 9 + a * 3 + m] = qa1[a] * v[c] * Ji[b * 3 + m] - qa2[a] * v[b] * Ji[c * 3 + m];
          o[kPtV + d * 3 + a] = v[a];
        }
        if constexpr (PtRec<T, NN>::ud) {
          o[kPtM2 + d * 3 + 0] = w2[0] * v[0] + w2[1] * v[1] + w2[2] * v[2];
          o[kPtM2 + d * 3 + 1] = w2[1] * v[0] + w2[3] * v[1] + w2[4] * v[2];
          o[kPtM2 + d * 3 + 2] = w2[2] * v[0] + w2[4] * v[1] + w2[5] * v[2];
        }
      }
#pragma unroll
      for (int c = 0; c < 6; c++) {
        o[kPtM1 + c] = wgt * m1[c];
        if constexpr (!PtRec<T, NN>::ud) o[kPtM2 + c] = w2[c];
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  // this lane's DOF direction in each tile (0 for padding DOFs, whose table
  // entries are zero)
  int dr[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) dr[t] = lt.dir[t * 16 + r];
  const int64_t base = le * (int64_t)NP;
  auto store = [&](T *out, const A4 *acc) {
    if (live) store_upper<T, NE, NT>(out + base, acc, l);
  };
  {  // K = C^T Q
    A4 acc[NPAIR];
#pragma unroll
    for (int p = 0; p < NPAIR; p++) acc[p] = A4{0, 0, 0, 0};
#pragma unroll 2
    for (int u = 0; u < NU; u++) {
      const int gg = 4 * u + q;
      const T *pd = gpd[wv][gg < NG ? gg : 0];
      T mw[6];
#pragma unroll
      for (int c = 0; c < 6; c++) mw[c] = pd[kPtM1 + c];
      // tiles from the last: the products of tile t need the Q columns of
      // tiles t..NT-1 only, so each tile's C columns die after its MFMAs
      T bv[NT][3];
#pragma unroll
      for (int t = NT - 1; t >= 0; t--) {
        const int li = (u * NT + t) * 64 + l;
        const T dp0 = lt.d0[li], dp1 = lt.d1[li], dp2 = lt.d2[li];
        const T *md = pd + dr[t] * 9;
        T av[3];
        av[0] = md[0] * dp0 + md[1] * dp1 + md[2] * dp2;
        av[1] = md[3] * dp0 + md[4] * dp1 + md[5] * dp2;
        av[2] = md[6] * dp0 + md[7] * dp1 + md[8] * dp2;
        bv[t][0] = mw[0] * av[0] + mw[1] * av[1] + mw[2] * av[2];
        bv[t][1] = mw[1] * av[0] + mw[3] * av[1] + mw[4] * av[2];
        bv[t][2] = mw[2] * av[0] + mw[4] * av[1] + mw[5] * av[2];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int t2 = t; t2 < NT; t2++) {
            const int p = t * NT - t * (t - 1) / 2 + (t2 - t);
            acc[p] = Acc4<T>::mma(av[a], bv[t2][a], acc[p]);
          }
      }
    }
    store(Kout, acc);
  }
  {  // M = V^T P
    A4 acc[NPAIR];
#pragma unroll
    for (int p = 0; p < NPAIR; p++) acc[p] = A4{0, 0, 0, 0};
#pragma unroll 2
    for (int u = 0; u < NU; u++) {
      const int gg = 4 * u + q;
      const T *pd = gpd[wv][gg < NG ? gg : 0];
      T mw[6] = {};
      if constexpr (!PtRec<T, NN>::ud) {
#pragma unroll
        for (int c = 0; c < 6; c++) mw[c] = pd[kPtM2 + c];
      }
      T bv[NT][3];
#pragma unroll
      for (int t = NT - 1; t >= 0; t--) {
        const T ph = lt.p[(u * NT + t) * 64 + l];
        const T *vd = pd + kPtV + dr[t] * 3;
        T av[3];
        av[0] = ph * vd[0];
        av[1] = ph * vd[1];
        av[2] = ph * vd[2];
        if constexpr (PtRec<T, NN>::ud) {
          const T *ud = pd + kPtM2 + dr[t] * 3;
          bv[t][0] = ph * ud[0];
          bv[t][1] = ph * ud[1];
          bv[t][2] = ph * ud[2];
        } else {
          bv[t][0] = mw[0] * av[0] + mw[1] * av[1] + mw[2] * av[2];
          bv[t][1] = mw[1] * av[0] + mw[3] * av[1] + mw[4] * av[2];
          bv[t][2] = mw[2] * av[0] + mw[4] * av[1] + mw[5] * av[2];
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int t2 = t; t2 < NT; t2++) {
            const int p = t * NT - t * (t - 1) / 2 + (t2 - t);
            acc[p] = Acc4<T>::mma(av[a], bv[t2][a], acc[p]);
          }
      }
    }
    store(Mout, acc);
  }
}

// k_elem_km_mfma_multi: k_elem_km_mfma with EPW elements per wave, for elements
// whose nodes and points fill a fraction of a wave (8-node: 8 nodes, 8 points,
// four elements per wave).  The node gather and the point setup then keep
// EPW * 8 lanes busy instead of 8; the MFMA passes run element by element.
// (One kernel for both costs the 20-node instantiation 11 registers and an
// occupancy step, measured: 72 -> 61M element matrices/s.)
template <typename T, int NN, int NE, int NG, int EPW>
__global__ __launch_bounds__(256, 2) void k_elem_km_mfma_multi(
    const BasisTab *__restrict__ tab, LaneTab<T> lt, const double *__restrict__ xp, const double *__restrict__ yp,
    const double *__restrict__ zp, const double *__restrict__ inmu, const double *__restrict__ sig, Geom g, Gpml gp,
    int prev, int first_flags, int64_t e0, int64_t n, T *__restrict__ Kout, T *__restrict__ Mout, int *bad) {
  constexpr int WPB = 4;                 // waves per workgroup
  static_assert(EPW * NN <= 64 && EPW * NG <= 64, "the elements of a wave share its 64 lanes");
  constexpr int NT = (NE + 15) / 16;     // DOF tiles
  constexpr int NU = (NG + 3) / 4;       // Gauss-point groups of four
  constexpr int NPAIR = NT * (NT + 1) / 2;
  constexpr int NP = NE * (NE + 1) / 2;
  typedef typename Acc4<T>::type A4;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  // first element of this wave within the launch
  const int64_t le0 = (xcd_block(blockIdx.x, gridDim.x) * WPB + wv) * EPW;
  __shared__ double xs64[WPB][EPW][NN][3];
  __shared__ T nm[WPB][EPW][NN][12];
  __shared__ T gpd[WPB][EPW][NG][PtRec<T, NN>::n];
  // lane -> (element of the wave, node); the element's lattice position
  const int ek0 = l / NN;
  const int64_t le_n = le0 + ek0;
  const int64_t e_n = e0 + (le_n < n ? le_n : 0);
  int i, j, k;
  elem_index(g, e_n, i, j, k);
  if (l < EPW * NN) {
    const int ek = ek0, nd = l - ek * NN;
    int64_t id;
    int ix, jy;
    load_node(tab, g, i, j, k, nd, id, ix, jy);
    xs64[wv][ek][nd][0] = xp[ix];
    xs64[wv][ek][nd][1] = yp[jy];
    xs64[wv][ek][nd][2] = zp[id];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      nm[wv][ek][nd][c] = (T)inmu[id * 6 + c];
      nm[wv][ek][nd][6 + c] = (T)sig[2 * (id * 6 + c)];
    }
  }
  // every array above is private to its wave: a wave barrier orders the
  // stores before the reads of other lanes
  __builtin_amdgcn_wave_barrier();
  if (l < EPW * NG) {
    const int ek = l / NG, pt = l - ek * NG;
    const int64_t le = le0 + ek;
    const bool live = le < n;
    const int64_t e = e0 + (live ? le : 0);
    // GPML stretch of the point first, while little else is live
    double h[3] = {1.0, 1.0, 1.0};
    if (gp.on) {
      // the node lanes' (i, j, k) belong to the same element when NN == NG
      int fi = i, fj = j, fk = k;
      if constexpr (NN != NG) elem_index(g, e, fi, fj, fk);
      int fl = elem_flags(g, gp, prev, first_flags, e, fi, fj, fk);
      double rr[3] = {0, 0, 0};
      for (int m = 0; m < NN; m++)
        for (int c = 0; c < 3; c++) rr[c] += tab->N[pt][m] * xs64[wv][ek][m][c];
      for (int c = 0; c < 3; c++) h[c] = gpml_stretch(gp, fl, c, rr[c]);
    }
    const double h3 = h[0] * h[1] * h[2];
    // GPML curl weights: component a takes q_a1 on u_b v_c and q_a2 on u_c v_b
    const T q0 = (T)sqrt(h[0] * h[2] / h[1]), q1 = (T)sqrt(h[0] * h[1] / h[2]), q2 = (T)sqrt(h[1] * h[2] / h[0]);
    // Jacobian from the point-fastest derivative table (coalesced across lanes)
    T J[3][3];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int c = 0; c < 3; c++) J[m][c] = 0;
#pragma unroll 2
    for (int m = 0; m < NN; m++) {
      const T x0 = (T)xs64[wv][ek][m][0], x1 = (T)xs64[wv][ek][m][1], x2 = (T)xs64[wv][ek][m][2];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const T dn = lt.dN[(a * NN + m) * 32 + pt];
        J[a][0] += dn * x0;
        J[a][1] += dn * x1;
        J[a][2] += dn * x2;
      }
    }
    T det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const T ad = det < 0 ? -det : det;
    T Ji[9];
    Ji[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / ad;
    Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / ad;
    Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / ad;
    Ji[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / ad;
    Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / ad;
    Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / ad;
    Ji[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / ad;
    Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / ad;
    Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / ad;
    if (live && det == (T)0) atomicExch(bad, 1);
    T wgt = det * lt.w[pt];
    T m1[6], m2[6];
#pragma unroll
    for (int c = 0; c < 6; c++) m1[c] = m2[c] = 0;
#pragma unroll 2
    for (int m = 0; m < NN; m++) {
      const T Nm = lt.N[m * 32 + pt];
#pragma unroll
      for (int c = 0; c < 6; c++) {
        m1[c] += Nm * nm[wv][ek][m][c];
        m2[c] += Nm * nm[wv][ek][m][6 + c];
      }
    }
    const T qa1[3] = {q0, q1, q2}, qa2[3] = {q1, q2, q0};
    T *o = gpd[wv][ek][pt];
    const T wh = (T)((double)wgt * h3);
    T w2[6];
#pragma unroll
    for (int c = 0; c < 6; c++) w2[c] = wh * m2[c];
#pragma unroll 1
    for (int d = 0; d < 3; d++) {
      const T v[3] = {Ji[0 * 3 + d], Ji[1 * 3 + d], Ji[2 * 3 + d]};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
#pragma unroll
        for (int m = 0; m < 3; m++) o[d * 9 + a * 3 + m] = qa1[a] * v[c] * Ji[b * 3 + m] - qa2[a] * v[b] * Ji[c * 3 + m];
        o[kPtV + d * 3 + a] = v[a];
      }
      if constexpr (PtRec<T, NN>::ud) {
        o[kPtM2 + d * 3 + 0] = w2[0] * v[0] + w2[1] * v[1] + w2[2] * v[2];
        o[kPtM2 + d * 3 + 1] = w2[1] * v[0] + w2[3] * v[1] + w2[4] * v[2];
        o[kPtM2 + d * 3 + 2] = w2[2] * v[0] + w2[4] * v[1] + w2[5] * v[2];
      }
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
      o[kPtM1 + c] = wgt * m1[c];
      if constexpr (!PtRec<T, NN>::ud) o[kPtM2 + c] = w2[c];
    }
  }
  __builtin_amdgcn_wave_barrier();
  // this lane's DOF direction in each tile (0 for padding DOFs, whose table
  // entries are zero)
  int dr[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) dr[t] = lt.dir[t * 16 + r];
  for (int ek = 0; ek < EPW; ek++) {
  const int64_t le = le0 + ek;
  const bool live = le < n;
  const int64_t base = le * (int64_t)NP;
  const T(*gk)[PtRec<T, NN>::n] = gpd[wv][ek];
  auto store = [&](T *out, const A4 *acc) {
    if (live) store_upper<T, NE, NT>(out + base, acc, l);
  };
  {  // K = C^T Q
    A4 acc[NPAIR];
#pragma unroll
    for (int p = 0; p < NPAIR; p++) acc[p] = A4{0, 0, 0, 0};
#pragma unroll 2
    for (int u = 0; u < NU; u++) {
      const int gg = 4 * u + q;
      const T *pd = gk[gg < NG ? gg : 0];
      T mw[6];
#pragma unroll
      for (int c = 0; c < 6; c++) mw[c] = pd[kPtM1 + c];
      // tiles from the last: the products of tile t need the Q columns of
      // tiles t..NT-1 only, so each tile's C columns die after its MFMAs
      T bv[NT][3];
#pragma unroll
      for (int t = NT - 1; t >= 0; t--) {
        const int li = (u * NT + t) * 64 + l;
        const T dp0 = lt.d0[li], dp1 = lt.d1[li], dp2 = lt.d2[li];
        const T *md = pd + dr[t] * 9;
        T av[3];
        av[0] = md[0] * dp0 + md[1] * dp1 + md[2] * dp2;
        av[1] = md[3] * dp0 + md[4] * dp1 + md[5] * dp2;
        av[2] = md[6] * dp0 + md[7] * dp1 + md[8] * dp2;
        bv[t][0] = mw[0] * av[0] + mw[1] * av[1] + mw[2] * av[2];
        bv[t][1] = mw[1] * av[0] + mw[3] * av[1] + mw[4] * av[2];
        bv[t][2] = mw[2] * av[0] + mw[4] * av[1] + mw[5] * av[2];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int t2 = t; t2 < NT; t2++) {
            const int p = t * NT - t * (t - 1) / 2 + (t2 - t);
            acc[p] = Acc4<T>::mma(av[a], bv[t2][a], acc[p]);
          }
      }
    }
    store(Kout, acc);
  }
  {  // M = V^T P
    A4 acc[NPAIR];
#pragma unroll
    for (int p = 0; p < NPAIR; p++) acc[p] = A4{0, 0, 0, 0};
#pragma unroll 2
    for (int u = 0; u < NU; u++) {
      const int gg = 4 * u + q;
      const T *pd = gk[gg < NG ? gg : 0];
      T mw[6] = {};
      if constexpr (!PtRec<T, NN>::ud) {
#pragma unroll
        for (int c = 0; c < 6; c++) mw[c] = pd[kPtM2 + c];
      }
      T bv[NT][3];
#pragma unroll
      for (int t = NT - 1; t >= 0; t--) {
        const T ph = lt.p[(u * NT + t) * 64 + l];
        const T *vd = pd + kPtV + dr[t] * 3;
        T av[3];
        av[0] = ph * vd[0];
        av[1] = ph * vd[1];
        av[2] = ph * vd[2];
        if constexpr (PtRec<T, NN>::ud) {
          const T *ud = pd + kPtM2 + dr[t] * 3;
          bv[t][0] = ph * ud[0];
          bv[t][1] = ph * ud[1];
          bv[t][2] = ph * ud[2];
        } else {
          bv[t][0] = mw[0] * av[0] + mw[1] * av[1] + mw[2] * av[2];
          bv[t][1] = mw[1] * av[0] + mw[3] * av[1] + mw[4] * av[2];
          bv[t][2] = mw[2] * av[0] + mw[4] * av[1] + mw[5] * av[2];
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int t2 = t; t2 < NT; t2++) {
            const int p = t * NT - t * (t - 1) / 2 + (t2 - t);
            acc[p] = Acc4<T>::mma(av[a], bv[t2][a], acc[p]);
          }
      }
    }
    store(Mout, acc);
  }
  }
}

// per-frequency element RHS: be[e][im][d] complex, d = polarization
template <int NN, int NE, int NG>
__global__ __launch_bounds__(64) void k_elem_rhs(const BasisTab *__restrict__ tab, const double *__restrict__ xp,
                                                 const double *__restrict__ yp, const double *__restrict__ zp,
                                                 const double *__restrict__ sig, const double *__restrict__ inmu,
                                                 const double *__restrict__ dmu, Geom g, Gpml gp, int prev,
                                                 int first_flags, double omega, int64_t e0,
                                                 double *__restrict__ be) {
  const int64_t e = e0 + blockIdx.x;
  const int t = threadIdx.x;
  __shared__ double xs[NN][3];
  __shared__ double zz[NN];                 // float32(omega*b0*z)
  __shared__ double ds[NN][12];             // complex delta sigma (6)
  __shared__ double mh[NN][2][3];           // (mu^-1 dmu) Hp per polarization (real)
  __shared__ double sJi[NG][9];
  __shared__ double swb[NG][2][6];          // w h1h2h3 src, complex 3 per polarization
  int i, j, k;
  elem_index(g, e, i, j, k);
  const double w32 = (double)(float)omega;
  const double psig = (double)(float)(omega * MVF_EPS0);
  const double hh = (double)(float)MVF_B0 / (4 * MVF_PI * 1.e-7);
  if (t < NN) {
    int64_t id;
    int ix, jy;
    load_node(tab, g, i, j, k, t, id, ix, jy);
    xs[t][0] = xp[ix];
    xs[t][1] = yp[jy];
    xs[t][2] = zp[id];
    zz[t] = (double)(float)(omega * MVF_B0 * zp[id]);
    for (int c = 0; c < 6; c++) {
      ds[t][2 * c] = sig[2 * (id * 6 + c)];
      ds[t][2 * c + 1] = sig[2 * (id * 6 + c) + 1];
    }
    ds[t][1] -= psig; ds[t][7] -= psig; ds[t][11] -= psig;
    const double *im = inmu + id * 6, *dm = dmu + id * 6;
    double M[3][3] = {{im[0] * dm[0] + im[1] * dm[1] + im[2] * dm[2], im[0] * dm[1] + im[1] * dm[3] + im[2] * dm[4],
                       im[0] * dm[2] + im[1] * dm[4] + im[2] * dm[5]},
                      {im[1] * dm[0] + im[3] * dm[1] + im[4] * dm[2], im[1] * dm[1] + im[3] * dm[3] + im[4] * dm[4],
                       im[1] * dm[2] + im[3] * dm[4] + im[4] * dm[5]},
                      {im[2] * dm[0] + im[4] * dm[1] + im[5] * dm[2], im[2] * dm[1] + im[4] * dm[3] + im[5] * dm[4],
                       im[2] * dm[2] + im[4] * dm[4] + im[5] * dm[5]}};
    // pol 0: Hp = (0, hh, 0); pol 1: Hp = (hh, 0, 0)
    for (int a = 0; a < 3; a++) {
      mh[t][0][a] = M[a][1] * hh;
      mh[t][1][a] = M[a][0] * hh;
    }
  }
  __syncthreads();
  if (t < NG) {
    double Ji[9];
    double det = jacobian<NN>(tab, t, xs, Ji);
    double wgt = det * tab->w[t];
    double h[3] = {1.0, 1.0, 1.0};
    if (gp.on) {
      int fl = elem_flags(g, gp, prev, first_flags, e, i, j, k);
      double r[3] = {0, 0, 0};
      for (int n = 0; n < NN; n++)
        for (int c = 0; c < 3; c++) r[c] += tab->N[t][n] * xs[n][c];
      for (int c = 0; c < 3; c++) h[c] = gpml_stretch(gp, fl, c, r[c]);
    }
    double wh = wgt * (h[0] * h[1] * h[2]);
    for (int d = 0; d < 2; d++) {
      double sr[3] = {0, 0, 0}, si[3] = {0, 0, 0};  // dmpf + pcrl
      for (int n = 0; n < NN; n++) {
        double N = tab->N[t][n];
        // dsigma . Ep: Ep has a single imaginary component (-i zz on x for pol 0, +i zz on y for pol 1)
        int c0 = d == 0 ? 0 : 1;
        double ep_im = d == 0 ? -zz[n] : zz[n];
        static const int T6c[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        for (int a = 0; a < 3; a++) {
          int q = T6c[a][c0];
          double dr = ds[n][2 * q], di = ds[n][2 * q + 1];
          // (dr + i di) * (i ep_im)
          sr[a] += N * (-di * ep_im);
          si[a] += N * (dr * ep_im);
        }
        // grad N x (mu^-1 dmu Hp)
        double dn[3];
        for (int a = 0; a < 3; a++)
          dn[a] = Ji[a * 3 + 0] * tab->dN[t][0][n] + Ji[a * 3 + 1] * tab->dN[t][1][n] + Ji[a * 3 + 2] * tab->dN[t][2][n];
        const double *v = mh[n][d];
        sr[0] += v[2] * dn[1] - v[1] * dn[2];
        sr[1] += v[0] * dn[2] - v[2] * dn[0];
        sr[2] += v[1] * dn[0] - v[0] * dn[1];
      }
      // times -i w32: (a + i b)(-i w) = b w - i a w
      for (int a = 0; a < 3; a++) {
        swb[t][d][2 * a] = wh * (si[a] * w32);
        swb[t][d][2 * a + 1] = wh * (-sr[a] * w32);
      }
    }
    for (int c = 0; c < 9; c++) sJi[t][c] = Ji[c];
  }
  __syncthreads();
  for (int ed = t; ed < NE; ed += 64) {
    int d = tab->dir[ed];
    double b[4] = {0, 0, 0, 0};
    for (int gg = 0; gg < NG; gg++) {
      double ph = tab->phi[gg][ed];
      double v0 = ph * sJi[gg][0 * 3 + d], v1 = ph * sJi[gg][1 * 3 + d], v2 = ph * sJi[gg][2 * 3 + d];
      for (int pol = 0; pol < 2; pol++) {
        const double *s = swb[gg][pol];
        b[2 * pol] += v0 * s[0] + v1 * s[2] + v2 * s[4];
        b[2 * pol + 1] += v0 * s[1] + v1 * s[3] + v2 * s[5];
      }
    }
    double *o = be + ((e - e0) * NE + ed) * 4;
    o[0] = b[0]; o[1] = b[1]; o[2] = b[2]; o[3] = b[3];
  }
}

template <typename T>
static LaneTab<T> lane_tab(const mvf_handle *h);
template <>
LaneTab<double> lane_tab<double>(const mvf_handle *h) {
  return h->lt64;
}
template <>
LaneTab<float> lane_tab<float>(const mvf_handle *h) {
  return h->lt32;
}

template <typename T, int NN, int NE, int NG, int GC>
static int launch_km(mvf_handle *h, int64_t e0, int64_t n, int first_flags, T *dK, T *dM, int *bad) {
  int prev = (h->emulate & MVF_EMU_GPML_PREV_FLAGS) ? 1 : 0;
  if (!h->elem_legacy) {
    // four waves per workgroup, one element per wave (four for 8-node); at most 2^22 elements per launch
    const int64_t chunk = (int64_t)1 << 22;
    for (int64_t s = 0; s < n; s += chunk) {
      int64_t c = (n - s) < chunk ? (n - s) : chunk;
      constexpr int64_t NP = NE * (NE + 1) / 2;
      if constexpr (NN == 8)
        hipLaunchKernelGGL((k_elem_km_mfma_multi<T, NN, NE, NG, 4>), dim3((unsigned)((c + 15) / 16)), dim3(256), 0,
                           h->stream, h->d_tab, lane_tab<T>(h), h->d_xp, h->d_yp, h->d_zp, h->d_inmu, h->d_sig, h->geo,
                           h->gp, prev, first_flags, e0 + s, c, dK + s * NP, dM + s * NP, bad);
      else
        hipLaunchKernelGGL((k_elem_km_mfma<T, NN, NE, NG>), dim3((unsigned)((c + 3) / 4)), dim3(256), 0, h->stream,
                           h->d_tab, lane_tab<T>(h), h->d_xp, h->d_yp, h->d_zp, h->d_inmu, h->d_sig, h->geo, h->gp,
                           prev, first_flags, e0 + s, c, dK + s * NP, dM + s * NP, bad);
    }
    MVF_HIP(h, hipGetLastError());
    return 0;
  }
  const int64_t chunk = 1 << 20;
  for (int64_t s = 0; s < n; s += chunk) {
    int64_t c = (n - s) < chunk ? (n - s) : chunk;
    constexpr int64_t NP = NE * (NE + 1) / 2;
    hipLaunchKernelGGL((k_elem_km<T, NN, NE, NG, GC>), dim3((unsigned)c), dim3(256), 0, h->stream, h->d_tab, h->d_xp,
                       h->d_yp, h->d_zp, h->d_inmu, h->d_sig, h->geo, h->gp, prev, first_flags, e0 + s, dK + s * NP,
                       dM + s * NP, bad);
  }
  MVF_HIP(h, hipGetLastError());
  return 0;
}

template <typename T>
static int integrate_t(mvf_handle *h, int64_t e0, int64_t e1, int first_flags, T *dK, T *dM) {
  int *bad;
  MVF_TRY(dalloc(h, &bad, 1));
  MVF_HIP(h, hipMemsetAsync(bad, 0, sizeof(int), h->stream));
  int64_t n = e1 - e0;
  int rc = 0;
  if (n > 0) {
    if (h->nodes == 8) rc = launch_km<T, 8, 12, 8, 8>(h, e0, n, first_flags, dK, dM, bad);
    else if (h->nodes == 20) rc = launch_km<T, 20, 36, 27, 9>(h, e0, n, first_flags, dK, dM, bad);
    else rc = launch_km<T, 27, 54, 27, 6>(h, e0, n, first_flags, dK, dM, bad);
  }
  int hb = 0;
  if (!rc) {
    MVF_HIP(h, hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    MVF_HIP(h, hipStreamSynchronize(h->stream));
  }
  dfree(bad);
  if (rc) return rc;
  if (hb) return set_err(h, MVF_E_SINGULAR, "zero Jacobian determinant in an element (degenerate coordinates)");
  return 0;
}

int integrate_elements(mvf_handle *h, int64_t e0, int64_t e1, int first_flags, double *dK, double *dM) {
  return integrate_t<double>(h, e0, e1, first_flags, dK, dM);
}

int integrate_elements32(mvf_handle *h, int64_t e0, int64_t e1, int first_flags, float *dK, float *dM) {
  return integrate_t<float>(h, e0, e1, first_flags, dK, dM);
}

int element_rhs_range(mvf_handle *h, double omega, int first_flags, int64_t e0, int64_t e1, double *d_be) {
  int prev = (h->emulate & MVF_EMU_GPML_PREV_FLAGS) ? 1 : 0;
  const int64_t chunk = 1 << 22;
  const int me = h->et.ne;
  for (int64_t s = e0; s < e1; s += chunk) {
    int64_t c = (e1 - s) < chunk ? (e1 - s) : chunk;
    double *out = d_be + (s - e0) * me * 4;
    if (h->nodes == 8)
      hipLaunchKernelGGL((k_elem_rhs<8, 12, 8>), dim3((unsigned)c), dim3(64), 0, h->stream, h->d_tab, h->d_xp, h->d_yp,
                         h->d_zp, h->d_sig, h->d_inmu, h->d_dmu, h->geo, h->gp, prev, first_flags, omega, s, out);
    else if (h->nodes == 20)
      hipLaunchKernelGGL((k_elem_rhs<20, 36, 27>), dim3((unsigned)c), dim3(64), 0, h->stream, h->d_tab, h->d_xp,
                         h->d_yp, h->d_zp, h->d_sig, h->d_inmu, h->d_dmu, h->geo, h->gp, prev, first_flags, omega, s, out);
    else
      hipLaunchKernelGGL((k_elem_rhs<27, 54, 27>), dim3((unsigned)c), dim3(64), 0, h->stream, h->d_tab, h->d_xp,
                         h->d_yp, h->d_zp, h->d_sig, h->d_inmu, h->d_dmu, h->geo, h->gp, prev, first_flags, omega, s, out);
  }
  MVF_HIP(h, hipGetLastError());
  return 0;
}

// update_sigma (geometry.f90:144-153): linear entries [0, min(npt+30, 6 npt))
// whose imaginary part is non-zero get imag = float32(eps*omega).  Without
// MVF_EMU_UPDATE_SIGMA every entry is refreshed.
__global__ void k_update_sigma(double *sig, int64_t lim, double im) {
  int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (L >= lim) return;
  if (sig[2 * L + 1] != 0.0) sig[2 * L + 1] = im;
}

int update_sigma(mvf_handle *h, double omega) {
  int64_t npt = h->geo.npt;
  int64_t lim = (h->emulate & MVF_EMU_UPDATE_SIGMA) ? npt + 30 : 6 * npt;
  if (lim > 6 * npt) lim = 6 * npt;
  double im = (double)(float)(MVF_EPS0 * omega);
  hipLaunchKernelGGL(k_update_sigma, dim3(grid1d(lim, 256)), dim3(256), 0, h->stream, h->d_sig, lim, im);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

}  // namespace mvf
