// DOF numbering, sparsity pattern and gather-list construction on the device.
//
// Numbering restates c_gne12/36/54 (global_assembly.f90:231-475): elements in
// reference order (z fastest); a local DOF is copied from the -z, -y, -x face
// neighbour through the copy maps (later copies overwrite, x last), Dirichlet
// boundary DOFs get -face (face order -x,-y,-z,+x,+y,+z as edge_bdary,
// boundary_conds.f90:262-390), every other DOF takes the next number.  The
// serial counter becomes a per-element count + exclusive scan, and copied DOFs
// follow the copy chain back to the element where the DOF is new.
//
// The pattern is the set of (gne(e,im), gne(e,jm)) pairs over all elements
// (both triangles), i.e. what the reference's slot map nzindx enumerates
// (ga_nzindx, :197-229).  It is built by one stable radix sort of 64-bit
// (row, col) keys carrying (element, im, jm) payloads; the sorted payloads are
// the per-entry gather lists, already in element order, so assembly sums every
// entry in the same order as the reference's element loop.
//
// Large models (the pair sort would need more than a quarter of free HBM, e.g.
// 4.59e9 pairs, ~147 GB of keys, for 27-node 128 x 128 x 96) build the pattern
// row by row instead: one wavefront per row gathers the DOFs of the (at most
// eight) elements holding the row's DOF, sorts them in LDS and keeps the
// distinct ones; assembly then sums each row's entries element by element in
// LDS.  Both paths give the same pattern and the same sums in the same order.
#include <hipcub/hipcub.hpp>

#include "mvf_internal.h"

namespace mvf {

struct CopyTab {
  unsigned char c[3][12][2];
  int nf, me;
};

static CopyTab make_copytab(const ElemType &et) {
  CopyTab t{};
  t.nf = et.nf;
  t.me = et.ne;
  for (int a = 0; a < 3; a++)
    for (int k = 0; k < et.nf; k++)
      for (int s = 0; s < 2; s++) t.c[a][k][s] = et.copy[(a * et.nf + k) * 2 + s];
  return t;
}

__device__ inline int in_list(const CopyTab &T, int axis, int col, int im) {
  for (int k = 0; k < T.nf; k++)
    if (T.c[axis][k][col] == im) return k;
  return -1;
}

// -face code if the DOF is removed by Dirichlet, 0 otherwise
__device__ inline int boundary_face(const CopyTab &T, const Geom &g, int dirichlet, int i, int j, int k, int im) {
  if (!dirichlet) return 0;
  int cond[6] = {i == 1, j == 1, k == 1, i == g.nx, j == g.ny, k == g.nz};
  for (int f = 0; f < 6; f++) {
    if (!cond[f]) continue;
    int axis = (f % 3 == 0) ? 2 : ((f % 3 == 1) ? 1 : 0);
    if (in_list(T, axis, f < 3 ? 0 : 1, im) >= 0) return f + 1;
  }
  return 0;
}

__device__ inline int is_copied(const CopyTab &T, int i, int j, int k, int im) {
  return (k > 1 && in_list(T, 0, 0, im) >= 0) || (j > 1 && in_list(T, 1, 0, im) >= 0) ||
         (i > 1 && in_list(T, 2, 0, im) >= 0);
}

__device__ inline void elem_ijk(const Geom &g, int64_t e, int &i, int &j, int &k) {
  i = (int)(e / ((int64_t)g.ny * g.nz)) + 1;
  j = (int)((e / g.nz) % g.ny) + 1;
  k = (int)(e % g.nz) + 1;
}

__global__ void k_newmask(CopyTab T, Geom g, int dirichlet, uint64_t *mask, int64_t *cnt) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= g.ne) return;
  int i, j, k;
  elem_ijk(g, e, i, j, k);
  uint64_t m = 0;
  for (int im = 1; im <= T.me; im++) {
    if (boundary_face(T, g, dirichlet, i, j, k, im)) continue;
    if (!is_copied(T, i, j, k, im)) m |= (1ull << (im - 1));
  }
  mask[e] = m;
  cnt[e] = __popcll(m);
}

__global__ void k_number(CopyTab T, Geom g, int dirichlet, const uint64_t *mask, const int64_t *base,
                         int32_t *gne) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= g.ne * T.me) return;
  int64_t e = t / T.me;
  int im = (int)(t % T.me) + 1;
  int i, j, k;
  elem_ijk(g, e, i, j, k);
  int f = boundary_face(T, g, dirichlet, i, j, k, im);
  if (f) {
    gne[t] = -f;
    return;
  }
  int ci = i, cj = j, ck = k, cm = im;
  for (int hop = 0; hop < 16; hop++) {
    int64_t ce = ((int64_t)(ci - 1) * g.ny + (cj - 1)) * g.nz + (ck - 1);
    int bf = boundary_face(T, g, dirichlet, ci, cj, ck, cm);
    if (bf) {
      gne[t] = -bf;
      return;
    }
    uint64_t m = mask[ce];
    if ((m >> (cm - 1)) & 1ull) {
      uint64_t lower = (cm - 1 == 0) ? 0ull : (m & ((1ull << (cm - 1)) - 1ull));
      gne[t] = (int32_t)(base[ce] + __popcll(lower) + 1);
      return;
    }
    // last applicable copy wins: x, then y, then z
    int p;
    if (ci > 1 && (p = in_list(T, 2, 0, cm)) >= 0) {
      ci -= 1; cm = T.c[2][p][1];
    } else if (cj > 1 && (p = in_list(T, 1, 0, cm)) >= 0) {
      cj -= 1; cm = T.c[1][p][1];
    } else if (ck > 1 && (p = in_list(T, 0, 0, cm)) >= 0) {
      ck -= 1; cm = T.c[0][p][1];
    } else {
      gne[t] = 0;  // unreachable: neither new nor copied
      return;
    }
  }
  gne[t] = 0;
}

int setup_numbering(mvf_handle *h) {
  DevScope ds;
  const Geom &g = h->geo;
  CopyTab T = make_copytab(h->et);
  uint64_t *mask;
  int64_t *cnt, *base;
  MVF_TRY(ds.alloc(h, &mask, g.ne));
  MVF_TRY(ds.alloc(h, &cnt, g.ne));
  MVF_TRY(ds.alloc(h, &base, g.ne + 1));
  MVF_TRY(dalloc(h, &h->d_gne, g.ne * h->et.ne));
  hipLaunchKernelGGL(k_newmask, dim3(grid1d(g.ne, 256)), dim3(256), 0, h->stream, T, g, h->bc.dirichlet, mask, cnt);
  size_t tb = 0;
  MVF_HIP(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, base, (int)g.ne, h->stream));
  void *tmp;
  MVF_TRY(ds.alloc(h, (char **)&tmp, tb));
  MVF_HIP(h, hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, base, (int)g.ne, h->stream));
  hipLaunchKernelGGL(k_number, dim3(grid1d(g.ne * h->et.ne, 256)), dim3(256), 0, h->stream, T, g, h->bc.dirichlet,
                     mask, base, h->d_gne);
  int64_t last_base = 0, last_cnt = 0;
  MVF_HIP(h, hipMemcpyAsync(&last_base, base + g.ne - 1, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipMemcpyAsync(&last_cnt, cnt + g.ne - 1, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  MVF_HIP(h, hipGetLastError());
  int64_t nne = last_base + last_cnt;
  ds.drop(mask); ds.drop(cnt); ds.drop(base); ds.drop(tmp);
  if (nne <= 0 || nne > INT32_MAX) return set_err(h, MVF_E_ARG, "DOF count %lld out of range", (long long)nne);
  h->nne = (int32_t)nne;
  return 0;
}

// ---------------------------------------------------------------- pattern --
__global__ void k_validcount(const int32_t *gne, int me, int64_t ne, int64_t *cnt, int64_t *rcnt) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int v = 0;
  for (int a = 0; a < me; a++) v += gne[e * me + a] >= 0;
  cnt[e] = (int64_t)v * v;
  rcnt[e] = v;
}

// one block per element: all valid (a, b) pairs in row-major local order
__global__ void k_pairs(const int32_t *gne, int me, int64_t nne, const int64_t *off, const int64_t *roff,
                        uint64_t *keys, int64_t *vals, uint64_t *rkeys, int64_t *rvals) {
  int64_t e = blockIdx.x;
  __shared__ int32_t sg[54];
  __shared__ int32_t pos[54];
  __shared__ int nv;
  if (threadIdx.x == 0) {
    int v = 0;
    for (int a = 0; a < me; a++) {
      sg[a] = gne[e * me + a];
      pos[a] = sg[a] >= 0 ? v++ : -1;
    }
    nv = v;
  }
  __syncthreads();
  int v = nv;
  for (int t = threadIdx.x; keys && t < me * me; t += blockDim.x) {
    int a = t / me, b = t % me;
    if (sg[a] < 0 || sg[b] < 0) continue;
    int64_t p = off[e] + (int64_t)pos[a] * v + pos[b];
    keys[p] = (uint64_t)(sg[a] - 1) * (uint64_t)nne + (uint64_t)(sg[b] - 1);
    vals[p] = (e * me + a) * me + b;
  }
  for (int a = threadIdx.x; a < me; a += blockDim.x) {
    if (sg[a] < 0) continue;
    int64_t p = roff[e] + pos[a];
    rkeys[p] = (uint64_t)(sg[a] - 1);
    rvals[p] = e * me + a;
  }
}

__global__ void k_heads(const uint64_t *keys, int64_t n, int64_t *head) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// scan holds the inclusive head count; entry index = scan-1 at each head
__global__ void k_entries(const uint64_t *keys, const int64_t *scan, const int64_t *head, int64_t n, int64_t nne,
                          int64_t *seg, int32_t *col, int64_t *rowptr, int64_t *ntriu) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || !head[i]) return;
  int64_t ent = scan[i] - 1;
  uint64_t key = keys[i];
  int64_t r = (int64_t)(key / (uint64_t)nne), c = (int64_t)(key % (uint64_t)nne);
  seg[ent] = i;
  col[ent] = (int32_t)c;
  if (c >= r) atomicAdd((unsigned long long *)ntriu, 1ull);
  int64_t prev_r = -1;
  if (ent > 0) {
    // the previous entry's row: the key of the previous head
    int64_t j = i - 1;
    prev_r = (int64_t)(keys[j] / (uint64_t)nne);
  }
  for (int64_t rr = prev_r + 1; rr <= r; rr++) rowptr[rr] = ent;
}

__global__ void k_rseg(const uint64_t *rkeys, int64_t n, int64_t *rseg, int64_t nne) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t r = (int64_t)rkeys[i];
  int64_t pr = (i == 0) ? -1 : (int64_t)rkeys[i - 1];
  for (int64_t rr = pr + 1; rr <= r; rr++) rseg[rr] = i;
  if (i == n - 1)
    for (int64_t rr = r + 1; rr <= nne; rr++) rseg[rr] = n;
}

__global__ void k_diagpos(const int64_t *rowptr, const int32_t *col, int64_t nne, int64_t *diag) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nne) return;
  int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1, pos = -1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    int32_t c = col[mid];
    if (c == r) {
      pos = mid;
      break;
    }
    if (c < r) lo = mid + 1;
    else hi = mid - 1;
  }
  diag[r] = pos;
}

static int setup_pattern_rows(mvf_handle *h);

static int end_bit_for(uint64_t maxkey) {
  int b = 1;
  while (b < 64 && (maxkey >> b)) b++;
  return b;
}

int setup_pattern(mvf_handle *h) {
  DevScope ds;
  const Geom &g = h->geo;
  int me = h->et.ne;
  int64_t nne = h->nne;
  int64_t *cnt, *off, *rcnt, *roff;
  MVF_TRY(ds.alloc(h, &cnt, g.ne));
  MVF_TRY(ds.alloc(h, &off, g.ne));
  MVF_TRY(ds.alloc(h, &rcnt, g.ne));
  MVF_TRY(ds.alloc(h, &roff, g.ne));
  hipLaunchKernelGGL(k_validcount, dim3(grid1d(g.ne, 256)), dim3(256), 0, h->stream, h->d_gne, me, g.ne, cnt, rcnt);
  size_t tb = 0, tb2 = 0;
  MVF_HIP(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, off, (int)g.ne, h->stream));
  char *tmp;
  MVF_TRY(ds.alloc(h, &tmp, tb));
  MVF_HIP(h, hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, off, (int)g.ne, h->stream));
  MVF_HIP(h, hipcub::DeviceScan::ExclusiveSum(tmp, tb, rcnt, roff, (int)g.ne, h->stream));
  int64_t lo = 0, lc = 0, rlo = 0, rlc = 0;
  MVF_HIP(h, hipMemcpyAsync(&lo, off + g.ne - 1, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipMemcpyAsync(&lc, cnt + g.ne - 1, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipMemcpyAsync(&rlo, roff + g.ne - 1, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipMemcpyAsync(&rlc, rcnt + g.ne - 1, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  ds.drop(tmp);
  int64_t np = lo + lc, nr = rlo + rlc;
  h->npairs = np;
  {
    // the pair sort holds two key arrays, the payloads and the gather lists
    // (32 bytes per pair) at once: by rows when that exceeds a quarter of the
    // free memory, or when MVF_SETUP says so
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    const char *sv = getenv("MVF_SETUP");
    h->setup_rows = (sv && strcmp(sv, "rows") == 0) ? 1
                    : (sv && strcmp(sv, "pairs") == 0) ? 0
                                                       : ((double)np * 32.0 > 0.25 * (double)fr ? 1 : 0);
  }
  if (h->setup_rows) {
    uint64_t *rk1, *rk2;
    int64_t *rv1;
    MVF_TRY(ds.alloc(h, &rk1, nr));
    MVF_TRY(ds.alloc(h, &rk2, nr));
    MVF_TRY(ds.alloc(h, &rv1, nr));
    MVF_TRY(dalloc(h, &h->d_rcontrib, nr));
    MVF_TRY(dalloc(h, &h->d_rseg, nne + 1));
    hipLaunchKernelGGL(k_pairs, dim3((unsigned)g.ne), dim3(256), 0, h->stream, h->d_gne, me, nne, off, roff, nullptr,
                       nullptr, rk1, rv1);
    MVF_HIP(h, hipGetLastError());
    ds.drop(cnt); ds.drop(off); ds.drop(rcnt); ds.drop(roff);
    int reb = end_bit_for((uint64_t)nne);
    MVF_HIP(h, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, rk1, rk2, rv1, h->d_rcontrib, (int64_t)nr, 0, reb,
                                                   h->stream));
    MVF_TRY(ds.alloc(h, &tmp, tb));
    MVF_HIP(h, hipcub::DeviceRadixSort::SortPairs(tmp, tb, rk1, rk2, rv1, h->d_rcontrib, (int64_t)nr, 0, reb,
                                                   h->stream));
    hipLaunchKernelGGL(k_rseg, dim3(grid1d(nr, 256)), dim3(256), 0, h->stream, rk2, nr, h->d_rseg, nne);
    MVF_HIP(h, hipStreamSynchronize(h->stream));
    ds.drop(tmp); ds.drop(rk1); ds.drop(rk2); ds.drop(rv1);
    return setup_pattern_rows(h);
  }
  uint64_t *k1, *k2, *rk1, *rk2;
  int64_t *v1, *rv1;
  MVF_TRY(ds.alloc(h, &k1, np));
  MVF_TRY(ds.alloc(h, &k2, np));
  MVF_TRY(ds.alloc(h, &v1, np));
  MVF_TRY(dalloc(h, &h->d_contrib, np));
  MVF_TRY(ds.alloc(h, &rk1, nr));
  MVF_TRY(ds.alloc(h, &rk2, nr));
  MVF_TRY(ds.alloc(h, &rv1, nr));
  MVF_TRY(dalloc(h, &h->d_rcontrib, nr));
  hipLaunchKernelGGL(k_pairs, dim3((unsigned)g.ne), dim3(256), 0, h->stream, h->d_gne, me, nne, off, roff, k1, v1,
                     rk1, rv1);
  MVF_HIP(h, hipGetLastError());
  ds.drop(cnt); ds.drop(off); ds.drop(rcnt); ds.drop(roff);
  int eb = end_bit_for((uint64_t)nne * (uint64_t)nne);
  int reb = end_bit_for((uint64_t)nne);
  MVF_HIP(h, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k1, k2, v1, h->d_contrib, (int64_t)np, 0, eb, h->stream));
  MVF_HIP(h, hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, rk1, rk2, rv1, h->d_rcontrib, (int64_t)nr, 0, reb,
                                                 h->stream));
  size_t tbm = tb > tb2 ? tb : tb2;
  MVF_TRY(ds.alloc(h, &tmp, tbm));
  MVF_HIP(h, hipcub::DeviceRadixSort::SortPairs(tmp, tbm, k1, k2, v1, h->d_contrib, (int64_t)np, 0, eb, h->stream));
  MVF_HIP(h, hipcub::DeviceRadixSort::SortPairs(tmp, tbm, rk1, rk2, rv1, h->d_rcontrib, (int64_t)nr, 0, reb,
                                                 h->stream));
  ds.drop(tmp);
  ds.drop(k1); ds.drop(v1); ds.drop(rk1); ds.drop(rv1);
  // unique entries
  int64_t *head, *scan;
  MVF_TRY(ds.alloc(h, &head, np));
  MVF_TRY(ds.alloc(h, &scan, np));
  hipLaunchKernelGGL(k_heads, dim3(grid1d(np, 256)), dim3(256), 0, h->stream, k2, np, head);
  MVF_HIP(h, hipcub::DeviceScan::InclusiveSum(nullptr, tb, head, scan, (int64_t)np, h->stream));
  MVF_TRY(ds.alloc(h, &tmp, tb));
  MVF_HIP(h, hipcub::DeviceScan::InclusiveSum(tmp, tb, head, scan, (int64_t)np, h->stream));
  int64_t nnz = 0;
  MVF_HIP(h, hipMemcpyAsync(&nnz, scan + np - 1, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  ds.drop(tmp);
  h->nnz = nnz;
  MVF_TRY(dalloc(h, &h->d_seg, nnz + 1));
  MVF_TRY(dalloc(h, &h->d_col, nnz));
  MVF_TRY(dalloc(h, &h->d_rowptr, nne + 1));
  MVF_TRY(dalloc(h, &h->d_diagpos, nne));
  MVF_TRY(dalloc(h, &h->d_rseg, nne + 1));
  int64_t *ntriu;
  MVF_TRY(ds.alloc(h, &ntriu, 1));
  MVF_HIP(h, hipMemsetAsync(ntriu, 0, 8, h->stream));
  hipLaunchKernelGGL(k_entries, dim3(grid1d(np, 256)), dim3(256), 0, h->stream, k2, scan, head, np, nne, h->d_seg,
                     h->d_col, h->d_rowptr, ntriu);
  MVF_HIP(h, hipMemcpyAsync(h->d_seg + nnz, &np, 8, hipMemcpyHostToDevice, h->stream));
  MVF_HIP(h, hipMemcpyAsync(h->d_rowptr + nne, &nnz, 8, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_diagpos, dim3(grid1d(nne, 256)), dim3(256), 0, h->stream, h->d_rowptr, h->d_col, nne,
                     h->d_diagpos);
  hipLaunchKernelGGL(k_rseg, dim3(grid1d(nr, 256)), dim3(256), 0, h->stream, rk2, nr, h->d_rseg, nne);
  MVF_HIP(h, hipMemcpyAsync(&h->nnz_triu, ntriu, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  MVF_HIP(h, hipGetLastError());
  ds.drop(head); ds.drop(scan); ds.drop(k2); ds.drop(rk2); ds.drop(ntriu);
  return 0;
}

// ------------------------------------------------------- per-row pattern --
// A DOF lies in at most four elements (an element edge), so a row couples at
// most 4 * me candidate columns: NP = 64 (8-node) or 256 (20/27-node).
constexpr int kRowEl = 4;

// candidates of row r: the DOFs (0-based) of every element that holds r,
// padded with INT32_MAX to NP and sorted ascending in LDS; all loads of the
// row are issued before any is used.  Returns the element count or -1.
template <int NP>
__device__ inline int row_candidates(const int32_t *gne, int me, const int64_t *rseg, const int64_t *rcontrib,
                                     int64_t r, int lane, int32_t *v) {
  const int64_t t0 = rseg[r];
  const int nel = (int)(rseg[r + 1] - t0);
  if (nel > kRowEl) return -1;
  int64_t ea[kRowEl];
#pragma unroll
  for (int q = 0; q < kRowEl; q++) ea[q] = q < nel ? rcontrib[t0 + q] : 0;
#pragma unroll
  for (int k = 0; k < NP / 64; k++) {
    const int idx = lane + 64 * k, q = idx / me, b = idx - q * me;
    int32_t x = INT32_MAX;
    if (q < nel) {
      int64_t e = 0;
#pragma unroll
      for (int qq = 0; qq < kRowEl; qq++)
        if (qq == q) e = ea[qq] / me;
      const int32_t g = gne[e * me + b];
      if (g > 0) x = g - 1;
    }
    v[idx] = x;
  }
  __builtin_amdgcn_wave_barrier();
  // bitonic sort of the NP values, ascending; each step's pairs are disjoint
  for (int k = 2; k <= NP; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int c = 0; c < NP / 64; c++) {
        const int i = lane + 64 * c, ij = i ^ j;
        if (ij > i) {
          const int32_t a = v[i], b = v[ij];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            v[i] = b;
            v[ij] = a;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  return nel;
}

// pass 0: cnt[r] = distinct columns of row r, ntriu += those >= r;
// pass 1: col[rowptr[r] + rank] for each distinct column in ascending order
template <int NP, int PASS>
__global__ __launch_bounds__(256) void k_row_pattern(const int32_t *gne, int me, const int64_t *rseg,
                                                     const int64_t *rcontrib, int64_t nne, int64_t *cnt,
                                                     unsigned long long *ntriu, const int64_t *rowptr, int32_t *col,
                                                     int *bad) {
  __shared__ int32_t buf[4][NP];
  __shared__ unsigned long long up4[4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int32_t *v = buf[wv];
  // each wave walks rows with a grid stride and sums its upper-triangle count;
  // one atomic per workgroup (an atomic per row on one counter serialized the
  // count pass of a 38M-row model)
  unsigned long long upper_all = 0;
  for (int64_t r = (int64_t)blockIdx.x * 4 + wv; r < nne; r += (int64_t)gridDim.x * 4) {
    if (row_candidates<NP>(gne, me, rseg, rcontrib, r, lane, v) < 0) {
      if (lane == 0) atomicExch(bad, 1);
      continue;
    }
    const int64_t base = PASS ? rowptr[r] : 0;
    int count = 0, upper = 0;
#pragma unroll
    for (int c0 = 0; c0 < NP; c0 += 64) {
      const int i = c0 + lane;
      const int32_t x = v[i];
      const bool head = x != INT32_MAX && (i == 0 || v[i - 1] != x);
      const unsigned long long m = __ballot(head);
      if (PASS) {
        const int rank = count + (int)__popcll(m & ((1ull << lane) - 1ull));
        if (head) col[base + rank] = x;
      } else {
        upper += (int)__popcll(__ballot(head && x >= r));
      }
      count += (int)__popcll(m);
    }
    if (!PASS && lane == 0) cnt[r] = count;
    upper_all += (unsigned long long)upper;
    __builtin_amdgcn_wave_barrier();  // the next row's candidates overwrite v
  }
  if (!PASS) {
    if (lane == 0) up4[wv] = upper_all;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(ntriu, up4[0] + up4[1] + up4[2] + up4[3]);
  }
}

// K and M of row r: every contribution (element q of the row, local column b)
// is loaded first; then, element by element in element order, each lane adds
// its contribution to the entry found by binary search in the row's columns
// (staged in LDS).  One element adds to distinct entries.
template <typename TE, int NP>
__global__ __launch_bounds__(256) void k_assemble_rows(const int32_t *gne, int me, const int64_t *rseg,
                                                       const int64_t *rcontrib, const int64_t *rowptr,
                                                       const int32_t *col, int64_t nne, const TE *Ke, const TE *Me,
                                                       double *K, double *M, float *K32, float *M32) {
  __shared__ int32_t scol[4][NP];
  __shared__ double acc[4][2][NP];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + wv;
  if (r >= nne) return;
  const int64_t p0 = rowptr[r], p1 = rowptr[r + 1];
  const int n = (int)(p1 - p0);
  const int64_t t0 = rseg[r];
  const int nel = (int)(rseg[r + 1] - t0);
  int32_t *cs = scol[wv];
  double *ak = acc[wv][0], *am = acc[wv][1];
  int64_t ea[kRowEl];
#pragma unroll
  for (int q = 0; q < kRowEl; q++) ea[q] = q < nel ? rcontrib[t0 + q] : 0;
#pragma unroll
  for (int k = 0; k < NP / 64; k++) {
    const int i = lane + 64 * k;
    if (i < n) {
      cs[i] = col[p0 + i];
      ak[i] = am[i] = 0.0;
    }
  }
  const int64_t npk = (int64_t)me * (me + 1) / 2;
  int32_t gc[NP / 64];
  double kv[NP / 64], mv[NP / 64];
#pragma unroll
  for (int k = 0; k < NP / 64; k++) {
    const int idx = lane + 64 * k, q = idx / me, b = idx - q * me;
    gc[k] = -1;
    kv[k] = mv[k] = 0.0;
    if (q < nel) {
      int64_t c = 0;
#pragma unroll
      for (int qq = 0; qq < kRowEl; qq++)
        if (qq == q) c = ea[qq];
      const int64_t e = c / me;
      const int a = (int)(c - e * me);
      const int32_t g = gne[e * me + b];
      if (g > 0) {
        const int i0 = a < b ? a : b, j0 = a < b ? b : a;
        const int64_t pk = e * npk + (int64_t)i0 * me - (int64_t)i0 * (i0 - 1) / 2 + (j0 - i0);
        gc[k] = g - 1;
        kv[k] = (double)Ke[pk];
        mv[k] = (double)Me[pk];
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  for (int q = 0; q < nel; q++) {
#pragma unroll
    for (int k = 0; k < NP / 64; k++) {
      const int idx = lane + 64 * k;
      if (idx / me == q && gc[k] >= 0) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (cs[mid] < gc[k]) lo = mid + 1;
          else hi = mid;
        }
        ak[lo] += kv[k];
        am[lo] += mv[k];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int k = 0; k < NP / 64; k++) {
    const int i = lane + 64 * k;
    if (i < n) {
      K[p0 + i] = ak[i];
      M[p0 + i] = am[i];
      K32[p0 + i] = (float)ak[i];
      M32[p0 + i] = (float)am[i];
    }
  }
}

static int np_for(int me) { return me <= 16 ? 64 : 256; }

// pattern by rows (rseg/rcontrib built): rowptr, col, diagpos, nnz, nnz_triu
static int setup_pattern_rows(mvf_handle *h) {
  DevScope ds;
  const int me = h->et.ne;
  const int64_t nne = h->nne;
  int64_t *cnt;
  unsigned long long *ntriu;
  int *bad;
  MVF_TRY(ds.alloc(h, &cnt, nne + 1));
  MVF_TRY(ds.alloc(h, &ntriu, 1));
  MVF_TRY(ds.alloc(h, &bad, 1));
  MVF_HIP(h, hipMemsetAsync(ntriu, 0, 8, h->stream));
  MVF_HIP(h, hipMemsetAsync(bad, 0, 4, h->stream));
  MVF_HIP(h, hipMemsetAsync(cnt + nne, 0, 8, h->stream));
  // four rows per workgroup and pass, at most 32768 workgroups (grid stride)
  const unsigned grid = (unsigned)std::min<int64_t>((nne + 3) / 4, 32768);
  if (np_for(me) == 64)
    hipLaunchKernelGGL((k_row_pattern<64, 0>), dim3(grid), dim3(256), 0, h->stream, h->d_gne, me, h->d_rseg,
                       h->d_rcontrib, nne, cnt, ntriu, nullptr, nullptr, bad);
  else
    hipLaunchKernelGGL((k_row_pattern<256, 0>), dim3(grid), dim3(256), 0, h->stream, h->d_gne, me, h->d_rseg,
                       h->d_rcontrib, nne, cnt, ntriu, nullptr, nullptr, bad);
  MVF_TRY(dalloc(h, &h->d_rowptr, nne + 1));
  size_t tb = 0;
  MVF_HIP(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, h->d_rowptr, (int64_t)nne + 1, h->stream));
  char *tmp;
  MVF_TRY(ds.alloc(h, &tmp, tb));
  MVF_HIP(h, hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, h->d_rowptr, (int64_t)nne + 1, h->stream));
  int64_t nnz = 0;
  int hb = 0;
  MVF_HIP(h, hipMemcpyAsync(&nnz, h->d_rowptr + nne, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipMemcpyAsync(&h->nnz_triu, ntriu, 8, hipMemcpyDeviceToHost, h->stream));
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  if (hb) return set_err(h, MVF_E_INDEX, "a DOF lies in more than four elements: not a structured hexahedral grid");
  h->nnz = nnz;
  MVF_TRY(dalloc(h, &h->d_col, nnz));
  MVF_TRY(dalloc(h, &h->d_diagpos, nne));
  if (np_for(me) == 64)
    hipLaunchKernelGGL((k_row_pattern<64, 1>), dim3(grid), dim3(256), 0, h->stream, h->d_gne, me, h->d_rseg,
                       h->d_rcontrib, nne, nullptr, nullptr, h->d_rowptr, h->d_col, bad);
  else
    hipLaunchKernelGGL((k_row_pattern<256, 1>), dim3(grid), dim3(256), 0, h->stream, h->d_gne, me, h->d_rseg,
                       h->d_rcontrib, nne, nullptr, nullptr, h->d_rowptr, h->d_col, bad);
  hipLaunchKernelGGL(k_diagpos, dim3(grid1d(nne, 256)), dim3(256), 0, h->stream, h->d_rowptr, h->d_col, nne,
                     h->d_diagpos);
  MVF_HIP(h, hipStreamSynchronize(h->stream));
  MVF_HIP(h, hipGetLastError());
  return 0;
}

template <typename TE>
static void assemble_rows(mvf_handle *h, const TE *Ke, const TE *Me) {
  const int me = h->et.ne;
  const unsigned grid = (unsigned)((h->nne + 3) / 4);
  if (np_for(me) == 64)
    hipLaunchKernelGGL((k_assemble_rows<TE, 64>), dim3(grid), dim3(256), 0, h->stream, h->d_gne, me, h->d_rseg,
                       h->d_rcontrib, h->d_rowptr, h->d_col, (int64_t)h->nne, Ke, Me, h->d_K, h->d_M, h->d_K32,
                       h->d_M32);
  else
    hipLaunchKernelGGL((k_assemble_rows<TE, 256>), dim3(grid), dim3(256), 0, h->stream, h->d_gne, me, h->d_rseg,
                       h->d_rcontrib, h->d_rowptr, h->d_col, (int64_t)h->nne, Ke, Me, h->d_K, h->d_M, h->d_K32,
                       h->d_M32);
}

// --------------------------------------------------------------- assembly --
__device__ inline int64_t packed_index(int me, int i, int j) {
  if (i > j) {
    int t = i; i = j; j = t;
  }
  return (int64_t)i * me - (int64_t)i * (i - 1) / 2 + (j - i);
}

// K and M entry n = sum over its contributions in element order
template <typename TE>
__global__ void k_assemble(const int64_t *seg, const int64_t *contrib, int64_t nnz, int me, const TE *Ke,
                           const TE *Me, double *K, double *M, float *K32, float *M32) {
  int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nnz) return;
  int64_t np = (int64_t)me * (me + 1) / 2;
  double k = 0.0, m = 0.0;
  for (int64_t t = seg[n]; t < seg[n + 1]; t++) {
    int64_t c = contrib[t];
    int64_t e = c / ((int64_t)me * me);
    int r = (int)(c % ((int64_t)me * me));
    int64_t p = e * np + packed_index(me, r / me, r % me);
    k += (double)Ke[p];
    m += (double)Me[p];
  }
  K[n] = k;
  M[n] = m;
  K32[n] = (float)k;
  M32[n] = (float)m;
}

int assemble_operators(mvf_handle *h) {
  int64_t nnz = h->nnz;
  if (!h->d_K) {
    MVF_TRY(dalloc(h, &h->d_K, nnz));
    MVF_TRY(dalloc(h, &h->d_M, nnz));
    MVF_TRY(dalloc(h, &h->d_K32, nnz));
    MVF_TRY(dalloc(h, &h->d_M32, nnz));
  }
  if (h->setup_rows)
    assemble_rows<double>(h, h->d_Ke, h->d_Me);
  else
    hipLaunchKernelGGL(k_assemble<double>, dim3(grid1d(nnz, 256)), dim3(256), 0, h->stream, h->d_seg, h->d_contrib,
                       nnz, h->et.ne, h->d_Ke, h->d_Me, h->d_K, h->d_M, h->d_K32, h->d_M32);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

// FP32-integrated element matrices into the operators
int assemble_operators32(mvf_handle *h, const float *Ke, const float *Me) {
  int64_t nnz = h->nnz;
  if (h->setup_rows)
    assemble_rows<float>(h, Ke, Me);
  else
    hipLaunchKernelGGL(k_assemble<float>, dim3(grid1d(nnz, 256)), dim3(256), 0, h->stream, h->d_seg, h->d_contrib,
                       nnz, h->et.ne, Ke, Me, h->d_K, h->d_M, h->d_K32, h->d_M32);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

// RHS column: b[dof*ncol + col0 + d] = sum of be[e][im][d] over the DOF's contributions
__global__ void k_assemble_rhs(const int64_t *rseg, const int64_t *rcontrib, int64_t nne, const double *be,
                               double *b, int ncol, int col0) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nne) return;
  double s[4] = {0, 0, 0, 0};
  for (int64_t t = rseg[r]; t < rseg[r + 1]; t++) {
    int64_t c = rcontrib[t];  // e*me + im
    const double *p = be + c * 4;
    s[0] += p[0]; s[1] += p[1]; s[2] += p[2]; s[3] += p[3];
  }
  double *o = b + 2 * (r * ncol + col0);
  o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; o[3] = s[3];
}

int assemble_rhs(mvf_handle *h, const double *d_be, double *d_b, int ncol, int col0) {
  hipLaunchKernelGGL(k_assemble_rhs, dim3(grid1d(h->nne, 256)), dim3(256), 0, h->stream, h->d_rseg, h->d_rcontrib,
                     (int64_t)h->nne, d_be, d_b, ncol, col0);
  MVF_HIP(h, hipGetLastError());
  return 0;
}

}  // namespace mvf
