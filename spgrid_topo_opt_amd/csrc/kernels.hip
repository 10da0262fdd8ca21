// kernels.hip -- gfx950 kernels of the narrow-band elasticity solve.
//
// Threads map to nodes (64 nodes = one 4x4x4 block = one wavefront).  Node
// neighbourhoods never leave the 3x3x3 ring of blocks addressed through the
// per-block 27-neighbour table; absent blocks read as zero, as unallocated
// SPGrid pages do in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstdlib>

#include "kernels.hpp"

namespace spg {

__constant__ double cK0d[576];
__constant__ float cK0f[576];

CellConsts cell_consts(double mu, double lambda) {
  return CellConsts{9.0 * mu / 72.0,  4.5 * lambda / 72.0, 4.5 * mu / 72.0,
                    1.5 * mu / 72.0,  1.5 * lambda / 72.0, (2.0 * mu + 0.5 * lambda) / 72.0};
}

template <class T>
struct KT;
template <>
struct KT<float> {
  __device__ static inline float k(int e) { return cK0f[e]; }
};
template <>
struct KT<double> {
  __device__ static inline double k(int e) { return cK0d[e]; }
};

// Halo staging tables, the same for every block: entry (spoke << 8) | local
// index of the block-relative node (x,y,z) a halo position stands for.
// cHalo6: x,y,z in [-1,4] (6^3, apply); cHalo5: x,y,z in [-1,3] (5^3 cells,
// apply); cHaloC[c]: x,y,z in [p-1, p+3] around the nodes of colour c.
__constant__ int cHalo6[216];
__constant__ int cHalo5[125];
__constant__ int cHaloC[8 * 125];

static int host_entry(int x, int y, int z) {
  auto side = [](int v) { return v < 0 ? 0 : (v > 3 ? 2 : 1); };
  const int sp = side(x) * 9 + side(y) * 3 + side(z);
  const int q = ((x & 3) << 4) | ((y & 3) << 2) | (z & 3);
  return (sp << 8) | q;
}

hipError_t upload_k0(const double* K0) {
  float f[576];
  for (int i = 0; i < 576; ++i) f[i] = (float)K0[i];
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cK0d), K0, sizeof(double) * 576, 0, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cK0f), f, sizeof(float) * 576, 0, hipMemcpyHostToDevice);
  int h6[216], h5[125], hc[8 * 125];
  for (int h = 0; h < 216; ++h) h6[h] = host_entry(h / 36 - 1, (h / 6) % 6 - 1, h % 6 - 1);
  for (int h = 0; h < 125; ++h) h5[h] = host_entry(h / 25 - 1, (h / 5) % 5 - 1, h % 5 - 1);
  for (int c = 0; c < 8; ++c) {
    const int px = (c >> 2) & 1, py = (c >> 1) & 1, pz = c & 1;
    for (int h = 0; h < 125; ++h) hc[c * 125 + h] = host_entry(px - 1 + h / 25, py - 1 + (h / 5) % 5, pz - 1 + h % 5);
  }
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cHalo6), h6, sizeof(h6), 0, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cHalo5), h5, sizeof(h5), 0, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cHaloC), hc, sizeof(hc), 0, hipMemcpyHostToDevice);
  return e;
}

// fused multiply-add in the working precision; row sums are written as FMA
// chains onto their accumulator (a*b + c*d + e*f + acc would cost a multiply
// and an add per term triple more)
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

static constexpr int TPB = 256;
static inline unsigned nblk(int64_t n, int t = TPB) { return (unsigned)((n + t - 1) / t); }

__device__ inline int vidx(int blk, int c, int l) { return (blk * 3 + c) * 64 + l; }

// Neighbour addressing: the node at local (a,b,c) in [-1,4]^3 of block blk
// lives in block nbr[blk*27 + s], s = (ox+1)*9+(oy+1)*3+(oz+1) with
// o = -1/0/+1 when the coordinate is below/inside/above [0,3].
struct Ring {
  const int* t;  // nbr row of the block
  int blk;
};
__device__ __forceinline__ Ring make_ring(const int* __restrict__ nbr, int blk, int, int, int) {
  Ring R;
  R.t = nbr + blk * 27;
  R.blk = blk;
  return R;
}
__device__ __forceinline__ int ring_block(const Ring& R, int a, int b, int c) {
  const int ox = a < 0 ? 0 : (a > 3 ? 2 : 1), oy = b < 0 ? 0 : (b > 3 ? 2 : 1), oz = c < 0 ? 0 : (c > 3 ? 2 : 1);
  const int s = ox * 9 + oy * 3 + oz;
  return s == 13 ? R.blk : R.t[s];
}
__device__ inline int loc3(int a, int b, int c) { return ((a & 3) << 4) | ((b & 3) << 2) | (c & 3); }

// Workgroups are dispatched round-robin over the 8 XCDs; this renumbering
// gives each XCD one contiguous range of logical workgroups (and thus of
// Morton-ordered blocks), so neighbouring blocks share an L2.
__device__ __forceinline__ int xcd_chunk(int h, int n) {
  const int q = n >> 3, r = n & 7, x = h & 7, i = h >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}
__device__ __forceinline__ int ring_side(int x) { return x < 0 ? 0 : (x > 3 ? 2 : 1); }
__device__ __forceinline__ int ring_spoke(int x, int y, int z) {
  return ring_side(x) * 9 + ring_side(y) * 3 + ring_side(z);
}

// ---------------------------------------------------------------- level 0
// Row of the matrix-free operator at one node (SPGrid_Linear_Elasticity.h:
// 22-94 spokes; K_c = E_c * K0, K0 = mu0 K_mu + la0 K_la), optionally with the
// 3x3 self block (SPGrid_Gauss_Seidel.h:765-875).  Not scaled by dx.
template <class T, bool DIAG>
__device__ inline void row0(const int* __restrict__ nbr, const T* __restrict__ E, const T* __restrict__ u, int blk,
                            int l, T f[3], T d[6]) {
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const Ring R = make_ring(nbr, blk, li, lj, lk);
  T uu[27][3];
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    const int a = li + s / 9 - 1, b = lj + (s / 3) % 3 - 1, c = lk + s % 3 - 1;
    const int bb = ring_block(R, a, b, c);
    const int q = loc3(a, b, c);
#pragma unroll
    for (int w = 0; w < 3; ++w) uu[s][w] = bb >= 0 ? u[vidx(bb, w, q)] : T(0);
  }
  T f0 = 0, f1 = 0, f2 = 0;
  T d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0;
#pragma unroll
  for (int I = 0; I < 8; ++I) {
    const int ix = (I >> 2) & 1, iy = (I >> 1) & 1, iz = I & 1;
    const int bc = ring_block(R, li - ix, lj - iy, lk - iz);
    const T e = bc >= 0 ? E[bc * 64 + loc3(li - ix, lj - iy, lk - iz)] : T(0);
    T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int J = 0; J < 8; ++J) {
      const int jx = (J >> 2) & 1, jy = (J >> 1) & 1, jz = J & 1;
      const int s = (jx - ix + 1) * 9 + (jy - iy + 1) * 3 + (jz - iz + 1);
      const int kb = (I * 8 + J) * 9;
      t0 = fmaT(KT<T>::k(kb + 0), uu[s][0], t0);
      t0 = fmaT(KT<T>::k(kb + 1), uu[s][1], t0);
      t0 = fmaT(KT<T>::k(kb + 2), uu[s][2], t0);
      t1 = fmaT(KT<T>::k(kb + 3), uu[s][0], t1);
      t1 = fmaT(KT<T>::k(kb + 4), uu[s][1], t1);
      t1 = fmaT(KT<T>::k(kb + 5), uu[s][2], t1);
      t2 = fmaT(KT<T>::k(kb + 6), uu[s][0], t2);
      t2 = fmaT(KT<T>::k(kb + 7), uu[s][1], t2);
      t2 = fmaT(KT<T>::k(kb + 8), uu[s][2], t2);
    }
    f0 = fmaT(e, t0, f0);
    f1 = fmaT(e, t1, f1);
    f2 = fmaT(e, t2, f2);
    if (DIAG) {
      const int kd = (I * 8 + I) * 9;
      d0 += e * KT<T>::k(kd + 0);
      d1 += e * KT<T>::k(kd + 4);
      d2 += e * KT<T>::k(kd + 8);
      d3 += e * KT<T>::k(kd + 1);
      d4 += e * KT<T>::k(kd + 2);
      d5 += e * KT<T>::k(kd + 5);
    }
  }
  f[0] = f0;
  f[1] = f1;
  f[2] = f2;
  if (DIAG) {
    d[0] = d0; d[1] = d1; d[2] = d2; d[3] = d3; d[4] = d4; d[5] = d5;
  }
}

template <class T, int MODE>
__global__ void __launch_bounds__(TPB) k_apply(int nb, const int* __restrict__ nbr, const uint8_t* __restrict__ flags,
                                               const T* __restrict__ E, T dx, const T* __restrict__ in,
                                               const T* __restrict__ b, T* __restrict__ out) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  T f[3];
  row0<T, false>(nbr, E, in, blk, l, f, nullptr);
  if (MODE == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[vidx(blk, c, l)] = dx * f[c];
  } else if (MODE == 2) {
    const uint8_t fl = flags[gid];
#pragma unroll
    for (int c = 0; c < 3; ++c) out[vidx(blk, c, l)] = (fl & (2 << c)) ? T(0) : dx * f[c];
  } else {
    const uint8_t fl = flags[gid];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T r = b[vidx(blk, c, l)] - dx * f[c];
      out[vidx(blk, c, l)] = (fl & (2 << c)) ? T(0) : r;
    }
  }
}

// LDS-staged apply: a wavefront owns one block; it stages the 6^3 nodes and
// the 5^3 cells around the block, then lane l evaluates the row of node l
// from LDS.  All lanes visit the cells I = 0..7 in step, so K0[I][J] are
// wave-uniform constants.
static constexpr int AP_WAVES = 4;
// T: storage type of the vectors and moduli; A: arithmetic type of the row
// sums (FP32 vectors may be summed in FP64, which gfx950 issues at the same
// rate as FP32 on the vector ALUs)
// MODE 0: out = K in; 1: out = P (b - K in); 2: out = P K in; 3: as 2, and
// each block's partial sum of in . out goes to dpart[block]; 4 and 5: as 2
// and 3 in difference form (below).
template <class T, class A, int MODE>
__global__ void __launch_bounds__(64 * AP_WAVES) k_apply_lds(int nb, const int* __restrict__ nbr,
                                                             const uint8_t* __restrict__ flags,
                                                             const T* __restrict__ E, T dx, const T* __restrict__ in,
                                                             const T* __restrict__ b, T* __restrict__ out,
                                                             double* __restrict__ dpart = nullptr) {
  // the halo stays in the input type; values widen to A when read
  __shared__ T hu[AP_WAVES][3][216];
  __shared__ T he[AP_WAVES][128];
  __shared__ int ht[AP_WAVES][28];
  __shared__ int htab6[216], htab5[128];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = xcd_chunk(blockIdx.x, gridDim.x) * AP_WAVES + w;
  const bool live = blk < nb;
  if (lane < 27) ht[w][lane] = !live ? -1 : (lane == 13 ? blk : nbr[blk * 27 + lane]);
  if (threadIdx.x < 216) htab6[threadIdx.x] = cHalo6[threadIdx.x];
  if (threadIdx.x < 125) htab5[threadIdx.x] = cHalo5[threadIdx.x];
  __syncthreads();
  // FP32 residual: the node's flags and b go out with the halo loads (the
  // FP64 forms keep them at the end, where they cost no registers)
  constexpr bool kPre = MODE == 1 && sizeof(T) == 4 && sizeof(A) == 4;
  uint8_t pfl = 0;
  T pb[3] = {T(0), T(0), T(0)};
  {
    // 6^3 node halo (block coordinates -1..4) and 5^3 cell halo (-1..3)
    // from the tables in LDS; every pass's loads are issued before the LDS
    // writes
    T v[4][3], e[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int h = lane + 64 * i;
      const int t = h < 216 ? htab6[h] : 0;
      const int bb = h < 216 ? ht[w][t >> 8] : -1;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[i][c] = bb >= 0 ? in[vidx(bb, c, t & 255)] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int h = lane + 64 * i;
      const int t = h < 125 ? htab5[h] : 0;
      const int bb = h < 125 ? ht[w][t >> 8] : -1;
      e[i] = bb >= 0 ? E[bb * 64 + (t & 255)] : T(0);
    }
    if (kPre && live) {
      pfl = flags[blk * 64 + lane];
#pragma unroll
      for (int c = 0; c < 3; ++c) pb[c] = b[vidx(blk, c, lane)];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (lane + 64 * i < 216)
#pragma unroll
        for (int c = 0; c < 3; ++c) hu[w][c][lane + 64 * i] = v[i][c];
#pragma unroll
    for (int i = 0; i < 2; ++i)
      if (lane + 64 * i < 125) he[w][lane + 64 * i] = e[i];
  }
  __syncthreads();
  const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
  A uu[27][3];
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    const int h = (li + s / 9) * 36 + (lj + (s / 3) % 3) * 6 + (lk + s % 3);
#pragma unroll
    for (int c = 0; c < 3; ++c) uu[s][c] = (A)hu[w][c][h];
  }
  if (MODE >= 4) {
    // difference form: every element's 3x3 block rows sum to zero (rigid
    // translations), so K u = sum_m K_nm (u_m - u_n).  Where u is smooth the
    // differences are small and the row sums no longer cancel large terms,
    // so FP32 sums keep the accuracy the Krylov operator needs.
#pragma unroll
    for (int s = 0; s < 27; ++s)
      if (s != 13)
#pragma unroll
        for (int c = 0; c < 3; ++c) uu[s][c] -= uu[13][c];
#pragma unroll
    for (int c = 0; c < 3; ++c) uu[13][c] = A(0);
  }
  A f0 = 0, f1 = 0, f2 = 0;
#pragma unroll
  for (int I = 0; I < 8; ++I) {
    const int ix = (I >> 2) & 1, iy = (I >> 1) & 1, iz = I & 1;
    const A e = (A)he[w][(li - ix + 1) * 25 + (lj - iy + 1) * 5 + (lk - iz + 1)];
    A t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int J = 0; J < 8; ++J) {
      const int jx = (J >> 2) & 1, jy = (J >> 1) & 1, jz = J & 1;
      const int sp = (jx - ix + 1) * 9 + (jy - iy + 1) * 3 + (jz - iz + 1);
      const int kb = (I * 8 + J) * 9;
      t0 = fmaT(KT<A>::k(kb + 0), uu[sp][0], t0);
      t0 = fmaT(KT<A>::k(kb + 1), uu[sp][1], t0);
      t0 = fmaT(KT<A>::k(kb + 2), uu[sp][2], t0);
      t1 = fmaT(KT<A>::k(kb + 3), uu[sp][0], t1);
      t1 = fmaT(KT<A>::k(kb + 4), uu[sp][1], t1);
      t1 = fmaT(KT<A>::k(kb + 5), uu[sp][2], t1);
      t2 = fmaT(KT<A>::k(kb + 6), uu[sp][0], t2);
      t2 = fmaT(KT<A>::k(kb + 7), uu[sp][1], t2);
      t2 = fmaT(KT<A>::k(kb + 8), uu[sp][2], t2);
    }
    f0 = fmaT(e, t0, f0);
    f1 = fmaT(e, t1, f1);
    f2 = fmaT(e, t2, f2);
  }
  if (!live) return;  // wave-uniform: a wave owns one block
  const T f[3] = {(T)((A)dx * f0), (T)((A)dx * f1), (T)((A)dx * f2)};
  if (MODE == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[vidx(blk, c, lane)] = f[c];
  } else {
    const uint8_t fl = kPre ? pfl : flags[blk * 64 + lane];
    double d = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v = MODE >= 2 ? f[c] : (kPre ? pb[c] : b[vidx(blk, c, lane)]) - f[c];
      const T o = (fl & (2 << c)) ? T(0) : v;
      out[vidx(blk, c, lane)] = o;
      if (MODE == 3 || MODE == 5) d += (double)hu[w][c][(li + 1) * 36 + (lj + 1) * 6 + (lk + 1)] * (double)o;
    }
    if (MODE == 3 || MODE == 5) {
      // one partial per wavefront (= per block)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) d += __shfl_down(d, o, 64);
      if (lane == 0) dpart[blk] = d;
    }
  }
}

// ------------------------------------------------ cell-wise (strain-form) apply
// K0 = mu0 K_mu + la0 K_la in the 2x2x2 Walsh basis of a cell (modes m =
// products of +-1 per axis: the 3 gradients, 3 twists, 1 trilinear and the
// mean, per component) is block diagonal with 45 non-zeros and no mean
// coupling.  Written in strain form it is, per unit modulus (constants
// below, all / 72):
//   gradients   out(d,d) = 9 mu0 g(d,d) + 4.5 la0 tr g;
//               out(d,v) = out(v,d) = 4.5 mu0 (g(d,v) + g(v,d))
//   twists      the triple {(yz,x), (xz,y), (xy,z)}: 1.5 mu0 (t_i + sum t);
//               the pairs {(yz,y),(xz,x)}, {(yz,z),(xy,x)}, {(xz,z),(xy,y)}:
//               4.5 mu0 a + 1.5 la0 (a + b)
//   trilinear   (2 mu0 + 0.5 la0) t3
// and the node contributions are the transposed transform.  Every mode is
// formed from edge differences (exact in FP32 for smooth fields), and the
// symmetric gradient sums pair the two edge differences of each edge pair
// before adding the pairs, so rigid translations and rotations cancel
// exactly and FP32 sums keep the accuracy of the row sums
// (SPGrid_Linear_Elasticity.h:22-213 computes the same product node-wise).
// (constants of the strain-form cell operator: CellConsts, per material in LevelDev::cell)
template <class T>
struct CellK {
  T d, t, s, w, q, c3;
};

// f[J][v] += (E K0 a)[J][v] for one cell (corners J = x*4+y*2+z)
template <class T>
__device__ __forceinline__ void cell_strain_apply(const T (&a)[8][3], const CellK<T>& k, T (&f)[8][3]) {
  // edge differences along x (index y*2+z), y (x*2+z), z (x*2+y), per component
  T Dx[4][3], Dy[4][3], Dz[4][3];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        Dx[p * 2 + q][v] = a[4 + p * 2 + q][v] - a[p * 2 + q][v];
        Dy[p * 2 + q][v] = a[p * 4 + 2 + q][v] - a[p * 4 + q][v];
        Dz[p * 2 + q][v] = a[p * 4 + q * 2 + 1][v] - a[p * 4 + q * 2][v];
      }
  const T gxx = (Dx[0][0] + Dx[1][0]) + (Dx[2][0] + Dx[3][0]);
  const T gyy = (Dy[0][1] + Dy[1][1]) + (Dy[2][1] + Dy[3][1]);
  const T gzz = (Dz[0][2] + Dz[1][2]) + (Dz[2][2] + Dz[3][2]);
  // symmetric sums, edge pair by edge pair: y-edge (x=a,z) with x-edge (y=a,z)
  const T Sxy = ((Dy[0][0] + Dx[0][1]) + (Dy[1][0] + Dx[1][1])) + ((Dy[2][0] + Dx[2][1]) + (Dy[3][0] + Dx[3][1]));
  // z-edge (x=a,y) with x-edge (y,z=a)
  const T Sxz = ((Dz[0][0] + Dx[0][2]) + (Dz[1][0] + Dx[2][2])) + ((Dz[2][0] + Dx[1][2]) + (Dz[3][0] + Dx[3][2]));
  // z-edge (x,y=a) with y-edge (x,z=a)
  const T Syz = ((Dz[0][1] + Dy[0][2]) + (Dz[1][1] + Dy[1][2])) + ((Dz[2][1] + Dy[2][2]) + (Dz[3][1] + Dy[3][2]));
  T Byz[3], Bxz[3], Bxy[3], B3[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const T zy0 = Dz[1][v] - Dz[0][v], zy1 = Dz[3][v] - Dz[2][v];  // along y of the z-edges, x = 0 / 1
    Byz[v] = zy0 + zy1;
    Bxz[v] = (Dz[2][v] - Dz[0][v]) + (Dz[3][v] - Dz[1][v]);
    Bxy[v] = (Dy[2][v] - Dy[0][v]) + (Dy[3][v] - Dy[1][v]);
    B3[v] = zy1 - zy0;
  }
  // operator in the mode basis
  const T tr = (gxx + gyy) + gzz;
  const T ttr = k.t * tr;
  T G[3][3];  // gradient-mode outputs G[d][v]
  G[0][0] = fmaT(k.d, gxx, ttr);
  G[1][1] = fmaT(k.d, gyy, ttr);
  G[2][2] = fmaT(k.d, gzz, ttr);
  G[0][1] = G[1][0] = k.s * Sxy;
  G[0][2] = G[2][0] = k.s * Sxz;
  G[1][2] = G[2][1] = k.s * Syz;
  const T st = (Byz[0] + Bxz[1]) + Bxy[2];
  T Oyz[3], Oxz[3], Oxy[3];
  Oyz[0] = k.w * (Byz[0] + st);
  Oxz[1] = k.w * (Bxz[1] + st);
  Oxy[2] = k.w * (Bxy[2] + st);
  {
    const T p1 = k.q * (Byz[1] + Bxz[0]);
    Oyz[1] = fmaT(k.s, Byz[1], p1);
    Oxz[0] = fmaT(k.s, Bxz[0], p1);
    const T p2 = k.q * (Byz[2] + Bxy[0]);
    Oyz[2] = fmaT(k.s, Byz[2], p2);
    Oxy[0] = fmaT(k.s, Bxy[0], p2);
    const T p3 = k.q * (Bxz[2] + Bxy[1]);
    Oxz[2] = fmaT(k.s, Bxz[2], p3);
    Oxy[1] = fmaT(k.s, Bxy[1], p3);
  }
  // back to the corners: f_J = sum_m c_m(J) out_m, c_m(J) = prod_{axis in m} (J_axis ? +1 : -1)
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const T O3 = k.c3 * B3[v];
    // along z first: per (x,y) corner pair, the modes with and without z
    //   lo = A - Bz, hi = A + Bz with A = sx gx + sy gy + sx sy gxy, Bz = gz + sx gxz + sy gyz + sx sy O3
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const T sxv = x ? T(1) : T(-1), syv = y ? T(1) : T(-1);
        const T A = fmaT(sxv, G[0][v], fmaT(syv, G[1][v], (x == y) ? Oxy[v] : -Oxy[v]));
        const T Bz = fmaT(sxv, Oxz[v], fmaT(syv, Oyz[v], G[2][v] + ((x == y) ? O3 : -O3)));
        f[x * 4 + y * 2 + 0][v] += A - Bz;
        f[x * 4 + y * 2 + 1][v] += A + Bz;
      }
  }
}

// Stiffness apply cell by cell: a wavefront owns a block, stages the 6^3
// node and 5^3 cell halos, evaluates the 125 cells around the block (two
// passes of 64 lanes), scatters their corner forces into an LDS copy of the
// 6^3 nodes and reads its own 64 rows.  Modes as k_apply_lds (0: K in;
// 1: P (b - K in); 2: P K in; 3: as 2 with per-block in . out partials).
template <class T, int MODE>
__global__ void __launch_bounds__(64 * AP_WAVES) k_apply_cell(int nb, const int* __restrict__ nbr,
                                                              const uint8_t* __restrict__ flags,
                                                              const T* __restrict__ E, CellConsts kc, T dx,
                                                              const T* __restrict__ in, const T* __restrict__ b,
                                                              T* __restrict__ out, double* __restrict__ dpart = nullptr) {
  __shared__ T hu[AP_WAVES][3][216];
  __shared__ T acc[AP_WAVES][3][216];
  __shared__ T he[AP_WAVES][128];
  __shared__ int ht[AP_WAVES][28];
  __shared__ int htab6[216], htab5[128];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = xcd_chunk(blockIdx.x, gridDim.x) * AP_WAVES + w;
  const bool live = blk < nb;
  if (lane < 27) ht[w][lane] = !live ? -1 : (lane == 13 ? blk : nbr[blk * 27 + lane]);
  if (threadIdx.x < 216) htab6[threadIdx.x] = cHalo6[threadIdx.x];
  if (threadIdx.x < 125) htab5[threadIdx.x] = cHalo5[threadIdx.x];
  __syncthreads();
  {
    T v[4][3], e[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int h = lane + 64 * i;
      const int t = h < 216 ? htab6[h] : 0;
      const int bb = h < 216 ? ht[w][t >> 8] : -1;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[i][c] = bb >= 0 ? in[vidx(bb, c, t & 255)] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int h = lane + 64 * i;
      const int t = h < 125 ? htab5[h] : 0;
      const int bb = h < 125 ? ht[w][t >> 8] : -1;
      e[i] = bb >= 0 ? E[bb * 64 + (t & 255)] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (lane + 64 * i < 216)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          hu[w][c][lane + 64 * i] = v[i][c];
          acc[w][c][lane + 64 * i] = T(0);
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
      if (lane + 64 * i < 125) he[w][lane + 64 * i] = e[i];
  }
  __syncthreads();
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const int ci = lane + 64 * pass;
    if (ci < 125) {
      const int cx = ci / 25, cy = (ci / 5) % 5, cz = ci % 5;  // halo cell = halo node of its corner J = 0
      const int n0 = cx * 36 + cy * 6 + cz;
      const T ee = he[w][ci] * dx;
      const CellK<T> k{ee * (T)kc.d, ee * (T)kc.t, ee * (T)kc.s, ee * (T)kc.w, ee * (T)kc.q, ee * (T)kc.c3};
      T a[8][3], f[8][3];
#pragma unroll
      for (int J = 0; J < 8; ++J) {
        const int o = ((J >> 2) & 1) * 36 + ((J >> 1) & 1) * 6 + (J & 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          a[J][c] = hu[w][c][n0 + o];
          f[J][c] = T(0);
        }
      }
      cell_strain_apply<T>(a, k, f);
      // corner J of the 64 cells of a pass lands on 64 distinct nodes, and a
      // wavefront's LDS instructions execute in order: a read-add-write per
      // corner, with the compiler kept from moving one corner's accesses
      // past another's (other lanes may touch the same node at another
      // corner), is exact.  LDS float atomics measured 8x slower here.
#pragma unroll
      for (int J = 0; J < 8; ++J) {
        const int o = ((J >> 2) & 1) * 36 + ((J >> 1) & 1) * 6 + (J & 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[w][c][n0 + o] += f[J][c];
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
      }
    }
  }
  __syncthreads();
  if (!live) return;  // wave-uniform
  const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
  const int hn = (li + 1) * 36 + (lj + 1) * 6 + (lk + 1);
  const T f[3] = {acc[w][0][hn], acc[w][1][hn], acc[w][2][hn]};
  if (MODE == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[vidx(blk, c, lane)] = f[c];
  } else {
    const uint8_t fl = flags[blk * 64 + lane];
    double d = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v = MODE >= 2 ? f[c] : b[vidx(blk, c, lane)] - f[c];
      const T o = (fl & (2 << c)) ? T(0) : v;
      out[vidx(blk, c, lane)] = o;
      if (MODE == 3) d += (double)hu[w][c][hn] * (double)o;
    }
    if (MODE == 3) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) d += __shfl_down(d, o, 64);
      if (lane == 0) dpart[blk] = d;
    }
  }
}

// symmetric 3x3 inverse by minors with Dirichlet rows/cols handled as in
// SPGrid_Gauss_Seidel.h:1093-1144 / Multigrid_Hierarchy.h:64-101.
template <class T>
__device__ inline void dirichlet_inverse(T d[6], uint8_t fl) {
  if (fl & 2) { d[0] = 1; d[3] = 0; d[4] = 0; }
  if (fl & 4) { d[1] = 1; d[3] = 0; d[5] = 0; }
  if (fl & 8) { d[2] = 1; d[4] = 0; d[5] = 0; }
  const T m11 = d[1] * d[2] - d[5] * d[5], m12 = d[5] * d[4] - d[2] * d[3], m13 = d[3] * d[5] - d[4] * d[1];
  const T m22 = d[2] * d[0] - d[4] * d[4], m23 = d[4] * d[3] - d[0] * d[5], m33 = d[0] * d[1] - d[3] * d[3];
  const T det = d[0] * m11 + d[3] * m12 + d[4] * m13;
  const T inv = T(1) / det;
  d[0] = m11 * inv; d[1] = m22 * inv; d[2] = m33 * inv; d[3] = m12 * inv; d[4] = m13 * inv; d[5] = m23 * inv;
  if (fl & 2) { d[0] = 0; d[3] = 0; d[4] = 0; }
  if (fl & 4) { d[1] = 0; d[3] = 0; d[5] = 0; }
  if (fl & 8) { d[2] = 0; d[4] = 0; d[5] = 0; }
}

// u <- u + D^-1 (b - f) on the active, non-Dirichlet components of one node;
// Dirichlet components are forced to 0 (also on inactive nodes)
template <class T>
__device__ inline void gs_node_update_b(uint8_t fl, T d[6], T f0, T f1, T f2, int blk, int l, T u0, T u1, T u2,
                                        T b0, T b1, T b2, T* __restrict__ u) {
  if (fl & 1) {
    dirichlet_inverse(d, fl);
    const T r0 = b0 - f0, r1 = b1 - f1, r2 = b2 - f2;
    const T n0 = u0 + d[0] * r0 + d[3] * r1 + d[4] * r2;
    const T n1 = u1 + d[3] * r0 + d[1] * r1 + d[5] * r2;
    const T n2 = u2 + d[4] * r0 + d[5] * r1 + d[2] * r2;
    u[vidx(blk, 0, l)] = (fl & 2) ? T(0) : n0;
    u[vidx(blk, 1, l)] = (fl & 4) ? T(0) : n1;
    u[vidx(blk, 2, l)] = (fl & 8) ? T(0) : n2;
  } else if (fl & 0x0E) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (fl & (2 << c)) u[vidx(blk, c, l)] = T(0);
  }
}
// the same with b read here
template <class T>
__device__ inline void gs_node_update(uint8_t fl, T d[6], T f0, T f1, T f2, int blk, int l, T u0, T u1, T u2,
                                      const T* __restrict__ b, T* __restrict__ u) {
  gs_node_update_b<T>(fl, d, f0, f1, f2, blk, l, u0, u1, u2, b[vidx(blk, 0, l)], b[vidx(blk, 1, l)],
                      b[vidx(blk, 2, l)], u);
}

// the k-th node (0..7) of colour c in a 4x4x4 block
__device__ inline int colour_node(int colour, int k) {
  const int px = (colour >> 2) & 1, py = (colour >> 1) & 1, pz = colour & 1;
  const int i = px + 2 * ((k >> 2) & 1), j = py + 2 * ((k >> 1) & 1), kk = pz + 2 * (k & 1);
  return (i << 4) | (j << 2) | kk;
}

// ------------------------------------------------------- lane-octet helpers
// Gauss-Seidel kernels give every updated node an aligned octet of 8 lanes
// (lane bits 0..2).  octet_sum adds a value over the octet with three DPP
// steps; all 8 lanes receive the same bits.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_t(float x) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_t(double x) {
  const long long v = __double_as_longlong(x);
  const int lo = dpp_i<CTRL>((int)(v & 0xffffffffll)), hi = dpp_i<CTRL>((int)(v >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
template <class T>
__device__ __forceinline__ T octet_sum(T x) {
  x += dpp_t<0xB1>(x);   // quad_perm(1,0,3,2): partner lane ^ 1
  x += dpp_t<0x4E>(x);   // quad_perm(2,3,0,1): partner lane ^ 2
  x += dpp_t<0x141>(x);  // row_half_mirror: partner lane 7 - i of the octet
  return x;
}


// ------------------------------------------------- level-0 Gauss-Seidel
// Row of corner I of one cell, by reflection: the Q1 cube is symmetric under
// the mirror x_d -> 1 - x_d, so K0[I][J]_{vw} = s_v(I) s_w(I) K0[0][I^J]_{vw}
// with s_d(I) = -1 if bit d of I is set.  Only the 72 entries K0[0][*] are
// needed and they are wave-uniform (scalar registers).  `U(J)` returns the 3
// components of the cell's node J.  Result: the 3 rows times e (no dx).
template <class T, class LoadU>
__device__ __forceinline__ void corner_row(int I, T e, const LoadU& U, T& f0, T& f1, T& f2) {
  const T sx = (I & 4) ? T(-1) : T(1), sy = (I & 2) ? T(-1) : T(1), sz = (I & 1) ? T(-1) : T(1);
  // partial sums per (row v, component w); the signs s_v s_w apply once at the end
  T a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int Jp = 0; Jp < 8; ++Jp) {
    T v[3];
    U(I ^ Jp, v[0], v[1], v[2]);
    const int kb = Jp * 9;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int w = 0; w < 3; ++w) a[r][w] = fmaT(KT<T>::k(kb + 3 * r + w), v[w], a[r][w]);
  }
  const T ex = sx * e, ey = sy * e, ez = sz * e;
  f0 = sx * fmaT(ex, a[0][0], fmaT(ey, a[0][1], ez * a[0][2]));
  f1 = sy * fmaT(ex, a[1][0], fmaT(ey, a[1][1], ez * a[1][2]));
  f2 = sz * fmaT(ex, a[2][0], fmaT(ey, a[2][1], ez * a[2][2]));
}

// Octet reduction of the row and of the 3x3 self block
// D = sum_I e_I K0[I][I] (SPGrid_Gauss_Seidel.h:765-875), scaled by dx.
template <class T>
__device__ __forceinline__ void octet_row_and_diag(int I, T e, T dx, T& f0, T& f1, T& f2, T d[6]) {
  const T sx = (I & 4) ? T(-1) : T(1), sy = (I & 2) ? T(-1) : T(1), sz = (I & 1) ? T(-1) : T(1);
  f0 = dx * octet_sum(f0);
  f1 = dx * octet_sum(f1);
  f2 = dx * octet_sum(f2);
  const T S = dx * octet_sum(e), Sxy = dx * octet_sum(sx * sy * e), Sxz = dx * octet_sum(sx * sz * e),
          Syz = dx * octet_sum(sy * sz * e);
  d[0] = KT<T>::k(0) * S;
  d[1] = KT<T>::k(4) * S;
  d[2] = KT<T>::k(8) * S;
  d[3] = KT<T>::k(1) * Sxy;
  d[4] = KT<T>::k(2) * Sxz;
  d[5] = KT<T>::k(5) * Syz;
}

static constexpr int GS_WAVES = 4;  // blocks (wavefronts) per workgroup in the colour sweeps

// One colour of the level-0 multicolour sweep over all blocks
// (SPGrid_Gauss_Seidel.h:765-875).  A wavefront owns one block: it stages
// the 5^3 nodes around the block's 8 colour nodes in LDS, then lane
// (k, I) = (lane>>3, lane&7) adds the row of colour node k in the cell of
// which the node is corner I; the octet sums, and lane I=0 updates the node.
// A halo never holds nodes of the sweep's colour outside its own block, so
// waves do not race.
template <class T>
__global__ void __launch_bounds__(64 * GS_WAVES) k_gs0_colour(int nb, int colour, const int* __restrict__ nbr,
                                                              const uint8_t* __restrict__ flags,
                                                              const T* __restrict__ E, T dx, T* __restrict__ u,
                                                              const T* __restrict__ b) {
  __shared__ T hu[GS_WAVES][3][128];
  __shared__ int ht[GS_WAVES][28];
  __shared__ int htab[128];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = xcd_chunk(blockIdx.x, gridDim.x) * GS_WAVES + w;
  const bool live = blk < nb;
  const int px = (colour >> 2) & 1, py = (colour >> 1) & 1, pz = colour & 1;
  if (lane < 27) ht[w][lane] = !live ? -1 : (lane == 13 ? blk : nbr[blk * 27 + lane]);
  if (threadIdx.x < 125) htab[threadIdx.x] = cHaloC[colour * 125 + threadIdx.x];
  __syncthreads();
  {
    // the 5^3 halo of colour (px,py,pz) from the table in LDS; both passes'
    // loads are issued before one wait
    T v[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int h = lane + 64 * i;
      const int t = h < 125 ? htab[h] : 0;
      const int bb = h < 125 ? ht[w][t >> 8] : -1;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[i][c] = bb >= 0 ? u[vidx(bb, c, t & 255)] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
      if (lane + 64 * i < 125)
#pragma unroll
        for (int c = 0; c < 3; ++c) hu[w][c][lane + 64 * i] = v[i][c];
  }
  __syncthreads();
  const int k = lane >> 3, I = lane & 7;
  const int nx = 2 * ((k >> 2) & 1) + 1, ny = 2 * ((k >> 1) & 1) + 1, nz = 2 * (k & 1) + 1;  // halo coords
  const int cx = nx - ((I >> 2) & 1), cy = ny - ((I >> 1) & 1), cz = nz - (I & 1);           // cell corner
  const int ex = px - 1 + cx, ey = py - 1 + cy, ez = pz - 1 + cz;                             // block coords
  const int eb = ht[w][ring_spoke(ex, ey, ez)];
  const T e = eb >= 0 ? E[eb * 64 + loc3(ex, ey, ez)] : T(0);
  const T* H = &hu[w][0][0];
  auto U = [&](int J, T& v0, T& v1, T& v2) {
    const int h = ((cx + ((J >> 2) & 1)) * 5 + (cy + ((J >> 1) & 1))) * 5 + (cz + (J & 1));
    v0 = H[h];
    v1 = H[128 + h];
    v2 = H[256 + h];
  };
  T f0, f1, f2, d[6];
  corner_row<T>(I, e, U, f0, f1, f2);
  octet_row_and_diag<T>(I, e, dx, f0, f1, f2, d);
  if (I == 0 && live) {
    const int l = loc3(px - 1 + nx, py - 1 + ny, pz - 1 + nz);
    const int hn = (nx * 5 + ny) * 5 + nz;
    gs_node_update<T>(flags[blk * 64 + l], d, f0, f1, f2, blk, l, H[hn], H[128 + hn], H[256 + hn], b, u);
  }
}

// One colour of the level-0 sweep with a corner per wavefront: a 512-thread
// workgroup owns the 64 colour-c nodes of 8 consecutive blocks and stages
// the 5^3 nodes and 4^3 cells around each block's colour nodes once; wave I
// adds the rows of cell (node - I) for all 64 nodes, so K0[I][J] is
// wave-uniform (scalar registers).  The 8 partial rows meet in LDS and wave 0
// updates the nodes (SPGrid_Gauss_Seidel.h:765-875).  Races: as for
// k_gs0_colour, a halo never holds a colour-c node of another block.
static constexpr int UC_BLOCKS = 8;
template <class T>
__global__ void __launch_bounds__(512) k_gs0_colour_uc(int nb, int colour, const int* __restrict__ nbr,
                                                       const uint8_t* __restrict__ flags, const T* __restrict__ E,
                                                       T dx, T* __restrict__ u, const T* __restrict__ b) {
  __shared__ T hu[3][UC_BLOCKS * 125];
  __shared__ T he[UC_BLOCKS * 64];
  __shared__ T pf[8][3][64];
  __shared__ int ht[UC_BLOCKS][27];
  const int tid = threadIdx.x;
  const int blk0 = xcd_chunk(blockIdx.x, gridDim.x) * UC_BLOCKS;
  const int px = (colour >> 2) & 1, py = (colour >> 1) & 1, pz = colour & 1;
  if (tid < UC_BLOCKS * 27) {
    const int bl = tid / 27, sp = tid - bl * 27, blk = blk0 + bl;
    ht[bl][sp] = blk < nb ? (sp == 13 ? blk : nbr[blk * 27 + sp]) : -1;
  }
  __syncthreads();
  for (int h = tid; h < UC_BLOCKS * 125; h += 512) {
    const int bl = h / 125, hh = h - bl * 125;
    const int t = cHaloC[colour * 125 + hh];
    const int bb = ht[bl][t >> 8];
    const int q = t & 255;
#pragma unroll
    for (int c = 0; c < 3; ++c) hu[c][h] = bb >= 0 ? u[vidx(bb, c, q)] : T(0);
  }
  {
    const int bl = tid >> 6, hh = tid & 63;
    const int x = px - 1 + (hh >> 4), y = py - 1 + ((hh >> 2) & 3), z = pz - 1 + (hh & 3);
    const int bb = ht[bl][ring_spoke(x, y, z)];
    he[tid] = bb >= 0 ? E[bb * 64 + loc3(x, y, z)] : T(0);
  }
  __syncthreads();
  const int lane = tid & 63;
  const int I = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int bl = lane >> 3, k = lane & 7;
  const int nx = 2 * ((k >> 2) & 1) + 1, ny = 2 * ((k >> 1) & 1) + 1, nz = 2 * (k & 1) + 1;  // halo coords
  {
    const int ix = (I >> 2) & 1, iy = (I >> 1) & 1, iz = I & 1;
    const int cx = nx - ix, cy = ny - iy, cz = nz - iz;  // cell min corner (halo coords)
    const T e = he[bl * 64 + (cx * 4 + cy) * 4 + cz];
    T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int J = 0; J < 8; ++J) {
      const int h = bl * 125 + ((cx + ((J >> 2) & 1)) * 5 + (cy + ((J >> 1) & 1))) * 5 + (cz + (J & 1));
      const T v0 = hu[0][h], v1 = hu[1][h], v2 = hu[2][h];
      const int kb = (I * 8 + J) * 9;
      t0 = fmaT(KT<T>::k(kb + 0), v0, t0);
      t0 = fmaT(KT<T>::k(kb + 1), v1, t0);
      t0 = fmaT(KT<T>::k(kb + 2), v2, t0);
      t1 = fmaT(KT<T>::k(kb + 3), v0, t1);
      t1 = fmaT(KT<T>::k(kb + 4), v1, t1);
      t1 = fmaT(KT<T>::k(kb + 5), v2, t1);
      t2 = fmaT(KT<T>::k(kb + 6), v0, t2);
      t2 = fmaT(KT<T>::k(kb + 7), v1, t2);
      t2 = fmaT(KT<T>::k(kb + 8), v2, t2);
    }
    pf[I][0][lane] = e * t0;
    pf[I][1][lane] = e * t1;
    pf[I][2][lane] = e * t2;
  }
  __syncthreads();
  if (I != 0) return;
  const int blk = blk0 + bl;
  if (blk >= nb) return;
  T f0 = 0, f1 = 0, f2 = 0, d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    f0 += pf[c][0][lane];
    f1 += pf[c][1][lane];
    f2 += pf[c][2][lane];
    const T e = he[bl * 64 + ((nx - ((c >> 2) & 1)) * 4 + (ny - ((c >> 1) & 1))) * 4 + (nz - (c & 1))];
    const int kd = (c * 8 + c) * 9;
    d0 += e * KT<T>::k(kd + 0);
    d1 += e * KT<T>::k(kd + 4);
    d2 += e * KT<T>::k(kd + 8);
    d3 += e * KT<T>::k(kd + 1);
    d4 += e * KT<T>::k(kd + 2);
    d5 += e * KT<T>::k(kd + 5);
  }
  const int l = loc3(px - 1 + nx, py - 1 + ny, pz - 1 + nz);
  const int hn = bl * 125 + (nx * 5 + ny) * 5 + nz;
  T d[6] = {dx * d0, dx * d1, dx * d2, dx * d3, dx * d4, dx * d5};
  gs_node_update<T>(flags[blk * 64 + l], d, dx * f0, dx * f1, dx * f2, blk, l, hu[0][hn], hu[1][hn], hu[2][hn], b, u);
}

// sum over the 4 lanes of an aligned quad (lane bits 0..1)
template <class T>
__device__ __forceinline__ T quad_sum(T x) {
  x += dpp_t<0xB1>(x);  // partner lane ^ 1
  x += dpp_t<0x4E>(x);  // partner lane ^ 2
  return x;
}

// One colour of the level-0 sweep with 4 lanes per node: a wavefront owns
// the colour-c nodes of 2 blocks; lane (k, q) adds the rows of colour node k
// in its cells with corners I = q and q|4 (the reflection form of
// corner_row), the quad sums, and lane q = 0 updates the node.  A 256-thread
// workgroup stages the 5^3 halos of its 8 blocks.  Same race argument as
// k_gs0_colour.
// With `dist` given, only nodes at distance >= dmin from the boundary band
// are updated (the far part of a sweep that runs beside the band phase).
template <class T, int QB>
__global__ void __launch_bounds__(32 * QB) k_gs0_colour4(int nb, int colour, const int* __restrict__ nbr,
                                                         const uint8_t* __restrict__ flags, const T* __restrict__ E,
                                                         T dx, T* __restrict__ u, const T* __restrict__ b,
                                                         const uint8_t* __restrict__ dist, int dmin,
                                                         const T* __restrict__ bcm = nullptr,
                                                         const uint8_t* __restrict__ fcm = nullptr) {
  static_assert(QB >= 4, "the 125-entry halo table is filled by one pass of the workgroup");
  __shared__ T hu[QB][3][128];
  __shared__ int ht[QB][27];
  __shared__ int htab[128];
  const int tid = threadIdx.x;
  const int blk0 = xcd_chunk(blockIdx.x, gridDim.x) * QB;
  const int px = (colour >> 2) & 1, py = (colour >> 1) & 1, pz = colour & 1;
  if (tid < QB * 27) {
    const int bl = tid / 27, sp = tid - bl * 27, blk = blk0 + bl;
    ht[bl][sp] = blk < nb ? (sp == 13 ? blk : nbr[blk * 27 + sp]) : -1;
  }
  if (tid < 125) htab[tid] = cHaloC[colour * 125 + tid];
  __syncthreads();
  // lane roles after staging: block bl, node k of the colour, corner pair q
  const int bl = tid >> 5, k = (tid >> 2) & 7, q = tid & 3;
  const int blk = blk0 + bl;
  const int nx = 2 * ((k >> 2) & 1) + 1, ny = 2 * ((k >> 1) & 1) + 1, nz = 2 * (k & 1) + 1;  // halo coords
  const int l = loc3(px - 1 + nx, py - 1 + ny, pz - 1 + nz);
  T ee[2], bv[3] = {T(0), T(0), T(0)};
  uint8_t fl = 0;
  bool skip = false;
  {
    // 32 threads per block stage its 5^3 halo (4 passes, table in LDS).  The
    // halo loads, this lane's two cell moduli and (lane q = 0) the node's
    // flags and b are all issued before one wait.
    const int rs = tid & 31;
    T v[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int hh = rs + 32 * i;
      const int t = hh < 125 ? htab[hh] : 0;
      const int bb = hh < 125 ? ht[bl][t >> 8] : -1;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[i][c] = bb >= 0 ? u[vidx(bb, c, t & 255)] : T(0);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int I = q | (r << 2);
      const int ex = px - 1 + nx - ((I >> 2) & 1), ey = py - 1 + ny - ((I >> 1) & 1), ez = pz - 1 + nz - (I & 1);
      const int eb = ht[bl][ring_spoke(ex, ey, ez)];
      ee[r] = eb >= 0 ? E[eb * 64 + loc3(ex, ey, ez)] : T(0);
    }
    if (q == 0 && blk < nb) {
      if (bcm) {
        // colour-major copies: the colour's 8 nodes of a block are adjacent
        const int64_t cb = (int64_t)colour * nb + blk;
        fl = fcm[cb * 8 + k];
#pragma unroll
        for (int c = 0; c < 3; ++c) bv[c] = bcm[(cb * 3 + c) * 8 + k];
      } else {
        fl = flags[blk * 64 + l];
#pragma unroll
        for (int c = 0; c < 3; ++c) bv[c] = b[vidx(blk, c, l)];
      }
      if (dist) skip = dist[blk * 64 + l] < dmin;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (rs + 32 * i < 125)
#pragma unroll
        for (int c = 0; c < 3; ++c) hu[bl][c][rs + 32 * i] = v[i][c];
  }
  __syncthreads();
  const T* H = &hu[bl][0][0];
  T f0 = 0, f1 = 0, f2 = 0;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int I = q | (r << 2);
    const int cx = nx - ((I >> 2) & 1), cy = ny - ((I >> 1) & 1), cz = nz - (I & 1);  // cell corner (halo)
    auto U = [&](int J, T& v0, T& v1, T& v2) {
      const int h = ((cx + ((J >> 2) & 1)) * 5 + (cy + ((J >> 1) & 1))) * 5 + (cz + (J & 1));
      v0 = H[h];
      v1 = H[128 + h];
      v2 = H[256 + h];
    };
    T g0, g1, g2;
    corner_row<T>(I, ee[r], U, g0, g1, g2);
    f0 += g0;
    f1 += g1;
    f2 += g2;
  }
  // self block: corners q and q|4 differ in the x mirror only
  const T sy = (q & 2) ? T(-1) : T(1), sz = (q & 1) ? T(-1) : T(1);
  const T es = ee[0] + ee[1], ed = ee[0] - ee[1];
  f0 = dx * quad_sum(f0);
  f1 = dx * quad_sum(f1);
  f2 = dx * quad_sum(f2);
  const T S = dx * quad_sum(es), Sxy = dx * quad_sum(sy * ed), Sxz = dx * quad_sum(sz * ed),
          Syz = dx * quad_sum(sy * sz * es);
  if (q != 0 || blk >= nb || skip) return;
  T d[6] = {KT<T>::k(0) * S, KT<T>::k(4) * S, KT<T>::k(8) * S, KT<T>::k(1) * Sxy, KT<T>::k(2) * Sxz, KT<T>::k(5) * Syz};
  const int hn = (nx * 5 + ny) * 5 + nz;
  gs_node_update_b<T>(fl, d, f0, f1, f2, blk, l, H[hn], H[128 + hn], H[256 + hn], bv[0], bv[1], bv[2], u);
}

// Level-0 update of list node `s` (blk*64+l) by the 8 lanes of an octet,
// reading straight from global memory (boundary lists).
template <class T>
__device__ __forceinline__ void gs0_octet_global(bool live, int s, int I, const int* __restrict__ nbr,
                                                 const uint8_t* __restrict__ flags, const T* __restrict__ E, T dx,
                                                 T* __restrict__ u, const T* __restrict__ b) {
  const int blk = live ? (s >> 6) : 0, l = s & 63;
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const int cx = li - ((I >> 2) & 1), cy = lj - ((I >> 1) & 1), cz = lk - (I & 1);  // block coords of the cell
  const int* t = nbr + blk * 27;
  auto ring = [&](int x, int y, int z) {
    const int sp = ring_spoke(x, y, z);
    return !live ? -1 : (sp == 13 ? blk : t[sp]);
  };
  const int eb = ring(cx, cy, cz);
  const T e = eb >= 0 ? E[eb * 64 + loc3(cx, cy, cz)] : T(0);
  auto U = [&](int J, T& v0, T& v1, T& v2) {
    const int x = cx + ((J >> 2) & 1), y = cy + ((J >> 1) & 1), z = cz + (J & 1);
    const int bb = ring(x, y, z);
    const int q = loc3(x, y, z);
    v0 = bb >= 0 ? u[vidx(bb, 0, q)] : T(0);
    v1 = bb >= 0 ? u[vidx(bb, 1, q)] : T(0);
    v2 = bb >= 0 ? u[vidx(bb, 2, q)] : T(0);
  };
  T f0, f1, f2, d[6];
  corner_row<T>(I, e, U, f0, f1, f2);
  octet_row_and_diag<T>(I, e, dx, f0, f1, f2, d);
  if (I == 0 && live)
    gs_node_update<T>(flags[blk * 64 + l], d, f0, f1, f2, blk, l, u[vidx(blk, 0, l)], u[vidx(blk, 1, l)],
                      u[vidx(blk, 2, l)], b, u);
}

// colour order of the sweeps: nibble i = i-th colour
struct ColourOrder {
  unsigned v;
  __host__ __device__ int at(int i) const { return (int)((v >> (4 * i)) & 7u); }
};
// boundary lists of one level, concatenated by colour
struct BndLists {
  const int* list;
  int off[9];
};

static constexpr int BND_THREADS = 1024;  // fused boundary smoother: one workgroup

// All boundary sweeps of one smoothing phase in a single workgroup: colour
// steps are separated by workgroup barriers, which order the global-memory
// updates between the waves of the workgroup (Multigrid_Hierarchy.h:577-586
// boundary pre/post sweeps).
template <class T>
__global__ void __launch_bounds__(BND_THREADS) k_gs0_boundary(BndLists L, ColourOrder ord, int sweeps,
                                                              const int* __restrict__ nbr,
                                                              const uint8_t* __restrict__ flags,
                                                              const T* __restrict__ E, T dx, T* __restrict__ u,
                                                              const T* __restrict__ b) {
  const int oct = threadIdx.x >> 3, I = threadIdx.x & 7;
  for (int sw = 0; sw < sweeps; ++sw)
    for (int ci = 0; ci < 8; ++ci) {
      const int c = ord.at(ci);
      const int n0 = L.off[c], n1 = L.off[c + 1];
      for (int base = n0; base < n1; base += BND_THREADS / 8) {
        const int i = base + oct;
        const bool live = i < n1;
        gs0_octet_global<T>(live, live ? L.list[i] : 0, I, nbr, flags, E, dx, u, b);
      }
      __syncthreads();
    }
}
// one colour of a boundary list over many workgroups (long lists)
// One level-0 colour step over a list of nodes (the boundary lists), one
// octet per node like gs0_octet_global, with the 27 neighbour slots of each
// list node read from a precomputed table (k_bnd_nslot) instead of walking
// the block-neighbour table.
template <class T>
__global__ void __launch_bounds__(256) k_gs0_list_tab(const int* __restrict__ list, const int* __restrict__ nslot,
                                                      int n, const uint8_t* __restrict__ flags,
                                                      const T* __restrict__ E, T dx, T* __restrict__ u,
                                                      const T* __restrict__ b) {
  const int i = blockIdx.x * 32 + (threadIdx.x >> 3), I = threadIdx.x & 7;
  const bool live = i < n;
  // the updating lane's own node data do not depend on the row: load first
  const bool upd = live && I == 0;
  const int node = upd ? list[i] : 0, nblk0 = node >> 6, nl = node & 63;
  const uint8_t nfl = upd ? flags[node] : 0;
  T s0 = 0, s1 = 0, s2 = 0, c0 = 0, c1 = 0, c2 = 0;
  if (upd) {
    s0 = u[vidx(nblk0, 0, nl)];
    s1 = u[vidx(nblk0, 1, nl)];
    s2 = u[vidx(nblk0, 2, nl)];
    c0 = b[vidx(nblk0, 0, nl)];
    c1 = b[vidx(nblk0, 1, nl)];
    c2 = b[vidx(nblk0, 2, nl)];
  }
  const int* ns = nslot + (int64_t)(live ? i : 0) * 27;
  const int ix = (I >> 2) & 1, iy = (I >> 1) & 1, iz = I & 1;
  const int cs = live ? ns[(1 - ix) * 9 + (1 - iy) * 3 + (1 - iz)] : -1;  // cell with min corner node - I
  const T e = cs >= 0 ? E[cs] : T(0);
  auto U = [&](int J, T& v0, T& v1, T& v2) {
    const int sp = (((J >> 2) & 1) - ix + 1) * 9 + (((J >> 1) & 1) - iy + 1) * 3 + ((J & 1) - iz + 1);
    const int q = live ? ns[sp] : -1;
    v0 = q >= 0 ? u[vidx(q >> 6, 0, q & 63)] : T(0);
    v1 = q >= 0 ? u[vidx(q >> 6, 1, q & 63)] : T(0);
    v2 = q >= 0 ? u[vidx(q >> 6, 2, q & 63)] : T(0);
  };
  T f0, f1, f2, d[6];
  corner_row<T>(I, e, U, f0, f1, f2);
  octet_row_and_diag<T>(I, e, dx, f0, f1, f2, d);
  if (upd) gs_node_update_b<T>(nfl, d, f0, f1, f2, nblk0, nl, s0, s1, s2, c0, c1, c2, u);
}
// slot (block*64 + local) of the 27 neighbours of every list node, -1 if absent
__global__ void k_bnd_nslot(int n, const int* __restrict__ list, const int* __restrict__ nbr, int* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 27) return;
  const int i = t / 27, sp0 = t - 27 * i;
  const int node = list[i], blk = node >> 6, l = node & 63;
  const int x = (l >> 4) + sp0 / 9 - 1, y = ((l >> 2) & 3) + (sp0 / 3) % 3 - 1, z = (l & 3) + sp0 % 3 - 1;
  const int sp = ring_spoke(x, y, z);
  const int bb = sp == 13 ? blk : nbr[blk * 27 + sp];
  out[t] = bb >= 0 ? bb * 64 + loc3(x, y, z) : -1;
}

// --------------------------------------------------------- explicit levels
// Stencils of a (block, colour) are stored contiguously (st_index); lane
// (k, g) = (lane>>3, lane&7) of an octet handles spokes g, g+8, g+16 and,
// for g < 3, g+24 of colour node k, so each spoke group reads 64 consecutive
// coefficients.  The spoke s = 13 (the 3x3 self block) falls to g = 5.
template <class T, class LoadU>
__device__ __forceinline__ void st_octet_row(int k, int g, const T* __restrict__ S, const LoadU& U, T& f0, T& f1,
                                             T& f2, T c9[9]) {
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (t == 3 && g >= 3) break;
    const int s = g + 8 * t;
    T x, y, z;
    U(s, x, y, z);
    T m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = t < 3 ? S[(t * 9 + q) * 64 + k * 8 + g] : S[1728 + q * 24 + k * 3 + g];
    a0 = fmaT(m[0], x, a0);
    a0 = fmaT(m[1], y, a0);
    a0 = fmaT(m[2], z, a0);
    a1 = fmaT(m[3], x, a1);
    a1 = fmaT(m[4], y, a1);
    a1 = fmaT(m[5], z, a1);
    a2 = fmaT(m[6], x, a2);
    a2 = fmaT(m[7], y, a2);
    a2 = fmaT(m[8], z, a2);
    if (t == 1) {
#pragma unroll
      for (int q = 0; q < 9; ++q) c9[q] = m[q];
    }
  }
  f0 = octet_sum(a0);
  f1 = octet_sum(a1);
  f2 = octet_sum(a2);
}

// MODE 0: one colour of the GS sweep over all blocks; MODE 1: residual
// r = P(b - A u) over all (block, colour) waves (Stencil_Multiply).
template <class T, int MODE>
__global__ void __launch_bounds__(64 * GS_WAVES) k_st_colour(int nb, int colour0, const int* __restrict__ nbr,
                                                             const uint8_t* __restrict__ flags,
                                                             const T* __restrict__ ST, T* __restrict__ u,
                                                             const T* __restrict__ b, T* __restrict__ r) {
  __shared__ T hu[GS_WAVES][3][128];
  __shared__ int ht[GS_WAVES][28];
  __shared__ int htab[GS_WAVES][128];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wid = xcd_chunk(blockIdx.x, gridDim.x) * GS_WAVES + w;
  const int blk = MODE == 0 ? wid : (wid >> 3);
  const int colour = MODE == 0 ? colour0 : (wid & 7);
  const bool live = blk < nb;
  const int px = (colour >> 2) & 1, py = (colour >> 1) & 1, pz = colour & 1;
  if (lane < 27) ht[w][lane] = !live ? -1 : (lane == 13 ? blk : nbr[blk * 27 + lane]);
  for (int h = lane; h < 125; h += 64) htab[w][h] = cHaloC[colour * 125 + h];
  __syncthreads();
  const int k = lane >> 3, g = lane & 7;
  const int nx = 2 * ((k >> 2) & 1) + 1, ny = 2 * ((k >> 1) & 1) + 1, nz = 2 * (k & 1) + 1;
  const int l = loc3(px - 1 + nx, py - 1 + ny, pz - 1 + nz);
  uint8_t fl = 0;
  T bv[3] = {T(0), T(0), T(0)};
  {
    // the 5^3 halo of colour (px,py,pz) from the table in LDS; both passes'
    // loads, and the updating lane's flags and b, are issued before one wait
    T v[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int h = lane + 64 * i;
      const int t = h < 125 ? htab[w][h] : 0;
      const int bb = h < 125 ? ht[w][t >> 8] : -1;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[i][c] = bb >= 0 ? u[vidx(bb, c, t & 255)] : T(0);
    }
    if (g == 5 && live) {
      fl = flags[blk * 64 + l];
#pragma unroll
      for (int c = 0; c < 3; ++c) bv[c] = b[vidx(blk, c, l)];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
      if (lane + 64 * i < 125)
#pragma unroll
        for (int c = 0; c < 3; ++c) hu[w][c][lane + 64 * i] = v[i][c];
  }
  __syncthreads();
  const T* H = &hu[w][0][0];
  auto U = [&](int s, T& x, T& y, T& z) {
    const int h = ((nx + s / 9 - 1) * 5 + (ny + (s / 3) % 3 - 1)) * 5 + (nz + s % 3 - 1);
    x = H[h];
    y = H[128 + h];
    z = H[256 + h];
  };
  const T* S = ST + ((int64_t)(live ? blk : 0) * 8 + colour) * 1944;
  T f0, f1, f2, c9[9];
  st_octet_row<T>(k, g, S, U, f0, f1, f2, c9);
  if (g == 5 && live) {
    if (MODE == 0) {
      const int hn = (nx * 5 + ny) * 5 + nz;
      T d[6] = {c9[0], c9[4], c9[8], c9[1], c9[2], c9[5]};
      gs_node_update_b<T>(fl, d, f0, f1, f2, blk, l, H[hn], H[128 + hn], H[256 + hn], bv[0], bv[1], bv[2], u);
    } else {
      const T f[3] = {f0, f1, f2};
#pragma unroll
      for (int c = 0; c < 3; ++c) r[vidx(blk, c, l)] = (fl & (2 << c)) ? T(0) : bv[c] - f[c];
    }
  }
}

template <class T>
__device__ __forceinline__ void st_octet_global(bool live, int s, int g, const int* __restrict__ nbr,
                                                const uint8_t* __restrict__ flags, const T* __restrict__ ST,
                                                T* __restrict__ u, const T* __restrict__ b) {
  const int blk = live ? (s >> 6) : 0, l = s & 63;
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const int colour = ((li & 1) << 2) | ((lj & 1) << 1) | (lk & 1);
  const int k = ((li >> 1) << 2) | ((lj >> 1) << 1) | (lk >> 1);
  const int* t = nbr + blk * 27;
  auto U = [&](int sp, T& x, T& y, T& z) {
    const int a = li + sp / 9 - 1, bq = lj + (sp / 3) % 3 - 1, c = lk + sp % 3 - 1;
    const int rs = ring_spoke(a, bq, c);
    const int bb = !live ? -1 : (rs == 13 ? blk : t[rs]);
    const int q = loc3(a, bq, c);
    x = bb >= 0 ? u[vidx(bb, 0, q)] : T(0);
    y = bb >= 0 ? u[vidx(bb, 1, q)] : T(0);
    z = bb >= 0 ? u[vidx(bb, 2, q)] : T(0);
  };
  const T* S = ST + ((int64_t)blk * 8 + colour) * 1944;
  T f0, f1, f2, c9[9];
  st_octet_row<T>(k, g, S, U, f0, f1, f2, c9);
  if (g == 5 && live) {
    T d[6] = {c9[0], c9[4], c9[8], c9[1], c9[2], c9[5]};
    gs_node_update<T>(flags[blk * 64 + l], d, f0, f1, f2, blk, l, u[vidx(blk, 0, l)], u[vidx(blk, 1, l)],
                      u[vidx(blk, 2, l)], b, u);
  }
}

template <class T>
__global__ void __launch_bounds__(BND_THREADS) k_st_boundary(BndLists L, ColourOrder ord, int sweeps,
                                                             const int* __restrict__ nbr,
                                                             const uint8_t* __restrict__ flags,
                                                             const T* __restrict__ ST, T* __restrict__ u,
                                                             const T* __restrict__ b) {
  const int oct = threadIdx.x >> 3, g = threadIdx.x & 7;
  for (int sw = 0; sw < sweeps; ++sw)
    for (int ci = 0; ci < 8; ++ci) {
      const int c = ord.at(ci);
      const int n0 = L.off[c], n1 = L.off[c + 1];
      for (int base = n0; base < n1; base += BND_THREADS / 8) {
        const int i = base + oct;
        const bool live = i < n1;
        st_octet_global<T>(live, live ? L.list[i] : 0, g, nbr, flags, ST, u, b);
      }
      __syncthreads();
    }
}
template <class T>
__global__ void __launch_bounds__(256) k_st_list(const int* __restrict__ list, int n, const int* __restrict__ nbr,
                                                 const uint8_t* __restrict__ flags, const T* __restrict__ ST,
                                                 T* __restrict__ u, const T* __restrict__ b) {
  const int i = blockIdx.x * 32 + (threadIdx.x >> 3);
  const bool live = i < n;
  st_octet_global<T>(live, live ? list[i] : 0, threadIdx.x & 7, nbr, flags, ST, u, b);
}

// ---------------------------------------------- LDS-resident boundary phase
// Boundary nodes are few (the 3 rings around Dirichlet nodes) and only they
// change during a boundary phase, so the phase runs on a compact system:
// b~ = b - (row part of the nodes outside the boundary set) is formed once,
// then every colour step of every sweep updates the boundary u held in LDS.
// Coefficient rows: A[i*288 + g*36 + t*9 + m] is entry m = v*3+w of the 3x3
// block of spoke s = g + 8t of boundary node i (zero for s >= 27), so lane g
// of an octet reads its 36 coefficients as one contiguous run.

// level 0: A_s = dx * sum over the cells holding the node (corner I) and its
// neighbour (corner I + o) of E_cell K0[I][I+o] (the rows of the
// matrix-free operator, SPGrid_Linear_Elasticity.h:22-94)
__global__ void k_bnd_coef0(int n, const int* __restrict__ list, const int* __restrict__ nbr,
                            const double* __restrict__ E, double dx, double* __restrict__ A) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * 32) return;
  const int i = gid >> 5, t = (gid >> 3) & 3, g = gid & 7, s = g + 8 * t;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < 27) {
    const int node = list[i], blk = node >> 6, l = node & 63;
    const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
    const int ox = s / 9 - 1, oy = (s / 3) % 3 - 1, oz = s % 3 - 1;
    for (int I = 0; I < 8; ++I) {
      const int ix = (I >> 2) & 1, iy = (I >> 1) & 1, iz = I & 1;
      const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
      if (jx < 0 || jx > 1 || jy < 0 || jy > 1 || jz < 0 || jz > 1) continue;
      const int cx = li - ix, cy = lj - iy, cz = lk - iz;
      const int rs = ring_spoke(cx, cy, cz);
      const int cb = rs == 13 ? blk : nbr[blk * 27 + rs];
      if (cb < 0) continue;
      const double e = E[cb * 64 + loc3(cx, cy, cz)];
      const int kb = (I * 8 + ((jx << 2) | (jy << 1) | jz)) * 9;
      for (int m = 0; m < 9; ++m) acc[m] += e * cK0d[kb + m];
    }
  }
  for (int m = 0; m < 9; ++m) A[(size_t)i * 288 + g * 36 + t * 9 + m] = dx * acc[m];
}
// explicit levels: the Galerkin stencil rows
__global__ void k_bnd_coef_st(int n, const int* __restrict__ list, const double* __restrict__ ST,
                              double* __restrict__ A) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * 32) return;
  const int i = gid >> 5, t = (gid >> 3) & 3, g = gid & 7, s = g + 8 * t;
  const int node = list[i];
  for (int m = 0; m < 9; ++m)
    A[(size_t)i * 288 + g * 36 + t * 9 + m] = s < 27 ? ST[st_index(node >> 6, node & 63, s * 9 + m)] : 0.0;
}

// b~_i = b_i - sum over spokes to nodes outside the boundary set (bt is
// padded to 4 entries per node).  With `ub` given it also copies the
// boundary iterate to ub ([3][n+1], zero sentinel at n) and clears the grid
// barrier counter `ctr` of the single-launch phase; with `xmap` given it
// writes the dense map's input [u_B (3n), b~ (3n)].
template <class T>
__global__ void __launch_bounds__(256) k_bnd_rhs(int n, const int* __restrict__ list, const int* __restrict__ bslot,
                                                 const T* __restrict__ A, const T* __restrict__ u,
                                                 const T* __restrict__ b, T* __restrict__ bt, T* __restrict__ ub = nullptr,
                                                 unsigned* __restrict__ ctr = nullptr, T* __restrict__ xmap = nullptr) {
  const int i = blockIdx.x * 32 + (threadIdx.x >> 3), g = threadIdx.x & 7;
  if (ub && blockIdx.x == 0 && threadIdx.x < 3) ub[threadIdx.x * (n + 1) + n] = T(0);
  if (ctr && blockIdx.x == 0 && threadIdx.x == 0) *ctr = 0u;
  const bool live = i < n;
  // the node's own b and u go out with the neighbour loads below
  const int node = live ? list[i] : 0, blk = node >> 6, l = node & 63;
  T bn[3] = {T(0), T(0), T(0)}, un[3] = {T(0), T(0), T(0)};
  if (g == 0 && live)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bn[c] = b[vidx(blk, c, l)];
      un[c] = u[vidx(blk, c, l)];
    }
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int sl = live ? bslot[i * 32 + g * 4 + t] : -1;
    if (sl >= 0) {
      const T x = u[vidx(sl >> 6, 0, sl & 63)], y = u[vidx(sl >> 6, 1, sl & 63)], z = u[vidx(sl >> 6, 2, sl & 63)];
      const T* M = A + (size_t)i * 288 + g * 36 + t * 9;
      a0 = fmaT(M[0], x, a0);
      a0 = fmaT(M[1], y, a0);
      a0 = fmaT(M[2], z, a0);
      a1 = fmaT(M[3], x, a1);
      a1 = fmaT(M[4], y, a1);
      a1 = fmaT(M[5], z, a1);
      a2 = fmaT(M[6], x, a2);
      a2 = fmaT(M[7], y, a2);
      a2 = fmaT(M[8], z, a2);
    }
  }
  a0 = octet_sum(a0);
  a1 = octet_sum(a1);
  a2 = octet_sum(a2);
  if (g == 0 && live) {
    bt[i * 4 + 0] = bn[0] - a0;
    bt[i * 4 + 1] = bn[1] - a1;
    bt[i * 4 + 2] = bn[2] - a2;
    bt[i * 4 + 3] = T(0);
    if (ub)
#pragma unroll
      for (int c = 0; c < 3; ++c) ub[c * (n + 1) + i] = un[c];
    if (xmap) {
#pragma unroll
      for (int c = 0; c < 3; ++c) xmap[3 * i + c] = un[c];
      xmap[3 * n + 3 * i + 0] = bn[0] - a0;
      xmap[3 * n + 3 * i + 1] = bn[1] - a1;
      xmap[3 * n + 3 * i + 2] = bn[2] - a2;
    }
  }
}

struct BndOff {
  int off[9];
};
// largest dynamic LDS of one workgroup on gfx950
static constexpr size_t kMaxLds = 163840;

// Per octet and pass: the 36 coefficients of the lane's spokes, their
// neighbour indices and b~ of the node, loaded one pass ahead (two register
// sets used in turn) so global-memory latency overlaps the previous pass.
// Spokes s >= 27 hold zero coefficients and index -1.
template <class T>
struct BndFetch {
  T m[36];
  int j[4];
  T r[4];
  int i;
  bool live;
};
template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double2;  // 16 bytes
};
template <class T>
__device__ __forceinline__ void bnd_fetch(BndFetch<T>& F, int i, bool live, int g, const int* __restrict__ bci,
                                          const T* __restrict__ A, const T* __restrict__ bt) {
  using V = typename Vec4<T>::type;
  constexpr int PER = 16 / sizeof(T);
  F.i = i;
  F.live = live;
  if (!live) {
    // octets past the end of the colour load nothing
#pragma unroll
    for (int q = 0; q < 36; ++q) F.m[q] = T(0);
#pragma unroll
    for (int t = 0; t < 4; ++t) F.j[t] = -1;
#pragma unroll
    for (int q = 0; q < 4; ++q) F.r[q] = T(0);
    return;
  }
  const V* M = reinterpret_cast<const V*>(A + (size_t)i * 288 + g * 36);
#pragma unroll
  for (int q = 0; q < 36 / PER; ++q) {
    const V v = M[q];
    if constexpr (PER == 4) {
      F.m[4 * q] = v.x;
      F.m[4 * q + 1] = v.y;
      F.m[4 * q + 2] = v.z;
      F.m[4 * q + 3] = v.w;
    } else {
      F.m[2 * q] = v.x;
      F.m[2 * q + 1] = v.y;
    }
  }
  const int4 jj = *reinterpret_cast<const int4*>(bci + (size_t)i * 32 + g * 4);
  F.j[0] = jj.x;
  F.j[1] = jj.y;
  F.j[2] = jj.z;
  F.j[3] = jj.w;
  if (g == 5 && bt) {
    const V* R = reinterpret_cast<const V*>(bt + (size_t)i * 4);
#pragma unroll
    for (int q = 0; q < 4 / PER; ++q) {
      const V v = R[q];
      if constexpr (PER == 4) {
        F.r[0] = v.x;
        F.r[1] = v.y;
        F.r[2] = v.z;
        F.r[3] = v.w;
      } else {
        F.r[2 * q] = v.x;
        F.r[2 * q + 1] = v.y;
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) F.r[q] = T(0);
  }
}
// one pass: the row of node F.i over its boundary neighbours (LDS), then the
// update of the node by lane g = 5, which holds spoke 13 (the self block)
template <class T>
__device__ __forceinline__ void bnd_pass(const BndFetch<T>& F, int g, int n, T* ub, const uint8_t* fb) {
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = F.j[t];
    if (j >= 0) {
      const T x = ub[j], y = ub[n + j], z = ub[2 * n + j];
      a0 = fmaT(F.m[t * 9 + 0], x, a0);
      a0 = fmaT(F.m[t * 9 + 1], y, a0);
      a0 = fmaT(F.m[t * 9 + 2], z, a0);
      a1 = fmaT(F.m[t * 9 + 3], x, a1);
      a1 = fmaT(F.m[t * 9 + 4], y, a1);
      a1 = fmaT(F.m[t * 9 + 5], z, a1);
      a2 = fmaT(F.m[t * 9 + 6], x, a2);
      a2 = fmaT(F.m[t * 9 + 7], y, a2);
      a2 = fmaT(F.m[t * 9 + 8], z, a2);
    }
  }
  a0 = octet_sum(a0);
  a1 = octet_sum(a1);
  a2 = octet_sum(a2);
  if (g == 5 && F.live) {
    const int i = F.i;
    const uint8_t fl = fb[i];
    if (fl & 1) {
      T d[6] = {F.m[9], F.m[13], F.m[17], F.m[10], F.m[11], F.m[14]};
      dirichlet_inverse(d, fl);
      const T r0 = F.r[0] - a0, r1 = F.r[1] - a1, r2 = F.r[2] - a2;
      const T u0 = ub[i] + d[0] * r0 + d[3] * r1 + d[4] * r2;
      const T u1 = ub[n + i] + d[3] * r0 + d[1] * r1 + d[5] * r2;
      const T u2 = ub[2 * n + i] + d[4] * r0 + d[5] * r1 + d[2] * r2;
      ub[i] = (fl & 2) ? T(0) : u0;
      ub[n + i] = (fl & 4) ? T(0) : u1;
      ub[2 * n + i] = (fl & 8) ? T(0) : u2;
    } else {
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (fl & (2 << cc)) ub[cc * n + i] = T(0);
    }
  }
}
// position in the pass sequence (sweep, colour step, pass within the colour)
struct BndCursor {
  int sw, ci, p, c;
};
template <int NO>
__device__ __forceinline__ BndCursor bnd_advance(const BndCursor& a, const BndOff& L, const ColourOrder& ord,
                                                 bool& step_end) {
  BndCursor b = a;
  b.p = a.p + 1;
  step_end = L.off[a.c] + b.p * NO >= L.off[a.c + 1];
  if (step_end) {
    b.p = 0;
    if (++b.ci == 8) {
      b.ci = 0;
      ++b.sw;
    }
    b.c = ord.at(b.ci);
  }
  return b;
}

template <class T, int NT>
__global__ void __launch_bounds__(NT) k_bnd_sweeps(int n, BndOff L, ColourOrder ord, int sweeps,
                                                   const int* __restrict__ list, const int* __restrict__ bci,
                                                   const T* __restrict__ A, const uint8_t* __restrict__ flags,
                                                   const T* __restrict__ bt, T* __restrict__ u) {
  constexpr int NO = NT / 8;  // octets = nodes per pass
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* ub = reinterpret_cast<T*>(smem_raw);                 // [3][n]
  uint8_t* fb = reinterpret_cast<uint8_t*>(ub + 3 * n);  // [n]
  for (int i = threadIdx.x; i < n; i += NT) {
    const int node = list[i], blk = node >> 6, l = node & 63;
#pragma unroll
    for (int c = 0; c < 3; ++c) ub[c * n + i] = u[vidx(blk, c, l)];
    fb[i] = flags[node];
  }
  const int oct = threadIdx.x >> 3, g = threadIdx.x & 7;
  auto fetch = [&](BndFetch<T>& F, const BndCursor& k) {
    const int i = L.off[k.c] + k.p * NO + oct;
    bnd_fetch<T>(F, i, i < L.off[k.c + 1], g, bci, A, bt);
  };
  BndCursor cur{0, 0, 0, ord.at(0)};
  BndFetch<T> X, Y;
  fetch(X, cur);
  __syncthreads();
  while (true) {
    bool end;
    BndCursor nx = bnd_advance<NO>(cur, L, ord, end);
    bool more = nx.sw < sweeps;
    fetch(Y, more ? nx : cur);
    bnd_pass<T>(X, g, n, ub, fb);
    if (end) __syncthreads();
    if (!more) break;
    cur = nx;
    nx = bnd_advance<NO>(cur, L, ord, end);
    more = nx.sw < sweeps;
    fetch(X, more ? nx : cur);
    bnd_pass<T>(Y, g, n, ub, fb);
    if (end) __syncthreads();
    if (!more) break;
    cur = nx;
  }
  for (int i = threadIdx.x; i < n; i += NT) {
    const int node = list[i], blk = node >> 6, l = node & 63;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) u[vidx(blk, cc, l)] = ub[cc * n + i];
  }
}

// Dense map [G H] of a boundary phase, u_B <- G u_B + H b~ (row-major,
// 3n x 6n with rows padded to bmap_stride(n), entry index node*3 + component).  Column-parallel: the phase is
// run on all unit vectors at once, lane = column (columns j < 3n start from
// u_B = e_j with b~ = 0, columns 3n + j from u_B = 0 with b~ = e_j), so
// every node row and its coefficients are wave-uniform.  The 16 waves of a
// workgroup share 64 columns and split the nodes of each colour; the
// matrix itself holds the iterate (nodes of one colour do not couple).
constexpr int BM_WAVES = 16;
__global__ void __launch_bounds__(64 * BM_WAVES) k_bmap_build(int n, BndOff L, ColourOrder ord, int sweeps,
                                                            const int* __restrict__ list,
                                                            const int* __restrict__ bci,
                                                            const double* __restrict__ A,
                                                            const uint8_t* __restrict__ flags,
                                                            double* __restrict__ M) {
  const int64_t K = bmap_stride(n);  // row stride; columns >= 6n stay zero
  const int col = blockIdx.x * 64 + (threadIdx.x & 63);
  const bool live = col < 6 * n;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int rows = 3 * n;
  double* Mc = M + col;
  if (live)
    for (int r = w; r < rows; r += BM_WAVES) Mc[(int64_t)r * K] = (col == r) ? 1.0 : 0.0;
  const int hj = col - rows;  // right-hand-side unit index of an H column
  for (int sw = 0; sw < sweeps; ++sw)
    for (int k = 0; k < 8; ++k) {
      __syncthreads();
      const int c = ord.at(k);
      for (int i = L.off[c] + w; i < L.off[c + 1]; i += BM_WAVES) {
        const uint8_t fl = flags[list[i]];
        double* Ui = Mc + (int64_t)(3 * i) * K;
        if (!live) continue;
        if (fl & 1) {
          const double* Ai = A + (size_t)i * 288;
          const int* Ji = bci + (size_t)i * 32;
          double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
          for (int g = 0; g < 8; ++g)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const int j = Ji[g * 4 + t];
              if (j < 0) continue;
              const double* m = Ai + g * 36 + t * 9;
              const double* Uj = Mc + (int64_t)(3 * j) * K;
              const double x = Uj[0], y = Uj[K], z = Uj[2 * K];
              a0 = fmaT(m[0], x, fmaT(m[1], y, fmaT(m[2], z, a0)));
              a1 = fmaT(m[3], x, fmaT(m[4], y, fmaT(m[5], z, a1)));
              a2 = fmaT(m[6], x, fmaT(m[7], y, fmaT(m[8], z, a2)));
            }
          // self block: spoke 13 = (g 5, t 1)
          const double* m = Ai + 5 * 36 + 9;
          double d[6] = {m[0], m[4], m[8], m[1], m[2], m[5]};
          dirichlet_inverse(d, fl);
          const double r0 = (hj == 3 * i ? 1.0 : 0.0) - a0;
          const double r1 = (hj == 3 * i + 1 ? 1.0 : 0.0) - a1;
          const double r2 = (hj == 3 * i + 2 ? 1.0 : 0.0) - a2;
          const double u0 = Ui[0] + d[0] * r0 + d[3] * r1 + d[4] * r2;
          const double u1 = Ui[K] + d[3] * r0 + d[1] * r1 + d[5] * r2;
          const double u2 = Ui[2 * K] + d[4] * r0 + d[5] * r1 + d[2] * r2;
          Ui[0] = (fl & 2) ? 0.0 : u0;
          Ui[K] = (fl & 4) ? 0.0 : u1;
          Ui[2 * K] = (fl & 8) ? 0.0 : u2;
        } else {
          if (fl & 2) Ui[0] = 0.0;
          if (fl & 4) Ui[K] = 0.0;
          if (fl & 8) Ui[2 * K] = 0.0;
        }
      }
    }
}

// Dense boundary map apply: x = [u_B; b~] (node-major), y = [G H] x.  Rows
// are padded to a multiple of 64 entries (zeros), one wavefront per row
// streams it with 16-byte loads, four in flight per lane, and reduces in a
// fixed order.
template <class T>
__device__ __forceinline__ T dot4(const typename Vec4<T>::type& a, const typename Vec4<T>::type& b, T acc);
template <>
__device__ __forceinline__ float dot4<float>(const float4& a, const float4& b, float acc) {
  return fmaf(a.w, b.w, fmaf(a.z, b.z, fmaf(a.y, b.y, fmaf(a.x, b.x, acc))));
}
template <>
__device__ __forceinline__ double dot4<double>(const double2& a, const double2& b, double acc) {
  return fma(a.y, b.y, fma(a.x, b.x, acc));
}
// y = M x, one wavefront per row; with `list` given, row 3i+c goes straight
// to u at boundary node list[i], component c
template <class T>
__global__ void __launch_bounds__(256) k_bmap_gemv(int rows, int Kp, const T* __restrict__ M, const T* __restrict__ x,
                                                   T* __restrict__ y, const int* __restrict__ list = nullptr,
                                                   T* __restrict__ u = nullptr) {
  using V = typename Vec4<T>::type;
  constexpr int PER = 16 / sizeof(T);
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= rows) return;
  const V* row = reinterpret_cast<const V*>(M + (int64_t)r * Kp);
  const V* xv = reinterpret_cast<const V*>(x);
  const int nv = Kp / PER;
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int k = lane;
  for (; k + 192 < nv; k += 256) {
    const V m0 = row[k], m1 = row[k + 64], m2 = row[k + 128], m3 = row[k + 192];
    a0 = dot4<T>(m0, xv[k], a0);
    a1 = dot4<T>(m1, xv[k + 64], a1);
    a2 = dot4<T>(m2, xv[k + 128], a2);
    a3 = dot4<T>(m3, xv[k + 192], a3);
  }
  for (; k < nv; k += 64) a0 = dot4<T>(row[k], xv[k], a0);
  T acc = (a0 + a1) + (a2 + a3);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) {
    if (list) {
      const int i = r / 3, c = r - 3 * i, node = list[i];
      u[vidx(node >> 6, c, node & 63)] = acc;
    } else {
      y[r] = acc;
    }
  }
}
// ------------------------------------- level-0 matrix-free boundary phase
// Level 0 needs no coefficient rows: boundary node i carries the moduli of
// its 8 cells (bE[I*n + i], cell = node - I), the boundary indices of each
// cell's 8 nodes (cnode, int16; -1 outside the set) and the inverse of its
// 3x3 self block (bD, 6 entries + 2 pad).  Wave w takes cell corner I = w&7
// of 64 nodes, so K0[I][J] is wave-uniform; the 8 corner rows meet in LDS
// and the waves with I = 0 update the nodes.
__global__ void k_bnd0_setup(int n, const int* __restrict__ list, const uint8_t* __restrict__ flags,
                             const int* __restrict__ cslot, const double* __restrict__ E, double dx,
                             double* __restrict__ bE, double* __restrict__ bD) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d[6] = {0, 0, 0, 0, 0, 0};
  for (int I = 0; I < 8; ++I) {
    const int sl = cslot[I * n + i];
    const double e = sl >= 0 ? E[sl] : 0.0;
    bE[I * n + i] = e;
    const int kd = (I * 8 + I) * 9;
    d[0] += e * cK0d[kd + 0];
    d[1] += e * cK0d[kd + 4];
    d[2] += e * cK0d[kd + 8];
    d[3] += e * cK0d[kd + 1];
    d[4] += e * cK0d[kd + 2];
    d[5] += e * cK0d[kd + 5];
  }
  for (int q = 0; q < 6; ++q) d[q] *= dx;
  dirichlet_inverse(d, flags[list[i]]);
  for (int q = 0; q < 6; ++q) bD[i * 8 + q] = d[q];
  bD[i * 8 + 6] = 0.0;
  bD[i * 8 + 7] = 0.0;
}

template <class T>
struct Bnd0Fetch {
  T e;
  int c[4];  // 8 int16 node indices, two per int
  T r[3];
  T di[6];
  int i;
  bool live;
};

static constexpr int B0_THREADS = 1024;
static constexpr int B0_NODES = 128;  // nodes per pass (2 groups of 64 x 8 corners)
template <class T>
__global__ void __launch_bounds__(B0_THREADS) k_bnd0_sweeps(int n, BndOff L, ColourOrder ord, int sweeps,
                                                            const int* __restrict__ list,
                                                            const uint8_t* __restrict__ flags,
                                                            const T* __restrict__ bE,
                                                            const int16_t* __restrict__ cnode,
                                                            const T* __restrict__ bD, const T* __restrict__ bt,
                                                            T dx, T* __restrict__ u) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n1 = n + 1;                                 // entry n of each component is a zero sentinel
  T* ub = reinterpret_cast<T*>(smem_raw);               // [3][n+1]
  T* pf = ub + 3 * n1;                                  // [2][8][3][128]
  uint8_t* fb = reinterpret_cast<uint8_t*>(pf + 2 * 8 * 3 * B0_NODES);  // [n]
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int I = __builtin_amdgcn_readfirstlane(w & 7);
  const int j = (w >> 3) * 64 + lane;  // node within the pass
  for (int i = tid; i < n; i += B0_THREADS) {
    const int node = list[i], blk = node >> 6, l = node & 63;
#pragma unroll
    for (int c = 0; c < 3; ++c) ub[c * n1 + i] = u[vidx(blk, c, l)];
    fb[i] = flags[node];
  }
  if (tid < 3) ub[tid * n1 + n] = T(0);
  auto fetch = [&](Bnd0Fetch<T>& F, const BndCursor& k) {
    const int i = L.off[k.c] + k.p * B0_NODES + j;
    F.i = i;
    F.live = i < L.off[k.c + 1];
    if (F.live) {
      F.e = bE[I * n + i];
      const int4 a = *reinterpret_cast<const int4*>(cnode + ((size_t)I * n + i) * 8);
      F.c[0] = a.x;
      F.c[1] = a.y;
      F.c[2] = a.z;
      F.c[3] = a.w;
    } else {
      F.e = T(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) F.c[q] = -1;
    }
    if (F.live && I == 0) {
#pragma unroll
      for (int q = 0; q < 3; ++q) F.r[q] = bt[i * 4 + q];
#pragma unroll
      for (int q = 0; q < 6; ++q) F.di[q] = bD[i * 8 + q];
    } else {
#pragma unroll
      for (int q = 0; q < 3; ++q) F.r[q] = T(0);
#pragma unroll
      for (int q = 0; q < 6; ++q) F.di[q] = T(0);
    }
  };
  auto pass = [&](const Bnd0Fetch<T>& F, int buf, bool end) {
    // row of corner I over the node's boundary neighbours (outside ones read the sentinel)
    T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
    for (int J = 0; J < 8; ++J) {
      const int wv = F.c[J >> 1];
      int q = (J & 1) ? (wv >> 16) : (int)(short)(wv & 0xffff);
      q = q < 0 ? n : q;
      const T v0 = ub[q], v1 = ub[n1 + q], v2 = ub[2 * n1 + q];
      const int kb = (I * 8 + J) * 9;
      a0 = fmaT(KT<T>::k(kb + 0), v0, a0);
      a0 = fmaT(KT<T>::k(kb + 1), v1, a0);
      a0 = fmaT(KT<T>::k(kb + 2), v2, a0);
      a1 = fmaT(KT<T>::k(kb + 3), v0, a1);
      a1 = fmaT(KT<T>::k(kb + 4), v1, a1);
      a1 = fmaT(KT<T>::k(kb + 5), v2, a1);
      a2 = fmaT(KT<T>::k(kb + 6), v0, a2);
      a2 = fmaT(KT<T>::k(kb + 7), v1, a2);
      a2 = fmaT(KT<T>::k(kb + 8), v2, a2);
    }
    T* P = pf + buf * (8 * 3 * B0_NODES);
    P[(I * 3 + 0) * B0_NODES + j] = F.e * a0;
    P[(I * 3 + 1) * B0_NODES + j] = F.e * a1;
    P[(I * 3 + 2) * B0_NODES + j] = F.e * a2;
    __syncthreads();
    if (I == 0 && F.live) {
      T f[3];
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        T acc = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) acc += P[(c * 3 + v) * B0_NODES + j];
        f[v] = dx * acc;
      }
      const int i = F.i;
      const uint8_t fl = fb[i];
      if (fl & 1) {
        const T r0 = F.r[0] - f[0], r1 = F.r[1] - f[1], r2 = F.r[2] - f[2];
        const T u0 = ub[i] + F.di[0] * r0 + F.di[3] * r1 + F.di[4] * r2;
        const T u1 = ub[n1 + i] + F.di[3] * r0 + F.di[1] * r1 + F.di[5] * r2;
        const T u2 = ub[2 * n1 + i] + F.di[4] * r0 + F.di[5] * r1 + F.di[2] * r2;
        ub[i] = (fl & 2) ? T(0) : u0;
        ub[n1 + i] = (fl & 4) ? T(0) : u1;
        ub[2 * n1 + i] = (fl & 8) ? T(0) : u2;
      } else {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
          if (fl & (2 << cc)) ub[cc * n1 + i] = T(0);
      }
    }
    if (end) __syncthreads();
  };
  BndCursor cur{0, 0, 0, ord.at(0)};
  Bnd0Fetch<T> X, Y;
  fetch(X, cur);
  __syncthreads();
  int buf = 0;
  while (true) {
    bool end;
    BndCursor nx = bnd_advance<B0_NODES>(cur, L, ord, end);
    bool more = nx.sw < sweeps;
    fetch(Y, more ? nx : cur);
    pass(X, buf, end);
    buf ^= 1;
    if (!more) break;
    cur = nx;
    nx = bnd_advance<B0_NODES>(cur, L, ord, end);
    more = nx.sw < sweeps;
    fetch(X, more ? nx : cur);
    pass(Y, buf, end);
    buf ^= 1;
    if (!more) break;
    cur = nx;
  }
  for (int i = tid; i < n; i += B0_THREADS) {
    const int node = list[i], blk = node >> 6, l = node & 63;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) u[vidx(blk, cc, l)] = ub[cc * n1 + i];
  }
}

// Level-0 boundary phase over several CUs: one launch of G <= 32
// workgroups (all resident at once), each of 8 waves; wave I computes cell
// corner I's row for 64 nodes of a colour, the corner rows meet in LDS and
// wave 0 updates the nodes.  The boundary iterate lives in `ub` ([3][n+1],
// entry n a zero sentinel, filled by k_bnd_rhs); colour steps are separated
// by a grid barrier.  Each node's update in the last sweep also goes to u.
// A barrier wait longer than `limit` wall-clock ticks sets coop[1] and ends
// the kernel, so every wave exits whatever happens.
static constexpr int BC_THREADS = 512;
static constexpr int BC_MAX_WG = 32;
// per-lane tables of one node group: the corner's modulus and node indices,
// and (wave 0) b~, the inverse self block, the node and its flags
template <class T>
struct CoopFetch {
  T e, r[3], di[6];
  int cw[4];
  int node, i;
  uint8_t fl;
  bool live;
};
template <class T>
__global__ void __launch_bounds__(BC_THREADS) k_bnd0_coop(int n, BndOff L, ColourOrder ord, int sweeps,
                                                          const int* __restrict__ list,
                                                          const uint8_t* __restrict__ flags,
                                                          const T* __restrict__ bE,
                                                          const int16_t* __restrict__ cnode,
                                                          const T* __restrict__ bD, const T* __restrict__ bt, T dx,
                                                          T* ub, T* __restrict__ u, unsigned* coop,
                                                          unsigned long long limit) {
  __shared__ T P[8][3][64];
  __shared__ int bad;
  const int n1 = n + 1;
  const int tid = threadIdx.x, lane = tid & 63;
  const int I = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int G = gridDim.x;
  const int steps = sweeps * 8;
  unsigned* ctr = coop;
  int* err = reinterpret_cast<int*>(coop + 1);
  // everything but the iterate: loaded ahead, the next step's first group
  // while the grid barrier is pending
  auto fetch = [&](CoopFetch<T>& F, int st, int q) {
    const int c = ord.at(st & 7), o0 = L.off[c], nc = L.off[c + 1] - o0;
    const int j = q * 64 + lane;
    F.live = j < nc;
    F.i = o0 + (F.live ? j : 0);
    F.e = T(0);
    F.cw[0] = F.cw[1] = F.cw[2] = F.cw[3] = -1;
    if (F.live) {
      F.e = bE[I * n + F.i];
      const int4 a = *reinterpret_cast<const int4*>(cnode + ((size_t)I * n + F.i) * 8);
      F.cw[0] = a.x;
      F.cw[1] = a.y;
      F.cw[2] = a.z;
      F.cw[3] = a.w;
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) F.r[v] = T(0);
#pragma unroll
    for (int v = 0; v < 6; ++v) F.di[v] = T(0);
    F.node = 0;
    F.fl = 0;
    if (I == 0 && F.live) {
#pragma unroll
      for (int v = 0; v < 3; ++v) F.r[v] = bt[F.i * 4 + v];
#pragma unroll
      for (int v = 0; v < 6; ++v) F.di[v] = bD[F.i * 8 + v];
      F.node = list[F.i];
      F.fl = flags[F.node];
    }
  };
  CoopFetch<T> F;
  fetch(F, 0, blockIdx.x);
  for (int st = 0; st < steps; ++st) {
    const int c = ord.at(st & 7);
    const bool last = st >= steps - 8;
    const int nc = L.off[c + 1] - L.off[c];
    for (int q = blockIdx.x; q * 64 < nc; q += G) {
      if (q != (int)blockIdx.x) fetch(F, st, q);
      T a0 = 0, a1 = 0, a2 = 0, x[3];
#pragma unroll
      for (int J = 0; J < 8; ++J) {
        const int wv = F.cw[J >> 1];
        int k = (J & 1) ? (wv >> 16) : (int)(short)(wv & 0xffff);
        k = k < 0 ? n : k;
        const T v0 = ub[k], v1 = ub[n1 + k], v2 = ub[2 * n1 + k];
        if (J == 0) {  // for corner 0, node 0 of the cell is the node itself
          x[0] = v0;
          x[1] = v1;
          x[2] = v2;
        }
        const int kb = (I * 8 + J) * 9;
        a0 = fmaT(KT<T>::k(kb + 0), v0, a0);
        a0 = fmaT(KT<T>::k(kb + 1), v1, a0);
        a0 = fmaT(KT<T>::k(kb + 2), v2, a0);
        a1 = fmaT(KT<T>::k(kb + 3), v0, a1);
        a1 = fmaT(KT<T>::k(kb + 4), v1, a1);
        a1 = fmaT(KT<T>::k(kb + 5), v2, a1);
        a2 = fmaT(KT<T>::k(kb + 6), v0, a2);
        a2 = fmaT(KT<T>::k(kb + 7), v1, a2);
        a2 = fmaT(KT<T>::k(kb + 8), v2, a2);
      }
      P[I][0][lane] = F.e * a0;
      P[I][1][lane] = F.e * a1;
      P[I][2][lane] = F.e * a2;
      __syncthreads();
      if (I == 0 && F.live) {
        const int i = F.i;
        T f[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          T acc = 0;
#pragma unroll
          for (int cc = 0; cc < 8; ++cc) acc += P[cc][v][lane];
          f[v] = dx * acc;
        }
        const uint8_t fl = F.fl;
        if (fl & 1) {
          const T r0 = F.r[0] - f[0], r1 = F.r[1] - f[1], r2 = F.r[2] - f[2];
          const T u0 = x[0] + F.di[0] * r0 + F.di[3] * r1 + F.di[4] * r2;
          const T u1 = x[1] + F.di[3] * r0 + F.di[1] * r1 + F.di[5] * r2;
          const T u2 = x[2] + F.di[4] * r0 + F.di[5] * r1 + F.di[2] * r2;
          x[0] = (fl & 2) ? T(0) : u0;
          x[1] = (fl & 4) ? T(0) : u1;
          x[2] = (fl & 8) ? T(0) : u2;
        } else {
#pragma unroll
          for (int v = 0; v < 3; ++v)
            if (fl & (2 << v)) x[v] = T(0);
        }
#pragma unroll
        for (int v = 0; v < 3; ++v) ub[v * n1 + i] = x[v];
        if (last) {
          const int blk = F.node >> 6, l = F.node & 63;
#pragma unroll
          for (int v = 0; v < 3; ++v) u[vidx(blk, v, l)] = x[v];
        }
      }
      __syncthreads();  // P is reused by the next group
    }
    if (st + 1 == steps) break;
    // Grid barrier.  Thread 0 (of wave 0, the only wave that stores to
    // global memory) publishes the step with a release add, the workgroup
    // then loads the next step's tables, and thread 0 waits for all G
    // arrivals; a wait longer than `limit` wall-clock ticks sets the error
    // word and ends the kernel, so every wave exits.
    const unsigned target = (unsigned)(G * (st + 1));
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    fetch(F, st + 1, blockIdx.x);
    if (tid == 0) {
      int b = 0;
      const unsigned long long t0 = wall_clock64();
      // relaxed polls (an acquire load would invalidate the L2 at every
      // poll), then one acquire fence for the data the other workgroups
      // published with their release add
      while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (wall_clock64() - t0 > limit) {
          __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          b = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      bad = b;
    }
    __syncthreads();
    if (bad) return;
  }
}

// workgroup size of the boundary phase: two register sets of BndFetch per lane
template <class T>
struct BndThreads {
  static constexpr int value = sizeof(T) == 4 ? 1024 : 512;
};

// ----------------------------------------------------------------- transfers
// Transfers, LDS-staged: a wavefront owns one block of the output level.
// Restriction R = P^T is full weighting {1/2,1,1/2}^3 around the fine node
// 2c-B (Restriction.h:52-92); the wave stages the 9^3 fine nodes
// x = 2l + d (d = -1..1) of its coarse block (fine block (x>>2)+1 of the
// 27-entry source table).  Prolongation adds the trilinear interpolant of
// the 3^3 coarse nodes around the fine block, then applies the Dirichlet and
// null-space projections (Prolongation.h:57-99, Multigrid_Hierarchy.h:600-603).
static constexpr int TR_WAVES = 2;
template <class T>
// With `czero` given, the kernel also clears that vector on the coarse
// level (the coarse iterate of the V-cycle), saving a launch.
__global__ void __launch_bounds__(64 * TR_WAVES) k_restrict_lds(int nbc, const int* __restrict__ rsrc,
                                                                const uint8_t* __restrict__ cflags,
                                                                const T* __restrict__ fine, T* __restrict__ coarse,
                                                                T* __restrict__ czero) {
  __shared__ T hf[TR_WAVES][3][729];
  __shared__ int ht[TR_WAVES][27];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = xcd_chunk(blockIdx.x, gridDim.x) * TR_WAVES + w;
  const bool live = blk < nbc;
  if (lane < 27) ht[w][lane] = live ? rsrc[blk * 27 + lane] : -1;
  const uint8_t fl = live ? cflags[blk * 64 + lane] : 0;  // with the source table, not after the sums
  __syncthreads();
  {
    // the 9^3 fine halo in 12 passes; all loads are issued before the LDS writes
    constexpr int NP = (729 + 63) / 64;
    T v[NP][3];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int h = lane + 64 * i;
      const int hh = h < 729 ? h : 0;
      const int x = hh / 81 - 1, y = (hh / 9) % 9 - 1, z = hh % 9 - 1;
      const int fb = h < 729 ? ht[w][((x >> 2) + 1) * 9 + ((y >> 2) + 1) * 3 + ((z >> 2) + 1)] : -1;
      const int q = loc3(x, y, z);
#pragma unroll
      for (int c = 0; c < 3; ++c) v[i][c] = fb >= 0 ? fine[vidx(fb, c, q)] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NP; ++i)
      if (lane + 64 * i < 729)
#pragma unroll
        for (int c = 0; c < 3; ++c) hf[w][c][lane + 64 * i] = v[i][c];
  }
  __syncthreads();
  if (!live) return;
  const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int dk = -1; dk <= 1; ++dk) {
        const T wgt = T((di == 0 ? 1.0 : 0.5) * (dj == 0 ? 1.0 : 0.5) * (dk == 0 ? 1.0 : 0.5));
        const int h = (2 * li + di + 1) * 81 + (2 * lj + dj + 1) * 9 + (2 * lk + dk + 1);
        a0 += wgt * hf[w][0][h];
        a1 += wgt * hf[w][1][h];
        a2 += wgt * hf[w][2][h];
      }
  coarse[vidx(blk, 0, lane)] = (fl & 2) ? T(0) : a0;
  coarse[vidx(blk, 1, lane)] = (fl & 4) ? T(0) : a1;
  coarse[vidx(blk, 2, lane)] = (fl & 8) ? T(0) : a2;
  if (czero)
#pragma unroll
    for (int c = 0; c < 3; ++c) czero[vidx(blk, c, lane)] = T(0);
}

// The 27 coarse nodes around fine block b: the 3^3 coarse nodes from the
// fine block's phase (((base + B) >> 1) & 3 per axis) in the up to 8 coarse
// blocks of its psrc row, as (coarse block << 6) | local index, -1 absent.
// Built once per topology, so the prolongation's coarse loads depend on one
// table load instead of the base, then psrc, then the data.
__device__ __forceinline__ int prolong_source(int b, int lane, const int* __restrict__ psrc,
                                              const int* __restrict__ fbase) {
  const int phx = ((fbase[3 * b] + 4) >> 1) & 3, phy = ((fbase[3 * b + 1] + 4) >> 1) & 3,
            phz = ((fbase[3 * b + 2] + 8) >> 1) & 3;
  const int x = phx + lane / 9, y = phy + (lane / 3) % 3, z = phz + lane % 3;
  const int cb = psrc[b * 8 + ((x >> 2) << 2) + ((y >> 2) << 1) + (z >> 2)];
  return cb >= 0 ? (cb << 6) | loc3(x, y, z) : -1;
}
__global__ void k_prolong_table(int nbf, const int* __restrict__ psrc, const int* __restrict__ fbase,
                                int* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nbf * 27) return;
  out[i] = prolong_source((int)(i / 27), (int)(i % 27), psrc, fbase);
}

template <class T>
__global__ void __launch_bounds__(64 * TR_WAVES) k_prolong_lds(int nbf, const int* __restrict__ psrc,
                                                               const int* __restrict__ fbase,
                                                               const int* __restrict__ psrc27,
                                                               const uint8_t* __restrict__ fflags,
                                                               const T* __restrict__ coarse, T* __restrict__ fine) {
  __shared__ T hc[TR_WAVES][3][27];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = xcd_chunk(blockIdx.x, gridDim.x) * TR_WAVES + w;
  const bool live = blk < nbf;
  const int b0 = live ? blk : 0;
  // the fine block's own values do not depend on the staged coarse halo:
  // their loads go out first
  const uint8_t fl = fflags[b0 * 64 + lane];
  T a0 = fine[vidx(b0, 0, lane)], a1 = fine[vidx(b0, 1, lane)], a2 = fine[vidx(b0, 2, lane)];
  if (lane < 27) {
    const int t = !live ? -1 : (psrc27 ? psrc27[blk * 27 + lane] : prolong_source(blk, lane, psrc, fbase));
#pragma unroll
    for (int c = 0; c < 3; ++c) hc[w][c][lane] = t >= 0 ? coarse[vidx(t >> 6, c, t & 63)] : T(0);
  }
  __syncthreads();
  if (!live) return;
  const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
#pragma unroll
  for (int oi = 0; oi < 2; ++oi)
#pragma unroll
    for (int oj = 0; oj < 2; ++oj)
#pragma unroll
      for (int ok = 0; ok < 2; ++ok) {
        const T wi = (li & 1) ? T(0.5) : (oi ? T(0) : T(1));
        const T wj = (lj & 1) ? T(0.5) : (oj ? T(0) : T(1));
        const T wk = (lk & 1) ? T(0.5) : (ok ? T(0) : T(1));
        const T wgt = wi * wj * wk;
        const int h = ((li >> 1) + oi) * 9 + ((lj >> 1) + oj) * 3 + ((lk >> 1) + ok);
        a0 += wgt * hc[w][0][h];
        a1 += wgt * hc[w][1][h];
        a2 += wgt * hc[w][2][h];
      }
  const bool act = fl & 1;
  fine[vidx(blk, 0, lane)] = (!act || (fl & 2)) ? T(0) : a0;
  fine[vidx(blk, 1, lane)] = (!act || (fl & 4)) ? T(0) : a1;
  fine[vidx(blk, 2, lane)] = (!act || (fl & 8)) ? T(0) : a2;
}

// ------------------------------------------------------------------- vectors
template <class T>
__global__ void k_zero(T* v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = T(0);
}
template <class T>
__global__ void k_copy(const T* a, T* b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void k_d2f(const double* a, float* b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_f2d(const float* a, double* b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
template <class T>
__global__ void k_axpy_d(double* u, const T* d, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] += (double)d[i];
}

static constexpr int RED_T = 256;
static constexpr int RED_MAX_WG = 2048;

int64_t reduce_partials(int64_t n) {
  (void)n;
  return RED_MAX_WG;
}
static inline unsigned red_grid(int64_t n) {
  int64_t g = (n + RED_T * 4 - 1) / (RED_T * 4);
  if (g < 1) g = 1;
  if (g > RED_MAX_WG) g = RED_MAX_WG;
  return (unsigned)g;
}

template <int OP>  // 0 sum, 1 max
__device__ inline double block_reduce(double v) {
  __shared__ double sh[RED_T / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = OP == 0 ? v + w : fmax(v, w);
  }
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int i = 1; i < RED_T / 64; ++i) r = OP == 0 ? r + sh[i] : fmax(r, sh[i]);
    sh[0] = r;
  }
  __syncthreads();
  const double r = sh[0];
  __syncthreads();
  return r;
}

template <class T>
__global__ void __launch_bounds__(RED_T) k_dot_part(const T* a, const T* b, int64_t n, double* part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[i] * (double)b[i];
  s = block_reduce<0>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
template <class T>
__global__ void __launch_bounds__(RED_T) k_linf_part(const T* a, int64_t n, double* part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s = fmax(s, fabs((double)a[i]));
  s = block_reduce<1>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void __launch_bounds__(RED_T) k_sumsq_part(const double* a, int64_t n, double* part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * a[i];
  s = block_reduce<0>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// PCG scalar updates, run by the thread that has just written the
// reduction they depend on (stage 0 after rho, 1 after s.q, 2 after ||r||)
__device__ inline void cg_stage(CgScalars* sc, int stage) {
  if (stage == 0) {
    sc->beta = sc->rho / sc->rho_old;
  } else if (stage == 1) {
    sc->singular = !(sc->sq > 0.0);
    sc->alpha = sc->sq != 0.0 ? sc->rho / sc->sq : (double)FLT_MAX;
  } else if (stage == 2) {
    sc->rho_old = sc->rho;
  } else if (stage == 3) {
    sc->rho_old = (double)FLT_MAX;
    sc->singular = 0;
  }
}
template <int OP>
__global__ void __launch_bounds__(RED_T) k_final(const double* part, int np, double* out, CgScalars* sc = nullptr,
                                                 int stage = -1) {
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s = OP == 0 ? s + part[i] : fmax(s, part[i]);
  s = block_reduce<OP>(s);
  if (threadIdx.x == 0) {
    *out = s;
    if (sc) cg_stage(sc, stage);
  }
}

__global__ void __launch_bounds__(RED_T) k_sum_part(const double* a, int64_t n, double* part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += a[i];
  s = block_reduce<0>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// sum of np partials in a fixed order by one 1024-thread workgroup
__global__ void __launch_bounds__(1024) k_sum_wide(const double* __restrict__ part, int np, double* out,
                                                   CgScalars* sc = nullptr, int stage = -1) {
  __shared__ double sh[16];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * 1024 < np; i += 4 * 1024) {
    a0 += part[i];
    a1 += part[i + 1024];
    a2 += part[i + 2048];
    a3 += part[i + 3072];
  }
  for (; i < np; i += 1024) a0 += part[i];
  double v = (a0 + a1) + (a2 + a3);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int q = 1; q < 16; ++q) t += sh[q];
    *out = t;
    if (sc) cg_stage(sc, stage);
  }
}

// ----------------------------------------------------------------------- PCG
template <class T>
__global__ void k_cg_direction(T* s, const T* z, const CgScalars* sc, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T beta = (T)sc->beta;  // CG_VECTOR::Copy(c,x,y) rounds c to T (CG_VECTOR.cpp:96-114)
  s[i] = beta * s[i] + z[i];
}
template <class T>
// x += alpha s, r -= alpha q, partial max |r|; with `zc` also clears zc
// (the next V-cycle's iterate)
__global__ void __launch_bounds__(RED_T) k_cg_update(T* x, T* r, const T* s, const T* q, const CgScalars* sc, int64_t n,
                                                      double* part, T* zc) {
  const T alpha = (T)sc->alpha;
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = alpha * s[i] + x[i];
    const T rv = -alpha * q[i] + r[i];
    r[i] = rv;
    m = fmax(m, fabs((double)rv));
    if (zc) zc[i] = T(0);
  }
  m = block_reduce<1>(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}
__global__ void k_cg_scalars(CgScalars* sc, int stage) {
  if (threadIdx.x == 0) cg_stage(sc, stage);
}

// v = P P_null v and partial sums of v . w, node by node (grid-stride)
template <class T>
__global__ void __launch_bounds__(RED_T) k_project_dot_part(int nb, const uint8_t* __restrict__ flags, T* v,
                                                            const T* __restrict__ w, double* part) {
  double s = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < (int64_t)nb * 64;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int blk = (int)(g >> 6), l = (int)(g & 63);
    const uint8_t fl = flags[g];
    const bool dead = !(fl & 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int k = vidx(blk, c, l);
      if (dead || (fl & (2 << c))) v[k] = T(0);
      else s += (double)v[k] * (double)w[k];
    }
  }
  s = block_reduce<0>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
template <class T>
__global__ void k_project(int nb, const uint8_t* flags, T* v, int nullspace) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  const uint8_t fl = flags[gid];
  const bool dead = nullspace && !(fl & 1);
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if (dead || (fl & (2 << c))) v[vidx(blk, c, l)] = T(0);
}

// ------------------------------------------------------------------ Galerkin
// G[cell][q] = E of the fine cell at position q inside the coarse cell
__global__ void k_galerkin_gather(int64_t ncell, int depth, const int* __restrict__ cfb, const int* __restrict__ cgi,
                                  const double* __restrict__ E0, double* __restrict__ G) {
  const int n = 1 << depth;
  const int64_t npos = (int64_t)n * n * n;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ncell * npos) return;
  const int64_t cell = gid / npos;
  const int q = (int)(gid % npos);
  if (depth == 0) {
    G[gid] = E0[cgi[cell]];
    return;
  }
  const int px = q / (n * n), py = (q / n) % n, pz = q % n;
  const int nper = n >> 2;
  const int fb = cfb[cell * nper * nper * nper + ((px >> 2) * nper + (py >> 2)) * nper + (pz >> 2)];
  G[gid] = fb >= 0 ? E0[fb * 64 + loc3(px, py, pz)] : 0.0;
}

// stencil of every node from the 24x24 matrices of its 8 incident cells
// (Galerkin_Coarsen.h:59-98), times dx, then Dirichlet rows -> identity and
// Dirichlet columns -> 0 (Delete_Dirichlet_Equations.h:24-51)
__global__ void __launch_bounds__(TPB) k_galerkin_assemble(int nb, const int* __restrict__ nbr,
                                                           const uint8_t* __restrict__ flags,
                                                           const int* __restrict__ cell_row,
                                                           const double* __restrict__ Cm, double dx,
                                                           double* __restrict__ ST) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const uint8_t fl = flags[gid];
  const Ring R = make_ring(nbr, blk, li, lj, lk);
  int row[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const int a = li + ((p >> 2) & 1) - 1, b = lj + ((p >> 1) & 1) - 1, c = lk + (p & 1) - 1;
    const int bb = ring_block(R, a, b, c);
    row[p] = ((fl & 1) && bb >= 0) ? cell_row[bb * 64 + loc3(a, b, c)] : -1;
  }
  for (int s = 0; s < 27; ++s) {
    const int o[3] = {s / 9, (s / 3) % 3, s % 3};
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int pp[3] = {(p >> 2) & 1, (p >> 1) & 1, p & 1};
      const int J[3] = {o[0] - pp[0], o[1] - pp[1], o[2] - pp[2]};
      if (J[0] < 0 || J[0] > 1 || J[1] < 0 || J[1] > 1 || J[2] < 0 || J[2] > 1) continue;
      if (row[p] < 0) continue;
      const int I = ((1 - pp[0]) << 2) | ((1 - pp[1]) << 1) | (1 - pp[2]);
      const int Jn = (J[0] << 2) | (J[1] << 1) | J[2];
      const double* M = Cm + (int64_t)row[p] * 576 + (I * 8 + Jn) * 9;
#pragma unroll
      for (int e = 0; e < 9; ++e) acc[e] += M[e];
    }
    // neighbour flags for column deletion
    const int a = li + o[0] - 1, b = lj + o[1] - 1, c = lk + o[2] - 1;
    const int nbk = ring_block(R, a, b, c);
    uint8_t mfl = nbk >= 0 ? flags[nbk * 64 + loc3(a, b, c)] : 0;
    if (!(mfl & 1)) mfl = 0;
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
      for (int w = 0; w < 3; ++w) {
        double val = acc[v * 3 + w] * dx;
        if (!(fl & 1)) val = 0.0;
        else if (fl & (2 << v)) val = (v == w && s == 13) ? 1.0 : 0.0;
        else if (mfl & (2 << w)) val = 0.0;
        ST[st_index(blk, l, s * 9 + v * 3 + w)] = val;
      }
  }
}

// ------------------------------------------------------------------- bottom
__global__ void k_dense_assemble(int nb, const int* __restrict__ nbr, const uint8_t* __restrict__ flags,
                                 const double* __restrict__ ST, const int* __restrict__ var_of, int64_t n,
                                 double* __restrict__ A) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  const int vi = var_of[gid];
  if (vi < 0) return;
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const Ring R = make_ring(nbr, blk, li, lj, lk);
  for (int s = 0; s < 27; ++s) {
    const int a = li + s / 9 - 1, b = lj + (s / 3) % 3 - 1, c = lk + s % 3 - 1;
    const int bb = ring_block(R, a, b, c);
    if (bb < 0) continue;
    const int vj = var_of[bb * 64 + loc3(a, b, c)];
    if (vj < 0) continue;
    for (int v = 0; v < 3; ++v)
      for (int w = 0; w < 3; ++w) A[(int64_t)(vi + v) + (int64_t)(vj + w) * n] = ST[st_index(blk, l, s * 9 + v * 3 + w)];
  }
}
__global__ void k_symmetrize_lower(double* A, int64_t n) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * n) return;
  const int64_t i = gid % n, j = gid / n;  // A[i + j n]
  if (i < j) A[i + j * n] = A[j + i * n];
}
__global__ void k_set_diagonal(double* A, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i + i * n] = v;
}
template <class T>
__global__ void k_bottom_gather(int nb, const int* __restrict__ var_of, const T* __restrict__ b, T* x) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int vi = var_of[gid];
  if (vi < 0) return;
  const int blk = gid >> 6, l = gid & 63;
  for (int c = 0; c < 3; ++c) x[vi + c] = b[vidx(blk, c, l)];
}
template <class T>
__global__ void k_bottom_scatter(int nb, const int* __restrict__ var_of, const T* y, const T* __restrict__ b, T* u) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int vi = var_of[gid];
  const int blk = gid >> 6, l = gid & 63;
  for (int c = 0; c < 3; ++c) u[vidx(blk, c, l)] = vi >= 0 ? y[vi + c] : b[vidx(blk, c, l)];
}

// The symmetric bottom inverse kept as its lower triangle in 64x64 tiles:
// tile (I, J), I >= J, at index I(I+1)/2 + J, row-major, zero padded.
template <class T>
__global__ void k_pack_sym_tiles(const double* __restrict__ A, int64_t nv, int nT, T* __restrict__ tiles) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ntile = (int64_t)nT * (nT + 1) / 2;
  if (gid >= ntile * 4096) return;
  const int64_t tile = gid >> 12;
  const int e = (int)(gid & 4095);
  int I = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((int64_t)(I + 1) * (I + 2) / 2 <= tile) ++I;
  while ((int64_t)I * (I + 1) / 2 > tile) --I;
  const int J = (int)(tile - (int64_t)I * (I + 1) / 2);
  const int64_t row = (int64_t)I * 64 + (e >> 6), col = (int64_t)J * 64 + (e & 63);
  tiles[gid] = (row < nv && col < nv) ? (T)A[row + col * nv] : T(0);
}

// y = A x with the packed tiles.  Workgroup (I, s) takes the tile columns
// J = s, s+S, ... of row block I: a stored tile (J <= I) contributes
// T(I,J) x_J, a mirrored one (J > I) contributes T(J,I)^T x_J.  Partial
// sums per s are added in a fixed order by k_symv_reduce (deterministic).
template <class T>
__global__ void __launch_bounds__(256) k_symv_tiles(int nT, int S, const T* __restrict__ tiles,
                                                    const T* __restrict__ x, T* __restrict__ part) {
  const int I = blockIdx.x, s = blockIdx.y;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  T accN[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) accN[k] = 0;
  T accT = 0;
  for (int J = s; J < nT; J += S) {
    if (J <= I) {
      const T* P = tiles + ((int64_t)I * (I + 1) / 2 + J) * 4096;
      const T xv = x[J * 64 + lane];
#pragma unroll
      for (int k = 0; k < 16; ++k) accN[k] += P[(w + 4 * k) * 64 + lane] * xv;
    } else {
      const T* P = tiles + ((int64_t)J * (J + 1) / 2 + I) * 4096;
#pragma unroll
      for (int k = 0; k < 16; ++k) accT += P[(w + 4 * k) * 64 + lane] * x[J * 64 + w + 4 * k];
    }
  }
  __shared__ T ys[64];
  __shared__ T yt[4][64];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    T v = accN[k];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) ys[w + 4 * k] = v;
  }
  yt[w][lane] = accT;
  __syncthreads();
  if (t < 64) part[((int64_t)s * nT + I) * 64 + t] = ys[t] + yt[0][t] + yt[1][t] + yt[2][t] + yt[3][t];
}
template <class T>
__global__ void k_symv_reduce(int n, int S, int64_t stride, const T* __restrict__ part, T* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T a = 0;
  for (int s = 0; s < S; ++s) a += part[s * stride + i];
  y[i] = a;
}

// ---------------------------------------------------- density -> SIMP moduli
// Device form of the density update (SIMP map, topo_opt/fem_interface.h:
// 336-353, applied as SPGrid_CPU_Solver.cpp:112-130): a read-only check that
// the active-cell set is unchanged, then the update of rho and E.  A slot
// holds an active cell exactly when its current E is non-zero.
__global__ void k_density_check(int64_t n, const int* __restrict__ slot, const double* __restrict__ dens,
                                const double* __restrict__ E64, int* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = slot[i];
  if (s < 0) return;
  const double d = dens[i];
  int e = 0;
  if (d != d) e = 1;                                 // NaN
  else if (E64[s] != 0.0 && !(d > 0)) e = 2;         // would deactivate a cell
  else if (E64[s] == 0.0 && d > 0) e = 4;            // would activate a cell
  if (e) atomicOr(err, e);
}
__global__ void k_density_update(int64_t n, const int* __restrict__ slot, const double* __restrict__ dens,
                                 double penalty, double emin, double rmin, double* __restrict__ E64,
                                 float* __restrict__ E32, double* __restrict__ rho) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = slot[i];
  if (s < 0) return;
  const double d = dens[i];
  rho[s] = d;
  if (E64[s] != 0.0) {
    const double e = penalty != 0.0 ? (1.0 - emin) * pow(fmax(rmin, d), penalty) + emin : d;
    E64[s] = e;
    E32[s] = (float)e;
  }
}

// -------------------------------------------------------------- sensitivity
// s_e = a_e^T K0 u_e over the 8 nodes of the cell at this slot
// (objective/minimal_compliance.cpp:86-146; K0 without dx)
__global__ void __launch_bounds__(TPB) k_sensitivity(int nb, const int* __restrict__ nbr, const double* __restrict__ rho,
                                                     const double* __restrict__ u, const double* __restrict__ a,
                                                     double penalty, double emin, double* __restrict__ dc,
                                                     double* __restrict__ objcell) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  const double d = rho[gid];
  if (!(d > 0.0)) {
    dc[gid] = 0.0;
    if (objcell) objcell[gid] = 0.0;
    return;
  }
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const Ring R = make_ring(nbr, blk, li, lj, lk);
  double ue[24], ae[24];
#pragma unroll
  for (int J = 0; J < 8; ++J) {
    const int x = li + ((J >> 2) & 1), y = lj + ((J >> 1) & 1), z = lk + (J & 1);
    const int bb = ring_block(R, x, y, z);
    const int q = loc3(x, y, z);
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      ue[3 * J + w] = bb >= 0 ? u[vidx(bb, w, q)] : 0.0;
      ae[3 * J + w] = bb >= 0 ? a[vidx(bb, w, q)] : 0.0;
    }
  }
  double s = 0.0;
#pragma unroll
  for (int I = 0; I < 8; ++I)
#pragma unroll
    for (int J = 0; J < 8; ++J)
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        const int kb = ((I * 8 + J) * 3 + v) * 3;
        s += ae[3 * I + v] * (cK0d[kb] * ue[3 * J] + cK0d[kb + 1] * ue[3 * J + 1] + cK0d[kb + 2] * ue[3 * J + 2]);
      }
  dc[gid] = -(s * pow(d, penalty - 1.0) * penalty * (1.0 - emin));
  if (objcell) objcell[gid] = s * (emin + pow(d, penalty) * (1.0 - emin));
}

// -------------------------------------------------------- input-order maps
__global__ void k_scatter_input_vec(const int* in_slot, int64_t nin, const double* in3, double* v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin) return;
  const int s = in_slot[i];
  if (s < 0) return;
  const int blk = s >> 6, l = s & 63;
  for (int c = 0; c < 3; ++c) v[vidx(blk, c, l)] = in3[i * 3 + c];
}
__global__ void k_gather_input_vec(const int* in_slot, int64_t nin, const double* v, double* out3) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin) return;
  const int s = in_slot[i];
  const int blk = s >> 6, l = s & 63;
  for (int c = 0; c < 3; ++c) out3[i * 3 + c] = s >= 0 ? v[vidx(blk, c, l)] : 0.0;
}
// Output records, one per input block: `hdr` 32-bit words of header (the
// block's base coordinates when hdr = 3), then its 64 nodes as
// {u_x, u_y, u_z, max_c |r_c|} doubles, zero for a node without a slot
// (SPGrid_CPU_Solver.cpp:345-384).  Stored as 32-bit words: after a 12-byte
// header the doubles are only 4-byte aligned.
__global__ void k_pack_out4(const int* __restrict__ in_slot, int64_t nin, const double* __restrict__ u,
                            const double* __restrict__ r, const int32_t* __restrict__ base, int hdr,
                            uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin) return;
  const int64_t b = i >> 6;
  const int l = (int)(i & 63);
  uint32_t* rec = out + b * (hdr + 512);
  if (l < hdr) rec[l] = (uint32_t)base[b * 3 + l];
  double o[4] = {0.0, 0.0, 0.0, 0.0};
  const int s = in_slot[i];
  if (s >= 0) {
    const int blk = s >> 6, sl = s & 63;
    for (int c = 0; c < 3; ++c) {
      o[c] = u[vidx(blk, c, sl)];
      o[3] = fmax(o[3], fabs(r[vidx(blk, c, sl)]));
    }
  }
  uint32_t* dst = rec + hdr + l * 8;
  for (int k = 0; k < 4; ++k) {
    const uint64_t w = (uint64_t)__double_as_longlong(o[k]);
    dst[2 * k] = (uint32_t)w;
    dst[2 * k + 1] = (uint32_t)(w >> 32);
  }
}
__global__ void k_gather_input_scalar(const int* in_slot, int64_t nin, const double* v, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin) return;
  const int s = in_slot[i];
  out[i] = s >= 0 ? v[s] : 0.0;
}
// the entries are distinct (the host keeps one per slot and component)
__global__ void k_set_entries(double* v, const int* __restrict__ idx, const double* __restrict__ val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = idx[i] / 3, c = idx[i] % 3;
  v[vidx(s >> 6, c, s & 63)] = val[i];
}

// ======================================================== launch wrappers
// The FP32 operator of the Krylov recurrence (apply mode 2, q = P K s):
// direct FP32 accumulation of the 576-term rows loses accuracy to
// cancellation on smooth search directions and cost the FP32 MG-PCG 23
// instead of 18 iterations on the 256^3 shell (the FP64 solve and the
// reference's float mode take 18).  The rows are therefore summed in FP64
// (SPG_ACC64_MODES, bit m = apply mode m).  The difference form with FP32
// sums (SPG_DIFF_SUMS=1) removes the cancellation of translations only: it
// keeps 18 iterations on the shell at 118 instead of 152-160 us per apply,
// but the 2048-wide ellipsoid, whose displacement is mostly rotation, needs
// 44 iterations in 3 passes instead of 32 in 2.  The V-cycle residual
// (mode 1) keeps plain FP32 sums: the preconditioner tolerates them.
static int acc64_modes() {
  const char* e = std::getenv("SPG_ACC64_MODES");
  return e ? std::atoi(e) : 4;
}
template <class T, class A>
static void launch_apply_t(hipStream_t st, const LevelDev& L, const T* E, T dx, const T* in, const T* b, T* out,
                           int mode) {
  const int nwg = (L.nb + AP_WAVES - 1) / AP_WAVES;
  if (mode == 0)
    k_apply_lds<T, A, 0><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, b, out);
  else if (mode == 2)
    k_apply_lds<T, A, 2><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, b, out);
  else
    k_apply_lds<T, A, 1><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, b, out);
}
// SPG_DIFF_SUMS=1: the FP32 Krylov operator P K s (mode 2) in difference
// form with FP32 sums instead of the direct form with FP64 sums
static bool diff_sums() {
  const char* e = std::getenv("SPG_DIFF_SUMS");
  return e && std::atoi(e) == 1;
}
// Which applies run cell by cell in strain form (k_apply_cell): by default
// the FP32 Krylov operator P K s (modes 2 and 3), whose FP32 row sums would
// otherwise need FP64 accumulation; SPG_APPLY_CELL=0 none, 1 every apply.
// 256^3 shell, per apply: FP32 operator 121 us (cell) vs 162 us (rows with
// FP64 sums); FP32 residual 116 vs 107 us and FP64 228 vs 179 us favour the
// rows there.
static bool apply_cell(bool fp32, int mode) {
  const char* e = std::getenv("SPG_APPLY_CELL");
  if (e) return std::atoi(e) == 1;
  return fp32 && mode >= 2;
}
template <class T>
static void launch_apply_cell(hipStream_t st, const LevelDev& L, const T* E, T dx, const T* in, const T* b, T* out,
                              int mode, double* dpart = nullptr) {
  const int nwg = (L.nb + AP_WAVES - 1) / AP_WAVES;
  const CellConsts k = L.cell;
  if (mode == 0)
    k_apply_cell<T, 0><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, k, dx, in, b, out);
  else if (mode == 1)
    k_apply_cell<T, 1><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, k, dx, in, b, out);
  else if (mode == 2)
    k_apply_cell<T, 2><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, k, dx, in, b, out);
  else
    k_apply_cell<T, 3><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, k, dx, in, b, out, dpart);
}
template <class T>
void launch_apply(hipStream_t st, const LevelDev& L, const T* E, T dx, const T* in, const T* b, T* out, int mode) {
  if (apply_cell(sizeof(T) == 4, mode)) {
    launch_apply_cell<T>(st, L, E, dx, in, b, out, mode);
  } else if (sizeof(T) == 4 && mode == 2 && diff_sums()) {
    const int nwg = (L.nb + AP_WAVES - 1) / AP_WAVES;
    k_apply_lds<T, T, 4><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, b, out);
  } else if (sizeof(T) == 4 && ((acc64_modes() >> mode) & 1)) {
    launch_apply_t<T, double>(st, L, E, dx, in, b, out, mode);
  } else {
    launch_apply_t<T, T>(st, L, E, dx, in, b, out, mode);
  }
}
int64_t apply_dot_partials(int64_t nb) { return nb > RED_MAX_WG ? nb : RED_MAX_WG; }
template <class T>
void launch_apply_dot(hipStream_t st, const LevelDev& L, const T* E, T dx, const T* in, T* out, double* bpart,
                      double* part, double* dot, CgScalars* sc, int stage) {
  // 256^3 shell, 18 iterations: mixed 42.8 vs 43.1 ms, FP64 60.0 vs 60.5 ms
  // against the apply followed by a separate dot
  const int nwg = (L.nb + AP_WAVES - 1) / AP_WAVES;
  if (apply_cell(sizeof(T) == 4, 3))
    launch_apply_cell<T>(st, L, E, dx, in, nullptr, out, 3, bpart);
  else if (sizeof(T) == 4 && diff_sums())
    k_apply_lds<T, T, 5><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, nullptr, out, bpart);
  else if (sizeof(T) == 4 && ((acc64_modes() >> 2) & 1))
    k_apply_lds<T, double, 3><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, nullptr, out, bpart);
  else
    k_apply_lds<T, T, 3><<<nwg, 64 * AP_WAVES, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, nullptr, out, bpart);
  // one workgroup sums up to 65536 partials; beyond that two stages
  // (SPG_SUM_WIDE_MAX lowers the limit so that tests reach both forms)
  int wide_max = 65536;
  if (const char* e = std::getenv("SPG_SUM_WIDE_MAX")) wide_max = std::atoi(e);
  if (L.nb <= wide_max) {
    k_sum_wide<<<1, 1024, 0, st>>>(bpart, L.nb, dot, sc, stage);
  } else {
    const unsigned g = red_grid(L.nb);
    k_sum_part<<<g, RED_T, 0, st>>>(bpart, L.nb, part);
    k_final<0><<<1, RED_T, 0, st>>>(part, (int)g, dot, sc, stage);
  }
}
static ColourOrder colour_order(bool level0, bool forward) {
  ColourOrder o{0};
  for (int i = 0; i < 8; ++i) o.v |= (unsigned)sweep_colour(level0, forward, i) << (4 * i);
  return o;
}
static BndLists bnd_lists(const LevelDev& L) {
  BndLists bl;
  bl.list = L.bnd;
  for (int c = 0; c < 9; ++c) bl.off[c] = L.bnd_off[c];
  return bl;
}
// a boundary phase runs in one workgroup while no colour list is longer than this
static constexpr int kFusedBoundaryMax = 4096;
// SPG_BOUNDARY_PATH selects the boundary smoother for tests: 0 (default)
// automatic (level 0: the cooperative matrix-free phase, else one launch per
// colour step; explicit levels: the dense map, else the LDS coefficient-row
// system), 1 the single-workgroup global-memory kernel, 2 one launch per
// colour step, 3 the coefficient-row LDS system also on level 0, 4 the
// matrix-free single-workgroup level-0 phase, 5 the cooperative level-0
// phase (automatic on the explicit levels).
static int boundary_path() {
  const char* e = std::getenv("SPG_BOUNDARY_PATH");
  return e ? std::atoi(e) : 0;
}

template <class T>
struct BndPtr;
template <>
struct BndPtr<float> {
  static const float* A(const LevelDev& L) { return L.bA32; }
  static float* bt(const LevelDev& L) { return L.bt32; }
};
template <>
struct BndPtr<double> {
  static const double* A(const LevelDev& L) { return L.bA64; }
  static double* bt(const LevelDev& L) { return L.bt64; }
};
// the boundary phase on the LDS-resident system; false if it does not fit
template <class T>
static bool launch_bnd_phase(hipStream_t st, const LevelDev& L, T* u, const T* b, int sweeps, bool level0,
                             bool forward) {
  const int n = L.bnd_off[8];
  const size_t lds = (size_t)3 * n * sizeof(T) + (size_t)n;
  const T* A = BndPtr<T>::A(L);
  T* bt = BndPtr<T>::bt(L);
  if (!A || !bt || !L.bci || !L.bslot || lds > kMaxLds) return false;
  constexpr int NT = BndThreads<T>::value;
  // the attribute is per device: set at every launch (a host-side call)
  (void)hipFuncSetAttribute((const void*)k_bnd_sweeps<T, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kMaxLds);
  k_bnd_rhs<T><<<(n + 31) / 32, 256, 0, st>>>(n, L.bnd, L.bslot, A, u, b, bt);
  BndOff o;
  for (int c = 0; c < 9; ++c) o.off[c] = L.bnd_off[c];
  k_bnd_sweeps<T, NT><<<1, NT, lds, st>>>(n, o, colour_order(level0, forward), sweeps, L.bnd, L.bci, A, L.flags, bt,
                                          u);
  return true;
}
template <class T>
struct Bnd0Ptr;
template <>
struct Bnd0Ptr<float> {
  static const float* E(const LevelDev& L) { return L.b0E32; }
  static const float* D(const LevelDev& L) { return L.b0D32; }
};
template <>
struct Bnd0Ptr<double> {
  static const double* E(const LevelDev& L) { return L.b0E64; }
  static const double* D(const LevelDev& L) { return L.b0D64; }
};
// level-0 boundary phase on the matrix-free LDS-resident system; false if
// the tables are absent or the LDS does not hold it
template <class T>
static bool launch_bnd0_phase(hipStream_t st, const LevelDev& L, T dx, T* u, const T* b, int sweeps, bool forward) {
  const int n = L.bnd_off[8];
  const size_t lds = (size_t)(3 * (n + 1) + 2 * 8 * 3 * B0_NODES) * sizeof(T) + (size_t)n;
  const T* A = BndPtr<T>::A(L);
  T* bt = BndPtr<T>::bt(L);
  const T* bE = Bnd0Ptr<T>::E(L);
  const T* bD = Bnd0Ptr<T>::D(L);
  if (!A || !bt || !bE || !bD || !L.b0node || !L.bslot || lds > kMaxLds) return false;
  (void)hipFuncSetAttribute((const void*)k_bnd0_sweeps<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
  k_bnd_rhs<T><<<(n + 31) / 32, 256, 0, st>>>(n, L.bnd, L.bslot, A, u, b, bt);
  BndOff o;
  for (int c = 0; c < 9; ++c) o.off[c] = L.bnd_off[c];
  k_bnd0_sweeps<T><<<1, B0_THREADS, lds, st>>>(n, o, colour_order(true, forward), sweeps, L.bnd, L.flags, bE,
                                               L.b0node, bD, bt, dx, u);
  return true;
}
// level-0 boundary phase as one launch over several CUs; false if the
// tables are absent or a cooperative launch is refused (the caller then runs
// another path; `loud` reports a refused launch on stderr).  SPG_COOP_WG
// caps the workgroup count (tests use it to make a workgroup take several
// node groups per colour step).
template <class T>
static bool launch_bnd0_coop(hipStream_t st, const LevelDev& L, T dx, T* u, const T* b, int sweeps, bool forward,
                             bool loud) {
  const int n = L.bnd_off[8];
  const T* A = BndPtr<T>::A(L);
  T* bt = BndPtr<T>::bt(L);
  const T* bE = Bnd0Ptr<T>::E(L);
  const T* bD = Bnd0Ptr<T>::D(L);
  if (!A || !bt || !bE || !bD || !L.b0node || !L.bslot || !L.b0ub || !L.coop) return false;
  int groups = 1;
  for (int c = 0; c < 8; ++c) groups = std::max(groups, (L.bnd_off[c + 1] - L.bnd_off[c] + 63) / 64);
  int G = std::min(groups, BC_MAX_WG);
  if (const char* e = std::getenv("SPG_COOP_WG"))
    if (std::atoi(e) > 0) G = std::min(G, std::atoi(e));
  static const unsigned long long wait_50ms = [] {
    int dev = 0, khz = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
    return (unsigned long long)khz * 50;
  }();
  // SPG_COOP_LIMIT_TICKS shortens the wait (tests force the time-out path)
  unsigned long long limit = wait_50ms;
  if (const char* e = std::getenv("SPG_COOP_LIMIT_TICKS"))
    if (std::atoll(e) > 0) limit = (unsigned long long)std::atoll(e);
  T* ub = reinterpret_cast<T*>(L.b0ub);
  k_bnd_rhs<T><<<(n + 31) / 32, 256, 0, st>>>(n, L.bnd, L.bslot, A, u, b, bt, ub, L.coop);
  BndOff o;
  for (int c = 0; c < 9; ++c) o.off[c] = L.bnd_off[c];
  ColourOrder ord = colour_order(true, forward);
  const int16_t* cn = L.b0node;
  const int* list = L.bnd;
  const uint8_t* flags = L.flags;
  unsigned* coop = L.coop;
  unsigned long long lim = limit;
  int nn = n, sw = sweeps;
  // An ordinary launch by default: the solver's stream runs one kernel at a
  // time, so the G <= 32 small workgroups all start at once on the empty
  // device (256 CUs).  The cooperative launch API guarantees that even
  // while other streams fill the device, but costs about 40 us per phase
  // (256^3 shell: phase 96 vs 115 us, mixed solve 40.2 vs 42.9 ms);
  // SPG_COOP_LAUNCH=1 selects it.  Either way a barrier wait ends after
  // `limit` and is reported.
  const char* cl = std::getenv("SPG_COOP_LAUNCH");
  if (!(cl && std::atoi(cl) == 1)) {
    k_bnd0_coop<T><<<G, BC_THREADS, 0, st>>>(nn, o, ord, sw, list, flags, bE, cn, bD, bt, dx, ub, u, coop, lim);
    return true;
  }
  void* args[] = {&nn, &o, &ord, &sw, &list, &flags, &bE, &cn, &bD, &bt, &dx, &ub, &u, &coop, &lim};
  const hipError_t e = hipLaunchCooperativeKernel((const void*)k_bnd0_coop<T>, dim3(G), dim3(BC_THREADS), args, 0, st);
  if (e == hipSuccess) return true;
  (void)hipGetLastError();
  if (loud)
    std::fprintf(stderr, "spgrid_topo_opt_amd: cooperative boundary phase not launched (%s); using the list path\n",
               hipGetErrorString(e));
  return false;
}
void launch_bnd0_setup(hipStream_t st, const LevelDev& L, const int* cslot, const double* E, double dx, double* bE,
                       double* bD) {
  const int n = L.bnd_off[8];
  if (n > 0) k_bnd0_setup<<<nblk(n), TPB, 0, st>>>(n, L.bnd, L.flags, cslot, E, dx, bE, bD);
}
void launch_bnd_coef0(hipStream_t st, const LevelDev& L, const double* E, double dx, double* A) {
  const int n = L.bnd_off[8];
  if (n > 0) k_bnd_coef0<<<nblk((int64_t)n * 32), TPB, 0, st>>>(n, L.bnd, L.nbr, E, dx, A);
}
void launch_bmap_build(hipStream_t st, const LevelDev& L, int sweeps, bool forward, double* M) {
  const int n = L.bnd_off[8];
  if (n == 0) return;
  BndOff o;
  for (int c = 0; c < 9; ++c) o.off[c] = L.bnd_off[c];
  k_bmap_build<<<(6 * n + 63) / 64, 64 * BM_WAVES, 0, st>>>(n, o, colour_order(false, forward), sweeps, L.bnd, L.bci,
                                                           L.bA64, L.flags, M);
}
void launch_bnd_coef_st(hipStream_t st, const LevelDev& L, const double* ST, double* A) {
  const int n = L.bnd_off[8];
  if (n > 0) k_bnd_coef_st<<<nblk((int64_t)n * 32), TPB, 0, st>>>(n, L.bnd, ST, A);
}

// colour-major copies: thread i = (blk, comp, l) of the node-major vector
__device__ __forceinline__ int colour_of(int l) { return (((l >> 4) & 1) << 2) | (((l >> 2) & 1) << 1) | (l & 1); }
__device__ __forceinline__ int colour_k(int l) { return (((l >> 5) & 1) << 2) | (((l >> 3) & 1) << 1) | ((l >> 1) & 1); }
template <class T>
__global__ void k_colour_major(int nb, const T* __restrict__ v, T* __restrict__ vcm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nb * 192) return;
  const int blk = (int)(i / 192), comp = (int)((i / 64) % 3), l = (int)(i & 63);
  vcm[(((int64_t)colour_of(l) * nb + blk) * 3 + comp) * 8 + colour_k(l)] = v[i];
}
__global__ void k_colour_major_flags(int nb, const uint8_t* __restrict__ flags, uint8_t* __restrict__ fcm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nb * 64) return;
  const int blk = (int)(i >> 6), l = (int)(i & 63);
  fcm[((int64_t)colour_of(l) * nb + blk) * 8 + colour_k(l)] = flags[i];
}
template <class T>
void launch_colour_major(hipStream_t st, const LevelDev& L, const T* v, T* vcm) {
  k_colour_major<T><<<nblk((int64_t)L.nb * 192), TPB, 0, st>>>(L.nb, v, vcm);
}
void launch_colour_major_flags(hipStream_t st, int nb, const uint8_t* flags, uint8_t* fcm) {
  k_colour_major_flags<<<nblk((int64_t)nb * 64), TPB, 0, st>>>(nb, flags, fcm);
}

// SPG_GS_KERNEL selects the level-0 colour kernel: 1 the octet kernel
// k_gs0_colour, 2 the corner-per-wave k_gs0_colour_uc, 3 the quad-lane
// k_gs0_colour4; 0 or unset picks per precision.  All are parity-tested.
template <class T>
void launch_gs0(hipStream_t st, const LevelDev& L, const T* E, T dx, T* u, const T* b, int colour,
                const uint8_t* dist, int dmin, const T* bcm) {
  if (!L.flags_cm) bcm = nullptr;
  if (dist) {
    // the far part of a sweep beside the band phase: the quad kernel
    k_gs0_colour4<T, 8><<<(L.nb + 7) / 8, 256, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b, dist, dmin);
    return;
  }
  const char* e = std::getenv("SPG_GS_KERNEL");
  // measured on gfx950 (256^3 shell, per 8-colour sweep): quad 374 / 535 us,
  // octet 455 / 580 us, corner-per-wave 533 / 898 us (FP32 / FP64); two
  // more layouts were measured against the quad's later 302 / 445 us and
  // dropped: a lane per node over 8 blocks per wavefront (456 / 737 us,
  // 3 waves per SIMD leave the halo loads exposed) and two lanes per node
  // with the x-mirror split (315-352 / 541-653 us).  Against the quad's
  // 297-311 / 451-482 us, also measured and dropped: the quad with 16% fewer
  // VALU instructions (precomputed LDS offsets, one address per halo entry)
  // 311-318 / 466-477 us; the same with b and flags read from a colour-major
  // copy (29% fewer fetched bytes) 310 / 469 us; stored inverse self blocks
  // 359 / 571 us; a persistent, software-pipelined quad kernel with buffer
  // loads 389 / 600 us (118 VGPRs: 4 waves per SIMD); a cell corner per
  // wavefront with table-driven staging 436 / 726 us.  The quad kernel keeps
  // the VALUs about 94% busy, and with fewer VALU instructions the latency of
  // its two dependent load phases (neighbour rows, then the halo) sets the
  // same time.
  const int kind = e && std::atoi(e) > 0 ? std::atoi(e) : 3;
  if (kind == 2) {
    k_gs0_colour_uc<T><<<(L.nb + UC_BLOCKS - 1) / UC_BLOCKS, 512, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b);
  } else if (kind == 3) {
    // SPG_GS_QB blocks per workgroup (4, 8 or 16); 256^3 shell, FP32 / FP64
    // sweep: 4 -> 284-308 / 497-509 us, 8 -> 279-281 / 468 us, 16 -> 300 / 489 us
    const char* q = std::getenv("SPG_GS_QB");
    const int qb = q ? std::atoi(q) : 8;
    const uint8_t* fcm = bcm ? L.flags_cm : nullptr;
    if (qb == 4)
      k_gs0_colour4<T, 4><<<(L.nb + 3) / 4, 128, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b, nullptr, 0,
                                                           bcm, fcm);
    else if (qb == 16)
      k_gs0_colour4<T, 16><<<(L.nb + 15) / 16, 512, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b, nullptr, 0,
                                                              bcm, fcm);
    else
      k_gs0_colour4<T, 8><<<(L.nb + 7) / 8, 256, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b, nullptr, 0,
                                                           bcm, fcm);
  } else {
    const int nwg = (L.nb + GS_WAVES - 1) / GS_WAVES;
    k_gs0_colour<T><<<nwg, 64 * GS_WAVES, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b);
  }
}
template <class T>
void launch_gs0_boundary(hipStream_t st, const LevelDev& L, const T* E, T dx, T* u, const T* b, int sweeps,
                         bool forward) {
  if (sweeps <= 0 || L.bnd_off[8] == 0) return;
  int path = boundary_path();
  // Automatic choice, from measurement on the 256^3 shell (3655 boundary
  // nodes, 3 sweeps), FP32 / FP64 per phase: cooperative launch 123 / 136
  // us, per-colour list launches 145 / 168 us, single workgroup 154 / 235
  // us (one CU's VALU work).  The lists remain for sets beyond the int16
  // node tables and for a refused cooperative launch.
  if (path == 0 || path == 5) {
    if (!L.coop_off && launch_bnd0_coop<T>(st, L, dx, u, b, sweeps, forward, path == 5)) return;
    path = 2;
  }
  if (path == 4 && launch_bnd0_phase<T>(st, L, dx, u, b, sweeps, forward)) return;
  if (path == 3 && launch_bnd_phase<T>(st, L, u, b, sweeps, true, forward)) return;
  const ColourOrder ord = colour_order(true, forward);
  if (path != 2 && L.bnd_max <= kFusedBoundaryMax) {
    k_gs0_boundary<T><<<1, BND_THREADS, 0, st>>>(bnd_lists(L), ord, sweeps, L.nbr, L.flags, E, dx, u, b);
    return;
  }
  for (int sw = 0; sw < sweeps; ++sw)
    for (int i = 0; i < 8; ++i) {
      const int c = ord.at(i), n = L.bnd_off[c + 1] - L.bnd_off[c];
      if (n > 0)
        k_gs0_list_tab<T><<<(n + 31) / 32, 256, 0, st>>>(L.bnd + L.bnd_off[c], L.bnslot + (int64_t)L.bnd_off[c] * 27,
                                                         n, L.flags, E, dx, u, b);
    }
}
template <class T>
void launch_st_residual(hipStream_t st, const LevelDev& L, const T* ST, const T* u, const T* b, T* r) {
  const int nwg = (L.nb * 8 + GS_WAVES - 1) / GS_WAVES;
  k_st_colour<T, 1><<<nwg, 64 * GS_WAVES, 0, st>>>(L.nb, 0, L.nbr, L.flags, ST, const_cast<T*>(u), b, r);
}
template <class T>
void launch_st_gs(hipStream_t st, const LevelDev& L, const T* ST, T* u, const T* b, int colour) {
  const int nwg = (L.nb + GS_WAVES - 1) / GS_WAVES;
  k_st_colour<T, 0><<<nwg, 64 * GS_WAVES, 0, st>>>(L.nb, colour, L.nbr, L.flags, ST, u, b, nullptr);
}
template <class T>
struct BmapPtr;
template <>
struct BmapPtr<float> {
  static const float* M(const LevelDev& L, bool fwd) { return L.bmap32[fwd ? 0 : 1]; }
  static float* x(const LevelDev& L) { return L.bmx32; }
};
template <>
struct BmapPtr<double> {
  static const double* M(const LevelDev& L, bool fwd) { return L.bmap64[fwd ? 0 : 1]; }
  static double* x(const LevelDev& L) { return L.bmx64; }
};
template <class T>
void launch_st_boundary(hipStream_t st, const LevelDev& L, const T* ST, T* u, const T* b, int sweeps, bool forward) {
  if (sweeps <= 0 || L.bnd_off[8] == 0) return;
  const int path = boundary_path() == 5 ? 0 : boundary_path();  // 5 selects a level-0 phase only
  const T* M = BmapPtr<T>::M(L, forward);
  if (path == 0 && M && sweeps == L.bmap_sweeps) {
    const int n = L.bnd_off[8];
    const int Kp = (int)bmap_stride(n);
    T* x = BmapPtr<T>::x(L);  // [Kp] input (zero past 6n), then [3n] output behind it
    T* bt = BndPtr<T>::bt(L);
    // two launches: b~ and the map input, then the map with the result
    // written to u (the separate gather / scatter kernels cost two launch
    // gaps per phase)
    k_bnd_rhs<T><<<(n + 31) / 32, 256, 0, st>>>(n, L.bnd, L.bslot, BndPtr<T>::A(L), u, b, bt, nullptr, nullptr, x);
    k_bmap_gemv<T><<<(3 * n + 3) / 4, 256, 0, st>>>(3 * n, Kp, M, x, x + Kp, L.bnd, u);
    return;
  }
  if (path == 0 && launch_bnd_phase<T>(st, L, u, b, sweeps, false, forward)) return;
  const ColourOrder ord = colour_order(false, forward);
  if (path != 2 && L.bnd_max <= kFusedBoundaryMax) {
    k_st_boundary<T><<<1, BND_THREADS, 0, st>>>(bnd_lists(L), ord, sweeps, L.nbr, L.flags, ST, u, b);
    return;
  }
  for (int sw = 0; sw < sweeps; ++sw)
    for (int i = 0; i < 8; ++i) {
      const int c = ord.at(i), n = L.bnd_off[c + 1] - L.bnd_off[c];
      if (n > 0) k_st_list<T><<<(n + 31) / 32, 256, 0, st>>>(L.bnd + L.bnd_off[c], n, L.nbr, L.flags, ST, u, b);
    }
}
template <class T>
void launch_restrict(hipStream_t st, const LevelDev& C, const T* fine, T* coarse, T* coarse_zero) {
  k_restrict_lds<T><<<(C.nb + TR_WAVES - 1) / TR_WAVES, 64 * TR_WAVES, 0, st>>>(C.nb, C.rsrc, C.flags, fine, coarse,
                                                                              coarse_zero);
}
template <class T>
void launch_prolongate(hipStream_t st, const LevelDev& F, const T* coarse, T* fine) {
  k_prolong_lds<T><<<(F.nb + TR_WAVES - 1) / TR_WAVES, 64 * TR_WAVES, 0, st>>>(F.nb, F.psrc, F.base, F.psrc27, F.flags,
                                                                              coarse, fine);
}
void launch_prolong_table(hipStream_t st, const LevelDev& F, int* out) {
  if (F.nb > 0) k_prolong_table<<<nblk((int64_t)F.nb * 27), TPB, 0, st>>>(F.nb, F.psrc, F.base, out);
}
template <class T>
void launch_zero(hipStream_t st, T* v, int64_t n) {
  if (n > 0) k_zero<T><<<nblk(n), TPB, 0, st>>>(v, n);
}
template <class T>
void launch_copy(hipStream_t st, const T* a, T* b, int64_t n) {
  if (n > 0) k_copy<T><<<nblk(n), TPB, 0, st>>>(a, b, n);
}
void launch_cast_d2f(hipStream_t st, const double* a, float* b, int64_t n) {
  if (n > 0) k_d2f<<<nblk(n), TPB, 0, st>>>(a, b, n);
}
void launch_cast_f2d(hipStream_t st, const float* a, double* b, int64_t n) {
  if (n > 0) k_f2d<<<nblk(n), TPB, 0, st>>>(a, b, n);
}
template <class T>
void launch_axpy_d(hipStream_t st, double* u, const T* d, int64_t n) {
  if (n > 0) k_axpy_d<T><<<nblk(n), TPB, 0, st>>>(u, d, n);
}
template <class T>
void launch_dot(hipStream_t st, const T* a, const T* b, int64_t n, double* partial, double* out) {
  const unsigned g = red_grid(n);
  k_dot_part<T><<<g, RED_T, 0, st>>>(a, b, n, partial);
  k_final<0><<<1, RED_T, 0, st>>>(partial, (int)g, out);
}
template <class T>
void launch_linf(hipStream_t st, const T* a, int64_t n, double* partial, double* out) {
  const unsigned g = red_grid(n);
  k_linf_part<T><<<g, RED_T, 0, st>>>(a, n, partial);
  k_final<1><<<1, RED_T, 0, st>>>(partial, (int)g, out);
}
void launch_sumsq(hipStream_t st, const double* a, int64_t n, double* partial, double* out) {
  const unsigned g = red_grid(n);
  k_sumsq_part<<<g, RED_T, 0, st>>>(a, n, partial);
  k_final<0><<<1, RED_T, 0, st>>>(partial, (int)g, out);
}
template <class T>
void launch_cg_direction(hipStream_t st, T* s, const T* z, const CgScalars* sc, int64_t n) {
  k_cg_direction<T><<<nblk(n), TPB, 0, st>>>(s, z, sc, n);
}
template <class T>
void launch_cg_update(hipStream_t st, T* x, T* r, const T* s, const T* q, CgScalars* sc, int64_t n, int stage,
                      T* zclear) {
  // caller passes the partial buffer right after the scalars block
  double* part = reinterpret_cast<double*>(sc + 1);
  const unsigned g = red_grid(n);
  k_cg_update<T><<<g, RED_T, 0, st>>>(x, r, s, q, sc, n, part, zclear);
  k_final<1><<<1, RED_T, 0, st>>>(part, (int)g, &sc->nrm, stage >= 0 ? sc : nullptr, stage);
}
void launch_cg_scalars(hipStream_t st, CgScalars* sc, int stage) { k_cg_scalars<<<1, 64, 0, st>>>(sc, stage); }
template <class T>
void launch_project_dot(hipStream_t st, const LevelDev& L, T* v, const T* w, double* partial, double* out,
                        CgScalars* sc, int stage) {
  const unsigned g = red_grid((int64_t)L.nb * 64);
  k_project_dot_part<T><<<g, RED_T, 0, st>>>(L.nb, L.flags, v, w, partial);
  k_final<0><<<1, RED_T, 0, st>>>(partial, (int)g, out, sc, stage);
}
template <class T>
void launch_project(hipStream_t st, const LevelDev& L, T* v, bool nullspace) {
  k_project<T><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.flags, v, nullspace ? 1 : 0);
}
void launch_galerkin_gather(hipStream_t st, int64_t ncell, int depth, const int* cfb, const int* cgi, const double* E0,
                            double* G) {
  const int64_t n = ncell * ((int64_t)1 << (3 * depth));
  if (n > 0) k_galerkin_gather<<<nblk(n), TPB, 0, st>>>(ncell, depth, cfb, cgi, E0, G);
}
void launch_galerkin_assemble(hipStream_t st, const LevelDev& L, const int* cell_row, const double* Cm, double dx,
                              double* ST) {
  k_galerkin_assemble<<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.nbr, L.flags, cell_row, Cm, dx, ST);
}
void launch_stencil_to_float(hipStream_t st, const double* a, float* b, int64_t n) { launch_cast_d2f(st, a, b, n); }
void launch_dense_assemble(hipStream_t st, const LevelDev& L, const double* ST, const int* var_of, int64_t n,
                           double* A) {
  k_dense_assemble<<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.nbr, L.flags, ST, var_of, n, A);
}
void launch_list_nslot(hipStream_t st, const LevelDev& L, int n, const int* list, int* out) {
  if (n > 0) k_bnd_nslot<<<nblk((int64_t)n * 27), TPB, 0, st>>>(n, list, L.nbr, out);
}
template <class T>
void launch_gs0_list(hipStream_t st, const LevelDev& L, const int* list, const int* nslot, int n, const T* E, T dx,
                     T* u, const T* b) {
  if (n > 0) k_gs0_list_tab<T><<<(n + 31) / 32, 256, 0, st>>>(list, nslot, n, L.flags, E, dx, u, b);
}
void launch_bnd_nslot(hipStream_t st, const LevelDev& L, int* out) {
  const int n = L.bnd_off[8];
  if (n > 0) k_bnd_nslot<<<nblk((int64_t)n * 27), TPB, 0, st>>>(n, L.bnd, L.nbr, out);
}
void launch_density_check(hipStream_t st, int64_t n, const int* slot, const double* dens, const double* E64,
                          int* err) {
  if (n > 0) k_density_check<<<nblk(n), TPB, 0, st>>>(n, slot, dens, E64, err);
}
void launch_density_update(hipStream_t st, int64_t n, const int* slot, const double* dens, double penalty,
                           double emin, double rmin, double* E64, float* E32, double* rho) {
  if (n > 0) k_density_update<<<nblk(n), TPB, 0, st>>>(n, slot, dens, penalty, emin, rmin, E64, E32, rho);
}
void launch_set_diagonal(hipStream_t st, double* A, int64_t n, double v) {
  k_set_diagonal<<<nblk(n), TPB, 0, st>>>(A, n, v);
}
void launch_symmetrize_lower(hipStream_t st, double* A, int64_t n) {
  k_symmetrize_lower<<<nblk(n * n), TPB, 0, st>>>(A, n);
}
template <class T>
void launch_bottom_gather(hipStream_t st, const LevelDev& L, const int* var_of, const T* b, T* x) {
  k_bottom_gather<T><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, var_of, b, x);
}
template <class T>
void launch_bottom_scatter(hipStream_t st, const LevelDev& L, const int* var_of, const T* y, const T* b, T* u) {
  k_bottom_scatter<T><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, var_of, y, b, u);
}
template <class T>
void launch_pack_sym_tiles(hipStream_t st, const double* A, int64_t nv, int nT, T* tiles) {
  const int64_t n = (int64_t)nT * (nT + 1) / 2 * 4096;
  k_pack_sym_tiles<T><<<nblk(n), TPB, 0, st>>>(A, nv, nT, tiles);
}
int symv_splits(int nT) { return std::max(1, std::min(nT, 2048 / std::max(nT, 1))); }
template <class T>
void launch_symv(hipStream_t st, int nT, const T* tiles, const T* x, T* part, T* y) {
  const int S = symv_splits(nT);
  k_symv_tiles<T><<<dim3(nT, S), 256, 0, st>>>(nT, S, tiles, x, part);
  k_symv_reduce<T><<<nblk((int64_t)nT * 64), TPB, 0, st>>>(nT * 64, S, (int64_t)nT * 64, part, y);
}
void launch_sensitivity(hipStream_t st, const LevelDev& L, const double* rho, const double* u, const double* a,
                        double penalty, double emin, double* dc, double* objcell) {
  k_sensitivity<<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.nbr, rho, u, a, penalty, emin, dc, objcell);
}
void launch_scatter_input_vec(hipStream_t st, const int* in_slot, int64_t nin, const double* in3, double* v) {
  if (nin > 0) k_scatter_input_vec<<<nblk(nin), TPB, 0, st>>>(in_slot, nin, in3, v);
}
void launch_gather_input_vec(hipStream_t st, const int* in_slot, int64_t nin, const double* v, double* out3) {
  if (nin > 0) k_gather_input_vec<<<nblk(nin), TPB, 0, st>>>(in_slot, nin, v, out3);
}
void launch_pack_out4(hipStream_t st, const int* in_slot, int64_t nin, const double* u, const double* r,
                      const int32_t* base, int hdr, uint32_t* out) {
  if (nin > 0) k_pack_out4<<<nblk(nin), TPB, 0, st>>>(in_slot, nin, u, r, base, hdr, out);
}
void launch_gather_input_scalar(hipStream_t st, const int* in_slot, int64_t nin, const double* v, double* out) {
  if (nin > 0) k_gather_input_scalar<<<nblk(nin), TPB, 0, st>>>(in_slot, nin, v, out);
}
void launch_set_entries(hipStream_t st, double* v, const int* idx, const double* val, int64_t n) {
  if (n > 0) k_set_entries<<<nblk(n), TPB, 0, st>>>(v, idx, val, n);
}

#define SPG_INST(T)                                                                                              \
  template void launch_apply<T>(hipStream_t, const LevelDev&, const T*, T, const T*, const T*, T*, int);         \
  template void launch_apply_dot<T>(hipStream_t, const LevelDev&, const T*, T, const T*, T*, double*, double*,   \
                                    double*, CgScalars*, int);                                                  \
  template void launch_gs0<T>(hipStream_t, const LevelDev&, const T*, T, T*, const T*, int, const uint8_t*, int,  \
                              const T*);                                                                        \
  template void launch_colour_major<T>(hipStream_t, const LevelDev&, const T*, T*);                              \
  template void launch_gs0_list<T>(hipStream_t, const LevelDev&, const int*, const int*, int, const T*, T, T*,      \
                                   const T*);                                                                       \
  template void launch_gs0_boundary<T>(hipStream_t, const LevelDev&, const T*, T, T*, const T*, int, bool);      \
  template void launch_st_residual<T>(hipStream_t, const LevelDev&, const T*, const T*, const T*, T*);           \
  template void launch_st_gs<T>(hipStream_t, const LevelDev&, const T*, T*, const T*, int);                      \
  template void launch_st_boundary<T>(hipStream_t, const LevelDev&, const T*, T*, const T*, int, bool);          \
  template void launch_restrict<T>(hipStream_t, const LevelDev&, const T*, T*, T*);                              \
  template void launch_prolongate<T>(hipStream_t, const LevelDev&, const T*, T*);                                \
  template void launch_zero<T>(hipStream_t, T*, int64_t);                                                        \
  template void launch_copy<T>(hipStream_t, const T*, T*, int64_t);                                              \
  template void launch_axpy_d<T>(hipStream_t, double*, const T*, int64_t);                                       \
  template void launch_dot<T>(hipStream_t, const T*, const T*, int64_t, double*, double*);                       \
  template void launch_linf<T>(hipStream_t, const T*, int64_t, double*, double*);                                \
  template void launch_cg_direction<T>(hipStream_t, T*, const T*, const CgScalars*, int64_t);                    \
  template void launch_cg_update<T>(hipStream_t, T*, T*, const T*, const T*, CgScalars*, int64_t, int, T*);    \
  template void launch_project<T>(hipStream_t, const LevelDev&, T*, bool);                                       \
  template void launch_project_dot<T>(hipStream_t, const LevelDev&, T*, const T*, double*, double*, CgScalars*,  \
                                      int);                                                                     \
  template void launch_bottom_gather<T>(hipStream_t, const LevelDev&, const int*, const T*, T*);                 \
  template void launch_bottom_scatter<T>(hipStream_t, const LevelDev&, const int*, const T*, const T*, T*);      \
  template void launch_pack_sym_tiles<T>(hipStream_t, const double*, int64_t, int, T*);                          \
  template void launch_symv<T>(hipStream_t, int, const T*, const T*, T*, T*);
SPG_INST(float)
SPG_INST(double)

}  // namespace spg
