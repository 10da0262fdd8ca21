// kernels.hip -- gfx950 kernels of the narrow-band elasticity solve.
//
// Threads map to nodes (64 nodes = one 4x4x4 block = one wavefront).  Node
// neighbourhoods never leave the 3x3x3 ring of blocks addressed through the
// per-block 27-neighbour table; absent blocks read as zero, as unallocated
// SPGrid pages do in the reference.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdio>

#include "kernels.hpp"

namespace spg {

__constant__ double cK0d[576];
__constant__ float cK0f[576];

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

void upload_k0(const double* K0, hipStream_t st) {
  float f[576];
  for (int i = 0; i < 576; ++i) f[i] = (float)K0[i];
  (void)st;
  hipMemcpyToSymbol(HIP_SYMBOL(cK0d), K0, sizeof(double) * 576, 0, hipMemcpyHostToDevice);
  hipMemcpyToSymbol(HIP_SYMBOL(cK0f), f, sizeof(float) * 576, 0, hipMemcpyHostToDevice);
}

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
      t0 += KT<T>::k(kb + 0) * uu[s][0] + KT<T>::k(kb + 1) * uu[s][1] + KT<T>::k(kb + 2) * uu[s][2];
      t1 += KT<T>::k(kb + 3) * uu[s][0] + KT<T>::k(kb + 4) * uu[s][1] + KT<T>::k(kb + 5) * uu[s][2];
      t2 += KT<T>::k(kb + 6) * uu[s][0] + KT<T>::k(kb + 7) * uu[s][1] + KT<T>::k(kb + 8) * uu[s][2];
    }
    f0 += e * t0;
    f1 += e * t1;
    f2 += e * t2;
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

// the k-th node (0..7) of colour c in a 4x4x4 block
__device__ inline int colour_node(int colour, int k) {
  const int px = (colour >> 2) & 1, py = (colour >> 1) & 1, pz = colour & 1;
  const int i = px + 2 * ((k >> 2) & 1), j = py + 2 * ((k >> 1) & 1), kk = pz + 2 * (k & 1);
  return (i << 4) | (j << 2) | kk;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_gs0(int nb, int colour, const int* __restrict__ nbr,
                                             const uint8_t* __restrict__ flags, const T* __restrict__ E, T dx, T* u,
                                             const T* __restrict__ b, const int* __restrict__ list, int nlist) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  int blk, l;
  if (list) {
    if (gid >= nlist) return;
    const int s = list[gid];
    blk = s >> 6;
    l = s & 63;
  } else {
    if (gid >= nb * 8) return;
    blk = gid >> 3;
    l = colour_node(colour, gid & 7);
  }
  const uint8_t fl = flags[blk * 64 + l];
  if (fl & 1) {
    T f[3], d[6];
    row0<T, true>(nbr, E, u, blk, l, f, d);
    T r0 = b[vidx(blk, 0, l)] - dx * f[0];
    T r1 = b[vidx(blk, 1, l)] - dx * f[1];
    T r2 = b[vidx(blk, 2, l)] - dx * f[2];
#pragma unroll
    for (int e = 0; e < 6; ++e) d[e] *= dx;
    dirichlet_inverse(d, fl);
    const T u0 = u[vidx(blk, 0, l)] + d[0] * r0 + d[3] * r1 + d[4] * r2;
    const T u1 = u[vidx(blk, 1, l)] + d[3] * r0 + d[1] * r1 + d[5] * r2;
    const T u2 = u[vidx(blk, 2, l)] + d[4] * r0 + d[5] * r1 + d[2] * r2;
    u[vidx(blk, 0, l)] = (fl & 2) ? T(0) : u0;
    u[vidx(blk, 1, l)] = (fl & 4) ? T(0) : u1;
    u[vidx(blk, 2, l)] = (fl & 8) ? T(0) : u2;
  } else if (fl & 0x0E) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (fl & (2 << c)) u[vidx(blk, c, l)] = T(0);
  }
}

// --------------------------------------------------------- explicit levels
__device__ inline int sidx(int blk, int e, int l) { return (blk * 243 + e) * 64 + l; }

template <class T>
__device__ inline void row_st(const int* __restrict__ nbr, const T* __restrict__ ST, const T* __restrict__ u, int blk,
                              int l, T f[3]) {
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const Ring R = make_ring(nbr, blk, li, lj, lk);
  T f0 = 0, f1 = 0, f2 = 0;
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    const int a = li + s / 9 - 1, b = lj + (s / 3) % 3 - 1, c = lk + s % 3 - 1;
    const int bb = ring_block(R, a, b, c);
    if (bb < 0) continue;
    const int q = loc3(a, b, c);
    const T x = u[vidx(bb, 0, q)], y = u[vidx(bb, 1, q)], z = u[vidx(bb, 2, q)];
    const int e = s * 9;
    f0 += ST[sidx(blk, e + 0, l)] * x + ST[sidx(blk, e + 1, l)] * y + ST[sidx(blk, e + 2, l)] * z;
    f1 += ST[sidx(blk, e + 3, l)] * x + ST[sidx(blk, e + 4, l)] * y + ST[sidx(blk, e + 5, l)] * z;
    f2 += ST[sidx(blk, e + 6, l)] * x + ST[sidx(blk, e + 7, l)] * y + ST[sidx(blk, e + 8, l)] * z;
  }
  f[0] = f0;
  f[1] = f1;
  f[2] = f2;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_st_residual(int nb, const int* __restrict__ nbr,
                                                     const uint8_t* __restrict__ flags, const T* __restrict__ ST,
                                                     const T* __restrict__ u, const T* __restrict__ b,
                                                     T* __restrict__ r) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  T f[3];
  row_st(nbr, ST, u, blk, l, f);
  const uint8_t fl = flags[gid];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T v = b[vidx(blk, c, l)] - f[c];
    r[vidx(blk, c, l)] = (fl & (2 << c)) ? T(0) : v;
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_st_gs(int nb, int colour, const int* __restrict__ nbr,
                                               const uint8_t* __restrict__ flags, const T* __restrict__ ST, T* u,
                                               const T* __restrict__ b, const int* __restrict__ list, int nlist) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  int blk, l;
  if (list) {
    if (gid >= nlist) return;
    const int s = list[gid];
    blk = s >> 6;
    l = s & 63;
  } else {
    if (gid >= nb * 8) return;
    blk = gid >> 3;
    l = colour_node(colour, gid & 7);
  }
  const uint8_t fl = flags[blk * 64 + l];
  if (fl & 1) {
    T f[3];
    row_st(nbr, ST, u, blk, l, f);
    const int e = 13 * 9;
    T d[6] = {ST[sidx(blk, e + 0, l)], ST[sidx(blk, e + 4, l)], ST[sidx(blk, e + 8, l)],
              ST[sidx(blk, e + 1, l)], ST[sidx(blk, e + 2, l)], ST[sidx(blk, e + 5, l)]};
    dirichlet_inverse(d, fl);
    const T r0 = b[vidx(blk, 0, l)] - f[0], r1 = b[vidx(blk, 1, l)] - f[1], r2 = b[vidx(blk, 2, l)] - f[2];
    const T u0 = u[vidx(blk, 0, l)] + d[0] * r0 + d[3] * r1 + d[4] * r2;
    const T u1 = u[vidx(blk, 1, l)] + d[3] * r0 + d[1] * r1 + d[5] * r2;
    const T u2 = u[vidx(blk, 2, l)] + d[4] * r0 + d[5] * r1 + d[2] * r2;
    u[vidx(blk, 0, l)] = (fl & 2) ? T(0) : u0;
    u[vidx(blk, 1, l)] = (fl & 4) ? T(0) : u1;
    u[vidx(blk, 2, l)] = (fl & 8) ? T(0) : u2;
  } else if (fl & 0x0E) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (fl & (2 << c)) u[vidx(blk, c, l)] = T(0);
  }
}

// ----------------------------------------------------------------- transfers
// R = P^T, full weighting {1/2,1,1/2}^3 around 2c-B (Restriction.h:52-92)
template <class T>
__global__ void __launch_bounds__(TPB) k_restrict(int nbc, const int* __restrict__ rsrc,
                                                  const uint8_t* __restrict__ cflags, const T* __restrict__ fine,
                                                  T* __restrict__ coarse) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nbc * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int dk = -1; dk <= 1; ++dk) {
        const int x = 2 * li + di, y = 2 * lj + dj, z = 2 * lk + dk;
        const int fb = rsrc[blk * 27 + ((x >> 2) + 1) * 9 + ((y >> 2) + 1) * 3 + ((z >> 2) + 1)];
        if (fb < 0) continue;
        const T w = T((di == 0 ? 1.0 : 0.5) * (dj == 0 ? 1.0 : 0.5) * (dk == 0 ? 1.0 : 0.5));
        const int q = loc3(x, y, z);
        a0 += w * fine[vidx(fb, 0, q)];
        a1 += w * fine[vidx(fb, 1, q)];
        a2 += w * fine[vidx(fb, 2, q)];
      }
  const uint8_t fl = cflags[gid];
  coarse[vidx(blk, 0, l)] = (fl & 2) ? T(0) : a0;
  coarse[vidx(blk, 1, l)] = (fl & 4) ? T(0) : a1;
  coarse[vidx(blk, 2, l)] = (fl & 8) ? T(0) : a2;
}

// fine += P coarse, then Dirichlet and null-space projection
// (Prolongation.h:57-99, Multigrid_Hierarchy.h:600-603)
template <class T>
__global__ void __launch_bounds__(TPB) k_prolong(int nbf, const int* __restrict__ psrc, const int* __restrict__ fbase,
                                                 const uint8_t* __restrict__ fflags, const T* __restrict__ coarse,
                                                 T* __restrict__ fine) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nbf * 64) return;
  const int blk = gid >> 6, l = gid & 63;
  const int li = l >> 4, lj = (l >> 2) & 3, lk = l & 3;
  const int phx = ((fbase[3 * blk] + 4) >> 1) & 3, phy = ((fbase[3 * blk + 1] + 4) >> 1) & 3,
            phz = ((fbase[3 * blk + 2] + 8) >> 1) & 3;
  const uint8_t fl = fflags[gid];
  T a0 = fine[vidx(blk, 0, l)], a1 = fine[vidx(blk, 1, l)], a2 = fine[vidx(blk, 2, l)];
#pragma unroll
  for (int oi = 0; oi < 2; ++oi)
#pragma unroll
    for (int oj = 0; oj < 2; ++oj)
#pragma unroll
      for (int ok = 0; ok < 2; ++ok) {
        const T wi = (li & 1) ? T(0.5) : (oi ? T(0) : T(1));
        const T wj = (lj & 1) ? T(0.5) : (oj ? T(0) : T(1));
        const T wk = (lk & 1) ? T(0.5) : (ok ? T(0) : T(1));
        const T w = wi * wj * wk;
        if (w == T(0)) continue;
        const int x = phx + (li >> 1) + oi, y = phy + (lj >> 1) + oj, z = phz + (lk >> 1) + ok;
        const int cb = psrc[blk * 8 + ((x >> 2) << 2) + ((y >> 2) << 1) + (z >> 2)];
        if (cb < 0) continue;
        const int q = loc3(x, y, z);
        a0 += w * coarse[vidx(cb, 0, q)];
        a1 += w * coarse[vidx(cb, 1, q)];
        a2 += w * coarse[vidx(cb, 2, q)];
      }
  const bool act = fl & 1;
  fine[vidx(blk, 0, l)] = (!act || (fl & 2)) ? T(0) : a0;
  fine[vidx(blk, 1, l)] = (!act || (fl & 4)) ? T(0) : a1;
  fine[vidx(blk, 2, l)] = (!act || (fl & 8)) ? T(0) : a2;
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
template <int OP>
__global__ void __launch_bounds__(RED_T) k_final(const double* part, int np, double* out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s = OP == 0 ? s + part[i] : fmax(s, part[i]);
  s = block_reduce<OP>(s);
  if (threadIdx.x == 0) *out = s;
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
__global__ void __launch_bounds__(RED_T) k_cg_update(T* x, T* r, const T* s, const T* q, const CgScalars* sc, int64_t n,
                                                      double* part) {
  const T alpha = (T)sc->alpha;
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = alpha * s[i] + x[i];
    const T rv = -alpha * q[i] + r[i];
    r[i] = rv;
    m = fmax(m, fabs((double)rv));
  }
  m = block_reduce<1>(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}
__global__ void k_cg_scalars(CgScalars* sc, int stage) {
  if (threadIdx.x != 0) return;
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
        ST[sidx(blk, s * 9 + v * 3 + w, l)] = val;
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
      for (int w = 0; w < 3; ++w) A[(int64_t)(vi + v) + (int64_t)(vj + w) * n] = ST[sidx(blk, s * 9 + v * 3 + w, l)];
  }
}
__global__ void k_symmetrize_lower(double* A, int64_t n) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * n) return;
  const int64_t i = gid % n, j = gid / n;  // A[i + j n]
  if (i < j) A[i + j * n] = A[j + i * n];
}
template <class T>
__global__ void k_bottom_gather(int nb, const int* __restrict__ var_of, const T* __restrict__ b, double* x) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int vi = var_of[gid];
  if (vi < 0) return;
  const int blk = gid >> 6, l = gid & 63;
  for (int c = 0; c < 3; ++c) x[vi + c] = (double)b[vidx(blk, c, l)];
}
template <class T>
__global__ void k_bottom_scatter(int nb, const int* __restrict__ var_of, const double* y, const T* __restrict__ b,
                                 T* u) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * 64) return;
  const int vi = var_of[gid];
  const int blk = gid >> 6, l = gid & 63;
  for (int c = 0; c < 3; ++c) u[vidx(blk, c, l)] = vi >= 0 ? (T)y[vi + c] : b[vidx(blk, c, l)];
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
__global__ void k_gather_input_scalar(const int* in_slot, int64_t nin, const double* v, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin) return;
  const int s = in_slot[i];
  out[i] = s >= 0 ? v[s] : 0.0;
}
__global__ void k_set_entries(double* v, const int* idx, const double* val, int64_t n) {
  // serial order (last write wins) across duplicates: one thread
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int64_t i = 0; i < n; ++i) {
    const int s = idx[i] / 3, c = idx[i] % 3;
    v[vidx(s >> 6, c, s & 63)] = val[i];
  }
}

// ======================================================== launch wrappers
template <class T>
void launch_apply(hipStream_t st, const LevelDev& L, const T* E, T dx, const T* in, const T* b, T* out, int mode) {
  if (mode == 0)
    k_apply<T, 0><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, b, out);
  else if (mode == 2)
    k_apply<T, 2><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, b, out);
  else
    k_apply<T, 1><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.nbr, L.flags, E, dx, in, b, out);
}
template <class T>
void launch_gs0(hipStream_t st, const LevelDev& L, const T* E, T dx, T* u, const T* b, int colour, bool boundary) {
  if (boundary) {
    const int n = L.nbnd[colour];
    if (n == 0) return;
    k_gs0<T><<<nblk(n), TPB, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b, L.bnd[colour], n);
  } else {
    k_gs0<T><<<nblk((int64_t)L.nb * 8), TPB, 0, st>>>(L.nb, colour, L.nbr, L.flags, E, dx, u, b, nullptr, 0);
  }
}
template <class T>
void launch_st_residual(hipStream_t st, const LevelDev& L, const T* ST, const T* u, const T* b, T* r) {
  k_st_residual<T><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, L.nbr, L.flags, ST, u, b, r);
}
template <class T>
void launch_st_gs(hipStream_t st, const LevelDev& L, const T* ST, T* u, const T* b, int colour, bool boundary) {
  if (boundary) {
    const int n = L.nbnd[colour];
    if (n == 0) return;
    k_st_gs<T><<<nblk(n), TPB, 0, st>>>(L.nb, colour, L.nbr, L.flags, ST, u, b, L.bnd[colour], n);
  } else {
    k_st_gs<T><<<nblk((int64_t)L.nb * 8), TPB, 0, st>>>(L.nb, colour, L.nbr, L.flags, ST, u, b, nullptr, 0);
  }
}
template <class T>
void launch_restrict(hipStream_t st, const LevelDev& C, const T* fine, T* coarse) {
  k_restrict<T><<<nblk((int64_t)C.nb * 64), TPB, 0, st>>>(C.nb, C.rsrc, C.flags, fine, coarse);
}
template <class T>
void launch_prolongate(hipStream_t st, const LevelDev& F, const T* coarse, T* fine) {
  k_prolong<T><<<nblk((int64_t)F.nb * 64), TPB, 0, st>>>(F.nb, F.psrc, F.base, F.flags, coarse, fine);
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
void launch_cg_update(hipStream_t st, T* x, T* r, const T* s, const T* q, const CgScalars* sc, int64_t n) {
  // caller passes the partial buffer right after the scalars block
  double* part = reinterpret_cast<double*>(const_cast<CgScalars*>(sc) + 1);
  const unsigned g = red_grid(n);
  k_cg_update<T><<<g, RED_T, 0, st>>>(x, r, s, q, sc, n, part);
  k_final<1><<<1, RED_T, 0, st>>>(part, (int)g, &const_cast<CgScalars*>(sc)->nrm);
}
void launch_cg_scalars(hipStream_t st, CgScalars* sc, int stage) { k_cg_scalars<<<1, 64, 0, st>>>(sc, stage); }
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
void launch_symmetrize_lower(hipStream_t st, double* A, int64_t n) {
  k_symmetrize_lower<<<nblk(n * n), TPB, 0, st>>>(A, n);
}
template <class T>
void launch_bottom_gather(hipStream_t st, const LevelDev& L, const int* var_of, const T* b, double* x) {
  k_bottom_gather<T><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, var_of, b, x);
}
template <class T>
void launch_bottom_scatter(hipStream_t st, const LevelDev& L, const int* var_of, const double* y, const T* b, T* u) {
  k_bottom_scatter<T><<<nblk((int64_t)L.nb * 64), TPB, 0, st>>>(L.nb, var_of, y, b, u);
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
void launch_gather_input_scalar(hipStream_t st, const int* in_slot, int64_t nin, const double* v, double* out) {
  if (nin > 0) k_gather_input_scalar<<<nblk(nin), TPB, 0, st>>>(in_slot, nin, v, out);
}
void launch_set_entries(hipStream_t st, double* v, const int* idx, const double* val, int64_t n) {
  if (n > 0) k_set_entries<<<1, 64, 0, st>>>(v, idx, val, n);
}

#define SPG_INST(T)                                                                                              \
  template void launch_apply<T>(hipStream_t, const LevelDev&, const T*, T, const T*, const T*, T*, int);         \
  template void launch_gs0<T>(hipStream_t, const LevelDev&, const T*, T, T*, const T*, int, bool);               \
  template void launch_st_residual<T>(hipStream_t, const LevelDev&, const T*, const T*, const T*, T*);           \
  template void launch_st_gs<T>(hipStream_t, const LevelDev&, const T*, T*, const T*, int, bool);                \
  template void launch_restrict<T>(hipStream_t, const LevelDev&, const T*, T*);                                  \
  template void launch_prolongate<T>(hipStream_t, const LevelDev&, const T*, T*);                                \
  template void launch_zero<T>(hipStream_t, T*, int64_t);                                                        \
  template void launch_copy<T>(hipStream_t, const T*, T*, int64_t);                                              \
  template void launch_axpy_d<T>(hipStream_t, double*, const T*, int64_t);                                       \
  template void launch_dot<T>(hipStream_t, const T*, const T*, int64_t, double*, double*);                       \
  template void launch_linf<T>(hipStream_t, const T*, int64_t, double*, double*);                                \
  template void launch_cg_direction<T>(hipStream_t, T*, const T*, const CgScalars*, int64_t);                    \
  template void launch_cg_update<T>(hipStream_t, T*, T*, const T*, const T*, const CgScalars*, int64_t);         \
  template void launch_project<T>(hipStream_t, const LevelDev&, T*, bool);                                       \
  template void launch_bottom_gather<T>(hipStream_t, const LevelDev&, const int*, const T*, double*);            \
  template void launch_bottom_scatter<T>(hipStream_t, const LevelDev&, const int*, const double*, const T*, T*);
SPG_INST(float)
SPG_INST(double)

}  // namespace spg
