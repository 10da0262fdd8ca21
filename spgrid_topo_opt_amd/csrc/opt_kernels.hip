// Optimizer-side kernels next to the solve: the sensitivity filter and the
// optimality-criterion (OC) update, over the same 4x4x4 cell blocks.
//
//   filter:  topo_opt/spgrid_topo_opt.cpp:527-574 (smooth_dc)
//   OC:      topo_opt/optimizer/optimality_criterion.cpp:19-146
//
// Cell index block*64 + i*16 + j*4 + k; a block's 27 neighbours are given as
// nbr[block*27 + s], s = (dx+1)*9 + (dy+1)*3 + (dz+1), -1 where absent.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "spgrid_topo_opt_amd.h"

namespace spg {
void set_last_error(const std::string& msg);
}

namespace {

constexpr int FW = 4;        // blocks (wavefronts) per workgroup in the filter
constexpr int MAX_HALF = 3;  // filter radius floor(r) supported (halo within the 27 neighbours)
constexpr int RT = 256;      // reduction workgroup size

__device__ inline int side(int x) { return x < 0 ? 0 : (x > 3 ? 2 : 1); }

// out(x) = sum_o w(o) v(x+o) / (sum_o w(o) m(x+o) (rho(x) + 1e-3)) with
// v = dc rho [inside], m = [inside], w(o) = max(0, r - |o|), |o_i| <= h.
// Cells outside the container take the same formula (smooth_dc writes every
// cell of the fat blocks); where no cell of the window is inside, the
// reference's 0/0 is written as 0.  The halo (4+2h)^3 of v and m is staged
// in LDS per block.
template <int H>
__global__ void __launch_bounds__(64 * FW) k_filter(int64_t nb, const int* __restrict__ nbr,
                                                     const uint8_t* __restrict__ inside, const double* __restrict__ dc,
                                                     const double* __restrict__ rho, double r, double* __restrict__ out) {
  constexpr int S = 4 + 2 * H, S3 = S * S * S;
  __shared__ double hv[FW][S3], hm[FW][S3];
  __shared__ int ht[FW][28];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t blk = (int64_t)blockIdx.x * FW + w;
  const bool live = blk < nb;
  if (lane < 27) ht[w][lane] = !live ? -1 : (lane == 13 ? (int)blk : nbr[blk * 27 + lane]);
  __syncthreads();
  for (int h = lane; h < S3; h += 64) {
    const int x = h / (S * S) - H, y = (h / S) % S - H, z = h % S - H;
    const int bb = ht[w][side(x) * 9 + side(y) * 3 + side(z)];
    double v = 0.0, m = 0.0;
    if (bb >= 0) {
      const int64_t c = (int64_t)bb * 64 + ((x & 3) << 4) + ((y & 3) << 2) + (z & 3);
      if (inside[c]) {
        v = dc[c] * rho[c];
        m = 1.0;
      }
    }
    hv[w][h] = v;
    hm[w][h] = m;
  }
  __syncthreads();
  if (!live) return;
  const int i = lane >> 4, j = (lane >> 2) & 3, k = lane & 3;
  const int64_t c = blk * 64 + lane;
  double num = 0.0, den = 0.0;
  for (int a = -H; a <= H; ++a)
    for (int b = -H; b <= H; ++b)
      for (int e = -H; e <= H; ++e) {
        const double wt = r - sqrt((double)(a * a + b * b + e * e));
        if (!(wt > 0.0)) continue;
        const int h = ((i + a + H) * S + (j + b + H)) * S + (k + e + H);
        num += wt * hv[w][h];
        den += wt * hm[w][h];
      }
  out[c] = den > 0.0 ? num / (den * (rho[c] + 1e-3)) : 0.0;
}

// mask bits of the OC kernels
constexpr uint8_t OC_INSIDE = 1;  // cell inside the container: updated
constexpr uint8_t OC_FIXED = 2;   // fixed-density cell (flags.get_fixed_density())

// new density of one cell at multiplier mid (optimality_criterion.cpp:84-101);
// shift = min(0, min over the container of -dc) with smart_oc, else 0
__device__ inline double oc_new(double rho, double dc, double mid, double step, double dmin, double shift) {
  double v = rho * sqrt(fmax(0.0, -dc - shift) / mid);
  v = fmin(fmax(v, rho - step), rho + step);
  return fmin(fmax(v, dmin), 1.0);
}

__device__ inline double block_sum(double s) {
  __shared__ double red[RT / 64];
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < RT / 64; ++q) t += red[q];
  return t;
}
__device__ inline double block_max(double s) {
  __shared__ double red[RT / 64];
  for (int o = 32; o >= 1; o >>= 1) s = fmax(s, __shfl_xor(s, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < RT / 64; ++q) t = fmax(t, red[q]);
  return t;
}

struct OcParams {
  double mid, step, dmin, fixed_density, shift;
  int exclude_fixed;
};

// per-workgroup partial volume (and max change, with the new densities
// written when out is given); the partials are added by k_oc_final in a
// fixed order, so a trial is deterministic.  Fixed cells take the fixed
// density and leave the volume sum when exclude_fixed is set
// (optimality_criterion.cpp:96-108).
__global__ void __launch_bounds__(RT) k_oc_part(int64_t n, const uint8_t* __restrict__ mask,
                                                const double* __restrict__ rho, const double* __restrict__ dc,
                                                OcParams q, double* __restrict__ out, double* __restrict__ part) {
  double s = 0.0, ch = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const double r = rho[i];
    const uint8_t m = mask[i];
    double v = r;
    if (m & OC_INSIDE) {
      const bool fixed = (m & OC_FIXED) != 0;
      v = fixed ? q.fixed_density : oc_new(r, dc[i], q.mid, q.step, q.dmin, q.shift);
      if (!(fixed && q.exclude_fixed)) s += v;
      ch = fmax(ch, fabs(v - r));
    }
    if (out) out[i] = v;
  }
  const double bs = block_sum(s);
  __syncthreads();
  const double bc = block_max(ch);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = bs;
    part[gridDim.x + blockIdx.x] = bc;
  }
}
__global__ void __launch_bounds__(RT) k_oc_final(int np, const double* __restrict__ part, double* __restrict__ res) {
  double s = 0.0, c = 0.0;
  for (int q = threadIdx.x; q < np; q += RT) {
    s += part[q];
    c = fmax(c, part[np + q]);
  }
  const double bs = block_sum(s);
  __syncthreads();
  const double bc = block_max(c);
  if (threadIdx.x == 0) {
    res[0] = bs;
    res[1] = bc;
  }
}

// per-device scratch of the OC reduction, allocated on first use
struct OcScratch {
  double* part = nullptr;
  double* res = nullptr;
  double* host = nullptr;
  int np = 0;
};
std::mutex g_oc_mu;
std::vector<OcScratch> g_oc;

int fail(int code, const std::string& msg) {
  spg::set_last_error(msg);
  return code;
}

}  // namespace

extern "C" int spg_filter_sensitivity(int64_t nblocks, const int32_t* nbr, const uint8_t* inside, const double* dc,
                                      const double* rho, double filter_size, double* out) {
  if (nblocks <= 0) return SPG_OK;
  if (!nbr || !inside || !dc || !rho || !out) return fail(SPG_E_INVALID, "null argument");
  const int h = (int)std::floor(filter_size);
  if (!(filter_size > 0) || h > MAX_HALF)
    return fail(SPG_E_INVALID, "filter_size must be in (0, " + std::to_string(MAX_HALF + 1) + ")");
  const unsigned g = (unsigned)((nblocks + FW - 1) / FW);
  switch (h) {
    case 0: k_filter<0><<<g, 64 * FW>>>(nblocks, nbr, inside, dc, rho, filter_size, out); break;
    case 1: k_filter<1><<<g, 64 * FW>>>(nblocks, nbr, inside, dc, rho, filter_size, out); break;
    case 2: k_filter<2><<<g, 64 * FW>>>(nblocks, nbr, inside, dc, rho, filter_size, out); break;
    default: k_filter<3><<<g, 64 * FW>>>(nblocks, nbr, inside, dc, rho, filter_size, out); break;
  }
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(SPG_E_HIP, std::string("filter: ") + hipGetErrorString(e));
  return SPG_OK;
}

extern "C" int spg_oc_trial(int64_t n, const uint8_t* mask, const double* rho, const double* dc, double mid,
                            double step, double min_density, double fixed_density, int32_t exclude_fixed,
                            double dc_shift, double* out, double* volume, double* change) {
  if (!mask || !rho || !dc || !volume) return fail(SPG_E_INVALID, "null argument");
  if (!(mid > 0)) return fail(SPG_E_INVALID, "multiplier must be positive");
  if (n < 0) return fail(SPG_E_INVALID, "negative cell count");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(SPG_E_HIP, "no HIP device");
  std::lock_guard<std::mutex> lock(g_oc_mu);
  if ((int)g_oc.size() <= dev) g_oc.resize(dev + 1);
  OcScratch& S = g_oc[dev];
  if (!S.part) {
    S.np = 1024;
    if (hipMalloc(&S.part, sizeof(double) * 2 * S.np) != hipSuccess ||
        hipMalloc(&S.res, sizeof(double) * 2) != hipSuccess ||
        hipHostMalloc(&S.host, sizeof(double) * 2) != hipSuccess)
      return fail(SPG_E_HIP, "OC scratch allocation failed");
  }
  const int g = (int)std::min<int64_t>(S.np, std::max<int64_t>(1, (n + RT - 1) / RT));
  const OcParams q{mid, step, min_density, fixed_density, dc_shift, exclude_fixed ? 1 : 0};
  k_oc_part<<<g, RT>>>(n, mask, rho, dc, q, out, S.part);
  k_oc_final<<<1, RT>>>(g, S.part, S.res);
  hipError_t e = hipMemcpy(S.host, S.res, sizeof(double) * 2, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(SPG_E_HIP, std::string("OC trial: ") + hipGetErrorString(e));
  *volume = S.host[0];
  if (change) *change = S.host[1];
  return SPG_OK;
}

// ------------------------------------------------ OC trial without allocation
// Scratch of spg_oc_trial_noalloc: the per-workgroup (volume, change)
// partials of k_oc_part, SPG_OC_NOALLOC_PARTS pairs (16 KiB).  Every GPU the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` only names the symbol: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is passed on.
static __device__ __align__(256) double scratch[2 * SPG_OC_NOALLOC_PARTS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
double* g_na_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

extern "C" int spg_oc_trial_noalloc(int64_t n, const uint8_t* mask, const double* rho, const double* dc, double mid,
                                    double step, double min_density, double fixed_density, int32_t exclude_fixed,
                                    double dc_shift, double* out, double* result, void* stream) {
  if (!mask || !rho || !dc || !result) return fail(SPG_E_INVALID, "null argument");
  if (!(mid > 0)) return fail(SPG_E_INVALID, "multiplier must be positive");
  if (n < 0) return fail(SPG_E_INVALID, "negative cell count");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(SPG_E_HIP, "no HIP device");
  if (dev < 0 || dev >= kMaxDevices) return fail(SPG_E_INVALID, "device ordinal beyond the scratch table");
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(SPG_E_HIP, std::string("OC scratch symbol address: ") + hipGetErrorString(e));
      g_na_ptr[dev] = static_cast<double*>(p);
    }
    part = g_na_ptr[dev];
  }
  // one partial pair per workgroup; the grid-stride loop of k_oc_part lets a
  // grid of at most SPG_OC_NOALLOC_PARTS workgroups cover any n
  const int64_t want = std::max<int64_t>(1, (n + RT - 1) / RT);
  const int g = (int)std::min<int64_t>(SPG_OC_NOALLOC_PARTS, want);
  if (g > SPG_OC_NOALLOC_PARTS) return fail(SPG_E_INVALID, "OC trial needs more partials than the scratch holds");
  const OcParams q{mid, step, min_density, fixed_density, dc_shift, exclude_fixed ? 1 : 0};
  hipStream_t st = static_cast<hipStream_t>(stream);
  // k_oc_part writes every partial k_oc_final reads: nothing to clear
  k_oc_part<<<g, RT, 0, st>>>(n, mask, rho, dc, q, out, part);
  k_oc_final<<<1, RT, 0, st>>>(g, part, result);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SPG_E_HIP, std::string("OC trial launch: ") + hipGetErrorString(e));
  return SPG_OK;
}
