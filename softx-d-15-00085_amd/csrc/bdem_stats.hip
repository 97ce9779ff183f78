/*
 * bdem_stats.hip - ensemble statistics of the current state, on the device.
 *
 * The reference computes these on the host from a full state download:
 * the particles left above the hatch (Silo_Stats, main.cpp:253-285), and the
 * mean and largest speed in m/s with the P_ID of the fastest particle
 * (Output_SystemState, Input/FileIO.h:286-318).  Here one grid-stride
 * kernel reduces them, plus the solid fraction of an axis-aligned box (no
 * reference counterpart), into a few words: per wave by shuffles, then one
 * atomic per wave.  Sums are in double, so the result does not depend on
 * the order the waves finish in except in the last bits.
 */
#include <mutex>
#include "bdem_kernels.hpp"

namespace bdem {

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_count(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_stats(int n, const float4* __restrict__ pos_r, const float4* __restrict__ vel_m,
                                               const float4* __restrict__ omg_t, float min_y, float3 lo, float3 hi,
                                               int want_volume, StatsWords* __restrict__ out) {
  uint32_t cnt = 0u, bcnt = 0u;
  double sum_v = 0.0, vol = 0.0;
  unsigned long long best = 0ull;   /* (speed bits << 32) | (P_ID + 1); 0: none */
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = pos_r[i], v = vel_m[i], o = omg_t[i];
    const int pid = packed_pid(o.w);
    if (pid < 0) continue;
    if (p.y > min_y) {
      /* length(v * 0.010) as the reference's host code writes it */
      const float sx = v.x * 0.010f, sy = v.y * 0.010f, sz = v.z * 0.010f;
      const float sp = __builtin_sqrtf(sx * sx + sy * sy + sz * sz);
      ++cnt;
      sum_v += (double)sp;
      const unsigned long long key = ((unsigned long long)__float_as_uint(sp) << 32) | (uint32_t)(pid + 1);
      best = key > best ? key : best;
    }
    if (p.x >= lo.x && p.x < hi.x && p.y >= lo.y && p.y < hi.y && p.z >= lo.z && p.z < hi.z) {
      ++bcnt;
      if (want_volume) {
        const double r = (double)p.w;
        vol += (4.0 / 3.0) * 3.14159265358979323846 * r * r * r;
      }
    }
  }
  cnt = wave_sum_count(cnt);
  bcnt = wave_sum_count(bcnt);
  sum_v = wave_sum_f64(sum_v);
  vol = wave_sum_f64(vol);
  best = wave_max_key(best);
  if ((threadIdx.x & 63) == 0) {
    if (cnt) atomicAdd(&out->count, cnt);
    if (bcnt) atomicAdd(&out->box_count, bcnt);
    if (sum_v != 0.0) atomicAdd(&out->sum_speed, sum_v);
    if (vol != 0.0) atomicAdd(&out->box_volume, vol);
    if (best) atomicMax(&out->best, best);
  }
}

void launch_stats(const StepBuffers& b, float min_y, float3 lo, float3 hi, bool want_volume, StatsWords* out,
                  hipStream_t s) {
  const int blocks = b.n <= 0 ? 1 : min((b.n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_stats, dim3(blocks), dim3(256), 0, s, b.n, b.pos_cur, b.vel_cur, b.omg_cur, min_y, lo, hi,
                     want_volume ? 1 : 0, out);
}

/* Free-surface profile: over bins of width 1/inv_w along `axis` (0 x, 2 z)
 * from lo, the highest particle top y + r of each bin (r: the bounding
 * radius for polyhedra), as an order-preserving key of the float (empty: 0).
 * Each block keeps its maxima in LDS and merges the bins it touched with
 * one global atomicMax each; the particles arrive in cell order, so a block
 * touches few bins. */
__device__ __forceinline__ uint32_t order_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_profile(int n, const float4* __restrict__ pos_r, const float4* __restrict__ omg_t,
                                                 int axis, float lo, float inv_w, int bins, uint32_t* __restrict__ keys) {
  __shared__ uint32_t sh[PROFILE_MAX_BINS];
  for (int k = threadIdx.x; k < bins; k += blockDim.x) sh[k] = 0u;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = pos_r[i];
    if (packed_pid(omg_t[i].w) < 0) continue;
    const float t = ((axis == 0 ? p.x : p.z) - lo) * inv_w;
    if (!(t >= 0.0f && t < (float)bins)) continue;
    const int k = min((int)t, bins - 1);
    atomicMax(&sh[k], order_key(p.y + p.w));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < bins; k += blockDim.x)
    if (sh[k]) atomicMax(&keys[k], sh[k]);
}

void launch_profile(const StepBuffers& b, int axis, float lo, float inv_w, int bins, uint32_t* keys, hipStream_t s) {
  const int blocks = b.n <= 0 ? 1 : min((b.n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_profile, dim3(blocks), dim3(256), 0, s, b.n, b.pos_cur, b.omg_cur, axis, lo, inv_w, bins, keys);
}

/* ------------------------------------------------------------------------ */
/* Allocation-free profile: the bin keys and the finished heights live in a  */
/* file-scope device array, one copy per GPU, its address taken once per     */
/* device with hipGetSymbolAddress.  Every call on a GPU shares that GPU's   */
/* array (calls that overlap on one GPU race); each call clears the keys it  */
/* uses with a kernel first.                                                 */
/* ------------------------------------------------------------------------ */
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch(int device, uint32_t** out) {
  *out = nullptr;
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;   /* the caller makes `device` current */
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

/* keys -> heights (NaN for an empty bin), the inverse of order_key */
__global__ __launch_bounds__(256) void k_profile_finish(const uint32_t* __restrict__ keys, int bins,
                                                        float* __restrict__ top) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= bins) return;
  const uint32_t u = keys[k];
  top[k] = u == 0u ? __uint_as_float(0x7FC00000u) : __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

void launch_profile_noalloc(const StepBuffers& b, int axis, float lo, float inv_w, int bins, uint32_t* scr,
                            hipStream_t s) {
  launch_fill_u32(scr, (size_t)bins, 0u, s);
  launch_profile(b, axis, lo, inv_w, bins, scr, s);
  hipLaunchKernelGGL(k_profile_finish, dim3((bins + 255) / 256), dim3(256), 0, s, scr, bins,
                     reinterpret_cast<float*>(scr + PROFILE_MAX_BINS));
}

}  // namespace bdem
