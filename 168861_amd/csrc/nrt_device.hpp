// Device helpers for gfx950 (wave64): DPP/cross-lane reductions, block
// reductions, and a bounded grid barrier for persistent kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace nrt {

constexpr int kWave = 64;

// 64-bit DPP move (two 32-bit halves). CTRL is a gfx9 DPP control word.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ double readlane_f64(double x, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_f32(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// Full-wave sum, uniform result. Within each row of 16 lanes: quad_perm xor1,
// quad_perm xor2, row_half_mirror, row_mirror (every lane of a row ends with the
// bit-identical row sum); the four row sums are combined through readlane in a
// fixed order. Requires all 64 lanes active.
__device__ __forceinline__ double wave_sum(double x) {
  x = x + dpp_f64<0xB1>(x);   // quad_perm(1,0,3,2)
  x = x + dpp_f64<0x4E>(x);   // quad_perm(2,3,0,1)
  x = x + dpp_f64<0x141>(x);  // row_half_mirror
  x = x + dpp_f64<0x140>(x);  // row_mirror
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}
__device__ __forceinline__ float wave_sum(float x) {
  x = x + dpp_f32<0xB1>(x);
  x = x + dpp_f32<0x4E>(x);
  x = x + dpp_f32<0x141>(x);
  x = x + dpp_f32<0x140>(x);
  return (readlane_f32(x, 0) + readlane_f32(x, 16)) + (readlane_f32(x, 32) + readlane_f32(x, 48));
}

// Full-wave maximum, uniform result (the reduction pattern of wave_sum).
__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp_f64<0xB1>(x));
  x = fmax(x, dpp_f64<0x4E>(x));
  x = fmax(x, dpp_f64<0x141>(x));
  x = fmax(x, dpp_f64<0x140>(x));
  return fmax(fmax(readlane_f64(x, 0), readlane_f64(x, 16)),
              fmax(readlane_f64(x, 32), readlane_f64(x, 48)));
}

// Error-free transformation a + b = s + e (Knuth two-sum).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// Compensated (s, c) pair sum of a wave: every lane's (s, c) is folded with
// two-sum so the result carries about twice the working precision.
__device__ __forceinline__ void wave_sum_comp(float& s, float& c) {
#define NRT_COMP_STEP(CTRL)                                  \
  {                                                          \
    float so = dpp_f32<CTRL>(s), co = dpp_f32<CTRL>(c), t, e; \
    two_sum(s, so, t, e);                                    \
    s = t;                                                   \
    c = c + co + e;                                          \
  }
  NRT_COMP_STEP(0xB1)
  NRT_COMP_STEP(0x4E)
  NRT_COMP_STEP(0x141)
  NRT_COMP_STEP(0x140)
#undef NRT_COMP_STEP
  float s0 = readlane_f32(s, 0), s1 = readlane_f32(s, 16), s2 = readlane_f32(s, 32),
        s3 = readlane_f32(s, 48);
  float c0 = readlane_f32(c, 0), c1 = readlane_f32(c, 16), c2 = readlane_f32(c, 32),
        c3 = readlane_f32(c, 48);
  float t, e, cc = (c0 + c1) + (c2 + c3);
  two_sum(s0, s1, t, e);
  cc = cc + e;
  two_sum(t, s2, t, e);
  cc = cc + e;
  two_sum(t, s3, t, e);
  cc = cc + e;
  s = t;
  c = cc;
}

// Sum of NV values over the block (blockDim.x a multiple of 64, <= 1024); the
// result is returned to every thread. lds must hold NV*16 doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) lds[k * 16 + wid] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0;
    for (int w = 0; w < nw; w++) s = s + lds[k * 16 + w];
    v[k] = s;
  }
  __syncthreads();
}

// Maximum of NV non-negative values over the block, as block_sum.
template <int NV>
__device__ __forceinline__ void block_max(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_max(v[k]);
    if (lane == 0) lds[k * 16 + wid] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0;
    for (int w = 0; w < nw; w++) s = fmax(s, lds[k * 16 + w]);
    v[k] = s;
  }
  __syncthreads();
}

// Orders LDS accesses between the lanes of one wave: the LDS executes a wave's
// operations in issue order, so keeping the compiler from moving accesses across
// this point (a wavefront-scope fence, which emits no instruction) is enough.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Grid barrier for a persistent launch (all blocks co-resident). bar[0] counts
// arrivals, bar[1] is the generation. The wait is bounded: after spin_limit
// polls, or once another block raised *err, it sets *err and returns so that
// every block drains; callers test *err after each barrier.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned nblocks, unsigned* err,
                                             unsigned long long spin_limit) {
  // The workgroup barrier orders every wave's stores before thread 0's agent-scope
  // release (fences are cumulative), so one fence per workgroup publishes them.
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned g = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned prev =
        __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == nblocks - 1) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar[1], g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long spins = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > spin_limit ||
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
          __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Deterministic sum of per-block partials part[b*stride + k], b < nblocks, k < NV,
// identical bits in every block (fixed lane mapping and reduction tree).
template <int NV>
__device__ __forceinline__ void grid_partials_sum(const double* part, int stride, int nblocks,
                                                  double (&out)[NV], double* lds) {
  if (threadIdx.x < kWave) {
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0;
    for (int b = threadIdx.x; b < nblocks; b += kWave) {
#pragma unroll
      for (int k = 0; k < NV; k++) acc[k] = acc[k] + part[b * stride + k];
    }
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = wave_sum(acc[k]);
      if (threadIdx.x == 0) lds[k] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) out[k] = lds[k];
  __syncthreads();
}

// Maximum of non-negative per-block partials, as grid_partials_sum.
template <int NV>
__device__ __forceinline__ void grid_partials_max(const double* part, int stride, int nblocks,
                                                  double (&out)[NV], double* lds) {
  if (threadIdx.x < kWave) {
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0;
    for (int b = threadIdx.x; b < nblocks; b += kWave) {
#pragma unroll
      for (int k = 0; k < NV; k++) acc[k] = fmax(acc[k], part[b * stride + k]);
    }
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = wave_max(acc[k]);
      if (threadIdx.x == 0) lds[k] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) out[k] = lds[k];
  __syncthreads();
}

}  // namespace nrt
