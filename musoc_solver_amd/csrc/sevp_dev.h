// Device-side helpers shared by the SEVP kernels (gfx950, wave64).
//
// Arithmetic rule for every kernel that claims parity with the reference:
// the library is compiled with -ffp-contract=off, every expression is written
// in the reference's left-to-right order, float coefficients are promoted to
// double exactly, and every dot product is a sequential sum from zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SEVP_WAVE 64

// Column-major (Fortran) 2-D index with 1-based i, j.
__device__ __forceinline__ long f2(int ld, int i, int j) {
  return (long)(i - 1) + (long)(j - 1) * (long)ld;
}

// Lane tiles of an (m, k) operand of the 16x16x4 MFMA steps: block
// (m / 16, k / 16) of 256 elements holds, at lane l = m % 16 + 16 q and slot
// t, the element (m, 16 (k / 16) + 4 q + t), so one wavefront's operand for
// 16 rows and 16 k is one contiguous run.  nchk = ceil(columns / 16).
__device__ __forceinline__ long lt_index(int m, int k, int nchk) {
  return ((long)((m >> 4) * nchk + (k >> 4)) * 64 + (m & 15) + 16 * ((k & 15) >> 2)) * 4 + (k & 3);
}

// Lane i receives lane i-1's value (lane 0 receives `fill`): DPP wave_shr:1.
__device__ __forceinline__ double wave_shr1(double v, double fill) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  int flo = __double2loint(fill), fhi = __double2hiint(fill);
  lo = __builtin_amdgcn_update_dpp(flo, lo, 0x138, 0xf, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(fhi, hi, 0x138, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Lane i receives lane i+1's value (lane 63 receives `fill`): DPP wave_shl:1.
__device__ __forceinline__ double wave_shl1(double v, double fill) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  int flo = __double2loint(fill), fhi = __double2hiint(fill);
  lo = __builtin_amdgcn_update_dpp(flo, lo, 0x130, 0xf, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(fhi, hi, 0x130, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Lane i receives lane i-1's value, lane 0 receives +0 (DPP wave_shr:1 with
// bound_ctrl): no fill operand, so no copy of one.
__device__ __forceinline__ double wave_shr1_zero(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Lane i receives lane i+1's value, lane 63 receives +0 (DPP wave_shl:1).
__device__ __forceinline__ double wave_shl1_zero(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Broadcast lane `src` (wave-uniform) of a double.
__device__ __forceinline__ double read_lane(double v, int src) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Kernel span from the kernel's own clock (profiling only, t null otherwise):
// workgroup w stores the 100 MHz real-time counter at its entry into t[2w]
// and, after its memory operations have completed, at its exit into
// t[2w + 1] (plain stores, one slot pair per workgroup, so timing adds no
// contention).  The launch's span is max(t[2w + 1]) - min(t[2w]).
__device__ __forceinline__ long span_wg() {
  return blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
}
__device__ __forceinline__ void span_begin(unsigned long long *t) {
  if (t && threadIdx.x == 0) t[2 * span_wg()] = __builtin_amdgcn_s_memrealtime();
}
__device__ __forceinline__ void span_end(unsigned long long *t) {
  if (!t) return;
  __syncthreads();
  if (threadIdx.x == 0) t[2 * span_wg() + 1] = __builtin_amdgcn_s_memrealtime();
}

// Grouped launches: one launch carries the same step of many independent
// problems.  Entry e of a launch owns workgroups start[e] .. start[e+1]-1
// (start[0] = 0, increasing); a workgroup's entry is the number of starts at
// or below its index, minus one, counted by wavefront ballots 64 starts at a
// time (one round of vector loads for up to 256 entries, then bit counts).
__device__ __forceinline__ int group_entry(const int *start, int n, int b) {
  const int lane = threadIdx.x & 63;
  int cnt = 0;
  for (int base = 0; base < n; base += 256) {
    int v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = base + 64 * u + lane;
      v[u] = i < n ? start[i] : 0x7fffffff;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) cnt += __popcll(__ballot(v[u] <= b));
  }
  return cnt - 1;
}
