// Deterministic block and grid reductions for 64-wide wavefronts.
//
// A reduction kernel leaves one partial per workgroup and value in `partials`,
// then the last workgroup to finish (ticket counter) sums the partials in a
// fixed order and writes the result.  The split of rows over workgroups
// depends only on the problem size and the grid, so every run gives the same
// bits no matter how workgroups are scheduled -- no floating-point atomics.
//
// Latency matters more than bandwidth here (a few hundred partials): every
// value is reduced inside each wave with shuffles, the waves meet in LDS behind
// a single barrier, and the last workgroup loads all partials in parallel
// (L lanes per value) before a shuffle tree, so a reduction costs about two
// global-memory round trips however many values it carries.
#pragma once

#include <hip/hip_runtime.h>

#include "pk_internal.hpp"

namespace pk {

__device__ __forceinline__ double wave_sum(double v) {
   // fixed butterfly order across the 64 lanes; every lane gets the sum
   for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
   return v;
}

// Wave sums of values 0 .. nv-1 into sh[k][wave] (lane 0 writes).  Values are
// taken kWaveBatch at a time: all their get() calls first (their loads issue
// together), then their butterflies, so that a batch costs one memory round
// trip and overlapping shuffles instead of one round trip per value.  Indices
// past nv - 1 are clamped (the value is computed again and not written).
constexpr int kWaveBatch = 8;

// The first batch's values come from first(c), c = 0 .. CH - 1 a compile-time
// index standing for value min(c, nv - 1): a caller that loaded what they need
// ahead gives them from registers.
template <int MAXV, class First, class Get>
__device__ __forceinline__ void wave_sums(int nv, First first, Get get, double (*sh)[kBlock / 64]) {
   constexpr int CH = MAXV < kWaveBatch ? MAXV : kWaveBatch;
   const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
   for (int k0 = 0; k0 < nv; k0 += CH) {
      double v[CH];
      if (k0 == 0) {
#pragma unroll
         for (int c = 0; c < CH; c++) v[c] = first(c);
      } else {
#pragma unroll
         for (int c = 0; c < CH; c++) v[c] = get(min(k0 + c, nv - 1));
      }
#pragma unroll
      for (int c = 0; c < CH; c++) v[c] = wave_sum(v[c]);
      if (lane == 0) {
#pragma unroll
         for (int c = 0; c < CH; c++)
            if (k0 + c < nv) sh[k0 + c][wid] = v[c];
      }
   }
}

// Sum of one value over the whole workgroup (blockDim.x == kBlock), result in
// every thread.  `sh` must hold kBlock/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
   const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
   v = wave_sum(v);
   __syncthreads();
   if (lane == 0) sh[wid] = v;
   __syncthreads();
   double s = 0.0;
#pragma unroll
   for (int w = 0; w < kBlock / 64; w++) s += sh[w];
   return s;
}

// Grid reduction of nv <= MAXV values (runtime count).  Each thread passes its
// contribution for value k through `get(k)`.  Writes out[0..nv-1] and returns
// true, in every thread of the last workgroup only, once out[] is complete and
// visible to the whole workgroup -- the place for a serial epilogue.
// `partials` must hold gridDim.x * nv doubles.
template <int MAXV, class Get>
__device__ bool grid_reduce(int nv, Get get, double* partials, unsigned* ticket, double* out) {
   __shared__ double sh[MAXV][kBlock / 64];
   __shared__ int is_last;
   wave_sums<MAXV>(
       nv, [&](int c) { return get(min(c, nv - 1)); }, get, sh);
   __syncthreads();
   // Partials are written with system-scope stores, which go through to memory
   // instead of sitting dirty in this XCD's L2, and waited for before the
   // ticket is taken; the last workgroup reads them with system-scope loads.
   // That gives cross-XCD visibility of the partials without the agent-scope
   // fence, whose L2 writeback in every workgroup dominated the kernel time.
   if ((int)threadIdx.x < nv) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kBlock / 64; w++) s += sh[threadIdx.x][w];
      __hip_atomic_store(&partials[(long)blockIdx.x * nv + threadIdx.x], s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
   }
   __builtin_amdgcn_s_waitcnt(0);
   __syncthreads();
   if (threadIdx.x == 0)
      is_last = (__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                 gridDim.x - 1);
   __syncthreads();
   if (!is_last) return false;
   // L lanes per value (L divides 64, so a value's lanes sit in one wave)
   int L = 64;
   while (L > 1 && L * nv > kBlock) L >>= 1;
   const int k = threadIdx.x / L, sub = threadIdx.x % L;
   const int nb = gridDim.x;
   double s = 0.0;
   auto ld = [&](long p) {
      return __hip_atomic_load(&partials[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
   };
   if (k < nv) {
      int b = sub;
      for (; b + 3 * L < nb; b += 4 * L) {
         double a0 = ld((long)b * nv + k);
         double a1 = ld((long)(b + L) * nv + k);
         double a2 = ld((long)(b + 2 * L) * nv + k);
         double a3 = ld((long)(b + 3 * L) * nv + k);
         s += a0;
         s += a1;
         s += a2;
         s += a3;
      }
      for (; b < nb; b += L) s += ld((long)b * nv + k);
   }
   for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
   if (k < nv && sub == 0) out[k] = s;
   if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   __threadfence_block();
   __syncthreads();
   return true;
}

// Sum of partial rows first .. first+count-1 (entry k of row e at
// partials[e * stride * nv + k]) by the whole workgroup, L lanes per value,
// fixed order; lanes sub == 0 return the sums of their value k.
__device__ __forceinline__ double sum_partials(const double* partials, int count, int stride, int nv, int& k,
                                               int& sub) {
   int L = 64;
   while (L > 1 && L * nv > kBlock) L >>= 1;
   k = threadIdx.x / L;
   sub = threadIdx.x % L;
   auto ld = [&](long p) {
      return __hip_atomic_load(&partials[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
   };
   double s = 0.0;
   if (k < nv) {
      int b = sub;
      for (; b + 3 * L < count; b += 4 * L) {
         double a0 = ld((long)b * stride * nv + k);
         double a1 = ld((long)(b + L) * stride * nv + k);
         double a2 = ld((long)(b + 2 * L) * stride * nv + k);
         double a3 = ld((long)(b + 3 * L) * stride * nv + k);
         s += a0;
         s += a1;
         s += a2;
         s += a3;
      }
      for (; b < count; b += L) s += ld((long)b * stride * nv + k);
   }
   for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
   return s;
}

// Segmented form: the workgroups of one segment (one batched problem) reduce
// among themselves.  `seg_block` is this workgroup's index inside its segment,
// `seg_blocks` the segment's workgroup count, `partials` the segment's base
// (seg_blocks * nv doubles), `ticket` the segment's counter and `gticket` the
// segment's group counters (one per workgroup).  Segments of more than
// 16 workgroups reduce in two levels: the last workgroup of each group sums
// its group's partials into the group's first row, and the last group to
// finish sums the group rows; the order of every sum is fixed.  The group size
// (16, 32 or 64 workgroups, about the square root of the segment's count)
// keeps both levels short.
__device__ __forceinline__ int seg_group(int seg_blocks) {
   return seg_blocks <= 256 ? 16 : (seg_blocks <= 1024 ? 32 : 64);
}

// `first` gives the first kWaveBatch values as wave_sums takes them.
template <int MAXV, class First, class Get>
__device__ bool seg_reduce(int nv, First first, Get get, double* partials, unsigned* ticket, unsigned* gticket,
                           double* out, int seg_block, int seg_blocks) {
   __shared__ double sh[MAXV][kBlock / 64];
   __shared__ int is_last;
   wave_sums<MAXV>(nv, first, get, sh);
   __syncthreads();
   if ((int)threadIdx.x < nv) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kBlock / 64; w++) s += sh[threadIdx.x][w];
      __hip_atomic_store(&partials[(long)seg_block * nv + threadIdx.x], s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
   }
   __builtin_amdgcn_s_waitcnt(0);
   __syncthreads();
   int count = seg_blocks, stride = 1;
   const int group = seg_group(seg_blocks);
   if (seg_blocks > group) {
      const int g0 = seg_block / group * group;
      const int gs = min(group, seg_blocks - g0);
      if (threadIdx.x == 0)
         is_last = (__hip_atomic_fetch_add(&gticket[g0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                    (unsigned)gs - 1);
      __syncthreads();
      if (!is_last) return false;
      int k, sub;
      const double s = sum_partials(partials + (long)g0 * nv, gs, 1, nv, k, sub);
      if (k < nv && sub == 0)
         __hip_atomic_store(&partials[(long)g0 * nv + k], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      if (threadIdx.x == 0) __hip_atomic_store(&gticket[g0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      count = (seg_blocks + group - 1) / group;
      stride = group;
   }
   if (threadIdx.x == 0)
      is_last = (__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                 (unsigned)count - 1);
   __syncthreads();
   if (!is_last) return false;
   int k, sub;
   const double s = sum_partials(partials, count, stride, nv, k, sub);
   if (k < nv && sub == 0) out[k] = s;
   if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   __threadfence_block();
   __syncthreads();
   return true;
}

template <int MAXV, class Get>
__device__ __forceinline__ bool seg_reduce(int nv, Get get, double* partials, unsigned* ticket, unsigned* gticket,
                                          double* out, int seg_block, int seg_blocks) {
   return seg_reduce<MAXV>(
       nv, [&](int c) { return get(min(c, nv - 1)); }, get, partials, ticket, gticket, out, seg_block, seg_blocks);
}

template <int MAXV, class Get>
__device__ __forceinline__ void grid_reduce_finish(int nv, Get get, double* partials,
                                                   unsigned* ticket, double* out) {
   (void)grid_reduce<MAXV>(nv, get, partials, ticket, out);
}

template <int NV>
__device__ __forceinline__ void block_reduce_finish(double (&v)[NV], double* partials,
                                                    unsigned* ticket, double* out) {
   (void)grid_reduce<NV>(
       NV, [&](int k) { return v[k]; }, partials, ticket, out);
}

}  // namespace pk
