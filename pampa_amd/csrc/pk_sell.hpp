// Row kernel of the sliced-ELL layout (slices of h = 32, 64 or 128 rows, each
// slice stored column-major: entry k of row r of a slice sits at the slice's
// offset + (r mod h) + h k, so the rows of a wavefront read contiguous runs).
#pragma once

#include <hip/hip_runtime.h>

namespace pk {

// sum_k a[p_k] x[col[p_k]] over the stored entries of one sliced-ELL row
// (p_k = base + h k, k < wd), in ascending k: four entries' loads are issued
// before their products are summed, so several gathers are in flight per lane.
// Padding entries (value 0) add exact zeros, so the sum does not depend on
// the slice height.
template <class T, class A>
__device__ __forceinline__ A sell_row_dot(const int* __restrict__ col, const T* __restrict__ val,
                                          const T* __restrict__ x, long base, int wd, A acc, int h = 64) {
   int k = 0;
   for (; k + 4 <= wd; k += 4) {
      const long p = base + (long)k * h;
      const int c0 = col[p], c1 = col[p + h], c2 = col[p + 2 * h], c3 = col[p + 3 * h];
      const T v0 = val[p], v1 = val[p + h], v2 = val[p + 2 * h], v3 = val[p + 3 * h];
      const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
      acc += (A)(v0 * x0);
      acc += (A)(v1 * x1);
      acc += (A)(v2 * x2);
      acc += (A)(v3 * x3);
   }
   // the last 1 to 3 entries: all loads first (indices clamped to the last
   // entry), then the products in order
   const int rem = wd - k;
   if (rem > 0) {
      int cc[3];
      T vv[3], xx[3];
#pragma unroll
      for (int u = 0; u < 3; u++) {
         const long p = base + (long)(k + min(u, rem - 1)) * h;
         cc[u] = col[p];
         vv[u] = val[p];
      }
#pragma unroll
      for (int u = 0; u < 3; u++) xx[u] = x[cc[u]];
#pragma unroll
      for (int u = 0; u < 3; u++)
         if (u < rem) acc += (A)(vv[u] * xx[u]);
   }
   return acc;
}

// A row's first four stored entries (indices clamped to its last entry),
// loaded ahead: a caller issues them with its other first loads instead of
// after the barriers that precede their use.
template <class T>
struct SellHead {
   int c[4];
   T v[4];
};

template <class T>
__device__ __forceinline__ SellHead<T> sell_row_head(const int* __restrict__ col, const T* __restrict__ val,
                                                     long base, int wd, int h) {
   SellHead<T> hd;
#pragma unroll
   for (int u = 0; u < 4; u++) {
      const long p = base + (long)min(u, max(wd - 1, 0)) * h;
      hd.c[u] = wd > 0 ? col[p] : 0;
      hd.v[u] = wd > 0 ? val[p] : T(0);
   }
   return hd;
}

}  // namespace pk
