// Row kernel of the SELL-64 layout (slices of 64 rows = one wavefront, each
// slice stored column-major: lane r reads entry k of its row at base + 64 k).
#pragma once

#include <hip/hip_runtime.h>

namespace pk {

// sum_k a[p_k] x[col[p_k]] over the stored entries of one SELL-64 row
// (p_k = base + 64 k, k < wd), in ascending k: four entries' loads are issued
// before their products are summed, so several gathers are in flight per lane
template <class T, class A>
__device__ __forceinline__ A sell_row_dot(const int* __restrict__ col, const T* __restrict__ val,
                                          const T* __restrict__ x, long base, int wd, A acc) {
   int k = 0;
   for (; k + 4 <= wd; k += 4) {
      const long p = base + (long)k * 64;
      const int c0 = col[p], c1 = col[p + 64], c2 = col[p + 128], c3 = col[p + 192];
      const T v0 = val[p], v1 = val[p + 64], v2 = val[p + 128], v3 = val[p + 192];
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
         const long p = base + (long)(k + min(u, rem - 1)) * 64;
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

}  // namespace pk
