// Discrete ordinates on the device: matrix-free application of the S_N
// operator of SNSolver::buildMatrices (SNSolver.cxx:344-631) in the regime
// where it is a pure upwind transport operator (mixed-face-interpolation 1,
// least-squares boundary correction off, vacuum boundaries).  There the
// streaming + collision part L is lower triangular in upwind order for every
// direction, so L^-1 is applied by a sweep instead of a factorisation:
//
//   (Sigma_t V + sum_out c w A + sum_vac,out w A) psi_i = q_i - sum_in c w A psi_i2
//
// with w = Omega_m . n_f, c = (r_if + r_i2f) / r_ii2 (SNSolver.cxx:193-198 at
// delta = 1), q_i the isotropic scattering + fission source (SNSolver.cxx:426-439
// moved to the right-hand side).
//
// Parallel structure: each direction has its own upwind level schedule (cells
// whose upwind neighbours all sit in earlier levels), and directions do not
// depend on each other within a sweep.  One workgroup sweeps one direction:
// it walks that direction's levels in order, a workgroup barrier between
// levels, a thread per (direction, cell) item looping over the groups.  The
// whole sweep is one launch with no inter-workgroup synchronisation (the
// level-by-level launches of earlier designs paid a kernel boundary per level).
// The device layout of psi is [m][cell][group] (the G values of an upwind
// neighbour are contiguous); the reference layout [cell][group][m] is produced
// only at the boundary of the API.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "pk_internal.hpp"
#include "pk_reduce.hpp"

namespace pk {

constexpr int kMaxSnGroups = 16;

struct SnGeom {
   const int* fptr;
   const int* fnb;
   const double* fnorm;
   const double* farea;
   const double* coef;  // (r_if + r_i2f) / r_ii2 on interior faces
   const double* vol;
   const int* mat;
   const double* dirs;
};

// pos[m C + i]: the item's position in direction-major level order, which is
// also where its G values sit in psi
template <class T>
__global__ void __launch_bounds__(256) sn_sweep_step_kernel(const int* __restrict__ items, long nitems,
                                                            int C, int G, SnGeom geo,
                                                            const int* __restrict__ pos,
                                                            const T* __restrict__ sigt,
                                                            const T* __restrict__ q, T* psi) {
#pragma clang fp contract(off)
   long it = blockIdx.x * 256L + threadIdx.x;
   if (it >= nitems) return;
   const int code = items[it];
   const int m = code / C, i = code - m * C;
   const double o0 = geo.dirs[3 * m], o1 = geo.dirs[3 * m + 1], o2 = geo.dirs[3 * m + 2];
   T lhs0 = 0;
   T inflow[kMaxSnGroups];
   for (int g = 0; g < G; g++) inflow[g] = 0;
   for (int f = geo.fptr[i]; f < geo.fptr[i + 1]; f++) {
      const double* n = geo.fnorm + 3L * f;
      T w = (T)(o0 * n[0] + o1 * n[1] + o2 * n[2]);
      const int i2 = geo.fnb[f];
      const T A = (T)geo.farea[f];
      if (i2 < 0) {
         if (w > (T)0) lhs0 += w * A;
      } else if (w > (T)0) {
         lhs0 += (T)geo.coef[f] * w * A;
      } else {
         const T c = (T)geo.coef[f] * w * A;
         const T* up = psi + (long)pos[m * C + i2] * G;
         for (int g = 0; g < G; g++) inflow[g] += c * up[g];
      }
   }
   const int mt = geo.mat[i];
   const T V = (T)geo.vol[i];
   T* out = psi + (long)pos[code] * G;
   const T* qi = q + (long)i * G;
   for (int g = 0; g < G; g++) out[g] = (qi[g] - inflow[g]) / (sigt[mt * G + g] * V + lhs0);
}

// Per-item records: what a sweep needs of (direction m, cell i), taken from the
// geometry and the cross sections once per precision in the step kernel's own
// arithmetic (same operations in the same order, so the same bits): the
// denominators den_g = Sigma_t,g V + lhs0 with lhs0 =
// sum_out coef w A + sum_vac,out w A, and up to kMaxIn incoming interior faces
// as (source of the upwind value, coef w A) in face order.  An upwind value is
// read from a slot of the workgroup's local memory when its producer was given
// one, else from psi in global memory: a source s >= 0 is slot s, s < 0 is the
// item at position -(s + 1).  nb packs the incoming count (3 bits) and 1 + the item's
// own output slot (0 = none).  Faces whose weight rounds to zero in T (coef w A
// = 0) are left out.  Unused incoming entries hold (slot 0, weight 0).  Stored
// in direction-major level order: nb, up and den structure-of-arrays, the
// weights of an item together ([n][kMaxIn], one vector load).  When every
// value has a slot, tp packs an item's topology for the per-direction kernel
// into 16 bytes (one load): the four source slots as 16-bit fields (slots are
// local-memory indices, < 65536), then the item's own slot (-1 = none).  The
// records hold Sigma_t: they are built from the cross sections the sweep was
// prepared with.
constexpr int kMaxIn = 4;

template <class T>
struct SnRec {
   long n;
   int* nb;
   T* den;   // [G][n]
   int* up;  // [kMaxIn][n]
   T* cw;    // [n][kMaxIn]
   uint4* tp;  // [n] {up0 | up1 << 16, up2 | up3 << 16, own slot, 0}, or null
};

// slot_of[m C + i]: local-memory slot of item (m, i)'s value, -1 if none
template <class T>
__global__ void sn_rec_kernel(const int* __restrict__ items, long n, int C, int G, SnGeom geo,
                              const T* __restrict__ sigt, const int* __restrict__ slot_of,
                              const int* __restrict__ pos, SnRec<T> R, int* overflow) {
#pragma clang fp contract(off)
   long it = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (it >= n) return;
   const int code = items[it];
   const int m = code / C, i = code - m * C;
   const double o0 = geo.dirs[3 * m], o1 = geo.dirs[3 * m + 1], o2 = geo.dirs[3 * m + 2];
   T lhs0 = 0;
   int k = 0;
   unsigned tslot[kMaxIn] = {0, 0, 0, 0};
   for (int f = geo.fptr[i]; f < geo.fptr[i + 1]; f++) {
      const double* nn = geo.fnorm + 3L * f;
      T w = (T)(o0 * nn[0] + o1 * nn[1] + o2 * nn[2]);
      const int i2 = geo.fnb[f];
      const T A = (T)geo.farea[f];
      if (i2 < 0) {
         if (w > (T)0) lhs0 += w * A;
      } else if (w > (T)0) {
         lhs0 += (T)geo.coef[f] * w * A;
      } else if (w < (T)0) {
         if (k < kMaxIn) {
            const int up = m * C + i2;
            const int sl = slot_of[up];
            R.up[(long)k * n + it] = sl >= 0 ? sl : -(pos[up] + 1);
            R.cw[it * kMaxIn + k] = (T)geo.coef[f] * w * A;
            if (sl >= 0) tslot[k] = (unsigned)sl;
         } else {
            *overflow = 1;
         }
         k++;
      }
   }
   for (int j = k; j < kMaxIn; j++) {
      R.up[(long)j * n + it] = 0;
      R.cw[it * kMaxIn + j] = 0;
   }
   const int mt = geo.mat[i];
   const T V = (T)geo.vol[i];
   for (int g = 0; g < G; g++) R.den[(long)g * n + it] = sigt[mt * G + g] * V + lhs0;
   R.nb[it] = (k < kMaxIn ? k : kMaxIn) | ((slot_of[code] + 1) << 3);
   if (R.tp)
      R.tp[it] = make_uint4((tslot[0] & 0xFFFFu) | (tslot[1] << 16), (tslot[2] & 0xFFFFu) | (tslot[3] << 16),
                            (unsigned)slot_of[code], 0u);
}

// One whole sweep, one workgroup per direction.  The records of direction m are
// stored level by level (lptr[m][l] .. lptr[m][l+1]); the workgroup walks its
// levels in order with a workgroup barrier between them.  Upwind values come
// from slots of the workgroup's local memory, assigned on the host so that a
// slot lives from its producer's level to its last consumer's level.
// Directions never read each other's values: no workgroup waits on another.
//
// Latency: everything a level loads from global memory (its records and its
// source) is loaded D levels ahead, while earlier levels compute; within a
// level only local memory is read.  Each thread owns P items of a level (P NT
// >= the widest level; all loads issued together), and every upwind value comes
// from a slot (the host checks both, else the kernel below runs).  Per item the
// arithmetic is that of the step kernel (same operations, same order;
// zero-weight faces add exactly zero there).
template <class T, int GT>
struct SnItem {
   uint4 tp;  // source slots and own slot
   T den[GT];
   T cw[kMaxIn];
   T qv[GT];  // the item's source
};

template <class T, int GT>
__device__ __forceinline__ void sn_load_item(const SnRec<T>& R, const T* __restrict__ qp, int G, int it, int g0,
                                             SnItem<T, GT>& r) {
   r.tp = R.tp[it];
#pragma unroll
   for (int g = 0; g < GT; g++) r.qv[g] = qp[(long)(g0 + g) * R.n + it];
#pragma unroll
   for (int g = 0; g < GT; g++) r.den[g] = R.den[(long)(g0 + g) * R.n + it];
   if constexpr (sizeof(T) == 4) {
      const float4 c = reinterpret_cast<const float4*>(R.cw)[it];
      r.cw[0] = c.x;
      r.cw[1] = c.y;
      r.cw[2] = c.z;
      r.cw[3] = c.w;
   } else {
      const double2 c0 = reinterpret_cast<const double2*>(R.cw)[2L * it];
      const double2 c1 = reinterpret_cast<const double2*>(R.cw)[2L * it + 1];
      r.cw[0] = c0.x;
      r.cw[1] = c0.y;
      r.cw[2] = c1.x;
      r.cw[3] = c1.y;
   }
}

// one level l: prefetch of level l + D (records and source) into `pre`, the
// arithmetic of `cur` (level l), the stores, the barrier.  Groups
// g0 .. g0+GT-1 of G.  Item slot j of a thread covers position a + tid + j NT;
// lanes past the level's end repeat the level's last item (clamped index: a
// wave past the end touches one cache line) and store exactly its values
// again, so that every load and store is unconditional and the waits stay
// precise.  The source qp (group-major: one plane per group) and psi are in
// level order, so both are read and written with contiguous accesses.
template <class T, int GT, int NT, int P, int D>
__device__ __forceinline__ void sn_level(const SnRec<T>& R, const int* lp, int l, int nl, int G, int g0, int tid,
                                         const SnItem<T, GT> (&cur)[P], SnItem<T, GT> (&pre)[P],
                                         const T* __restrict__ qp, T* psi, T* slots) {
#pragma clang fp contract(off)
   const int a = lp[l], e = lp[l + 1];
   {
      const bool more = l + D < nl;
      const int ap = more ? lp[l + D] : a, ep = more ? lp[l + D + 1] : a + 1;
#pragma unroll
      for (int j = 0; j < P; j++) sn_load_item<T, GT>(R, qp, G, min(ap + tid + j * NT, ep - 1), g0, pre[j]);
   }
   __builtin_amdgcn_sched_barrier(0);
#pragma unroll
   for (int j = 0; j < P; j++) {
      const SnItem<T, GT>& r = cur[j];
      // every entry, used or not: an unused one is (slot 0, weight +0) and
      // adds an exact zero after the real faces (slots hold finite values:
      // zero at the start, then fluxes), so the sum has the step kernel's bits
      // and the local-memory reads issue together
      const int up[kMaxIn] = {(int)(r.tp.x & 0xFFFFu), (int)(r.tp.x >> 16), (int)(r.tp.y & 0xFFFFu),
                              (int)(r.tp.y >> 16)};
      T v[kMaxIn][GT];
#pragma unroll
      for (int k = 0; k < kMaxIn; k++)
#pragma unroll
         for (int g = 0; g < GT; g++) v[k][g] = slots[up[k] * GT + g];
      T inflow[GT];
#pragma unroll
      for (int g = 0; g < GT; g++) inflow[g] = 0;
#pragma unroll
      for (int k = 0; k < kMaxIn; k++)
#pragma unroll
         for (int g = 0; g < GT; g++) inflow[g] += r.cw[k] * v[k][g];
      T out[GT];
#pragma unroll
      for (int g = 0; g < GT; g++) out[g] = (r.qv[g] - inflow[g]) / r.den[g];
      const long it = min(a + tid + j * NT, e - 1);
      T* o = psi + it * G + g0;
#pragma unroll
      for (int g = 0; g < GT; g++) o[g] = out[g];
      const int os = (int)r.tp.z;
      if (os >= 0) {
#pragma unroll
         for (int g = 0; g < GT; g++) slots[os * GT + g] = out[g];
      }
   }
   __syncthreads();
}

// Workgroup b sweeps direction m for groups g0 .. g0+GT-1, G = S GT.  When M is
// a multiple of 8 the S workgroups of a direction get ids of one residue mod 8
// (the residue picks the chiplet a workgroup is dispatched to), so that they
// share the direction's records through one L2.  The records and source of
// step l + D are loaded during step l (D + 1 register sets in turn): a deeper
// D hides a longer load latency at a higher register cost.  The sets rotate
// with the step loop unrolled D + 1 times, so that no register copy forces an
// early wait.
template <class T, int GT, int NT, int P, int D>
__global__ void __launch_bounds__(NT) sn_sweep_lds_kernel(int G, int M, int lstride, const int* __restrict__ lptr,
                                                          const int* __restrict__ nlev, SnRec<T> R,
                                                          const T* __restrict__ qp, T* psi, int nslot_values) {
   static_assert(D >= 1 && D <= 4, "prefetch distance 1 to 4");
   extern __shared__ __align__(16) unsigned char sn_smem[];
   T* slots = reinterpret_cast<T*>(sn_smem);
   const int S = G / GT;
   const int b = blockIdx.x;
   int m, sidx;
   if (M % 8 == 0) {
      const int r = b >> 3;
      sidx = r % S;
      m = (b & 7) + 8 * (r / S);
   } else {
      m = b / S;
      sidx = b - m * S;
   }
   const int g0 = sidx * GT;
   const int* lp = lptr + (long)m * lstride;
   const int nl = nlev[m];
   const int tid = threadIdx.x;
   // the value slots start at zero (read with weight 0 before any write)
   for (int l = tid; l < nslot_values; l += NT) slots[l] = 0;
   __syncthreads();
   // D + 1 record sets in rotation: step l computes set l mod (D+1) and
   // prefetches step l + D into the set step l - 1 used
   SnItem<T, GT> X[D + 1][P];
#pragma unroll
   for (int u = 0; u < D; u++) {
      const bool in = u < nl;
      const int a = in ? lp[u] : lp[0], e = in ? lp[u + 1] : lp[0] + 1;
#pragma unroll
      for (int j = 0; j < P; j++) sn_load_item<T, GT>(R, qp, G, min(a + tid + j * NT, e - 1), g0, X[u][j]);
   }
   int l = 0;
   for (; l + D < nl; l += D + 1) {
#pragma unroll
      for (int u = 0; u <= D; u++)
         sn_level<T, GT, NT, P, D>(R, lp, l + u, nl, G, g0, tid, X[u], X[(u + D) % (D + 1)], qp, psi, slots);
   }
#pragma unroll
   for (int u = 0; u < D; u++)
      if (l + u < nl)
         sn_level<T, GT, NT, P, D>(R, lp, l + u, nl, G, g0, tid, X[u], X[(u + D) % (D + 1)], qp, psi, slots);
}

// Any configuration: the same sweep, one item at a time, upwind values from
// the slots or psi (level-ordered qp and psi as above).
template <class T, int NT>
__global__ void __launch_bounds__(NT) sn_sweep_lds_any_kernel(int G, int lstride, const int* __restrict__ lptr,
                                                              const int* __restrict__ nlev, SnRec<T> R,
                                                              const T* __restrict__ qp, T* psi) {
#pragma clang fp contract(off)
   extern __shared__ __align__(16) unsigned char sn_smem[];
   T* slots = reinterpret_cast<T*>(sn_smem);
   const int m = blockIdx.x;
   const int* lp = lptr + (long)m * lstride;
   const int nl = nlev[m];
   for (int l = 0; l < nl; l++) {
      const int a = lp[l], e = lp[l + 1];
      for (int it = a + (int)threadIdx.x; it < e; it += NT) {
         const int nb = R.nb[it];
         const int nin = nb & 7, os = (nb >> 3) - 1;
         T inflow[kMaxSnGroups];
         for (int g = 0; g < G; g++) inflow[g] = 0;
         for (int k = 0; k < nin; k++) {
            const int s = R.up[(long)k * R.n + it];
            const T c = R.cw[(long)it * kMaxIn + k];
            const T* up = s >= 0 ? slots + (long)s * G : psi + (long)(-(s + 1)) * G;
            for (int g = 0; g < G; g++) inflow[g] += c * up[g];
         }
         T* o = psi + (long)it * G;
         for (int g = 0; g < G; g++) {
            const T v = (qp[(long)g * R.n + it] - inflow[g]) / R.den[(long)g * R.n + it];
            o[g] = v;
            if (os >= 0) slots[(long)os * G + g] = v;
         }
      }
      __syncthreads();
   }
}

// qp[g][p] = q[i][g] for the item (m, i) at position p: the source in the
// sweep's level order, one plane per group, so that a workgroup sweeping
// group g reads its items' sources as one contiguous run.  One thread per
// item: the item's G values are read together (one gather) and written to
// the G planes (coalesced across the threads).
template <class T, int GT>
__global__ void __launch_bounds__(kBlock) sn_permute_kernel(const int* __restrict__ items, long n, int C, int G_,
                                                            const T* __restrict__ q, T* __restrict__ qp) {
   const int G = GT > 0 ? GT : G_;
   const long it = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (it >= n) return;
   const long i = items[it] % C;
   constexpr int GM = GT > 0 ? GT : kMaxSnGroups;
   T v[GM];
#pragma unroll
   for (int g = 0; g < GM; g++)
      if (GT > 0 || g < G) v[g] = q[i * G + g];
#pragma unroll
   for (int g = 0; g < GM; g++)
      if (GT > 0 || g < G) qp[(long)g * n + it] = v[g];
}

// phiw[i][g] = sum_m w_m psi(m, i, g) (FP64 accumulation, m ascending); the
// values of item (m, i) sit at position pos[m C + i] of psi.  One thread per
// cell: the positions are read coalesced, each item's G values as one gather.
template <class T, int GT>
__global__ void __launch_bounds__(kBlock) sn_moment_kernel(const T* __restrict__ psi, const int* __restrict__ pos,
                                                           const double* __restrict__ wts, int M, int C, int G_,
                                                           double* phiw) {
   const int G = GT > 0 ? GT : G_;
   const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i >= C) return;
   constexpr int GM = GT > 0 ? GT : kMaxSnGroups;
   double s[GM];
#pragma unroll
   for (int g = 0; g < GM; g++) s[g] = 0.0;
   for (int m = 0; m < M; m++) {
      const T* v = psi + (long)pos[(long)m * C + i] * G;
      const double w = wts[m];
#pragma unroll
      for (int g = 0; g < GM; g++)
         if (GT > 0 || g < G) s[g] += w * (double)v[g];
   }
#pragma unroll
   for (int g = 0; g < GM; g++)
      if (GT > 0 || g < G) phiw[i * G + g] = s[g];
}

template <class T>
static void sn_moment(const T* psi, const int* pos, const double* wts, int M, int C, int G, double* phiw,
                      hipStream_t st) {
   switch (G) {
      case 1: sn_moment_kernel<T, 1><<<grid_for(C), kBlock, 0, st>>>(psi, pos, wts, M, C, G, phiw); break;
      case 2: sn_moment_kernel<T, 2><<<grid_for(C), kBlock, 0, st>>>(psi, pos, wts, M, C, G, phiw); break;
      case 4: sn_moment_kernel<T, 4><<<grid_for(C), kBlock, 0, st>>>(psi, pos, wts, M, C, G, phiw); break;
      default: sn_moment_kernel<T, 0><<<grid_for(C), kBlock, 0, st>>>(psi, pos, wts, M, C, G, phiw); break;
   }
}

// q[i][g] = V_i (sum_g2 Sigma_s(g2 -> g) phiw[i][g2] + chi_g / k sum_g2 nuSf_g2
// phiw[i][g2]), one thread per cell, sums in ascending g2 (the oracle's order).
// GT > 0: the group count at compile time, the cell's moment held in registers.
template <class T, int GT>
__global__ void __launch_bounds__(kBlock) sn_source_kernel(int C, int G_, const double* __restrict__ vol,
                                                           const int* __restrict__ mat,
                                                           const double* __restrict__ sigs,
                                                           const double* __restrict__ nusf,
                                                           const double* __restrict__ chi,
                                                           const double* __restrict__ phiw,
                                                           const double* __restrict__ kk, T* __restrict__ q) {
#pragma clang fp contract(off)
   const int G = GT > 0 ? GT : G_;
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= C) return;
   const int mt = mat[i];
   const double inv_k = 1.0 / kk[0];
   constexpr int GM = GT > 0 ? GT : kMaxSnGroups;
   double ph[GM];
#pragma unroll
   for (int g2 = 0; g2 < GM; g2++) ph[g2] = (GT > 0 || g2 < G) ? phiw[(long)i * G + g2] : 0.0;
   double fis = 0.0;
#pragma unroll
   for (int g2 = 0; g2 < GM; g2++)
      if (GT > 0 || g2 < G) fis += nusf[mt * G + g2] * ph[g2];
   for (int g = 0; g < G; g++) {
      double s = 0.0;
#pragma unroll
      for (int g2 = 0; g2 < GM; g2++)
         if (GT > 0 || g2 < G) s += sigs[(mt * G + g2) * G + g] * ph[g2];
      q[(long)i * G + g] = (T)(vol[i] * (s + chi[mt * G + g] * inv_k * fis));
   }
}

// The same source for 16 groups on the matrix cores: with the per-material
// matrix K_m[g2][g] = Sigma_s,m(g2 -> g) + chi_m(g) nuSigma_f,m(g2) / k the
// source of 16 cells is the product Q(16 cells x 16 groups) = Phi(16 x 16) K_m,
// four V_MFMA_F64_16X16X4_F64 steps, then q = V Q.  A wavefront takes 16
// consecutive cells.  Lane l supplies row l % 16 of A and column l % 16 of B
// at inner index l / 16, and holds D rows 4 v + l / 16, v < 4, column l % 16
// (so the 64 lanes of item v hold 4 consecutive cells' 16 groups).
// Product step c (c < 4) pairs inner index kq with group 4 kq + c, so a lane
// reads its cell's 4 consecutive groups 4 kq .. 4 kq + 3 once.  Cells of a
// tile with different materials: one pass per distinct material, the other
// rows of A zero (they add exact zeros).  Sums in another order than the
// vector kernel, so the two agree to rounding, not bitwise.
typedef double sn_d4 __attribute__((ext_vector_type(4)));

template <class T>
__global__ void __launch_bounds__(256) sn_source_mfma_kernel(int C, const double* __restrict__ vol,
                                                             const int* __restrict__ mat,
                                                             const double* __restrict__ sigs,
                                                             const double* __restrict__ nusf,
                                                             const double* __restrict__ chi,
                                                             const double* __restrict__ phiw,
                                                             const double* __restrict__ kk, T* __restrict__ q) {
   constexpr int G = 16;
   const int lane = threadIdx.x & 63;
   const long c0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
   if (c0 >= C) return;  // the whole wavefront
   const int i = lane & 15, kq = lane >> 4;
   const long cell = c0 + i;
   const bool valid = cell < C;
   double a[4] = {0.0, 0.0, 0.0, 0.0};
   int mine = -1;
   if (valid) {
      const double2* p = (const double2*)(phiw + cell * G + 4 * kq);
      const double2 p0 = p[0], p1 = p[1];
      a[0] = p0.x;
      a[1] = p0.y;
      a[2] = p1.x;
      a[3] = p1.y;
      mine = mat[cell];
   }
   const double inv_k = 1.0 / kk[0];
   sn_d4 d = {0.0, 0.0, 0.0, 0.0};
   bool pending = valid;
   for (;;) {
      const unsigned long long bal = __ballot(pending);
      if (!bal) break;
      const int m = __shfl(mine, __ffsll((long long)bal) - 1);
      const bool take = pending && mine == m;
      const double cm = chi[m * G + i] * inv_k;
#pragma unroll
      for (int c = 0; c < 4; c++) {
         const int g2 = 4 * kq + c;
         const double b = sigs[(m * G + g2) * G + i] + cm * nusf[m * G + g2];
         d = __builtin_amdgcn_mfma_f64_16x16x4f64(take ? a[c] : 0.0, b, d, 0, 0, 0);
      }
      pending = pending && !take;
   }
#pragma unroll
   for (int v = 0; v < 4; v++) {
      const long cr = c0 + 4 * v + kq;
      if (cr < C) q[cr * G + i] = (T)(vol[cr] * d[v]);
   }
}

// Which contraction computes the source: the matrix cores for 16 groups
// (mode 0 = by group count, 2 = matrix cores), else the vector kernel.
static inline bool sn_use_mfma(int mode, int G) { return G == 16 && mode != 1; }

template <class T>
static void sn_source(int mode, int C, int G, const pk_mesh_s* mesh, const pk_xs_s* xs, const double* phiw,
                      const double* kk, T* q, hipStream_t st) {
   if (sn_use_mfma(mode, G)) {
      const long tiles = ((long)C + 15) / 16;
      sn_source_mfma_kernel<T><<<(unsigned)((tiles + 3) / 4), 256, 0, st>>>(C, mesh->vol.p, mesh->mat.p, xs->sigs.p,
                                                                          xs->nusf.p, xs->chi.p, phiw, kk, q);
      return;
   }
   const double *vol = mesh->vol.p, *sigs = xs->sigs.p, *nusf = xs->nusf.p, *chi = xs->chi.p;
   const int* mat = mesh->mat.p;
   switch (G) {
      case 1: sn_source_kernel<T, 1><<<grid_for(C), kBlock, 0, st>>>(C, G, vol, mat, sigs, nusf, chi, phiw, kk, q); break;
      case 2: sn_source_kernel<T, 2><<<grid_for(C), kBlock, 0, st>>>(C, G, vol, mat, sigs, nusf, chi, phiw, kk, q); break;
      case 4: sn_source_kernel<T, 4><<<grid_for(C), kBlock, 0, st>>>(C, G, vol, mat, sigs, nusf, chi, phiw, kk, q); break;
      case 8: sn_source_kernel<T, 8><<<grid_for(C), kBlock, 0, st>>>(C, G, vol, mat, sigs, nusf, chi, phiw, kk, q); break;
      case 16: sn_source_kernel<T, 16><<<grid_for(C), kBlock, 0, st>>>(C, G, vol, mat, sigs, nusf, chi, phiw, kk, q); break;
      default: sn_source_kernel<T, 0><<<grid_for(C), kBlock, 0, st>>>(C, G, vol, mat, sigs, nusf, chi, phiw, kk, q); break;
   }
}

// production rate sum nuSf phiw V
__global__ void __launch_bounds__(kBlock) sn_prod_kernel(int C, int G, const double* vol, const int* mat,
                                                         const double* nusf, const double* phiw,
                                                         double* partials, unsigned* ticket,
                                                         double* out) {
   double s = 0.0;
   for (long l = blockIdx.x * (long)kBlock + threadIdx.x; l < (long)C * G; l += (long)gridDim.x * kBlock) {
      long i = l / G;
      s += nusf[mat[i] * G + (l - i * G)] * phiw[l] * vol[i];
   }
   double v[1] = {s};
   block_reduce_finish<1>(v, partials, ticket, out);
}

// k update and rescaling: st = {k, prod_old, prod_new}; phiw *= prod_old/prod_new;
// out = {|phiw_new - phiw_old|^2, |phiw_new|^2}
__global__ void __launch_bounds__(kBlock) sn_update_kernel(long CG, double* phiw, const double* old,
                                                           const double* st, double* partials,
                                                           unsigned* ticket, double* out) {
   const double sc = st[1] / st[2];
   double d = 0.0, nn = 0.0;
   for (long l = blockIdx.x * (long)kBlock + threadIdx.x; l < CG; l += (long)gridDim.x * kBlock) {
      double v = phiw[l] * sc;
      phiw[l] = v;
      double e = v - old[l];
      d += e * e;
      nn += v * v;
   }
   double v2[2] = {d, nn};
   block_reduce_finish<2>(v2, partials, ticket, out);
}

// psi_ref[i][g][m] = scale * psi[m][i][g]
// Anderson acceleration of the fixed-point map phi -> g(phi) (one source
// iteration, production-normalised): with residuals f = g - phi and the last
// columns df = f_n - f_{n-1}, dg = g_n - g_{n-1} (a ring of depth slots), the
// next iterate is g_n - sum_j gamma_j dg_j with gamma minimising
// |f_n - sum_j gamma_j df_j| (Walker and Ni, SIAM J. Numer. Anal. 49 (2011)).
// The fixed point, so the eigenpair, is that of the plain iteration.
constexpr int kSnAAMax = 8;

// f = g - phi; when have_prev: df = f - fprev, dg = g - gprev; then fprev = f, gprev = g
__global__ void __launch_bounds__(kBlock) sn_aa_diff_kernel(long n, const double* __restrict__ g,
                                                            const double* __restrict__ phi, double* fprev,
                                                            double* gprev, double* df, double* dg, int have_prev) {
   const long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (l >= n) return;
   const double f = g[l] - phi[l], gl = g[l];
   if (have_prev) {
      df[l] = f - fprev[l];
      dg[l] = gl - gprev[l];
   }
   fprev[l] = f;
   gprev[l] = gl;
}

// out[j] = <df_j, df_slot>, out[depth + j] = <df_j, f> for every slot j (FP64,
// deterministic grid reduction)
__global__ void __launch_bounds__(kBlock) sn_aa_dots_kernel(long n, int depth, const double* __restrict__ df,
                                                            int slot, const double* __restrict__ f,
                                                            double* partials, unsigned* ticket, double* out) {
   double v[2 * kSnAAMax];
#pragma unroll
   for (int j = 0; j < 2 * kSnAAMax; j++) v[j] = 0.0;
   for (long l = blockIdx.x * (long)kBlock + threadIdx.x; l < n; l += (long)gridDim.x * kBlock) {
      const double a = df[(long)slot * n + l], b = f[l];
#pragma unroll
      for (int j = 0; j < kSnAAMax; j++) {
         if (j < depth) {
            const double d = df[(long)j * n + l];
            v[j] += d * a;
            v[kSnAAMax + j] += d * b;
         }
      }
   }
   (void)grid_reduce<2 * kSnAAMax>(
       2 * kSnAAMax, [&](int k) { return v[k]; }, partials, ticket, out);
}

struct SnAACoef {
   double g[kSnAAMax];
};

// phi = g - sum_j gamma_j dg_j (in place on g)
__global__ void __launch_bounds__(kBlock) sn_aa_mix_kernel(long n, int depth, double* phi,
                                                           const double* __restrict__ dg, SnAACoef c) {
   const long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (l >= n) return;
   double v = phi[l];
   // slots outside the window have gamma 0 and may hold no values yet
   for (int j = 0; j < depth; j++)
      if (c.g[j] != 0.0) v -= c.g[j] * dg[(long)j * n + l];
   phi[l] = v;
}

template <class T>
__global__ void sn_to_reference_kernel(const T* psi, const int* __restrict__ pos, int C, int G, int M, double scale,
                                       double* out) {
   long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   long n = (long)C * G * M;
   if (l >= n) return;
   int m = (int)(l % M);
   long ig = l / M;  // i*G + g
   const long i = ig / G;
   const int g = (int)(ig - i * G);
   out[l] = scale * (double)psi[(long)pos[(long)m * C + i] * G + g];
}

template <class T>
__global__ void sn_fill_kernel(T* x, long n, T v) {
   long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (l < n) x[l] = v;
}

}  // namespace pk

using namespace pk;

struct pk_sn_s {
   // execution context of this handle: its own stream, reduction scratch and
   // pinned scalars, so that solves of different handles can run concurrently
   // (from different host threads) with each other and with the library stream
   struct {
      hipStream_t stream = nullptr;
      double* partials = nullptr;  // [kMaxRedBlocks * 2 kSnAAMax]
      unsigned* ticket = nullptr;
      double* host_scalars = nullptr;  // pinned [32]
   } sc;
   ~pk_sn_s() {
      if (sc.stream) (void)hipStreamSynchronize(sc.stream);
      if (sc.partials) (void)hipFree(sc.partials);
      if (sc.ticket) (void)hipFree(sc.ticket);
      if (sc.host_scalars) (void)hipHostFree(sc.host_scalars);
      if (sc.stream) (void)hipStreamDestroy(sc.stream);
   }
   pk_mesh_s* mesh = nullptr;
   pk_xs_s* xs = nullptr;
   int C = 0, G = 0, M = 0;
   int precision = PK_PREC_FP64;
   double tol_k = 1e-9, tol_phi = 1e-8;
   int max_outer = 100000;
   std::vector<double> h_wts;
   std::vector<int> nlev, level;  // [M], [M][C]
   std::vector<long> step_ptr;    // nsteps + 1
   int nsteps = 0;
   DBuf<int> items;
   // per-item records in direction-major level order (one set per precision,
   // built on first use); geometry = true sweeps from the geometry instead
   bool geometry = false;
   DBuf<int> ditems;  // [M C]: items in direction-major level order
   DBuf<int> dpos;    // [M C]: position of item (m, i) in ditems and psi
   int lstride = 0;   // row stride of h_dlptr
   std::vector<int> h_ditems, h_dlptr;  // items, and per direction the offsets of its upwind levels
   std::vector<double> h_dirs;
   // records of one precision; slots = local-memory value slots they use
   struct RecBufs {
      DBuf<int> nb, up, slot_of;
      DBuf<uint4> tp;  // packed topology, per-direction kernel only
      DBuf<double> den64, cw64;
      DBuf<float> den32, cw32;
      int slots = 0;
      int gt = 0;             // groups per workgroup the slots are sized for
      int P = 0, D = 0;       // items per thread and prefetch distance of that kernel
      // sweep steps of each direction: its upwind levels, a level wider than
      // P 1024 split into independent sub-steps (items of one level do not
      // read each other); [M][lstride] (+2 padding) and [M]
      DBuf<int> lptr, nsteps;
      int lstride = 0;
      bool complete = false;  // every value that is read has a slot
      bool fast = false;      // the per-direction kernel applies
      bool built = false;
   } rec32, rec64;
   DBuf<int> rec_err;
   DBuf<double> dirs, wts, coef;
   DBuf<double> psi64, q64, sigt64, qp64;  // qp: the source in level order
   DBuf<float> psi32, q32, sigt32, qp32;
   DBuf<double> phiw, old, st;
   // value accesses of one sweep: per (i,g,m) one psi write plus one upwind psi
   // read per incoming interior face, per (i,g) one source read and one moment write
   double value_units = 0.0;
   // Anderson acceleration: depth (0 = plain source iteration) and history
   int aa_depth = 8;
   int contraction = 0;  // source contraction: 0 by group count, 1 vector, 2 matrix cores
   // mixed precision: FP32 sweeps until the iterate's relative change falls
   // below switch64 (the FP32 round-off floor is near 1e-7), then FP64 sweeps
   bool finish64 = true;
   DBuf<double> aa_df, aa_dg, aa_f, aa_g;  // [depth][C G] x 2, [C G] x 2
   // state kept between solves (warm start)
   bool has_state = false;
   double keff = 1.0;
   int outer = 0;
   long sweeps = 0;
   double seconds = 0.0;
};

static SnGeom geom(pk_sn_s* S) {
   pk_mesh_s* m = S->mesh;
   return SnGeom{m->fptr.p, m->fnb.p, m->fnorm.p, m->farea.p, S->coef.p, m->vol.p, m->mat.p, S->dirs.p};
}

static double sweep_bytes(const pk_sn_s* S, double s);
static int sn_prepare(pk_sn_s* S);

template <class T>
static SnRec<T> recs(pk_sn_s* S);
template <>
SnRec<double> recs<double>(pk_sn_s* S) {
   auto& B = S->rec64;
   return SnRec<double>{(long)S->M * S->C, B.nb.p, B.den64.p, B.up.p, B.cw64.p, B.tp.p};
}
template <>
SnRec<float> recs<float>(pk_sn_s* S) {
   auto& B = S->rec32;
   return SnRec<float>{(long)S->M * S->C, B.nb.p, B.den32.p, B.up.p, B.cw32.p, B.tp.p};
}

// Local-memory bytes a workgroup may use (the device's per-workgroup limit).
static int sn_lds_limit() {
   static const int lim = [] {
      int v = 0;
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx().device) != hipSuccess) v = 0;
      return v > 0 ? v : 65536;
   }();
   return lim;
}

// Value slots: item (m, i) gets a slot when some later item reads its value
// (a downwind neighbour: w < 0 seen from the neighbour, the schedule's own
// predicate) and a slot is free; the slot is held from the item's step to the
// step of its last reader and is reused from the step after.  Steps run in
// order inside a workgroup, so a slot is never written while a reader of its
// previous value can still run.  Returns the peak number of slots in use;
// complete = every value with a reader got a slot.
static int sn_assign_slots(const pk_sn_s* S, const std::vector<int>& lptr, const std::vector<int>& nsteps,
                           int lstride, int cap, std::vector<int>& slot_of, bool& complete) {
   const int C = S->C, M = S->M;
   const pk_mesh_s* mesh = S->mesh;
   const auto& fptr = mesh->h_fptr;
   const auto& fnb = mesh->h_fnb;
   const auto& fn = mesh->h_fnorm;
   slot_of.assign((size_t)M * C, -1);
   std::vector<int> last(C), step(C), fr;
   std::vector<std::vector<int>> rel;
   int peak = 0;
   complete = true;
   for (int m = 0; m < M; m++) {
      const double* om = &S->h_dirs[3L * m];
      const int* lp = &lptr[(size_t)m * lstride];
      const int ns = nsteps[m];
      for (int st = 0; st < ns; st++)
         for (int k = lp[st]; k < lp[st + 1]; k++) step[S->h_ditems[k] - m * C] = st;
      std::fill(last.begin(), last.end(), -1);
      for (int i2 = 0; i2 < C; i2++)
         for (int f = fptr[i2]; f < fptr[i2 + 1]; f++) {
            const int i = fnb[f];
            if (i < 0) continue;
            const double w = om[0] * fn[3L * f] + om[1] * fn[3L * f + 1] + om[2] * fn[3L * f + 2];
            if (w < 0.0 && step[i2] > last[i]) last[i] = step[i2];
         }
      fr.clear();
      for (int sl = cap - 1; sl >= 0; sl--) fr.push_back(sl);
      rel.assign(ns + 1, {});
      for (int st = 0; st < ns; st++) {
         for (int sl : rel[st]) fr.push_back(sl);
         for (int k = lp[st]; k < lp[st + 1]; k++) {
            const int i = S->h_ditems[k] - m * C;
            if (last[i] < 0) continue;
            if (fr.empty()) {
               complete = false;
               continue;
            }
            const int sl = fr.back();
            fr.pop_back();
            slot_of[(size_t)m * C + i] = sl;
            rel[last[i] + 1].push_back(sl);
            peak = std::max(peak, cap - (int)fr.size());
         }
      }
   }
   return peak;
}

// Steps of every direction from its upwind levels, levels wider than `cap`
// split into near-equal sub-steps (cap <= 0: the levels as they are).
static void sn_steps(const pk_sn_s* S, int cap, std::vector<int>& lptr, std::vector<int>& nsteps, int& lstride) {
   const int M = S->M;
   std::vector<std::vector<int>> b(M);
   int maxs = 0;
   for (int m = 0; m < M; m++) {
      const int* lp = &S->h_dlptr[(size_t)m * S->lstride];
      b[m].push_back(lp[0]);
      for (int l = 0; l < S->nlev[m]; l++) {
         const int a = lp[l], e = lp[l + 1], w = e - a;
         const int parts = cap > 0 ? (w + cap - 1) / cap : 1;
         for (int q = 1; q <= parts; q++) b[m].push_back(a + (int)((long)w * q / parts));
      }
      maxs = std::max(maxs, (int)b[m].size() - 1);
   }
   lstride = maxs + 1;
   lptr.assign((size_t)M * lstride + 2, 0);
   nsteps.assign(M, 0);
   for (int m = 0; m < M; m++) {
      nsteps[m] = (int)b[m].size() - 1;
      for (int st = 0; st <= maxs; st++)
         lptr[(size_t)m * lstride + st] = b[m][std::min(st, nsteps[m])];
   }
}

// Configuration of the per-direction kernel: GT groups per workgroup (G / GT
// workgroups per direction), P items per thread (steps of at most P 1024
// items) and prefetch distance D, within the register budget of 1024-thread
// workgroups (128 VGPRs).  GT is the smallest divisor of G in {1, 2, 4} whose
// M G / GT workgroups fit on the compute units at once (one per unit);
// (P, D) the measured best that fits: P = 1, D = 2 (FP32) or the deepest of
// 3 (GT = 1), 2, 1 (GT = 4) in FP64.  On the S6 lattice (FP32, GT = 1):
// 0.68 ms per sweep at P = 1 (D = 2, 3 or 4 alike), 0.73 ms at P = 2.
static bool sn_fits(bool f32, int gt, int P, int D) {
   if (P < 1 || P > 2 || D < 1 || D > 4) return false;
   if (f32) return gt == 1 ? (P == 1 || D <= 3) : gt == 2 ? (P == 1 ? D <= 3 : D <= 2) : (P == 1 && D <= 2);
   return gt == 1 ? (P == 1 ? D <= 3 : D <= 2) : gt == 2 ? (P == 1 && D <= 2) : (P == 1 && D == 1);
}

template <class T>
static bool sn_plan(const pk_sn_s* S, int& gt, int& P, int& D) {
   constexpr bool f32 = sizeof(T) == sizeof(float);
   const char* env = getenv("PK_SWEEP_GT");
   const int force_gt = env ? atoi(env) : 0;
   env = getenv("PK_SWEEP_P");
   const int force_p = env ? atoi(env) : 0;
   env = getenv("PK_SWEEP_D");
   const int force_d = env ? atoi(env) : 0;
   const int cus = std::max(1, ctx().num_cus);
   gt = 0;
   for (int cand : {1, 2, 4}) {
      if (S->G % cand || (force_gt && cand != force_gt)) continue;
      gt = cand;
      if ((long)S->M * (S->G / cand) <= cus) break;
   }
   if (!gt) return false;
   // one item per thread (steps of at most 1024 items) measured fastest; the
   // prefetch depth matters little beyond 2
   P = 1;
   D = f32 ? 2 : (gt == 1 ? 3 : gt == 2 ? 2 : 1);
   if (force_p || force_d) {
      const int p2 = force_p ? force_p : P, d2 = force_d ? force_d : D;
      if (sn_fits(f32, gt, p2, d2)) {
         P = p2;
         D = d2;
      }
   }
   return sn_fits(f32, gt, P, D);
}

// records of precision T (with their value slots), built once
template <class T>
static int sn_records(pk_sn_s* S) {
   auto& c = S->sc;
   const long n = (long)S->M * S->C;
   const bool f32 = sizeof(T) == sizeof(float);
   auto& B = f32 ? S->rec32 : S->rec64;
   if (B.built) return 0;
   // first build: the arrays and the value slots (geometry only, kept for
   // rebuilds after a cross-section change, which rerun the record kernel only)
   if (!B.nb.p) {
      PK_REQUIRE(!B.nb.alloc(n) && !B.up.alloc((size_t)kMaxIn * n), "device allocation failed");
      if (f32)
         PK_REQUIRE(!B.den32.alloc((size_t)S->G * n) && !B.cw32.alloc((size_t)kMaxIn * n),
                    "device allocation failed");
      else
         PK_REQUIRE(!B.den64.alloc((size_t)S->G * n) && !B.cw64.alloc((size_t)kMaxIn * n),
                    "device allocation failed");
   }
   if (!S->rec_err.p) PK_REQUIRE(!S->rec_err.alloc(1), "device allocation failed");
   if (!B.slot_of.p) {
      // steps and slots for the per-direction kernel; when some value gets no
      // slot there, the one-item kernel's (Kahn levels, slots of all G groups)
      int gt = S->G, P = 1, D = 1;
      const bool fast = sn_plan<T>(S, gt, P, D);
      auto cap_for = [&](int groups) {
         return std::max(1, std::min(sn_lds_limit() / (groups * (int)sizeof(T)), (1 << 27)));
      };
      std::vector<int> slot_of, lptr, nsteps;
      int lstride = 0;
      // PK_SWEEP_STEP caps the items of a step below P 1024 (tests of the split)
      const char* env = getenv("PK_SWEEP_STEP");
      const int cap = env && atoi(env) > 0 ? std::min(P * 1024, atoi(env)) : P * 1024;
      sn_steps(S, fast ? cap : 0, lptr, nsteps, lstride);
      B.slots = std::max(1, sn_assign_slots(S, lptr, nsteps, lstride, cap_for(fast ? gt : S->G), slot_of,
                                            B.complete));
      // (slots are 16-bit fields of the packed topology)
      B.fast = fast && B.complete && B.slots <= 65536;
      if (!B.fast) {
         gt = S->G;
         sn_steps(S, 0, lptr, nsteps, lstride);
         B.slots = std::max(1, sn_assign_slots(S, lptr, nsteps, lstride, cap_for(gt), slot_of, B.complete));
      }
      if (B.fast) PK_REQUIRE(!B.tp.alloc(n), "device allocation failed");
      B.gt = gt;
      B.P = P;
      B.D = D;
      B.lstride = lstride;
      PK_REQUIRE(!B.slot_of.alloc(slot_of.size()) && !B.lptr.alloc(lptr.size()) && !B.nsteps.alloc(nsteps.size()),
                 "device allocation failed");
      PK_HIP(hipMemcpyAsync(B.slot_of.p, slot_of.data(), sizeof(int) * slot_of.size(), hipMemcpyHostToDevice,
                            c.stream));
      PK_HIP(hipMemcpyAsync(B.lptr.p, lptr.data(), sizeof(int) * lptr.size(), hipMemcpyHostToDevice, c.stream));
      PK_HIP(hipMemcpyAsync(B.nsteps.p, nsteps.data(), sizeof(int) * nsteps.size(), hipMemcpyHostToDevice,
                            c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
   }
   DBuf<int>& dslot = B.slot_of;
   PK_HIP(hipMemsetAsync(S->rec_err.p, 0, sizeof(int), c.stream));
   const T* sigt = f32 ? (const T*)S->sigt32.p : (const T*)S->sigt64.p;
   sn_rec_kernel<T><<<grid_for(n), kBlock, 0, c.stream>>>(S->ditems.p, n, S->C, S->G, geom(S), sigt, dslot.p,
                                                        S->dpos.p, recs<T>(S), S->rec_err.p);
   PK_HIP(hipGetLastError());
   int h = 0;
   PK_HIP(hipMemcpyAsync(&h, S->rec_err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   PK_REQUIRE(h == 0, "more incoming faces per cell than the sweep records hold");
   B.built = true;
   return 0;
}

template <class T, int GT, int P, int D>
static int sn_launch_lds(pk_sn_s* S, const T* qp, T* psi) {
   auto& c = S->sc;
   constexpr int NT = 1024;
   const auto& B = sizeof(T) == sizeof(float) ? S->rec32 : S->rec64;
   const size_t shm = (size_t)B.slots * GT * sizeof(T);
   static const hipError_t attr = hipFuncSetAttribute((const void*)sn_sweep_lds_kernel<T, GT, NT, P, D>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, sn_lds_limit());
   PK_HIP(attr);
   PK_LAUNCH_SHM("sn_sweep", profiling() ? sweep_bytes(S, sizeof(T)) : 0.0, (sn_sweep_lds_kernel<T, GT, NT, P, D>),
                 S->M * (S->G / GT), NT, shm, c.stream, S->G, S->M, B.lstride, B.lptr.p, B.nsteps.p, recs<T>(S),
                 qp, psi, B.slots * GT);
   return 0;
}

template <class T>
static int sn_launch_lds_any(pk_sn_s* S, const T* qp, T* psi) {
   auto& c = S->sc;
   constexpr int NT = 1024;
   const auto& B = sizeof(T) == sizeof(float) ? S->rec32 : S->rec64;
   const size_t shm = (size_t)B.slots * S->G * sizeof(T);
   static const hipError_t attr = hipFuncSetAttribute((const void*)sn_sweep_lds_any_kernel<T, NT>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, sn_lds_limit());
   PK_HIP(attr);
   PK_LAUNCH_SHM("sn_sweep", profiling() ? sweep_bytes(S, sizeof(T)) : 0.0, (sn_sweep_lds_any_kernel<T, NT>), S->M,
                 NT, shm, c.stream, S->G, B.lstride, B.lptr.p, B.nsteps.p, recs<T>(S), qp, psi);
   return 0;
}

// One sweep psi = L^-1 q: one launch of the per-direction kernel (or, with
// geometry set, one launch per global level step of the geometry kernel).
template <class T>
static int sn_sweep(pk_sn_s* S, const T* sigt, const T* q, T* psi) {
   auto& c = S->sc;
   if (S->geometry) {
      SnGeom g = geom(S);
      const double per_item = profiling() ? sweep_bytes(S, sizeof(T)) / (double)S->step_ptr[S->nsteps] : 0.0;
      for (int s = 0; s < S->nsteps; s++) {
         long a = S->step_ptr[s], n = S->step_ptr[s + 1] - a;
         PK_LAUNCH("sn_sweep_step", per_item * n, sn_sweep_step_kernel<T>, grid_for(n, 256), 256, c.stream,
                   S->items.p + a, n, S->C, S->G, g, S->dpos.p, sigt, q, psi);
      }
      PK_HIP(hipGetLastError());
      return 0;
   }
   PK_CALL(sn_records<T>(S));
   constexpr bool f32 = sizeof(T) == sizeof(float);
   const auto& B = f32 ? S->rec32 : S->rec64;
   // the source in level order (one gather over all items, all compute units)
   T* qp = f32 ? (T*)S->qp32.p : (T*)S->qp64.p;
   {
      const long n = (long)S->M * S->C;
      const int G = S->G;
      const double traffic = profiling() ? (double)n * G * sizeof(T) * 2 + n * 4.0 : 0.0;
      switch (G) {
         case 1:
            PK_LAUNCH("sn_permute", traffic, (sn_permute_kernel<T, 1>), grid_for(n), kBlock, c.stream, S->ditems.p,
                      n, S->C, G, q, qp);
            break;
         case 2:
            PK_LAUNCH("sn_permute", traffic, (sn_permute_kernel<T, 2>), grid_for(n), kBlock, c.stream, S->ditems.p,
                      n, S->C, G, q, qp);
            break;
         case 4:
            PK_LAUNCH("sn_permute", traffic, (sn_permute_kernel<T, 4>), grid_for(n), kBlock, c.stream, S->ditems.p,
                      n, S->C, G, q, qp);
            break;
         default:
            PK_LAUNCH("sn_permute", traffic, (sn_permute_kernel<T, 0>), grid_for(n), kBlock, c.stream, S->ditems.p,
                      n, S->C, G, q, qp);
            break;
      }
   }
   int done = -1;
   if (B.fast) {
      switch (B.gt * 100 + B.P * 10 + B.D) {
         case 111: done = sn_launch_lds<T, 1, 1, 1>(S, qp, psi); break;
         case 112: done = sn_launch_lds<T, 1, 1, 2>(S, qp, psi); break;
         case 113: done = sn_launch_lds<T, 1, 1, 3>(S, qp, psi); break;
         case 114: if constexpr (f32) done = sn_launch_lds<T, 1, 1, 4>(S, qp, psi); break;
         case 121: done = sn_launch_lds<T, 1, 2, 1>(S, qp, psi); break;
         case 122: done = sn_launch_lds<T, 1, 2, 2>(S, qp, psi); break;
         case 123: if constexpr (f32) done = sn_launch_lds<T, 1, 2, 3>(S, qp, psi); break;
         case 211: done = sn_launch_lds<T, 2, 1, 1>(S, qp, psi); break;
         case 212: done = sn_launch_lds<T, 2, 1, 2>(S, qp, psi); break;
         case 213: if constexpr (f32) done = sn_launch_lds<T, 2, 1, 3>(S, qp, psi); break;
         case 221: if constexpr (f32) done = sn_launch_lds<T, 2, 2, 1>(S, qp, psi); break;
         case 222: if constexpr (f32) done = sn_launch_lds<T, 2, 2, 2>(S, qp, psi); break;
         case 411: done = sn_launch_lds<T, 4, 1, 1>(S, qp, psi); break;
         case 412: if constexpr (f32) done = sn_launch_lds<T, 4, 1, 2>(S, qp, psi); break;
         default: break;
      }
      if (done > 0) return 1;
   }
   if (done != 0) PK_CALL(sn_launch_lds_any<T>(S, qp, psi));
   PK_HIP(hipGetLastError());
   return 0;
}

// algorithmic bytes of one sweep: value accesses x value size, plus the
// geometry once: per face normal (24 B), area (8), coefficient (8),
// neighbour (4); per cell volume (8), material (4), face pointer (4)
static double sweep_bytes(const pk_sn_s* S, double s) {
   return S->value_units * s + (double)S->mesh->F * 44.0 + (double)S->C * 16.0;
}

extern "C" {

int pk_sn_create(pk_mesh mesh, pk_xs xs, int M, const double* dirs, const double* wts, double delta,
                 int num_bc, const int* bc_type, pk_sn* out) {
   PK_CALL(ensure_init());
   PK_REQUIRE(mesh && xs && dirs && wts && M > 0, "null argument");
   PK_REQUIRE(delta == 1.0, "sweeps need pure upwind face interpolation (mixed-face-interpolation 1)");
   PK_REQUIRE(xs->G <= kMaxSnGroups, "too many energy groups for the sweep kernel");
   PK_REQUIRE(!xs->feedback, "sweeps need constant cross sections per material");
   const int C = mesh->C;
   PK_REQUIRE((long)M * C < (1L << 31), "mesh too large for the sweep schedule");
   const auto& fptr = mesh->h_fptr;
   const auto& fnb = mesh->h_fnb;
   const auto& fn = mesh->h_fnorm;
   for (int f = 0; f < mesh->F; f++)
      if (fnb[f] < 0) {
         int id = -fnb[f];
         PK_REQUIRE(id < num_bc && bc_type[id] == PK_BC_VACUUM,
                    "sweeps need vacuum boundaries (reflective boundaries couple directions)");
      }
   auto* S = new pk_sn_s();
   {
      auto& sc = S->sc;
      if (hipStreamCreateWithFlags(&sc.stream, hipStreamNonBlocking) != hipSuccess ||
          hipMalloc(&sc.partials, sizeof(double) * kMaxRedBlocks * 2 * kSnAAMax) != hipSuccess ||
          hipMalloc(&sc.ticket, sizeof(unsigned) * 4) != hipSuccess ||
          hipHostMalloc(&sc.host_scalars, sizeof(double) * 32, hipHostMallocDefault) != hipSuccess ||
          hipMemsetAsync(sc.ticket, 0, sizeof(unsigned) * 4, sc.stream) != hipSuccess ||
          hipStreamSynchronize(sc.stream) != hipSuccess) {
         delete S;
         PK_FAIL("unable to set up the sweep's stream and scratch");
      }
   }
   S->mesh = mesh;
   S->xs = xs;
   S->C = C;
   S->G = xs->G;
   S->M = M;
   S->h_wts.assign(wts, wts + M);
   // face coefficients (SNSolver.cxx:193-198, delta = 1)
   std::vector<double> coef(mesh->F, 0.0);
   const auto& cen = mesh->h_cen;
   const auto& fcen = mesh->h_fcen;
   auto dist = [](const double* a, const double* b) {
      double d = 0.0;
      for (int k = 0; k < 3; k++) d += (b[k] - a[k]) * (b[k] - a[k]);
      return std::sqrt(d);
   };
   for (int i = 0; i < C; i++)
      for (int f = fptr[i]; f < fptr[i + 1]; f++) {
         int i2 = fnb[f];
         if (i2 < 0) continue;
         double rif = dist(&fcen[3L * f], &cen[3L * i]);
         double ri2f = dist(&fcen[3L * f], &cen[3L * i2]);
         double rii2 = dist(&cen[3L * i], &cen[3L * i2]);
         coef[f] = (ri2f + delta * rif) / rii2;
      }
   // upwind level schedule per direction (Kahn), ascending cells inside a level
   S->nlev.assign(M, 0);
   S->level.assign((size_t)M * C, 0);
   std::vector<int> indeg(C), cur, nxt;
   std::vector<std::vector<int>> lists(M);
   long nin_total = 0;
   int maxlev = 0;
   for (int m = 0; m < M; m++) {
      const double* om = dirs + 3L * m;
      for (int i = 0; i < C; i++) {
         int d = 0;
         for (int f = fptr[i]; f < fptr[i + 1]; f++) {
            if (fnb[f] < 0) continue;
            double w = om[0] * fn[3L * f] + om[1] * fn[3L * f + 1] + om[2] * fn[3L * f + 2];
            if (w < 0.0) d++;
         }
         indeg[i] = d;
         nin_total += d;
      }
      cur.clear();
      for (int i = 0; i < C; i++)
         if (indeg[i] == 0) cur.push_back(i);
      int lev = 0, done = 0;
      auto& L = lists[m];
      L.clear();
      L.reserve(C);
      std::vector<int> lev_ptr{0};
      while (!cur.empty()) {
         for (int i : cur) {
            S->level[(size_t)m * C + i] = lev;
            L.push_back(i);
         }
         done += (int)cur.size();
         lev_ptr.push_back((int)L.size());
         nxt.clear();
         for (int i : cur)
            for (int f = fptr[i]; f < fptr[i + 1]; f++) {
               int i2 = fnb[f];
               if (i2 < 0) continue;
               double w = om[0] * fn[3L * f] + om[1] * fn[3L * f + 1] + om[2] * fn[3L * f + 2];
               if (w > 0.0 && --indeg[i2] == 0) nxt.push_back(i2);
            }
         std::sort(nxt.begin(), nxt.end());
         cur.swap(nxt);
         lev++;
      }
      if (done != C) {
         delete S;
         PK_FAIL("cyclic upwind dependencies: the mesh cannot be swept");
      }
      S->nlev[m] = lev;
      maxlev = std::max(maxlev, lev);
   }
   // steps: level s of every direction
   S->nsteps = maxlev;
   S->step_ptr.assign(maxlev + 1, 0);
   std::vector<int> items;
   items.reserve((size_t)M * C);
   {
      std::vector<std::vector<int>> lptr(M);
      for (int m = 0; m < M; m++) {
         lptr[m].assign(S->nlev[m] + 1, 0);
         const auto& L = lists[m];
         // rebuild level boundaries from the level array (lists are level-major)
         int l = 0;
         for (size_t k = 0; k < L.size(); k++) {
            int lv = S->level[(size_t)m * C + L[k]];
            while (l < lv) lptr[m][++l] = (int)k;
         }
         while (l < S->nlev[m]) lptr[m][++l] = (int)L.size();
      }
      for (int s = 0; s < maxlev; s++) {
         for (int m = 0; m < M; m++) {
            if (s >= S->nlev[m]) continue;
            for (int k = lptr[m][s]; k < lptr[m][s + 1]; k++) items.push_back(m * C + lists[m][k]);
         }
         S->step_ptr[s + 1] = (long)items.size();
      }
   }
   // direction-major copy for the per-direction kernel: the items of direction
   // m level by level, with absolute level offsets padded to maxlev + 1
   std::vector<int> ditems;
   ditems.reserve((size_t)M * C);
   S->lstride = maxlev + 1;
   std::vector<int> dlptr((size_t)M * S->lstride, 0);
   for (int m = 0; m < M; m++) {
      const auto& L = lists[m];
      int l = 0;
      const int base = (int)ditems.size();
      dlptr[(size_t)m * S->lstride] = base;
      for (size_t k = 0; k < L.size(); k++) {
         const int lv = S->level[(size_t)m * C + L[k]];
         while (l < lv) dlptr[(size_t)m * S->lstride + (++l)] = base + (int)k;
         ditems.push_back(m * C + L[k]);
      }
      while (l < maxlev) dlptr[(size_t)m * S->lstride + (++l)] = (int)ditems.size();
   }
   const char* ge = getenv("PK_SWEEP_GEOMETRY");
   S->geometry = ge && *ge && *ge != '0';
   const int G = S->G;
   S->value_units = (double)((long)C * G * M + nin_total * G) + 2.0 * C * G;
   int e = 0;
   e |= S->items.alloc(items.size()) | S->ditems.alloc(ditems.size());
   e |= S->dirs.alloc(3L * M) | S->wts.alloc(M) | S->coef.alloc(mesh->F ? mesh->F : 1);
   e |= S->phiw.alloc((size_t)C * G) | S->old.alloc((size_t)C * G) | S->st.alloc(8 + 2 * kSnAAMax);
   if (e) {
      delete S;
      PK_FAIL("device allocation failed");
   }
   PK_HIP(hipMemcpy(S->items.p, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice));
   PK_HIP(hipMemcpy(S->ditems.p, ditems.data(), sizeof(int) * ditems.size(), hipMemcpyHostToDevice));
   {
      std::vector<int> pos((size_t)M * C);
      for (size_t k = 0; k < ditems.size(); k++) pos[ditems[k]] = (int)k;
      PK_REQUIRE(!S->dpos.alloc(pos.size()), "device allocation failed");
      PK_HIP(hipMemcpy(S->dpos.p, pos.data(), sizeof(int) * pos.size(), hipMemcpyHostToDevice));
   }
   S->h_ditems.swap(ditems);
   S->h_dlptr.swap(dlptr);
   S->h_dirs.assign(dirs, dirs + 3L * M);

   PK_HIP(hipMemcpy(S->dirs.p, dirs, sizeof(double) * 3 * M, hipMemcpyHostToDevice));
   PK_HIP(hipMemcpy(S->wts.p, wts, sizeof(double) * M, hipMemcpyHostToDevice));
   if (mesh->F) PK_HIP(hipMemcpy(S->coef.p, coef.data(), sizeof(double) * mesh->F, hipMemcpyHostToDevice));
   *out = S;
   return 0;
}

int pk_sn_set_xs(pk_sn S, const double* sigt, const double* sigs, const double* nusf, const double* kapf,
                 const double* chi) {
   PK_REQUIRE(S && sigt && sigs && nusf && kapf && chi, "null argument");
   pk_xs_s* xs = S->xs;
   auto& c = S->sc;
   const size_t n1 = (size_t)xs->nm * xs->G, n2 = n1 * xs->G;
   xs->h_sigt.assign(sigt, sigt + n1);
   xs->h_sigs.assign(sigs, sigs + n2);
   xs->h_nusf.assign(nusf, nusf + n1);
   xs->h_kapf.assign(kapf, kapf + n1);
   xs->h_chi.assign(chi, chi + n1);
   // on the handle's stream, after its own kernels that read the tables
   PK_HIP(hipMemcpyAsync(xs->sigt.p, xs->h_sigt.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->sigs.p, xs->h_sigs.data(), sizeof(double) * n2, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->nusf.p, xs->h_nusf.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->kapf.p, xs->h_kapf.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->chi.p, xs->h_chi.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   const int nm = (int)n1;
   if (S->sigt64.p)
      PK_HIP(hipMemcpyAsync(S->sigt64.p, xs->h_sigt.data(), sizeof(double) * nm, hipMemcpyHostToDevice, c.stream));
   std::vector<float> st(xs->h_sigt.begin(), xs->h_sigt.end());
   if (S->sigt32.p) PK_HIP(hipMemcpyAsync(S->sigt32.p, st.data(), sizeof(float) * nm, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   // the records hold Sigma_t V: rebuilt on the next sweep
   S->rec32.built = false;
   S->rec64.built = false;
   return 0;
}

int pk_sn_prepare(pk_sn S) {
   PK_REQUIRE(S, "null sweep");
   PK_CALL(sn_prepare(S));
   // one sweep of a zero source: records, value slots and kernel attributes in
   // place, so that later solves allocate nothing
   auto& c = S->sc;
   const long CG = (long)S->C * S->G;
   if (S->precision == PK_PREC_MIXED) {
      sn_fill_kernel<float><<<grid_for(CG), kBlock, 0, c.stream>>>(S->q32.p, CG, 0.0f);
      PK_CALL(sn_sweep<float>(S, S->sigt32.p, S->q32.p, S->psi32.p));
   }
   if (S->precision != PK_PREC_MIXED || S->finish64) {
      sn_fill_kernel<double><<<grid_for(CG), kBlock, 0, c.stream>>>(S->q64.p, CG, 0.0);
      PK_CALL(sn_sweep<double>(S, S->sigt64.p, S->q64.p, S->psi64.p));
   }
   PK_HIP(hipGetLastError());
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

int pk_sn_get_moment(pk_sn S, double* phiw) {
   PK_REQUIRE(S && phiw, "null argument");
   auto& c = S->sc;
   PK_HIP(hipMemcpyAsync(phiw, S->phiw.p, sizeof(double) * S->C * S->G, hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

int pk_sn_set_acceleration(pk_sn S, int depth) {
   PK_REQUIRE(S && depth >= 0 && depth <= kSnAAMax, "wrong acceleration depth");
   S->aa_depth = depth;
   return 0;
}

int pk_sn_reset(pk_sn S) {
   PK_REQUIRE(S, "null sweep");
   S->has_state = false;
   return 0;
}

int pk_sn_get_schedule(pk_sn S, int* num_levels, int* level_of_cell) {
   PK_REQUIRE(S, "null sweep");
   if (num_levels) std::copy(S->nlev.begin(), S->nlev.end(), num_levels);
   if (level_of_cell) std::copy(S->level.begin(), S->level.end(), level_of_cell);
   return 0;
}

int pk_sn_set_precision(pk_sn S, int precision) {
   PK_REQUIRE(S && (precision == PK_PREC_FP64 || precision == PK_PREC_MIXED), "wrong precision");
   S->precision = precision;
   return 0;
}

int pk_sn_set_tolerances(pk_sn S, double tol_k, double tol_phi, int max_outer) {
   PK_REQUIRE(S && tol_k > 0.0 && tol_phi > 0.0 && max_outer > 0, "wrong tolerances");
   S->tol_k = tol_k;
   S->tol_phi = tol_phi;
   S->max_outer = max_outer;
   return 0;
}

// Buffers of the current precision (allocated once).  Every copy of the
// sweep handle goes through its own stream: no operation on the legacy NULL
// stream, which would synchronise with (and break a graph capture on) the
// library stream of another thread.
static int sn_prepare(pk_sn_s* S) {
   auto& c = S->sc;
   const long CGM = (long)S->C * S->G * S->M, CG = (long)S->C * S->G;
   if (S->aa_depth > 0 && !S->aa_df.p) {
      PK_REQUIRE(!S->aa_df.alloc((size_t)kSnAAMax * CG) && !S->aa_dg.alloc((size_t)kSnAAMax * CG) &&
                     !S->aa_f.alloc(CG) && !S->aa_g.alloc(CG),
                 "device allocation failed");
      // the history starts as zeros (slots outside the window are never read
      // with a nonzero weight, but hold defined values)
      PK_HIP(hipMemsetAsync(S->aa_df.p, 0, sizeof(double) * kSnAAMax * CG, c.stream));
      PK_HIP(hipMemsetAsync(S->aa_dg.p, 0, sizeof(double) * kSnAAMax * CG, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
   }
   const int nm = S->xs->nm * S->G;
   // mixed precision finishes its solves with FP64 sweeps: both sets
   if (S->precision == PK_PREC_MIXED) {
      if (!S->psi32.p) {
         PK_REQUIRE(!S->psi32.alloc(CGM) && !S->q32.alloc(CG) && !S->sigt32.alloc(nm) && !S->qp32.alloc(CGM),
                    "device allocation failed");
         std::vector<float> st(S->xs->h_sigt.begin(), S->xs->h_sigt.end());
         PK_HIP(hipMemcpyAsync(S->sigt32.p, st.data(), sizeof(float) * nm, hipMemcpyHostToDevice, c.stream));
         PK_HIP(hipStreamSynchronize(c.stream));
      }
   }
   if (S->precision != PK_PREC_MIXED || S->finish64) {
      if (!S->psi64.p) {
         PK_REQUIRE(!S->psi64.alloc(CGM) && !S->q64.alloc(CG) && !S->sigt64.alloc(nm) && !S->qp64.alloc(CGM),
                    "device allocation failed");
         PK_HIP(hipMemcpyAsync(S->sigt64.p, S->xs->h_sigt.data(), sizeof(double) * nm, hipMemcpyHostToDevice,
                               c.stream));
         PK_HIP(hipStreamSynchronize(c.stream));
      }
   }
   return 0;
}

int pk_sn_sweep_host(pk_sn S, const double* q, double* psi, double* phiw) {
   PK_REQUIRE(S && q, "null argument");
   PK_CALL(sn_prepare(S));
   auto& c = S->sc;
   const long CG = (long)S->C * S->G, CGM = CG * S->M;
   DBuf<double> out;
   PK_REQUIRE(!out.alloc(CGM), "device allocation failed");
   if (S->precision == PK_PREC_MIXED) {
      std::vector<float> qf(q, q + CG);
      PK_HIP(hipMemcpyAsync(S->q32.p, qf.data(), sizeof(float) * CG, hipMemcpyHostToDevice, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      PK_CALL(sn_sweep<float>(S, S->sigt32.p, S->q32.p, S->psi32.p));
      sn_moment<float>(S->psi32.p, S->dpos.p, S->wts.p, S->M, S->C, S->G, S->phiw.p, c.stream);
      sn_to_reference_kernel<float><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi32.p, S->dpos.p, S->C, S->G, S->M, 1.0, out.p);
   } else {
      PK_HIP(hipMemcpyAsync(S->q64.p, q, sizeof(double) * CG, hipMemcpyHostToDevice, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      PK_CALL(sn_sweep<double>(S, S->sigt64.p, S->q64.p, S->psi64.p));
      sn_moment<double>(S->psi64.p, S->dpos.p, S->wts.p, S->M, S->C, S->G, S->phiw.p, c.stream);
      sn_to_reference_kernel<double><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi64.p, S->dpos.p, S->C, S->G, S->M, 1.0, out.p);
   }
   PK_HIP(hipGetLastError());
   PK_HIP(hipStreamSynchronize(c.stream));
   if (psi) PK_HIP(hipMemcpyAsync(psi, out.p, sizeof(double) * CGM, hipMemcpyDeviceToHost, c.stream));
   if (phiw) PK_HIP(hipMemcpyAsync(phiw, S->phiw.p, sizeof(double) * CG, hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

int pk_sn_solve(pk_sn S, double* keff, pk_vec psi, pk_vec phiw_out) {
   PK_REQUIRE(S && keff, "null argument");
   PK_CALL(sn_prepare(S));
   auto& c = S->sc;
   auto t0 = std::chrono::steady_clock::now();
   const int C = S->C, G = S->G, M = S->M;
   const long CG = (long)C * G, CGM = CG * M;
   PK_REQUIRE(!psi || psi->n == CGM, "angular-flux vector size mismatch");
   PK_REQUIRE(!phiw_out || phiw_out->n == CG, "moment vector size mismatch");
   pk_mesh_s* mesh = S->mesh;
   pk_xs_s* xs = S->xs;
   const int gb = grid_for(CG, kBlock, kMaxRedBlocks);
   // warm start from the previous solution of this operator when there is one
   if (!S->has_state) {
      sn_fill_kernel<double><<<grid_for(CG), kBlock, 0, c.stream>>>(S->phiw.p, CG, 1.0);
      S->keff = 1.0;
   }
   double* st = S->st.p;  // {k, prod_old, prod_new, |d|^2, |phi|^2}
   c.host_scalars[0] = S->keff;
   PK_HIP(hipMemcpyAsync(st, c.host_scalars, sizeof(double), hipMemcpyHostToDevice, c.stream));
   sn_prod_kernel<<<gb, kBlock, 0, c.stream>>>(C, G, mesh->vol.p, mesh->mat.p, xs->nusf.p, S->phiw.p,
                                              c.partials, c.ticket, st + 1);
   double k = S->keff;
   double last_scale = 1.0;
   int it = 0;
   S->sweeps = 0;
   bool conv = false;
   // Anderson history: ncol difference columns in a ring of `depth` slots (head
   // = next slot), their Gram matrix H on the host; have_prev once f, g of an
   // earlier iterate are stored
   const int depth = std::min(S->aa_depth, kSnAAMax);
   int ncol = 0, head = 0;
   bool have_prev = false;
   double H[kSnAAMax][kSnAAMax] = {};
   double fbest = 0.0;
   bool use32 = S->precision == PK_PREC_MIXED;
   const double switch64 = std::max(100.0 * S->tol_phi, 1e-7);
   bool last32 = use32;
   for (it = 0; it < S->max_outer; it++) {
      PK_HIP(hipMemcpyAsync(S->old.p, S->phiw.p, sizeof(double) * CG, hipMemcpyDeviceToDevice, c.stream));
      last32 = use32;
      if (use32) {
         sn_source<float>(S->contraction, C, G, mesh, xs, S->phiw.p, st, S->q32.p, c.stream);
         PK_CALL(sn_sweep<float>(S, S->sigt32.p, S->q32.p, S->psi32.p));
         sn_moment<float>(S->psi32.p, S->dpos.p, S->wts.p, M, C, G, S->phiw.p, c.stream);
      } else {
         sn_source<double>(S->contraction, C, G, mesh, xs, S->phiw.p, st, S->q64.p, c.stream);
         PK_CALL(sn_sweep<double>(S, S->sigt64.p, S->q64.p, S->psi64.p));
         sn_moment<double>(S->psi64.p, S->dpos.p, S->wts.p, M, C, G, S->phiw.p, c.stream);
      }
      S->sweeps++;
      sn_prod_kernel<<<gb, kBlock, 0, c.stream>>>(C, G, mesh->vol.p, mesh->mat.p, xs->nusf.p, S->phiw.p,
                                                 c.partials, c.ticket, st + 2);
      sn_update_kernel<<<gb, kBlock, 0, c.stream>>>(CG, S->phiw.p, S->old.p, st, c.partials, c.ticket, st + 3);
      // phiw = g_n (this iterate's image), old = phi_n
      const int slot = head;
      if (depth > 0) {
         sn_aa_diff_kernel<<<grid_for(CG), kBlock, 0, c.stream>>>(CG, S->phiw.p, S->old.p, S->aa_f.p, S->aa_g.p,
                                                                S->aa_df.p + (long)slot * CG,
                                                                S->aa_dg.p + (long)slot * CG, have_prev ? 1 : 0);
         if (have_prev)
            sn_aa_dots_kernel<<<gb, kBlock, 0, c.stream>>>(CG, depth, S->aa_df.p, slot, S->aa_f.p, c.partials,
                                                          c.ticket, st + 8);
      }
      PK_HIP(hipGetLastError());
      PK_HIP(hipMemcpyAsync(c.host_scalars, st, sizeof(double) * (depth > 0 ? 8 + 2 * kSnAAMax : 5),
                            hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      double prod_old = c.host_scalars[1], prod_new = c.host_scalars[2];
      PK_REQUIRE(prod_new > 0.0 && prod_old > 0.0, "zero fission source in the S_N eigen solve");
      double knew = k * prod_new / prod_old;
      last_scale = prod_old / prod_new;
      const double fnorm2 = c.host_scalars[3];
      double dphi = sqrt(fnorm2 / c.host_scalars[4]);
      double dk = fabs(knew - k);
      k = knew;
      c.host_scalars[0] = k;  // prod stays prod_old after rescaling
      PK_HIP(hipMemcpyAsync(st, c.host_scalars, sizeof(double), hipMemcpyHostToDevice, c.stream));
      if (!(k == k)) PK_FAIL("S_N eigen iteration diverged");
      if (dk < S->tol_k && dphi < S->tol_phi) {
         conv = true;
         it++;
         break;
      }
      if (use32 && S->finish64 && dphi < switch64) use32 = false;
      if (depth > 0) {
         // restart the history when the residual grows well past its best
         if (have_prev && fnorm2 > 100.0 * fbest) ncol = 0;
         fbest = (have_prev && ncol > 0) ? std::min(fbest, fnorm2) : fnorm2;
         if (have_prev) {
            const double* v = c.host_scalars + 8;  // <df_j, df_slot>, then <df_j, f>
            for (int j = 0; j < depth; j++) H[j][slot] = H[slot][j] = v[j];
            head = (head + 1) % depth;
            ncol = std::min(ncol + 1, depth);
            // least squares over the ncol newest columns: (H + eps I) gamma = <df, f>
            int idx[kSnAAMax];
            for (int j = 0; j < ncol; j++) idx[j] = (slot - j + depth) % depth;
            double A[kSnAAMax][kSnAAMax + 1];
            double dmax = 0.0;
            for (int a = 0; a < ncol; a++) dmax = std::max(dmax, H[idx[a]][idx[a]]);
            for (int a = 0; a < ncol; a++) {
               for (int b = 0; b < ncol; b++) A[a][b] = H[idx[a]][idx[b]];
               A[a][a] += 1e-12 * dmax;
               A[a][ncol] = v[kSnAAMax + idx[a]];
            }
            bool ok = dmax > 0.0;
            for (int col = 0; ok && col < ncol; col++) {
               int pr = col;
               for (int r = col + 1; r < ncol; r++)
                  if (fabs(A[r][col]) > fabs(A[pr][col])) pr = r;
               if (fabs(A[pr][col]) <= 1e-300) {
                  ok = false;
                  break;
               }
               if (pr != col)
                  for (int q = 0; q <= ncol; q++) std::swap(A[pr][q], A[col][q]);
               for (int r = col + 1; r < ncol; r++) {
                  const double fct = A[r][col] / A[col][col];
                  for (int q = col; q <= ncol; q++) A[r][q] -= fct * A[col][q];
               }
            }
            SnAACoef cf{};
            for (int r = ncol - 1; ok && r >= 0; r--) {
               double acc = A[r][ncol];
               for (int q = r + 1; q < ncol; q++) acc -= A[r][q] * cf.g[idx[q]];
               cf.g[idx[r]] = acc / A[r][r];
               if (!(fabs(cf.g[idx[r]]) < 1e6)) ok = false;
            }
            if (ok)
               sn_aa_mix_kernel<<<grid_for(CG), kBlock, 0, c.stream>>>(CG, depth, S->phiw.p, S->aa_dg.p, cf);
            else
               ncol = 0;
         }
         have_prev = true;
      }
   }
   PK_HIP(hipStreamSynchronize(c.stream));
   S->keff = k;
   S->outer = it;
   S->has_state = true;
   *keff = k;
   // psi of the last sweep, rescaled like phiw (reference layout)
   if (psi) {
      if (last32)
         sn_to_reference_kernel<float><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi32.p, S->dpos.p, C, G, M, last_scale, psi->d.p);
      else
         sn_to_reference_kernel<double><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi64.p, S->dpos.p, C, G, M, last_scale, psi->d.p);
   }
   if (phiw_out) PK_HIP(hipMemcpyAsync(phiw_out->d.p, S->phiw.p, sizeof(double) * CG, hipMemcpyDeviceToDevice, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   S->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   PK_REQUIRE(conv, "S_N eigen iteration did not converge");
   return 0;
}

int pk_sn_get_stats(pk_sn S, int* outer, long* sweeps, double* seconds) {
   PK_REQUIRE(S, "null sweep");
   if (outer) *outer = S->outer;
   if (sweeps) *sweeps = S->sweeps;
   if (seconds) *seconds = S->seconds;
   return 0;
}

int pk_sn_set_contraction(pk_sn S, int mode) {
   PK_REQUIRE(S, "null handle");
   PK_REQUIRE(mode >= 0 && mode <= 2, "contraction mode must be 0, 1 or 2");
   PK_REQUIRE(mode != 2 || S->G == 16, "the matrix-core source contraction takes 16 groups");
   S->contraction = mode;
   return 0;
}

int pk_sn_source_host(pk_sn S, const double* phiw, double keff, double* q) {
   PK_REQUIRE(S && phiw && q, "null argument");
   PK_REQUIRE(keff != 0.0, "keff must be nonzero");
   auto& c = S->sc;
   const long CG = (long)S->C * S->G;
   DBuf<double> dphi, dq, dk;
   PK_REQUIRE(!dphi.alloc(CG) && !dq.alloc(CG) && !dk.alloc(1), "device allocation failed");
   PK_HIP(hipMemcpyAsync(dphi.p, phiw, sizeof(double) * CG, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(dk.p, &keff, sizeof(double), hipMemcpyHostToDevice, c.stream));
   sn_source<double>(S->contraction, S->C, S->G, S->mesh, S->xs, dphi.p, dk.p, dq.p, c.stream);
   PK_HIP(hipGetLastError());
   PK_HIP(hipMemcpyAsync(q, dq.p, sizeof(double) * CG, hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

int pk_sn_time_source(pk_sn S, int reps, double* avg_seconds, double* bytes_per_call) {
   PK_REQUIRE(S && reps > 0, "wrong arguments");
   PK_CALL(sn_prepare(S));
   auto& c = S->sc;
   const int C = S->C, G = S->G;
   const long CG = (long)C * G;
   const bool f32 = S->precision == PK_PREC_MIXED;
   sn_fill_kernel<double><<<grid_for(CG), kBlock, 0, c.stream>>>(S->phiw.p, CG, 1.0);
   sn_fill_kernel<double><<<1, 1, 0, c.stream>>>(S->st.p, 1, 1.0);
   EventPair ev;
   PK_HIP(ev.create());
   for (int r = -1; r < reps; r++) {
      if (r == 0) PK_HIP(hipEventRecord(ev.e0, c.stream));
      if (f32)
         sn_source<float>(S->contraction, C, G, S->mesh, S->xs, S->phiw.p, S->st.p, S->q32.p, c.stream);
      else
         sn_source<double>(S->contraction, C, G, S->mesh, S->xs, S->phiw.p, S->st.p, S->q64.p, c.stream);
   }
   PK_HIP(hipEventRecord(ev.e1, c.stream));
   PK_HIP(hipGetLastError());
   float ms = 0.f;
   PK_HIP(ev.ms(&ms));
   if (avg_seconds) *avg_seconds = ms * 1e-3 / reps;
   // the moment read, the source written, volume and material per cell
   if (bytes_per_call) *bytes_per_call = (double)CG * (8.0 + (f32 ? 4.0 : 8.0)) + (double)C * 12.0;
   S->has_state = false;  // the moment was overwritten: the next solve starts cold
   return 0;
}

int pk_sn_time_sweep(pk_sn S, int reps, double* avg_seconds, double* bytes_per_sweep) {
   PK_REQUIRE(S && reps > 0, "wrong arguments");
   PK_CALL(sn_prepare(S));
   auto& c = S->sc;
   const long CG = (long)S->C * S->G;
   const bool f32 = S->precision == PK_PREC_MIXED;
   if (f32)
      sn_fill_kernel<float><<<grid_for(CG), kBlock, 0, c.stream>>>(S->q32.p, CG, 1.0f);
   else
      sn_fill_kernel<double><<<grid_for(CG), kBlock, 0, c.stream>>>(S->q64.p, CG, 1.0);
   EventPair ev;
   PK_HIP(ev.create());
   for (int r = -1; r < reps; r++) {
      if (r == 0) PK_HIP(hipEventRecord(ev.e0, c.stream));
      if (f32)
         PK_CALL(sn_sweep<float>(S, S->sigt32.p, S->q32.p, S->psi32.p));
      else
         PK_CALL(sn_sweep<double>(S, S->sigt64.p, S->q64.p, S->psi64.p));
   }
   PK_HIP(hipEventRecord(ev.e1, c.stream));
   PK_HIP(hipGetLastError());
   float ms = 0.f;
   PK_HIP(ev.ms(&ms));
   if (avg_seconds) *avg_seconds = ms * 1e-3 / reps;
   if (bytes_per_sweep) *bytes_per_sweep = sweep_bytes(S, f32 ? 4.0 : 8.0);
   return 0;
}

int pk_sn_destroy(pk_sn* S) {
   if (!S || !*S) return 0;
   delete *S;
   *S = nullptr;
   return 0;
}

}  // extern "C"
