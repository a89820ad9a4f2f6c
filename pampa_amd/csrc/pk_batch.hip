// Pooled k-eigenvalue engine: many independent problems R_b x = lambda F_b x
// resident on the device at once, every kernel launch advancing all of them,
// problems admitted and released while others run (continuous batching).
//
// Layout.  The pool owns one row space (in units of 256 rows, one workgroup
// each) and entry spaces for the sliced-ELL values of R, F and the shifted
// operator A = R - F/ks (slice height per problem, from the autotuner).  A problem in slot b owns a range of row units
// (n_b rounded up to 256, so a workgroup never holds two problems) and ranges
// of entries; the columns of its SELL arrays are moved by its row offset, so
// the pool is one block-diagonal system.  unit_slot[] maps a row unit to its
// slot (-1 when free) and kernels return at once for free units.
//
// Every decision is per slot and on the device.  Reductions are segmented
// (pk_reduce.hpp seg_reduce): the workgroups of one slot reduce among
// themselves and the last one runs that slot's serial epilogue.  A slot walks
// a state machine, one "tick" at a time:
//   OUTER  k = <Fx,Fx>/<Fx,Rx>, phi = x/|x|, r = F phi/k - R phi,
//          res = |r|/|F phi/k|; res <= tol -> DONE; after the estimate settles
//          shift ks = k + delta (rebuild A and its block inverses) -> INNER
//   INNER  right-preconditioned GMRES(m) on A e = r (preconditioner: cell-block
//          Jacobi with a sweep over each workgroup's block, in a two-term
//          polynomial), one restart cycle per tick, to |r - A e| <= reduction
//          |r|; then x = phi + e -> OUTER
// A tick is a fixed kernel sequence, captured once as a HIP graph (per grid
// size) and relaunched; the host reads the slots' phases once per tick (a
// fixed batch queues the next tick first).  FP64 outer quantities; the GMRES
// basis, operator copy and block inverses in FP32 (mixed) or FP64, with FP64
// reductions and Givens rotations.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

#include "pk_internal.hpp"
#include "pk_reduce.hpp"
#include "pk_sell.hpp"

namespace pk {

constexpr int kRows = 256;   // rows per unit (= workgroup)
constexpr int kBMaxM = 64;   // largest GMRES restart
constexpr int kAA = 3;       // deepest Anderson history of the outer iterates
constexpr int kAANv = kAA * (kAA + 1) / 2 + kAA;  // its reduced values: Gram entries, right-hand sides
constexpr int kBMaxBs = 16;  // largest preconditioner block

enum Phase { kFree = 0, kOuter = 1, kInner = 2, kDone = 3 };

struct BLayout {
   const int* unit_slot;   // [units]
   const int* wg_first;    // [S] first unit
   const int* wg_count;    // [S] units
   const long* row_off;    // [S]
   const int* nrows;       // [S]
   const int* pbs;         // [S] preconditioner block
   const int* fbs;         // [S] cell block of F
   const long* binv_off;   // [S]
   const int* mlen;        // [S] restart length of the slot (<= the pool's)
   int sweep;              // 1: S includes the workgroup-block sweep, 0: S = B^-1
};

// Sliced-ELL operator of the pool.  Each problem keeps the slice height h of
// its matrix (32, 64 or 128 rows, picked per size class by the autotuner
// below); the table has one 16-byte record per 32-row group g of the pool: the
// offset of the group's first row in its slice, the slice width and h, so
// entry k of row r sits at grp[r / 32].off + r % 32 + h k.
struct alignas(16) SellGroup {
   long off;
   int width, h;
};

struct BSell {
   const SellGroup* grp;
   const int* col;
   const int* rlen;
};

struct SellRef {
   long base;  // position of entry 0 of the row
   int wd, h;  // slice width and height: entry k at base + h k
};

__device__ __forceinline__ SellRef sell_ref(const BSell& A, long row) {
   const SellGroup g = A.grp[row >> 5];
   return SellRef{g.off + (row & 31), g.width, g.h};
}

// The same for a row of the calling wavefront's own 64 rows (row = 64 w +
// lane): the wavefront's two group records are loaded once, at a uniform
// address (two 16-byte loads), and each lane picks its half.
__device__ __forceinline__ SellRef sell_ref_wave(const BSell& A, long row) {
   const long g0 = 2L * __builtin_amdgcn_readfirstlane((int)(row >> 6));
   const SellGroup a = A.grp[g0], b = A.grp[g0 + 1];
   const bool hi = (row & 32) != 0;
   return SellRef{(hi ? b.off : a.off) + (row & 31), hi ? b.width : a.width, hi ? b.h : a.h};
}

struct BState {  // per-slot state (arrays over slots)
   int m;
   // GMRES
   double* H;      // [S][(m+1) m] column-major
   double* cs;     // [S][m]
   double* sn;     // [S][m]
   double* g;      // [S][m+1]
   double* y;      // [S][m]
   double* vs;     // [S][m+1] 1/norm of each stored (unnormalised) basis vector
   double* hd;     // [S][m+2] the step's reduced values (dots and the new vector's norm)
   double* resid;  // [S][2]: estimate, target
   double* ww;     // [S]
   int* flags;     // [S][4]: stop, steps in cycle, breakdown, steps in this solve
   int* jcur;      // [S] the cycle's next step (0 at the cycle start); a cycle may span ticks
   // outer
   double* dots;   // [S][3]: <Fx,Rx>, <Fx,Fx>, <x,x>
   double* rq;     // [S][2]: |r|^2, |F phi/k|^2
   double* k;      // [S]
   double* kprev;  // [S]
   double* ks;     // [S] 0 = no shift
   double* res;    // [S]
   double* dl;     // [S] shift distance of the slot (k_s = k + dl)
   int* phase;     // [S]
   int* start;     // [S] 1: the inner solve starts this tick
   int* rebuild;   // [S] 1: rebuild A and its block inverses this tick
   int* shifted;   // [S]
   int* it;        // [S] corrections done
   int* inner;     // [S] GMRES steps in all corrections
   int* status;    // [S] 0 running, 1 converged, 2 failed
   int* slow;      // [S] consecutive shifted outer updates that reduced the residual by less than 0.8
   int* klast;     // [S][2] GMRES steps of the last two inner solves (the host sizes ticks from them)
   // Anderson mixing of the outer iterates
   double* aa_red;   // [S][kAANv] the normal equations' entries, reduced
   double* aa_gam;   // [S][kAA] mixing coefficients of the last update
   double* aa_best;  // [S] smallest outer residual since the history started
   int* aa_cnt;      // [S] iterates recorded in the history (0: none)
   int* aa_nc;       // [S] coefficients in aa_gam (-1: singular system, the history restarts)
};

struct BOpts {
   double tol, delta, reduction;
   int max_outer, max_inner;
   int poly;  // 1: preconditioner S (2 I - A S) (one Richardson correction), 0: S (precond_apply)
   int sweep = 1;  // 1: S = B^-1 followed by the workgroup-block sweep; 0: S = B^-1
   int aa = kAA;   // Anderson depth of the outer iterates (0: plain update x = phi + e)
};

__device__ __forceinline__ int slot_of_wg(const BLayout& L) { return L.unit_slot[blockIdx.x]; }

__device__ __forceinline__ int slot_of_row(const BLayout& L, long row) { return L.unit_slot[row / kRows]; }



// ---- admission: place a problem's SELL arrays in the pool --------------------------

// table entries of the 32-row groups of a problem from its matrix's SELL-h
// tables: group j (rows 32 j .. 32 j + 31 of the problem) lies in source slice
// 32 j / h at row 32 j mod h of that slice; groups past the matrix's slices
// are empty
__global__ void place_groups_kernel(long ns_src, int h, long ng_dst, const long* soff, const int* swidth,
                                    long dst_g0, long base, SellGroup* dst) {
   const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (j >= ng_dst) return;
   const long s = 32 * j / h;
   SellGroup g;
   if (s < ns_src) {
      g.off = base + (soff[s] - soff[0]) + (32 * j - s * h);
      g.width = swidth[s];
   } else {
      g.off = base;
      g.width = 0;
   }
   g.h = h;
   dst[dst_g0 + j] = g;
}

__global__ void place_cols_kernel(long n, const int* scol, int shift, int* dcol) {
   long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p < n) dcol[p] = scol[p] + shift;
}

__global__ void place_rlen_kernel(long n_src, long n_dst, const int* src, int* dst) {
   long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (r < n_dst) dst[r] = r < n_src ? src[r] : 0;
}

// initial iterate of a slot: x0 (n values) or ones; zero on padding rows
__global__ void place_x_kernel(long n, long padded, const double* x0, double* x) {
   long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (r >= padded) return;
   x[r] = r < n ? (x0 ? x0[r] : 1.0) : 0.0;
}

// ---- shifted operator and block inverses (slots flagged for rebuild) -------------

__global__ void bshift_kernel(long N, BLayout L, BSell R, const double* rval, BSell F,
                              const double* fval, BState S, double* aval, float* aval32) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= N) return;
   const int b = slot_of_row(L, row);
   if (b < 0 || !S.rebuild[b]) return;
   const double inv_ks = S.ks[b] > 0.0 ? 1.0 / S.ks[b] : 0.0;
   const long local = row - L.row_off[b];
   const SellRef ra = sell_ref_wave(R, row), fa = sell_ref_wave(F, row);
   const int fb = L.fbs[b];
   const long cell = local / fb;
   const int len = R.rlen[row];
   for (int k = 0; k < ra.wd; k++) {
      long p = ra.base + (long)k * ra.h;
      double v = rval[p];
      if (k < len && inv_ks != 0.0) {
         long c = (long)R.col[p] - L.row_off[b];
         if (c / fb == cell) v -= fval[fa.base + (c % fb) * fa.h] * inv_ks;
      }
      aval[p] = v;
      if (aval32) aval32[p] = (float)v;
   }
}

// Inverse of each cell block (bs x bs, bs a power of two <= 16, so a block's
// rows are bs consecutive lanes of one wavefront): lane r of the group holds
// row r of [A_cell | I] in registers, and Gauss-Jordan with partial pivoting
// runs across the group's lanes (pivot search, row exchange and elimination
// by cross-lane shuffles; the pivot is the first row of largest magnitude).
// Launched with one wavefront per 64 rows.
template <class T>
__global__ void __launch_bounds__(64) bblock_inverse_kernel(long N, BLayout L, BSell A, const double* aval,
                                                            const int* rebuild, T* binv, int* err) {
   const long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= N) return;
   const int b = slot_of_row(L, row);
   if (b < 0 || !rebuild[b]) return;
   const long local = row - L.row_off[b];
   if (local >= L.nrows[b]) return;  // whole blocks: nrows is a multiple of bs
   const int bs = L.pbs[b];
   const int r = (int)(local & (bs - 1));
   const long cell = local / bs;
   const int base = (int)(threadIdx.x & 63) - r;  // lane of the block's first row
   double a[kBMaxBs], m[kBMaxBs];
#pragma unroll
   for (int q = 0; q < kBMaxBs; q++) {
      a[q] = 0.0;
      m[q] = q == r ? 1.0 : 0.0;
   }
   {
      const SellRef ar = sell_ref(A, row);
      for (int k = 0; k < A.rlen[row]; k++) {
         const long p = ar.base + (long)k * ar.h;
         const long c = (long)A.col[p] - L.row_off[b];
         if (c / bs == cell) {
            const int q0 = (int)(c - cell * bs);
            const double v = aval[p];
#pragma unroll
            for (int q = 0; q < kBMaxBs; q++)
               if (q == q0) a[q] += v;
         }
      }
   }
   bool singular = false;
#pragma unroll
   for (int k = 0; k < kBMaxBs; k++) {
      if (k >= bs) break;
      // pivot: the first row >= k of largest |a[k]|
      double best = r >= k ? fabs(a[k]) : -1.0;
      int piv = r;
      for (int o = 1; o < bs; o <<= 1) {
         const double ob = __shfl_xor(best, o, 64);
         const int op = __shfl_xor(piv, o, 64);
         if (ob > best || (ob == best && op < piv)) {
            best = ob;
            piv = op;
         }
      }
      if (best == 0.0) singular = true;
      // exchange rows k and piv, then scale row k
      const int src = base + (r == k ? piv : (r == piv ? k : r));
#pragma unroll
      for (int q = 0; q < kBMaxBs; q++) {
         a[q] = __shfl(a[q], src, 64);
         m[q] = __shfl(m[q], src, 64);
      }
      if (r == k && !singular) {
         const double inv = 1.0 / a[k];
#pragma unroll
         for (int q = 0; q < kBMaxBs; q++) {
            a[q] *= inv;
            m[q] *= inv;
         }
      }
      // eliminate column k from the other rows
      const double f = a[k];
#pragma unroll
      for (int q = 0; q < kBMaxBs; q++) {
         const double pa = __shfl(a[q], base + k, 64);
         const double pm = __shfl(m[q], base + k, 64);
         if (r != k && f != 0.0) {
            a[q] -= f * pa;
            m[q] -= f * pm;
         }
      }
   }
   if (singular) {
      if (r == 0) *err = 1;
      return;
   }
   T* dst = binv + L.binv_off[b] + cell * bs * bs + (long)r * bs;
#pragma unroll
   for (int q = 0; q < kBMaxBs; q++)
      if (q < bs) dst[q] = (T)m[q];
}

// ---- outer step ------------------------------------------------------------------------

// y1 = R x, y2 = F x and <y2,y1>, <y2,y2>, <x,x> per slot
__global__ void __launch_bounds__(kRows) brf_kernel(BLayout L, BSell R, const double* rval, BSell F,
                                                     const double* fval, const double* x, double* y1,
                                                     double* y2, double* partials, unsigned* tickets, unsigned* gtickets,
                                                     BState S) {
   const int b = slot_of_wg(L);
   if (b < 0 || S.phase[b] != kOuter) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const SellRef ra = sell_ref_wave(R, row), fa = sell_ref_wave(F, row);
   const double a = sell_row_dot<double, double>(R.col, rval, x, ra.base, ra.wd, 0.0, ra.h);
   const double c = sell_row_dot<double, double>(F.col, fval, x, fa.base, fa.wd, 0.0, fa.h);
   y1[row] = a;
   y2[row] = c;
   const double xv = x[row];
   seg_reduce<3>(
       3, [&](int k) { return k == 0 ? c * a : (k == 1 ? c * c : xv * xv); },
       partials + (long)L.wg_first[b] * 3, tickets + b, gtickets + L.wg_first[b], S.dots + 3 * b, (int)blockIdx.x - L.wg_first[b],
       L.wg_count[b]);
}

// phi = x/|x|, r = (y2/k - y1)/|x|; residual; convergence, shift and phase decisions
__global__ void __launch_bounds__(kRows) bresid_kernel(BLayout L, const double* x, const double* y1,
                                                        const double* y2, double* phi, double* r,
                                                        BOpts o, double* partials, unsigned* tickets, unsigned* gtickets,
                                                        BState S) {
   const int b = slot_of_wg(L);
   if (b < 0 || S.phase[b] != kOuter) return;
   const double* d = S.dots + 3 * b;
   const double inv_k = d[0] / d[1];
   const double inv_nx = 1.0 / sqrt(d[2]);
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const double q = y2[row] * inv_k * inv_nx;
   const double v = q - y1[row] * inv_nx;
   phi[row] = x[row] * inv_nx;
   r[row] = v;
   const double rr = v * v, qq = q * q;
   bool last = seg_reduce<2>(
       2, [&](int k) { return k == 0 ? rr : qq; }, partials + (long)L.wg_first[b] * 2, tickets + b, gtickets + L.wg_first[b],
       S.rq + 2 * b, (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   if (!last || threadIdx.x != 0) return;
   // the slot's state, read at once before any store (the stores may alias it)
   const double d0 = d[0], d1 = d[1], rq0 = S.rq[2 * b], rq1 = S.rq[2 * b + 1];
   const double rprev = S.res[b];  // 0 before the first update
   const double dl0 = S.dl[b], kprev = S.kprev[b], ks0 = S.ks[b];
   const int it = S.it[b], shifted = S.shifted[b], slow0 = S.slow[b];
   const double k = d1 / d0;
   const double res = sqrt(rq0 / rq1);
   S.k[b] = k;
   S.res[b] = res;
   if (res <= o.tol) {
      S.status[b] = 1;
      S.phase[b] = kDone;
      return;
   }
   if (!(res == res) || it >= o.max_outer) {
      S.status[b] = 2;
      S.phase[b] = kDone;
      return;
   }
   if (o.aa > 0) {
      // the Anderson history restarts when the residual grows 100x past its best
      const int cnt = S.aa_cnt[b];
      const double best = S.aa_best[b];
      if (cnt == 0 || res < best) S.aa_best[b] = res;
      if (cnt > 0 && res > 100.0 * best) {
         S.aa_cnt[b] = 0;
         S.aa_best[b] = res;
      }
   }
   double ks = ks0;
   if (o.delta > 0.0) {
      // Adaptive Wielandt shift: the outer residual falls by about
      // (1/k - 1/k_s) / (1/k_2 - 1/k_s) per update, which tends to 1 when the
      // dominance ratio k_2/k does; a closer shift takes fewer updates and
      // harder inner solves.  Once shifted, a problem whose last inner solve
      // took at most 24 GMRES steps (cheap) halves its shift distance once,
      // down to 0.05; three shifted updates in a row that reduce the residual
      // by less than 0.8 halve it too (down to 0.004)
      double dl = dl0;
      const int klast = S.klast[2 * b];
      if (shifted && dl > 0.05 && klast > 0 && klast <= 24) {
         dl = fmax(0.5 * dl, 0.05);
         S.dl[b] = dl;
         ks = k + dl;
         S.ks[b] = ks;
         S.rebuild[b] = 1;
      }
      if (shifted && rprev > 0.0) {
         int slow = res > 0.8 * rprev ? slow0 + 1 : 0;
         if (slow >= 3 && dl > 0.004) {
            dl *= 0.5;
            S.dl[b] = dl;
            slow = 0;
            ks = k + dl;
            S.ks[b] = ks;
            S.rebuild[b] = 1;
         }
         S.slow[b] = slow;
      }
      const bool settle = it >= 2 && fabs(k - kprev) < 1e-2 * k;
      if (settle && (!shifted || fabs(k + dl - ks) > 0.2 * dl)) {
         S.ks[b] = k + dl;
         S.shifted[b] = 1;
         S.rebuild[b] = 1;
      }
   }
   S.kprev[b] = k;
   S.it[b] = it + 1;
   S.phase[b] = kInner;
   S.start[b] = 1;
}

// ---- GMRES ------------------------------------------------------------------------------

// The next step of slot b's cycle, or -1 when the slot is not iterating: its
// phase, stop flag, step index and restart length (*mlen) are loaded together
// (one memory round trip).
__device__ __forceinline__ int gmres_step(const BLayout& L, const BState& S, int b, int* mlen) {
   if (b < 0) return -1;
   const int ph = S.phase[b], stop = S.flags[4 * b], j = S.jcur[b];
   *mlen = L.mlen[b];
   return ph == kInner && !stop ? j : -1;
}

// z = B^-1 t for this workgroup's rows of slot b: t of the whole workgroup is
// staged in LDS (a cell block never straddles a workgroup); every thread calls
// it (it holds a barrier), padding rows get 0
// A thread's part in the preconditioner of slot b, looked up once per kernel
// (with the kernel's other first loads, not behind the applies' barriers):
// the row of its cell block's inverse (null for a padding row), the block
// size, the thread of the block's first row, and the first entries of its
// row of A for the workgroup-block sweep.
template <class T>
struct PrecondRow {
   const T* bi;
   int bs, first;
   SellHead<T> head;
};

template <class T>
__device__ __forceinline__ PrecondRow<T> precond_row(const BLayout& L, int b, const T* binv, const BSell& A,
                                                     const T* aval, const SellRef& ar) {
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const long local = row - L.row_off[b];
   const int bs = L.pbs[b], nrows = L.nrows[b];
   const long boff = L.binv_off[b];
   const long cell = local / bs;
   const int gr = (int)(local - cell * bs);
   PrecondRow<T> pr;
   pr.bi = local < nrows ? binv + boff + cell * bs * bs + (long)gr * bs : nullptr;
   pr.bs = bs;
   pr.first = (int)(threadIdx.x - gr);
   pr.head = sell_row_head<T>(A.col, aval, ar.base, L.sweep ? ar.wd : 0, ar.h);
   return pr;
}

template <class T>
__device__ __forceinline__ T cell_block_apply(const PrecondRow<T>& pr, T t, T* tl) {
   tl[threadIdx.x] = t;
   __syncthreads();
   T acc = 0;
   if (pr.bi)
      for (int q = 0; q < pr.bs; q++) acc += pr.bi[q] * tl[pr.first + q];
   return acc;
}

// S t, the preconditioner's building block: z = B^-1 t (cell blocks), then one
// block-Jacobi sweep over the workgroup's own diagonal block of A (the entries
// whose column is one of the workgroup's rows): z += B^-1 (t - A_loc z).  A
// fixed linear map of t (the same for a problem alone or in a mix: its
// workgroups start at its first row).  Every thread calls it (it holds
// barriers); padding rows get 0.  ar: the thread's row of A (sell_ref_wave),
// which the step kernels look up at their start, beside the slot lookup; pr:
// its precond_row.

template <class T>
__device__ __forceinline__ T precond_apply(const BLayout& L, const BSell& A, const T* aval, const SellRef& ar,
                                           const PrecondRow<T>& pr, T t, T* tl, T* xl) {
   const T x = cell_block_apply<T>(pr, t, tl);
   if (!L.sweep) {
      __syncthreads();  // tl is staged again by the caller's next apply
      return x;
   }
   xl[threadIdx.x] = x;
   __syncthreads();
   const long row0 = blockIdx.x * (long)kRows;
   const int wd = ar.wd;
   T acc = 0;
   // entries 4 at a time: their loads first, then the sum in order
   for (int k0 = 0; k0 < wd; k0 += 4) {
      int cc[4];
      T aa[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
         if (k0 == 0) {
            cc[u] = pr.head.c[u];
            aa[u] = pr.head.v[u];
         } else {
            const long p = ar.base + (long)min(k0 + u, wd - 1) * ar.h;
            cc[u] = A.col[p];
            aa[u] = aval[p];
         }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
         const long lc = (long)cc[u] - row0;
         if (k0 + u < wd && lc >= 0 && lc < kRows) acc += aa[u] * xl[lc];
      }
   }
   __syncthreads();  // tl is staged again below
   return x + cell_block_apply<T>(pr, t - acc, tl);
}

// V0 = rhs = r (in T), z = S V0, e = 0, beta = |V0|, target = reduction beta.
// The basis is stored unnormalised: vs[i] = 1/|V_i| scales it where it is used.
template <class T>
__global__ void __launch_bounds__(kRows) bstart_kernel(BLayout L, BSell A, const T* aval, const double* r, T* rhs,
                                                        T* e, T* V0, T* z, const T* binv, BOpts o, double* partials,
                                                        unsigned* tickets, unsigned* gtickets, BState S) {
   const int b = slot_of_wg(L);
   if (b < 0 || S.phase[b] != kInner || !S.start[b]) return;
   __shared__ T tl[kRows], xl[kRows];
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const SellRef ar = sell_ref_wave(A, row);
   const PrecondRow<T> pr = precond_row<T>(L, b, binv, A, aval, ar);
   const T v = (T)r[row];
   rhs[row] = v;
   e[row] = 0;
   V0[row] = v;
   z[row] = precond_apply<T>(L, A, aval, ar, pr, v, tl, xl);
   const double u = (double)v;
   bool last = seg_reduce<1>(
       1, [&](int) { return u * u; }, partials + L.wg_first[b], tickets + b, gtickets + L.wg_first[b], S.ww + b,
       (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   if (!last || threadIdx.x != 0) return;
   const double beta = sqrt(S.ww[b]);
   int* fl = S.flags + 4 * b;
   S.resid[2 * b + 1] = o.reduction * beta;
   S.resid[2 * b] = beta;
   S.g[(long)b * (S.m + 1)] = beta;
   S.vs[(long)b * (S.m + 1)] = beta > 0.0 ? 1.0 / beta : 0.0;
   fl[0] = beta > 0.0 ? 0 : 1;
   fl[1] = 0;
   fl[2] = 0;
   fl[3] = 0;
   S.jcur[b] = 0;
   S.start[b] = 0;
}

// Step j of the Arnoldi process with delayed normalisation.  V_j is stored
// unnormalised and its norm is not known yet: w = A z (z = M^-1 V_j) is stored
// unscaled, and one reduction gives d_i = <vs_i V_i, w> (i < j), d_j = <V_j, w>
// and n_j = <V_j, V_j>.  The last workgroup sets vs_j = 1/sqrt(n_j) (j >= 1;
// vs_0 comes from the cycle start), finishes column j-1 (its subdiagonal is
// sqrt(n_j): Givens rotation, residual estimate, stop test), and unless the
// cycle stops there stores column j: h(i, j) = vs_j d_i, h(j, j) = vs_j^2 d_j.
// At j = mlen (the column past the cycle's last) only n_j is reduced.  What
// the epilogue reads is loaded by every workgroup into local memory before the
// reduction (one value per thread; the last workgroup then has it at hand),
// the reduced values stay in local memory, and the stores run in parallel.
// Profiling of the GMRES step kernels: every workgroup stores its first and
// last device wall-clock reading at its own place (stamps[2 wg], no atomics);
// a one-workgroup kernel then takes the launch's span (earliest start, latest
// end) into a record.  stamps == null outside profiling.
__device__ __forceinline__ unsigned long long stamp_begin(const unsigned long long* stamps) {
   return stamps ? (unsigned long long)wall_clock64() : 0ull;
}
__device__ __forceinline__ void stamp_end(unsigned long long* stamps, unsigned long long t0) {
   if (!stamps) return;
   __syncthreads();
   if (threadIdx.x == 0) {
      stamps[2 * (long)blockIdx.x] = t0;
      stamps[2 * (long)blockIdx.x + 1] = (unsigned long long)wall_clock64();
   }
}

__global__ void __launch_bounds__(256) stamp_span_kernel(const unsigned long long* stamps, int n,
                                                         unsigned long long* rec) {
   unsigned long long lo = ~0ull, hi = 0ull;
   for (int i = threadIdx.x; i < n; i += 256) {
      lo = min(lo, stamps[2 * (long)i]);
      hi = max(hi, stamps[2 * (long)i + 1]);
   }
   __shared__ unsigned long long slo[256], shi[256];
   slo[threadIdx.x] = lo;
   shi[threadIdx.x] = hi;
   __syncthreads();
   for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
         slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + w]);
         shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + w]);
      }
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      rec[0] = slo[0];
      rec[1] = shi[0];
   }
}

// Activity of a slot in the GMRES step kernels, counted by the slot's first
// workgroup when profiling (act != null): [0] full SpMV/dot launches,
// [1] sum of j + 1 over them (basis vectors read), [2] norm-only launches,
// [3] orthogonalisation launches, [4] sum of j + 4 over them, [5] second
// preconditioner terms.  The host turns them into algorithmic bytes.
constexpr int kAct = 6;
constexpr int kRecs = 1 << 14;  // launch spans held on the device before a read-back
__device__ __forceinline__ void count_act(const BLayout& L, int b, long long* act, int k0, long long v0, int k1,
                                          long long v1) {
   if (act && blockIdx.x == (unsigned)L.wg_first[b] && threadIdx.x == 0) {
      act[(long)b * kAct + k0] += v0;
      if (k1 >= 0) act[(long)b * kAct + k1] += v1;
   }
}

template <class T>
__device__ __forceinline__ void spmv_dot_step(long N, const BLayout& L, const BSell& A, const T* aval, const T* z,
                                              T* w, const T* V, double* partials, unsigned* tickets, unsigned* gtickets,
                                              const BState& S, long long* act) {
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const SellRef ar = sell_ref_wave(A, row);  // issued beside the slot lookup
   const int b = slot_of_wg(L);
   int mlen = 0;
   const int j = gmres_step(L, S, b, &mlen);  // read before the reduction; the epilogue advances it
   if (j < 0) return;
   const int m = S.m;
   const double* vs = S.vs + (long)b * (m + 1);
   __shared__ double hd[kBMaxM + 2];
   const bool edge = j == mlen;
   if (edge)
      count_act(L, b, act, 2, 1, -1, 0);
   else
      count_act(L, b, act, 0, 1, 1, j + 1);
   const double vj = (double)V[(long)j * N + row];
   __shared__ double hc[kBMaxM + 2], cc[kBMaxM], ss[kBMaxM], sc[4];
   __shared__ int stop_sh;
   const int ld = m + 1;
   const int tid = threadIdx.x;
   double* Hb = S.H + (long)b * (m + 1) * m;
   const int jj = j - 1;  // the column finished here (none at j = 0)
   if (jj >= 0) {
      if (tid <= jj) hc[tid] = Hb[tid + (long)jj * ld];
      if (tid < jj) {
         cc[tid] = S.cs[(long)b * m + tid];
         ss[tid] = S.sn[(long)b * m + tid];
      }
      if (tid == 128) sc[0] = S.g[(long)b * (m + 1) + jj];
      if (tid == 129) sc[1] = S.resid[2 * b + 1];
   }
   if (tid == 130) sc[2] = vs[0];
   // Every slot reduces in a window of m + 2 partials per workgroup at its own
   // units, whatever its restart length: slots at the edge step (one value)
   // and slots mid-cycle (j + 2 values) share this launch, and windows scaled
   // by the value count would overlap between them.
   double* const part = partials + (long)L.wg_first[b] * (m + 2);
   bool last;
   if (edge) {
      last = seg_reduce<1>(
          1, [&](int) { return vj * vj; }, part, tickets + b, gtickets + L.wg_first[b], hd + j + 1,
          (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   } else {
      // the first batch of the dots' basis entries and scales, requested with
      // the row's entries (ahead of the SpMV's gathers, not after them)
      T vpre[kWaveBatch];
      double spre[kWaveBatch];
#pragma unroll
      for (int u = 0; u < kWaveBatch; u++) {
         const int i = min(u, j);
         vpre[u] = V[(long)i * N + row];
         spre[u] = vs[i];
      }
      const T acc = sell_row_dot<T, T>(A.col, aval, z, ar.base, ar.wd, (T)0, ar.h);
      w[row] = acc;
      const double wr = (double)acc;
      const int nv = j + 2;
      last = seg_reduce<kBMaxM + 2>(
          nv,
          [&](int c) { return c < j ? (double)vpre[c] * spre[c] * wr : (c == j ? vj * wr : vj * vj); },
          [&](int k) {
             return k < j ? (double)V[(long)k * N + row] * vs[k] * wr : (k == j ? vj * wr : vj * vj);
          },
          part, tickets + b, gtickets + L.wg_first[b], hd,
          (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   }
   if (!last) return;
   int* fl = S.flags + 4 * b;
   if (tid == 0) {
      double vsj = sc[2];
      int stop = 0;
      if (jj >= 0) {
         // Givens rotations of column jj (the rotated entry carried in a
         // register, the next one read from local memory), then its own
         const double hn = sqrt(fmax(hd[j + 1], 0.0));
         vsj = hn > 0.0 ? 1.0 / hn : 0.0;
         S.vs[(long)b * (m + 1) + j] = vsj;
         const double gj = sc[0];
         double cur = hc[0];
         for (int i = 0; i < jj; i++) {
            const double nxt = hc[i + 1];
            hc[i] = cc[i] * cur + ss[i] * nxt;
            cur = -ss[i] * cur + cc[i] * nxt;
         }
         const double a = cur, c2 = hn;
         const double rr = sqrt(a * a + c2 * c2);
         const double c = rr > 0.0 ? a / rr : 1.0, sn = rr > 0.0 ? c2 / rr : 0.0;
         hc[jj] = rr;
         hc[jj + 1] = 0.0;
         S.cs[(long)b * m + jj] = c;
         S.sn[(long)b * m + jj] = sn;
         S.g[(long)b * (m + 1) + jj + 1] = -sn * gj;
         S.g[(long)b * (m + 1) + jj] = c * gj;
         const double res = fabs(sn * gj);
         S.resid[2 * b] = res;
         fl[1] = jj + 1;
         const bool breakdown = !(hn > 1e-300) || !(rr > 0.0);
         if (breakdown) fl[2] = 1;
         if (res <= sc[1] || breakdown || jj + 1 == mlen) {
            fl[0] = 1;
            stop = 1;
         }
      }
      sc[3] = vsj;
      stop_sh = stop;
      S.jcur[b] = j + 1;
   }
   __syncthreads();
   // the rotated column jj, then (unless the cycle stops) column j
   if (jj >= 0 && tid <= jj + 1) Hb[tid + (long)jj * ld] = hc[tid];
   if (stop_sh || edge) return;
   const double vsj = sc[3];
   if (tid < j) Hb[tid + (long)j * ld] = vsj * hd[tid];
   if (tid == j) Hb[j + (long)j * ld] = vsj * vsj * hd[j];
}

template <class T>
__global__ void __launch_bounds__(kRows) bspmv_dot_kernel(long N, BLayout L, BSell A, const T* aval,
                                                          const T* z, T* w, const T* V,
                                                          double* partials, unsigned* tickets, unsigned* gtickets, BState S,
                                                          long long* act, unsigned long long* stamps) {
   const unsigned long long t0 = stamp_begin(stamps);
   spmv_dot_step<T>(N, L, A, aval, z, w, V, partials, tickets, gtickets, S, act);
   stamp_end(stamps, t0);
}


// V[j+1] = vs_j w - sum_i h(i,j) vs_i V_i (stored unnormalised), z = S V[j+1];
// its norm is reduced by the next step (spmv_dot_step)
template <class T>
__device__ __forceinline__ void orth_step(long N, const BLayout& L, const BSell& A, const T* aval, const T* w, T* V,
                                          T* z, const T* binv, const BState& S, long long* act) {
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const SellRef ar = sell_ref_wave(A, row);  // issued beside the slot lookup
   const int b = slot_of_wg(L);
   int mlen = 0;
   const int j = gmres_step(L, S, b, &mlen) - 1;  // the step whose SpMV and dots ran last
   if (j < 0) return;
   count_act(L, b, act, 3, 1, 4, j + 4);
   const PrecondRow<T> pr = precond_row<T>(L, b, binv, A, aval, ar);
   __shared__ T tl[kRows], xl[kRows];
   const int m = S.m;
   const double* h = S.H + (long)b * (m + 1) * m + (long)j * (m + 1);
   const double* vs = S.vs + (long)b * (m + 1);
   T t = (T)vs[j] * w[row];
   // basis vectors 8 at a time: their loads issue together, the sum runs in order
   for (int i0 = 0; i0 <= j; i0 += 8) {
      T cf[8], vv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
         const int i = min(i0 + u, j);
         cf[u] = (T)(h[i] * vs[i]);
         vv[u] = V[(long)i * N + row];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
         if (i0 + u <= j) t -= cf[u] * vv[u];
   }
   V[(long)(j + 1) * N + row] = t;
   z[row] = precond_apply<T>(L, A, aval, ar, pr, t, tl, xl);
}

template <class T>
__global__ void __launch_bounds__(kRows) borth_kernel(long N, BLayout L, BSell A, const T* aval, const T* w, T* V,
                                                      T* z, const T* binv, BState S, long long* act,
                                                      unsigned long long* stamps) {
   const unsigned long long t0 = stamp_begin(stamps);
   orth_step<T>(N, L, A, aval, w, V, z, binv, S, act);
   stamp_end(stamps, t0);
}

// y = H^-1 g, one workgroup per slot that took steps this cycle.  H is loaded
// into local memory 4 values per thread at a time; the triangular solve runs
// in one wave, lane r owning row r: from the last column back, the column's
// unknown is solved by its row's lane and broadcast, and the rows above
// subtract its contribution (column order: a fixed order of operations).
__global__ void __launch_bounds__(kRows) btrisolve_kernel(int nslots, BState S) {
   const int b = blockIdx.x;
   if (b >= nslots || S.phase[b] != kInner || !S.flags[4 * b]) return;  // mid-cycle slots go on
   const int m = S.m;
   const int k = S.flags[4 * b + 1];
   if (k == 0) return;
   __shared__ double h[kBMaxM * kBMaxM];
   const double* Hb = S.H + (long)b * (m + 1) * m;
   const int kk = k * k;
   for (int p0 = threadIdx.x; p0 < kk; p0 += 4 * kRows) {
      double v[4];
      int pp[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
         pp[u] = min(p0 + u * kRows, kk - 1);
         const int i = pp[u] % k, q = pp[u] / k;
         v[u] = Hb[i + (long)q * (m + 1)];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
         if (p0 + u * kRows < kk) h[pp[u] % k + (pp[u] / k) * kBMaxM] = v[u];
   }
   const int r = threadIdx.x;
   const double gr0 = r < k ? S.g[(long)b * (m + 1) + r] : 0.0;
   __syncthreads();
   if (r >= 64) return;
   double gr = gr0, yr = 0.0;
   for (int i = k - 1; i >= 0; i--) {
      const double dd = h[i + i * kBMaxM];
      const double gi = __shfl(gr, i, 64);
      const double yi = dd != 0.0 ? gi / dd : 0.0;
      if (r == i) yr = yi;
      if (r < i) gr -= h[r + i * kBMaxM] * yi;
   }
   if (r < k) S.y[(long)b * m + r] = yr;
}

// Second term of the preconditioner M^-1 = S (2 I - A S):
// z = z0 + S (v - A z0), z0 = S v given.  mode 0: slots iterating in
// this GMRES step, v = V_j at the slot's step j (v points at the basis, N apart);
// mode 1: slots whose cycle stopped after taking steps (the update: also e += z).
template <class T>
__device__ __forceinline__ void poly_step(long N, const BLayout& L, const BSell& A, const T* aval, const T* v,
                                          const T* z0, const T* binv, T* z, const BState& S, int mode, T* e,
                                          long long* act);

template <class T>
__global__ void __launch_bounds__(kRows) bpoly_kernel(long N, BLayout L, BSell A, const T* aval, const T* v,
                                                      const T* z0, const T* binv, T* z, BState S, int mode, T* e,
                                                      long long* act, unsigned long long* stamps) {
   const unsigned long long t0 = stamp_begin(stamps);
   poly_step<T>(N, L, A, aval, v, z0, binv, z, S, mode, e, act);
   stamp_end(stamps, t0);
}

template <class T>
__device__ __forceinline__ void poly_step(long N, const BLayout& L, const BSell& A, const T* aval, const T* v,
                                          const T* z0, const T* binv, T* z, const BState& S, int mode, T* e,
                                          long long* act) {
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const SellRef ar = sell_ref_wave(A, row);  // issued beside the slot lookup
   const int b = slot_of_wg(L);
   if (mode == 0) {
      int mlen = 0;
      const int j = gmres_step(L, S, b, &mlen);
      if (j < 0 || j >= mlen) return;  // at j = mlen (the column past the cycle's last): norm only
      v += (long)j * N;
      count_act(L, b, act, 5, 1, -1, 0);
   } else if (b < 0 || S.phase[b] != kInner || !S.flags[4 * b] || S.flags[4 * b + 1] == 0) {
      return;
   }
   const PrecondRow<T> pr = precond_row<T>(L, b, binv, A, aval, ar);
   __shared__ T tl[kRows], xl[kRows];
   const T az = sell_row_dot<T, T>(A.col, aval, z0, ar.base, ar.wd, (T)0, ar.h);
   const T c = precond_apply<T>(L, A, aval, ar, pr, v[row] - az, tl, xl);
   const T zr = z0[row] + c;
   z[row] = zr;
   if (mode == 1) e[row] += zr;
}

// u = sum_i y_i vs_i V[i]; accumulate: e += S u, else u and S u are
// written out for the second preconditioner term (e += M^-1 u in two kernels)
template <class T>
__global__ void __launch_bounds__(kRows) bupdate_kernel(long N, BLayout L, BSell A, const T* aval, const T* V,
                                                        const T* binv, BState S, T* e, int accumulate, T* uout,
                                                        T* zout) {
   const int b = slot_of_wg(L);
   if (b < 0 || S.phase[b] != kInner || !S.flags[4 * b]) return;  // mid-cycle slots go on
   const int k = S.flags[4 * b + 1];
   if (k == 0) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const long local = row - L.row_off[b];
   const SellRef ar = sell_ref_wave(A, row);
   const PrecondRow<T> pr = precond_row<T>(L, b, binv, A, aval, ar);
   __shared__ T u[kRows], xl[kRows];
   T acc = 0;
   const double* vs = S.vs + (long)b * (S.m + 1);
   for (int i0 = 0; i0 < k; i0 += 8) {
      T cf[8], vv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
         const int i = min(i0 + u, k - 1);
         cf[u] = (T)(S.y[(long)b * S.m + i] * vs[i]);
         vv[u] = V[(long)i * N + row];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
         if (i0 + u < k) acc += cf[u] * vv[u];
   }
   const T s = precond_apply<T>(L, A, aval, ar, pr, acc, u, xl);
   if (accumulate) {
      if (local < L.nrows[b]) e[row] += s;
   } else {
      uout[row] = acc;
      zout[row] = s;
   }
}

// true residual V[0] = rhs - A e after the cycle, z = S V[0], and the stop decision
template <class T>
__global__ void __launch_bounds__(kRows) brestart_kernel(BLayout L, BSell A, const T* aval, const T* e,
                                                         const T* rhs, T* V0, T* z, const T* binv, BOpts o,
                                                         double* partials, unsigned* tickets, unsigned* gtickets, BState S) {
   const int b = slot_of_wg(L);
   // only slots whose cycle stopped after taking steps: the others are mid-cycle
   // (it goes on next tick) or stopped before this cycle took a step (their
   // flags stay)
   if (b < 0 || S.phase[b] != kInner || !S.flags[4 * b] || S.flags[4 * b + 1] == 0) return;
   __shared__ T tl[kRows], xl[kRows];
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const SellRef ar = sell_ref_wave(A, row);
   const PrecondRow<T> pr = precond_row<T>(L, b, binv, A, aval, ar);
   const T acc = sell_row_dot<T, T>(A.col, aval, e, ar.base, ar.wd, (T)0, ar.h);
   const T rv = rhs[row] - acc;
   V0[row] = rv;
   z[row] = precond_apply<T>(L, A, aval, ar, pr, rv, tl, xl);
   const double u = (double)rv;
   bool last = seg_reduce<1>(
       1, [&](int) { return u * u; }, partials + L.wg_first[b], tickets + b, gtickets + L.wg_first[b], S.ww + b,
       (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   if (!last || threadIdx.x != 0) return;
   const double beta = sqrt(S.ww[b]);
   int* fl = S.flags + 4 * b;
   fl[3] += fl[1];
   S.g[(long)b * (S.m + 1)] = beta;
   S.resid[2 * b] = beta;
   S.vs[(long)b * (S.m + 1)] = beta > 0.0 ? 1.0 / beta : 0.0;
   fl[0] = (beta <= S.resid[2 * b + 1] || !(beta > 0.0) || fl[3] >= o.max_inner) ? 1 : 0;
   fl[1] = 0;
   fl[2] = 0;
   S.jcur[b] = 0;
}

// x = phi + e for slots whose inner solve ended
template <class T>
__global__ void bfinish_kernel(long N, BLayout L, const double* phi, const T* e, BState S, double* x) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= N) return;
   const int b = slot_of_row(L, row);
   if (b < 0 || S.phase[b] != kInner || !S.flags[4 * b]) return;
   x[row] = phi[row] + (double)e[row];
}

// ---- Anderson mixing of the outer iterates ----------------------------------------------
// The outer update is the fixed-point map phi -> g = phi + e (e the inner
// solve's correction), whose residual is f = g - phi = e.  Anderson mixing
// (Walker and Ni 2011) of depth d takes x = g - sum_i gamma_i dG_i, gamma the
// least-squares solution of min |f - sum_i gamma_i dF_i| over the last d
// differences dF_i, dG_i of successive f and g.  A slot mixes only while
// shifted, and its history restarts when the outer residual grows 100x past
// its best (bresid) or the small system is singular.  The fixed point, so
// the eigenpair, is that of the plain update.

struct BAnderson {
   double* df;  // [kAA][N] differences of f, a ring filled in order
   double* dg;  // [kAA][N] differences of g
   double* f;   // [N] f of the last recorded iterate
   double* g;   // [N] g of the last recorded iterate
   long N;
};

// index of Gram entry (i, j), i <= j < kAA, among the reduced values
__host__ __device__ constexpr int aa_pair(int i, int j) { return i * kAA - i * (i - 1) / 2 + (j - i); }

// For slots whose inner solve ended this tick (shifted ones): record f and g,
// write the newest differences into the ring, reduce the normal equations
// dF^T dF gamma = dF^T f; the slot's last workgroup solves them (Cholesky, in
// a fixed order).  A pivot below 1e-12 of the largest diagonal entry marks
// the system singular: no mixing this time, and the history restarts.
template <class T>
__global__ void __launch_bounds__(kRows) baa_dots_kernel(BLayout L, const double* phi, const T* e, BAnderson H,
                                                         int depth, double* partials, unsigned* tickets,
                                                         unsigned* gtickets, BState S) {
   const int b = slot_of_wg(L);
   if (b < 0 || S.phase[b] != kInner || !S.flags[4 * b] || !S.shifted[b]) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const int cnt = S.aa_cnt[b];
   const double f = (double)e[row], g = phi[row] + f;
   const bool mix = cnt >= 1;
   const int nd = mix ? min(cnt, depth) : 0;
   const int q = mix ? (cnt - 1) % depth : 0;
   double d[kAA];
#pragma unroll
   for (int i = 0; i < kAA; i++) d[i] = 0.0;
   if (mix) {
      const double dfn = f - H.f[row], dgn = g - H.g[row];
#pragma unroll
      for (int i = 0; i < kAA; i++)
         if (i < nd) d[i] = i == q ? dfn : H.df[(long)i * H.N + row];
      H.df[(long)q * H.N + row] = dfn;
      H.dg[(long)q * H.N + row] = dgn;
   }
   H.f[row] = f;
   H.g[row] = g;
   if (!mix) return;
   const bool last = seg_reduce<kAANv>(
       kAANv,
       [&](int k) {
          if (k >= kAANv - kAA) return d[k - (kAANv - kAA)] * f;
          int i = 0;
          while (k >= kAA - i) {
             k -= kAA - i;
             i++;
          }
          return d[i] * d[i + k];
       },
       partials + (long)L.wg_first[b] * kAANv, tickets + b, gtickets + L.wg_first[b], S.aa_red + (long)b * kAANv,
       (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   if (!last || threadIdx.x != 0) return;
   const double* red = S.aa_red + (long)b * kAANv;
   double a[kAA][kAA], r[kAA], gam[kAA];
   double dmax = 0.0;
   for (int i = 0; i < nd; i++) {
      for (int j = 0; j < nd; j++) a[i][j] = red[aa_pair(min(i, j), max(i, j))];
      r[i] = red[kAANv - kAA + i];
      dmax = fmax(dmax, a[i][i]);
   }
   // Cholesky a = L L^T in place (lower triangle), then the two triangular solves
   bool ok = dmax > 0.0 && dmax < 1e300;
   for (int j = 0; j < nd && ok; j++) {
      double djj = a[j][j];
      for (int p = 0; p < j; p++) djj -= a[j][p] * a[j][p];
      if (!(djj > 1e-12 * dmax)) {
         ok = false;
         break;
      }
      a[j][j] = sqrt(djj);
      for (int i = j + 1; i < nd; i++) {
         double v = a[i][j];
         for (int p = 0; p < j; p++) v -= a[i][p] * a[j][p];
         a[i][j] = v / a[j][j];
      }
   }
   if (ok) {
      for (int i = 0; i < nd; i++) {
         double v = r[i];
         for (int p = 0; p < i; p++) v -= a[i][p] * gam[p];
         gam[i] = v / a[i][i];
      }
      for (int i = nd - 1; i >= 0; i--) {
         double v = gam[i];
         for (int p = i + 1; p < nd; p++) v -= a[p][i] * gam[p];
         gam[i] = v / a[i][i];
      }
      for (int i = 0; i < nd; i++) ok = ok && gam[i] == gam[i] && fabs(gam[i]) < 1e300;
   }
   for (int i = 0; i < kAA; i++) S.aa_gam[(long)b * kAA + i] = ok && i < nd ? gam[i] : 0.0;
   S.aa_nc[b] = ok ? nd : -1;
}

// x = g - sum_i gamma_i dG_i (x = g where the slot does not mix), for slots
// whose inner solve ended this tick
template <class T>
__global__ void __launch_bounds__(kRows) baa_mix_kernel(BLayout L, const double* phi, const T* e, BAnderson H,
                                                        BState S, double* x) {
   const int b = slot_of_wg(L);
   if (b < 0 || S.phase[b] != kInner || !S.flags[4 * b]) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   double v = phi[row] + (double)e[row];
   if (S.shifted[b] && S.aa_cnt[b] >= 1) {
      const int nc = S.aa_nc[b];
      for (int i = 0; i < nc; i++) v -= S.aa_gam[(long)b * kAA + i] * H.dg[(long)i * H.N + row];
   }
   x[row] = v;
}

// y = A x over the units of live slots (timing of the pool's operator)
template <class T>
__global__ void __launch_bounds__(kRows) bspmv_kernel(BLayout L, BSell A, const T* aval, const T* x, T* y) {
   const int b = slot_of_wg(L);
   if (b < 0) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const SellRef ar = sell_ref_wave(A, row);
   y[row] = sell_row_dot<T, T>(A.col, aval, x, ar.base, ar.wd, (T)0, ar.h);
}

__global__ void bphase_kernel(int nslots, BState S) {
   int b = blockIdx.x * blockDim.x + threadIdx.x;
   if (b >= nslots) return;
   S.rebuild[b] = 0;  // consumed by this tick's shift and block inverses
   if (S.phase[b] != kInner || !S.flags[4 * b]) return;
   S.inner[b] += S.flags[4 * b + 3];
   S.klast[2 * b + 1] = S.klast[2 * b];
   S.klast[2 * b] = S.flags[4 * b + 3];
   // the Anderson history: one more iterate recorded while shifted (restarted
   // at this one after a singular system), none before the shift
   const int c = S.aa_cnt[b];
   S.aa_cnt[b] = S.shifted[b] ? (c >= 1 && S.aa_nc[b] < 0 ? 1 : c + 1) : 0;
   S.phase[b] = kOuter;
}

// A slot's tables and initial state, written at its admission, then its
// units mapped to it.
struct SlotInit {
   int first, count, nrows, fbs, pbs, mlen;
   long row_off, binv_off;
   double delta;
};

__global__ void __launch_bounds__(kRows) slot_init_kernel(int b, SlotInit v, int* wg_first, int* wg_count,
                                                          int* nrows, int* fbs, int* pbs, int* mlen, long* row_off,
                                                          long* binv_off, BState S, int* unit_slot, long u0,
                                                          long units) {
   if (threadIdx.x == 0) {
      wg_first[b] = v.first;
      wg_count[b] = v.count;
      nrows[b] = v.nrows;
      fbs[b] = v.fbs;
      pbs[b] = v.pbs;
      mlen[b] = v.mlen;
      row_off[b] = v.row_off;
      binv_off[b] = v.binv_off;
      for (int q = 0; q < 4; q++) S.flags[4 * b + q] = 0;
      S.start[b] = 0;
      S.shifted[b] = 0;
      S.it[b] = 0;
      S.inner[b] = 0;
      S.status[b] = 0;
      S.ks[b] = 0.0;
      S.kprev[b] = 0.0;
      S.res[b] = 0.0;
      S.dl[b] = v.delta;
      S.slow[b] = 0;
      S.jcur[b] = 0;
      S.klast[2 * b] = 0;
      S.klast[2 * b + 1] = 0;
      S.aa_cnt[b] = 0;
      S.aa_nc[b] = 0;
      S.aa_best[b] = 0.0;
      S.rebuild[b] = 1;
      S.phase[b] = kOuter;
   }
   for (long u = threadIdx.x; u < units; u += kRows) unit_slot[u0 + u] = b;
}

// A released slot: FREE, its units unmapped.
__global__ void __launch_bounds__(kRows) slot_release_kernel(int b, BState S, int* unit_slot, long u0, long units) {
   if (threadIdx.x == 0) S.phase[b] = kFree;
   for (long u = threadIdx.x; u < units; u += kRows) unit_slot[u0 + u] = -1;
}

// Every slot's phase, cycle position (next step, steps of the inner solve's
// earlier cycles) and the longer of its last two inner solves, then the error
// flag, into pinned host memory (one workgroup; kStatus values per slot), for a
// host that reads them after an event instead of a copy.
constexpr int kStatus = 4;
__global__ void __launch_bounds__(kRows) bstatus_kernel(int nslots, BState S, const int* err, int* host) {
   for (int b = threadIdx.x; b < nslots; b += kRows) {
      const int v[kStatus] = {S.phase[b], S.jcur[b], S.flags[4 * b + 3],
                              max(S.klast[2 * b], S.klast[2 * b + 1])};
      for (int q = 0; q < kStatus; q++)
         __hip_atomic_store(host + kStatus * b + q, v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
   }
   if (threadIdx.x == 0)
      __hip_atomic_store(host + kStatus * nslots, *err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace pk

using namespace pk;

namespace {

// First-fit allocator of [0, cap) in whole units, with coalescing.
struct RangeAlloc {
   long cap = 0;
   std::map<long, long> free_;  // start -> length
   void reset(long c) {
      cap = c;
      free_.clear();
      if (c > 0) free_[0] = c;
   }
   long alloc(long len) {
      if (len <= 0) return 0;
      for (auto it = free_.begin(); it != free_.end(); ++it) {
         if (it->second >= len) {
            long start = it->first, rest = it->second - len;
            free_.erase(it);
            if (rest > 0) free_[start + len] = rest;
            return start;
         }
      }
      return -1;
   }
   void release(long start, long len) {
      if (len <= 0) return;
      auto it = free_.emplace(start, len).first;
      auto nx = std::next(it);
      if (nx != free_.end() && it->first + it->second == nx->first) {
         it->second += nx->second;
         free_.erase(nx);
      }
      if (it != free_.begin()) {
         auto pv = std::prev(it);
         if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
         }
      }
   }
   long high_water() const {  // end of the highest allocated unit
      if (free_.empty()) return cap;
      auto last = std::prev(free_.end());
      return last->first + last->second == cap ? last->first : cap;
   }
   long largest_free() const {
      long m = 0;
      for (auto& kv : free_) m = std::max(m, kv.second);
      return m;
   }
};

struct Slot {
   bool used = false, harvested = false;
   long n = 0;
   int bs = 1, pbs = 1;
   int mlen = 0;               // restart length
   long unit0 = 0, units = 0;  // row units
   long re0 = 0, re = 0;       // R/A entries (x64)
   long fe0 = 0, fe = 0;       // F entries (x64)
   long bi0 = 0, bi = 0;       // block-inverse values (x64)
   long nnz = 0;               // stored entries of R (algorithmic count)
   long ticket = 0;            // admission order
};

template <class T>
struct Inner {
   DBuf<T> V, z, z0, w, e, rhs, binv, aval;
};

// The captured tick: one executable per (grid, step launches).
struct Graphs {
   std::map<std::pair<long, int>, hipGraphExec_t> exec;
   void clear() {
      for (auto& e : exec) (void)hipGraphExecDestroy(e.second);
      exec.clear();
   }
   ~Graphs() { clear(); }
};

}  // namespace

struct pk_batch_s {
   // capacity
   long max_units = 0, max_rentries = 0, max_fentries = 0, max_binv = 0;
   int max_slots = 0;
   int m = 60;
   int precision = PK_PREC_FP64;
   BOpts opts{1e-8, 0.1, 0.1, 10000, 2000, 1};
   // device arrays
   DBuf<int> unit_slot, wg_first, wg_count, nrows, pbs, fbs, mlen;
   DBuf<long> row_off, binv_off;
   DBuf<SellGroup> r_grp, f_grp;  // per 32-row group
   DBuf<int> r_col, f_col, rlen;
   DBuf<double> r_val, f_val, a_val;
   DBuf<double> x, phi, y1, y2, r;
   DBuf<double> aa_df, aa_dg, aa_f, aa_g;  // Anderson history of the outer iterates
   DBuf<double> sd;
   DBuf<int> si;
   DBuf<double> part;
   DBuf<unsigned> tick;
   DBuf<unsigned> gtick;  // group counters of the two-level reductions, one per row unit
   DBuf<int> err;
   Inner<double> i64;
   Inner<float> i32;
   BState S{};
   // host bookkeeping
   RangeAlloc ua, ra, fa, ba;
   std::vector<Slot> slots;
   std::vector<int> h_unit_slot;
   std::vector<int> h_phase;
   long admitted = 0;
   Graphs graph;
   // pinned read-back of the phases and the error flag (two sets in turn) and
   // their events, for the fixed-batch run
   int* h_status = nullptr;
   hipEvent_t status_ev[2] = {nullptr, nullptr};
   ~pk_batch_s() {
      for (auto e : status_ev)
         if (e) (void)hipEventDestroy(e);
      if (h_status) (void)hipHostFree(h_status);
   }
   DBuf<long long> act;  // per-slot activity of the GMRES step kernels (profiling)
   // profiled launch spans: per-workgroup stamps of the current launch, and
   // records (start, end) of launches not yet read back, with their kernel
   DBuf<unsigned long long> stamps, recs;
   std::vector<int> rec_kind;
   int nrec = 0;
   double clock_hz = 0.0;
   double tick_seconds = 0.0;
   long ticks = 0;
   // fixed-batch use (pk_batch_create): members and optional initial vectors
   std::vector<pk_mat_s*> fixR, fixF;
   std::vector<DBuf<double>> fix_init;
   std::vector<char> fix_has_init;
   std::vector<int> fix_slot;
   double seconds = 0.0;

   BLayout layout() const {
      return BLayout{unit_slot.p, wg_first.p, wg_count.p, row_off.p, nrows.p, pbs.p, fbs.p, binv_off.p, mlen.p,
                     opts.sweep};
   }
   BSell rsell() const { return BSell{r_grp.p, r_col.p, rlen.p}; }
   BSell fsell() const { return BSell{f_grp.p, f_col.p, rlen.p}; }
};

static int preconditioner_block(int bs) {
   // largest power of two <= 16 dividing the cell block, or the block itself
   // when 256-row workgroups hold whole cells of it
   int p = 1;
   while (p * 2 <= kBMaxBs && bs % (p * 2) == 0) p *= 2;
   if (bs <= kBMaxBs && kRows % bs == 0) p = bs;
   return p;
}

static inline long units_of(long n) { return std::max(1L, (n + kRows - 1) / kRows); }
static inline long blocks64(long n) { return (n + 63) / 64; }

// device bytes of a pool of the given capacity (units of 256 rows, entries and
// block-inverse values in units of 64)
static double pool_bytes(long units, long rent, long fent, long binv, int slots, int m, int precision) {
   const double rows = (double)units * kRows, ng = rows / 32.0;
   const double t = precision == PK_PREC_MIXED ? 4.0 : 8.0;
   double b = 0.0;
   b += units * 4.0 + ng * 2 * (8.0 + 4.0 + 4.0) + rows * 4.0;  // unit map, group tables, rlen
   b += rent * 64.0 * (4.0 + 8.0 + 8.0) + fent * 64.0 * 12.0;    // R cols+vals, A (FP64), F
   b += rows * 8.0 * 5;                                          // x, phi, y1, y2, r
   b += rows * 8.0 * (2 * kAA + 2);                              // Anderson history
   b += rows * t * (m + 1 + 5) + binv * 64.0 * t;                // basis, z z0 w e rhs, block inverses
   if (precision == PK_PREC_MIXED) b += rent * 64.0 * t;         // A copy in FP32
   b += (double)units * std::max(m + 4, kAANv) * 8.0;            // reduction partials
   b += (double)slots * ((m + 1.0) * m + 6.0 * m + 40.0) * 8.0;  // per-slot state
   return b;
}

static int pool_alloc(pk_batch_s* P) {
   Ctx& c = ctx();
   const long units = P->max_units, rows = units * kRows, sl = rows / 32;
   const int S = P->max_slots, m = P->m;
   const long rent = P->max_rentries * 64, fent = P->max_fentries * 64, binv = P->max_binv * 64;
   bool bad = P->unit_slot.alloc(units) || P->wg_first.alloc(S) || P->wg_count.alloc(S) || P->nrows.alloc(S) ||
              P->mlen.alloc(S) ||
              P->pbs.alloc(S) || P->fbs.alloc(S) || P->row_off.alloc(S) || P->binv_off.alloc(S) ||
              P->r_grp.alloc(sl) || P->f_grp.alloc(sl) ||
              P->rlen.alloc(rows) || P->r_col.alloc(rent) || P->r_val.alloc(rent) || P->a_val.alloc(rent) ||
              P->f_col.alloc(fent) || P->f_val.alloc(fent) || P->x.alloc(rows) || P->phi.alloc(rows) ||
              P->y1.alloc(rows) || P->y2.alloc(rows) || P->r.alloc(rows) ||
              P->aa_df.alloc((size_t)kAA * rows) || P->aa_dg.alloc((size_t)kAA * rows) || P->aa_f.alloc(rows) ||
              P->aa_g.alloc(rows) ||
              P->part.alloc((size_t)units * std::max(m + 4, kAANv)) || P->tick.alloc(S) || P->gtick.alloc(units) ||
              P->err.alloc(1);
   PK_REQUIRE(!bad, "device allocation failed for the eigen pool");
   auto alloc_inner = [&](auto& I, bool own_a) {
      return I.V.alloc((size_t)(m + 1) * rows) || I.z.alloc(rows) || I.z0.alloc(rows) || I.w.alloc(rows) ||
             I.e.alloc(rows) ||
             I.rhs.alloc(rows) || I.binv.alloc(binv) || (own_a && I.aval.alloc(rent));
   };
   if (P->precision == PK_PREC_MIXED)
      PK_REQUIRE(!alloc_inner(P->i32, true), "device allocation failed for the eigen pool");
   else
      PK_REQUIRE(!alloc_inner(P->i64, false), "device allocation failed for the eigen pool");
   // per-slot state
   const size_t nd = (size_t)S * ((size_t)(m + 1) * m + 6 * (size_t)m + 40);
   const size_t ni = (size_t)S * 20;
   PK_REQUIRE(!P->sd.alloc(nd) && !P->si.alloc(ni), "device allocation failed for the eigen pool");
   PK_HIP(hipMemsetAsync(P->sd.p, 0, sizeof(double) * nd, c.stream));
   PK_HIP(hipMemsetAsync(P->si.p, 0, sizeof(int) * ni, c.stream));
   double* d = P->sd.p;
   BState& B = P->S;
   B.m = m;
   auto take = [&](size_t k) {
      double* p = d;
      d += k;
      return p;
   };
   B.H = take((size_t)S * (m + 1) * m);
   B.cs = take((size_t)S * m);
   B.sn = take((size_t)S * m);
   B.g = take((size_t)S * (m + 1));
   B.y = take((size_t)S * m);
   B.vs = take((size_t)S * (m + 1));
   B.hd = take((size_t)S * (m + 2));
   B.resid = take(2 * (size_t)S);
   B.ww = take(S);
   B.dots = take(3 * (size_t)S);
   B.rq = take(2 * (size_t)S);
   B.k = take(S);
   B.kprev = take(S);
   B.ks = take(S);
   B.res = take(S);
   B.dl = take(S);
   B.aa_red = take((size_t)S * kAANv);
   B.aa_gam = take((size_t)S * kAA);
   B.aa_best = take(S);
   int* q = P->si.p;
   auto takei = [&](size_t k) {
      int* p = q;
      q += k;
      return p;
   };
   B.flags = takei(4 * (size_t)S);
   B.phase = takei(S);
   B.start = takei(S);
   B.rebuild = takei(S);
   B.shifted = takei(S);
   B.it = takei(S);
   B.inner = takei(S);
   B.status = takei(S);
   B.slow = takei(S);
   B.jcur = takei(S);
   B.klast = takei(2 * (size_t)S);
   B.aa_cnt = takei(S);
   B.aa_nc = takei(S);
   // tables: every unit free, every slice empty
   P->h_unit_slot.assign(units, -1);
   PK_HIP(hipMemcpyAsync(P->unit_slot.p, P->h_unit_slot.data(), sizeof(int) * units, hipMemcpyHostToDevice,
                         c.stream));
   // zero width everywhere (the height is read only where a width is set)
   PK_HIP(hipMemsetAsync(P->r_grp.p, 0, sizeof(SellGroup) * sl, c.stream));
   PK_HIP(hipMemsetAsync(P->f_grp.p, 0, sizeof(SellGroup) * sl, c.stream));
   PK_HIP(hipMemsetAsync(P->rlen.p, 0, sizeof(int) * rows, c.stream));
   PK_HIP(hipMemsetAsync(P->x.p, 0, sizeof(double) * rows, c.stream));
   PK_HIP(hipMemsetAsync(P->phi.p, 0, sizeof(double) * rows, c.stream));
   PK_HIP(hipMemsetAsync(P->tick.p, 0, sizeof(unsigned) * S, c.stream));
   PK_HIP(hipMemsetAsync(P->gtick.p, 0, sizeof(unsigned) * units, c.stream));
   PK_HIP(hipMemsetAsync(P->err.p, 0, sizeof(int), c.stream));
   auto zero_inner = [&](auto& I, size_t t) {
      return hipMemsetAsync(I.V.p, 0, t * (m + 1) * rows, c.stream) ||
             hipMemsetAsync(I.z.p, 0, t * rows, c.stream) || hipMemsetAsync(I.z0.p, 0, t * rows, c.stream) ||
             hipMemsetAsync(I.w.p, 0, t * rows, c.stream) ||
             hipMemsetAsync(I.e.p, 0, t * rows, c.stream) || hipMemsetAsync(I.rhs.p, 0, t * rows, c.stream);
   };
   if (P->precision == PK_PREC_MIXED)
      PK_REQUIRE(!zero_inner(P->i32, 4), "clearing the pool failed");
   else
      PK_REQUIRE(!zero_inner(P->i64, 8), "clearing the pool failed");
   P->ua.reset(units);
   P->ra.reset(P->max_rentries);
   P->fa.reset(P->max_fentries);
   P->ba.reset(P->max_binv);
   P->slots.assign(S, Slot());
   P->h_phase.assign(S, 0);
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

// ---- admission / release ---------------------------------------------------------------

static int check_member(pk_mat_s* R, pk_mat_s* F) {
   PK_REQUIRE(R && F && R->assembled && F->assembled, "matrices not assembled");
   PK_REQUIRE(R->n == F->n && R->bs == F->bs, "R and F differ in size or block size");
   PK_REQUIRE(F->nnz == R->n * (long)R->bs, "F must hold one full cell block per row");
   PK_REQUIRE(R->n > 0, "empty problem");
   return 0;
}

// ---- autotuner: slice height and SpMV workgroup per size class ---------------------
//
// The first problem of a size class (rows, cell block, nonzeros, precision)
// admitted to a pool times the sliced-ELL SpMV of its operator (FP32 values in
// mixed precision, the inner solve's type) for slice heights 32, 64 and 128
// and workgroups of 128, 256, 512 and 1024 rows, 1 + 20 launches each on the
// library stream; later problems of the class take the stored choice.  The
// slice height is chosen from the 256-row workgroups, the shape of the pool's
// kernels (their workgroups stay 256 rows: a workgroup is also the block of
// the preconditioner's sweep and the unit of the segmented reductions, so its
// size is part of the arithmetic); the workgroup size, for the plain SpMV of
// the matrix, from the chosen height's row.  A candidate replaces the default
// (64 rows, 256 rows) only when it is at least 3 % faster: small operators
// time at the launch latency, where the differences are noise.  A height
// whose padded size exceeds that of the 64-row layout is skipped (pool
// capacities are sized with that layout).  Padding entries add exact zeros:
// every choice gives the same bits.  PK_TUNE=0 turns the tuner off;
// PK_SLICE=32|64|128 forces a height.
struct TuneKey {
   long n, nnz;
   int bs, precision;
   bool operator<(const TuneKey& o) const {
      if (n != o.n) return n < o.n;
      if (nnz != o.nnz) return nnz < o.nnz;
      if (bs != o.bs) return bs < o.bs;
      return precision < o.precision;
   }
};
struct TuneChoice {
   int h = 64, block = 256;
   double us[3][4] = {};  // [height 32/64/128][workgroup 128/256/512/1024], 0 = skipped
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tune;
static std::vector<TuneKey> g_tune_order;

static int env_int(const char* name, int dflt) {
   const char* e = getenv(name);
   return e && *e ? atoi(e) : dflt;
}

static int tune_operator(pk_mat_s* R, int precision) {
   const int forced = env_int("PK_SLICE", 0);
   if (forced) {
      PK_REQUIRE(forced == 32 || forced == 64 || forced == 128, "PK_SLICE must be 32, 64 or 128");
      if (R->sell.h != forced) PK_CALL(build_sell(R, forced));
      if (R->sell.padded > R->sell.pad64) PK_CALL(build_sell(R, 64));
      return 0;
   }
   if (!env_int("PK_TUNE", 1)) return 0;
   static const int hs[3] = {32, 64, 128}, blocks[4] = {128, 256, 512, 1024};
   const TuneKey key{R->n, R->nnz, R->bs, precision};
   std::lock_guard<std::mutex> lock(g_tune_mu);
   auto it = g_tune.find(key);
   if (it == g_tune.end()) {
      const bool f32 = precision == PK_PREC_MIXED;
      TuneChoice ch;
      const long pad64 = R->sell.pad64;
      for (int a = 0; a < 3; a++) {
         if (R->sell.h != hs[a]) PK_CALL(build_sell(R, hs[a]));
         if (R->sell.padded > pad64) continue;
         for (int q = 0; q < 4; q++) {
            double s = 0.0;
            PK_CALL(time_spmv(R, f32, blocks[q], 20, &s));
            ch.us[a][q] = s * 1e6;
         }
      }
      constexpr double kGain = 0.97;
      int ah = 1;  // 64 rows
      for (int a = 0; a < 3; a++)
         if (ch.us[a][1] > 0.0 && ch.us[a][1] < kGain * ch.us[1][1] && ch.us[a][1] < ch.us[ah][1]) ah = a;
      int qb = 1;  // 256 rows
      for (int q = 0; q < 4; q++)
         if (ch.us[ah][q] < kGain * ch.us[ah][1] && ch.us[ah][q] < ch.us[ah][qb]) qb = q;
      ch.h = hs[ah];
      ch.block = blocks[qb];
      it = g_tune.emplace(key, ch).first;
      g_tune_order.push_back(key);
   }
   if (R->sell.h != it->second.h) PK_CALL(build_sell(R, it->second.h));
   R->sell.spmv_block = it->second.block;
   return 0;
}

// 0: admitted into *slot; 2: no room now (nothing changed); 1: error
static int pool_admit(pk_batch_s* P, pk_mat_s* R, pk_mat_s* F, const double* x0_dev, int* slot) {
   Ctx& c = ctx();
   PK_CALL(check_member(R, F));
   PK_CALL(tune_operator(R, P->precision));
   const long n = R->n;
   const int bs = R->bs, pbs = preconditioner_block(bs);
   const long units = units_of(n);
   const long re = blocks64(R->sell.padded), fe = blocks64(F->sell.padded), bi = blocks64(n * pbs);
   PK_REQUIRE(units <= P->max_units && re <= P->max_rentries && fe <= P->max_fentries && bi <= P->max_binv,
              "problem larger than the eigen pool");
   int sidx = -1;
   for (int s = 0; s < P->max_slots; s++)
      if (!P->slots[s].used) {
         sidx = s;
         break;
      }
   if (sidx < 0) return 2;
   long u0 = P->ua.alloc(units);
   if (u0 < 0) return 2;
   long r0 = P->ra.alloc(re);
   if (r0 < 0) {
      P->ua.release(u0, units);
      return 2;
   }
   long f0 = P->fa.alloc(fe);
   if (f0 < 0) {
      P->ua.release(u0, units);
      P->ra.release(r0, re);
      return 2;
   }
   long b0 = P->ba.alloc(bi);
   if (b0 < 0) {
      P->ua.release(u0, units);
      P->ra.release(r0, re);
      P->fa.release(f0, fe);
      return 2;
   }
   Slot& S = P->slots[sidx];
   S = Slot();
   S.used = true;
   S.n = n;
   S.bs = bs;
   S.pbs = pbs;
   S.unit0 = u0;
   S.units = units;
   S.re0 = r0;
   S.re = re;
   S.fe0 = f0;
   S.fe = fe;
   S.bi0 = b0;
   S.bi = bi;
   S.nnz = R->sell.nnz;
   S.ticket = P->admitted++;
   const long row0 = u0 * kRows, padded = units * kRows;
   const long g0 = row0 / 32, ng = padded / 32;
   // SELL arrays of R and F, columns moved by the row offset
   place_groups_kernel<<<grid_for(ng), kBlock, 0, c.stream>>>(R->sell.nslices, R->sell.h, ng, R->sell.off.p,
                                                             R->sell.width.p, g0, r0 * 64, P->r_grp.p);
   place_groups_kernel<<<grid_for(ng), kBlock, 0, c.stream>>>(F->sell.nslices, F->sell.h, ng, F->sell.off.p,
                                                             F->sell.width.p, g0, f0 * 64, P->f_grp.p);
   if (R->sell.padded) {
      place_cols_kernel<<<grid_for(R->sell.padded), kBlock, 0, c.stream>>>(R->sell.padded, R->sell.col.p,
                                                                          (int)row0, P->r_col.p + r0 * 64);
      PK_HIP(hipMemcpyAsync(P->r_val.p + r0 * 64, R->sell.val.p, sizeof(double) * R->sell.padded,
                            hipMemcpyDeviceToDevice, c.stream));
   }
   if (F->sell.padded) {
      place_cols_kernel<<<grid_for(F->sell.padded), kBlock, 0, c.stream>>>(F->sell.padded, F->sell.col.p,
                                                                          (int)row0, P->f_col.p + f0 * 64);
      PK_HIP(hipMemcpyAsync(P->f_val.p + f0 * 64, F->sell.val.p, sizeof(double) * F->sell.padded,
                            hipMemcpyDeviceToDevice, c.stream));
   }
   place_rlen_kernel<<<grid_for(padded), kBlock, 0, c.stream>>>(n, padded, R->sell.rlen.p, P->rlen.p + row0);
   place_x_kernel<<<grid_for(padded), kBlock, 0, c.stream>>>(n, padded, x0_dev, P->x.p + row0);
   PK_HIP(hipGetLastError());
   // slot tables
   const int i_first = (int)u0, i_count = (int)units, i_n = (int)n, i_bs = bs, i_pbs = pbs;
   // restart length: the pool's, except 30 for cell blocks of 16 groups and
   // more, whose stronger block-Jacobi blocks gain too few iterations from a
   // longer basis to pay for it (measured on the 256 x 256 instances)
   const int i_mlen = bs >= 16 ? std::min(P->m, 30) : P->m;
   P->slots[sidx].mlen = i_mlen;
   const long l_row0 = row0, l_b0 = b0 * 64;
   // the slot's tables and state (OUTER, rebuild A = R and the block inverses
   // on the first tick), then its units, in one launch
   SlotInit v;
   v.first = i_first;
   v.count = i_count;
   v.nrows = i_n;
   v.fbs = i_bs;
   v.pbs = i_pbs;
   v.mlen = i_mlen;
   v.row_off = l_row0;
   v.binv_off = l_b0;
   v.delta = P->opts.delta;
   slot_init_kernel<<<1, kRows, 0, c.stream>>>(sidx, v, P->wg_first.p, P->wg_count.p, P->nrows.p, P->fbs.p, P->pbs.p,
                                              P->mlen.p, P->row_off.p, P->binv_off.p, P->S, P->unit_slot.p, u0, units);
   PK_HIP(hipGetLastError());
   for (long u = u0; u < u0 + units; u++) P->h_unit_slot[u] = sidx;
   // the admission reads the caller's operators and initial vector: done
   // before returning
   PK_HIP(hipStreamSynchronize(c.stream));
   P->h_phase[sidx] = kOuter;
   *slot = sidx;
   return 0;
}

static const char* const kStepKernel[3] = {"gmres_spmv_dot", "gmres_orth", "gmres_poly"};

// Profiling: the spans recorded so far into the profiler (launches, seconds).
static int flush_spans(pk_batch_s* P) {
   if (P->nrec == 0) return 0;
   Ctx& c = ctx();
   std::vector<unsigned long long> h(2 * (size_t)P->nrec);
   PK_HIP(hipMemcpyAsync(h.data(), P->recs.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                         c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   for (int r = 0; r < P->nrec; r++)
      prof_add(kStepKernel[P->rec_kind[r]], 1, (double)(h[2 * r + 1] - h[2 * r]) / P->clock_hz, 0.0);
   P->nrec = 0;
   return 0;
}

// Profiling: the slots' GMRES activity since the last fold turned into the
// step kernels' algorithmic bytes (value and column of every stored entry of
// the inner operator, the basis vectors each step reads and writes, the
// block inverses), added to the profiler's records; counters cleared.
static int fold_activity(pk_batch_s* P) {
   if (!profiling() || !P->act.p) return 0;
   PK_CALL(flush_spans(P));
   Ctx& c = ctx();
   std::vector<long long> h((size_t)P->max_slots * kAct);
   PK_HIP(hipMemcpyAsync(h.data(), P->act.p, sizeof(long long) * h.size(), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipMemsetAsync(P->act.p, 0, sizeof(long long) * h.size(), c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   const double s = P->precision == PK_PREC_MIXED ? 4.0 : 8.0;
   double bdot = 0.0, borth = 0.0, bpoly = 0.0;
   for (int b = 0; b < P->max_slots; b++) {
      const long long* a = &h[(size_t)b * kAct];
      if (!P->slots[b].used) continue;
      const double n = (double)P->slots[b].n, nnz = (double)P->slots[b].nnz, bsn = n * P->slots[b].pbs;
      const double ent = nnz * (s + 4.0);
      bdot += a[0] * (ent + 2.0 * n * s) + (a[1] + a[2]) * n * s;
      borth += a[3] * (bsn * s + (P->opts.sweep ? ent : 0.0)) + a[4] * n * s;
      bpoly += a[5] * (ent + 3.0 * n * s + bsn * s);
   }
   prof_add("gmres_spmv_dot", 0, 0.0, bdot);
   prof_add("gmres_orth", 0, 0.0, borth);
   if (P->opts.poly) prof_add("gmres_poly", 0, 0.0, bpoly);
   return 0;
}

static int pool_release(pk_batch_s* P, int sidx) {
   Ctx& c = ctx();
   PK_REQUIRE(sidx >= 0 && sidx < P->max_slots && P->slots[sidx].used, "no such slot");
   PK_CALL(fold_activity(P));
   Slot& S = P->slots[sidx];
   for (long u = S.unit0; u < S.unit0 + S.units; u++) P->h_unit_slot[u] = -1;
   slot_release_kernel<<<1, kRows, 0, c.stream>>>(sidx, P->S, P->unit_slot.p, S.unit0, S.units);
   PK_HIP(hipGetLastError());
   P->ua.release(S.unit0, S.units);
   P->ra.release(S.re0, S.re);
   P->fa.release(S.fe0, S.fe);
   P->ba.release(S.bi0, S.bi);
   S = Slot();
   P->h_phase[sidx] = kFree;
   return 0;
}

// ---- one tick ------------------------------------------------------------------------------

template <class T>
static Inner<T>& inner_of(pk_batch_s* P);
template <>
Inner<double>& inner_of<double>(pk_batch_s* P) {
   return P->i64;
}
template <>
Inner<float>& inner_of<float>(pk_batch_s* P) {
   return P->i32;
}

// Launches one tick on the stream for row units [0, units): the outer part,
// `steps` GMRES step launches (1 .. m + 1; each slot takes the next step of its
// cycle, and a cycle longer than the tick goes on in the next one), then the
// end of the cycles that stopped.  With `prof`, the GMRES step kernels carry
// profiler events and activity counters (the same kernels, arguments and
// launches as the captured graph).
template <class T>
static int launch_tick(pk_batch_s* P, long units, int steps, bool prof) {
   Ctx& c = ctx();
   Inner<T>& I = inner_of<T>(P);
   const BLayout L = P->layout();
   const BSell Rs = P->rsell(), Fs = P->fsell();
   const long N = P->max_units * kRows;  // vector stride of the basis
   const long rows = units * kRows;
   const int nwg = (int)units, m = P->m;
   const int ns = P->max_slots;
   const BOpts o = P->opts;
   const BState S = P->S;
   double* part = P->part.p;
   unsigned* tk = P->tick.p;
   unsigned* gtk = P->gtick.p;
   const bool fp32 = sizeof(T) != sizeof(double);
   const T* aval = fp32 ? I.aval.p : (const T*)P->a_val.p;
   brf_kernel<<<nwg, kRows, 0, c.stream>>>(L, Rs, P->r_val.p, Fs, P->f_val.p, P->x.p, P->y1.p, P->y2.p, part, tk, gtk,
                                           S);
   bresid_kernel<<<nwg, kRows, 0, c.stream>>>(L, P->x.p, P->y1.p, P->y2.p, P->phi.p, P->r.p, o, part, tk, gtk, S);
   // the shifted operator (and its FP32 copy in mixed pools) of slots being
   // rebuilt; their rebuild flags are cleared by the tick's last kernel
   float* a32 = nullptr;
   if constexpr (sizeof(T) == sizeof(float)) a32 = I.aval.p;
   bshift_kernel<<<grid_for(rows), kBlock, 0, c.stream>>>(rows, L, Rs, P->r_val.p, Fs, P->f_val.p, S,
                                                          P->a_val.p, a32);
   bblock_inverse_kernel<T><<<grid_for(rows, 64), 64, 0, c.stream>>>(rows, L, Rs, P->a_val.p, S.rebuild,
                                                                     I.binv.p, P->err.p);
   // with the polynomial preconditioner the first term S t goes to z0 and
   // the step's bpoly kernel completes z; without it, straight into z
   T* zfirst = o.poly ? I.z0.p : I.z.p;
   bstart_kernel<T><<<nwg, kRows, 0, c.stream>>>(L, Rs, aval, P->r.p, I.rhs.p, I.e.p, I.V.p, zfirst, I.binv.p, o,
                                                 part, tk, gtk, S);
   // `steps` step launches; a slot at step j of its cycle runs the
   // preconditioner, the SpMV and dots, and the orthogonalisation of step j,
   // or at j = mlen only the norm of V_mlen that finishes column mlen - 1.  At
   // launch m every slot is at its cycle's last column or stopped, so only the
   // SpMV/dot kernel is issued there.  With `prof` the step kernels are
   // launched with events that record their own start and end, and count each
   // slot's activity (algorithmic bytes, summed by the host afterwards); every
   // step is issued, as in the captured graph.
   long long* act = prof ? P->act.p : nullptr;
   unsigned long long* st = prof ? P->stamps.p : nullptr;
   // the span of a profiled launch into the next record
   auto span = [&](int which) -> int {
      if (!prof) return 0;
      if (P->nrec == kRecs) PK_CALL(flush_spans(P));
      stamp_span_kernel<<<1, 256, 0, c.stream>>>(P->stamps.p, nwg, P->recs.p + 2 * (long)P->nrec);
      P->rec_kind[P->nrec++] = which;
      return 0;
   };
   for (int q = 0; q < steps; q++) {
      const bool full = q < m;
      if (o.poly && full) {
         bpoly_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Rs, aval, I.V.p, I.z0.p, I.binv.p, I.z.p, S, 0, nullptr,
                                                      act, st);
         PK_CALL(span(2));
      }
      bspmv_dot_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Rs, aval, I.z.p, I.w.p, I.V.p, part, tk, gtk, S, act,
                                                       st);
      PK_CALL(span(0));
      if (full) {
         borth_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Rs, aval, I.w.p, I.V.p, zfirst, I.binv.p, S, act, st);
         PK_CALL(span(1));
      }
   }
   btrisolve_kernel<<<ns, kRows, 0, c.stream>>>(ns, S);
   if (o.poly) {
      // e += M^-1 u: u and S u, then the second term, then the sum
      bupdate_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Rs, aval, I.V.p, I.binv.p, S, I.e.p, 0, I.w.p, I.z0.p);
      bpoly_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Rs, aval, I.w.p, I.z0.p, I.binv.p, I.z.p, S, 1, I.e.p,
                                                   nullptr, nullptr);
   } else {
      bupdate_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Rs, aval, I.V.p, I.binv.p, S, I.e.p, 1, nullptr,
                                                     nullptr);
   }
   brestart_kernel<T><<<nwg, kRows, 0, c.stream>>>(L, Rs, aval, I.e.p, I.rhs.p, I.V.p, zfirst, I.binv.p, o, part,
                                                   tk, gtk, S);
   if (o.aa > 0) {
      const BAnderson H{P->aa_df.p, P->aa_dg.p, P->aa_f.p, P->aa_g.p, N};
      baa_dots_kernel<T><<<nwg, kRows, 0, c.stream>>>(L, P->phi.p, I.e.p, H, o.aa, part, tk, gtk, S);
      baa_mix_kernel<T><<<nwg, kRows, 0, c.stream>>>(L, P->phi.p, I.e.p, H, S, P->x.p);
   } else {
      bfinish_kernel<T><<<grid_for(rows), kBlock, 0, c.stream>>>(rows, L, P->phi.p, I.e.p, S, P->x.p);
   }
   bphase_kernel<<<grid_for(ns), kBlock, 0, c.stream>>>(ns, S);
   PK_HIP(hipGetLastError());
   return 0;
}

// PK_NO_GRAPH=1 in the environment launches the tick's kernels one by one
// instead of through the captured graph (same kernels and arguments), for
// tools that trace individual launches.
static bool graphs_disabled() {
   static const bool off = [] {
      const char* e = getenv("PK_NO_GRAPH");
      return e && *e && *e != '0';
   }();
   return off;
}

template <class T>
static int run_tick(pk_batch_s* P, long units, int steps) {
   Ctx& c = ctx();
   if (profiling()) {
      if (!P->act.p) {
         int khz = 0;
         PK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device));
         PK_REQUIRE(khz > 0, "no device wall clock for profiling");
         P->clock_hz = khz * 1e3;
         PK_REQUIRE(!P->act.alloc((size_t)P->max_slots * kAct) && !P->stamps.alloc(2 * (size_t)P->max_units) &&
                        !P->recs.alloc(2 * (size_t)kRecs),
                    "device allocation failed");
         P->rec_kind.assign(kRecs, 0);
         P->nrec = 0;
         PK_HIP(hipMemsetAsync(P->act.p, 0, sizeof(long long) * P->max_slots * kAct, c.stream));
      }
      return launch_tick<T>(P, units, steps, true);
   }
   if (graphs_disabled()) return launch_tick<T>(P, units, steps, false);
   const std::pair<long, int> key(units, steps);
   auto it = P->graph.exec.find(key);
   if (it == P->graph.exec.end()) {
      // every quantity that varies between ticks is read from device memory,
      // so one graph serves every tick of this grid and length; a batch whose
      // members change size keeps at most kMaxGraphs of them, dropped once the
      // stream has finished the ticks in flight
      constexpr size_t kMaxGraphs = 64;
      if (P->graph.exec.size() >= kMaxGraphs) {
         PK_HIP(hipStreamSynchronize(c.stream));
         P->graph.clear();
      }
      hipGraph_t g = nullptr;
      PK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      int rc = launch_tick<T>(P, units, steps, false);
      hipError_t e = hipStreamEndCapture(c.stream, &g);
      if (rc) {
         if (g) (void)hipGraphDestroy(g);
         return rc;
      }
      PK_HIP(e);
      hipGraphExec_t x = nullptr;
      hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      PK_HIP(ie);
      it = P->graph.exec.emplace(key, x).first;
   }
   PK_HIP(hipGraphLaunch(it->second, c.stream));
   return 0;
}

// Runs `ticks` ticks (none when nothing is live), then reads every slot's
// phase.  *live = slots still iterating, *done = DONE slots not yet collected.
static int pool_step(pk_batch_s* P, int ticks, int* live, int* done) {
   Ctx& c = ctx();
   const long units = P->ua.high_water();
   int nlive = 0;
   for (int s = 0; s < P->max_slots; s++)
      if (P->slots[s].used && (P->h_phase[s] == kOuter || P->h_phase[s] == kInner)) nlive++;
   auto t0 = std::chrono::steady_clock::now();
   if (nlive > 0 && units > 0)
      for (int t = 0; t < ticks; t++) {
         if (P->precision == PK_PREC_MIXED)
            PK_CALL(run_tick<float>(P, units, P->m + 1));
         else
            PK_CALL(run_tick<double>(P, units, P->m + 1));
         P->ticks++;
      }
   int herr = 0;
   PK_HIP(hipMemcpyAsync(P->h_phase.data(), P->S.phase, sizeof(int) * P->max_slots, hipMemcpyDeviceToHost,
                         c.stream));
   PK_HIP(hipMemcpyAsync(&herr, P->err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   P->tick_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   PK_REQUIRE(herr == 0, "singular diagonal block in the block-Jacobi preconditioner");
   int nl = 0, nd = 0;
   for (int s = 0; s < P->max_slots; s++) {
      if (!P->slots[s].used) continue;
      if (P->h_phase[s] == kDone) {
         if (!P->slots[s].harvested) nd++;
      } else {
         nl++;
      }
   }
   if (live) *live = nl;
   if (done) *done = nd;
   return 0;
}

struct SlotResult {
   double k = 0.0, res = 0.0;
   int it = 0, inner = 0, status = 0;
};

static int slot_result(pk_batch_s* P, int sidx, SlotResult& R, pk_vec_s* x) {
   Ctx& c = ctx();
   PK_REQUIRE(sidx >= 0 && sidx < P->max_slots && P->slots[sidx].used, "no such slot");
   PK_REQUIRE(P->h_phase[sidx] == kDone, "slot has not finished");
   PK_HIP(hipMemcpyAsync(&R.k, P->S.k + sidx, sizeof(double), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipMemcpyAsync(&R.res, P->S.res + sidx, sizeof(double), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipMemcpyAsync(&R.it, P->S.it + sidx, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipMemcpyAsync(&R.inner, P->S.inner + sidx, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipMemcpyAsync(&R.status, P->S.status + sidx, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   if (x) {
      PK_REQUIRE(x->n == P->slots[sidx].n, "vector size mismatch");
      PK_HIP(hipMemcpyAsync(x->d.p, P->phi.p + P->slots[sidx].unit0 * kRows, sizeof(double) * x->n,
                            hipMemcpyDeviceToDevice, c.stream));
   }
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

// ---- fixed batch: every member admitted at once and run to the end ------------------

// Step launches tick t + 1 needs, from every live slot's state after tick
// t - 1 (st: kStatus values per slot) and the launches lt of tick t, assuming
// that each inner solve takes as many GMRES steps as the longer of the slot's
// last two (a whole cycle while none is known).  A cycle that stops after c
// columns is seen by its launch c; a slot whose cycle ends in tick t starts its
// next cycle (a restart, or a new inner solve after the outer update) in tick
// t + 1.  The most any slot needs, plus two, at most m + 1.  Only the tick
// length depends on it: a cycle longer than its tick goes on in the next.
static int predict_steps(const pk_batch_s* P, const int* st, int lt) {
   constexpr int kMargin = 2;
   const int m1 = P->m + 1;
   int need = 1;
   for (int b = 0; b < P->max_slots; b++) {
      const Slot& sl = P->slots[b];
      const int* v = st + kStatus * b;
      if (!sl.used || (v[0] != kOuter && v[0] != kInner)) continue;
      const int mlen = sl.mlen, K = v[3];
      if (K <= 0) return m1;
      // an outer slot starts a new inner solve in tick t
      int j = v[0] == kOuter ? 0 : v[1];
      int done = v[0] == kOuter ? 0 : v[2];
      auto stop = [&](int d) { return std::min(mlen, std::max(1, K - d)); };
      // the column the current cycle stops at: not before its next launch
      const int c = std::min(mlen, std::max(stop(done), j));
      const int left = c - j + 1;  // launches to the end of the current cycle
      int n;
      if (lt < left) {
         n = left - lt;
      } else {
         done += c;
         n = stop(done >= K ? 0 : done) + 1;
      }
      need = std::max(need, n + kMargin);
   }
   return std::min(need, m1);
}

// Ticks run one ahead of the host's read of the slots' phases: tick t + 1 is
// queued before the host waits for tick t's phases (a tick over slots that are
// all DONE changes nothing), so the device does not idle during the read-back.
// The phases come back through pinned memory, two sets in turn, with what the
// length of the next tick is predicted from.
static int fixed_run(pk_batch_s* P) {
   Ctx& c = ctx();
   const int ns = P->max_slots;
   const long units = P->ua.high_water();
   if (units <= 0) return 0;
   const int set = kStatus * ns + 1;
   if (!P->h_status) {
      PK_HIP(hipHostMalloc((void**)&P->h_status, sizeof(int) * 2 * set));
      for (auto& e : P->status_ev) PK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
   }
   auto t0 = std::chrono::steady_clock::now();
   auto issue = [&](int k, int steps) -> int {
      int* h = P->h_status + k * set;
      for (int q = 0; q < set; q++) h[q] = -1;
      if (P->precision == PK_PREC_MIXED)
         PK_CALL(run_tick<float>(P, units, steps));
      else
         PK_CALL(run_tick<double>(P, units, steps));
      P->ticks++;
      bstatus_kernel<<<1, kRows, 0, c.stream>>>(ns, P->S, P->err.p, h);
      PK_HIP(hipGetLastError());
      PK_HIP(hipEventRecord(P->status_ev[k], c.stream));
      return 0;
   };
   const int m1 = P->m + 1;
   // PK_TICK_STEPS=n runs every tick with n step launches (1 .. m + 1; the same
   // bits for every n; for tests)
   const int forced = std::min(env_int("PK_TICK_STEPS", 0), m1);
   std::vector<int> st;  // the last status read
   int last_steps = forced > 0 ? forced : m1;  // launches of the last tick queued
   PK_CALL(issue(0, last_steps));
   for (int cur = 0;; cur ^= 1) {
      const int steps = forced > 0 ? forced : st.empty() ? m1 : predict_steps(P, st.data(), last_steps);
      PK_CALL(issue(cur ^ 1, steps));
      last_steps = steps;
      PK_HIP(hipEventSynchronize(P->status_ev[cur]));
      const int* h = P->h_status + cur * set;
      st.resize(set);
      for (int q = 0; q < set; q++) st[q] = __atomic_load_n(&h[q], __ATOMIC_ACQUIRE);
      const int herr = st[kStatus * ns];
      int nl = 0;
      for (int q = 0; q < ns; q++) {
         const int ph = st[kStatus * q];
         P->h_phase[q] = ph;
         if (P->slots[q].used && ph != kDone) nl++;
      }
      if (herr != 0 || nl == 0) {
         PK_HIP(hipStreamSynchronize(c.stream));
         P->tick_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
         PK_REQUIRE(herr == 0, "singular diagonal block in the block-Jacobi preconditioner");
         return fold_activity(P);
      }
   }
}


static int fixed_solve(pk_batch_s* P) {
   const int B = (int)P->fixR.size();
   for (int s = 0; s < P->max_slots; s++)
      if (P->slots[s].used) PK_CALL(pool_release(P, s));
   P->fix_slot.assign(B, -1);
   for (int b = 0; b < B; b++) {
      int slot = -1;
      const double* x0 = P->fix_has_init[b] ? P->fix_init[b].p : nullptr;
      int rc = pool_admit(P, P->fixR[b], P->fixF[b], x0, &slot);
      PK_REQUIRE(rc != 2, "batch does not fit its pool");
      PK_CALL(rc);
      P->fix_slot[b] = slot;
   }
   return fixed_run(P);
}

template <class T>
static int time_pool_spmv(pk_batch_s* P, int reps, double* avg_seconds) {
   Ctx& c = ctx();
   Inner<T>& I = inner_of<T>(P);
   const T* aval = sizeof(T) == sizeof(double) ? (const T*)P->a_val.p : I.aval.p;
   const long units = P->ua.high_water();
   EventPair ev;
   PK_HIP(ev.create());
   // one untimed launch, then `reps` between the events
   bspmv_kernel<T><<<(int)units, kRows, 0, c.stream>>>(P->layout(), P->rsell(), aval, I.z.p, I.w.p);
   PK_HIP(hipEventRecord(ev.e0, c.stream));
   for (int r = 0; r < reps; r++)
      bspmv_kernel<T><<<(int)units, kRows, 0, c.stream>>>(P->layout(), P->rsell(), aval, I.z.p, I.w.p);
   PK_HIP(hipEventRecord(ev.e1, c.stream));
   PK_HIP(hipGetLastError());
   float ms = 0.f;
   PK_HIP(ev.ms(&ms));
   *avg_seconds = ms * 1e-3 / reps;
   return 0;
}

extern "C" {

int pk_batch_create(int num, const pk_mat* R, const pk_mat* F, pk_batch* out) {
   PK_CALL(ensure_init());
   PK_REQUIRE(num >= 1 && R && F && out, "wrong batch arguments");
   long units = 0, re = 0, fe = 0, bi = 0;
   for (int b = 0; b < num; b++) {
      if (check_member(R[b], F[b])) {
         std::string msg = std::string("batch member ") + std::to_string(b) + ": " + pk_last_error();
         PK_FAIL(msg);
      }
      units += units_of(R[b]->n);
      re += blocks64(R[b]->sell.pad64);
      fe += blocks64(F[b]->sell.padded);
      bi += blocks64(R[b]->n * preconditioner_block(R[b]->bs));
   }
   PK_REQUIRE(units * kRows < (1L << 31), "batch too large for 32-bit columns");
   auto* P = new pk_batch_s();
   P->max_units = units;
   P->max_rentries = std::max(1L, re);
   P->max_fentries = std::max(1L, fe);
   P->max_binv = std::max(1L, bi);
   P->max_slots = num;
   for (int b = 0; b < num; b++) {
      P->fixR.push_back(R[b]);
      P->fixF.push_back(F[b]);
   }
   P->fix_init.resize(num);
   P->fix_has_init.assign(num, 0);
   *out = P;
   return 0;
}

int pk_batch_set_member(pk_batch P, int index, pk_mat R, pk_mat F) {
   PK_REQUIRE(P && index >= 0 && index < (int)P->fixR.size(), "no such batch member");
   PK_CALL(check_member(R, F));
   long units = 0, re = 0, fe = 0, bi = 0;
   for (int b = 0; b < (int)P->fixR.size(); b++) {
      pk_mat_s* r = b == index ? R : P->fixR[b];
      pk_mat_s* f = b == index ? F : P->fixF[b];
      units += units_of(r->n);
      re += blocks64(r->sell.pad64);
      fe += blocks64(f->sell.padded);
      bi += blocks64(r->n * preconditioner_block(r->bs));
   }
   PK_REQUIRE(units <= P->max_units && re <= P->max_rentries && fe <= P->max_fentries && bi <= P->max_binv,
              "the new member does not fit the batch's pool");
   P->fixR[index] = R;
   P->fixF[index] = F;
   return 0;
}

// The options reach the tick's kernels as launch arguments: a change drops the
// captured tick (the next tick captures it again); setting the same values
// keeps it, so a caller that sets its options before every solve reuses it.
int pk_batch_set_tolerances(pk_batch P, double tol, int max_outer) {
   PK_REQUIRE(P && tol > 0.0 && max_outer > 0, "wrong tolerances");
   if (tol == P->opts.tol && max_outer == P->opts.max_outer) return 0;
   P->opts.tol = tol;
   P->opts.max_outer = max_outer;
   P->graph.clear();
   return 0;
}

int pk_batch_set_precision(pk_batch P, int precision) {
   PK_REQUIRE(P && (precision == PK_PREC_FP64 || precision == PK_PREC_MIXED), "wrong precision");
   PK_REQUIRE(!P->sd.p || precision == P->precision, "precision must be set before the first solve");
   P->precision = precision;
   return 0;
}

int pk_batch_set_inner(pk_batch P, int restart, double reduction, int max_inner) {
   PK_REQUIRE(P && restart >= 1 && restart <= kBMaxM && reduction > 0.0 && reduction < 1.0 && max_inner > 0,
              "wrong inner-solver options");
   PK_REQUIRE(!P->sd.p || restart == P->m, "the restart length must be set before the first solve");
   if (restart == P->m && reduction == P->opts.reduction && max_inner == P->opts.max_inner) return 0;
   P->m = restart;
   P->opts.reduction = reduction;
   P->opts.max_inner = max_inner;
   P->graph.clear();
   return 0;
}

int pk_batch_set_preconditioner(pk_batch P, int poly) {
   PK_REQUIRE(P && (poly == 0 || poly == 1 || poly == 2), "wrong preconditioner");
   if ((poly == 1) == (P->opts.poly != 0) && (poly != 2) == (P->opts.sweep != 0)) return 0;
   P->opts.poly = poly == 1;
   P->opts.sweep = poly != 2;
   P->graph.clear();
   return 0;
}

int pk_batch_set_acceleration(pk_batch P, int depth) {
   PK_REQUIRE(P && depth >= 0 && depth <= kAA, "wrong Anderson depth (0 to 3)");
   if (depth == P->opts.aa) return 0;
   P->opts.aa = depth;
   P->graph.clear();
   return 0;
}

int pk_batch_set_shift(pk_batch P, double delta) {
   PK_REQUIRE(P && delta >= 0.0, "wrong shift");
   if (delta == P->opts.delta) return 0;
   P->opts.delta = delta;
   P->graph.clear();
   return 0;
}

int pk_batch_set_initial(pk_batch P, int index, pk_vec x) {
   PK_REQUIRE(P && index >= 0 && index < (int)P->fixR.size() && x && x->n == P->fixR[index]->n,
              "wrong initial vector");
   // an all-zero vector means "no initial space" (petsc.cxx:136-140)
   double nrm = 0.0;
   PK_CALL(pk_vec_dot(x, x, &nrm));
   P->fix_has_init[index] = nrm > 0.0;
   if (nrm > 0.0) {
      DBuf<double>& d = P->fix_init[index];
      if (!d.p) PK_REQUIRE(!d.alloc(x->n), "device allocation failed");
      PK_HIP(hipMemcpyAsync(d.p, x->d.p, sizeof(double) * x->n, hipMemcpyDeviceToDevice, ctx().stream));
   }
   return 0;
}

int pk_batch_solve(pk_batch P) {
   PK_REQUIRE(P && !P->fixR.empty(), "not a fixed batch");
   auto t0 = std::chrono::steady_clock::now();
   if (!P->sd.p) PK_CALL(pool_alloc(P));
   int rc = fixed_solve(P);
   P->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   PK_CALL(rc);
   for (int b = 0; b < (int)P->fixR.size(); b++) {
      SlotResult R;
      PK_CALL(slot_result(P, P->fix_slot[b], R, nullptr));
      PK_REQUIRE(R.status == 1, "eigen iteration of batch member " + std::to_string(b) + " did not converge");
   }
   return 0;
}

int pk_batch_get_eigenpair(pk_batch P, int index, double* lambda, pk_vec x) {
   PK_REQUIRE(P && index >= 0 && index < (int)P->fix_slot.size() && P->fix_slot[index] >= 0,
              "no solution for this batch member");
   SlotResult R;
   PK_CALL(slot_result(P, P->fix_slot[index], R, x));
   if (lambda) *lambda = 1.0 / R.k;
   return 0;
}

int pk_batch_get_stats(pk_batch P, int index, int* outer, long* inner, double* residual, double* seconds) {
   PK_REQUIRE(P && index >= 0 && index < (int)P->fix_slot.size() && P->fix_slot[index] >= 0,
              "no solution for this batch member");
   SlotResult R;
   PK_CALL(slot_result(P, P->fix_slot[index], R, nullptr));
   if (outer) *outer = R.it;
   if (inner) *inner = R.inner;
   if (residual) *residual = R.res;
   if (seconds) *seconds = P->seconds;
   return 0;
}

int pk_batch_destroy(pk_batch* P) {
   if (!P || !*P) return 0;
   (void)hipStreamSynchronize(ctx().stream);
   delete *P;
   *P = nullptr;
   return 0;
}

// ---- pool (continuous batching) ----------------------------------------------------

double pk_pool_bytes(long max_rows, long max_entries, long max_fentries, long max_binv, int max_slots,
                     int precision, int restart) {
   return pool_bytes(units_of(max_rows), blocks64(max_entries), blocks64(max_fentries), blocks64(max_binv),
                     max_slots, restart, precision);
}

int pk_pool_create(long max_rows, long max_entries, long max_fentries, long max_binv, int max_slots,
                   int precision, int restart, pk_batch* out) {
   PK_CALL(ensure_init());
   PK_REQUIRE(out && max_rows > 0 && max_slots > 0 && restart >= 1 && restart <= kBMaxM &&
                  (precision == PK_PREC_FP64 || precision == PK_PREC_MIXED),
              "wrong eigen pool arguments");
   auto* P = new pk_batch_s();
   P->max_units = units_of(max_rows);
   P->max_rentries = std::max(1L, blocks64(max_entries));
   P->max_fentries = std::max(1L, blocks64(max_fentries));
   P->max_binv = std::max(1L, blocks64(max_binv));
   P->max_slots = max_slots;
   P->precision = precision;
   P->m = restart;
   if (P->max_units * kRows >= (1L << 31)) {
      delete P;
      PK_FAIL("pool too large for 32-bit columns");
   }
   if (pool_alloc(P)) {
      delete P;
      return 1;
   }
   *out = P;
   return 0;
}

int pk_pool_set_options(pk_batch P, double tol, int max_outer, double reduction, int max_inner, double delta) {
   PK_REQUIRE(P && tol > 0.0 && max_outer > 0 && reduction > 0.0 && reduction < 1.0 && max_inner > 0 &&
                  delta >= 0.0,
              "wrong pool options");
   P->opts = BOpts{tol, delta, reduction, max_outer, max_inner, P->opts.poly, P->opts.sweep, P->opts.aa};
   P->graph.clear();
   return 0;
}

int pk_pool_admit(pk_batch P, pk_mat R, pk_mat F, pk_vec x0, int* slot) {
   PK_REQUIRE(P && slot && P->sd.p, "wrong pool arguments");
   const double* x0d = nullptr;
   if (x0) {
      PK_REQUIRE(R && x0->n == R->n, "initial vector size mismatch");
      x0d = x0->d.p;
   }
   *slot = -1;
   return pool_admit(P, R, F, x0d, slot);
}

int pk_pool_fits(pk_batch P, pk_mat R, pk_mat F, int* fits) {
   PK_REQUIRE(P && fits && P->sd.p, "wrong pool arguments");
   PK_CALL(check_member(R, F));
   const long units = units_of(R->n);
   const long re = blocks64(R->sell.pad64), fe = blocks64(F->sell.padded);
   const long bi = blocks64(R->n * preconditioner_block(R->bs));
   bool slot = false;
   for (auto& s : P->slots) slot |= !s.used;
   *fits = slot && P->ua.largest_free() >= units && P->ra.largest_free() >= re && P->fa.largest_free() >= fe &&
           P->ba.largest_free() >= bi;
   return 0;
}

int pk_pool_footprint(pk_mat R, pk_mat F, int precision, int restart, double* bytes) {
   PK_REQUIRE(bytes && restart >= 1 && restart <= kBMaxM, "wrong arguments");
   PK_CALL(check_member(R, F));
   const long units = units_of(R->n);
   *bytes = pool_bytes(units, blocks64(R->sell.pad64), blocks64(F->sell.padded),
                       blocks64(R->n * preconditioner_block(R->bs)), 1, restart, precision);
   return 0;
}

int pk_pool_step(pk_batch P, int ticks, int* live, int* done) {
   PK_REQUIRE(P && P->sd.p && ticks >= 0, "wrong pool arguments");
   return pool_step(P, ticks, live, done);
}

int pk_pool_collect(pk_batch P, int max, int* slots, int* count) {
   PK_REQUIRE(P && slots && count && max >= 0, "wrong pool arguments");
   // DONE slots in admission order
   std::vector<std::pair<long, int>> d;
   for (int s = 0; s < P->max_slots; s++)
      if (P->slots[s].used && P->h_phase[s] == kDone && !P->slots[s].harvested)
         d.push_back({P->slots[s].ticket, s});
   std::sort(d.begin(), d.end());
   int n = std::min<int>(max, (int)d.size());
   for (int i = 0; i < n; i++) {
      slots[i] = d[i].second;
      P->slots[d[i].second].harvested = true;
   }
   *count = n;
   return 0;
}

int pk_pool_result(pk_batch P, int slot, double* keff, pk_vec x, int* outer, long* inner, double* residual,
                   int* converged) {
   PK_REQUIRE(P, "null pool");
   SlotResult R;
   PK_CALL(slot_result(P, slot, R, x));
   if (keff) *keff = R.k;
   if (outer) *outer = R.it;
   if (inner) *inner = R.inner;
   if (residual) *residual = R.res;
   if (converged) *converged = R.status == 1;
   return 0;
}

int pk_pool_release(pk_batch P, int slot) {
   PK_REQUIRE(P, "null pool");
   return pool_release(P, slot);
}

int pk_pool_time_spmv(pk_batch P, int reps, double* avg_seconds, double* bytes) {
   PK_REQUIRE(P && P->sd.p && reps > 0 && avg_seconds && bytes, "wrong pool arguments");
   // every live slot's shifted operator must exist: run after a tick
   double n = 0.0, nnz = 0.0;
   for (auto& sl : P->slots)
      if (sl.used) {
         n += sl.n;
         nnz += (double)sl.nnz;
      }
   PK_REQUIRE(n > 0.0, "empty pool");
   const double s = P->precision == PK_PREC_MIXED ? 4.0 : 8.0;
   // algorithmic bytes of the CSR SpMV of every live problem
   *bytes = nnz * (s + 4.0) + (n + 1.0) * 4.0 + 2.0 * n * s;
   if (P->precision == PK_PREC_MIXED) return time_pool_spmv<float>(P, reps, avg_seconds);
   return time_pool_spmv<double>(P, reps, avg_seconds);
}

int pk_pool_capacity(pk_batch P, long* free_rows, long* free_entries, int* free_slots, long* ticks) {
   PK_REQUIRE(P, "null pool");
   if (free_rows) *free_rows = P->ua.largest_free() * kRows;
   if (free_entries) *free_entries = P->ra.largest_free() * 64;
   int fs = 0;
   for (auto& s : P->slots) fs += !s.used;
   if (free_slots) *free_slots = fs;
   if (ticks) *ticks = P->ticks;
   return 0;
}

int pk_tune_count(int* n) {
   PK_REQUIRE(n, "null argument");
   std::lock_guard<std::mutex> lock(g_tune_mu);
   *n = (int)g_tune_order.size();
   return 0;
}

int pk_tune_get(int i, long* rows, int* block, long* nnz, int* precision, int* slice_height, int* spmv_block,
                double* us) {
   std::lock_guard<std::mutex> lock(g_tune_mu);
   PK_REQUIRE(i >= 0 && i < (int)g_tune_order.size(), "no such tuned class");
   const TuneKey& k = g_tune_order[i];
   const TuneChoice& c = g_tune.at(k);
   if (rows) *rows = k.n;
   if (block) *block = k.bs;
   if (nnz) *nnz = k.nnz;
   if (precision) *precision = k.precision;
   if (slice_height) *slice_height = c.h;
   if (spmv_block) *spmv_block = c.block;
   if (us)
      for (int a = 0; a < 3; a++)
         for (int q = 0; q < 4; q++) us[a * 4 + q] = c.us[a][q];
   return 0;
}

int pk_tune_reset(void) {
   std::lock_guard<std::mutex> lock(g_tune_mu);
   g_tune.clear();
   g_tune_order.clear();
   return 0;
}

}  // extern "C"
