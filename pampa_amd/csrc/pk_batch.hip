// Batched k-eigenvalue solver: B independent problems R_b x = lambda F_b x
// advanced together, one kernel launch per step for all of them.
//
// Layout.  The instances are concatenated into one block-diagonal system.
// Instance b owns rows row_off[b] .. row_off[b] + n_b - 1 and is padded to a
// multiple of 256 rows, so a 256-thread workgroup never holds two instances;
// wg_inst[] maps a workgroup to its instance.  The SELL-64 arrays of R, F and
// of the shifted operator A = R - F/ks are concatenated on the device with the
// columns moved by row_off[b]; padding slices have width 0.
//
// Reductions are segmented: the workgroups of one instance reduce among
// themselves (pk_reduce.hpp seg_reduce), and the last one runs that
// instance's serial epilogue (Givens rotations, eigenvalue update, stop
// decisions).  Every decision is taken on the device, per instance; converged
// or stopped instances skip their workgroups at kernel entry.  The host syncs
// once per GMRES restart cycle and once per outer iteration, only to learn
// whether every instance is done.  A restart cycle (m GMRES steps and the
// update) is captured once as a HIP graph and relaunched.
//
// Algorithm per instance (FP64 outer, FP32 or FP64 inner):
//   k = <Fx,Fx>/<Fx,Rx>,  phi = x/|x|,  r = F phi/k - R phi,  res = |r|/|F phi/k|
//   stop when res <= tol; after the estimate settles, shift ks = k + delta;
//   solve (R - F/ks) e = r by right-preconditioned GMRES(m) with cell-block
//   Jacobi to a relative reduction, x = phi + e.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "pk_internal.hpp"
#include "pk_reduce.hpp"

namespace pk {

constexpr int kRows = 256;
constexpr int kBMaxM = 64;
constexpr int kBMaxBs = 16;

struct BLayout {
   const int* wg_inst;    // [nwg]
   const int* wg_first;   // [B]
   const int* wg_count;   // [B]
   const long* row_off;   // [B]
   const int* nrows;      // [B]
   const int* pbs;        // [B] preconditioner block
   const int* fbs;        // [B] cell block of F
   const long* binv_off;  // [B]
};

struct BSell {
   const long* off;
   const int* width;
   const int* col;
   const int* rlen;
};

struct BG {  // per-instance GMRES state (arrays of B)
   double* H;      // [B][(m+1) m] column-major per instance
   double* cs;     // [B][m]
   double* sn;     // [B][m]
   double* g;      // [B][m+1]
   double* y;      // [B][m]
   double* scale;  // [B]
   double* resid;  // [B][2]: estimate, target
   int* flags;     // [B][4]: stop, steps in cycle, breakdown, steps total
   double* ww;     // [B]
   const int* active;  // [B]
   int m;
};

struct BE {  // per-instance outer state
   double* dots;   // [B][3]: <y2,y1>, <y2,y2>, <x,x>
   double* rq;     // [B][2]: |r|^2, |q|^2
   double* k;      // [B]
   double* kprev;  // [B]
   double* ks;     // [B] 0 = no shift
   double* res;    // [B]
   int* it;        // [B]
   int* conv;      // [B]
   int* active;    // [B]
   int* rebuild;   // [B]
   int* shifted;   // [B]
};

__device__ __forceinline__ int seg_inst(const BLayout& L) { return L.wg_inst[blockIdx.x]; }

// ---- SELL concatenation ---------------------------------------------------------

__global__ void cat_slices_kernel(long ns_src, long ns_dst, const long* soff, const int* swidth,
                                  long dst_slice0, long base, long* doff, int* dwidth) {
   long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (s >= ns_dst) return;
   if (s < ns_src) {
      doff[dst_slice0 + s] = base + (soff[s] - soff[0]);
      dwidth[dst_slice0 + s] = swidth[s];
   } else {
      doff[dst_slice0 + s] = base + (soff[ns_src] - soff[0]);
      dwidth[dst_slice0 + s] = 0;
   }
}

__global__ void cat_entries_kernel(long n, const int* scol, int shift, int* dcol) {
   long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p < n) dcol[p] = scol[p] + shift;
}

__global__ void cat_rlen_kernel(long n_src, long n_dst, const int* src, int* dst) {
   long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (r < n_dst) dst[r] = r < n_src ? src[r] : 0;
}

__global__ void set_last_off_kernel(long* off, long ns, long total) {
   if (blockIdx.x == 0 && threadIdx.x == 0) off[ns] = total;
}

template <class T>
__global__ void to_t_kernel(const double* a, T* b, long n) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) b[i] = (T)a[i];
}

// ---- shifted operator and block inverses ------------------------------------------

// A = R - F/ks for instances flagged for rebuild (all instances when force)
__global__ void bshift_kernel(long N, BLayout L, BSell R, const double* rval, BSell F,
                              const double* fval, const double* ks, const int* rebuild, int force,
                              double* aval) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= N) return;
   const int b = L.wg_inst[row / kRows];
   if (!force && !rebuild[b]) return;
   const double inv_ks = ks[b] > 0.0 ? 1.0 / ks[b] : 0.0;
   const long local = row - L.row_off[b];
   const long s = row >> 6;
   const int lane = (int)(row & 63);
   const long base = R.off[s] + lane;
   const int fb = L.fbs[b];
   const long cell = local / fb;
   const long fbase = F.off[s] + lane;
   for (int k = 0; k < R.width[s]; k++) {
      long p = base + (long)k * 64;
      double v = rval[p];
      if (k < R.rlen[row]) {
         long c = (long)R.col[p] - L.row_off[b];
         if (c / fb == cell && inv_ks != 0.0) v -= fval[fbase + (c % fb) * 64] * inv_ks;
      }
      aval[p] = v;
   }
}

template <class T>
__global__ void bblock_inverse_kernel(long N, BLayout L, BSell A, const double* aval,
                                      const int* rebuild, int force, T* binv, int* err) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= N) return;
   const int b = L.wg_inst[row / kRows];
   if (!force && !rebuild[b]) return;
   const long local = row - L.row_off[b];
   const int bs = L.pbs[b];
   if (local >= L.nrows[b] || local % bs != 0) return;
   const long cell = local / bs;
   double a[kBMaxBs * kBMaxBs], m[kBMaxBs * kBMaxBs];
   for (int p = 0; p < bs * bs; p++) {
      a[p] = 0.0;
      m[p] = 0.0;
   }
   for (int r = 0; r < bs; r++) {
      m[r * bs + r] = 1.0;
      long rr = row + r;
      long s = rr >> 6;
      int lane = (int)(rr & 63);
      for (int k = 0; k < A.rlen[rr]; k++) {
         long p = A.off[s] + (long)k * 64 + lane;
         long c = (long)A.col[p] - L.row_off[b];
         if (c / bs == cell) a[r * bs + (c % bs)] += aval[p];
      }
   }
   for (int k = 0; k < bs; k++) {
      int piv = k;
      double best = fabs(a[k * bs + k]);
      for (int r = k + 1; r < bs; r++)
         if (fabs(a[r * bs + k]) > best) {
            best = fabs(a[r * bs + k]);
            piv = r;
         }
      if (best == 0.0) {
         *err = 1;
         return;
      }
      if (piv != k)
         for (int q = 0; q < bs; q++) {
            double t = a[k * bs + q];
            a[k * bs + q] = a[piv * bs + q];
            a[piv * bs + q] = t;
            t = m[k * bs + q];
            m[k * bs + q] = m[piv * bs + q];
            m[piv * bs + q] = t;
         }
      double inv = 1.0 / a[k * bs + k];
      for (int q = 0; q < bs; q++) {
         a[k * bs + q] *= inv;
         m[k * bs + q] *= inv;
      }
      for (int r = 0; r < bs; r++) {
         if (r == k) continue;
         double f = a[r * bs + k];
         if (f == 0.0) continue;
         for (int q = 0; q < bs; q++) {
            a[r * bs + q] -= f * a[k * bs + q];
            m[r * bs + q] -= f * m[k * bs + q];
         }
      }
   }
   T* dst = binv + L.binv_off[b] + cell * bs * bs;
   for (int p = 0; p < bs * bs; p++) dst[p] = (T)m[p];
}

// ---- GMRES kernels ------------------------------------------------------------------

template <class T>
__global__ void __launch_bounds__(kRows) bprec_kernel(long N, BLayout L, T* Vj, const T* binv, T* z, BG S) {
   const int b = seg_inst(L);
   if (S.flags[4 * b]) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const long local = row - L.row_off[b];
   const int bs = L.pbs[b];
   const T sc = (T)S.scale[b];
   T own = 0, acc = 0;
   if (local < L.nrows[b]) {
      const long cell = local / bs;
      const int gr = (int)(local - cell * bs);
      const T* bi = binv + L.binv_off[b] + cell * bs * bs + (long)gr * bs;
      const T* v = Vj + L.row_off[b] + cell * bs;
      for (int q = 0; q < bs; q++) acc += bi[q] * (v[q] * sc);
      own = Vj[row] * sc;
   }
   __syncthreads();
   if (row < N) {
      z[row] = acc;
      if (local < L.nrows[b]) Vj[row] = own;
   }
}

template <class T>
__global__ void __launch_bounds__(kRows) bspmv_dot_kernel(long N, BLayout L, BSell A, const T* aval,
                                                          const T* z, T* w, const T* V, int j,
                                                          double* partials, unsigned* tickets, BG S) {
   const int b = seg_inst(L);
   if (S.flags[4 * b]) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   T acc = 0;
   if (row < N) {
      const long s = row >> 6;
      const long base = A.off[s] + (threadIdx.x & 63);
      const int wd = A.width[s];
      for (int k = 0; k < wd; k++) {
         long p = base + (long)k * 64;
         acc += aval[p] * z[A.col[p]];
      }
      w[row] = acc;
   }
   const double wr = (double)acc;
   const int nv = j + 1;
   seg_reduce<kBMaxM>(
       nv, [&](int k) { return row < N ? (double)V[(long)k * N + row] * wr : 0.0; },
       partials + (long)L.wg_first[b] * nv, tickets + b, S.H + (long)b * (S.m + 1) * S.m + (long)j * (S.m + 1),
       (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
}

template <class T>
__global__ void __launch_bounds__(kRows) borth_kernel(long N, BLayout L, const T* w, T* V, int j,
                                                      double* partials, unsigned* tickets, BG S) {
   const int b = seg_inst(L);
   if (S.flags[4 * b]) return;
   const int m = S.m;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   double* Hb = S.H + (long)b * (m + 1) * m;
   double u = 0.0;
   if (row < N) {
      const double* h = Hb + (long)j * (m + 1);
      T t = w[row];
      for (int i = 0; i <= j; i++) t -= (T)h[i] * V[(long)i * N + row];
      V[(long)(j + 1) * N + row] = t;
      u = (double)t;
   }
   bool last = seg_reduce<1>(
       1, [&](int) { return u * u; }, partials + L.wg_first[b], tickets + b, S.ww + b,
       (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   if (!last) return;
   __shared__ double hc[kBMaxM + 2], cc[kBMaxM], ss[kBMaxM], gj;
   const int ld = m + 1;
   const int t = threadIdx.x;
   if (t <= j) hc[t] = Hb[t + (long)j * ld];
   if (t < j) {
      cc[t] = S.cs[(long)b * m + t];
      ss[t] = S.sn[(long)b * m + t];
   }
   if (t == 0) gj = S.g[(long)b * (m + 1) + j];
   __syncthreads();
   if (t == 0) {
      double hn = sqrt(fmax(S.ww[b], 0.0));
      hc[j + 1] = hn;
      for (int i = 0; i < j; i++) {
         double a = hc[i], c2 = hc[i + 1];
         hc[i] = cc[i] * a + ss[i] * c2;
         hc[i + 1] = -ss[i] * a + cc[i] * c2;
      }
      double a = hc[j], c2 = hc[j + 1];
      double rr = sqrt(a * a + c2 * c2);
      double c = rr > 0.0 ? a / rr : 1.0, s = rr > 0.0 ? c2 / rr : 0.0;
      hc[j] = rr;
      hc[j + 1] = 0.0;
      for (int i = 0; i <= j + 1; i++) Hb[i + (long)j * ld] = hc[i];
      S.cs[(long)b * m + j] = c;
      S.sn[(long)b * m + j] = s;
      S.g[(long)b * (m + 1) + j + 1] = -s * gj;
      S.g[(long)b * (m + 1) + j] = c * gj;
      double res = fabs(s * gj);
      S.resid[2 * b] = res;
      int* fl = S.flags + 4 * b;
      fl[1] = j + 1;
      bool breakdown = !(hn > 1e-300) || !(rr > 0.0);
      if (breakdown) fl[2] = 1;
      if (res <= S.resid[2 * b + 1] || breakdown || j + 1 == m) fl[0] = 1;
      S.scale[b] = hn > 0.0 ? 1.0 / hn : 0.0;
   }
}

// y = H^-1 g per instance (one workgroup per instance)
__global__ void __launch_bounds__(kRows) btrisolve_kernel(int B, BG S) {
   const int b = blockIdx.x;
   if (b >= B) return;
   const int m = S.m;
   const int k = S.flags[4 * b + 1];
   if (k == 0) return;
   __shared__ double h[kBMaxM * kBMaxM], gg[kBMaxM], yy[kBMaxM];
   const double* Hb = S.H + (long)b * (m + 1) * m;
   for (int p = threadIdx.x; p < k * k; p += blockDim.x) {
      int i = p % k, q = p / k;
      h[i + q * kBMaxM] = Hb[i + (long)q * (m + 1)];
   }
   if ((int)threadIdx.x < k) gg[threadIdx.x] = S.g[(long)b * (m + 1) + threadIdx.x];
   __syncthreads();
   if (threadIdx.x == 0) {
      for (int i = k - 1; i >= 0; i--) {
         double s = gg[i];
         for (int q = i + 1; q < k; q++) s -= h[i + q * kBMaxM] * yy[q];
         double d = h[i + i * kBMaxM];
         yy[i] = d != 0.0 ? s / d : 0.0;
      }
   }
   __syncthreads();
   if ((int)threadIdx.x < k) S.y[(long)b * m + threadIdx.x] = yy[threadIdx.x];
}

// e += B^-1 (sum_i y_i V[i]) for instances that took steps in this cycle
template <class T>
__global__ void __launch_bounds__(kRows) bupdate_kernel(long N, BLayout L, const T* V, const T* binv,
                                                        BG S, T* e) {
   const int b = seg_inst(L);
   const int k = S.flags[4 * b + 1];
   if (k == 0) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   const long local = row - L.row_off[b];
   const int bs = L.pbs[b];
   __shared__ T u[kRows];
   T acc = 0;
   if (row < N)
      for (int i = 0; i < k; i++) acc += (T)S.y[(long)b * S.m + i] * V[(long)i * N + row];
   u[threadIdx.x] = acc;
   __syncthreads();
   if (local < L.nrows[b]) {
      const long cell = local / bs;
      const int gr = (int)(local - cell * bs);
      const T* bi = binv + L.binv_off[b] + cell * bs * bs + (long)gr * bs;
      const int first = (int)(threadIdx.x - gr);  // cells never straddle a workgroup
      T s = 0;
      for (int q = 0; q < bs; q++) s += bi[q] * u[first + q];
      e[row] += s;
   }
}

// V[0] = rhs - A e; beta; (re)arm the instance's cycle.  first != 0: e = 0,
// the target becomes max(reduction beta, abs_target) and the step totals restart.
template <class T>
__global__ void __launch_bounds__(kRows) brestart_kernel(long N, BLayout L, BSell A, const T* aval,
                                                         const T* e, const T* rhs, T* V, int first,
                                                         double reduction, double abs_target,
                                                         int max_it, double* partials,
                                                         unsigned* tickets, BG S) {
   const int b = seg_inst(L);
   // finished instances (converged outer loop, or stopped before this cycle
   // took a step) keep the flags and residual they have
   if (!first && (!S.active[b] || (S.flags[4 * b] && S.flags[4 * b + 1] == 0))) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   double u = 0.0;
   if (row < N) {
      T acc = 0;
      if (!first) {
         const long s = row >> 6;
         const long base = A.off[s] + (threadIdx.x & 63);
         const int wd = A.width[s];
         for (int k = 0; k < wd; k++) {
            long p = base + (long)k * 64;
            acc += aval[p] * e[A.col[p]];
         }
      }
      T r = rhs[row] - acc;
      V[row] = r;
      u = (double)r;
   }
   bool last = seg_reduce<1>(
       1, [&](int) { return u * u; }, partials + L.wg_first[b], tickets + b, S.ww + b,
       (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   if (last && threadIdx.x == 0) {
      double beta = sqrt(S.ww[b]);
      int* fl = S.flags + 4 * b;
      if (first) {
         S.resid[2 * b + 1] = fmax(reduction * beta, abs_target);
         fl[3] = 0;
      } else {
         fl[3] += fl[1];
      }
      S.g[(long)b * (S.m + 1)] = beta;
      S.resid[2 * b] = beta;
      S.scale[b] = beta > 0.0 ? 1.0 / beta : 0.0;
      fl[0] = (!S.active[b] || beta <= S.resid[2 * b + 1] || !(beta > 0.0) || fl[3] >= max_it) ? 1 : 0;
      fl[1] = 0;
      fl[2] = 0;
   }
}

// ---- outer (eigen) kernels ----------------------------------------------------------

// y1 = R x, y2 = F x, dots per instance
__global__ void __launch_bounds__(kRows) brf_kernel(long N, BLayout L, BSell R, const double* rval,
                                                     BSell F, const double* fval, const double* x,
                                                     double* y1, double* y2, double* partials,
                                                     unsigned* tickets, BE E) {
   const int b = seg_inst(L);
   if (E.conv[b]) return;
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   double a = 0.0, c = 0.0, xv = 0.0;
   if (row < N) {
      const long s = row >> 6;
      const int lane = threadIdx.x & 63;
      long base = R.off[s] + lane;
      for (int k = 0; k < R.width[s]; k++) {
         long p = base + (long)k * 64;
         a += rval[p] * x[R.col[p]];
      }
      base = F.off[s] + lane;
      for (int k = 0; k < F.width[s]; k++) {
         long p = base + (long)k * 64;
         c += fval[p] * x[F.col[p]];
      }
      y1[row] = a;
      y2[row] = c;
      xv = x[row];
   }
   seg_reduce<3>(
       3, [&](int k) { return k == 0 ? c * a : (k == 1 ? c * c : xv * xv); },
       partials + (long)L.wg_first[b] * 3, tickets + b, E.dots + 3 * b, (int)blockIdx.x - L.wg_first[b],
       L.wg_count[b]);
}

// phi = x/|x|, r = (y2/k - y1)/|x|; residual; convergence and shift decisions
__global__ void __launch_bounds__(kRows) bresid_kernel(long N, BLayout L, const double* x,
                                                        const double* y1, const double* y2,
                                                        double* phi, double* r, double tol,
                                                        double delta, double* partials,
                                                        unsigned* tickets, BE E) {
   const int b = seg_inst(L);
   if (E.conv[b]) return;
   const double* d = E.dots + 3 * b;
   const double inv_k = d[0] / d[1];
   const double inv_nx = 1.0 / sqrt(d[2]);
   const long row = blockIdx.x * (long)kRows + threadIdx.x;
   double rr = 0.0, qq = 0.0;
   if (row < N) {
      double q = y2[row] * inv_k * inv_nx;
      double v = q - y1[row] * inv_nx;
      phi[row] = x[row] * inv_nx;
      r[row] = v;
      rr = v * v;
      qq = q * q;
   }
   bool last = seg_reduce<2>(
       2, [&](int k) { return k == 0 ? rr : qq; }, partials + (long)L.wg_first[b] * 2, tickets + b,
       E.rq + 2 * b, (int)blockIdx.x - L.wg_first[b], L.wg_count[b]);
   if (last && threadIdx.x == 0) {
      const double k = d[1] / d[0];
      const double res = sqrt(E.rq[2 * b] / E.rq[2 * b + 1]);
      E.k[b] = k;
      E.res[b] = res;
      if (res <= tol || !(res == res)) {
         E.conv[b] = 1;
         E.active[b] = 0;
      } else if (delta > 0.0) {
         const int it = E.it[b];
         const bool settle = it >= 2 && fabs(k - E.kprev[b]) < 1e-2 * k;
         if (settle && (!E.shifted[b] || fabs(k + delta - E.ks[b]) > 0.2 * delta)) {
            E.ks[b] = k + delta;
            E.shifted[b] = 1;
            E.rebuild[b] = 1;
         }
      }
      E.kprev[b] = k;
      E.it[b] += 1;
   }
}

template <class T>
__global__ void badd_kernel(long N, BLayout L, const double* phi, const T* e, const int* active,
                            double* x) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= N) return;
   const int b = L.wg_inst[row / kRows];
   x[row] = active[b] ? phi[row] + (double)e[row] : phi[row];
}

__global__ void bfill_kernel(double* x, long N, BLayout L, double v) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= N) return;
   const int b = L.wg_inst[row / kRows];
   x[row] = (row - L.row_off[b] < L.nrows[b]) ? v : 0.0;
}

}  // namespace pk

using namespace pk;

// A concatenated SELL matrix (device).
struct CatSell {
   long nslices = 0, entries = 0;
   DBuf<long> off;
   DBuf<int> width, col, rlen;
   DBuf<double> val;
   BSell view() const { return BSell{off.p, width.p, col.p, rlen.p}; }
};

template <class T>
struct BGmresT {
   int m = 30;
   DBuf<T> V, z, w, e, rhs, binv, aval;
   hipGraphExec_t graph = nullptr;
   ~BGmresT() {
      if (graph) (void)hipGraphExecDestroy(graph);
   }
};

struct pk_batch_s {
   int B = 0;
   std::vector<pk_mat_s*> R, F;
   std::vector<long> row_off;
   std::vector<int> n, pbs, fbs, wg_first, wg_count;
   long N = 0;
   int nwg = 0;
   // layout
   DBuf<int> d_wg_inst, d_wg_first, d_wg_count, d_nrows, d_pbs, d_fbs, d_active_g;
   DBuf<long> d_row_off, d_binv_off;
   long binv_total = 0;
   CatSell Rc, Fc;
   DBuf<double> aval64;  // shifted operator values (FP64)
   // reduction scratch of its own: captured graphs keep these addresses
   DBuf<double> part;
   DBuf<unsigned> tick;
   DBuf<int> err;
   // options
   double tol = 1e-8;
   int max_outer = 10000;
   int precision = PK_PREC_FP64;
   int restart = 30;
   double reduction = 0.1;
   int max_inner = 2000;
   double delta = 0.1;
   // state
   DBuf<double> x, phi, y1, y2, r;
   DBuf<double> st_d;  // H, cs, sn, g, y, scale, resid, ww, dots, rq, k, kprev, ks, res
   DBuf<int> st_i;     // flags, it, conv, active, rebuild, shifted
   BG gs{};
   BE es{};
   BGmresT<double> g64;
   BGmresT<float> g32;
   bool ready = false;
   std::vector<char> has_init;
   DBuf<double> init;
   // stats
   std::vector<int> outer;
   std::vector<long> inner;
   std::vector<double> keff, resid;
   double seconds = 0.0;

   BLayout layout() const {
      return BLayout{d_wg_inst.p, d_wg_first.p, d_wg_count.p, d_row_off.p,
                     d_nrows.p,   d_pbs.p,      d_fbs.p,      d_binv_off.p};
   }
};

static int cat_sell(pk_batch_s* Bt, const std::vector<pk_mat_s*>& mats, CatSell& out) {
   Ctx& c = ctx();
   const int B = Bt->B;
   long total_slices = Bt->N / 64;
   std::vector<long> base(B + 1, 0);
   for (int b = 0; b < B; b++) base[b + 1] = base[b] + mats[b]->sell.padded;
   out.nslices = total_slices;
   out.entries = base[B];
   PK_REQUIRE(!out.off.alloc(total_slices + 1) && !out.width.alloc(total_slices) &&
                  !out.col.alloc(out.entries ? out.entries : 1) && !out.val.alloc(out.entries ? out.entries : 1) &&
                  !out.rlen.alloc(Bt->N),
              "device allocation failed");
   for (int b = 0; b < B; b++) {
      const SellData& S = mats[b]->sell;
      long ns_dst = (b + 1 < B ? Bt->row_off[b + 1] : Bt->N) / 64 - Bt->row_off[b] / 64;
      cat_slices_kernel<<<grid_for(ns_dst), kBlock, 0, c.stream>>>(S.nslices, ns_dst, S.off.p, S.width.p,
                                                                  Bt->row_off[b] / 64, base[b], out.off.p,
                                                                  out.width.p);
      if (S.padded) {
         cat_entries_kernel<<<grid_for(S.padded), kBlock, 0, c.stream>>>(S.padded, S.col.p,
                                                                        (int)Bt->row_off[b],
                                                                        out.col.p + base[b]);
         PK_HIP(hipMemcpyAsync(out.val.p + base[b], S.val.p, sizeof(double) * S.padded,
                               hipMemcpyDeviceToDevice, c.stream));
      }
      long n_dst = (b + 1 < B ? Bt->row_off[b + 1] : Bt->N) - Bt->row_off[b];
      cat_rlen_kernel<<<grid_for(n_dst), kBlock, 0, c.stream>>>(mats[b]->n, n_dst, S.rlen.p,
                                                               out.rlen.p + Bt->row_off[b]);
   }
   set_last_off_kernel<<<1, 1, 0, c.stream>>>(out.off.p, total_slices, out.entries);
   PK_HIP(hipGetLastError());
   return 0;
}

template <class T>
static BGmresT<T>& gm(pk_batch_s* Bt);
template <>
BGmresT<double>& gm<double>(pk_batch_s* Bt) {
   return Bt->g64;
}
template <>
BGmresT<float>& gm<float>(pk_batch_s* Bt) {
   return Bt->g32;
}

static int batch_setup(pk_batch_s* Bt) {
   if (Bt->ready) return 0;
   Ctx& c = ctx();
   const int B = Bt->B;
   const int m = Bt->restart;
   // layout
   Bt->row_off.assign(B, 0);
   Bt->wg_first.assign(B, 0);
   Bt->wg_count.assign(B, 0);
   long rows = 0;
   std::vector<int> wg_inst;
   std::vector<long> binv_off(B, 0);
   long binv = 0;
   for (int b = 0; b < B; b++) {
      Bt->row_off[b] = rows;
      long padded = ((long)Bt->n[b] + kRows - 1) / kRows * kRows;
      if (padded == 0) padded = kRows;
      Bt->wg_first[b] = (int)(rows / kRows);
      Bt->wg_count[b] = (int)(padded / kRows);
      for (long w = 0; w < padded / kRows; w++) wg_inst.push_back(b);
      rows += padded;
      binv_off[b] = binv;
      binv += (long)Bt->n[b] * Bt->pbs[b];
   }
   Bt->N = rows;
   Bt->nwg = (int)wg_inst.size();
   Bt->binv_total = binv;
   PK_REQUIRE(rows < (1L << 31), "batch too large for 32-bit columns");
   auto up_i = [&](DBuf<int>& d, const std::vector<int>& h) {
      return d.alloc(h.size()) || hipMemcpy(d.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice);
   };
   auto up_l = [&](DBuf<long>& d, const std::vector<long>& h) {
      return d.alloc(h.size()) || hipMemcpy(d.p, h.data(), sizeof(long) * h.size(), hipMemcpyHostToDevice);
   };
   PK_REQUIRE(!up_i(Bt->d_wg_inst, wg_inst) && !up_i(Bt->d_wg_first, Bt->wg_first) &&
                  !up_i(Bt->d_wg_count, Bt->wg_count) && !up_i(Bt->d_nrows, Bt->n) &&
                  !up_i(Bt->d_pbs, Bt->pbs) && !up_i(Bt->d_fbs, Bt->fbs) && !up_l(Bt->d_row_off, Bt->row_off) &&
                  !up_l(Bt->d_binv_off, binv_off),
              "device allocation failed");
   PK_CALL(cat_sell(Bt, Bt->R, Bt->Rc));
   PK_CALL(cat_sell(Bt, Bt->F, Bt->Fc));
   PK_REQUIRE(!Bt->aval64.alloc(Bt->Rc.entries ? Bt->Rc.entries : 1), "device allocation failed");
   const long N = Bt->N;
   PK_REQUIRE(!Bt->x.alloc(N) && !Bt->phi.alloc(N) && !Bt->y1.alloc(N) && !Bt->y2.alloc(N) && !Bt->r.alloc(N),
              "device allocation failed");
   // per-instance state
   const size_t nd = (size_t)B * ((size_t)(m + 1) * m + 2 * m + (m + 1) + m + 1 + 2 + 1 + 3 + 2 + 4);
   PK_REQUIRE(!Bt->st_d.alloc(nd) && !Bt->st_i.alloc((size_t)B * 10), "device allocation failed");
   PK_HIP(hipMemsetAsync(Bt->st_d.p, 0, sizeof(double) * nd, c.stream));
   PK_HIP(hipMemsetAsync(Bt->st_i.p, 0, sizeof(int) * B * 10, c.stream));
   double* p = Bt->st_d.p;
   BG& g = Bt->gs;
   g.m = m;
   g.H = p;
   p += (size_t)B * (m + 1) * m;
   g.cs = p;
   p += (size_t)B * m;
   g.sn = p;
   p += (size_t)B * m;
   g.g = p;
   p += (size_t)B * (m + 1);
   g.y = p;
   p += (size_t)B * m;
   g.scale = p;
   p += B;
   g.resid = p;
   p += 2 * (size_t)B;
   g.ww = p;
   p += B;
   BE& e = Bt->es;
   e.dots = p;
   p += 3 * (size_t)B;
   e.rq = p;
   p += 2 * (size_t)B;
   e.k = p;
   p += B;
   e.kprev = p;
   p += B;
   e.ks = p;
   p += B;
   e.res = p;
   p += B;
   int* q = Bt->st_i.p;
   g.flags = q;
   q += 4 * B;
   e.it = q;
   q += B;
   e.conv = q;
   q += B;
   e.active = q;
   q += B;
   e.rebuild = q;
   q += B;
   e.shifted = q;
   q += B;
   g.active = e.active;
   // GMRES buffers
   auto alloc_g = [&](auto& G) {
      return G.V.alloc((size_t)(m + 1) * N) || G.z.alloc(N) || G.w.alloc(N) || G.e.alloc(N) ||
             G.rhs.alloc(N) || G.binv.alloc(Bt->binv_total ? Bt->binv_total : 1) ||
             G.aval.alloc(Bt->Rc.entries ? Bt->Rc.entries : 1);
   };
   if (Bt->precision == PK_PREC_MIXED) {
      PK_REQUIRE(!alloc_g(Bt->g32), "device allocation failed");
      Bt->g32.m = m;
   } else {
      PK_REQUIRE(!alloc_g(Bt->g64), "device allocation failed");
      Bt->g64.m = m;
   }
   PK_HIP(hipMemsetAsync(Bt->x.p, 0, sizeof(double) * N, c.stream));
   PK_REQUIRE(!Bt->part.alloc((size_t)Bt->nwg * (m + 4)) && !Bt->tick.alloc(B), "device allocation failed");
   PK_HIP(hipMemsetAsync(Bt->tick.p, 0, sizeof(unsigned) * B, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   Bt->ready = true;
   return 0;
}

// A values and block inverses for the instances flagged for rebuild (or all)
template <class T>
static int batch_rebuild(pk_batch_s* Bt, int force) {
   Ctx& c = ctx();
   const BLayout L = Bt->layout();
   const long N = Bt->N;
   bshift_kernel<<<grid_for(N), kBlock, 0, c.stream>>>(N, L, Bt->Rc.view(), Bt->Rc.val.p, Bt->Fc.view(),
                                                      Bt->Fc.val.p, Bt->es.ks, Bt->es.rebuild, force,
                                                      Bt->aval64.p);
   auto& G = gm<T>(Bt);
   if (sizeof(T) == sizeof(double))
      PK_HIP(hipMemcpyAsync(G.aval.p, Bt->aval64.p, sizeof(double) * Bt->Rc.entries, hipMemcpyDeviceToDevice,
                            c.stream));
   else
      to_t_kernel<T><<<grid_for(Bt->Rc.entries), kBlock, 0, c.stream>>>(Bt->aval64.p, G.aval.p, Bt->Rc.entries);
   DBuf<int>& err = Bt->err;
   if (!err.p) PK_REQUIRE(!err.alloc(1), "device allocation failed");
   PK_HIP(hipMemsetAsync(err.p, 0, sizeof(int), c.stream));
   bblock_inverse_kernel<T><<<grid_for(N, 64), 64, 0, c.stream>>>(N, L, Bt->Rc.view(), Bt->aval64.p,
                                                                 Bt->es.rebuild, force, G.binv.p, err.p);
   PK_HIP(hipGetLastError());
   int h = 0;
   PK_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipMemsetAsync(Bt->es.rebuild, 0, sizeof(int) * Bt->B, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   PK_REQUIRE(h == 0, "singular diagonal block in the block-Jacobi preconditioner");
   return 0;
}

// one restart cycle (m steps + update + true residual) on the stream
template <class T>
static void launch_cycle(pk_batch_s* Bt) {
   Ctx& c = ctx();
   auto& G = gm<T>(Bt);
   const BLayout L = Bt->layout();
   const long N = Bt->N;
   const int nwg = Bt->nwg, m = G.m;
   for (int j = 0; j < m; j++) {
      bprec_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, G.V.p + (size_t)j * N, G.binv.p, G.z.p, Bt->gs);
      bspmv_dot_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Bt->Rc.view(), G.aval.p, G.z.p, G.w.p, G.V.p, j,
                                                       Bt->part.p, Bt->tick.p, Bt->gs);
      borth_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, G.w.p, G.V.p, j, Bt->part.p, Bt->tick.p, Bt->gs);
   }
   btrisolve_kernel<<<Bt->B, kRows, 0, c.stream>>>(Bt->B, Bt->gs);
   bupdate_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, G.V.p, G.binv.p, Bt->gs, G.e.p);
   brestart_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Bt->Rc.view(), G.aval.p, G.e.p, G.rhs.p, G.V.p, 0,
                                                   Bt->reduction, 0.0, Bt->max_inner, Bt->part.p, Bt->tick.p,
                                                   Bt->gs);
}

// Profiled restart cycle: every kernel bracketed by events, and only steps in
// which some instance still works are issued.  Algorithmic bytes count the
// instances that run the step: prec reads V[j] and the block inverses and
// writes z and V[j]; spmv_dot is a SELL SpMV (values + columns, gathered z,
// w written) plus j+1 basis vectors; orth reads w and j+1 basis vectors and
// writes V[j+1].
template <class T>
static int launch_cycle_profiled(pk_batch_s* Bt) {
   Ctx& c = ctx();
   auto& G = gm<T>(Bt);
   const BLayout L = Bt->layout();
   const long N = Bt->N;
   const int nwg = Bt->nwg, m = G.m, B = Bt->B;
   const double s = sizeof(T);
   std::vector<int> fl(4 * (size_t)B);
   for (int j = 0; j < m; j++) {
      PK_HIP(hipMemcpyAsync(fl.data(), Bt->gs.flags, sizeof(int) * 4 * B, hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      double n = 0.0, nnz = 0.0, bsn = 0.0;
      for (int b = 0; b < B; b++)
         if (!fl[4 * b]) {
            n += Bt->n[b];
            nnz += (double)Bt->R[b]->nnz;
            bsn += (double)Bt->n[b] * Bt->pbs[b];
         }
      if (n == 0.0) break;
      prof_begin("gmres_prec", n * s * 3.0 + bsn * s);
      bprec_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, G.V.p + (size_t)j * N, G.binv.p, G.z.p, Bt->gs);
      prof_end();
      prof_begin("gmres_spmv_dot", nnz * (s + 4.0) + (n + B) * 4.0 + 2.0 * n * s + (j + 1.0) * n * s);
      bspmv_dot_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Bt->Rc.view(), G.aval.p, G.z.p, G.w.p, G.V.p, j,
                                                       Bt->part.p, Bt->tick.p, Bt->gs);
      prof_end();
      prof_begin("gmres_orth", n * s * (j + 3.0));
      borth_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, G.w.p, G.V.p, j, Bt->part.p, Bt->tick.p, Bt->gs);
      prof_end();
   }
   btrisolve_kernel<<<Bt->B, kRows, 0, c.stream>>>(Bt->B, Bt->gs);
   bupdate_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, G.V.p, G.binv.p, Bt->gs, G.e.p);
   brestart_kernel<T><<<nwg, kRows, 0, c.stream>>>(N, L, Bt->Rc.view(), G.aval.p, G.e.p, G.rhs.p, G.V.p, 0,
                                                   Bt->reduction, 0.0, Bt->max_inner, Bt->part.p, Bt->tick.p,
                                                   Bt->gs);
   PK_HIP(hipGetLastError());
   return 0;
}

// GMRES on A e = r (r = Bt->r, FP64) for every active instance
template <class T>
static int batch_gmres(pk_batch_s* Bt, std::vector<int>& hflags) {
   Ctx& c = ctx();
   auto& G = gm<T>(Bt);
   const BLayout L = Bt->layout();
   const long N = Bt->N;
   const int B = Bt->B;
   to_t_kernel<T><<<grid_for(N), kBlock, 0, c.stream>>>(Bt->r.p, G.rhs.p, N);
   PK_HIP(hipMemsetAsync(G.e.p, 0, sizeof(T) * N, c.stream));
   brestart_kernel<T><<<Bt->nwg, kRows, 0, c.stream>>>(N, L, Bt->Rc.view(), G.aval.p, G.e.p, G.rhs.p, G.V.p, 1,
                                                       Bt->reduction, 0.0, Bt->max_inner, Bt->part.p,
                                                       Bt->tick.p, Bt->gs);
   PK_HIP(hipGetLastError());
   if (!G.graph && !profiling()) {
      // capture one restart cycle; its kernels read every varying quantity
      // from device memory, so the graph is reused for all cycles and solves
      hipGraph_t g = nullptr;
      PK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      launch_cycle<T>(Bt);
      PK_HIP(hipStreamEndCapture(c.stream, &g));
      PK_HIP(hipGraphInstantiate(&G.graph, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
   }
   hflags.assign(4 * (size_t)B, 0);
   for (int cyc = 0; cyc < 100000; cyc++) {
      PK_HIP(hipMemcpyAsync(hflags.data(), Bt->gs.flags, sizeof(int) * 4 * B, hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      bool all = true;
      for (int b = 0; b < B; b++) all &= hflags[4 * b] != 0;
      if (all) break;
      if (profiling())
         PK_CALL(launch_cycle_profiled<T>(Bt));
      else if (G.graph)
         PK_HIP(hipGraphLaunch(G.graph, c.stream));
      else
         launch_cycle<T>(Bt);
      PK_HIP(hipGetLastError());
   }
   return 0;
}

template <class T>
static int batch_solve_t(pk_batch_s* Bt) {
   Ctx& c = ctx();
   const BLayout L = Bt->layout();
   const long N = Bt->N;
   const int B = Bt->B;
   const int gb = grid_for(N);
   // initial iterates and state
   bfill_kernel<<<gb, kBlock, 0, c.stream>>>(Bt->x.p, N, L, 1.0);
   if (Bt->init.p)
      for (int b = 0; b < B; b++)
         if (Bt->has_init[b])
            PK_HIP(hipMemcpyAsync(Bt->x.p + Bt->row_off[b], Bt->init.p + Bt->row_off[b],
                                  sizeof(double) * Bt->n[b], hipMemcpyDeviceToDevice, c.stream));
   PK_HIP(hipMemsetAsync(Bt->st_i.p, 0, sizeof(int) * B * 10, c.stream));
   PK_HIP(hipMemsetAsync(Bt->es.ks, 0, sizeof(double) * B, c.stream));
   std::vector<int> ones(B, 1);
   PK_HIP(hipMemcpyAsync(Bt->es.active, ones.data(), sizeof(int) * B, hipMemcpyHostToDevice, c.stream));
   PK_CALL(batch_rebuild<T>(Bt, 1));
   Bt->inner.assign(B, 0);
   std::vector<int> hconv(B), hreb(B), hflags;
   for (int it = 0; it < Bt->max_outer; it++) {
      brf_kernel<<<Bt->nwg, kRows, 0, c.stream>>>(N, L, Bt->Rc.view(), Bt->Rc.val.p, Bt->Fc.view(), Bt->Fc.val.p,
                                                  Bt->x.p, Bt->y1.p, Bt->y2.p, Bt->part.p, Bt->tick.p, Bt->es);
      bresid_kernel<<<Bt->nwg, kRows, 0, c.stream>>>(N, L, Bt->x.p, Bt->y1.p, Bt->y2.p, Bt->phi.p, Bt->r.p, Bt->tol,
                                                     Bt->delta, Bt->part.p, Bt->tick.p, Bt->es);
      PK_HIP(hipGetLastError());
      PK_HIP(hipMemcpyAsync(hconv.data(), Bt->es.conv, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipMemcpyAsync(hreb.data(), Bt->es.rebuild, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      bool all = true, reb = false;
      for (int b = 0; b < B; b++) {
         all &= hconv[b] != 0;
         reb |= hreb[b] != 0;
      }
      if (all) break;
      if (reb) PK_CALL(batch_rebuild<T>(Bt, 0));
      PK_CALL(batch_gmres<T>(Bt, hflags));
      for (int b = 0; b < B; b++) Bt->inner[b] += hflags[4 * b + 3];
      badd_kernel<T><<<gb, kBlock, 0, c.stream>>>(N, L, Bt->phi.p, gm<T>(Bt).e.p, Bt->es.active, Bt->x.p);
      PK_HIP(hipGetLastError());
   }
   // per-instance results
   std::vector<double> hk(B), hres(B);
   std::vector<int> hit(B);
   PK_HIP(hipMemcpy(hk.data(), Bt->es.k, sizeof(double) * B, hipMemcpyDeviceToHost));
   PK_HIP(hipMemcpy(hres.data(), Bt->es.res, sizeof(double) * B, hipMemcpyDeviceToHost));
   PK_HIP(hipMemcpy(hit.data(), Bt->es.it, sizeof(int) * B, hipMemcpyDeviceToHost));
   PK_HIP(hipMemcpy(hconv.data(), Bt->es.conv, sizeof(int) * B, hipMemcpyDeviceToHost));
   Bt->keff = hk;
   Bt->resid = hres;
   Bt->outer.assign(B, 0);
   for (int b = 0; b < B; b++) Bt->outer[b] = hit[b] > 0 ? hit[b] - 1 : 0;
   for (int b = 0; b < B; b++)
      PK_REQUIRE(hres[b] <= Bt->tol, "eigen iteration of batch member " + std::to_string(b) + " did not converge");
   return 0;
}

extern "C" {

int pk_batch_create(int num, const pk_mat* R, const pk_mat* F, pk_batch* out) {
   PK_CALL(ensure_init());
   PK_REQUIRE(num >= 1 && R && F && out, "wrong batch arguments");
   auto* Bt = new pk_batch_s();
   Bt->B = num;
   for (int b = 0; b < num; b++) {
      pk_mat_s* r = R[b];
      pk_mat_s* f = F[b];
      if (!r || !f || !r->assembled || !f->assembled || r->n != f->n || r->bs != f->bs ||
          f->nnz != r->n * (long)r->bs) {
         delete Bt;
         PK_FAIL("batch member " + std::to_string(b) +
                 ": R and F must be assembled, of one size, and F must hold one cell block per row");
      }
      Bt->R.push_back(r);
      Bt->F.push_back(f);
      Bt->n.push_back((int)r->n);
      Bt->fbs.push_back(r->bs);
      int p = 1;
      while (p * 2 <= kBMaxBs && r->bs % (p * 2) == 0) p *= 2;
      if (r->bs <= kBMaxBs && kRows % r->bs == 0) p = r->bs;
      Bt->pbs.push_back(p);
   }
   Bt->has_init.assign(num, 0);
   *out = Bt;
   return 0;
}

int pk_batch_set_tolerances(pk_batch Bt, double tol, int max_outer) {
   PK_REQUIRE(Bt && tol > 0.0 && max_outer > 0, "wrong tolerances");
   Bt->tol = tol;
   Bt->max_outer = max_outer;
   return 0;
}

int pk_batch_set_precision(pk_batch Bt, int precision) {
   PK_REQUIRE(Bt && (precision == PK_PREC_FP64 || precision == PK_PREC_MIXED), "wrong precision");
   PK_REQUIRE(!Bt->ready, "precision must be set before the first solve");
   Bt->precision = precision;
   return 0;
}

int pk_batch_set_inner(pk_batch Bt, int restart, double reduction, int max_inner) {
   PK_REQUIRE(Bt && restart >= 1 && restart <= kBMaxM && reduction > 0.0 && reduction < 1.0 && max_inner > 0,
              "wrong inner-solver options");
   PK_REQUIRE(!Bt->ready, "inner options must be set before the first solve");
   Bt->restart = restart;
   Bt->reduction = reduction;
   Bt->max_inner = max_inner;
   return 0;
}

int pk_batch_set_shift(pk_batch Bt, double delta) {
   PK_REQUIRE(Bt && delta >= 0.0, "wrong shift");
   Bt->delta = delta;
   return 0;
}

int pk_batch_set_initial(pk_batch Bt, int index, pk_vec x) {
   PK_REQUIRE(Bt && index >= 0 && index < Bt->B && x && x->n == Bt->n[index], "wrong initial vector");
   PK_CALL(batch_setup(Bt));
   if (!Bt->init.p) {
      PK_REQUIRE(!Bt->init.alloc(Bt->N), "device allocation failed");
      PK_HIP(hipMemsetAsync(Bt->init.p, 0, sizeof(double) * Bt->N, ctx().stream));
   }
   // an all-zero vector means "no initial space" (petsc.cxx:136-140)
   double nrm = 0.0;
   PK_CALL(pk_vec_dot(x, x, &nrm));
   Bt->has_init[index] = nrm > 0.0;
   if (nrm > 0.0)
      PK_HIP(hipMemcpyAsync(Bt->init.p + Bt->row_off[index], x->d.p, sizeof(double) * x->n, hipMemcpyDeviceToDevice,
                            ctx().stream));
   return 0;
}

int pk_batch_solve(pk_batch Bt) {
   PK_REQUIRE(Bt, "null batch");
   auto t0 = std::chrono::steady_clock::now();
   PK_CALL(batch_setup(Bt));
   int rc = Bt->precision == PK_PREC_MIXED ? batch_solve_t<float>(Bt) : batch_solve_t<double>(Bt);
   Bt->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   return rc;
}

int pk_batch_get_eigenpair(pk_batch Bt, int index, double* lambda, pk_vec x) {
   PK_REQUIRE(Bt && index >= 0 && index < Bt->B && !Bt->keff.empty(), "no solution for this batch member");
   if (lambda) *lambda = 1.0 / Bt->keff[index];
   if (x) {
      PK_REQUIRE(x->n == Bt->n[index], "vector size mismatch");
      PK_HIP(hipMemcpyAsync(x->d.p, Bt->phi.p + Bt->row_off[index], sizeof(double) * x->n,
                            hipMemcpyDeviceToDevice, ctx().stream));
      PK_HIP(hipStreamSynchronize(ctx().stream));
   }
   return 0;
}

int pk_batch_get_stats(pk_batch Bt, int index, int* outer, long* inner, double* residual, double* seconds) {
   PK_REQUIRE(Bt && index >= 0 && index < Bt->B && !Bt->keff.empty(), "no solution for this batch member");
   if (outer) *outer = Bt->outer[index];
   if (inner) *inner = Bt->inner[index];
   if (residual) *residual = Bt->resid[index];
   if (seconds) *seconds = Bt->seconds;
   return 0;
}

int pk_batch_destroy(pk_batch* Bt) {
   if (!Bt || !*Bt) return 0;
   (void)hipStreamSynchronize(ctx().stream);
   delete *Bt;
   *Bt = nullptr;
   return 0;
}

}  // extern "C"
