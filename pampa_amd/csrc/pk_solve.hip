// Krylov solvers of the pk_* layer: restarted GMRES with right cell-block
// Jacobi preconditioning for the linear (transient) solves, and the pk_eig
// interface, which runs one problem in the pooled eigen engine (pk_batch.hip,
// the same GMRES steps per slot).
//
// GMRES step j of a restart cycle is three kernels and no host round trip:
//   prec     V[j] *= 1/h(j,j-1) (normalise), z = B^-1 V[j]      (B = cell blocks)
//   spmv_dot w = A z, then h(i,j) = <V[i], w> for i <= j          (grid reduction)
//   orth     V[j+1] = w - sum_i h(i,j) V[i], h(j+1,j) = |V[j+1]|  (grid reduction)
// The last workgroup of `orth` applies the Givens rotations, updates the
// residual estimate and raises a stop flag that later kernels of the cycle
// read at entry, so a cycle is launched as a whole and synchronised once.
// Vectors/matrix are FP64 (plain drop-in) or FP32 (mixed precision; every
// reduction is still accumulated in FP64 across workgroups).
#include <algorithm>
#include <chrono>
#include <cmath>

#include "pk_internal.hpp"
#include "pk_reduce.hpp"

namespace pk {

// ---- cell-block inverses ------------------------------------------------------

// Gather the bs x bs diagonal block of cell c from SELL storage and invert it
// by Gauss-Jordan with partial pivoting (one thread per cell).
constexpr int kMaxBs = 16;
constexpr int kMaxRestart = 64;

template <class T>
__global__ void block_inverse_kernel(long ncell, int bs, const long* off, const int* rlen,
                                     const int* col, const double* val, T* binv, int* err) {
   long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (c >= ncell) return;
   double a[kMaxBs * kMaxBs], b[kMaxBs * kMaxBs];
   for (int p = 0; p < bs * bs; p++) {
      a[p] = 0.0;
      b[p] = 0.0;
   }
   for (int r = 0; r < bs; r++) {
      b[r * bs + r] = 1.0;
      long row = c * bs + r;
      long s = row >> 6;
      int lane = (int)(row & 63);
      for (int k = 0; k < rlen[row]; k++) {
         long p = off[s] + (long)k * 64 + lane;
         int cc = col[p];
         if (cc / bs == c) a[r * bs + (cc % bs)] += val[p];
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
            t = b[k * bs + q];
            b[k * bs + q] = b[piv * bs + q];
            b[piv * bs + q] = t;
         }
      double inv = 1.0 / a[k * bs + k];
      for (int q = 0; q < bs; q++) {
         a[k * bs + q] *= inv;
         b[k * bs + q] *= inv;
      }
      for (int r = 0; r < bs; r++) {
         if (r == k) continue;
         double f = a[r * bs + k];
         if (f == 0.0) continue;
         for (int q = 0; q < bs; q++) {
            a[r * bs + q] -= f * a[k * bs + q];
            b[r * bs + q] -= f * b[k * bs + q];
         }
      }
   }
   for (int p = 0; p < bs * bs; p++) binv[c * bs * bs + p] = (T)b[p];
}

// ---- GMRES kernels --------------------------------------------------------------

struct GmresState {
   double* H;      // (m+1) x m, column-major: H[i + j*(m+1)]
   double* cs;     // m
   double* sn;     // m
   double* g;      // m+1
   double* y;      // m
   double* scale;  // [0]: 1/h(j,j-1) for the next prec kernel
   double* resid;  // [0]: residual estimate |g(j+1)|; [1]: target
   int* flags;     // [0]: stop, [1]: steps used, [2]: breakdown
};

template <class T>
__global__ void __launch_bounds__(kBlock) prec_kernel(long n, int bs, T* __restrict__ Vj,
                                                      const T* __restrict__ binv, T* __restrict__ z,
                                                      GmresState S) {
   if (S.flags[0]) return;  // uniform: the flag only changes between kernels
   long r = blockIdx.x * (long)kBlock + threadIdx.x;
   const T sc = (T)S.scale[0];
   T own = 0;
   if (r < n) {
      long c = r / bs;
      int gr = (int)(r - c * bs);
      const T* bi = binv + c * bs * bs + (long)gr * bs;
      const T* v = Vj + c * bs;
      T acc = 0;
      for (int q = 0; q < bs; q++) acc += bi[q] * (v[q] * sc);
      z[r] = acc;
      own = Vj[r] * sc;
   }
   // every thread read its whole cell block of V[j]; normalise in place only
   // after the workgroup is done reading (cells never straddle workgroups since
   // kBlock % bs == 0, checked at setup)
   __syncthreads();
   if (r < n) Vj[r] = own;
}

template <class T>
__global__ void __launch_bounds__(kBlock) spmv_dot_kernel(long n, const long* __restrict__ off,
                                                          const int* __restrict__ width,
                                                          const int* __restrict__ col,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ z, T* __restrict__ w,
                                                          const T* __restrict__ V, int j, int m,
                                                          double* partials, unsigned* ticket,
                                                          GmresState S) {
   if (S.flags[0]) return;
   long row = blockIdx.x * (long)kBlock + threadIdx.x;
   T acc = 0;
   if (row < n) {
      long s = row >> 6;
      long base = off[s] + (threadIdx.x & 63);
      int wd = width[s];
      for (int k = 0; k < wd; k++) {
         long p = base + (long)k * 64;
         acc += val[p] * z[col[p]];
      }
      w[row] = acc;
   }
   const double wr = (double)acc;
   grid_reduce<kMaxRestart>(
       j + 1, [&](int k) { return row < n ? (double)V[(long)k * n + row] * wr : 0.0; }, partials,
       ticket, S.H + (long)j * (m + 1));
}

template <class T>
__global__ void __launch_bounds__(kBlock) orth_kernel(long n, const T* __restrict__ w, T* __restrict__ V,
                                                      int j, int m, double* partials,
                                                      unsigned* ticket, double* ww, GmresState S) {
   if (S.flags[0]) return;
   long row = blockIdx.x * (long)kBlock + threadIdx.x;
   double u = 0.0;
   if (row < n) {
      const double* h = S.H + (long)j * (m + 1);
      T t = w[row];
      for (int i = 0; i <= j; i++) t -= (T)h[i] * V[(long)i * n + row];
      V[(long)(j + 1) * n + row] = t;
      u = (double)t;
   }
   bool last = grid_reduce<1>(
       1, [&](int) { return u * u; }, partials, ticket, ww);
   if (!last) return;
   // Givens update of column j, operands staged in LDS by parallel loads
   __shared__ double hc[kMaxRestart + 2], cc[kMaxRestart], ss[kMaxRestart], gj;
   const int ld = m + 1;
   const int t = threadIdx.x;
   if (t <= j) hc[t] = S.H[t + (long)j * ld];
   if (t < j) {
      cc[t] = S.cs[t];
      ss[t] = S.sn[t];
   }
   if (t == 0) gj = S.g[j];
   __syncthreads();
   if (t == 0) {
      double hn = sqrt(fmax(ww[0], 0.0));
      hc[j + 1] = hn;
      for (int i = 0; i < j; i++) {
         double a = hc[i], b = hc[i + 1];
         hc[i] = cc[i] * a + ss[i] * b;
         hc[i + 1] = -ss[i] * a + cc[i] * b;
      }
      double a = hc[j], b = hc[j + 1];
      double rr = sqrt(a * a + b * b);
      double c = rr > 0.0 ? a / rr : 1.0, s = rr > 0.0 ? b / rr : 0.0;
      hc[j] = rr;
      hc[j + 1] = 0.0;
      for (int i = 0; i <= j + 1; i++) S.H[i + (long)j * ld] = hc[i];
      S.cs[j] = c;
      S.sn[j] = s;
      S.g[j + 1] = -s * gj;
      S.g[j] = c * gj;
      double res = fabs(s * gj);
      S.resid[0] = res;
      S.flags[1] = j + 1;
      bool breakdown = !(hn > 1e-300) || !(rr > 0.0);
      if (breakdown) S.flags[2] = 1;
      if (res <= S.resid[1] || breakdown || j + 1 == m) S.flags[0] = 1;
      S.scale[0] = hn > 0.0 ? 1.0 / hn : 0.0;
   }
}

// y = H(0:k,0:k)^-1 g(0:k) (upper triangular, k = steps used); operands are
// staged in LDS by the whole workgroup, then one thread back-substitutes
__global__ void __launch_bounds__(kBlock) gmres_trisolve_kernel(int m, GmresState S) {
   __shared__ double h[kMaxRestart * kMaxRestart], gg[kMaxRestart], yy[kMaxRestart];
   const int k = S.flags[1];
   const int ld = m + 1;
   for (int p = threadIdx.x; p < k * k; p += blockDim.x) {
      int i = p % k, q = p / k;
      h[i + q * kMaxRestart] = S.H[i + (long)q * ld];
   }
   if ((int)threadIdx.x < k) gg[threadIdx.x] = S.g[threadIdx.x];
   __syncthreads();
   if (threadIdx.x == 0) {
      for (int i = k - 1; i >= 0; i--) {
         double s = gg[i];
         for (int q = i + 1; q < k; q++) s -= h[i + q * kMaxRestart] * yy[q];
         double d = h[i + i * kMaxRestart];
         yy[i] = d != 0.0 ? s / d : 0.0;
      }
   }
   __syncthreads();
   if ((int)threadIdx.x < k) S.y[threadIdx.x] = yy[threadIdx.x];
}

// u = sum_i y_i V[i]
template <class T>
__global__ void combine_kernel(long n, const T* V, GmresState S, T* u) {
   long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (r >= n) return;
   int k = S.flags[1];
   T acc = 0;
   for (int i = 0; i < k; i++) acc += (T)S.y[i] * V[(long)i * n + r];
   u[r] = acc;
}

// e += B^-1 u
template <class T>
__global__ void block_apply_add_kernel(long n, int bs, const T* binv, const T* u, T* e) {
   long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (r >= n) return;
   long c = r / bs;
   int gr = (int)(r - c * bs);
   const T* bi = binv + c * bs * bs + (long)gr * bs;
   T acc = 0;
   for (int q = 0; q < bs; q++) acc += bi[q] * u[c * bs + q];
   e[r] += acc;
}

// V[0] = rhs - A e (restart residual), |V[0]|^2 -> out; sets up the cycle state
template <class T>
__global__ void __launch_bounds__(kBlock) restart_kernel(long n, const long* __restrict__ off,
                                                         const int* __restrict__ width,
                                                         const int* __restrict__ col,
                                                         const T* __restrict__ val,
                                                         const T* __restrict__ e,
                                                         const T* __restrict__ rhs, T* V, int use_e,
                                                         double* partials, unsigned* ticket,
                                                         double* out, GmresState S, double target) {
   long row = blockIdx.x * (long)kBlock + threadIdx.x;
   double u = 0.0;
   if (row < n) {
      T acc = 0;
      if (use_e) {
         long s = row >> 6;
         long base = off[s] + (threadIdx.x & 63);
         int wd = width[s];
         for (int k = 0; k < wd; k++) {
            long p = base + (long)k * 64;
            acc += val[p] * e[col[p]];
         }
      }
      T r = rhs[row] - acc;
      V[row] = r;
      u = (double)r;
   }
   bool last = grid_reduce<1>(
       1, [&](int) { return u * u; }, partials, ticket, out);
   if (last && threadIdx.x == 0) {
      double beta = sqrt(out[0]);
      S.g[0] = beta;
      S.resid[0] = beta;
      S.resid[1] = target;
      S.scale[0] = beta > 0.0 ? 1.0 / beta : 0.0;
      S.flags[0] = (beta <= target || !(beta > 0.0)) ? 1 : 0;
      S.flags[3] += S.flags[1];  // steps of the cycle that just ended
      S.flags[1] = 0;
      S.flags[2] = 0;
   }
}

template <class T>
__global__ void convert_kernel(const double* a, T* b, long n) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) b[i] = (T)a[i];
}

// ---- GMRES engine ----------------------------------------------------------------

// Solves A e = r (r FP64, e in precision T, e0 = 0) to |r - A e| <= reduction |r|
// or abs_target, within max_it iterations.
template <class T>
struct Gmres {
   long n = 0;
   int bs = 1, m = 30;
   const SellData* A = nullptr;
   const T* Aval = nullptr;
   DBuf<T> V, z, w, e, u, rhs, binv;
   DBuf<double> st;  // H, cs, sn, g, y, scale, resid, ww/out scratch
   DBuf<int> fl;
   GmresState S{};
   double* ww = nullptr;
   long last_iters = 0;
   double last_resid = 0.0, last_r0 = 0.0;

   int setup(const SellData* Amat, const T* vals, long n_, int bs_, int m_) {
      A = Amat;
      Aval = vals;
      n = n_;
      bs = bs_;
      m = m_;
      PK_REQUIRE(m >= 1 && m <= kMaxRestart, "wrong GMRES restart length");
      PK_REQUIRE(bs <= kMaxBs && kBlock % bs == 0, "cell block too large");
      PK_REQUIRE(!V.alloc((size_t)(m + 1) * n) && !z.alloc(n) && !w.alloc(n) && !e.alloc(n) &&
                     !u.alloc(n) && !rhs.alloc(n) && !binv.alloc((size_t)n * bs),
                 "device allocation failed");
      size_t ns = (size_t)(m + 1) * m + 3 * m + 2 + m + 8;
      PK_REQUIRE(!st.alloc(ns) && !fl.alloc(4), "device allocation failed");
      double* p = st.p;
      S.H = p;
      p += (size_t)(m + 1) * m;
      S.cs = p;
      p += m;
      S.sn = p;
      p += m;
      S.g = p;
      p += m + 1;
      S.y = p;
      p += m;
      S.scale = p;
      p += 1;
      S.resid = p;
      p += 2;
      ww = p;
      S.flags = fl.p;
      PK_CALL(reserve_partials((size_t)grid_for(n) * (m + 2)));
      return 0;
   }

   // binv from an FP64 SELL matrix (same pattern as A)
   int factor(const SellData& Ad) {
      Ctx& c = ctx();
      DBuf<int> err;
      PK_REQUIRE(!err.alloc(1), "device allocation failed");
      PK_HIP(hipMemsetAsync(err.p, 0, sizeof(int), c.stream));
      long ncell = n / bs;
      block_inverse_kernel<T><<<grid_for(ncell, 64), 64, 0, c.stream>>>(
          ncell, bs, Ad.off.p, Ad.rlen.p, Ad.col.p, Ad.val.p, binv.p, err.p);
      PK_HIP(hipGetLastError());
      int h = 0;
      PK_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      PK_REQUIRE(h == 0, "singular diagonal block in the block-Jacobi preconditioner");
      return 0;
   }

   // r: FP64 right-hand side on the device.  Result in e (precision T).
   int solve(const double* r, double reduction, double abs_target, int max_it) {
      Ctx& c = ctx();
      const int gb = grid_for(n);
      convert_kernel<T><<<gb, kBlock, 0, c.stream>>>(r, rhs.p, n);
      PK_HIP(hipMemsetAsync(e.p, 0, sizeof(T) * n, c.stream));
      PK_HIP(hipMemsetAsync(S.flags, 0, sizeof(int) * 4, c.stream));
      long its = 0;
      // first residual = rhs (e = 0); its norm fixes the target of this solve
      restart_kernel<T><<<gb, kBlock, 0, c.stream>>>(n, A->off.p, A->width.p, A->col.p, Aval, e.p,
                                                     rhs.p, V.p, 0, c.partials, c.ticket, ww, S, 0.0);
      PK_HIP(hipGetLastError());
      PK_HIP(hipMemcpyAsync(c.host_scalars, ww, sizeof(double), hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      const double r0 = sqrt(c.host_scalars[0]);
      last_r0 = r0;
      last_resid = r0;
      last_iters = 0;
      const double target = std::max(reduction * r0, abs_target);
      if (!(r0 > target)) return 0;
      c.host_scalars[0] = target;
      PK_HIP(hipMemcpyAsync(S.resid + 1, c.host_scalars, sizeof(double), hipMemcpyHostToDevice, c.stream));
      while (true) {
         int steps = (int)std::min<long>(m, max_it - its);
         const bool prof = profiling();
         const double s = sizeof(T);
         for (int j = 0; j < steps; j++) {
            // algorithmic bytes: prec reads V[j], the block inverses and writes z, V[j];
            // spmv_dot is a CSR SpMV plus the j+1 basis vectors; orth reads w and
            // j+1 basis vectors and writes V[j+1]
            PK_LAUNCH("lin_prec", (double)n * s * (3.0 + bs), prec_kernel<T>, gb, kBlock, c.stream, n, bs,
                      V.p + (size_t)j * n, binv.p, z.p, S);
            PK_LAUNCH("lin_spmv_dot",
                      (double)A->nnz * (s + 4.0) + (double)(n + 1) * 4.0 + 2.0 * n * s + (double)(j + 1) * n * s,
                      spmv_dot_kernel<T>, gb, kBlock, c.stream, n, A->off.p, A->width.p, A->col.p, Aval, z.p,
                      w.p, V.p, j, m, c.partials, c.ticket, S);
            PK_LAUNCH("lin_orth", (double)n * s * (j + 3.0), orth_kernel<T>, gb, kBlock, c.stream, n, w.p, V.p,
                      j, m, c.partials, c.ticket, ww, S);
            if (prof) {
               // profiling: issue only steps that do work
               int f0 = 0;
               PK_HIP(hipMemcpyAsync(&f0, S.flags, sizeof(int), hipMemcpyDeviceToHost, c.stream));
               PK_HIP(hipStreamSynchronize(c.stream));
               if (f0) break;
            }
         }
         PK_HIP(hipGetLastError());
         gmres_trisolve_kernel<<<1, kBlock, 0, c.stream>>>(m, S);
         combine_kernel<T><<<gb, kBlock, 0, c.stream>>>(n, V.p, S, u.p);
         block_apply_add_kernel<T><<<gb, kBlock, 0, c.stream>>>(n, bs, binv.p, u.p, e.p);
         // true residual of the updated correction
         restart_kernel<T><<<gb, kBlock, 0, c.stream>>>(n, A->off.p, A->width.p, A->col.p, Aval, e.p,
                                                        rhs.p, V.p, 1, c.partials, c.ticket, ww, S,
                                                        target);
         PK_HIP(hipGetLastError());
         int hf[4];
         PK_HIP(hipMemcpyAsync(c.host_scalars, ww, sizeof(double), hipMemcpyDeviceToHost, c.stream));
         PK_HIP(hipMemcpyAsync(hf, S.flags, sizeof(int) * 4, hipMemcpyDeviceToHost, c.stream));
         PK_HIP(hipStreamSynchronize(c.stream));
         its = hf[3];  // steps actually taken (accumulated by the restart kernel)
         double res = sqrt(c.host_scalars[0]);
         last_resid = res;
         last_iters = its;
         if (res <= target || its >= max_it || steps == 0) break;
      }
      return 0;
   }
};

}  // namespace pk

// ---- eigen solver -------------------------------------------------------------------
// One problem in the pooled engine (pk_batch.hip): the same kernels and tick
// graph that serve batches, with a pool of one slot kept for repeated solves.

using namespace pk;

struct pk_eig_s {
   pk_mat_s* R = nullptr;
   pk_mat_s* F = nullptr;
   pk_batch B = nullptr;
   bool solved = false;
};

extern "C" {

int pk_eig_create(pk_mat R, pk_mat F, pk_eig* E) {
   PK_CALL(ensure_init());
   PK_REQUIRE(E, "null output");
   pk_batch B = nullptr;
   PK_CALL(pk_batch_create(1, &R, &F, &B));
   auto* e = new pk_eig_s();
   e->R = R;
   e->F = F;
   e->B = B;
   *E = e;
   return 0;
}

int pk_eig_set_operators(pk_eig E, pk_mat R, pk_mat F) {
   PK_REQUIRE(E, "null eigen solver");
   PK_CALL(pk_batch_set_member(E->B, 0, R, F));
   E->R = R;
   E->F = F;
   E->solved = false;
   return 0;
}

int pk_eig_set_tolerances(pk_eig E, double tol, int max_outer) {
   PK_REQUIRE(E, "null eigen solver");
   return pk_batch_set_tolerances(E->B, tol, max_outer);
}

int pk_eig_set_precision(pk_eig E, int precision) {
   PK_REQUIRE(E, "null eigen solver");
   return pk_batch_set_precision(E->B, precision);
}

int pk_eig_set_inner(pk_eig E, int restart, double reduction, int max_inner) {
   PK_REQUIRE(E, "null eigen solver");
   return pk_batch_set_inner(E->B, restart, reduction, max_inner);
}

int pk_eig_set_preconditioner(pk_eig E, int poly) {
   PK_REQUIRE(E, "null eigen solver");
   return pk_batch_set_preconditioner(E->B, poly);
}

int pk_eig_set_shift(pk_eig E, double delta) {
   PK_REQUIRE(E, "null eigen solver");
   return pk_batch_set_shift(E->B, delta);
}

int pk_eig_set_acceleration(pk_eig E, int depth) {
   PK_REQUIRE(E, "null eigen solver");
   return pk_batch_set_acceleration(E->B, depth);
}

int pk_eig_set_initial(pk_eig E, pk_vec x) {
   PK_REQUIRE(E, "null eigen solver");
   return pk_batch_set_initial(E->B, 0, x);
}

int pk_eig_solve(pk_eig E) {
   PK_REQUIRE(E, "null eigen solver");
   E->solved = false;
   PK_CALL(pk_batch_solve(E->B));
   E->solved = true;
   return 0;
}

int pk_eig_get_eigenpair(pk_eig E, double* lambda, pk_vec x) {
   PK_REQUIRE(E && E->solved, "no eigen solution");
   return pk_batch_get_eigenpair(E->B, 0, lambda, x);
}

int pk_eig_get_stats(pk_eig E, int* outer, long* inner, double* residual, double* seconds) {
   PK_REQUIRE(E && E->solved, "no eigen solution");
   return pk_batch_get_stats(E->B, 0, outer, inner, residual, seconds);
}

int pk_eig_destroy(pk_eig* E) {
   if (!E || !*E) return 0;
   (void)pk_batch_destroy(&(*E)->B);
   delete *E;
   *E = nullptr;
   return 0;
}

}  // extern "C"

// ---- linear solver ---------------------------------------------------------------------

struct pk_lin_s {
   pk_mat_s* A = nullptr;
   double rtol = 1e-5, atol = 1e-50;  // PETSc KSP defaults
   int max_it = 10000;
   Gmres<double> g;
   DBuf<double> r;
   bool ready = false;
   long iters = 0;
   double resid = 0.0;
};

namespace pk {
__global__ void lin_resid_kernel(long n, const long* off, const int* width, const int* col,
                                 const double* val, const double* x, const double* b, double* r) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= n) return;
   long s = row >> 6;
   long base = off[s] + (row & 63);
   double acc = 0.0;
   for (int k = 0; k < width[s]; k++) {
      long p = base + (long)k * 64;
      acc += val[p] * x[col[p]];
   }
   r[row] = b[row] - acc;
}
__global__ void add_inplace_kernel(double* x, const double* e, long n) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) x[i] += e[i];
}
}  // namespace pk

extern "C" {

int pk_lin_create(pk_mat A, pk_lin* L) {
   PK_CALL(ensure_init());
   PK_REQUIRE(A && A->assembled, "matrix not assembled");
   auto* l = new pk_lin_s();
   l->A = A;
   *L = l;
   return 0;
}

int pk_lin_set_tolerances(pk_lin L, double rtol, double atol, int max_it) {
   PK_REQUIRE(L && rtol > 0.0 && atol >= 0.0 && max_it > 0, "wrong tolerances");
   L->rtol = rtol;
   L->atol = atol;
   L->max_it = max_it;
   return 0;
}

int pk_lin_solve(pk_lin L, pk_vec b, pk_vec x) {
   PK_REQUIRE(L && b && x && b->n == L->A->n && x->n == L->A->n, "vector size mismatch");
   Ctx& c = ctx();
   long n = L->A->n;
   // the solver's kernels read the 64-row slice layout; a matrix laid out
   // otherwise (by the eigen pool's autotuner) is laid out again, and the
   // solver set up again when its operator arrays moved
   if (L->A->sell.h != 64) {
      PK_CALL(build_sell(L->A, 64));
      L->ready = false;
   }
   if (L->ready && L->g.Aval != L->A->sell.val.p) L->ready = false;
   if (!L->ready) {
      PK_REQUIRE(!L->r.alloc(n), "device allocation failed");
      // block-Jacobi blocks: the cell block when whole cells fit a workgroup,
      // else its largest power-of-two divisor up to 16 (a block-diagonal part
      // of the cell block; point Jacobi for odd group counts)
      int pb = 1;
      while (pb * 2 <= kMaxBs && L->A->bs % (pb * 2) == 0) pb *= 2;
      if (L->A->bs <= kMaxBs && kBlock % L->A->bs == 0) pb = L->A->bs;
      PK_CALL(L->g.setup(&L->A->sell, L->A->sell.val.p, n, pb, 30));
      PK_CALL(L->g.factor(L->A->sell));
      L->ready = true;
   }
   double bn = 0.0;
   PK_CALL(pk_vec_norm2(b, &bn));
   double target = std::max(L->rtol * bn, L->atol);
   L->iters = 0;
   // defect correction: r = b - A x (FP64), GMRES on A e = r, x += e
   for (int pass = 0; pass < 100; pass++) {
      lin_resid_kernel<<<grid_for(n), kBlock, 0, c.stream>>>(n, L->A->sell.off.p, L->A->sell.width.p,
                                                             L->A->sell.col.p, L->A->sell.val.p,
                                                             x->d.p, b->d.p, L->r.p);
      PK_CALL(reduce_dot(L->r.p, L->r.p, n, c.scalars));
      PK_HIP(hipMemcpyAsync(c.host_scalars, c.scalars, sizeof(double), hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
      L->resid = sqrt(c.host_scalars[0]);
      if (L->resid <= target || L->iters >= L->max_it) break;
      PK_CALL(L->g.solve(L->r.p, 1e-30, target, (int)(L->max_it - L->iters)));
      L->iters += L->g.last_iters;
      add_inplace_kernel<<<grid_for(n), kBlock, 0, c.stream>>>(x->d.p, L->g.e.p, n);
   }
   PK_REQUIRE(L->resid <= target, "linear solve did not converge");
   return 0;
}

int pk_lin_get_stats(pk_lin L, long* iterations, double* residual) {
   PK_REQUIRE(L, "null linear solver");
   if (iterations) *iterations = L->iters;
   if (residual) *residual = L->resid;
   return 0;
}

int pk_lin_destroy(pk_lin* L) {
   if (!L || !*L) return 0;
   (void)hipStreamSynchronize(ctx().stream);
   delete *L;
   *L = nullptr;
   return 0;
}

}  // extern "C"
