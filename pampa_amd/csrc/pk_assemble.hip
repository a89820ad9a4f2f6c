// Device assembly of the finite-volume diffusion operators R and F from the
// face connectivity, and on-device power normalisation.
//
// The pattern is built on the device in closed form: row (i, g) holds the G
// columns (i, g2) of the cell block and one column (i2, g) per distinct
// physical neighbour, sorted by column; explicit zeros are kept.  That is the
// pattern PETSc's AIJ storage ends up with for the reference's
// MatSetValues(INSERT_VALUES) rows (DiffusionSolver.cxx:293-296).  Values are
// computed with the reference's operation order and without FMA contraction,
// so FP64 entries can be compared bit-for-bit.
#include <rocprim/device/device_scan.hpp>

#include "pk_internal.hpp"
#include "pk_reduce.hpp"

namespace pk {

// surface leakage factor w = (p2-p1).n / |p2-p1|^2 (math.cxx:150-161)
__device__ __forceinline__ double slf(const double* p1, const double* p2, const double* n) {
#pragma clang fp contract(off)
   double d0 = p2[0] - p1[0], d1 = p2[1] - p1[1], d2 = p2[2] - p1[2];
   double num = 0.0;
   num += d0 * n[0];
   num += d1 * n[1];
   num += d2 * n[2];
   double den = 0.0;
   den += d0 * d0;
   den += d1 * d1;
   den += d2 * d2;
   return num / den;
}

struct MeshView {
   int C;
   const int* fptr;
   const double* vol;
   const double* cen;
   const int* mat;
   const double* farea;
   const double* fcen;
   const double* fnorm;
   const int* fnb;
};

struct XSView {
   int G;
   const double* sigt;
   const double* sigs;
   const double* nusf;
   const double* kapf;
   const double* chi;
   const double* D;
   const int* set0;
   const int* nset;
   const double* tref;
   const double* T;  // cell temperatures (null: 0)
};

// Interpolation between the sets of material m at temperature T
// (FeedbackNuclearData.hxx:62-90): below the first or above the last
// tabulated temperature the end set; between two, linear with weight f on the
// upper one.  One set: its values as they are.
struct XSAt {
   int i1, i2;
   double f;
   bool one;
};

__device__ __forceinline__ XSAt xs_at(const XSView& X, int m, double T) {
#pragma clang fp contract(off)
   const int n = X.nset[m], s0 = X.set0[m];
   if (n == 1) return XSAt{s0, s0, 0.0, true};
   const double* Tr = X.tref + s0;
   if (T <= Tr[0]) return XSAt{s0, s0, 0.0, false};
   if (T > Tr[n - 1]) return XSAt{s0 + n - 1, s0 + n - 1, 0.0, false};
   int k = 1;  // first tabulated temperature >= T
   while (Tr[k] < T) k++;
   return XSAt{s0 + k - 1, s0 + k, (T - Tr[k - 1]) / (Tr[k] - Tr[k - 1]), false};
}

// value g of a [sets][G] table (or entry (from, to) of sigs with ld = G, g = from * G + to)
__device__ __forceinline__ double xv(const XSAt& a, const double* tab, int ld, int g) {
#pragma clang fp contract(off)
   if (a.one) return tab[(long)a.i1 * ld + g];
   return (1.0 - a.f) * tab[(long)a.i1 * ld + g] + a.f * tab[(long)a.i2 * ld + g];
}

__device__ __forceinline__ double cell_T(const XSView& X, int i) { return X.T ? X.T[i] : 0.0; }

MeshView view(pk_mesh_s* m) {
   return MeshView{m->C, m->fptr.p, m->vol.p, m->cen.p, m->mat.p, m->farea.p, m->fcen.p, m->fnorm.p,
                   m->fnb.p};
}
XSView view(pk_xs_s* x) {
   return XSView{x->G,      x->sigt.p, x->sigs.p, x->nusf.p, x->kapf.p, x->chi.p,
                 x->D.p,    x->set0.p, x->nset.p, x->tref.p, x->T};
}

// Distinct physical neighbours of a cell in face order (first occurrence).
__device__ __forceinline__ bool first_occurrence(const int* fnb, int f0, int f) {
   int i2 = fnb[f];
   for (int f2 = f0; f2 < f; f2++)
      if (fnb[f2] == i2) return false;
   return true;
}

__global__ void diff_rowlen_kernel(MeshView M, int G, int* rlen) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= M.C) return;
   int u = 0;
   for (int f = M.fptr[i]; f < M.fptr[i + 1]; f++)
      if (M.fnb[f] >= 0 && first_occurrence(M.fnb, M.fptr[i], f)) u++;
   for (int g = 0; g < G; g++) rlen[(long)i * G + g] = G + u;
}

constexpr int kMaxFaces = 32;

// One thread per row (i, g): values of R (and F when fcol != nullptr).
__global__ void diff_fill_kernel(MeshView M, XSView X, const int* bct, const double* bcv,
                                 const int* rptr, int* rcol, double* rval, int* fcol,
                                 double* fval, const double* vel, double dt, const double* chip,
                                 const double* beta, double keff, int* err) {
#pragma clang fp contract(off)
   long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   const int G = X.G;
   if (l >= (long)M.C * G) return;
   int i = (int)(l / G), g = (int)(l % G);
   int m = M.mat[i];
   double V = M.vol[i];
   const double* ci = M.cen + 3 * (long)i;
   // every coefficient at this cell's temperature, the neighbour's diffusion
   // coefficient included (DiffusionSolver.cxx:275 evaluates it at T_i)
   const double Ti = cell_T(X, i);
   const XSAt xa = xs_at(X, m, Ti);
   // diagonal: total reaction, then self-scattering (DiffusionSolver.cxx:130-131,157-158)
   double diag = xv(xa, X.sigt, G, g) * V;
   if (vel) diag += V / (xv(xa, vel, G, g) * dt);
   // neighbour entries in face order; a repeated neighbour keeps the last value
   int ncol[kMaxFaces];
   double nval[kMaxFaces];
   int nn = 0;
   // scattering / prompt-fission off-diagonals of the cell block come later;
   // the diagonal receives the self terms in g2 order
   for (int g2 = 0; g2 < G; g2++) {
      if (g2 == g) {
         diag += -xv(xa, X.sigs, G * G, g2 * G + g) * V;
         if (vel) diag += -(1.0 - beta[m]) * xv(xa, chip, G, g) * xv(xa, X.nusf, G, g2) * V / keff;
      }
   }
   for (int f = M.fptr[i]; f < M.fptr[i + 1]; f++) {
      int i2 = M.fnb[f];
      double A = M.farea[f];
      if (i2 < 0) {
         int id = -i2;
         int t = bct[id];
         if (t == PK_BC_VACUUM) {
            double w = slf(ci, M.fcen + 3 * (long)f, M.fnorm + 3 * (long)f);
            diag += w * xv(xa, X.D, G, g) * A;
         } else if (t == PK_BC_REFLECTIVE) {
         } else if (t == PK_BC_ROBIN) {
            diag -= bcv[id] * A;
         } else {
            *err = 1;
         }
      } else {
         int m2 = M.mat[i2];
         double r;
         if (m2 == m) {
            double w = slf(ci, M.cen + 3 * (long)i2, M.fnorm + 3 * (long)f);
            r = -w * xv(xa, X.D, G, g) * A;
         } else {
            double w1 = slf(ci, M.fcen + 3 * (long)f, M.fnorm + 3 * (long)f);
            w1 *= xv(xa, X.D, G, g) * A;
            double w2 = slf(M.cen + 3 * (long)i2, M.fcen + 3 * (long)f, M.fnorm + 3 * (long)f);
            w2 *= -xv(xs_at(X, m2, Ti), X.D, G, g) * A;
            r = -(w1 * w2) / (w1 + w2);
         }
         diag -= r;
         int k;
         for (k = 0; k < nn; k++)
            if (ncol[k] == i2) break;
         if (k == nn) {
            if (nn < kMaxFaces) {
               ncol[nn] = i2;
               nval[nn] = r;
               nn++;
            } else {
               *err = 2;
            }
         } else {
            nval[k] = r;
         }
      }
   }
   // sort neighbours by cell index (insertion sort, nn <= faces of the cell)
   for (int a = 1; a < nn; a++) {
      int c = ncol[a];
      double v = nval[a];
      int b = a - 1;
      while (b >= 0 && ncol[b] > c) {
         ncol[b + 1] = ncol[b];
         nval[b + 1] = nval[b];
         b--;
      }
      ncol[b + 1] = c;
      nval[b + 1] = v;
   }
   long p = rptr[l];
   int a = 0;
   for (; a < nn && ncol[a] < i; a++, p++) {
      rcol[p] = ncol[a] * G + g;
      rval[p] = nval[a];
   }
   for (int g2 = 0; g2 < G; g2++, p++) {
      rcol[p] = i * G + g2;
      if (g2 == g) {
         rval[p] = diag;
      } else {
         double v = -xv(xa, X.sigs, G * G, g2 * G + g) * V;
         if (vel) v += -(1.0 - beta[m]) * xv(xa, chip, G, g) * xv(xa, X.nusf, G, g2) * V / keff;
         rval[p] = v;
      }
   }
   for (; a < nn; a++, p++) {
      rcol[p] = ncol[a] * G + g;
      rval[p] = nval[a];
   }
   if (fcol) {
      long q = l * G;
      for (int g2 = 0; g2 < G; g2++) {
         fcol[q + g2] = i * G + g2;
         fval[q + g2] = xv(xa, X.chi, G, g) * xv(xa, X.nusf, G, g2) * V;
      }
   }
}

// b = chi_d S + V/(v dt) phi0 (DiffusionSolver.cxx:134-148)
__global__ void diff_rhs_kernel(MeshView M, XSView X, const double* chid, const double* vel, double dt,
                                const double* S, const double* phi0, double* b) {
#pragma clang fp contract(off)
   const int G = X.G;
   long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (l >= (long)M.C * G) return;
   int i = (int)(l / G), g = (int)(l % G);
   const XSAt a = xs_at(X, M.mat[i], cell_T(X, i));
   double v = xv(a, chid, G, g) * S[i];
   double d = M.vol[i] / (xv(a, vel, G, g) * dt);
   v += d * phi0[l];
   b[l] = v;
}

int diffusion_pattern(pk_mesh_s* mesh, int G, DBuf<int>& ptr, long& nnz) {
   Ctx& c = ctx();
   long N = (long)mesh->C * G;
   DBuf<int> rlen;
   PK_REQUIRE(!rlen.alloc(N + 1) && !ptr.alloc(N + 1), "device allocation failed");
   diff_rowlen_kernel<<<grid_for(mesh->C), kBlock, 0, c.stream>>>(view(mesh), G, rlen.p);
   PK_HIP(hipGetLastError());
   PK_HIP(hipMemsetAsync(rlen.p + N, 0, sizeof(int), c.stream));
   size_t tmp_bytes = 0;
   PK_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, rlen.p, ptr.p, 0, (size_t)(N + 1), rocprim::plus<int>(),
                                  c.stream));
   DBuf<char> tmp;
   PK_REQUIRE(!tmp.alloc(tmp_bytes ? tmp_bytes : 1), "device allocation failed");
   PK_HIP(rocprim::exclusive_scan(tmp.p, tmp_bytes, rlen.p, ptr.p, 0, (size_t)(N + 1), rocprim::plus<int>(),
                                  c.stream));
   int h = 0;
   PK_HIP(hipMemcpyAsync(&h, ptr.p + N, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   nnz = h;
   tmp.free();
   rlen.free();
   return 0;
}

__global__ void seq_ptr_kernel(int* ptr, long n, int w) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i <= n) ptr[i] = (int)(i * w);
}

// p0 = sum phi kapf V (NeutronicSolver.cxx:55-60)
__global__ void __launch_bounds__(kBlock) power_kernel(MeshView M, XSView X, const double* phi,
                                                        double* partials, unsigned* ticket,
                                                        double* out) {
   double s = 0.0;
   const int G = X.G;
   for (long l = blockIdx.x * (long)kBlock + threadIdx.x; l < (long)M.C * G;
        l += (long)gridDim.x * kBlock) {
      int i = (int)(l / G), g = (int)(l % G);
      s += phi[l] * xv(xs_at(X, M.mat[i], cell_T(X, i)), X.kapf, G, g) * M.vol[i];
   }
   double v[1] = {s};
   block_reduce_finish<1>(v, partials, ticket, out);
}

// phi *= f; q_i, P_i; count of negative fluxes (NeutronicSolver.cxx:63-68,88-104)
__global__ void __launch_bounds__(kBlock) scale_power_kernel(MeshView M, XSView X, double* phi,
                                                              const double* p0, double power,
                                                              double inv_keff, double* q,
                                                              double* P, double* partials,
                                                              unsigned* ticket, double* out) {
   const int G = X.G;
   double f = power >= 0.0 ? power / p0[0] : 1.0;  // power < 0: rates only, no scaling
   double neg = 0.0, tot = 0.0;
   for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < M.C; i += (long)gridDim.x * kBlock) {
      const XSAt a = xs_at(X, M.mat[i], cell_T(X, i));
      double qi = 0.0, Pi = 0.0;
      for (int g = 0; g < G; g++) {
         double v = phi[i * G + g] * f;
         phi[i * G + g] = v;
         if (v < 0.0) neg += 1.0;
         qi += v * xv(a, X.kapf, G, g) * M.vol[i];
         Pi += v * xv(a, X.nusf, G, g) * M.vol[i] * inv_keff;
      }
      if (q) q[i] = qi;
      if (P) P[i] = Pi;
      tot += qi;
   }
   double v[2] = {neg, tot};
   block_reduce_finish<2>(v, partials, ticket, out);
}

// ---- reference-form S_N operator (SNSolver.cxx:159-208,344-631) ----------------------

// Face-interpolation weights of interior faces (SNSolver.cxx:193-198):
// coef[4 f + 0..1] outgoing (w_i, w_i2), coef[4 f + 2..3] incoming.
__global__ void sn_gauss_kernel(MeshView M, double delta, double* coef) {
#pragma clang fp contract(off)
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= M.C) return;
   const double* ci = M.cen + 3 * (long)i;
   for (int f = M.fptr[i]; f < M.fptr[i + 1]; f++) {
      const int i2 = M.fnb[f];
      if (i2 < 0) continue;
      const double* xf = M.fcen + 3 * (long)f;
      const double* c2 = M.cen + 3 * (long)i2;
      // math::distance: squares summed in coordinate order (math.cxx:82-92)
      double rif = 0.0, ri2f = 0.0, rii2 = 0.0;
      for (int k = 0; k < 3; k++) {
         double a = ci[k] - xf[k], b = c2[k] - xf[k], d = c2[k] - ci[k];
         rif += a * a;
         ri2f += b * b;
         rii2 += d * d;
      }
      rif = sqrt(rif);
      ri2f = sqrt(ri2f);
      rii2 = sqrt(rii2);
      coef[4 * (long)f + 0] = (ri2f + delta * rif) / rii2;
      coef[4 * (long)f + 1] = (1.0 - delta) * rif / rii2;
      coef[4 * (long)f + 2] = (1.0 - delta) * ri2f / rii2;
      coef[4 * (long)f + 3] = (rif + delta * ri2f) / rii2;
   }
}

struct SnRefView {
   int M;               // directions
   const double* dirs;  // [M][3]
   const double* wts;   // [M]
   const int* refl;     // [M][3] reflected direction per axis
   const double* coef;  // [faces][4]
   const double* lsc;   // [faces][3] least-squares weights of boundary cells, or null (LS off)
   const int* bct;      // boundary type by bc id
};

constexpr int kMaxRefl = 8;

// One thread per row l = (i G + g) M + m of R (and of F when fcol != null).
// R accumulates in the reference's order: per face, the MatSetValues(ADD)
// calls of the neighbour, least-squares and reflected entries
// (SNSolver.cxx:512,551,595), then the block row (:602), whose diagonal holds
// the total reaction, the self-scattering and then every face term in face
// order (:397,426-428,479-597).  A neighbour reached through several faces
// keeps one entry (sums in face order); the row is written sorted by column.
__global__ void sn_fill_kernel(MeshView M, XSView X, SnRefView Q, const int* rptr, int* rcol, double* rval,
                               int* fcol, double* fval, int* err) {
#pragma clang fp contract(off)
   const int G = X.G, Md = Q.M, GM = G * Md;
   const long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (l >= (long)M.C * GM) return;
   const int i = (int)(l / GM), gm = (int)(l % GM), g = gm / Md, m = gm % Md;
   const int mt = M.mat[i];
   const double V = M.vol[i];
   const double* om = Q.dirs + 3 * m;
   const double* ci = M.cen + 3 * (long)i;
   const XSAt xa = xs_at(X, mt, cell_T(X, i));
   double diag = xv(xa, X.sigt, G, g) * V;
   diag += -xv(xa, X.sigs, G * G, g * G + g) * Q.wts[m] * V;
   int ncol[kMaxFaces];
   double nval[kMaxFaces];
   int nn = 0;
   int rdir[kMaxRefl];
   double rv[kMaxRefl];
   int nr = 0;
   auto add_nb = [&](int c, double v) {
      int k = 0;
      while (k < nn && ncol[k] != c) k++;
      if (k == nn) {
         if (nn == kMaxFaces) {
            *err = 2;
            return;
         }
         ncol[nn] = c;
         nval[nn++] = v;
      } else {
         nval[k] += v;
      }
   };
   const int f0 = M.fptr[i], f1 = M.fptr[i + 1];
   for (int f = f0; f < f1; f++) {
      const int i2 = M.fnb[f];
      const double* nf = M.fnorm + 3 * (long)f;
      double w = 0.0;
      w += om[0] * nf[0];
      w += om[1] * nf[1];
      w += om[2] * nf[2];
      const double A = M.farea[f];
      if (i2 < 0) {
         const int t = Q.bct[-i2];
         if (t == PK_BC_VACUUM) {
            if (w > 0.0) {
               diag += w * A;
               if (Q.lsc) {
                  const double* xf = M.fcen + 3 * (long)f;
                  const double dp[3] = {xf[0] - ci[0], xf[1] - ci[1], xf[2] - ci[2]};
                  for (int f2 = f0; f2 < f1; f2++) {
                     const int i3 = M.fnb[f2];
                     if (i3 < 0) continue;
                     const double* gc = Q.lsc + 3 * (long)f2;
                     double wi3 = 0.0;
                     wi3 += dp[0] * gc[0];
                     wi3 += dp[1] * gc[1];
                     wi3 += dp[2] * gc[2];
                     diag += -wi3 * w * A;
                     add_nb(i3, wi3 * w * A);
                  }
               }
            }
         } else if (t == PK_BC_REFLECTIVE) {
            if (w > 0.0) {
               diag += w * A;
            } else {
               // reflection axis from the normal of face (f - f0) of cells 0, 1, 2
               // (the cell index is shadowed there, SNSolver.cxx:541-545)
               const int lf = f - f0;
               int m2 = -1;
               for (int a = 0; a < 3; a++) {
                  if (a >= M.C || lf >= M.fptr[a + 1] - M.fptr[a]) {
                     *err = 3;
                     return;
                  }
                  const double d = M.fnorm[3 * (long)(M.fptr[a] + lf) + a];
                  if (fabs(d) > 1.0 - 1.0e-6) m2 = Q.refl[3 * m + a];
               }
               if (m2 < 0) {
                  *err = 3;
                  return;
               }
               int k = 0;
               while (k < nr && rdir[k] != m2) k++;
               if (k == nr) {
                  if (nr == kMaxRefl) {
                     *err = 2;
                     return;
                  }
                  rdir[nr] = m2;
                  rv[nr++] = w * A;
               } else {
                  rv[k] += w * A;
               }
            }
         } else {
            *err = 1;
         }
      } else {
         const double* cf = Q.coef + 4 * (long)f;
         const double wi = w > 0.0 ? cf[0] : cf[2];
         const double wi2 = w > 0.0 ? cf[1] : cf[3];
         diag += wi * w * A;
         add_nb(i2, wi2 * w * A);
      }
   }
   // neighbours by cell index
   for (int a = 1; a < nn; a++) {
      const int c = ncol[a];
      const double v = nval[a];
      int b = a - 1;
      while (b >= 0 && ncol[b] > c) {
         ncol[b + 1] = ncol[b];
         nval[b + 1] = nval[b];
         b--;
      }
      ncol[b + 1] = c;
      nval[b + 1] = v;
   }
   long p = rptr[l];
   int a = 0;
   for (; a < nn && ncol[a] < i; a++, p++) {
      rcol[p] = ncol[a] * GM + gm;
      rval[p] = nval[a];
   }
   const long c0 = (long)i * GM;
   for (int g2 = 0; g2 < G; g2++) {
      const double sv = -xv(xa, X.sigs, G * G, g2 * G + g);
      for (int m2 = 0; m2 < Md; m2++, p++) {
         rcol[p] = (int)(c0 + g2 * Md + m2);
         double v = (g2 == g && m2 == m) ? diag : sv * Q.wts[m2] * V;
         if (g2 == g)
            for (int k = 0; k < nr; k++)
               if (rdir[k] == m2) v = rv[k] + v;
         rval[p] = v;
      }
   }
   for (; a < nn; a++, p++) {
      rcol[p] = ncol[a] * GM + gm;
      rval[p] = nval[a];
   }
   if (fcol) {
      const long q = l * GM;
      const double cn = xv(xa, X.chi, G, g);
      for (int g2 = 0; g2 < G; g2++)
         for (int m2 = 0; m2 < Md; m2++) {
            fcol[q + g2 * Md + m2] = (int)(c0 + g2 * Md + m2);
            fval[q + g2 * Md + m2] = cn * xv(xa, X.nusf, G, g2) * Q.wts[m2] * V;
         }
   }
}

// counters that must start at zero (a fill kernel, not a memset node: the
// call may sit inside a captured graph)
__global__ void zero_u32_kernel(unsigned* p, int n) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = 0u;
}

// The two kernels of the power normalisation with reduction scratch at `scr`
// (partials, scalars, tickets at the offsets of the static-scratch layout).
static int normalize_flux_launch(pk_mesh_s* mesh, pk_xs_s* xs, pk_vec_s* phi, double power, double keff,
                                 pk_vec_s* q, pk_vec_s* P, double* partials, unsigned* ticket, double* scalars) {
   Ctx& c = ctx();
   int g = grid_for((long)mesh->C * xs->G, kBlock, kMaxRedBlocks);
   power_kernel<<<g, kBlock, 0, c.stream>>>(view(mesh), view(xs), phi->d.p, partials, ticket, scalars);
   PK_HIP(hipGetLastError());
   int g2 = grid_for(mesh->C, kBlock, kMaxRedBlocks);
   scale_power_kernel<<<g2, kBlock, 0, c.stream>>>(view(mesh), view(xs), phi->d.p, scalars, power, 1.0 / keff,
                                                   q ? q->d.p : nullptr, P ? P->d.p : nullptr, partials, ticket,
                                                   scalars + 1);
   PK_HIP(hipGetLastError());
   return 0;
}

}  // namespace pk

using namespace pk;

template <class T>
static int upload(DBuf<T>& d, const T* h, size_t n) {
   if (d.alloc(n ? n : 1)) return 1;
   if (n && hipMemcpy(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return 1;
   return 0;
}

extern "C" {

int pk_mesh_create(int num_cells, const int* face_ptr, const double* vol, const double* cen,
                   const int* mat, const double* farea, const double* fcen, const double* fnorm,
                   const int* fnb, pk_mesh* mesh) {
   PK_CALL(ensure_init());
   PK_REQUIRE(num_cells >= 0 && face_ptr, "wrong mesh arguments");
   auto* m = new pk_mesh_s();
   m->C = num_cells;
   m->F = face_ptr[num_cells];
   int fmax = 0;
   for (int i = 0; i < num_cells; i++) fmax = std::max(fmax, face_ptr[i + 1] - face_ptr[i]);
   m->fmax = fmax;
   if (fmax > kMaxFaces) {
      delete m;
      PK_FAIL("too many faces per cell");
   }
   m->h_fptr.assign(face_ptr, face_ptr + num_cells + 1);
   m->h_vol.assign(vol, vol + num_cells);
   m->h_cen.assign(cen, cen + 3L * num_cells);
   m->h_mat.assign(mat, mat + num_cells);
   m->h_farea.assign(farea, farea + m->F);
   m->h_fcen.assign(fcen, fcen + 3L * m->F);
   m->h_fnorm.assign(fnorm, fnorm + 3L * m->F);
   m->h_fnb.assign(fnb, fnb + m->F);
   int e = 0;
   e |= upload(m->fptr, face_ptr, num_cells + 1);
   e |= upload(m->vol, vol, num_cells);
   e |= upload(m->cen, cen, 3L * num_cells);
   e |= upload(m->mat, mat, num_cells);
   e |= upload(m->farea, farea, m->F);
   e |= upload(m->fcen, fcen, 3L * m->F);
   e |= upload(m->fnorm, fnorm, 3L * m->F);
   e |= upload(m->fnb, fnb, m->F);
   if (e) {
      pk_mesh_destroy(&m);
      PK_FAIL("device allocation failed");
   }
   *mesh = m;
   return 0;
}

int pk_mesh_destroy(pk_mesh* mesh) {
   if (!mesh || !*mesh) return 0;
   pk_mesh_s* m = *mesh;
   m->fptr.free();
   m->vol.free();
   m->cen.free();
   m->mat.free();
   m->farea.free();
   m->fcen.free();
   m->fnorm.free();
   m->fnb.free();
   delete m;
   *mesh = nullptr;
   return 0;
}

int pk_xs_create(int num_mat, int G, const double* sigt, const double* sigs, const double* nusf,
                 const double* kapf, const double* chi, const double* D, pk_xs* xs) {
   return pk_xs_create_sets(num_mat, G, nullptr, nullptr, sigt, sigs, nusf, kapf, chi, D, xs);
}

int pk_xs_create_sets(int num_mat, int G, const int* num_sets, const double* temps, const double* sigt,
                      const double* sigs, const double* nusf, const double* kapf, const double* chi,
                      const double* D, pk_xs* xs) {
   PK_CALL(ensure_init());
   PK_REQUIRE(num_mat > 0 && G > 0, "wrong cross-section table size");
   std::vector<int> set0(num_mat), nset(num_mat, 1);
   int ns = 0;
   for (int m = 0; m < num_mat; m++) {
      if (num_sets) nset[m] = num_sets[m];
      PK_REQUIRE(nset[m] >= 1, "a material needs at least one cross-section set");
      set0[m] = ns;
      ns += nset[m];
   }
   std::vector<double> tref(ns, 0.0);
   if (temps) {
      tref.assign(temps, temps + ns);
      for (int m = 0; m < num_mat; m++)
         for (int k = 1; k < nset[m]; k++)
            PK_REQUIRE(tref[set0[m] + k] > tref[set0[m] + k - 1], "set temperatures must increase");
   }
   auto* x = new pk_xs_s();
   x->nm = num_mat;
   x->G = G;
   x->ns = ns;
   x->feedback = ns > num_mat;
   if (upload(x->set0, set0.data(), num_mat) || upload(x->nset, nset.data(), num_mat) ||
       upload(x->tref, tref.data(), ns)) {
      pk_xs_destroy(&x);
      PK_FAIL("device allocation failed");
   }
   size_t n1 = (size_t)ns * G, n2 = n1 * G;
   x->h_sigt.assign(sigt, sigt + n1);
   x->h_sigs.assign(sigs, sigs + n2);
   x->h_nusf.assign(nusf, nusf + n1);
   x->h_kapf.assign(kapf, kapf + n1);
   x->h_chi.assign(chi, chi + n1);
   if (D)
      x->h_D.assign(D, D + n1);
   else
      x->h_D.assign(n1, 0.0);
   int e = 0;
   e |= upload(x->sigt, x->h_sigt.data(), n1);
   e |= upload(x->sigs, x->h_sigs.data(), n2);
   e |= upload(x->nusf, x->h_nusf.data(), n1);
   e |= upload(x->kapf, x->h_kapf.data(), n1);
   e |= upload(x->chi, x->h_chi.data(), n1);
   e |= upload(x->D, x->h_D.data(), n1);
   if (e) {
      pk_xs_destroy(&x);
      PK_FAIL("device allocation failed");
   }
   *xs = x;
   return 0;
}

int pk_xs_destroy(pk_xs* xs) {
   if (!xs || !*xs) return 0;
   pk_xs_s* x = *xs;
   x->sigt.free();
   x->sigs.free();
   x->nusf.free();
   x->kapf.free();
   x->chi.free();
   x->D.free();
   delete x;
   *xs = nullptr;
   return 0;
}

static int diffusion_assemble_impl(pk_mesh mesh, pk_xs xs, int num_bc, const int* bc_type,
                                   const double* bc_value, const double* vel, double dt,
                                   const double* chip, const double* beta, double keff,
                                   pk_mat* R, pk_mat* F) {
   PK_REQUIRE(mesh && xs && R, "null argument");
   Ctx& c = ctx();
   const int G = xs->G;
   long N = (long)mesh->C * G;
   PK_REQUIRE(N < (1L << 31), "operator too large for 32-bit indices");
   DBuf<int> dbct, derr;
   DBuf<double> dbcv, dvdt, dchip, dbeta;
   PK_REQUIRE(!upload(dbct, bc_type, num_bc) && !upload(dbcv, bc_value, num_bc) && !derr.alloc(1),
              "device allocation failed");
   if (vel) {
      PK_REQUIRE(!upload(dvdt, vel, (size_t)xs->ns * G) && !upload(dchip, chip, (size_t)xs->ns * G) &&
                     !upload(dbeta, beta, (size_t)xs->nm),
                 "device allocation failed");
   }
   PK_HIP(hipMemsetAsync(derr.p, 0, sizeof(int), c.stream));
   pk_mat_s* r = new pk_mat_s();
   r->n = N;
   r->bs = G;
   PK_CALL(diffusion_pattern(mesh, G, r->ptr, r->nnz));
   PK_REQUIRE(!r->col.alloc(r->nnz) && !r->val.alloc(r->nnz), "device allocation failed");
   pk_mat_s* f = nullptr;
   if (F) {
      f = new pk_mat_s();
      f->n = N;
      f->bs = G;
      f->nnz = N * G;
      PK_REQUIRE(!f->ptr.alloc(N + 1) && !f->col.alloc(f->nnz) && !f->val.alloc(f->nnz),
                 "device allocation failed");
      seq_ptr_kernel<<<grid_for(N + 1), kBlock, 0, c.stream>>>(f->ptr.p, N, G);
   }
   diff_fill_kernel<<<grid_for(N), kBlock, 0, c.stream>>>(
       view(mesh), view(xs), dbct.p, dbcv.p, r->ptr.p, r->col.p, r->val.p, f ? f->col.p : nullptr,
       f ? f->val.p : nullptr, vel ? dvdt.p : nullptr, dt, dchip.p, dbeta.p, keff, derr.p);
   PK_HIP(hipGetLastError());
   int herr = 0;
   PK_HIP(hipMemcpyAsync(&herr, derr.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   if (herr) {
      pk_mat_destroy(&r);
      if (f) pk_mat_destroy(&f);
      PK_FAIL(herr == 1 ? "boundary condition not implemented" : "too many neighbours per cell");
   }
   PK_CALL(finalize_device_csr(r));
   if (f) PK_CALL(finalize_device_csr(f));
   if (!vel) {
      r->asm_bct = std::move(dbct);
      r->asm_bcv = std::move(dbcv);
   }
   *R = r;
   if (F) *F = f;
   return 0;
}

int pk_diffusion_assemble(pk_mesh mesh, pk_xs xs, int num_bc, const int* bc_type,
                          const double* bc_value, pk_mat* R, pk_mat* F) {
   return diffusion_assemble_impl(mesh, xs, num_bc, bc_type, bc_value, nullptr, 0.0, nullptr,
                                  nullptr, 1.0, R, F);
}

int pk_sn_assemble(pk_mesh mesh, pk_xs xs, int M, const double* dirs, const double* wts, const int* refl,
                   double delta, const double* ls_coef, int num_bc, const int* bc_type, pk_mat* R, pk_mat* F) {
   PK_REQUIRE(mesh && xs && R && F && dirs && wts && refl && bc_type && M > 0, "null argument");
   Ctx& c = ctx();
   const int G = xs->G, GM = G * M;
   const long N = (long)mesh->C * GM;
   PK_REQUIRE(N < (1L << 31) && N * (long)GM < (1L << 40), "operator too large for 32-bit indices");
   DBuf<double> ddir, dwts, dcoef, dlsc;
   DBuf<int> drefl, dbct, derr;
   PK_REQUIRE(!upload(ddir, dirs, (size_t)3 * M) && !upload(dwts, wts, (size_t)M) &&
                  !upload(drefl, refl, (size_t)3 * M) && !upload(dbct, bc_type, (size_t)num_bc) &&
                  !dcoef.alloc((size_t)4 * (mesh->F ? mesh->F : 1)) && !derr.alloc(1),
              "device allocation failed");
   if (ls_coef) PK_REQUIRE(!upload(dlsc, ls_coef, (size_t)3 * mesh->F), "device allocation failed");
   PK_HIP(hipMemsetAsync(derr.p, 0, sizeof(int), c.stream));
   PK_HIP(hipMemsetAsync(dcoef.p, 0, sizeof(double) * 4 * (mesh->F ? mesh->F : 1), c.stream));
   sn_gauss_kernel<<<grid_for(mesh->C), kBlock, 0, c.stream>>>(view(mesh), delta, dcoef.p);
   PK_HIP(hipGetLastError());
   pk_mat_s* r = new pk_mat_s();
   r->n = N;
   r->bs = GM;
   pk_mat_s* f = new pk_mat_s();
   f->n = N;
   f->bs = GM;
   f->nnz = N * GM;
   // R: the cell's G M block and one entry per distinct physical neighbour
   // (the diffusion pattern with blocks of G M rows)
   if (diffusion_pattern(mesh, GM, r->ptr, r->nnz) || r->col.alloc(r->nnz) || r->val.alloc(r->nnz) ||
       f->ptr.alloc(N + 1) || f->col.alloc(f->nnz) || f->val.alloc(f->nnz)) {
      pk_mat_destroy(&r);
      pk_mat_destroy(&f);
      PK_FAIL("device allocation failed for the S_N operator");
   }
   seq_ptr_kernel<<<grid_for(N + 1), kBlock, 0, c.stream>>>(f->ptr.p, N, GM);
   SnRefView Q{M, ddir.p, dwts.p, drefl.p, dcoef.p, ls_coef ? dlsc.p : nullptr, dbct.p};
   sn_fill_kernel<<<grid_for(N), kBlock, 0, c.stream>>>(view(mesh), view(xs), Q, r->ptr.p, r->col.p, r->val.p,
                                                        f->col.p, f->val.p, derr.p);
   PK_HIP(hipGetLastError());
   int herr = 0;
   PK_HIP(hipMemcpyAsync(&herr, derr.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   if (herr) {
      pk_mat_destroy(&r);
      pk_mat_destroy(&f);
      PK_FAIL(herr == 1   ? "boundary condition not implemented"
              : herr == 2 ? "too many neighbours per cell"
                          : "reflected direction not found");
   }
   PK_CALL(finalize_device_csr(r));
   PK_CALL(finalize_device_csr(f));
   *R = r;
   *F = f;
   return 0;
}

int pk_diffusion_refill(pk_mesh mesh, pk_xs xs, pk_mat R, pk_mat F) {
   PK_REQUIRE(mesh && xs && R && F && R->assembled && F->assembled, "null or unassembled argument");
   PK_REQUIRE(R->asm_bct.p, "R does not come from a steady diffusion assembly");
   const int G = xs->G;
   const long N = (long)mesh->C * G;
   PK_REQUIRE(R->n == N && F->n == N && R->bs == G && F->nnz == N * G,
              "mesh or group count differ from the assembly of R and F");
   Ctx& c = ctx();
   DBuf<int>& derr = R->asm_err;
   if (!derr.p) PK_REQUIRE(!derr.alloc(1), "device allocation failed");
   PK_HIP(hipMemsetAsync(derr.p, 0, sizeof(int), c.stream));
   // same mesh and boundary table: the pattern (and SELL layout) is unchanged
   diff_fill_kernel<<<grid_for(N), kBlock, 0, c.stream>>>(view(mesh), view(xs), R->asm_bct.p, R->asm_bcv.p,
                                                          R->ptr.p, R->col.p, R->val.p, F->col.p, F->val.p,
                                                          nullptr, 0.0, nullptr, nullptr, 1.0, derr.p);
   PK_HIP(hipGetLastError());
   PK_CALL(refresh_sell_values(R));
   PK_CALL(refresh_sell_values(F));
   int herr = 0;
   PK_HIP(hipMemcpyAsync(&herr, derr.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   PK_REQUIRE(herr == 0, herr == 1 ? "boundary condition not implemented" : "too many neighbours per cell");
   return 0;
}

int pk_xs_set_temperature(pk_xs xs, pk_vec T) {
   PK_REQUIRE(xs, "null argument");
   xs->T = T ? T->d.p : nullptr;
   return 0;
}

int pk_xs_set(pk_xs xs, const double* sigt, const double* sigs, const double* nusf, const double* kapf,
              const double* chi, const double* D) {
   PK_REQUIRE(xs && sigt && sigs && nusf && kapf && chi, "null argument");
   Ctx& c = ctx();
   const size_t n1 = (size_t)xs->ns * xs->G, n2 = n1 * xs->G;
   xs->h_sigt.assign(sigt, sigt + n1);
   xs->h_sigs.assign(sigs, sigs + n2);
   xs->h_nusf.assign(nusf, nusf + n1);
   xs->h_kapf.assign(kapf, kapf + n1);
   xs->h_chi.assign(chi, chi + n1);
   if (D) xs->h_D.assign(D, D + n1);
   // ordered on the library stream after any kernel still reading the tables
   PK_HIP(hipMemcpyAsync(xs->sigt.p, xs->h_sigt.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->sigs.p, xs->h_sigs.data(), sizeof(double) * n2, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->nusf.p, xs->h_nusf.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->kapf.p, xs->h_kapf.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->chi.p, xs->h_chi.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipMemcpyAsync(xs->D.p, xs->h_D.data(), sizeof(double) * n1, hipMemcpyHostToDevice, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

int pk_diffusion_assemble_transient(pk_mesh mesh, pk_xs xs, int num_bc, const int* bc_type,
                                    const double* bc_value, const double* velocity, double dt,
                                    const double* chip, const double* chid, const double* beta,
                                    double keff, pk_vec S, pk_vec phi0, pk_mat* R, pk_vec b) {
   PK_REQUIRE(velocity && chip && chid && beta && S && phi0 && b, "null argument");
   PK_CALL(diffusion_assemble_impl(mesh, xs, num_bc, bc_type, bc_value, velocity, dt, chip, beta,
                                   keff, R, nullptr));
   Ctx& c = ctx();
   const int G = xs->G;
   DBuf<double> dchid, dvdt;
   PK_REQUIRE(!upload(dchid, chid, (size_t)xs->ns * G) && !upload(dvdt, velocity, (size_t)xs->ns * G),
              "device allocation failed");
   diff_rhs_kernel<<<grid_for((long)mesh->C * G), kBlock, 0, c.stream>>>(
       view(mesh), view(xs), dchid.p, dvdt.p, dt, S->d.p, phi0->d.p, b->d.p);
   PK_HIP(hipGetLastError());
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

int pk_normalize_flux(pk_mesh mesh, pk_xs xs, pk_vec phi, double power, double keff, pk_vec q,
                      pk_vec P, double* total_power) {
   PK_REQUIRE(mesh && xs && phi && phi->n == (long)mesh->C * xs->G, "wrong arguments");
   Ctx& c = ctx();
   PK_CALL(normalize_flux_launch(mesh, xs, phi, power, keff, q, P, c.partials, c.ticket, c.scalars));
   PK_HIP(hipMemcpyAsync(c.host_scalars, c.scalars, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   if (total_power) *total_power = c.host_scalars[2];
   PK_REQUIRE(c.host_scalars[1] == 0.0, "negative values in the scalar-flux solution");
   return 0;
}

int pk_normalize_flux_static(pk_mesh mesh, pk_xs xs, pk_vec phi, double power, double keff, pk_vec q,
                             pk_vec P, double* total_power) {
   Ctx& c = ctx();
   PK_REQUIRE(c.stream, "pk_initialize must run before the allocation-free entry points");
   double* scr = static_scratch();
   PK_REQUIRE(scr, "no static scratch on this device");
   PK_REQUIRE(mesh && xs && phi && phi->n == (long)mesh->C * xs->G, "wrong arguments");
   PK_REQUIRE(q == nullptr || q->n == mesh->C, "wrong size of the power vector");
   PK_REQUIRE(P == nullptr || P->n == mesh->C, "wrong size of the production vector");
   // the reductions use at most kMaxRedBlocks workgroups x 2 values of partials;
   // the kernels index cells and rows with 32-bit integers
   PK_REQUIRE((long)mesh->C * xs->G <= 0x7fffffffL,
              "flux larger than the static-scratch path supports (cells x groups must be below 2^31)");
   unsigned* tickets = (unsigned*)(scr + kScratchTickets);
   zero_u32_kernel<<<1, 64, 0, c.stream>>>(tickets, 16);
   PK_HIP(hipGetLastError());
   PK_CALL(normalize_flux_launch(mesh, xs, phi, power, keff, q, P, scr, tickets, scr + kScratchScalars));
   double h[3] = {0.0, 0.0, 0.0};
   PK_HIP(hipMemcpyAsync(h, scr + kScratchScalars, sizeof(h), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   if (total_power) *total_power = h[2];
   PK_REQUIRE(h[1] == 0.0, "negative values in the scalar-flux solution");
   return 0;
}

}  // extern "C"
