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
// whose upwind neighbours all sit in earlier levels).  Global step s sweeps level
// s of every direction at once, so one launch covers sum_m |level_s(m)| cells x
// all groups; a thread owns one (direction, cell) pair and loops over groups,
// reading the cell's faces once.  The device layout of psi is [m][cell][group]
// (the G values of an upwind neighbour are contiguous); the reference layout
// [cell][group][m] is produced only at the boundary of the API.
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

template <class T>
__global__ void __launch_bounds__(256) sn_sweep_step_kernel(const int* __restrict__ items, long nitems,
                                                            int C, int G, SnGeom geo,
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
         const T* up = psi + ((long)m * C + i2) * G;
         for (int g = 0; g < G; g++) inflow[g] += c * up[g];
      }
   }
   const int mt = geo.mat[i];
   const T V = (T)geo.vol[i];
   T* out = psi + ((long)m * C + i) * G;
   const T* qi = q + (long)i * G;
   for (int g = 0; g < G; g++) out[g] = (qi[g] - inflow[g]) / (sigt[mt * G + g] * V + lhs0);
}

// Per-item records: what a sweep step needs of (direction m, cell i), taken
// from the geometry once per precision in the step kernel's own arithmetic
// (same operations in the same order, so the same bits): the flat index m C + i,
// the outgoing part of the left-hand side lhs0 = sum_out coef w A + sum_vac,out w A,
// and up to kMaxIn incoming interior faces as (flat upwind index, coef w A) in
// face order.  Stored structure-of-arrays in step order, so a step reads its
// records with coalesced loads and has one dependent gather (the upwind values)
// left, instead of the chain cell -> faces -> normals -> neighbours -> values.
constexpr int kMaxIn = 4;

template <class T>
struct SnRec {
   long n;
   int* code;
   int* nin;
   T* lhs;
   int* up;  // [kMaxIn][n]
   T* cw;    // [kMaxIn][n]
};

template <class T>
__global__ void sn_rec_kernel(const int* __restrict__ items, long n, int C, SnGeom geo, SnRec<T> R, int* overflow) {
#pragma clang fp contract(off)
   long it = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (it >= n) return;
   const int code = items[it];
   const int m = code / C, i = code - m * C;
   const double o0 = geo.dirs[3 * m], o1 = geo.dirs[3 * m + 1], o2 = geo.dirs[3 * m + 2];
   T lhs0 = 0;
   int k = 0;
   for (int f = geo.fptr[i]; f < geo.fptr[i + 1]; f++) {
      const double* nn = geo.fnorm + 3L * f;
      T w = (T)(o0 * nn[0] + o1 * nn[1] + o2 * nn[2]);
      const int i2 = geo.fnb[f];
      const T A = (T)geo.farea[f];
      if (i2 < 0) {
         if (w > (T)0) lhs0 += w * A;
      } else if (w > (T)0) {
         lhs0 += (T)geo.coef[f] * w * A;
      } else {
         if (k < kMaxIn) {
            R.up[(long)k * n + it] = m * C + i2;
            R.cw[(long)k * n + it] = (T)geo.coef[f] * w * A;
         } else {
            *overflow = 1;
         }
         k++;
      }
   }
   R.code[it] = code;
   R.nin[it] = k < kMaxIn ? k : kMaxIn;
   R.lhs[it] = lhs0;
}

// one step from the records: items a .. a+count-1
template <class T>
__global__ void __launch_bounds__(256) sn_sweep_rec_kernel(long a, long count, int C, int G, SnRec<T> R,
                                                           const int* __restrict__ mat,
                                                           const double* __restrict__ vol,
                                                           const T* __restrict__ sigt, const T* __restrict__ q,
                                                           T* psi) {
#pragma clang fp contract(off)
   const long t = blockIdx.x * 256L + threadIdx.x;
   if (t >= count) return;
   const long it = a + t;
   const int code = R.code[it];
   const int nin = R.nin[it];
   const T lhs0 = R.lhs[it];
   const int i = code % C;
   T inflow[kMaxSnGroups];
   for (int g = 0; g < G; g++) inflow[g] = 0;
   for (int k = 0; k < nin; k++) {
      const T c = R.cw[(long)k * R.n + it];
      const T* up = psi + (long)R.up[(long)k * R.n + it] * G;
      for (int g = 0; g < G; g++) inflow[g] += c * up[g];
   }
   const int mt = mat[i];
   const T V = (T)vol[i];
   T* out = psi + (long)code * G;
   const T* qi = q + (long)i * G;
   for (int g = 0; g < G; g++) out[g] = (qi[g] - inflow[g]) / (sigt[mt * G + g] * V + lhs0);
}

// phiw[i][g] = sum_m w_m psi[m][i][g] (FP64 accumulation)
template <class T>
__global__ void sn_moment_kernel(const T* __restrict__ psi, const double* __restrict__ wts, int M,
                                 long CG, double* phiw) {
   long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (l >= CG) return;
   double s = 0.0;
   for (int m = 0; m < M; m++) s += wts[m] * (double)psi[(long)m * CG + l];
   phiw[l] = s;
}

// q = V (sum_g2 Sigma_s(g2->g) phiw(g2) + chi_g/k sum_g2 nuSf(g2) phiw(g2))
template <class T>
__global__ void sn_source_kernel(int C, int G, const double* vol, const int* mat, const double* sigs,
                                 const double* nusf, const double* chi, const double* phiw,
                                 const double* kk, T* q) {
#pragma clang fp contract(off)
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= C) return;
   const int mt = mat[i];
   const double inv_k = 1.0 / kk[0];
   double fis = 0.0;
   for (int g2 = 0; g2 < G; g2++) fis += nusf[mt * G + g2] * phiw[(long)i * G + g2];
   for (int g = 0; g < G; g++) {
      double s = 0.0;
      for (int g2 = 0; g2 < G; g2++) s += sigs[(mt * G + g2) * G + g] * phiw[(long)i * G + g2];
      q[(long)i * G + g] = (T)(vol[i] * (s + chi[mt * G + g] * inv_k * fis));
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
template <class T>
__global__ void sn_to_reference_kernel(const T* psi, int C, int G, int M, double scale, double* out) {
   long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   long n = (long)C * G * M;
   if (l >= n) return;
   int m = (int)(l % M);
   long ig = l / M;  // i*G + g
   out[l] = scale * (double)psi[(long)m * C * G + ig];
}

template <class T>
__global__ void sn_fill_kernel(T* x, long n, T v) {
   long l = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (l < n) x[l] = v;
}

}  // namespace pk

using namespace pk;

struct pk_sn_s {
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
   // per-item records (one set per precision, built on first use) and the
   // captured step sequences; geometry = true sweeps from the geometry instead
   bool geometry = false;
   DBuf<int> rec_code, rec_nin, rec_up, rec_err;
   DBuf<double> rec_lhs64, rec_cw64;
   DBuf<float> rec_lhs32, rec_cw32;
   hipGraphExec_t graph32 = nullptr, graph64 = nullptr;
   const void* graph_ptrs[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
   ~pk_sn_s() {
      if (graph32) (void)hipGraphExecDestroy(graph32);
      if (graph64) (void)hipGraphExecDestroy(graph64);
   }
   DBuf<double> dirs, wts, coef;
   DBuf<double> psi64, q64, sigt64;
   DBuf<float> psi32, q32, sigt32;
   DBuf<double> phiw, old, st;
   // value accesses of one sweep: per (i,g,m) one psi write plus one upwind psi
   // read per incoming interior face, per (i,g) one source read and one moment write
   double value_units = 0.0;
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

template <class T>
static SnRec<T> recs(pk_sn_s* S);
template <>
SnRec<double> recs<double>(pk_sn_s* S) {
   const long n = S->step_ptr[S->nsteps];
   return SnRec<double>{n, S->rec_code.p, S->rec_nin.p, S->rec_lhs64.p, S->rec_up.p, S->rec_cw64.p};
}
template <>
SnRec<float> recs<float>(pk_sn_s* S) {
   const long n = S->step_ptr[S->nsteps];
   return SnRec<float>{n, S->rec_code.p, S->rec_nin.p, S->rec_lhs32.p, S->rec_up.p, S->rec_cw32.p};
}

// records of precision T, built once
template <class T>
static int sn_records(pk_sn_s* S) {
   Ctx& c = ctx();
   const long n = S->step_ptr[S->nsteps];
   const bool f32 = sizeof(T) == sizeof(float);
   if (f32 ? S->rec_lhs32.p != nullptr : S->rec_lhs64.p != nullptr) return 0;
   if (!S->rec_code.p)
      PK_REQUIRE(!S->rec_code.alloc(n) && !S->rec_nin.alloc(n) && !S->rec_up.alloc((size_t)kMaxIn * n) &&
                     !S->rec_err.alloc(1),
                 "device allocation failed");
   if (f32)
      PK_REQUIRE(!S->rec_lhs32.alloc(n) && !S->rec_cw32.alloc((size_t)kMaxIn * n), "device allocation failed");
   else
      PK_REQUIRE(!S->rec_lhs64.alloc(n) && !S->rec_cw64.alloc((size_t)kMaxIn * n), "device allocation failed");
   PK_HIP(hipMemsetAsync(S->rec_err.p, 0, sizeof(int), c.stream));
   sn_rec_kernel<T><<<grid_for(n), kBlock, 0, c.stream>>>(S->items.p, n, S->C, geom(S), recs<T>(S), S->rec_err.p);
   PK_HIP(hipGetLastError());
   int h = 0;
   PK_HIP(hipMemcpyAsync(&h, S->rec_err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   PK_REQUIRE(h == 0, "more incoming faces per cell than the sweep records hold");
   return 0;
}

template <class T>
static void sn_launch_steps(pk_sn_s* S, const T* sigt, const T* q, T* psi, bool prof) {
   Ctx& c = ctx();
   const SnRec<T> R = recs<T>(S);
   const double per_item = prof ? sweep_bytes(S, sizeof(T)) / (double)S->step_ptr[S->nsteps] : 0.0;
   for (int s = 0; s < S->nsteps; s++) {
      const long a = S->step_ptr[s], n = S->step_ptr[s + 1] - a;
      PK_LAUNCH("sn_sweep_step", per_item * n, sn_sweep_rec_kernel<T>, grid_for(n, 256), 256, c.stream, a, n,
                S->C, S->G, R, S->mesh->mat.p, S->mesh->vol.p, sigt, q, psi);
   }
}

// One sweep psi = L^-1 q.  The step launches are captured once per precision
// as a HIP graph (the sweep's buffers are fixed per handle and precision) and
// relaunched; with the profiler on they are launched one by one.
template <class T>
static int sn_sweep(pk_sn_s* S, const T* sigt, const T* q, T* psi) {
   Ctx& c = ctx();
   if (S->geometry) {
      SnGeom g = geom(S);
      const double per_item = profiling() ? sweep_bytes(S, sizeof(T)) / (double)S->step_ptr[S->nsteps] : 0.0;
      for (int s = 0; s < S->nsteps; s++) {
         long a = S->step_ptr[s], n = S->step_ptr[s + 1] - a;
         PK_LAUNCH("sn_sweep_step", per_item * n, sn_sweep_step_kernel<T>, grid_for(n, 256), 256, c.stream,
                   S->items.p + a, n, S->C, S->G, g, sigt, q, psi);
      }
      PK_HIP(hipGetLastError());
      return 0;
   }
   PK_CALL(sn_records<T>(S));
   if (profiling()) {
      sn_launch_steps<T>(S, sigt, q, psi, true);
      PK_HIP(hipGetLastError());
      return 0;
   }
   const int pi = sizeof(T) == sizeof(float) ? 0 : 1;
   hipGraphExec_t& G = pi == 0 ? S->graph32 : S->graph64;
   const void** gp = S->graph_ptrs[pi];
   if (G && (gp[0] != sigt || gp[1] != q || gp[2] != psi)) {
      (void)hipGraphExecDestroy(G);
      G = nullptr;
   }
   if (!G) {
      gp[0] = sigt;
      gp[1] = q;
      gp[2] = psi;
      hipGraph_t g = nullptr;
      PK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      sn_launch_steps<T>(S, sigt, q, psi, false);
      hipError_t e = hipStreamEndCapture(c.stream, &g);
      PK_HIP(e);
      hipError_t ie = hipGraphInstantiate(&G, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      PK_HIP(ie);
   }
   PK_HIP(hipGraphLaunch(G, c.stream));
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
   const char* ge = getenv("PK_SWEEP_GEOMETRY");
   S->geometry = ge && *ge && *ge != '0';
   const int G = S->G;
   S->value_units = (double)((long)C * G * M + nin_total * G) + 2.0 * C * G;
   int e = 0;
   e |= S->items.alloc(items.size());
   e |= S->dirs.alloc(3L * M) | S->wts.alloc(M) | S->coef.alloc(mesh->F ? mesh->F : 1);
   e |= S->phiw.alloc((size_t)C * G) | S->old.alloc((size_t)C * G) | S->st.alloc(8);
   if (e) {
      delete S;
      PK_FAIL("device allocation failed");
   }
   PK_HIP(hipMemcpy(S->items.p, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice));

   PK_HIP(hipMemcpy(S->dirs.p, dirs, sizeof(double) * 3 * M, hipMemcpyHostToDevice));
   PK_HIP(hipMemcpy(S->wts.p, wts, sizeof(double) * M, hipMemcpyHostToDevice));
   if (mesh->F) PK_HIP(hipMemcpy(S->coef.p, coef.data(), sizeof(double) * mesh->F, hipMemcpyHostToDevice));
   *out = S;
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

static int sn_prepare(pk_sn_s* S) {
   const long CGM = (long)S->C * S->G * S->M, CG = (long)S->C * S->G;
   const int nm = S->xs->nm * S->G;
   if (S->precision == PK_PREC_MIXED) {
      if (!S->psi32.p) {
         PK_REQUIRE(!S->psi32.alloc(CGM) && !S->q32.alloc(CG) && !S->sigt32.alloc(nm),
                    "device allocation failed");
         std::vector<float> st(S->xs->h_sigt.begin(), S->xs->h_sigt.end());
         PK_HIP(hipMemcpy(S->sigt32.p, st.data(), sizeof(float) * nm, hipMemcpyHostToDevice));
      }
   } else {
      if (!S->psi64.p) {
         PK_REQUIRE(!S->psi64.alloc(CGM) && !S->q64.alloc(CG) && !S->sigt64.alloc(nm),
                    "device allocation failed");
         PK_HIP(hipMemcpy(S->sigt64.p, S->xs->h_sigt.data(), sizeof(double) * nm, hipMemcpyHostToDevice));
      }
   }
   return 0;
}

int pk_sn_sweep_host(pk_sn S, const double* q, double* psi, double* phiw) {
   PK_REQUIRE(S && q, "null argument");
   PK_CALL(sn_prepare(S));
   Ctx& c = ctx();
   const long CG = (long)S->C * S->G, CGM = CG * S->M;
   DBuf<double> out;
   PK_REQUIRE(!out.alloc(CGM), "device allocation failed");
   if (S->precision == PK_PREC_MIXED) {
      std::vector<float> qf(q, q + CG);
      PK_HIP(hipMemcpy(S->q32.p, qf.data(), sizeof(float) * CG, hipMemcpyHostToDevice));
      PK_CALL(sn_sweep<float>(S, S->sigt32.p, S->q32.p, S->psi32.p));
      sn_moment_kernel<float><<<grid_for(CG), kBlock, 0, c.stream>>>(S->psi32.p, S->wts.p, S->M, CG, S->phiw.p);
      sn_to_reference_kernel<float><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi32.p, S->C, S->G, S->M, 1.0, out.p);
   } else {
      PK_HIP(hipMemcpy(S->q64.p, q, sizeof(double) * CG, hipMemcpyHostToDevice));
      PK_CALL(sn_sweep<double>(S, S->sigt64.p, S->q64.p, S->psi64.p));
      sn_moment_kernel<double><<<grid_for(CG), kBlock, 0, c.stream>>>(S->psi64.p, S->wts.p, S->M, CG, S->phiw.p);
      sn_to_reference_kernel<double><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi64.p, S->C, S->G, S->M, 1.0, out.p);
   }
   PK_HIP(hipGetLastError());
   PK_HIP(hipStreamSynchronize(c.stream));
   if (psi) PK_HIP(hipMemcpy(psi, out.p, sizeof(double) * CGM, hipMemcpyDeviceToHost));
   if (phiw) PK_HIP(hipMemcpy(phiw, S->phiw.p, sizeof(double) * CG, hipMemcpyDeviceToHost));
   return 0;
}

int pk_sn_solve(pk_sn S, double* keff, pk_vec psi, pk_vec phiw_out) {
   PK_REQUIRE(S && keff, "null argument");
   PK_CALL(sn_prepare(S));
   Ctx& c = ctx();
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
   for (it = 0; it < S->max_outer; it++) {
      PK_HIP(hipMemcpyAsync(S->old.p, S->phiw.p, sizeof(double) * CG, hipMemcpyDeviceToDevice, c.stream));
      if (S->precision == PK_PREC_MIXED) {
         sn_source_kernel<float><<<grid_for(C), kBlock, 0, c.stream>>>(C, G, mesh->vol.p, mesh->mat.p, xs->sigs.p,
                                                                    xs->nusf.p, xs->chi.p, S->phiw.p, st, S->q32.p);
         PK_CALL(sn_sweep<float>(S, S->sigt32.p, S->q32.p, S->psi32.p));
         sn_moment_kernel<float><<<grid_for(CG), kBlock, 0, c.stream>>>(S->psi32.p, S->wts.p, M, CG, S->phiw.p);
      } else {
         sn_source_kernel<double><<<grid_for(C), kBlock, 0, c.stream>>>(C, G, mesh->vol.p, mesh->mat.p, xs->sigs.p,
                                                                     xs->nusf.p, xs->chi.p, S->phiw.p, st, S->q64.p);
         PK_CALL(sn_sweep<double>(S, S->sigt64.p, S->q64.p, S->psi64.p));
         sn_moment_kernel<double><<<grid_for(CG), kBlock, 0, c.stream>>>(S->psi64.p, S->wts.p, M, CG, S->phiw.p);
      }
      S->sweeps++;
      sn_prod_kernel<<<gb, kBlock, 0, c.stream>>>(C, G, mesh->vol.p, mesh->mat.p, xs->nusf.p, S->phiw.p,
                                                 c.partials, c.ticket, st + 2);
      sn_update_kernel<<<gb, kBlock, 0, c.stream>>>(CG, S->phiw.p, S->old.p, st, c.partials, c.ticket, st + 3);
      PK_HIP(hipGetLastError());
      PK_HIP(hipMemcpyAsync(c.host_scalars, st, sizeof(double) * 5, hipMemcpyDeviceToHost, c.stream));
      PK_HIP(hipStreamSynchronize(c.stream));
         double prod_old = c.host_scalars[1], prod_new = c.host_scalars[2];
      PK_REQUIRE(prod_new > 0.0 && prod_old > 0.0, "zero fission source in the S_N eigen solve");
      double knew = k * prod_new / prod_old;
      last_scale = prod_old / prod_new;
      double dphi = sqrt(c.host_scalars[3] / c.host_scalars[4]);
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
   }
   PK_HIP(hipStreamSynchronize(c.stream));
   S->keff = k;
   S->outer = it;
   S->has_state = true;
   *keff = k;
   // psi of the last sweep, rescaled like phiw (reference layout)
   if (psi) {
      if (S->precision == PK_PREC_MIXED)
         sn_to_reference_kernel<float><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi32.p, C, G, M, last_scale, psi->d.p);
      else
         sn_to_reference_kernel<double><<<grid_for(CGM), kBlock, 0, c.stream>>>(S->psi64.p, C, G, M, last_scale, psi->d.p);
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

int pk_sn_time_sweep(pk_sn S, int reps, double* avg_seconds, double* bytes_per_sweep) {
   PK_REQUIRE(S && reps > 0, "wrong arguments");
   PK_CALL(sn_prepare(S));
   Ctx& c = ctx();
   const long CG = (long)S->C * S->G;
   const bool f32 = S->precision == PK_PREC_MIXED;
   if (f32)
      sn_fill_kernel<float><<<grid_for(CG), kBlock, 0, c.stream>>>(S->q32.p, CG, 1.0f);
   else
      sn_fill_kernel<double><<<grid_for(CG), kBlock, 0, c.stream>>>(S->q64.p, CG, 1.0);
   hipEvent_t e0, e1;
   PK_HIP(hipEventCreate(&e0));
   PK_HIP(hipEventCreate(&e1));
   for (int r = -1; r < reps; r++) {
      if (r == 0) PK_HIP(hipEventRecord(e0, c.stream));
      if (f32)
         PK_CALL(sn_sweep<float>(S, S->sigt32.p, S->q32.p, S->psi32.p));
      else
         PK_CALL(sn_sweep<double>(S, S->sigt64.p, S->q64.p, S->psi64.p));
   }
   PK_HIP(hipEventRecord(e1, c.stream));
   PK_HIP(hipEventSynchronize(e1));
   float ms = 0.f;
   PK_HIP(hipEventElapsedTime(&ms, e0, e1));
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   if (avg_seconds) *avg_seconds = ms * 1e-3 / reps;
   if (bytes_per_sweep) *bytes_per_sweep = sweep_bytes(S, f32 ? 4.0 : 8.0);
   return 0;
}

int pk_sn_destroy(pk_sn* S) {
   if (!S || !*S) return 0;
   (void)hipStreamSynchronize(ctx().stream);
   delete *S;
   *S = nullptr;
   return 0;
}

}  // extern "C"
