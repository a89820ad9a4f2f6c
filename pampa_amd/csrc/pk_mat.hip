// Sparse matrices of the pk_* layer: PETSc-AIJ row semantics on the host side
// of MatSetValues, CSR + sliced-ELL storage on the device, and SpMV kernels.
//
// Device formats
//   CSR        row_ptr/col/val, the exported pattern (bit-exact parity object).
//   SELL-h     slices of h = 32, 64 (default) or 128 consecutive rows, each
//              slice padded to its longest row and stored column-major inside
//              the slice, so row (h s + r) reads element k at
//              sell_off[s] + h k + r: the rows of a wavefront read contiguous
//              runs of col/val.  Padding entries carry column = row and value 0
//              and are not counted as work.  The slice height is picked per
//              size class by the autotuner of the eigen pool (pk_batch.hip).
#include <algorithm>
#include <numeric>

#include <rocprim/device/device_scan.hpp>

#include "pk_internal.hpp"
#include "pk_reduce.hpp"
#include "pk_sell.hpp"

namespace pk {

__global__ void to_f32_kernel(const double* a, float* b, long n) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) b[i] = (float)a[i];
}

// width of slice s = longest row among its h rows, as entries x h
__global__ void sell_width_kernel(long n, long ns, int h, const int* ptr, int* width, long* wl) {
   long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (s > ns) return;
   if (s == ns) {
      wl[ns] = 0;
      return;
   }
   int w = 0;
   const long r1 = min(n, (s + 1) * (long)h);
   for (long r = s * h; r < r1; r++) w = max(w, ptr[r + 1] - ptr[r]);
   width[s] = w;
   wl[s] = (long)w * h;
}

// one thread per row: its entries (or padding) at off[s] + h k + lane;
// with cols == false only the values are written (same pattern)
__global__ void sell_fill_kernel(long n, long nrows_pad, int h, const int* ptr, const int* col,
                                 const double* val, const long* off, const int* width, int* scol, double* sval,
                                 int* rlen, bool cols) {
   long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= nrows_pad) return;
   const long s = row / h;
   const int lane = (int)(row % h);
   const int w = width[s];
   const long a = row < n ? ptr[row] : 0;
   const int len = row < n ? ptr[row + 1] - ptr[row] : 0;
   for (int k = 0; k < w; k++) {
      const long p = off[s] + (long)k * h + lane;
      if (k < len) {
         if (cols) scol[p] = col[a + k];
         sval[p] = val[a + k];
      } else {
         if (cols) scol[p] = (int)(row < n ? row : 0);
         sval[p] = 0.0;
      }
   }
   if (cols && row < n) rlen[row] = len;
}

int build_sell(pk_mat_s* A, int h) {
   // device CSR -> SELL-h: slice widths, offsets by a scan, then the fill
   PK_REQUIRE(h == 32 || h == 64 || h == 128, "slice height must be 32, 64 or 128");
   Ctx& c = ctx();
   const long n = A->n;
   const long ns = (n + h - 1) / h;
   SellData& S = A->sell;
   S.h = h;
   S.nslices = ns;
   S.nnz = A->nnz;
   DBuf<long> wl;
   PK_REQUIRE(!S.off.alloc(ns + 1) && !S.width.alloc(ns ? ns : 1) && !S.rlen.alloc(n ? n : 1) &&
                  !wl.alloc(ns + 1),
              "device allocation failed");
   sell_width_kernel<<<grid_for(ns + 1), kBlock, 0, c.stream>>>(n, ns, h, A->ptr.p, S.width.p, wl.p);
   PK_HIP(hipGetLastError());
   size_t tmp_bytes = 0;
   PK_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, wl.p, S.off.p, 0L, (size_t)(ns + 1), rocprim::plus<long>(),
                                  c.stream));
   DBuf<char> tmp;
   PK_REQUIRE(!tmp.alloc(tmp_bytes ? tmp_bytes : 1), "device allocation failed");
   PK_HIP(rocprim::exclusive_scan(tmp.p, tmp_bytes, wl.p, S.off.p, 0L, (size_t)(ns + 1), rocprim::plus<long>(),
                                  c.stream));
   long tot = 0;
   PK_HIP(hipMemcpyAsync(&tot, S.off.p + ns, sizeof(long), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   S.padded = tot;
   if (h == 64) S.pad64 = tot;
   PK_REQUIRE(!S.col.alloc(tot ? tot : 1) && !S.val.alloc(tot ? tot : 1), "device allocation failed");
   if (n) {
      sell_fill_kernel<<<grid_for(ns * h), kBlock, 0, c.stream>>>(n, ns * h, h, A->ptr.p, A->col.p, A->val.p,
                                                                S.off.p, S.width.p, S.col.p, S.val.p, S.rlen.p,
                                                                true);
      PK_HIP(hipGetLastError());
   }
   PK_HIP(hipStreamSynchronize(c.stream));
   S.val32.free();
   return 0;
}

// new CSR values on the same pattern: refresh the SELL values (and drop the FP32 copy)
int refresh_sell_values(pk_mat_s* A) {
   Ctx& c = ctx();
   SellData& S = A->sell;
   if (A->n) {
      sell_fill_kernel<<<grid_for(S.nslices * S.h), kBlock, 0, c.stream>>>(
          A->n, S.nslices * S.h, S.h, A->ptr.p, A->col.p, A->val.p, S.off.p, S.width.p, nullptr, S.val.p,
          nullptr, false);
      PK_HIP(hipGetLastError());
   }
   if (S.val32.p) {
      long tot = S.padded ? S.padded : 1;
      to_f32_kernel<<<grid_for(tot), kBlock, 0, c.stream>>>(S.val.p, S.val32.p, S.padded);
      PK_HIP(hipGetLastError());
   }
   return 0;
}

// ---- SpMV kernels ------------------------------------------------------------

template <class T>
__global__ void __launch_bounds__(1024) sell_spmv_kernel(long n, int h, const long* __restrict__ off,
                                                         const int* __restrict__ width,
                                                         const int* __restrict__ col,
                                                         const T* __restrict__ val,
                                                         const T* __restrict__ x, T* __restrict__ y) {
   const long row = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (row >= n) return;
   const long s = row / h;
   y[row] = sell_row_dot<T, T>(col, val, x, off[s] + (row - s * h), width[s], (T)0, h);
}


// algorithmic bytes of a CSR SpMV: nnz (value + column) + row pointers + x, y
static double spmv_bytes(const pk_mat_s* A, double s) {
   return (double)A->sell.nnz * (s + 4.0) + (double)(A->n + 1) * 4.0 + 2.0 * A->n * s;
}

int spmv(pk_mat_s* A, const double* x, double* y) {
   if (A->n == 0) return 0;
   SellData& S = A->sell;
   PK_LAUNCH("spmv", spmv_bytes(A, 8.0), sell_spmv_kernel<double>, grid_for(A->n, S.spmv_block), S.spmv_block,
             ctx().stream, A->n, S.h, S.off.p, S.width.p, S.col.p, S.val.p, x, y);
   PK_HIP(hipGetLastError());
   return 0;
}

int spmv_f(pk_mat_s* A, const float* x, float* y) {
   if (A->n == 0) return 0;
   PK_CALL(A->ensure_fp32());
   SellData& S = A->sell;
   PK_LAUNCH("spmv_fp32", spmv_bytes(A, 4.0), sell_spmv_kernel<float>, grid_for(A->n, S.spmv_block),
             S.spmv_block, ctx().stream, A->n, S.h, S.off.p, S.width.p, S.col.p, S.val32.p, x, y);
   PK_HIP(hipGetLastError());
   return 0;
}

template <class T>
__global__ void fill_kernel(T* x, long n, T v) {
   const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) x[i] = v;
}

// average device time of the plain SpMV with the matrix's current layout and a
// workgroup of `block` rows (1 + reps launches on the library stream, x = 1)
int time_spmv(pk_mat_s* A, bool f32, int block, int reps, double* seconds) {
   Ctx& c = ctx();
   SellData& S = A->sell;
   const long n = A->n;
   PK_REQUIRE(n > 0 && reps > 0 && block >= 64 && block <= 1024 && block % 64 == 0, "wrong arguments");
   if (f32) PK_CALL(A->ensure_fp32());
   const size_t t = f32 ? sizeof(float) : sizeof(double);
   DBuf<char> x, y;
   PK_REQUIRE(!x.alloc(n * t) && !y.alloc(n * t), "device allocation failed");
   if (f32)
      fill_kernel<float><<<grid_for(n), kBlock, 0, c.stream>>>((float*)x.p, n, 1.0f);
   else
      fill_kernel<double><<<grid_for(n), kBlock, 0, c.stream>>>((double*)x.p, n, 1.0);
   EventPair ev;
   PK_HIP(ev.create());
   for (int r = -1; r < reps; r++) {
      if (r == 0) PK_HIP(hipEventRecord(ev.e0, c.stream));
      if (f32)
         sell_spmv_kernel<float><<<grid_for(n, block), block, 0, c.stream>>>(
             n, S.h, S.off.p, S.width.p, S.col.p, S.val32.p, (const float*)x.p, (float*)y.p);
      else
         sell_spmv_kernel<double><<<grid_for(n, block), block, 0, c.stream>>>(
             n, S.h, S.off.p, S.width.p, S.col.p, S.val.p, (const double*)x.p, (double*)y.p);
   }
   PK_HIP(hipEventRecord(ev.e1, c.stream));
   PK_HIP(hipGetLastError());
   float ms = 0.f;
   PK_HIP(ev.ms(&ms));
   *seconds = ms * 1e-3 / reps;
   return 0;
}

int finalize_device_csr(pk_mat_s* A) {
   PK_CALL(build_sell(A));
   A->assembled = true;
   return 0;
}

}  // namespace pk

int pk_mat_s::ensure_fp32() {
   if (sell.val32.p) return 0;
   long tot = sell.padded ? sell.padded : 1;
   if (sell.val32.alloc(tot)) PK_FAIL("device allocation failed");
   pk::to_f32_kernel<<<pk::grid_for(tot), pk::kBlock, 0, pk::ctx().stream>>>(sell.val.p, sell.val32.p,
                                                                           sell.padded);
   PK_HIP(hipGetLastError());
   return 0;
}

using namespace pk;

extern "C" {

int pk_mat_create(long n, int prealloc, pk_mat* M) {
   PK_CALL(ensure_init());
   PK_REQUIRE(n >= 0, "negative matrix size");
   auto* A = new pk_mat_s();
   A->n = n;
   A->prealloc = prealloc;
   A->stage.resize(n);
   A->stage_mode.resize(n);
   *M = A;
   return 0;
}

int pk_mat_set_values(pk_mat M, int nrows, const int* rows, int ncols, const int* cols,
                      const double* vals, int mode) {
   PK_REQUIRE(M, "null matrix");
   PK_REQUIRE(!M->assembled, "matrix already assembled");
   PK_REQUIRE(mode == PK_INSERT_VALUES || mode == PK_ADD_VALUES, "wrong insert mode");
   for (int a = 0; a < nrows; a++) {
      int r = rows[a];
      PK_REQUIRE(r >= 0 && r < M->n, "row index out of range");
      for (int b = 0; b < ncols; b++) {
         int c = cols[b];
         PK_REQUIRE(c >= 0 && c < M->n, "column index out of range");
         M->stage[r].push_back({c, vals[(long)a * ncols + b]});
         M->stage_mode[r].push_back(mode);
      }
   }
   return 0;
}

int pk_mat_assemble(pk_mat M) {
   PK_REQUIRE(M, "null matrix");
   if (M->assembled) return 0;
   long n = M->n;
   std::vector<int> ptr(n + 1, 0);
   std::vector<int> col;
   std::vector<double> val;
   for (long r = 0; r < n; r++) {
      auto& st = M->stage[r];
      auto& md = M->stage_mode[r];
      std::vector<int> idx(st.size());
      std::iota(idx.begin(), idx.end(), 0);
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return st[a].first < st[b].first; });
      for (size_t k = 0; k < idx.size(); k++) {
         const auto& e = st[idx[k]];
         if (!col.empty() && (long)col.size() > ptr[r] && col.back() == e.first) {
            if (md[idx[k]] == PK_ADD_VALUES)
               val.back() += e.second;
            else
               val.back() = e.second;
         } else {
            col.push_back(e.first);
            val.push_back(e.second);
         }
      }
      ptr[r + 1] = (int)col.size();
      std::vector<std::pair<int, double>>().swap(st);
      std::vector<int>().swap(md);
   }
   M->nnz = (long)col.size();
   PK_REQUIRE(!M->ptr.alloc(n + 1) && !M->col.alloc(M->nnz ? M->nnz : 1) &&
                  !M->val.alloc(M->nnz ? M->nnz : 1),
              "device allocation failed");
   PK_HIP(hipMemcpy(M->ptr.p, ptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
   if (M->nnz) {
      PK_HIP(hipMemcpy(M->col.p, col.data(), sizeof(int) * M->nnz, hipMemcpyHostToDevice));
      PK_HIP(hipMemcpy(M->val.p, val.data(), sizeof(double) * M->nnz, hipMemcpyHostToDevice));
   }
   return finalize_device_csr(M);
}

int pk_mat_destroy(pk_mat* M) {
   if (!M || !*M) return 0;
   (void)hipStreamSynchronize(ctx().stream);
   pk_mat_s* A = *M;
   A->ptr.free();
   A->col.free();
   A->val.free();
   A->sell.off.free();
   A->sell.width.free();
   A->sell.col.free();
   A->sell.rlen.free();
   A->sell.val.free();
   A->sell.val32.free();
   delete A;
   *M = nullptr;
   return 0;
}

int pk_mat_get_size(pk_mat M, long* n, long* nnz) {
   PK_REQUIRE(M, "null matrix");
   if (n) *n = M->n;
   if (nnz) *nnz = M->nnz;
   return 0;
}

int pk_mat_get_csr(pk_mat M, int* row_ptr, int* col_idx, double* val) {
   PK_REQUIRE(M && M->assembled, "matrix not assembled");
   PK_HIP(hipStreamSynchronize(ctx().stream));
   if (row_ptr) PK_HIP(hipMemcpy(row_ptr, M->ptr.p, sizeof(int) * (M->n + 1), hipMemcpyDeviceToHost));
   if (M->nnz) {
      if (col_idx) PK_HIP(hipMemcpy(col_idx, M->col.p, sizeof(int) * M->nnz, hipMemcpyDeviceToHost));
      if (val) PK_HIP(hipMemcpy(val, M->val.p, sizeof(double) * M->nnz, hipMemcpyDeviceToHost));
   }
   return 0;
}

int pk_mat_mult(pk_mat M, pk_vec x, pk_vec y) {
   PK_REQUIRE(M && M->assembled, "matrix not assembled");
   PK_REQUIRE(x && y && x->n == M->n && y->n == M->n, "vector size mismatch");
   return spmv(M, x->d.p, y->d.p);
}

int pk_mat_set_block_size(pk_mat M, int bs) {
   PK_REQUIRE(M && bs >= 1 && M->n % bs == 0, "wrong block size");
   M->bs = bs;
   return 0;
}

int pk_mat_time_mult(pk_mat M, pk_vec x, pk_vec y, int reps, int fp32, double* avg_seconds) {
   PK_REQUIRE(M && M->assembled && x && y && reps > 0, "wrong arguments");
   Ctx& c = ctx();
   DBuf<float> xb, yb;
   float* xf = nullptr;
   float* yf = nullptr;
   if (fp32) {
      PK_CALL(M->ensure_fp32());
      PK_REQUIRE(!xb.alloc(M->n) && !yb.alloc(M->n), "device allocation failed");
      xf = xb.p;
      yf = yb.p;
      to_f32_kernel<<<grid_for(M->n), kBlock, 0, c.stream>>>(x->d.p, xf, M->n);
   }
   EventPair ev;
   PK_HIP(ev.create());
   // warm-up
   if (fp32)
      PK_CALL(spmv_f(M, xf, yf));
   else
      PK_CALL(spmv(M, x->d.p, y->d.p));
   PK_HIP(hipEventRecord(ev.e0, c.stream));
   for (int r = 0; r < reps; r++) {
      if (fp32)
         PK_CALL(spmv_f(M, xf, yf));
      else
         PK_CALL(spmv(M, x->d.p, y->d.p));
   }
   PK_HIP(hipEventRecord(ev.e1, c.stream));
   float ms = 0.f;
   PK_HIP(ev.ms(&ms));
   *avg_seconds = (double)ms * 1e-3 / reps;
   return 0;
}

}  // extern "C"
