// Internal declarations of the pk_* numerics library (HIP, gfx950).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pk.h"

namespace pk {

// ---- errors ---------------------------------------------------------------
// Every failure records "file:line:function(): error: message." (the reference's
// convention, utils.hxx:41-47) and the entry point returns a positive code.
void set_error(const char* file, int line, const char* func, const std::string& msg);

#define PK_FAIL(msg)                                                  \
   do {                                                              \
      ::pk::set_error(__FILE__, __LINE__, __func__, (msg));          \
      return 1;                                                      \
   } while (0)

#define PK_REQUIRE(cond, msg) \
   do {                       \
      if (!(cond)) PK_FAIL(msg); \
   } while (0)

#define PK_HIP(call)                                                                   \
   do {                                                                                \
      hipError_t e_ = (call);                                                          \
      if (e_ != hipSuccess) PK_FAIL(std::string(#call) + ": " + hipGetErrorString(e_)); \
   } while (0)

#define PK_CALL(call)      \
   do {                    \
      if ((call) != 0) return 1; \
   } while (0)

// ---- device context --------------------------------------------------------
constexpr int kBlock = 256;          // threads per workgroup for 1-D kernels
constexpr int kMaxRedBlocks = 1024;  // workgroups of a reduction kernel
constexpr int kMaxRedVals = 80;      // values reduced together (GMRES restart + 2)

struct Ctx {
   int device = -1;
   hipStream_t stream = nullptr;
   double* partials = nullptr;   // [kMaxRedBlocks * kMaxRedVals]
   unsigned* ticket = nullptr;   // last-block-done counters
   double* scalars = nullptr;    // small device-resident scalars [256]
   double* host_scalars = nullptr;  // pinned mirror
   int num_cus = 0;
   size_t partials_cap = 0;      // doubles available in `partials`
};

Ctx& ctx();
int ensure_init();

// Kernel profiler: with profiling on, instrumented launches carry a pair of
// HIP events on the library stream; durations are resolved lazily.  `bytes`
// is the launch's algorithmic traffic.
bool profiling();
// With profiling on: a pending record (name, algorithmic bytes) whose start and
// stop events the launch itself records (hipExtLaunchKernelGGL: the events
// take the kernel's own start and end, not those of separate markers).
bool prof_events(const char* name, double bytes, hipEvent_t* e0, hipEvent_t* e1);
// Add measured launches directly (profiling on only).
void prof_add(const char* name, long launches, double seconds, double bytes);

// kernel<<<grid, block, 0, stream>>>(args...), timed when profiling is on
#define PK_LAUNCH(name, bytes, kernel, grid, block, stream, ...)                                        \
   do {                                                                                                 \
      hipEvent_t e0_ = nullptr, e1_ = nullptr;                                                          \
      if (::pk::prof_events((name), (bytes), &e0_, &e1_))                                               \
         hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, (stream), e0_, e1_, 0, __VA_ARGS__); \
      else                                                                                              \
         kernel<<<(grid), (block), 0, (stream)>>>(__VA_ARGS__);                                         \
   } while (0)

// the same with `shmem` bytes of dynamic local memory
#define PK_LAUNCH_SHM(name, bytes, kernel, grid, block, shmem, stream, ...)                                   \
   do {                                                                                                       \
      hipEvent_t e0_ = nullptr, e1_ = nullptr;                                                                \
      if (::pk::prof_events((name), (bytes), &e0_, &e1_))                                                     \
         hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shmem), (stream), e0_, e1_, 0, __VA_ARGS__); \
      else                                                                                                    \
         kernel<<<(grid), (block), (shmem), (stream)>>>(__VA_ARGS__);                                         \
   } while (0)

// Grow the reduction scratch to at least `count` doubles (setup time only).
int reserve_partials(size_t count);

// Static scratch of the allocation-free entry points: kScratchDoubles doubles
// (16.9 KB) per device, a file-scope __device__ array whose device address
// pk_initialize resolves per device.  Layout: reduction partials of up to
// kMaxRedBlocks workgroups x 2 values, then 8 scalars, then 16 ticket
// counters (unsigned, in the last doubles).  static_scratch() is the current
// device's copy (nullptr before pk_initialize).
constexpr int kScratchPartials = kMaxRedBlocks * 2;
constexpr int kScratchScalars = kScratchPartials;       // 8 doubles
constexpr int kScratchTickets = kScratchPartials + 8;   // 16 unsigned (8 doubles)
constexpr int kScratchDoubles = kScratchPartials + 64;
double* static_scratch();

inline int grid_for(long n, int block = kBlock, int cap = 1 << 30) {
   long g = (n + block - 1) / block;
   if (g < 1) g = 1;
   if (g > cap) g = cap;
   return (int)g;
}

// Device buffer owned by a handle (move-only; freed on destruction).  A buffer
// that borrows another's pointer must clear `p` before it is destroyed.
template <class T>
struct DBuf {
   T* p = nullptr;
   size_t n = 0;
   DBuf() = default;
   DBuf(const DBuf&) = delete;
   DBuf& operator=(const DBuf&) = delete;
   DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
      o.p = nullptr;
      o.n = 0;
   }
   DBuf& operator=(DBuf&& o) noexcept {
      if (this != &o) {
         free();
         p = o.p;
         n = o.n;
         o.p = nullptr;
         o.n = 0;
      }
      return *this;
   }
   ~DBuf() { free(); }
   int alloc(size_t count) {
      free();
      n = count;
      if (count == 0) return 0;
      if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
         p = nullptr;
         n = 0;
         return 1;
      }
      return 0;
   }
   void free() {
      if (p) (void)hipFree(p);
      p = nullptr;
      n = 0;
   }
};

// Two timing events on one stream, destroyed with the scope (error returns
// included).  ms() waits for the second one.
struct EventPair {
   hipEvent_t e0 = nullptr, e1 = nullptr;
   EventPair() = default;
   EventPair(const EventPair&) = delete;
   EventPair& operator=(const EventPair&) = delete;
   ~EventPair() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
   }
   hipError_t create() {
      const hipError_t e = hipEventCreate(&e0);
      return e != hipSuccess ? e : hipEventCreate(&e1);
   }
   hipError_t ms(float* out) {
      const hipError_t e = hipEventSynchronize(e1);
      return e != hipSuccess ? e : hipEventElapsedTime(out, e0, e1);
   }
};

// ---- deterministic reductions ----------------------------------------------
// Sum over n of f(i) for `nv` values at once; result written to out[0..nv-1]
// (device memory).  Fixed work split and fixed summation order, so the result
// does not depend on timing or on what else runs on the device.
int reduce_dot(const double* x, const double* y, long n, double* out_dev);
int reduce_dot_f(const float* x, const float* y, long n, double* out_dev);

}  // namespace pk

// Sliced-ELL (slice height 64 = one wavefront) device copy of a CSR matrix.
struct SellData {
   int h = 64;          // slice height (32, 64 or 128 rows)
   int spmv_block = 256;  // workgroup size of the plain SpMV kernel (64 .. 1024)
   long nslices = 0, padded = 0;
   long pad64 = 0;  // padded size of the 64-row layout: no layout in use is larger (pool capacities)
   long nnz = 0;  // unpadded nonzeros (the algorithmic count)
   pk::DBuf<long> off;
   pk::DBuf<int> width, col;
   pk::DBuf<int> rlen;  // unpadded row lengths (padding sits at k >= rlen[row])
   pk::DBuf<double> val;
   pk::DBuf<float> val32;
};

// Handle layouts shared across translation units.
struct pk_vec_s {
   long n = 0;
   pk::DBuf<double> d;
};

struct pk_mat_s {
   long n = 0;
   int prealloc = 0;
   int bs = 1;  // cell block size (energy groups)
   // host staging of MatSetValues rows: (col, val, seq) triples per row
   std::vector<std::vector<std::pair<int, double>>> stage;
   std::vector<std::vector<int>> stage_mode;
   bool assembled = false;
   long nnz = 0;
   pk::DBuf<int> ptr, col;
   pk::DBuf<double> val;
   SellData sell;  // SpMV copy; FP32 values built on demand for mixed precision
   // boundary table of a steady diffusion assembly (kept for value refills)
   pk::DBuf<int> asm_bct;
   pk::DBuf<double> asm_bcv;
   pk::DBuf<int> asm_err;
   int ensure_fp32();
};

struct pk_mesh_s {
   int C = 0;
   int F = 0;  // total faces
   int fmax = 0;
   pk::DBuf<int> fptr, mat, fnb;
   pk::DBuf<double> vol, cen, farea, fcen, fnorm;
   // host copies (schedules and host-side setup)
   std::vector<int> h_fptr, h_mat, h_fnb;
   std::vector<double> h_vol, h_cen, h_farea, h_fcen, h_fnorm;
};

// Cross-section tables in sets: material m owns nset[m] sets from set0[m],
// tabulated at increasing temperatures tref (one set: constant data).  Tables
// are [sets][G] and sigs [sets][G from][G to]; T (device, one value per cell,
// may be null = 0) is where temperature-dependent materials are evaluated.
struct pk_xs_s {
   int nm = 0, G = 0, ns = 0;
   bool feedback = false;  // some material has more than one set
   pk::DBuf<double> sigt, sigs, nusf, kapf, chi, D, tref;
   pk::DBuf<int> set0, nset;
   const double* T = nullptr;
   std::vector<double> h_sigt, h_sigs, h_nusf, h_kapf, h_chi, h_D;
};

// SpMV entry points used across files.
namespace pk {
int spmv(pk_mat_s* A, const double* x, double* y);
int spmv_f(pk_mat_s* A, const float* x, float* y);
int build_sell(pk_mat_s* A, int h = 64);
int time_spmv(pk_mat_s* A, bool f32, int block, int reps, double* seconds);
int refresh_sell_values(pk_mat_s* A);
int finalize_device_csr(pk_mat_s* A);
}  // namespace pk
