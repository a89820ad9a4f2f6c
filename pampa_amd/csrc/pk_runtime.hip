// Runtime, errors, vectors and deterministic reductions of the pk_* layer.
#include <cstring>
#include <mutex>

#include "pk_internal.hpp"
#include "pk_reduce.hpp"

namespace pk {

static thread_local std::string g_last_error;
static Ctx g_ctx;

// ---- kernel profiler ------------------------------------------------------------
struct ProfRec {
   std::string name;
   double bytes;
   hipEvent_t e0, e1;
};
struct ProfAcc {
   long launches = 0;
   double seconds = 0.0, bytes = 0.0;
};
static bool g_prof = false;
static std::vector<ProfRec> g_pending;
static std::vector<hipEvent_t> g_event_pool;
static std::vector<std::pair<std::string, ProfAcc>> g_prof_acc;

static hipEvent_t take_event() {
   if (!g_event_pool.empty()) {
      hipEvent_t e = g_event_pool.back();
      g_event_pool.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   (void)hipEventCreate(&e);
   return e;
}

bool profiling() { return g_prof; }

static void prof_resolve();

bool prof_events(const char* name, double bytes, hipEvent_t* e0, hipEvent_t* e1) {
   if (!g_prof) return false;
   // resolve in batches so that long profiled runs reuse their events
   if (g_pending.size() >= 8192) prof_resolve();
   ProfRec r{name, bytes, take_event(), take_event()};
   g_pending.push_back(r);
   *e0 = r.e0;
   *e1 = r.e1;
   return true;
}

void prof_add(const char* name, long launches, double seconds, double bytes) {
   if (!g_prof) return;
   ProfAcc* acc = nullptr;
   for (auto& p : g_prof_acc)
      if (p.first == name) acc = &p.second;
   if (!acc) {
      g_prof_acc.push_back({name, ProfAcc()});
      acc = &g_prof_acc.back().second;
   }
   acc->launches += launches;
   acc->seconds += seconds;
   acc->bytes += bytes;
}

static void prof_resolve() {
   if (g_pending.empty()) return;
   (void)hipStreamSynchronize(g_ctx.stream);
   for (auto& r : g_pending) {
      float ms = 0.f;
      (void)hipEventSynchronize(r.e1);  // launches on other streams (sweep handles)
      (void)hipEventElapsedTime(&ms, r.e0, r.e1);
      ProfAcc* acc = nullptr;
      for (auto& p : g_prof_acc)
         if (p.first == r.name) acc = &p.second;
      if (!acc) {
         g_prof_acc.push_back({r.name, ProfAcc()});
         acc = &g_prof_acc.back().second;
      }
      acc->launches++;
      acc->seconds += ms * 1e-3;
      acc->bytes += r.bytes;
      g_event_pool.push_back(r.e0);
      g_event_pool.push_back(r.e1);
   }
   g_pending.clear();
}

void set_error(const char* file, int line, const char* func, const std::string& msg) {
   const char* base = std::strrchr(file, '/');
   base = base ? base + 1 : file;
   g_last_error = std::string(base) + ":" + std::to_string(line) + ":" + func + "(): error: " + msg + ".";
}

Ctx& ctx() { return g_ctx; }

// Fixed device scratch of the allocation-free entry points (each device that
// loads the library holds one copy until the process ends).  Host code never
// uses this name as an address: it is a host-side stand-in.  pk_initialize
// resolves the copy of the current device with hipGetSymbolAddress and keeps
// that address per device; kernels receive it as an argument.
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {};

double* static_scratch() {
   const int d = g_ctx.device;
   return d >= 0 && d < kMaxDevices ? g_scratch_dev[d] : nullptr;
}

int ensure_init() {
   if (g_ctx.stream) return 0;
   return pk_initialize(-1);
}

int reserve_partials(size_t count) {
   Ctx& c = g_ctx;
   if (count <= c.partials_cap) return 0;
   PK_HIP(hipStreamSynchronize(c.stream));
   (void)hipFree(c.partials);
   c.partials = nullptr;
   PK_HIP(hipMalloc(&c.partials, sizeof(double) * count));
   c.partials_cap = count;
   return 0;
}

// ---- reductions ------------------------------------------------------------

template <class T>
__global__ void __launch_bounds__(kBlock) dot_kernel(const T* __restrict__ x, const T* __restrict__ y,
                                                     long n, double* partials, unsigned* ticket,
                                                     double* out) {
   double s = 0.0;
   for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
      s += (double)x[i] * (double)y[i];
   double v[1] = {s};
   block_reduce_finish<1>(v, partials, ticket, out);
}

int reduce_dot(const double* x, const double* y, long n, double* out_dev) {
   Ctx& c = ctx();
   int g = grid_for(n, kBlock, kMaxRedBlocks);
   dot_kernel<double><<<g, kBlock, 0, c.stream>>>(x, y, n, c.partials, c.ticket, out_dev);
   PK_HIP(hipGetLastError());
   return 0;
}

int reduce_dot_f(const float* x, const float* y, long n, double* out_dev) {
   Ctx& c = ctx();
   int g = grid_for(n, kBlock, kMaxRedBlocks);
   dot_kernel<float><<<g, kBlock, 0, c.stream>>>(x, y, n, c.partials, c.ticket, out_dev);
   PK_HIP(hipGetLastError());
   return 0;
}

// ---- elementwise vector kernels ---------------------------------------------

__global__ void set_kernel(double* x, long n, double a) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) x[i] = a;
}
// max |x| as the bit pattern of a non-negative double (the order of such bit
// patterns is the order of the values, so the atomic max is exact and does not
// depend on timing)
__global__ void __launch_bounds__(kBlock) absmax_kernel(const double* __restrict__ x, long n,
                                                        unsigned long long* out) {
   double m = 0.0;
   for (long l = blockIdx.x * (long)kBlock + threadIdx.x; l < n; l += (long)gridDim.x * kBlock)
      m = fmax(m, fabs(x[l]));
   for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
   __shared__ double sh[kBlock / 64];
   if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0) {
      for (int w = 1; w < kBlock / 64; w++) m = fmax(m, sh[w]);
      atomicMax(out, (unsigned long long)__double_as_longlong(m));
   }
}

// x in [-rel max|x|, 0) -> 0; count the entries below -rel max|x|
__global__ void __launch_bounds__(kBlock) clip_noise_kernel(double* x, long n, double rel,
                                                            const unsigned long long* amax,
                                                            unsigned long long* negatives) {
   const double fl = rel * __longlong_as_double((long long)*amax);
   unsigned long long cnt = 0;
   for (long l = blockIdx.x * (long)kBlock + threadIdx.x; l < n; l += (long)gridDim.x * kBlock) {
      const double v = x[l];
      if (v < 0.0) {
         if (-v <= fl)
            x[l] = 0.0;
         else
            cnt++;
      }
   }
   if (cnt) atomicAdd(negatives, cnt);
}

__global__ void scale_kernel(double* x, long n, double a) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) x[i] *= a;
}
__global__ void pointwise_mult_kernel(double* w, long n, const double* x, const double* y) {
   const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) w[i] = x[i] * y[i];
}

__global__ void axpy_kernel(double* y, long n, double a, const double* x) {
   long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (i < n) y[i] += a * x[i];
}

}  // namespace pk

using namespace pk;

extern "C" {

const char* pk_last_error(void) { return g_last_error.c_str(); }

int pk_initialize(int device) {
   Ctx& c = ctx();
   if (c.stream) return 0;
   int n = 0;
   PK_HIP(hipGetDeviceCount(&n));
   PK_REQUIRE(n > 0, "no HIP device available");
   if (device < 0) device = 0;
   PK_REQUIRE(device < n, "device index out of range");
   PK_HIP(hipSetDevice(device));
   c.device = device;
   hipDeviceProp_t prop;
   PK_HIP(hipGetDeviceProperties(&prop, device));
   c.num_cus = prop.multiProcessorCount;
   PK_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
   PK_HIP(hipMalloc(&c.partials, sizeof(double) * kMaxRedBlocks * kMaxRedVals));
   c.partials_cap = (size_t)kMaxRedBlocks * kMaxRedVals;
   PK_HIP(hipMalloc(&c.ticket, sizeof(unsigned) * 64));
   PK_HIP(hipMemsetAsync(c.ticket, 0, sizeof(unsigned) * 64, c.stream));
   PK_HIP(hipMalloc(&c.scalars, sizeof(double) * 256));
   PK_HIP(hipMemsetAsync(c.scalars, 0, sizeof(double) * 256, c.stream));
   PK_HIP(hipHostMalloc(&c.host_scalars, sizeof(double) * 256, hipHostMallocDefault));
   if (device < kMaxDevices && !g_scratch_dev[device]) {
      void* sp = nullptr;
      PK_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
      PK_REQUIRE(sp != nullptr, "no device address for the static scratch");
      g_scratch_dev[device] = (double*)sp;
   }
   PK_HIP(hipStreamSynchronize(c.stream));
   return 0;
}

}  // extern "C"

namespace pk {
// y = x over n floats, 4 per thread in one 16-byte access each
__global__ void __launch_bounds__(256) stream_copy_kernel(const float4* __restrict__ x, float4* __restrict__ y,
                                                          long n4) {
   long i = blockIdx.x * 256L + threadIdx.x;
   if (i < n4) y[i] = x[i];
}
}  // namespace pk

extern "C" {

int pk_stream_copy(long n, int reps, double* avg_seconds) {
   PK_CALL(ensure_init());
   PK_REQUIRE(n > 0 && n % 4 == 0 && reps > 0 && avg_seconds, "wrong arguments");
   Ctx& c = ctx();
   DBuf<float> x, y;
   PK_REQUIRE(!x.alloc(n) && !y.alloc(n), "device allocation failed");
   PK_HIP(hipMemsetAsync(x.p, 0, sizeof(float) * n, c.stream));
   const long n4 = n / 4;
   EventPair ev;
   PK_HIP(ev.create());
   stream_copy_kernel<<<grid_for(n4, 256), 256, 0, c.stream>>>((const float4*)x.p, (float4*)y.p, n4);
   PK_HIP(hipEventRecord(ev.e0, c.stream));
   for (int r = 0; r < reps; r++)
      stream_copy_kernel<<<grid_for(n4, 256), 256, 0, c.stream>>>((const float4*)x.p, (float4*)y.p, n4);
   PK_HIP(hipEventRecord(ev.e1, c.stream));
   PK_HIP(hipGetLastError());
   float ms = 0.f;
   PK_HIP(ev.ms(&ms));
   *avg_seconds = ms * 1e-3 / reps;
   return 0;
}

int pk_device_memory(double* free_bytes, double* total_bytes) {
   PK_CALL(ensure_init());
   size_t f = 0, t = 0;
   PK_HIP(hipMemGetInfo(&f, &t));
   if (free_bytes) *free_bytes = (double)f;
   if (total_bytes) *total_bytes = (double)t;
   return 0;
}

int pk_finalize(void) {
   Ctx& c = ctx();
   if (!c.stream) return 0;
   PK_HIP(hipStreamSynchronize(c.stream));
   (void)hipFree(c.partials);
   (void)hipFree(c.ticket);
   (void)hipFree(c.scalars);
   (void)hipHostFree(c.host_scalars);
   (void)hipStreamDestroy(c.stream);
   c = Ctx();
   return 0;
}

int pk_synchronize(void) {
   PK_CALL(ensure_init());
   PK_HIP(hipStreamSynchronize(ctx().stream));
   return 0;
}

int pk_vec_create(long n, pk_vec* v) {
   PK_CALL(ensure_init());
   PK_REQUIRE(n >= 0, "negative vector size");
   auto* p = new pk_vec_s();
   p->n = n;
   if (p->d.alloc(n ? n : 1)) {
      delete p;
      PK_FAIL("device allocation failed");
   }
   PK_HIP(hipMemsetAsync(p->d.p, 0, sizeof(double) * (n ? n : 1), ctx().stream));
   *v = p;
   return 0;
}

int pk_vec_destroy(pk_vec* v) {
   if (!v || !*v) return 0;
   (void)hipStreamSynchronize(ctx().stream);
   (*v)->d.free();
   delete *v;
   *v = nullptr;
   return 0;
}

long pk_vec_size(pk_vec v) { return v ? v->n : -1; }

int pk_vec_set(pk_vec v, double x) {
   PK_REQUIRE(v, "null vector");
   if (v->n == 0) return 0;
   set_kernel<<<grid_for(v->n), kBlock, 0, ctx().stream>>>(v->d.p, v->n, x);
   PK_HIP(hipGetLastError());
   return 0;
}

int pk_vec_set_values(pk_vec v, const double* host) {
   PK_REQUIRE(v && host, "null argument");
   PK_HIP(hipMemcpyAsync(v->d.p, host, sizeof(double) * v->n, hipMemcpyHostToDevice, ctx().stream));
   PK_HIP(hipStreamSynchronize(ctx().stream));
   return 0;
}

int pk_vec_get_values(pk_vec v, double* host) {
   PK_REQUIRE(v && host, "null argument");
   PK_HIP(hipMemcpyAsync(host, v->d.p, sizeof(double) * v->n, hipMemcpyDeviceToHost, ctx().stream));
   PK_HIP(hipStreamSynchronize(ctx().stream));
   return 0;
}

int pk_vec_copy(pk_vec src, pk_vec dst) {
   PK_REQUIRE(src && dst && src->n == dst->n, "vector size mismatch");
   PK_HIP(hipMemcpyAsync(dst->d.p, src->d.p, sizeof(double) * src->n, hipMemcpyDeviceToDevice,
                         ctx().stream));
   return 0;
}

int pk_vec_scale(pk_vec v, double a) {
   PK_REQUIRE(v, "null vector");
   if (v->n == 0) return 0;
   scale_kernel<<<grid_for(v->n), kBlock, 0, ctx().stream>>>(v->d.p, v->n, a);
   PK_HIP(hipGetLastError());
   return 0;
}

int pk_vec_clip_noise(pk_vec v, double rel, long* negatives) {
   PK_REQUIRE(v && rel >= 0.0, "wrong arguments");
   Ctx& c = ctx();
   if (negatives) *negatives = 0;
   if (v->n == 0) return 0;
   DBuf<unsigned long long> w;
   PK_REQUIRE(!w.alloc(2), "device allocation failed");
   PK_HIP(hipMemsetAsync(w.p, 0, sizeof(unsigned long long) * 2, c.stream));
   const int gb = grid_for(v->n, kBlock, 4 * kMaxRedBlocks);
   absmax_kernel<<<gb, kBlock, 0, c.stream>>>(v->d.p, v->n, w.p);
   clip_noise_kernel<<<gb, kBlock, 0, c.stream>>>(v->d.p, v->n, rel, w.p, w.p + 1);
   PK_HIP(hipGetLastError());
   unsigned long long h = 0;
   PK_HIP(hipMemcpyAsync(&h, w.p + 1, sizeof(h), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   if (negatives) *negatives = (long)h;
   return 0;
}

int pk_vec_pointwise_mult(pk_vec w, pk_vec x, pk_vec y) {
   PK_REQUIRE(w && x && y && x->n == w->n && y->n == w->n, "vector size mismatch");
   if (w->n == 0) return 0;
   pointwise_mult_kernel<<<grid_for(w->n), kBlock, 0, ctx().stream>>>(w->d.p, w->n, x->d.p, y->d.p);
   PK_HIP(hipGetLastError());
   return 0;
}

int pk_vec_axpy(pk_vec y, double a, pk_vec x) {
   PK_REQUIRE(x && y && x->n == y->n, "vector size mismatch");
   if (y->n == 0) return 0;
   axpy_kernel<<<grid_for(y->n), kBlock, 0, ctx().stream>>>(y->d.p, y->n, a, x->d.p);
   PK_HIP(hipGetLastError());
   return 0;
}

int pk_vec_dot(pk_vec x, pk_vec y, double* r) {
   PK_REQUIRE(x && y && x->n == y->n && r, "vector size mismatch");
   Ctx& c = ctx();
   PK_CALL(reduce_dot(x->d.p, y->d.p, x->n, c.scalars));
   PK_HIP(hipMemcpyAsync(c.host_scalars, c.scalars, sizeof(double), hipMemcpyDeviceToHost, c.stream));
   PK_HIP(hipStreamSynchronize(c.stream));
   *r = c.host_scalars[0];
   return 0;
}

int pk_vec_norm2(pk_vec x, double* r) {
   double d = 0.0;
   PK_CALL(pk_vec_dot(x, x, &d));
   *r = sqrt(d);
   return 0;
}

int pk_vec_device_ptr(pk_vec v, void** dev) {
   PK_REQUIRE(v && dev, "null argument");
   *dev = v->d.p;
   return 0;
}

int pk_profile_enable(int on) {
   PK_CALL(ensure_init());
   prof_resolve();
   g_prof = on != 0;
   return 0;
}

int pk_profile_reset(void) {
   prof_resolve();
   g_prof_acc.clear();
   return 0;
}

int pk_profile_get(const char* name, long* launches, double* seconds, double* bytes) {
   PK_REQUIRE(name, "null argument");
   prof_resolve();
   for (auto& p : g_prof_acc)
      if (p.first == name) {
         if (launches) *launches = p.second.launches;
         if (seconds) *seconds = p.second.seconds;
         if (bytes) *bytes = p.second.bytes;
         return 0;
      }
   if (launches) *launches = 0;
   if (seconds) *seconds = 0.0;
   if (bytes) *bytes = 0.0;
   return 0;
}

}  // extern "C"
