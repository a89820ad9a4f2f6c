// Serving loop: eigenvalue requests on resident base cases.  Diffusion
// requests are continuously batched into one device eigen pool (pk_pool_*);
// discrete-ordinates requests are solved by the sweep engine (pk_sn_*), one
// worker thread per resident S_N case, each on the sweep handle's own stream,
// so that they run on the device beside the pool.
//
// A case is a diffusion input (.pmp) parsed and uploaded once: its mesh, its
// base cross-section tables and its R/F pattern stay on the device.  A request
// names a case and a seeded cross-section perturbation; its operators are the
// case's R and F with values recomputed on the device for the perturbed tables
// (same pattern, nothing allocated), copied into the pool at admission.  The
// loop is open: requests arrive at given times; between ticks of the pool the
// loop admits waiting requests in arrival order while the pool has room (slots
// and device memory, first come first served), and collects finished ones.  A
// case remembers the flux of its latest converged request and can start the
// next one from it (warm start).  An S_N case takes its requests one at a time
// in arrival order; its sweep handle keeps the previous solution for the warm
// start.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <deque>
#include <memory>
#include <thread>

#include "../../../include/pampa.h"
#include "driver.hpp"

namespace pampa {

namespace {

// SplitMix64 with u = (x >> 11) 2^-53: portable seeded draws.
struct SplitMix64 {
   unsigned long long s;
   explicit SplitMix64(unsigned long long seed) : s(seed) {}
   unsigned long long next() {
      s += 0x9E3779B97F4A7C15ULL;
      unsigned long long z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      return z ^ (z >> 31);
   }
   double uniform(double a, double b) {
#pragma clang fp contract(off)
      const double u = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
      return a + (b - a) * u;
   }
};

// Every total, scattering, diffusion and nu-fission value of every material
// scaled by its own U[1-rel, 1+rel] factor, drawn material by material in the
// order sigma_t[g], sigma_s[g][g'] (row-major), D[g], nu sigma_f[g]; then
// sigma_t[g] = max(sigma_t[g], sum_g' sigma_s[g][g']) keeps absorption >= 0.
// Seed 0 leaves the tables as they are.
void perturb(const XSTables& base, unsigned long long seed, double rel, XSTables& out) {
#pragma clang fp contract(off)
   out = base;
   if (seed == 0 || rel == 0.0) return;
   SplitMix64 rng(seed);
   const int G = base.G;
   const double lo = 1.0 - rel, hi = 1.0 + rel;
   for (int m = 0; m < base.nm; m++) {
      for (int g = 0; g < G; g++) out.st[m * G + g] *= rng.uniform(lo, hi);
      for (int g = 0; g < G * G; g++) out.ss[(size_t)m * G * G + g] *= rng.uniform(lo, hi);
      if (!out.D.empty())
         for (int g = 0; g < G; g++) out.D[m * G + g] *= rng.uniform(lo, hi);
      for (int g = 0; g < G; g++) out.nf[m * G + g] *= rng.uniform(lo, hi);
      for (int g = 0; g < G; g++) {
         double sum = 0.0;
         for (int h = 0; h < G; h++) sum += out.ss[(size_t)m * G * G + (size_t)g * G + h];
         out.st[m * G + g] = std::max(out.st[m * G + g], sum);
      }
   }
}

struct Case {
   std::unique_ptr<Driver> driver;  // owns the mesh, the materials and the solver
   DiffusionSolver* solver = nullptr;
   XSTables base, work;
   pk_xs xs = nullptr;  // tables of the request being admitted
   pk_mat R = nullptr, F = nullptr;
   pk_vec warm = nullptr, tmp = nullptr;
   bool has_warm = false;
   long n = 0;
   double footprint = 0.0;
   // S_N case: workers, each with its own solver (sweep handle, stream and
   // previous solution), and the requests waiting for one of them
   struct Worker {
      std::unique_ptr<Driver> driver;
      pk_sn sn = nullptr;
      XSTables work;
      std::thread thread;
      std::atomic<bool> busy{false};
      int req = -1;  // request on the worker
      double k = 0.0, t_done = 0.0;
      int outer = 0, conv = 0, rc = 0;
      long sweeps = 0;
      std::string err;
      std::vector<double> phiw;
      ~Worker() {
         if (thread.joinable()) thread.join();
         if (driver) driver->finalize();
      }
   };
   std::vector<std::unique_ptr<Worker>> workers;
   bool sn = false;
   std::deque<int> queue;  // waiting requests, arrival order
   ~Case() {
      workers.clear();
      pk_mat_destroy(&R);
      pk_mat_destroy(&F);
      pk_vec_destroy(&warm);
      pk_vec_destroy(&tmp);
      pk_xs_destroy(&xs);
      if (driver) driver->finalize();
   }
};

}  // namespace

struct Server {
   pk_batch pool = nullptr;
   int precision = PK_PREC_FP64, restart = 30, max_slots = 0;
   double pool_bytes = 0.0;
   double tol = 1e-8, reduction = 0.1, shift = 0.1;
   std::vector<std::unique_ptr<Case>> cases;
   double seconds = 0.0;
   long ticks = 0;
   int max_live = 0;
   double progress = 0.0;  // > 0: a progress line on stderr every `progress` seconds of a run
   ~Server() { pk_batch_destroy(&pool); }

   int sn_workers = 1;  // workers (sweep handles) per S_N case

   // One S_N request on a worker thread: perturbed tables, a solve (warm from
   // the worker's previous one unless `warm` is 0), the moment to the host.
   static void solveSN(const Case* c, Case::Worker* w, unsigned long long seed, double rel, int warm,
                       std::chrono::steady_clock::time_point t0) {
      w->err.clear();
      perturb(c->base, seed, rel, w->work);
      int rc = pk_sn_set_xs(w->sn, w->work.st.data(), w->work.ss.data(), w->work.nf.data(), w->work.kf.data(),
                            w->work.ch.data());
      if (!rc && !warm) rc = pk_sn_reset(w->sn);
      if (!rc) rc = pk_sn_solve(w->sn, &w->k, nullptr, nullptr);
      w->conv = rc == 0;
      // a solve that stops at its iteration limit is reported, not an error
      if (rc && std::string(pk_last_error()).find("did not converge") != std::string::npos) rc = 0;
      if (!rc) rc = pk_sn_get_stats(w->sn, &w->outer, &w->sweeps, nullptr);
      if (!rc) rc = pk_sn_get_moment(w->sn, w->phiw.data());
      if (rc) w->err = pk_last_error();
      w->rc = rc;
      w->t_done = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      w->busy.store(false, std::memory_order_release);
   }

   int requestOperators(Case& c, unsigned long long seed, double rel) {
      perturb(c.base, seed, rel, c.work);
      PK_CHECK(pk_xs_set(c.xs, c.work.st.data(), c.work.ss.data(), c.work.nf.data(), c.work.kf.data(),
                         c.work.ch.data(), c.work.D.empty() ? nullptr : c.work.D.data()));
      PK_CHECK(pk_diffusion_refill(c.solver->deviceMesh(), c.xs, c.R, c.F));
      return 0;
   }
};

}  // namespace pampa

using namespace pampa;

struct pampa_server_s : Server {};

extern "C" {

void pampa_serve_create(double budget_bytes, int max_slots, int precision, pampa_server* S, int* error) {
   *error = 0;
   auto fail = [&](const char* msg) {
      report_error(__FILE__, __LINE__, __func__, msg);
      *error = 1;
   };
   if (!S || max_slots < 1 || (precision != PK_PREC_FP64 && precision != PK_PREC_MIXED)) return fail("wrong arguments");
   if (pk_initialize(0)) return fail(pk_last_error());
   if (budget_bytes <= 0.0) {
      double free_b = 0.0, total_b = 0.0;
      if (pk_device_memory(&free_b, &total_b)) return fail(pk_last_error());
      budget_bytes = 0.8 * free_b;
   }
   // capacity from the budget: per row of the row space, room for 24 stored
   // entries of R (and A), 16 of F and 16 block-inverse values (16-group 3-D
   // operators); rows scale down until the pool fits the budget
   // GMRES restart length of the pool (one tick = one cycle): PETSc's 30, as
   // in the drop-in.  On warm-started request streams 60 gave the same solves/s
   // with a higher median latency, and slowed the one-slot FP64 engine by 9 %.
   // PK_SERVE_RESTART overrides it (an integer in 1..64; anything else is
   // ignored) for tuning.
   int restart = 30;
   if (const char* e = getenv("PK_SERVE_RESTART")) {
      char* end = nullptr;
      const long v = strtol(e, &end, 10);
      if (end != e && *end == '\0' && v >= 1 && v <= 64) restart = (int)v;
   }
   long rows = 1L << 30;
   while (rows > 256 && pk_pool_bytes(rows, rows * 24, rows * 16, rows * 16, max_slots, precision, restart) >
                            budget_bytes)
      rows = rows * 3 / 4;
   rows = std::min(rows, (long)((1L << 31) - 256));
   auto* s = new pampa_server_s();
   s->precision = precision;
   s->restart = restart;
   s->max_slots = max_slots;
   s->pool_bytes = pk_pool_bytes(rows, rows * 24, rows * 16, rows * 16, max_slots, precision, restart);
   if (pk_pool_create(rows, rows * 24, rows * 16, rows * 16, max_slots, precision, restart, &s->pool)) {
      delete s;
      return fail(pk_last_error());
   }
   *S = s;
}

void pampa_serve_set_options(pampa_server S, double tol, double reduction, double shift, int* error) {
   *error = (!S || pk_pool_set_options(S->pool, tol, 100000, reduction, 2000, shift)) ? 1 : 0;
   if (*error) {
      report_error(__FILE__, __LINE__, __func__, S ? pk_last_error() : "null server");
      return;
   }
   S->tol = tol;
   S->reduction = reduction;
   S->shift = shift;
}

void pampa_serve_add_case(pampa_server S, const char* input, int* case_id, long* rows, int* groups,
                          int* error) {
   *error = 1;
   if (!S || !input || !case_id) return;
   auto c = std::make_unique<Case>();
   c->driver = std::make_unique<Driver>();
   std::string in(input), prog("pampa"), silent("-silent");
   std::vector<char*> argv = {&prog[0], &in[0], &silent[0]};
   std::vector<double> dt;
   if (c->driver->initialize((int)argv.size(), argv.data(), dt)) return;
   c->solver = dynamic_cast<DiffusionSolver*>(c->driver->mainSolver());
   if (!c->solver) {
      // S_N input: sn_workers solvers of the input, each a worker
      c->sn = true;
      for (int k = 0; k < S->sn_workers; k++) {
         auto w = std::make_unique<Case::Worker>();
         if (k == 0) {
            w->driver = std::move(c->driver);
         } else {
            w->driver = std::make_unique<Driver>();
            if (w->driver->initialize((int)argv.size(), argv.data(), dt)) return;
         }
         auto* sn = dynamic_cast<SNSolver*>(w->driver->mainSolver());
         if (!sn || !sn->sweepHandle()) {
            report_error(__FILE__, __LINE__, __func__,
                         "serving takes diffusion inputs and sweepable S_N inputs (pure upwind, LS off, vacuum)");
            return;
         }
         if (k == 0) {
            if (sn->materialTables(false, c->base)) return;
            c->n = sn->numCells() * (long)sn->groups();
            if (groups) *groups = sn->groups();
         }
         w->sn = sn->sweepHandle();
         w->phiw.assign(c->n, 0.0);
         if (pk_sn_set_precision(w->sn, S->precision) || pk_sn_prepare(w->sn)) {
            report_error(__FILE__, __LINE__, __func__, pk_last_error());
            return;
         }
         c->workers.push_back(std::move(w));
      }
      *case_id = (int)S->cases.size();
      if (rows) *rows = c->n;
      S->cases.push_back(std::move(c));
      *error = 0;
      return;
   }
   if (c->solver->materialTables(true, c->base)) return;
   if (c->base.feedback()) {
      report_error(__FILE__, __LINE__, __func__, "served cases need constant cross sections per material");
      return;
   }
   const int G = c->solver->groups();
   c->n = c->solver->numCells() * (long)G;
   if (pk_xs_create(c->base.nm, G, c->base.st.data(), c->base.ss.data(), c->base.nf.data(), c->base.kf.data(),
                    c->base.ch.data(), c->base.D.data(), &c->xs) ||
       c->solver->assembleBase(&c->R, &c->F) || pk_vec_create(c->n, &c->warm) || pk_vec_create(c->n, &c->tmp) ||
       pk_pool_footprint(c->R, c->F, S->precision, S->restart, &c->footprint)) {
      report_error(__FILE__, __LINE__, __func__, pk_last_error());
      return;
   }
   *case_id = (int)S->cases.size();
   if (rows) *rows = c->n;
   if (groups) *groups = G;
   S->cases.push_back(std::move(c));
   *error = 0;
}

void pampa_serve_case_tables(pampa_server S, int case_id, unsigned long long seed, double rel, double* sigt,
                             double* sigs, double* D, double* nusf, int* error) {
   *error = 1;
   if (!S || case_id < 0 || case_id >= (int)S->cases.size()) return;
   Case& c = *S->cases[case_id];
   XSTables t;
   perturb(c.base, seed, rel, t);
   if (sigt) std::copy(t.st.begin(), t.st.end(), sigt);
   if (sigs) std::copy(t.ss.begin(), t.ss.end(), sigs);
   if (D) std::copy(t.D.begin(), t.D.end(), D);
   if (nusf) std::copy(t.nf.begin(), t.nf.end(), nusf);
   *error = 0;
}

void pampa_serve_footprint(pampa_server S, int case_id, double* bytes, double* pool_bytes, int* error) {
   *error = 1;
   if (!S || case_id < 0 || case_id >= (int)S->cases.size()) return;
   if (bytes) *bytes = S->cases[case_id]->footprint;
   if (pool_bytes) *pool_bytes = S->pool_bytes;
   *error = 0;
}

void pampa_serve_run(pampa_server S, int n, const double* arrival, const int* case_id,
                     const unsigned long long* seed, double rel, int warm, double* keff, double* latency,
                     int* outer, long* inner, int* converged, double* const* flux, int* error) {
   *error = 1;
   if (!S || n < 0 || (n && (!arrival || !case_id || !seed))) return;
   for (int r = 0; r < n; r++)
      if (case_id[r] < 0 || case_id[r] >= (int)S->cases.size()) {
         report_error(__FILE__, __LINE__, __func__, "request for an unknown case");
         return;
      }
   for (int r = 1; r < n; r++)
      if (arrival[r] < arrival[r - 1]) {
         report_error(__FILE__, __LINE__, __func__, "arrival times must not decrease");
         return;
      }
   using clk = std::chrono::steady_clock;
   const auto t0 = clk::now();
   auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
   std::deque<int> pending;
   std::vector<int> slot_req(S->max_slots, -1);
   std::vector<int> ready(S->max_slots);
   int next = 0, done = 0, live = 0;
   long ticks0 = 0;
   pk_pool_capacity(S->pool, nullptr, nullptr, nullptr, &ticks0);
   auto check = [&](int rc) {
      if (rc) report_error(__FILE__, __LINE__, "pampa_serve_run", pk_last_error());
      return rc != 0;
   };
   double next_report = S->progress;
   int sn_busy = 0;  // S_N workers running
   // no worker carries a request of an earlier run into this one
   for (auto& cp : S->cases)
      for (auto& w : cp->workers) {
         if (w->thread.joinable()) w->thread.join();
         w->req = -1;
      }
   // a worker that has finished: its results into the request's outputs
   auto finish_sn = [&](Case::Worker& w) {
      w.thread.join();
      sn_busy--;
      const int r = w.req;
      if (w.rc) {
         report_error(__FILE__, __LINE__, "pampa_serve_run", w.err.c_str());
         return true;
      }
      if (keff) keff[r] = w.k;
      if (latency) latency[r] = w.t_done - arrival[r];
      if (outer) outer[r] = w.outer;
      if (inner) inner[r] = w.sweeps;
      if (converged) converged[r] = w.conv;
      if (flux && flux[r]) std::copy(w.phiw.begin(), w.phiw.end(), flux[r]);
      w.req = -1;
      done++;
      return false;
   };
   auto join_all = [&] {
      for (auto& cp : S->cases)
         for (auto& w : cp->workers)
            if (w->thread.joinable()) w->thread.join();
   };
   while (done < n) {
      const double t = now();
      if (S->progress > 0.0 && t >= next_report) {
         long tk = 0;
         pk_pool_capacity(S->pool, nullptr, nullptr, nullptr, &tk);
         fprintf(stderr, "[serve] %.1f s: %d of %d done, %d in flight (%d S_N), %zu waiting, %ld ticks\n", t, done,
                 n, live + sn_busy, sn_busy, pending.size(), tk - ticks0);
         fflush(stderr);
         next_report = t + S->progress;
      }
      while (next < n && arrival[next] <= t) {
         Case& c = *S->cases[case_id[next]];
         if (c.sn)
            c.queue.push_back(next++);
         else
            pending.push_back(next++);
      }
      // S_N requests: collect finished workers, give waiting requests to idle ones
      for (auto& cp : S->cases) {
         Case& c = *cp;
         if (!c.sn) continue;
         for (auto& wp : c.workers) {
            Case::Worker& w = *wp;
            if (w.thread.joinable() && !w.busy.load(std::memory_order_acquire)) {
               if (finish_sn(w)) {
                  join_all();
                  return;
               }
            }
            if (!w.thread.joinable() && !c.queue.empty()) {
               const int r = c.queue.front();
               c.queue.pop_front();
               w.req = r;
               w.busy.store(true, std::memory_order_release);
               w.thread = std::thread(Server::solveSN, &c, &w, seed[r], rel, warm, t0);
               sn_busy++;
            }
         }
      }
      // admission in arrival order while the pool has room
      while (!pending.empty()) {
         const int r = pending.front();
         Case& c = *S->cases[case_id[r]];
         int fits = 0;
         if (check(pk_pool_fits(S->pool, c.R, c.F, &fits))) {
            join_all();
            return;
         }
         if (!fits && live == 0) {
            report_error(__FILE__, __LINE__, __func__, "request larger than the eigen pool");
            join_all();
            return;
         }
         if (!fits) break;
         if (S->requestOperators(c, seed[r], rel)) {
            join_all();
            return;
         }
         int slot = -1;
         int rc = pk_pool_admit(S->pool, c.R, c.F, (warm && c.has_warm) ? c.warm : nullptr, &slot);
         if (rc == PK_POOL_FULL) break;
         if (check(rc)) {
            join_all();
            return;
         }
         slot_req[slot] = r;
         pending.pop_front();
         live++;
      }
      S->max_live = std::max(S->max_live, live + sn_busy);
      if (live == 0) {
         // no diffusion request in the pool: wait for the next arrival or a worker
         double wait = next < n ? arrival[next] - now() : 1.0;
         if (sn_busy) wait = std::min(wait, 2e-4);
         if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
         continue;
      }
      int nlive = 0, ndone = 0;
      if (check(pk_pool_step(S->pool, 1, &nlive, &ndone))) {
         join_all();
         return;
      }
      if (ndone == 0) continue;
      int cnt = 0;
      if (check(pk_pool_collect(S->pool, S->max_slots, ready.data(), &cnt))) {
         join_all();
         return;
      }
      const double tc = now();
      for (int i = 0; i < cnt; i++) {
         const int slot = ready[i];
         const int r = slot_req[slot];
         Case& c = *S->cases[case_id[r]];
         double k = 0.0, res = 0.0;
         int o = 0, conv = 0;
         long in = 0;
         if (check(pk_pool_result(S->pool, slot, &k, c.tmp, &o, &in, &res, &conv))) {
            join_all();
            return;
         }
         if (keff) keff[r] = k;
         if (latency) latency[r] = tc - arrival[r];
         if (outer) outer[r] = o;
         if (inner) inner[r] = in;
         if (converged) converged[r] = conv;
         if ((flux && flux[r] && check(pk_vec_get_values(c.tmp, flux[r]))) ||
             (conv && check(pk_vec_copy(c.tmp, c.warm))) || check(pk_pool_release(S->pool, slot))) {
            join_all();
            return;
         }
         if (conv) c.has_warm = true;
         slot_req[slot] = -1;
         live--;
         done++;
      }
   }
   long ticks1 = 0;
   pk_pool_capacity(S->pool, nullptr, nullptr, nullptr, &ticks1);
   S->ticks += ticks1 - ticks0;
   S->seconds += now();
   *error = 0;
}

void pampa_serve_time_spmv(pampa_server S, int reps, double* seconds, double* bytes, double* rows, int* error) {
   *error = 1;
   if (!S || reps < 1 || S->cases.empty()) return;
   // every case once (unperturbed) in the pool, one tick to build the inner
   // operators, the timed SpMV launch over all of them, then release
   std::vector<int> slots;
   double nrows = 0.0;
   auto cleanup = [&] {
      for (int sl : slots) pk_pool_release(S->pool, sl);
   };
   for (auto& cp : S->cases) {
      Case& c = *cp;
      if (c.sn) continue;
      int slot = -1;
      if (S->requestOperators(c, 0, 0.0)) return cleanup();
      int rc = pk_pool_admit(S->pool, c.R, c.F, nullptr, &slot);
      if (rc) {
         report_error(__FILE__, __LINE__, __func__, rc == PK_POOL_FULL ? "cases do not fit the pool together"
                                                                     : pk_last_error());
         return cleanup();
      }
      slots.push_back(slot);
      nrows += (double)c.n;
   }
   int live = 0, done = 0;
   if (pk_pool_step(S->pool, 1, &live, &done) || pk_pool_time_spmv(S->pool, reps, seconds, bytes)) {
      report_error(__FILE__, __LINE__, __func__, pk_last_error());
      return cleanup();
   }
   if (rows) *rows = nrows;
   cleanup();
   *error = 0;
}

void pampa_serve_set_sn_workers(pampa_server S, int workers, int* error) {
   *error = (S && workers >= 1) ? 0 : 1;
   if (!*error) S->sn_workers = workers;
}

void pampa_serve_set_progress(pampa_server S, double seconds, int* error) {
   *error = S ? 0 : 1;
   if (S) S->progress = seconds;
}

void pampa_serve_forget_warm(pampa_server S, int* error) {
   *error = S ? 0 : 1;
   if (S)
      for (auto& c : S->cases) {
         c->has_warm = false;
         for (auto& w : c->workers)
            if (pk_sn_reset(w->sn)) *error = 1;
      }
}

void pampa_serve_stats(pampa_server S, double* seconds, long* ticks, int* max_live, double* pool_bytes,
                       int* error) {
   *error = S ? 0 : 1;
   if (!S) return;
   if (seconds) *seconds = S->seconds;
   if (ticks) *ticks = S->ticks;
   if (max_live) *max_live = S->max_live;
   if (pool_bytes) *pool_bytes = S->pool_bytes;
}

void pampa_serve_destroy(pampa_server* S, int* error) {
   *error = 0;
   if (!S || !*S) return;
   delete *S;
   *S = nullptr;
}

}  // extern "C"
