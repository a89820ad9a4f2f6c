// pampa command-line driver (cxx/main.cxx:4-28): initialise, time-step loop
// n = 0..ndt calling the C-ABI, finalise.
#include <cstdio>

#include "../../../include/pampa.h"

int main(int argc, char* argv[]) {
   double* dt = nullptr;
   int ndt = 0, error = 0;
   pampa_initialize(argc, argv, &dt, &ndt, &error);
   if (error) return 1;
   double t = 0.0;
   for (int n = 0; n < ndt + 1; n++) {
      double dtn = (n == 0) ? 0.0 : dt[n - 1];
      t += dtn;
      pampa_solve(n, dtn, t, &error);
      if (error) return 1;
   }
   pampa_finalize(&dt, &error);
   return error ? 1 : 0;
}
