// Reference-form S_N operator for inputs outside the sweepable regime
// (mixed-face-interpolation < 1, least-squares boundary correction, reflective
// boundaries): R and F of SNSolver::buildMatrices (SNSolver.cxx:344-631),
// assembled on the device by pk_sn_assemble and solved by the pk_eig eigen
// solver.  The host computes the least-squares weights of the boundary cells
// (SNSolver.cxx:212-268), including that routine's flat-index reads
// d(jg, f) = d[jg * num_dims + f], which decide the reference's numbers; the
// face-interpolation weights (SNSolver.cxx:159-208) and the reflective
// branch's axis lookup through the normals of cells 0, 1, 2
// (SNSolver.cxx:541-545) are computed by the device kernels.
#include "solver.hpp"

namespace pampa {

int SNSolver::assembleReferenceForm() {
   const Mesh& m = *mesh;
   const int C = m.num_cells;
   PAMPA_CHECK(C < 3, "the reference-form S_N operator needs at least 3 cells");
   // least-squares gradient weights of boundary cells: lsc[f][3] per face slot
   std::vector<double> lsc;
   if (ls) {
      lsc.assign(3 * m.fnb.size(), 0.0);
      const int nd = m.num_dims;
      PAMPA_CHECK(nd < 1 || nd > 3, "wrong number of dimensions");
      for (int i = 0; i < C; i++) {
         bool bnd = false;
         for (int f = m.face_ptr[i]; f < m.face_ptr[i + 1]; f++) bnd |= m.fnb[f] < 0;
         if (!bnd) continue;
         const int f0 = m.face_ptr[i], nf = m.face_ptr[i + 1] - f0;
         std::vector<double> d((size_t)nf * nd + 1, 0.0);
         for (int f = 0; f < nf; f++) {
            int i2 = m.fnb[f0 + f];
            const double* c2 = i2 < 0 ? &m.fcen[3L * (f0 + f)] : &m.cen[3L * i2];
            for (int id = 0; id < nd; id++) d[f * nd + id] = c2[id] - m.cen[3L * i + id];
         }
         double Gm[9] = {0}, Gi[9] = {0};
         for (int jg = 0; jg < nd; jg++)
            for (int ig = 0; ig < nd; ig++)
               for (int f = 0; f < nf; f++) Gm[jg * 3 + ig] += d[jg * nd + f] * d[f * nd + ig];
         if (nd == 1) {
            Gi[0] = 1.0 / Gm[0];
         } else if (nd == 2) {
            double det = Gm[0] * Gm[4] - Gm[1] * Gm[3];
            Gi[0] = Gm[4] / det;
            Gi[1] = -Gm[1] / det;
            Gi[3] = -Gm[3] / det;
            Gi[4] = Gm[0] / det;
         } else {
            double a = Gm[0], b = Gm[1], c = Gm[2], e = Gm[3], f_ = Gm[4], g = Gm[5], h = Gm[6],
                   k = Gm[7], l = Gm[8];
            double det = a * (f_ * l - g * k) - b * (e * l - g * h) + c * (e * k - f_ * h);
            Gi[0] = (f_ * l - g * k) / det;
            Gi[1] = (c * k - b * l) / det;
            Gi[2] = (b * g - c * f_) / det;
            Gi[3] = (g * h - e * l) / det;
            Gi[4] = (a * l - c * h) / det;
            Gi[5] = (c * e - a * g) / det;
            Gi[6] = (e * k - f_ * h) / det;
            Gi[7] = (b * h - a * k) / det;
            Gi[8] = (a * f_ - b * e) / det;
         }
         for (int f = 0; f < nf; f++)
            for (int id = 0; id < nd; id++)
               for (int jd = 0; jd < nd; jd++) lsc[3L * (f0 + f) + id] += Gi[id * 3 + jd] * d[jd * nd + f];
      }
   }
   // R and F assembled on the device from the mesh, the tables and these weights
   PK_CHECK(pk_mat_destroy(&R));
   PK_CHECK(pk_mat_destroy(&F));
   PK_CHECK(pk_sn_assemble(dmesh, dxs, M, dirs.data(), wts.data(), refl.data(), delta, ls ? lsc.data() : nullptr,
                           (int)bc_type.size(), bc_type.data(), &R, &F));
   return 0;
}

}  // namespace pampa
