// Reference-form S_N operator for inputs outside the sweepable regime
// (mixed-face-interpolation < 1, least-squares boundary correction, reflective
// boundaries): R and F of SNSolver::buildMatrices (SNSolver.cxx:344-631)
// assembled row by row through pk_mat_set_values, then solved by the pk_eig
// eigen solver.  The face-interpolation weights follow SNSolver.cxx:159-208 and
// the least-squares weights SNSolver.cxx:212-268, including that routine's
// flat-index reads d(jg, f) = d[jg * num_dims + f] and the reflective branch's
// lookup of the reflection axis through the normals of cells 0, 1, 2
// (SNSolver.cxx:541-545), which decide the reference's numbers.
#include <cmath>

#include "solver.hpp"

namespace pampa {

static double dist3(const double* a, const double* b) {
   double d = 0.0;
   for (int k = 0; k < 3; k++) d += std::pow(b[k] - a[k], 2);
   return std::sqrt(d);
}

static double dot3(const double* a, const double* b) {
   double x = 0.0;
   for (int k = 0; k < 3; k++) x += a[k] * b[k];
   return x;
}

int SNSolver::assembleReferenceForm() {
   const Mesh& m = *mesh;
   const int C = m.num_cells, GM = G * M;
   PAMPA_CHECK(C < 3, "the reference-form S_N operator needs at least 3 cells");
   // face-interpolation weights for interior faces
   std::vector<double> coef(4 * m.fnb.size(), 0.0);
   for (int i = 0; i < C; i++)
      for (int f = m.face_ptr[i]; f < m.face_ptr[i + 1]; f++) {
         int i2 = m.fnb[f];
         if (i2 < 0) continue;
         double rif = dist3(&m.fcen[3L * f], &m.cen[3L * i]);
         double ri2f = dist3(&m.fcen[3L * f], &m.cen[3L * i2]);
         double rii2 = dist3(&m.cen[3L * i], &m.cen[3L * i2]);
         coef[4 * f + 0] = (ri2f + delta * rif) / rii2;
         coef[4 * f + 1] = (1.0 - delta) * rif / rii2;
         coef[4 * f + 2] = (1.0 - delta) * ri2f / rii2;
         coef[4 * f + 3] = (rif + delta * ri2f) / rii2;
      }
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
   const long N = (long)C * GM;
   PK_CHECK(pk_mat_destroy(&R));
   PK_CHECK(pk_mat_destroy(&F));
   PK_CHECK(pk_mat_create(N, GM + m.num_faces_max, &R));
   PK_CHECK(pk_mat_create(N, GM, &F));
   std::vector<int> rc, fc;
   std::vector<double> rv, fv;
   for (int i = 0; i < C; i++) {
      const Material& mt = materials[m.mat[i]];
      const double V = m.vol[i];
      for (int g = 0; g < G; g++)
         for (int dm = 0; dm < M; dm++) {
            const int l = (i * G + g) * M + dm;
            const double* om = &dirs[3L * dm];
            double diag = mt.sigma_total[g] * V;
            for (int f = m.face_ptr[i]; f < m.face_ptr[i + 1]; f++) {
               int i2 = m.fnb[f];
               double w = dot3(om, &m.fnorm[3L * f]);
               double A = m.farea[f];
               if (i2 < 0) {
                  int t = bc_type[-i2];
                  if (t == PK_BC_VACUUM) {
                     if (w > 0.0) {
                        diag += w * A;
                        if (ls) {
                           double dp[3] = {m.fcen[3L * f] - m.cen[3L * i], m.fcen[3L * f + 1] - m.cen[3L * i + 1],
                                           m.fcen[3L * f + 2] - m.cen[3L * i + 2]};
                           for (int f2 = m.face_ptr[i]; f2 < m.face_ptr[i + 1]; f2++) {
                              int i3 = m.fnb[f2];
                              if (i3 < 0) continue;
                              double wi3 = dot3(dp, &lsc[3L * f2]);
                              diag += -wi3 * w * A;
                              int l3 = (i3 * G + g) * M + dm;
                              double r = wi3 * w * A;
                              PK_CHECK(pk_mat_set_values(R, 1, &l, 1, &l3, &r, PK_ADD_VALUES));
                           }
                        }
                     }
                  } else if (t == PK_BC_REFLECTIVE) {
                     if (w > 0.0) {
                        diag += w * A;
                     } else {
                        int m2 = -1;
                        const int lf = f - m.face_ptr[i];
                        for (int a = 0; a < 3; a++) {
                           // normal of face lf of cell a (the reference's lookup)
                           PAMPA_CHECK(lf >= m.face_ptr[a + 1] - m.face_ptr[a], "reflected direction not found");
                           double d = m.fnorm[3L * (m.face_ptr[a] + lf) + a];
                           if (std::fabs(d) > 1.0 - 1.0e-6) m2 = refl[3 * dm + a];
                        }
                        PAMPA_CHECK(m2 == -1, "reflected direction not found");
                        int l2 = (i * G + g) * M + m2;
                        double r = w * A;
                        PK_CHECK(pk_mat_set_values(R, 1, &l, 1, &l2, &r, PK_ADD_VALUES));
                     }
                  } else {
                     PAMPA_CHECK(true, "boundary condition not implemented");
                  }
               } else {
                  double wi = w > 0.0 ? coef[4 * f + 0] : coef[4 * f + 2];
                  double wi2 = w > 0.0 ? coef[4 * f + 1] : coef[4 * f + 3];
                  diag += wi * w * A;
                  int l2 = (i2 * G + g) * M + dm;
                  double r = wi2 * w * A;
                  PK_CHECK(pk_mat_set_values(R, 1, &l, 1, &l2, &r, PK_ADD_VALUES));
               }
            }
            rc.clear();
            rv.clear();
            fc.clear();
            fv.clear();
            rc.push_back(l);
            rv.push_back(diag);
            for (int g2 = 0; g2 < G; g2++)
               for (int m2 = 0; m2 < M; m2++) {
                  int l2 = (i * G + g2) * M + m2;
                  double s = -mt.sigma_scattering[(size_t)g2 * G + g] * wts[m2] * V;
                  if (l2 == l)
                     rv[0] += s;
                  else {
                     rc.push_back(l2);
                     rv.push_back(s);
                  }
                  fc.push_back(l2);
                  fv.push_back(mt.chi_effective[g] * mt.nu_sigma_fission[g2] * wts[m2] * V);
               }
            PK_CHECK(pk_mat_set_values(R, 1, &l, (int)rc.size(), rc.data(), rv.data(), PK_ADD_VALUES));
            PK_CHECK(pk_mat_set_values(F, 1, &l, (int)fc.size(), fc.data(), fv.data(), PK_INSERT_VALUES));
         }
   }
   PK_CHECK(pk_mat_assemble(R));
   PK_CHECK(pk_mat_assemble(F));
   PK_CHECK(pk_mat_set_block_size(R, GM));
   PK_CHECK(pk_mat_set_block_size(F, GM));
   return 0;
}

}  // namespace pampa
