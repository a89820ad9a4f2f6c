#include "mesh.hpp"

#include <algorithm>
#include <cmath>

#include "input.hpp"
#include "vtk.hpp"

namespace pampa {

int Mesh::readBC(const std::vector<std::string>& line, std::istream& in) {
   PAMPA_CHECK(line.size() < 3, "wrong number of arguments for keyword '" + line[0] + "'");
   if (bcs.empty()) bcs.resize(1 + boundaries.size());
   if (bcs.size() < 1 + boundaries.size()) bcs.resize(1 + boundaries.size());
   int b = findBoundary(line[1]);
   PAMPA_CHECK(b < 0, "wrong boundary name");
   unsigned i = 2;
   PAMPA_CHECK(input::read(bcs[b + 1], line, i, in), "wrong boundary condition");
   return 0;
}

int Mesh::toDevice(pk_mesh* m) const {
   PK_CHECK(pk_mesh_create(num_cells, face_ptr.data(), vol.data(), cen.data(), mat.data(),
                           farea.data(), fcen.data(), fnorm.data(), fnb.data(), m));
   return 0;
}

int Mesh::writeVTK(const std::string& prefix, int n) const {
   return vtk::write_mesh(prefix, n, points, cpt_ptr, cpts, mat, nodal);
}

// ---- Cartesian (CartesianMesh.cxx:4-143 read, :146-414 build) -------------------------

static int read_axis(const std::vector<std::string>& line, int& n, std::vector<double>& d,
                     std::istream& in) {
   PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + line[0] + "'");
   PAMPA_CHECK(input::read(n, -INT_MAX, INT_MAX, line[1]), "wrong number of cells");
   if (n > 0) {
      PAMPA_CHECK(input::read(d, n, 0.0, DBL_MAX, in), "wrong spacing data");
   } else {
      PAMPA_CHECK(input::read(d, 1, 0.0, DBL_MAX, in), "wrong spacing data");
      n = -n;
      d.assign(n, d[0]);
   }
   return 0;
}

int CartesianMesh::read(const std::string& filename, const std::string& base_dir) {
   std::ifstream in;
   PAMPA_CHECK(!input::open(in, filename, base_dir), "unable to open " + filename);
   while (true) {
      auto line = input::next_line(in);
      if (line.empty()) break;
      const std::string& k = line[0];
      if (k == "dx") {
         PAMPA_CHECK(read_axis(line, nx, dx, in), "wrong dx data");
         num_dims++;
         boundaries.push_back("-x");
         boundaries.push_back("+x");
      } else if (k == "dy") {
         PAMPA_CHECK(read_axis(line, ny, dy, in), "wrong dy data");
         if (ny > 1) num_dims++;
         boundaries.push_back("-y");
         boundaries.push_back("+y");
      } else if (k == "dz") {
         PAMPA_CHECK(read_axis(line, nz, dz, in), "wrong dz data");
         if (nz > 1) num_dims++;
         boundaries.push_back("-z");
         boundaries.push_back("+z");
      } else if (k == "bc") {
         PAMPA_CHECK(readBC(line, in), "wrong boundary condition");
      } else if (k == "materials") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int n = nx * std::max(ny, 1) * std::max(nz, 1), m = 0;
         PAMPA_CHECK(input::read(m, n, n, line[1]), "wrong number of materials");
         PAMPA_CHECK(input::read(mat_in, n, 0, INT_MAX, in), "wrong material data");
         for (auto& v : mat_in) v--;
      } else if (k == "nodal-indices") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int n = nx * std::max(ny, 1) * std::max(nz, 1), m = 0;
         PAMPA_CHECK(input::read(m, n, n, line[1]), "wrong number of nodal indices");
         PAMPA_CHECK(input::read(nodal, n, -INT_MAX, INT_MAX, in), "wrong nodal-index data");
      } else {
         PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
      }
   }
   if (bcs.size() < 1 + boundaries.size()) bcs.resize(1 + boundaries.size());
   return 0;
}

int CartesianMesh::build() {
   PAMPA_CHECK(mat_in.empty(), "missing material data");
   const int NY = std::max(ny, 1), NZ = std::max(nz, 1);
   std::vector<double> x(nx + 1, 0.0), y(NY + 1, 0.0), z(NZ + 1, 0.0);
   for (int i = 0; i < nx; i++) x[i + 1] = x[i] + dx[i];
   for (int j = 0; j < ny; j++) y[j + 1] = y[j] + dy[j];
   for (int k = 0; k < nz; k++) z[k + 1] = z[k] + dz[k];
   // tensor grid of points (the layout the VTK output refers to)
   num_points = (nx + 1) * (ny + 1) * (nz + 1);
   points.assign(3L * num_points, 0.0);
   for (int ip = 0, k = 0; k < nz + 1; k++)
      for (int j = 0; j < ny + 1; j++)
         for (int i = 0; i < nx + 1; i++, ip++) {
            points[3L * ip] = x[i];
            points[3L * ip + 1] = y[j];
            points[3L * ip + 2] = z[k];
         }
   // voids: same in every layer; per row, voids before/after the first cell
   std::vector<int> lvoid(NY, 0), rvoid(NY, 0);
   int nxy = 0;
   for (int k = 0, im = 0; k < NZ; k++)
      for (int j = 0; j < NY; j++) {
         bool seen = false;
         for (int i = 0; i < nx; i++, im++) {
            bool phys = mat_in[im] != -1;
            if (k == 0) {
               if (phys) {
                  seen = true;
                  nxy++;
               } else
                  (seen ? rvoid[j] : lvoid[j])++;
            } else {
               bool phys0 = mat_in[im - nx * NY] != -1;
               PAMPA_CHECK(phys != phys0, "wrong material definition");
            }
         }
      }
   num_cells = nxy * NZ;
   const int npc = (nz > 0) ? 8 : (ny > 0) ? 4 : 2;
   num_faces_max = (nz > 0) ? 6 : (ny > 0) ? 4 : 2;
   auto bidx = [&](const char* name) { return findBoundary(name); };
   int b_mx = bidx("-x"), b_px = bidx("+x"), b_my = bidx("-y"), b_py = bidx("+y"), b_mz = bidx("-z"),
       b_pz = bidx("+z");
   PAMPA_CHECK(b_mx < 0 || b_px < 0, "wrong boundary name");
   PAMPA_CHECK(ny > 0 && (b_my < 0 || b_py < 0), "wrong boundary name");
   PAMPA_CHECK(nz > 0 && (b_mz < 0 || b_pz < 0), "wrong boundary name");
   vol.resize(num_cells);
   cen.resize(3L * num_cells);
   mat.resize(num_cells);
   cpt_ptr.assign(1, 0);
   cpts.clear();
   face_ptr.assign(1, 0);
   farea.clear();
   fcen.clear();
   fnorm.clear();
   fnb.clear();
   std::vector<int> nodal_new;
   const int sx = 1, sy = nx + 1, sz = (nx + 1) * (ny + 1);
   for (int ic = 0, im = 0, k = 0; k < NZ; k++)
      for (int j = 0; j < NY; j++)
         for (int i = 0; i < nx; i++, im++) {
            if (mat_in[im] == -1) continue;
            const double dyj = dy[j], dzk = dz[k];
            // gmsh point order
            int p0 = i * sx + j * sy + k * sz;
            cpts.push_back(p0);
            cpts.push_back(p0 + sx);
            if (npc >= 4) {
               cpts.push_back(p0 + sx + sy);
               cpts.push_back(p0 + sy);
            }
            if (npc == 8) {
               cpts.push_back(p0 + sz);
               cpts.push_back(p0 + sx + sz);
               cpts.push_back(p0 + sx + sy + sz);
               cpts.push_back(p0 + sy + sz);
            }
            cpt_ptr.push_back((int)cpts.size());
            vol[ic] = (nz > 0) ? dx[i] * dyj * dzk : (ny > 0) ? dx[i] * dyj : dx[i];
            cen[3L * ic] = x[i] + 0.5 * dx[i];
            cen[3L * ic + 1] = y[j] + 0.5 * dyj;
            cen[3L * ic + 2] = z[k] + 0.5 * dzk;
            mat[ic] = mat_in[im];
            if (!nodal.empty()) nodal_new.push_back(nodal[im]);
            auto face = [&](double a, double cx, double cy, double cz, double nx_, double ny_,
                            double nz_, int nb) {
               farea.push_back(a);
               fcen.insert(fcen.end(), {cx, cy, cz});
               fnorm.insert(fnorm.end(), {nx_, ny_, nz_});
               fnb.push_back(nb);
            };
            const double axy = (nz > 0) ? dx[i] * dzk : dx[i];
            const double ax = (nz > 0) ? dyj * dzk : (ny > 0) ? dyj : 1.0;
            if (ny > 0) {
               int nb = (j == 0 || mat_in[im - nx] == -1) ? -b_my - 1
                                                          : ic - nx + lvoid[j] + rvoid[j - 1];
               face(axy, x[i] + 0.5 * dx[i], y[j], z[k] + 0.5 * dzk, 0.0, -1.0, 0.0, nb);
            }
            {
               int nb = (i == nx - 1 || mat_in[im + 1] == -1) ? -b_px - 1 : ic + 1;
               face(ax, x[i] + dx[i], y[j] + 0.5 * dyj, z[k] + 0.5 * dzk, 1.0, 0.0, 0.0, nb);
            }
            if (ny > 0) {
               int nb = (j == ny - 1 || mat_in[im + nx] == -1) ? -b_py - 1
                                                               : ic + nx - rvoid[j] - lvoid[j + 1];
               face(axy, x[i] + 0.5 * dx[i], y[j] + dyj, z[k] + 0.5 * dzk, 0.0, 1.0, 0.0, nb);
            }
            {
               int nb = (i == 0 || mat_in[im - 1] == -1) ? -b_mx - 1 : ic - 1;
               face(ax, x[i], y[j] + 0.5 * dyj, z[k] + 0.5 * dzk, -1.0, 0.0, 0.0, nb);
            }
            if (nz > 0) {
               face(dx[i] * dyj, x[i] + 0.5 * dx[i], y[j] + 0.5 * dyj, z[k], 0.0, 0.0, -1.0,
                    k == 0 ? -b_mz - 1 : ic - nxy);
               face(dx[i] * dyj, x[i] + 0.5 * dx[i], y[j] + 0.5 * dyj, z[k] + dzk, 0.0, 0.0, 1.0,
                    k == nz - 1 ? -b_pz - 1 : ic + nxy);
            }
            face_ptr.push_back((int)fnb.size());
            ic++;
         }
   nodal.swap(nodal_new);
   return 0;
}

// ---- unstructured extruded (UnstructuredExtrudedMesh.cxx:4-149 read, :153-364 build) ----

static double shoelace(const std::vector<double>& xy, const int* ids, int n) {
   double a = 0.0;
   for (int k = 0; k < n; k++) {
      const double* p1 = &xy[2L * ids[k]];
      const double* p2 = &xy[2L * ids[(k + 1) % n]];
      a += p1[0] * p2[1] - p2[0] * p1[1];
   }
   return 0.5 * a;
}

static void poly_centroid(double* c, const std::vector<double>& xy, const int* ids, int n, double a) {
   c[0] = c[1] = 0.0;
   for (int k = 0; k < n; k++) {
      const double* p1 = &xy[2L * ids[k]];
      const double* p2 = &xy[2L * ids[(k + 1) % n]];
      double da = p1[0] * p2[1] - p2[0] * p1[1];
      c[0] += (p1[0] + p2[0]) * da;
      c[1] += (p1[1] + p2[1]) * da;
   }
   c[0] *= (1.0 / (6.0 * a));
   c[1] *= (1.0 / (6.0 * a));
}

int UnstructuredExtrudedMesh::read(const std::string& filename, const std::string& base_dir) {
   std::ifstream in;
   PAMPA_CHECK(!input::open(in, filename, base_dir), "unable to open " + filename);
   while (true) {
      auto line = input::next_line(in);
      if (line.empty()) break;
      const std::string& k = line[0];
      if (k == "points") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(nxy_points, 1, INT_MAX, line[1]), "wrong number of points");
         PAMPA_CHECK(input::read_rows(xy, nxy_points, 2, -DBL_MAX, DBL_MAX, in), "wrong point data");
      } else if (k == "cells") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         int nt = 0;
         PAMPA_CHECK(input::read(nxy_cells, 1, INT_MAX, line[1]), "wrong number of cells");
         PAMPA_CHECK(input::read(nt, 1, INT_MAX, line[2]), "wrong number of cell points");
         PAMPA_CHECK(input::read_ragged(xy_ptr, xy_pts, nxy_cells, nt, 0, INT_MAX, in), "wrong cell data");
         num_dims += 2;
      } else if (k == "dz") {
         PAMPA_CHECK(read_axis(line, nz, dz, in), "wrong dz data");
         if (nz > 1) num_dims++;
         boundaries.push_back("-z");
         boundaries.push_back("+z");
      } else if (k == "boundary") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         boundaries.push_back(line[1]);
         xy_bnames.push_back(line[1]);
         int n = 0;
         PAMPA_CHECK(input::read(n, 0, INT_MAX, line[2]), "wrong number of boundary points");
         std::vector<int> pts;
         if (n > 0)
            PAMPA_CHECK(input::read(pts, n, 0, INT_MAX, in), "wrong boundary data");
         else
            default_boundary = (int)xy_bpoints.size();
         xy_bpoints.push_back(pts);
      } else if (k == "bc") {
         PAMPA_CHECK(readBC(line, in), "wrong boundary condition");
      } else if (k == "materials") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int n = nxy_cells * std::max(nz, 1), m = 0;
         PAMPA_CHECK(input::read(m, n, n, line[1]), "wrong number of materials");
         PAMPA_CHECK(input::read(mat_in, n, 0, INT_MAX, in), "wrong material data");
         for (auto& v : mat_in) v--;
      } else if (k == "nodal-indices") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int n = nxy_cells * std::max(nz, 1), m = 0;
         PAMPA_CHECK(input::read(m, n, n, line[1]), "wrong number of nodal indices");
         PAMPA_CHECK(input::read(nodal, n, -INT_MAX, INT_MAX, in), "wrong nodal-index data");
      } else {
         PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
      }
   }
   if (bcs.size() < 1 + boundaries.size()) bcs.resize(1 + boundaries.size());
   return 0;
}

int UnstructuredExtrudedMesh::build() {
   PAMPA_CHECK(mat_in.empty(), "missing material data");
   const int NZ = std::max(nz, 1);
   std::vector<double> z(NZ + 1, 0.0);
   for (int k = 0; k < nz; k++) z[k + 1] = z[k] + dz[k];
   num_points = nxy_points * (nz + 1);
   points.assign(3L * num_points, 0.0);
   for (int ip = 0, k = 0; k < nz + 1; k++)
      for (int i = 0; i < nxy_points; i++, ip++) {
         points[3L * ip] = xy[2L * i];
         points[3L * ip + 1] = xy[2L * i + 1];
         points[3L * ip + 2] = z[k];
      }
   num_cells = nxy_cells * NZ;
   // point -> (cells, then boundary codes) lists in the reference's order
   std::vector<std::vector<int>> pc(nxy_points);
   for (int i = 0; i < nxy_cells; i++)
      for (int q = xy_ptr[i]; q < xy_ptr[i + 1]; q++) pc[xy_pts[q]].push_back(i);
   for (size_t b = 0; b < xy_bpoints.size(); b++) {
      int gb = findBoundary(xy_bnames[b]);
      PAMPA_CHECK(gb < 0, "wrong boundary name");
      for (int p : xy_bpoints[b]) pc[p].push_back(-gb - 1);
   }
   std::vector<int> xyn(xy_pts.size(), 0);
   int nmax = 0;
   for (int i = 0; i < nxy_cells; i++) {
      int n = xy_ptr[i + 1] - xy_ptr[i];
      nmax = std::max(nmax, n);
      for (int f = 0; f < n; f++) {
         int ip1 = xy_pts[xy_ptr[i] + f], ip2 = xy_pts[xy_ptr[i] + (f + 1) % n];
         bool found = false;
         for (int i1 : pc[ip1])
            for (int i2 : pc[ip2])
               if (i1 == i2 && i1 != i) {
                  PAMPA_CHECK(xyn[xy_ptr[i] + f] > 0, "wrong mesh connectivity");
                  xyn[xy_ptr[i] + f] = i1;
                  found = true;
               }
         if (!found && default_boundary >= 0) {
            xyn[xy_ptr[i] + f] = -findBoundary(xy_bnames[default_boundary]) - 1;
            found = true;
         }
         PAMPA_CHECK(!found, "wrong mesh connectivity");
      }
   }
   num_faces_max = (nz > 0) ? nmax + 2 : nmax;
   int bz0 = -1, bz1 = -1;
   if (nz > 0) {
      bz0 = findBoundary("-z");
      bz1 = findBoundary("+z");
      PAMPA_CHECK(bz0 < 0 || bz1 < 0, "wrong boundary name");
   }
   vol.resize(num_cells);
   cen.resize(3L * num_cells);
   mat = mat_in;
   cpt_ptr.assign(1, 0);
   cpts.clear();
   face_ptr.assign(1, 0);
   farea.clear();
   fcen.clear();
   fnorm.clear();
   fnb.clear();
   for (int ic = 0, k = 0; k < NZ; k++) {
      const double dzk = dz[k];
      for (int i = 0; i < nxy_cells; i++, ic++) {
         const int* ids = &xy_pts[xy_ptr[i]];
         int n = xy_ptr[i + 1] - xy_ptr[i];
         for (int l = 0; l < n; l++) cpts.push_back(ids[l] + k * nxy_points);
         if (nz > 0)
            for (int l = 0; l < n; l++) cpts.push_back(ids[l] + (k + 1) * nxy_points);
         cpt_ptr.push_back((int)cpts.size());
         double a = shoelace(xy, ids, n);
         vol[ic] = (nz > 0) ? a * dzk : a;
         double c[2];
         poly_centroid(c, xy, ids, n, a);
         cen[3L * ic] = c[0];
         cen[3L * ic + 1] = c[1];
         cen[3L * ic + 2] = z[k] + 0.5 * dzk;
         for (int f = 0; f < n; f++) {
            const double* p1 = &xy[2L * ids[f]];
            const double* p2 = &xy[2L * ids[(f + 1) % n]];
            double ddx = p2[0] - p1[0], ddy = p2[1] - p1[1];
            double len = std::sqrt(std::pow(ddx, 2) + std::pow(ddy, 2));
            farea.push_back((nz > 0) ? len * dzk : len);
            fcen.insert(fcen.end(), {0.5 * (p1[0] + p2[0]), 0.5 * (p1[1] + p2[1]), z[k] + 0.5 * dzk});
            double nx_ = p2[1] - p1[1], ny_ = p1[0] - p2[0];
            double nn = std::sqrt(nx_ * nx_ + ny_ * ny_);
            fnorm.insert(fnorm.end(), {nx_ / nn, ny_ / nn, 0.0});
            int nb = xyn[xy_ptr[i] + f];
            fnb.push_back(nb >= 0 ? nb + k * nxy_cells : nb);
         }
         if (nz > 0) {
            double fa = shoelace(xy, ids, n);
            double fc[2];
            poly_centroid(fc, xy, ids, n, fa);
            farea.push_back(fa);
            fcen.insert(fcen.end(), {fc[0], fc[1], z[k]});
            fnorm.insert(fnorm.end(), {0.0, 0.0, -1.0});
            fnb.push_back(k == 0 ? -bz0 - 1 : ic - nxy_cells);
            farea.push_back(fa);
            fcen.insert(fcen.end(), {fc[0], fc[1], z[k] + dzk});
            fnorm.insert(fnorm.end(), {0.0, 0.0, 1.0});
            fnb.push_back(k == nz - 1 ? -bz1 - 1 : ic + nxy_cells);
         }
         face_ptr.push_back((int)fnb.size());
      }
   }
   return 0;
}

}  // namespace pampa
