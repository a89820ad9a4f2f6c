// Mesh data read from a plain-text file (PartitionedMesh.cxx:4-263) for
// single-domain use, and the mesh edits the parser applies before the first
// solver: split materials (Mesh.cxx:9-60) and the removal of boundary-condition
// materials (Mesh.cxx:63-204).
#include <algorithm>

#include "input.hpp"
#include "material.hpp"
#include "mesh.hpp"

namespace pampa {

void Mesh::addBoundary(const std::string& name) {
   boundaries.push_back(name);
   if (bcs.size() < 1 + boundaries.size()) bcs.resize(1 + boundaries.size());
}

// n lines of doubles, any length each, at most nt in all
static int read_ragged_d(std::vector<int>& ptr, std::vector<double>& v, int n, int nt, double lo, double hi,
                         std::istream& in) {
   ptr.assign(1, 0);
   v.clear();
   for (int l = 0; l < n; l++) {
      auto line = input::next_line(in);
      PAMPA_CHECK(line.empty(), "missing data");
      PAMPA_CHECK((long)v.size() + (long)line.size() > nt, "out-of-bounds data");
      for (auto& w : line) {
         double x = 0.0;
         PAMPA_CHECK(input::read(x, lo, hi, w), "wrong data");
         v.push_back(x);
      }
      ptr.push_back((int)v.size());
   }
   return 0;
}

// the number of values of each row must be the cell's number of faces
static int check_rows(const std::vector<int>& ptr, const std::vector<int>& fptr) {
   for (size_t i = 0; i + 1 < ptr.size(); i++)
      PAMPA_CHECK(ptr[i + 1] - ptr[i] != fptr[i + 1] - fptr[i], "wrong number of face values for a cell");
   return 0;
}

int PartitionedMesh::read(const std::string& filename, const std::string& base_dir) {
   std::ifstream in;
   PAMPA_CHECK(!input::open(in, filename, base_dir), "unable to open " + filename);
   int num_cell_faces = 0;
   std::vector<int> ptr;
   while (true) {
      auto line = input::next_line(in);
      if (line.empty()) break;
      const std::string& k = line[0];
      if (k == "points") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(num_points, 1, INT_MAX, line[1]), "wrong number of points");
         PAMPA_CHECK(input::read_rows(points, num_points, 3, -DBL_MAX, DBL_MAX, in), "wrong point data");
      } else if (k == "cells") {
         PAMPA_CHECK(line.size() != 4, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(num_cells, 1, INT_MAX, line[1]), "wrong number of cells");
         PAMPA_CHECK(input::read(num_ghost, 1, INT_MAX, line[2]), "wrong number of ghost cells");
         PAMPA_CHECK(input::read(num_cells_global, 1, INT_MAX, line[3]), "wrong global number of cells");
      } else if (k == "cell-points") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         int rows = 0, total = 0;
         PAMPA_CHECK(input::read(rows, num_cells, num_cells, line[1]), "wrong number of cells");
         PAMPA_CHECK(input::read(total, 1, INT_MAX, line[2]), "wrong number of cell points");
         PAMPA_CHECK(input::read_ragged(cpt_ptr, cpts, num_cells, total, 0, INT_MAX, in), "wrong cell-point data");
      } else if (k == "cell-volumes") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int rows = 0;
         PAMPA_CHECK(input::read(rows, num_cells, num_cells, line[1]), "wrong number of cell volumes");
         PAMPA_CHECK(input::read(vol, num_cells, 0.0, DBL_MAX, in), "wrong cell-volume data");
      } else if (k == "cell-centroids") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         const int tot = num_cells + num_ghost;
         int rows = 0;
         PAMPA_CHECK(input::read(rows, tot, tot, line[1]), "wrong number of cell centroids");
         PAMPA_CHECK(input::read_rows(cen, tot, 3, -DBL_MAX, DBL_MAX, in), "wrong cell-centroid data");
         // dimensions from the spread of the centroids (PartitionedMesh.cxx:84-97)
         num_dims = 1;
         for (int i = 1; i < num_cells; i++)
            if (cen[3L * i + 1] != cen[1]) {
               num_dims++;
               break;
            }
         for (int i = 1; i < num_cells; i++)
            if (cen[3L * i + 2] != cen[2]) {
               num_dims++;
               break;
            }
      } else if (k == "cell-materials") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         const int tot = num_cells + num_ghost;
         int rows = 0;
         PAMPA_CHECK(input::read(rows, tot, tot, line[1]), "wrong number of cell materials");
         PAMPA_CHECK(input::read(mat, tot, 0, INT_MAX, in), "wrong cell-material data");
         for (auto& v : mat) v--;  // 1-based in the file
      } else if (k == "cell-nodal-indices") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int rows = 0;
         PAMPA_CHECK(input::read(rows, num_cells, num_cells, line[1]), "wrong number of cell nodal indices");
         PAMPA_CHECK(input::read(nodal, num_cells, -INT_MAX, INT_MAX, in), "wrong nodal-index data");
      } else if (k == "cell-global-indices") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         const int tot = num_cells + num_ghost;
         int rows = 0;
         PAMPA_CHECK(input::read(rows, tot, tot, line[1]), "wrong number of cell global indices");
         PAMPA_CHECK(input::read(global, tot, 0, INT_MAX, in), "wrong global-index data");
      } else if (k == "faces") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int rows = 0;
         PAMPA_CHECK(input::read(rows, num_cells, num_cells, line[1]), "wrong number of cell faces");
         std::vector<int> nf;
         PAMPA_CHECK(input::read(nf, num_cells, 1, INT_MAX, in), "wrong cell-face data");
         face_ptr.assign(num_cells + 1, 0);
         num_faces_max = 0;
         for (int i = 0; i < num_cells; i++) {
            face_ptr[i + 1] = face_ptr[i] + nf[i];
            num_faces_max = std::max(num_faces_max, nf[i]);
         }
         num_cell_faces = face_ptr[num_cells];
      } else if (k == "face-areas") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         int rows = 0, n = 0;
         PAMPA_CHECK(input::read(rows, num_cells, num_cells, line[1]), "wrong number of cell faces");
         PAMPA_CHECK(input::read(n, num_cell_faces, num_cell_faces, line[2]), "wrong number of face areas");
         PAMPA_CHECK(read_ragged_d(ptr, farea, num_cells, num_cell_faces, 0.0, DBL_MAX, in),
                     "wrong face-area data");
         PAMPA_CHECK(check_rows(ptr, face_ptr), "wrong face-area data");
      } else if (k == "face-centroids" || k == "face-normals") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         int rows = 0, n = 0;
         PAMPA_CHECK(input::read(rows, num_cells, num_cells, line[1]), "wrong number of cell faces");
         PAMPA_CHECK(input::read(n, 3 * num_cell_faces, 3 * num_cell_faces, line[2]),
                     "wrong number of face coordinates");
         PAMPA_CHECK(input::read_rows(k == "face-centroids" ? fcen : fnorm, num_cell_faces, 3, -DBL_MAX, DBL_MAX, in),
                     "wrong " + k + " data");
      } else if (k == "face-neighbors") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         int rows = 0, n = 0;
         PAMPA_CHECK(input::read(rows, num_cells, num_cells, line[1]), "wrong number of cell faces");
         PAMPA_CHECK(input::read(n, num_cell_faces, num_cell_faces, line[2]), "wrong number of face neighbors");
         PAMPA_CHECK(input::read_ragged(ptr, fnb, num_cells, num_cell_faces, -INT_MAX, INT_MAX, in),
                     "wrong face-neighbor data");
         PAMPA_CHECK(check_rows(ptr, face_ptr), "wrong face-neighbor data");
      } else if (k == "boundary") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         boundaries.push_back(line[1]);
      } else if (k == "bc") {
         PAMPA_CHECK(readBC(line, in), "wrong boundary condition");
      } else {
         PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
      }
   }
   if (bcs.size() < 1 + boundaries.size()) bcs.resize(1 + boundaries.size());
   return 0;
}

int PartitionedMesh::build() {
   // the data is used as read (PartitionedMesh.hxx: nothing to build); this
   // build runs one domain, so no face may lead to a ghost cell
   PAMPA_CHECK(num_cells <= 0 || face_ptr.size() != (size_t)num_cells + 1, "missing cell faces");
   const size_t F = face_ptr[num_cells];
   PAMPA_CHECK(vol.size() != (size_t)num_cells || cen.size() < 3 * (size_t)num_cells ||
                   mat.size() < (size_t)num_cells,
               "missing cell data");
   PAMPA_CHECK(farea.size() != F || fcen.size() != 3 * F || fnorm.size() != 3 * F || fnb.size() != F,
               "missing face data");
   for (size_t f = 0; f < F; f++) {
      PAMPA_CHECK(fnb[f] >= num_cells, "a face leads to a ghost cell: partitioned meshes of several domains "
                                       "need a run with one process per domain");
      PAMPA_CHECK(fnb[f] < 0 && -fnb[f] - 1 >= (int)boundaries.size(), "wrong boundary index of a face");
   }
   cen.resize(3 * (size_t)num_cells);
   mat.resize(num_cells);
   if (cpt_ptr.empty()) cpt_ptr.assign(num_cells + 1, 0);
   return 0;
}

// ---- material edits before the first solver (Parser.cxx:107-139) ------------------------

int split_materials(Mesh& mesh, std::vector<Material>& materials) {
   std::vector<int> subs(materials.size(), 0), stack;
   for (int i = 0; i < mesh.num_cells; i++) {
      const int im = mesh.mat[i];
      if (!materials[im].split) continue;
      // the connected region of material im around cell i gets a new material
      const int im2 = (int)materials.size();
      stack.assign(1, i);
      mesh.mat[i] = im2;
      while (!stack.empty()) {
         const int c = stack.back();
         stack.pop_back();
         for (int f = mesh.face_ptr[c]; f < mesh.face_ptr[c + 1]; f++) {
            const int c2 = mesh.fnb[f];
            if (c2 >= 0 && mesh.mat[c2] == im) {
               mesh.mat[c2] = im2;
               stack.push_back(c2);
            }
         }
      }
      Material sub = materials[im];
      sub.name = materials[im].name + "-" + std::to_string(subs[im]++);
      sub.split = false;
      // a boundary-condition material's regions become boundaries of their own
      if (sub.bc) mesh.addBoundary(sub.name);
      materials.push_back(sub);
      subs.push_back(0);
   }
   return 0;
}

int remove_bc_materials(Mesh& mesh, const std::vector<Material>& materials) {
   PAMPA_CHECK(mesh.partitioned, "unable to remove materials from a partitioned mesh");
   const int C = mesh.num_cells;
   std::vector<int> to_new(C, -1);
   int nc = 0;
   for (int i = 0; i < C; i++)
      if (!materials[mesh.mat[i]].bc) to_new[i] = nc++;
   std::vector<int> mat_bc(materials.size(), -1);
   for (size_t m = 0; m < materials.size(); m++)
      if (materials[m].bc) {
         mat_bc[m] = mesh.findBoundary(materials[m].name);
         PAMPA_CHECK(mat_bc[m] < 0, "wrong boundary name");
      }
   // points used by the remaining cells
   std::vector<int> used(mesh.num_points, 0), pnew(mesh.num_points, -1);
   const bool has_pts = mesh.cpt_ptr.size() == (size_t)C + 1;
   if (has_pts)
      for (int i = 0; i < C; i++)
         if (to_new[i] >= 0)
            for (int p = mesh.cpt_ptr[i]; p < mesh.cpt_ptr[i + 1]; p++) used[mesh.cpts[p]] = 1;
   std::vector<double> points;
   int np = 0;
   for (int p = 0; p < mesh.num_points; p++)
      if (used[p]) {
         pnew[p] = np++;
         points.insert(points.end(), mesh.points.begin() + 3L * p, mesh.points.begin() + 3L * p + 3);
      }
   std::vector<int> cpt_ptr(1, 0), cpts, fptr(1, 0), fnb, mat, nodal;
   std::vector<double> vol, cen, farea, fcen, fnorm;
   for (int i = 0; i < C; i++) {
      if (to_new[i] < 0) continue;
      if (has_pts) {
         for (int p = mesh.cpt_ptr[i]; p < mesh.cpt_ptr[i + 1]; p++) cpts.push_back(pnew[mesh.cpts[p]]);
         cpt_ptr.push_back((int)cpts.size());
      }
      vol.push_back(mesh.vol[i]);
      cen.insert(cen.end(), mesh.cen.begin() + 3L * i, mesh.cen.begin() + 3L * i + 3);
      mat.push_back(mesh.mat[i]);
      if (!mesh.nodal.empty()) nodal.push_back(mesh.nodal[i]);
      for (int f = mesh.face_ptr[i]; f < mesh.face_ptr[i + 1]; f++) {
         farea.push_back(mesh.farea[f]);
         fcen.insert(fcen.end(), mesh.fcen.begin() + 3L * f, mesh.fcen.begin() + 3L * f + 3);
         fnorm.insert(fnorm.end(), mesh.fnorm.begin() + 3L * f, mesh.fnorm.begin() + 3L * f + 3);
         const int i2 = mesh.fnb[f];
         if (i2 >= 0)
            fnb.push_back(to_new[i2] >= 0 ? to_new[i2] : -mat_bc[mesh.mat[i2]] - 1);
         else
            fnb.push_back(i2);
      }
      fptr.push_back((int)fnb.size());
   }
   mesh.num_cells = nc;
   mesh.num_points = np;
   mesh.points.swap(points);
   if (has_pts) {
      mesh.cpt_ptr.swap(cpt_ptr);
      mesh.cpts.swap(cpts);
   }
   mesh.vol.swap(vol);
   mesh.cen.swap(cen);
   mesh.mat.swap(mat);
   mesh.nodal.swap(nodal);
   mesh.face_ptr.swap(fptr);
   mesh.farea.swap(farea);
   mesh.fcen.swap(fcen);
   mesh.fnorm.swap(fnorm);
   mesh.fnb.swap(fnb);
   mesh.num_faces_max = 0;
   for (int i = 0; i < nc; i++) mesh.num_faces_max = std::max(mesh.num_faces_max, mesh.face_ptr[i + 1] - mesh.face_ptr[i]);
   return 0;
}

}  // namespace pampa
