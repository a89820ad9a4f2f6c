// Single-domain finite-volume meshes of the drop-in (Mesh.hxx:12-156,
// CartesianMesh.*, UnstructuredExtrudedMesh.*): cells in structure-of-arrays
// form and faces in a ragged layout, face f of cell i at face_ptr[i] + f.
#pragma once

#include "common.hpp"

namespace pampa {

class Mesh {
  public:
   virtual ~Mesh() {}
   virtual int read(const std::string& filename, const std::string& base_dir) = 0;
   virtual int build() = 0;
   int writeVTK(const std::string& prefix, int n) const;
   // Upload to the device (pk_mesh).
   int toDevice(pk_mesh* m) const;

   int num_dims = 0, num_points = 0, num_cells = 0, num_faces_max = 0;
   std::vector<double> points;               // [num_points][3]
   std::vector<int> cpt_ptr, cpts;           // cell points (ragged), VTK only
   std::vector<double> vol, cen;             // [C], [C][3]
   std::vector<int> mat;                     // 0-based material index
   std::vector<int> nodal;                   // nodal indices (optional)
   std::vector<int> face_ptr;                // [C+1]
   std::vector<double> farea, fcen, fnorm;   // [F], [F][3], [F][3]
   std::vector<int> fnb;                     // >= 0 cell, < 0 -(boundary+1)
   std::vector<std::string> boundaries;
   std::vector<BoundaryCondition> bcs;       // indexed by boundary + 1

   bool partitioned = false;                 // read from a partitioned-mesh data file
   // a boundary named after a boundary-condition material (Parser.cxx:93)
   void addBoundary(const std::string& name);

   int findBoundary(const std::string& name) const {
      for (size_t b = 0; b < boundaries.size(); b++)
         if (boundaries[b] == name) return (int)b;
      return -1;
   }

  protected:
   int readBC(const std::vector<std::string>& line, std::istream& in);
};

class CartesianMesh : public Mesh {
  public:
   int read(const std::string& filename, const std::string& base_dir) override;
   int build() override;

  private:
   int nx = 1, ny = 0, nz = 0;
   std::vector<double> dx{1.0}, dy{0.0}, dz{0.0};
   std::vector<int> mat_in;  // 0-based, -1 = void
};

class UnstructuredExtrudedMesh : public Mesh {
  public:
   int read(const std::string& filename, const std::string& base_dir) override;
   int build() override;

  private:
   int nxy_points = 0, nxy_cells = 0, nz = 0;
   std::vector<double> xy;
   std::vector<int> xy_ptr, xy_pts;
   std::vector<double> dz{0.0};
   std::vector<std::string> xy_bnames;
   std::vector<std::vector<int>> xy_bpoints;
   int default_boundary = -1;
   std::vector<int> mat_in;
};

// Mesh data of one domain in the plain-text format of PartitionedMesh.cxx:4-263
// (points, cells with ghost and global counts, cell points, volumes,
// centroids, materials, nodal and global indices, faces, face areas,
// centroids, normals and neighbours, boundaries and conditions).
class PartitionedMesh : public Mesh {
  public:
   PartitionedMesh() { partitioned = true; }
   int read(const std::string& filename, const std::string& base_dir) override;
   int build() override;

  private:
   int num_ghost = 0, num_cells_global = 0;
   std::vector<int> global;
};

struct Material;
// Each connected region of a split material becomes a material of its own,
// named <material>-<k> (Mesh.cxx:9-60).
int split_materials(Mesh& mesh, std::vector<Material>& materials);
// Cells of boundary-condition materials removed; their faces become faces of
// the boundary named after the material (Mesh.cxx:63-204).
int remove_bc_materials(Mesh& mesh, const std::vector<Material>& materials);

}  // namespace pampa
