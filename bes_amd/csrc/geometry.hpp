// Host-side mesh topology and metrics for the BES solver core.
//
// Builds the reference's face numbering and metric definitions from an
// in-memory mesh (the element lists of an MSH2 file), then derives the
// cell-centred, structure-of-arrays device layout.  Runs once per mesh.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bes {

constexpr int kInternal = 1, kSlipWall = 2, kFreestream = 3;

struct MeshInput {
  int ndim = 0;
  std::vector<double> coords;   // [nnode][ndim]
  std::vector<int64_t> cells;   // [ncell][ndim+1]
  std::vector<int64_t> bfaces;  // [nbface][ndim]
  std::vector<int> btype;       // [nbface]
  int64_t nnode() const { return ndim ? (int64_t)coords.size() / ndim : 0; }
  int64_t ncell() const { return ndim ? (int64_t)cells.size() / (ndim + 1) : 0; }
  int64_t nbface() const { return ndim ? (int64_t)bfaces.size() / ndim : 0; }
};

// Reference-order geometry (faces numbered as Mesh.cpp:236-534 numbers them).
struct Geometry {
  int nd = 0, k = 0, mode = 0;
  int64_t ncell = 0, nface = 0;
  // faces
  std::vector<int64_t> owner, nbr;       // nbr = -1 on boundaries
  std::vector<int> btype;
  std::vector<double> S, n, area;        // area vector, unit normal, area
  std::vector<double> center;
  // cells
  std::vector<double> vol, centroid;     // vol as the reference defines it
  std::vector<int64_t> cface, cadj;      // [ncell][k]
  std::vector<double> w, r;              // IDW weight [ncell][k], centroid->face [ncell][k][nd]
  std::vector<int> rsel;                 // local face whose vector the reconstruction uses
  std::vector<double> sproj;             // projected lengths [ncell][nd]
};

// Build the reference geometry.  Throws std::runtime_error on malformed input.
Geometry build_geometry(const MeshInput &in, int mode);

// Read the reference's MSH2 subset (Mesh.cpp:95-558).  Throws on failure.
MeshInput read_msh2(const std::string &path);

// Locality order of the cells (Morton order of the centroids): perm[new] = old.
std::vector<int64_t> locality_order(const Geometry &g);

}  // namespace bes
