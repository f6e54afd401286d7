// Shared state of the host-side region builder (mesh.cpp) and the grid and input file
// readers (grid_io.cpp, inputs.cpp).  Not part of the C-ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

// A grid as a grid file holds it (the reference's t_grid after RFLU_ConvCENTAUR2ROCFLU,
// utilities/part/RFLU_ModCENTAUR.F90:375-669): vertex coordinates, cell-to-vertex lists
// per cell kind in the reference's vertex order (tetrahedra, hexahedra, prisms,
// pyramids; modflu/RFLU_ModGrid.F90:92-102 gives the local faces), and per patch its
// boundary triangles and then quadrilaterals with their vertex cycles.  Indices are
// 0-based.
struct rflu_grid {
  std::string title;
  int n_verts = 0;
  std::vector<double> xyz;                            // (3, V)
  std::vector<int32_t> tet2v, hex2v, pri2v, pyr2v;    // (4|8|6|5, n)
  std::vector<int32_t> patch_type;                    // the grid file's patch code
  std::vector<int32_t> patch_ntri, patch_nquad;       // faces per patch
  std::vector<std::string> patch_name;
  std::vector<int32_t> btri2v, bquad2v;               // (3|4, n), patch by patch
};

struct rflu_mesh {
  int kind = 0, nx = 0, ny = 0, nz = 0;
  int nvpc = 8;  // vertices per cell when every cell has the same kind, else 0
  int n_verts = 0, n_cells = 0, n_faces = 0, n_bfaces = 0, n_patches = 0;
  int n_tets = 0, n_hexs = 0, n_pris = 0, n_pyrs = 0;
  std::vector<double> xyz;
  std::vector<int32_t> c2v_ptr, c2v;   // CSR cell -> vertices, cells tets first
  std::vector<int32_t> f2c, f2v;
  std::vector<double> fn, fc;
  std::vector<int32_t> bpatch_ptr, bf2c, bf2v;
  std::vector<double> bfn, bfc;
  std::vector<int32_t> patch_bc, patch_opt;
  std::vector<double> patch_pout;
  std::vector<double> vol, cofg;
  std::vector<int32_t> c2cs_ptr, c2cs_idx;
  std::vector<int32_t> c2cs_bptr, c2cs_bface;   // wall faces of each stencil: (patch, face)
  std::vector<double> xyz_moms;
  std::vector<int32_t> c2f_ptr, c2f_idx;
  int stencil_min = 0, stencil_max = 0;
  // symmetry/periodic virtual cells [n_cells - virt_src.size(), n_cells): the actual
  // cell each mirrors, the index of the transform of its data (xforms, 9 per transform,
  // row-major), the grid patch whose copy made it, and per region patch the grid patch
  // it came from (-1: the outer faces of the virtual layers)
  std::vector<int32_t> virt_src, virt_xf, virt_patch, patch_map;
  std::vector<double> xforms;
};

namespace rflu {

// Local faces of the four cell kinds, 0-based, in the reference's order and vertex
// cycles (f2vTet, f2vHex, f2vPri, f2vPyr of modflu/RFLU_ModGrid.F90:92-102): a cycle's
// right-hand normal (libfloflu/FaceVector.F90) points out of the cell.  -1 ends a
// triangle.
inline constexpr int LF_TET[4][4] = {{0, 1, 2, -1}, {1, 3, 2, -1}, {0, 2, 3, -1}, {0, 3, 1, -1}};
inline constexpr int LF_HEX[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                     {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
inline constexpr int LF_PRI[5][4] = {{0, 2, 1, -1}, {0, 1, 4, 3}, {1, 2, 5, 4},
                                     {0, 3, 5, 2}, {3, 4, 5, -1}};
inline constexpr int LF_PYR[5][4] = {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1},
                                     {2, 3, 4, -1}, {0, 4, 3, -1}};
struct CellKind { int nv, nf; const int (*lf)[4]; };
inline constexpr CellKind KINDS[4] = {{4, 4, LF_TET}, {8, 6, LF_HEX}, {6, 5, LF_PRI},
                                      {5, 5, LF_PYR}};

// Error text of the calling thread's last failed call.
void set_error(const std::string& msg);

// Builds a region from a grid: face list, geometry, cell-to-face lists and, when
// want_stencils is set, the cell stencils with their R-factors and wall-face lists.
// Returns 0 or an error code (the message through set_error).
int build_region(rflu_mesh& m, const rflu_grid& g, const int32_t* bc_type,
                 const int32_t* bc_opt, const double* bc_pout, int want_stencils);

}  // namespace rflu
