/*
 * rflu_mesh.h -- host-side region builder for the explicit residual/update path.
 *
 * The reference builds its unstructured region on the host once, before the time
 * loop: grid reading (utilities/part/RFLU_ModCENTAUR.F90:776-1315 for CENTAUR ASCII
 * files, RFLU_ConvCENTAUR2ROCFLU :375-669), the face list
 * (modflu/RFLU_ModFaceList.F90:784-1585), geometry (modflu/RFLU_ModGeometry.F90:307-665
 * with libfloflu/FaceVector.F90 and FaceCentroid.F90), the cell-to-cell stencils
 * (modflu/RFLU_ModStencilsCells.F90:750-1182,1207-1416, RFLU_ModStencilsUtils.F90:696-949)
 * and their R-factors (modflu/RFLU_ModWeights.F90:1770-1894).  This library restates
 * those steps for grids of tetrahedra, hexahedra, prisms and pyramids, read from
 * CENTAUR ASCII grid files or generated as jittered boxes, so that the GPU path and
 * its tests have exactly the arrays a t_region would hold.
 *
 * All arrays follow the reference's column-major layout with the component index
 * fastest (f2c(2,nFaces) is int32 f2c[2*f+s]), and cell/face indices are 0-based.
 * The Fortran host passes its own 1-based arrays straight to rocflu_amd.h instead.
 */
#ifndef RFLU_MESH_H
#define RFLU_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rflu_mesh rflu_mesh;
typedef struct rflu_grid rflu_grid;

/* Cell kinds of the box generator, and the kind of a region built from a grid. */
#define RFLU_MESH_HEX  0   /* one hexahedron per box cell                        */
#define RFLU_MESH_TET6 1   /* six tetrahedra per box cell (Kuhn split on 0-7 axis) */
#define RFLU_MESH_GRID 2   /* built by rflu_mesh_build from a grid                */

/* Boundary types (modfloflu/ModParameters.F90:397-441). */
#define RFLU_BC_INFLOW_TOTANG  10
#define RFLU_BC_INFLOW_VELTEMP 11
#define RFLU_BC_OUTFLOW  20
#define RFLU_BC_SLIPWALL 60
#define RFLU_BC_FARFIELD 80
#define RFLU_BC_INJECTION 90
#define RFLU_BC_SYMMETRY 100
#define RFLU_BC_PERIODIC 110
/* The outer faces of symmetry/periodic virtual layers (rflu_mesh_build_sype). */
#define RFLU_BC_VIRTUAL  130

/* Outflow options (modfloflu/ModParameters.F90:501-503). */
#define RFLU_BCOPT_SUPERSONIC 0
#define RFLU_BCOPT_SUBSONIC   1
#define RFLU_BCOPT_MIXED      2

/*
 * Grids: what a grid file holds (vertices, cells per kind, patch face lists).
 *
 * Cells are listed per kind with the reference's local vertex order (hexahedra: the
 * bottom quadrilateral 1-2-3-4 counter-clockwise seen from above, then 5-8 above
 * them; the local faces of modflu/RFLU_ModGrid.F90:92-102 then point out of the cell),
 * and a patch lists its triangles, then its quadrilaterals, with vertex cycles whose
 * right-hand normals point out of the domain.
 */

/* Read a CENTAUR ASCII grid file (.hyb.asc) as the reference's
 * RFLU_ReadGridCENTAURASCII does; empty patches are dropped and the rest renumbered.
 * Returns 0, 7 (cannot open) or 8 (format error, with the line number). */
int rflu_grid_read_centaur(const char* path, rflu_grid** out);

/* Write a grid as a CENTAUR ASCII file with 16- or 24-character coordinate fields
 * (24 characters carry 17 significant digits: every double reads back exactly). */
int rflu_grid_write_centaur(const rflu_grid* g, const char* path, int field_width);

/* The grid of a jittered box (the same nodes, cells and patches rflu_mesh_box uses). */
int rflu_grid_box(int kind, int nx, int ny, int nz,
                  double lx, double ly, double lz,
                  double x0, double y0, double z0,
                  double jitter, uint64_t seed, rflu_grid** out);

/* A grid from caller arrays, 0-based (xyz(3,V), tet2v(4,n), hex2v(8,n), pri2v(6,n),
 * pyr2v(5,n); per patch its code and face counts; btri2v(3,*) and bquad2v(4,*)
 * patch by patch). */
int rflu_grid_create(int n_verts, const double* xyz,
                     int64_t n_tets, const int32_t* tet2v,
                     int64_t n_hexs, const int32_t* hex2v,
                     int64_t n_pris, const int32_t* pri2v,
                     int64_t n_pyrs, const int32_t* pyr2v,
                     int n_patches, const int32_t* patch_type,
                     const int32_t* patch_ntri, const int32_t* patch_nquad,
                     const int32_t* btri2v, const int32_t* bquad2v, rflu_grid** out);

void rflu_grid_destroy(rflu_grid* g);

/* out[0]=nVert, [1..4]=tetrahedra, hexahedra, prisms, pyramids, [5]=patches,
 * [6]=patch triangles, [7]=patch quadrilaterals. */
int rflu_grid_sizes(const rflu_grid* g, int64_t out[8]);

/* Borrow a grid array by name: "xyz" (double), "tet2v", "hex2v", "pri2v", "pyr2v",
 * "patch_type", "patch_ntri", "patch_nquad", "btri2v", "bquad2v" (int32, 0-based). */
int rflu_grid_array(const rflu_grid* g, const char* name,
                    const void** ptr, int64_t* count, int* dtype);

/*
 * Build the region of a grid.  Patch p (0-based, grid order) gets boundary type
 * bc_type[p], outflow option bc_opt[p] and outflow pressure bc_pout[p] (the .bc
 * file's values; the grid file's own patch codes are not used).  Faces are created in
 * cell order (tetrahedra, hexahedra, prisms, pyramids) and local face order; each
 * keeps the local vertex cycle of the cell that created it.  Boundary faces keep the
 * grid's cycles; a cycle pointing into its cell, a cell face on no patch, a patch
 * face that is not a cell face, a face of three cells or a non-positive volume is an
 * error.  build_stencils as in rflu_mesh_box.
 */
int rflu_mesh_build(const rflu_grid* g, const int32_t* bc_type, const int32_t* bc_opt,
                    const double* bc_pout, int build_stencils, rflu_mesh** out);

/*
 * The region of a grid with symmetry (100) and periodic (110) patches, made as the
 * reference's partitioner makes it (utilities/part/rflupart.F90:361-372,
 * modflu/RFLU_ModSymmetryPeriodic.F90:148-1421): for each such patch in patch order,
 * the cells touching the source patch's vertices (the patch itself for symmetry, the
 * related patch for periodic) and two further vertex-neighbour layers -- virtual cells
 * of earlier patches included, those of a periodic patch's own related patch not --
 * are copied across the patch (mirrored in its plane, or moved by the period along
 * AXIS) as virtual cells numbered after the actual ones.  related (1-based), axis
 * (1..3) and angle (radians; 0 only) are the .bc file's RELPATCH, AXIS and ANGLE per
 * patch.  The symmetry and periodic patches leave the region (their faces join
 * actual and virtual cells); the outer faces of the virtual layers form a last patch
 * (patch_map -1, patch_bc 60: its fluxes reach virtual cells only, and it is not a
 * wall of the stencils' wall-face lists).  The grid must have cells of one kind.
 * Extra arrays: "virt_src"(Nv) the actual cell each virtual cell mirrors (0-based),
 * "virt_xf"(Nv) its transform, "virt_patch"(Nv) the grid patch whose copy made it, "xforms"(9,nT) the transforms (row-major 3x3: momentum
 * and gradients of a virtual cell are those of its actual cell turned by it), and
 * "patch_map"(P') the grid patch of each region patch.  The virtual cells are the last
 * Nv of nCells (rflu_mesh_sizes out[0] counts them).
 */
int rflu_mesh_build_sype(const rflu_grid* g, const int32_t* bc_type, const int32_t* bc_opt,
                         const double* bc_pout, const int32_t* related, const int32_t* axis,
                         const double* angle, int build_stencils, rflu_mesh** out);

/*
 * Build a box region [x0,x0+lx]x[y0,y0+ly]x[z0,z0+lz] with nx*ny*nz box cells.
 * Interior nodes are displaced by U(-jitter,jitter)*spacing per coordinate, drawn
 * from a splitmix64 stream seeded with `seed`, in node order.
 * The six patches are x-, x+, y-, y+, z-, z+ in that order; patch p gets boundary
 * type bc_type[p], outflow option bc_opt[p] and outflow pressure bc_pout[p].
 * build_stencils = 1 also builds the cell-to-cell stencils and R-factors (needed for
 * second-order runs).  The region is rflu_mesh_build of rflu_grid_box's grid.
 * Returns 0 on success, nonzero on error (see rflu_mesh_last_error).
 */
int rflu_mesh_box(int kind, int nx, int ny, int nz,
                  double lx, double ly, double lz,
                  double x0, double y0, double z0,
                  double jitter, uint64_t seed,
                  const int32_t bc_type[6], const int32_t bc_opt[6],
                  const double bc_pout[6], int build_stencils,
                  rflu_mesh** out);

void rflu_mesh_destroy(rflu_mesh* m);

/*
 * Sizes: out[0]=nCells, [1]=nFaces (interior), [2]=nBFaces (all patches),
 * [3]=nPatches, [4]=nVerts, [5]=total stencil members, [6]=vertices per cell (0 when
 * the cells are of several kinds), [7]=max stencil size, [8]=min stencil size.
 */
int rflu_mesh_sizes(const rflu_mesh* m, int64_t out[9]);

/*
 * Borrow a pointer to one of the mesh arrays by name.  *dtype is 0 for int32 and
 * 1 for double; *count is the number of elements.  Names:
 *   int32:  "f2c"(2,F) "f2v"(4,F) "bf2c"(B) "bf2v"(4,B) "bpatch_ptr"(P+1)
 *           "patch_bc"(P) "patch_opt"(P) "c2v"(CSR) "c2v_ptr"(N+1) "c2cs_ptr"(N+1)
 *           "c2cs_idx"(S)
 *           "c2cs_bptr"(N+1) "c2cs_bface"(2,Sb) "c2f_ptr"(N+1) "c2f_idx"(k*N)
 *   double: "xyz"(3,V) "fn"(4,F) "fc"(3,F) "bfn"(4,B) "bfc"(3,B) "patch_pout"(P)
 *           "vol"(N) "cofg"(3,N) "xyz_moms"(10,N)
 * c2f lists, per cell, the faces of the cell as signed ids: f+1 for interior face f
 * with the cell as c1, -(f+1) as c2, and F+b+1 for boundary face b.
 * c2cs_bface holds, per cell (CSR c2cs_bptr), the wall faces of its stencil as
 * (patch, face within the patch), 0-based, nearest first, at most 9
 * (c2cs(icg)%bFaceMembs).
 */
int rflu_mesh_array(const rflu_mesh* m, const char* name,
                    const void** ptr, int64_t* count, int* dtype);

const char* rflu_mesh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
