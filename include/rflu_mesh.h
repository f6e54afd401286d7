/*
 * rflu_mesh.h -- host-side region builder for the explicit residual/update path.
 *
 * The reference builds its unstructured region on the host once, before the time
 * loop: grid generation/reading (utilities/part, tests/mesh_generator), the face list
 * (modflu/RFLU_ModFaceList.F90:784-1585), geometry (modflu/RFLU_ModGeometry.F90:307-665
 * with libfloflu/FaceVector.F90 and FaceCentroid.F90), the cell-to-cell stencils
 * (modflu/RFLU_ModStencilsCells.F90:750-1182,1207-1416, RFLU_ModStencilsUtils.F90:696-949)
 * and their R-factors (modflu/RFLU_ModWeights.F90:1770-1894).  This library restates
 * those steps for jittered box meshes of hexahedra or of Kuhn-split tetrahedra, so that
 * the GPU path and its tests have exactly the arrays a t_region would hold.
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

/* Cell kinds of the box generator. */
#define RFLU_MESH_HEX  0   /* one hexahedron per box cell                        */
#define RFLU_MESH_TET6 1   /* six tetrahedra per box cell (Kuhn split on 0-7 axis) */

/* Boundary types (modfloflu/ModParameters.F90:397-438). */
#define RFLU_BC_INFLOW_TOTANG  10
#define RFLU_BC_INFLOW_VELTEMP 11
#define RFLU_BC_OUTFLOW  20
#define RFLU_BC_SLIPWALL 60
#define RFLU_BC_FARFIELD 80
#define RFLU_BC_INJECTION 90

/* Outflow options (modfloflu/ModParameters.F90:501-503). */
#define RFLU_BCOPT_SUPERSONIC 0
#define RFLU_BCOPT_SUBSONIC   1
#define RFLU_BCOPT_MIXED      2

/*
 * Build a box region [x0,x0+lx]x[y0,y0+ly]x[z0,z0+lz] with nx*ny*nz box cells.
 * Interior nodes are displaced by U(-jitter,jitter)*spacing per coordinate, drawn
 * from a splitmix64 stream seeded with `seed`, in node order.
 * The six patches are x-, x+, y-, y+, z-, z+ in that order; patch p gets boundary
 * type bc_type[p], outflow option bc_opt[p] and outflow pressure bc_pout[p].
 * build_stencils = 1 also builds the cell-to-cell stencils and R-factors (needed for
 * second-order runs).  Returns 0 on success, nonzero on error (see rflu_mesh_last_error).
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
 * [3]=nPatches, [4]=nVerts, [5]=total stencil members, [6]=vertices per cell,
 * [7]=max stencil size, [8]=min stencil size.
 */
int rflu_mesh_sizes(const rflu_mesh* m, int64_t out[9]);

/*
 * Borrow a pointer to one of the mesh arrays by name.  *dtype is 0 for int32 and
 * 1 for double; *count is the number of elements.  Names:
 *   int32:  "f2c"(2,F) "f2v"(4,F) "bf2c"(B) "bf2v"(4,B) "bpatch_ptr"(P+1)
 *           "patch_bc"(P) "patch_opt"(P) "c2v"(nv,N) "c2cs_ptr"(N+1) "c2cs_idx"(S)
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
