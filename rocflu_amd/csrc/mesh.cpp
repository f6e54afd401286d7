// Host-side region builder: face list, geometry, cell-to-cell stencils and their
// R-factors of a grid (read from a grid file or generated as a jittered box of
// hexahedra or Kuhn tetrahedra).  See include/rflu_mesh.h.
//
// Everything here runs once per region on the host, as the reference's
// RFLU_InitFlowSolver does (rocflu/RFLU_InitFlowSolver.F90:495-946).
#include "rflu_mesh.h"
#include "mesh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // uniform in [-1, 1)
  double sym() { return (double)(next() >> 11) * (2.0 / 9007199254740992.0) - 1.0; }
};

}  // namespace

void rflu::set_error(const std::string& msg) { g_err = msg; }

namespace {

// Face vector of a triangle (libfloflu/FaceVector.F90:79-104).
void face_vector_tria(const double* p1, const double* p2, const double* p3, double* v) {
  double dxa = p1[0] - p2[0], dya = p1[1] - p2[1], dza = p1[2] - p2[2];
  double dxb = p1[0] - p3[0], dyb = p1[1] - p3[1], dzb = p1[2] - p3[2];
  v[0] = 0.5 * (dya * dzb - dyb * dza);
  v[1] = -0.5 * (dxa * dzb - dxb * dza);
  v[2] = 0.5 * (dxa * dyb - dxb * dya);
}

// Face vector of a quadrilateral from its diagonals (libfloflu/FaceVector.F90:110-135).
void face_vector_quad(const double* p1, const double* p2, const double* p3, const double* p4,
                      double* v) {
  double dxa = p3[0] - p1[0], dya = p3[1] - p1[1], dza = p3[2] - p1[2];
  double dxb = p2[0] - p4[0], dyb = p2[1] - p4[1], dzb = p2[2] - p4[2];
  v[0] = 0.5 * (dza * dyb - dya * dzb);
  v[1] = 0.5 * (dxa * dzb - dza * dxb);
  v[2] = 0.5 * (dya * dxb - dxa * dyb);
}

// Face centroids (libfloflu/FaceCentroid.F90:77-121).
void face_centroid_tria(const double* p1, const double* p2, const double* p3, double* c) {
  const double thrd = 1.0 / 3.0;
  for (int d = 0; d < 3; ++d) c[d] = thrd * (p1[d] + p2[d] + p3[d]);
}
void face_centroid_quad(const double* p1, const double* p2, const double* p3, const double* p4,
                        double* c) {
  for (int d = 0; d < 3; ++d) c[d] = 0.25 * (p1[d] + p2[d] + p3[d] + p4[d]);
}

void face_geom(const rflu_mesh& m, const int32_t* fv, double* vec, double* cen) {
  const double* p1 = &m.xyz[3 * fv[0]];
  const double* p2 = &m.xyz[3 * fv[1]];
  const double* p3 = &m.xyz[3 * fv[2]];
  if (fv[3] < 0) {
    face_vector_tria(p1, p2, p3, vec);
    face_centroid_tria(p1, p2, p3, cen);
  } else {
    const double* p4 = &m.xyz[3 * fv[3]];
    face_vector_quad(p1, p2, p3, p4, vec);
    face_centroid_quad(p1, p2, p3, p4, cen);
  }
}

using rflu::CellKind;
using rflu::KINDS;

struct FaceKey {
  int32_t v[4];
  bool operator==(const FaceKey& o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
};
struct FaceKeyHash {
  size_t operator()(const FaceKey& k) const {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < 4; ++i) {
      h ^= (uint64_t)(uint32_t)k.v[i];
      h *= 1099511628211ULL;
    }
    return (size_t)h;
  }
};

// Singular-value ratio test of the stencil's design matrix, the acceptance rule of
// RFLU_InvertMatrixSVD (libflu/RFLU_InvertMatrixSVD.F90:150-189): singular when
// s1 <= 100*eps or some s(i) <= 0.01*s(i-1).  The singular values come from a cyclic
// Jacobi eigen-solve of the 4x4 normal matrix of the column-scaled rows
// (1/d, dx/d, dy/d, dz/d) (modflu/RFLU_ModStencilsCells.F90:863-900).
bool stencil_singular(const rflu_mesh& m, int icg, const std::vector<int32_t>& memb, int degr) {
  std::vector<double> a((size_t)degr * 4);
  for (int isl = 0; isl < degr; ++isl) {
    int c2 = memb[isl];
    double dx = m.cofg[3 * c2] - m.cofg[3 * icg];
    double dy = m.cofg[3 * c2 + 1] - m.cofg[3 * icg + 1];
    double dz = m.cofg[3 * c2 + 2] - m.cofg[3 * icg + 2];
    double term = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
    a[4 * isl + 0] = term;
    a[4 * isl + 1] = term * dx;
    a[4 * isl + 2] = term * dy;
    a[4 * isl + 3] = term * dz;
  }
  for (int c = 0; c < 4; ++c) {
    double cm = 0.0;
    for (int r = 0; r < degr; ++r) cm = std::max(cm, std::fabs(a[4 * r + c]));
    for (int r = 0; r < degr; ++r) a[4 * r + c] /= cm;
  }
  double g[4][4] = {};
  for (int r = 0; r < degr; ++r)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) g[i][j] += a[4 * r + i] * a[4 * r + j];
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < 4; ++i)
      for (int j = i + 1; j < 4; ++j) off += g[i][j] * g[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) {
        if (std::fabs(g[p][q]) < 1e-300) continue;
        double theta = (g[q][q] - g[p][p]) / (2.0 * g[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {
          double gkp = g[k][p], gkq = g[k][q];
          g[k][p] = c * gkp - s * gkq;
          g[k][q] = s * gkp + c * gkq;
        }
        for (int k = 0; k < 4; ++k) {
          double gpk = g[p][k], gqk = g[q][k];
          g[p][k] = c * gpk - s * gqk;
          g[q][k] = s * gpk + c * gqk;
        }
      }
  }
  double sv[4];
  for (int i = 0; i < 4; ++i) sv[i] = std::sqrt(std::max(g[i][i], 0.0));
  std::sort(sv, sv + 4, [](double x, double y) { return x > y; });
  if (!(sv[0] > 100.0 * 2.220446049250313e-16)) return true;
  for (int i = 1; i < 4; ++i)
    if (!(sv[i] > 0.01 * sv[i - 1])) return true;
  return false;
}

// Sorts a ascending carrying b along, by the iterative partition-exchange sort of
// QuickSortRFREALInteger (modfloflu/ModSortSearch.F90): middle-element pivot, the
// larger partition pushed on an explicit stack, the smaller one processed next.
// Equal keys end in the same relative order as in the reference.
void rank_sort(std::vector<double>& a, std::vector<int32_t>& b) {
  const int n = (int)a.size();
  if (n < 2) return;
  int stl[64], str[64];
  int s = 0;
  stl[0] = 1;
  str[0] = n;
  // 1-based indices as in the reference: element k is a[k-1]
  auto A = [&](int k) -> double& { return a[k - 1]; };
  auto Bv = [&](int k) -> int32_t& { return b[k - 1]; };
  s = 1;
  while (s != 0) {
    int l = stl[s - 1], r = str[s - 1];
    --s;
    for (;;) {
      int i = l, j = r;
      int k = (l + r) / 2;
      double x = A(k);
      for (;;) {
        while (A(i) < x) ++i;
        while (x < A(j)) --j;
        if (i <= j) {
          std::swap(A(i), A(j));
          std::swap(Bv(i), Bv(j));
          ++i;
          --j;
          if (i > j) break;
        } else {
          break;
        }
      }
      if (j - l >= r - i) {
        if (l < j) { stl[s] = l; str[s] = j; ++s; }
        l = i;
      } else {
        if (i < r) { stl[s] = i; str[s] = r; ++s; }
        r = j;
      }
      if (!(l < r)) break;
    }
  }
}

void build_stencils(rflu_mesh& m) {
  const int N = m.n_cells;
  const int32_t* cp = m.c2v_ptr.data();
  const int32_t* cvx = m.c2v.data();
  // vertex -> cells, ascending cell order
  std::vector<int32_t> v2c_ptr(m.n_verts + 1, 0), v2c;
  for (int c = 0; c < N; ++c)
    for (int k = cp[c]; k < cp[c + 1]; ++k) v2c_ptr[cvx[k] + 1]++;
  for (int v = 0; v < m.n_verts; ++v) v2c_ptr[v + 1] += v2c_ptr[v];
  v2c.resize(v2c_ptr[m.n_verts]);
  {
    std::vector<int32_t> fill(v2c_ptr.begin(), v2c_ptr.end() - 1);
    for (int c = 0; c < N; ++c)
      for (int k = cp[c]; k < cp[c + 1]; ++k) v2c[fill[cvx[k]]++] = c;
  }
  // Stencil size rules: RFLU_SetInfoC2CStencil calls the four-argument
  // RFLU_ComputeStencilSize(global, dimens=3, factor=1, order=1)
  // (modflu/RFLU_ModStencilsCells.F90:2870-2873, RFLU_ModStencilsUtils.F90:1093-1130),
  // so min = (1+1)(1+2)(1+3)/6 = 4, max = 10*min = 40, at most 6 layers.
  const int smin = 4, smax = 40, nlayers_max = 6;
  m.stencil_min = smin;
  m.stencil_max = smax;
  m.c2cs_ptr.assign(N + 1, 0);
  m.c2cs_idx.clear();
  m.xyz_moms.assign((size_t)10 * N, 0.0);

  std::vector<std::vector<int32_t>> lists(N);
#pragma omp parallel
  {
  std::vector<int32_t> tmp, memb, help;
  std::vector<double> frac;
  std::vector<int32_t> fid;
#pragma omp for schedule(dynamic, 256)
  for (int icg = 0; icg < N; ++icg) {
    // Basic stencil: vertex neighbours, ranked by shared-vertex fraction, best
    // stencilSizeMax kept, then sorted by cell index
    // (modflu/RFLU_ModStencilsCells.F90:1207-1416).
    // the shared-vertex fraction is counted against the cell's own vertex count
    const int nv = cp[icg + 1] - cp[icg];
    tmp.clear();
    for (int k = 0; k < nv; ++k) {
      int v = cvx[cp[icg] + k];
      for (int j = v2c_ptr[v]; j < v2c_ptr[v + 1]; ++j) tmp.push_back(v2c[j]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    tmp.erase(std::remove(tmp.begin(), tmp.end(), icg), tmp.end());
    frac.clear();
    fid.clear();
    int vs[8];
    for (int k = 0; k < nv; ++k) vs[k] = cvx[cp[icg] + k];
    std::sort(vs, vs + nv);
    for (int c2 : tmp) {
      int vs2[8];
      const int nv2 = cp[c2 + 1] - cp[c2];
      for (int k = 0; k < nv2; ++k) vs2[k] = cvx[cp[c2] + k];
      std::sort(vs2, vs2 + nv2);
      int common = 0;
      for (int a = 0, b = 0; a < nv && b < nv2;) {
        if (vs[a] == vs2[b]) { ++common; ++a; ++b; }
        else if (vs[a] < vs2[b]) ++a;
        else ++b;
      }
      frac.push_back(common * (1.0 / nv));
      fid.push_back(c2);
    }
    // The ranking sort decides which tied candidates survive the cut at smax, so it
    // keeps the reference's partition-exchange order.
    rank_sort(frac, fid);
    memb.clear();
    for (int i = (int)frac.size() - 1; i >= std::max((int)frac.size() - smax, 0); --i)
      memb.push_back(fid[i]);
    std::sort(memb.begin(), memb.end());
    int degr = (int)memb.size();
    int lay_beg = 0, lay_end = degr;  // last layer [beg, end)

    // Extra layers while too small or singular
    // (modflu/RFLU_ModStencilsCells.F90:849-925, RFLU_ModStencilsUtils.F90:696-949).
    for (int layer = 2; layer <= nlayers_max; ++layer) {
      bool singular = false;
      if (degr >= smin) singular = stencil_singular(m, icg, memb, degr);
      if (!(singular || degr < smin)) break;
      std::vector<int32_t> sorted_memb(memb.begin(), memb.end());
      std::sort(sorted_memb.begin(), sorted_memb.end());
      help.clear();
      bool full = false;
      for (int isl = lay_beg; isl < lay_end && !full; ++isl) {
        int c = memb[isl];
        for (int k = cp[c]; k < cp[c + 1] && !full; ++k) {
          int v = cvx[k];
          for (int j = v2c_ptr[v]; j < v2c_ptr[v + 1]; ++j) {
            int c2 = v2c[j];
            if (c2 == icg) continue;
            if (std::binary_search(sorted_memb.begin(), sorted_memb.end(), c2)) continue;
            auto it = std::lower_bound(help.begin(), help.end(), c2);
            if (it != help.end() && *it == c2) continue;
            if ((int)help.size() >= smax) { full = true; break; }
            help.insert(it, c2);
          }
        }
      }
      lay_beg = degr;
      for (int c2 : help)
        if (degr < smax) { memb.push_back(c2); ++degr; }
      lay_end = degr;
    }
    lists[icg] = memb;

    // R-factors of the inverse-distance-weighted normal matrix
    // (modflu/RFLU_ModWeights.F90:1770-1894).
    double s11 = 0, s12 = 0, s22 = 0, s13 = 0, s23 = 0, s33 = 0, s14 = 0, s24 = 0, s34 = 0, s44 = 0;
    for (int isl = 0; isl < degr; ++isl) {
      int c2 = memb[isl];
      double dx = m.cofg[3 * c2] - m.cofg[3 * icg];
      double dy = m.cofg[3 * c2 + 1] - m.cofg[3 * icg + 1];
      double dz = m.cofg[3 * c2 + 2] - m.cofg[3 * icg + 2];
      double wt = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
      dx = wt * dx; dy = wt * dy; dz = wt * dz;
      s11 = s11 + dx * dx; s12 = s12 + dx * dy; s22 = s22 + dy * dy;
      s13 = s13 + dx * dz; s23 = s23 + dy * dz; s33 = s33 + dz * dz;
      s14 = s14 + wt * dx; s24 = s24 + wt * dy; s34 = s34 + wt * dz; s44 = s44 + wt * wt;
    }
    double r11 = std::sqrt(s11);
    double ir11 = 1.0 / r11;
    double r12 = ir11 * s12;
    double r22 = std::sqrt(s22 - r12 * r12);
    double ir22 = 1.0 / r22;
    double r13 = ir11 * s13;
    double r23 = ir22 * (s23 - r12 * r13);
    double r33 = std::sqrt(s33 - (r13 * r13 + r23 * r23));
    double ir33 = 1.0 / r33;
    double r14 = ir11 * s14;
    double r24 = ir22 * (s24 - r12 * r14);
    double r34 = ir33 * (s34 - (r13 * r14 + r23 * r24));
    double r44 = std::sqrt(s44 - (r14 * r14 + r24 * r24 + r34 * r34));
    double* mo = &m.xyz_moms[10 * (size_t)icg];
    mo[0] = r11; mo[1] = r12; mo[2] = r22; mo[3] = r13; mo[4] = r23;
    mo[5] = r33; mo[6] = r14; mo[7] = r24; mo[8] = r34; mo[9] = r44;
  }
  }  // omp parallel
  for (int icg = 0; icg < N; ++icg)
    m.c2cs_ptr[icg + 1] = m.c2cs_ptr[icg] + (int32_t)lists[icg].size();
  m.c2cs_idx.resize(m.c2cs_ptr[N]);
  for (int icg = 0; icg < N; ++icg) {
    std::copy(lists[icg].begin(), lists[icg].end(), m.c2cs_idx.begin() + m.c2cs_ptr[icg]);
    std::vector<int32_t>().swap(lists[icg]);
  }
}

// Wall faces of each cell's stencil (the constrained-gradient rows), as the
// reference builds c2cs(icg)%bFaceMembs: RFLU_AddBFaces
// (modflu/RFLU_ModStencilsUtils.F90:134-298) walks the cell itself and then its
// stencil members; for every wall patch in patch order, the member's face on that
// patch (found through the patch's cell-sorted face list) joins, up to 2*9 faces;
// RFLU_SortBFaces (:1501-1585) ranks them by squared distance from the cell centroid
// with the same partition-exchange sort as the stencil ranking, and
// RFLU_ModStencilsCells.F90:1059-1079 keeps the first nBFaceMembsMax = 9 (:2866).
// Slip walls are the wall type of this builder.
void build_bface_stencils(rflu_mesh& m) {
  const int N = m.n_cells, P = m.n_patches;
  const int nmax = 9, nmax_tmp = 2 * nmax;
  // per wall patch: cell -> face within the patch (-1: none)
  std::vector<std::vector<int32_t>> face_of(P);
  for (int p = 0; p < P; ++p) {
    if (m.patch_bc[p] != RFLU_BC_SLIPWALL) continue;
    face_of[p].assign(N, -1);
    for (int b = m.bpatch_ptr[p]; b < m.bpatch_ptr[p + 1]; ++b)
      face_of[p][m.bf2c[b]] = b - m.bpatch_ptr[p];
  }
  m.c2cs_bptr.assign(N + 1, 0);
  m.c2cs_bface.clear();
  std::vector<std::vector<int32_t>> lists(N);
#pragma omp parallel for schedule(dynamic, 256)
  for (int icg = 0; icg < N; ++icg) {
    std::vector<int32_t> pf;   // (patch, face) pairs
    auto visit = [&](int isg) {
      for (int p = 0; p < P; ++p) {
        if (face_of[p].empty()) continue;
        int f = face_of[p][isg];
        if (f < 0) continue;
        if ((int)pf.size() / 2 >= nmax_tmp) return false;
        pf.push_back(p);
        pf.push_back(f);
      }
      return true;
    };
    bool more = visit(icg);
    for (int s = m.c2cs_ptr[icg]; more && s < m.c2cs_ptr[icg + 1]; ++s) more = visit(m.c2cs_idx[s]);
    const int nb = (int)pf.size() / 2;
    if (nb == 0) continue;
    std::vector<double> dist(nb);
    std::vector<int32_t> key(nb);
    const double* xc = &m.cofg[3 * (size_t)icg];
    for (int k = 0; k < nb; ++k) {
      const double* fcp = &m.bfc[3 * (size_t)(m.bpatch_ptr[pf[2 * k]] + pf[2 * k + 1])];
      double dx = fcp[0] - xc[0], dy = fcp[1] - xc[1], dz = fcp[2] - xc[2];
      dist[k] = dx * dx + dy * dy + dz * dz;
      key[k] = k;
    }
    rank_sort(dist, key);
    const int keep = std::min(nb, nmax);
    lists[icg].resize(2 * (size_t)keep);
    for (int k = 0; k < keep; ++k) {
      lists[icg][2 * k] = pf[2 * key[k]];
      lists[icg][2 * k + 1] = pf[2 * key[k] + 1];
    }
  }
  for (int icg = 0; icg < N; ++icg)
    m.c2cs_bptr[icg + 1] = m.c2cs_bptr[icg] + (int32_t)lists[icg].size() / 2;
  m.c2cs_bface.reserve(2 * (size_t)m.c2cs_bptr[N]);
  for (int icg = 0; icg < N; ++icg)
    m.c2cs_bface.insert(m.c2cs_bface.end(), lists[icg].begin(), lists[icg].end());
}

// A jittered box as a grid: nodes in i-fastest order, interior nodes displaced by
// U(-jitter, jitter) x spacing per coordinate from a splitmix64 stream; one hexahedron
// or six Kuhn tetrahedra (axis order a, b, c: corner 0 -> e_a -> e_a+e_b -> 7) per box
// cell, with their vertices ordered as the reference's local faces expect (the right-
// hand normal of every local face points out of the cell on the undisplaced nodes);
// patches x-, x+, y-, y+, z-, z+ holding the cell faces on each side, in cell order.
int grid_box(rflu_grid& g, int kind, int nx, int ny, int nz, double lx, double ly, double lz,
             double x0, double y0, double z0, double jitter, uint64_t seed) {
  if (nx < 1 || ny < 1 || nz < 1) { g_err = "box dimensions must be >= 1"; return 1; }
  if (kind != RFLU_MESH_HEX && kind != RFLU_MESH_TET6) { g_err = "unknown cell kind"; return 1; }
  if (!(jitter >= 0.0 && jitter < 0.25)) { g_err = "jitter must be in [0, 0.25)"; return 1; }
  const int64_t ncell64 = (int64_t)nx * ny * nz * (kind == RFLU_MESH_TET6 ? 6 : 1);
  if (ncell64 > 400000000LL) { g_err = "mesh too large"; return 1; }
  const int NX = nx + 1, NY = ny + 1, NZ = nz + 1;
  g.n_verts = NX * NY * NZ;
  const double dx = lx / nx, dy = ly / ny, dz = lz / nz;

  std::vector<double> xyz0((size_t)3 * g.n_verts);
  g.xyz.resize((size_t)3 * g.n_verts);
  SplitMix64 rng(seed);
  for (int k = 0; k < NZ; ++k)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        size_t v = (size_t)i + (size_t)NX * (j + (size_t)NY * k);
        double px = x0 + i * dx, py = y0 + j * dy, pz = z0 + k * dz;
        if (i == nx) px = x0 + lx;
        if (j == ny) py = y0 + ly;
        if (k == nz) pz = z0 + lz;
        xyz0[3 * v] = px; xyz0[3 * v + 1] = py; xyz0[3 * v + 2] = pz;
        bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
        if (interior && jitter > 0.0) {
          px += jitter * dx * rng.sym();
          py += jitter * dy * rng.sym();
          pz += jitter * dz * rng.sym();
        }
        g.xyz[3 * v] = px; g.xyz[3 * v + 1] = py; g.xyz[3 * v + 2] = pz;
      }

  // corner index b = bx + 2*by + 4*bz; the reference's hexahedron is the bottom
  // quadrilateral counter-clockwise seen from +z, then the top one above it
  const int HEX_ORDER[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::vector<int32_t>& c2v = kind == RFLU_MESH_HEX ? g.hex2v : g.tet2v;
  const int nv = kind == RFLU_MESH_HEX ? 8 : 4;
  c2v.resize((size_t)nv * ncell64);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int corner[8];
        for (int b = 0; b < 8; ++b) {
          int bi = b & 1, bj = (b >> 1) & 1, bk = (b >> 2) & 1;
          corner[b] = (i + bi) + NX * ((j + bj) + NY * (k + bk));
        }
        size_t box = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
        if (kind == RFLU_MESH_HEX) {
          for (int b = 0; b < 8; ++b) c2v[8 * box + b] = corner[HEX_ORDER[b]];
          continue;
        }
        for (int t = 0; t < 6; ++t) {
          int b1 = 1 << perms[t][0];
          int b2 = b1 | (1 << perms[t][1]);
          int32_t* tv = &c2v[4 * (6 * box + t)];
          tv[0] = corner[0]; tv[1] = corner[b1]; tv[2] = corner[b2]; tv[3] = corner[7];
          // local face (0,1,2) must face away from vertex 3
          const double* p0 = &xyz0[3 * tv[0]];
          double a[3], c[3], e[3];
          for (int d = 0; d < 3; ++d) {
            a[d] = xyz0[3 * tv[1] + d] - p0[d];
            c[d] = xyz0[3 * tv[2] + d] - p0[d];
            e[d] = xyz0[3 * tv[3] + d] - p0[d];
          }
          double det = (a[1] * c[2] - a[2] * c[1]) * e[0] + (a[2] * c[0] - a[0] * c[2]) * e[1] +
                       (a[0] * c[1] - a[1] * c[0]) * e[2];
          if (det > 0) std::swap(tv[1], tv[2]);
        }
      }

  // patch faces: the cell faces whose vertices all lie on one box side
  auto on_side = [&](int v, int s) {
    int i = v % NX, j = (v / NX) % NY, k = v / (NX * NY);
    switch (s) {
      case 0: return i == 0;
      case 1: return i == nx;
      case 2: return j == 0;
      case 3: return j == ny;
      case 4: return k == 0;
      default: return k == nz;
    }
  };
  const CellKind& ck = KINDS[kind == RFLU_MESH_HEX ? 1 : 0];
  std::vector<std::vector<int32_t>> side_faces(6);
  for (int64_t c = 0; c < ncell64; ++c) {
    const int32_t* cv = &c2v[(size_t)nv * c];
    for (int lf = 0; lf < ck.nf; ++lf) {
      const int nvf = ck.lf[lf][3] < 0 ? 3 : 4;
      for (int s = 0; s < 6; ++s) {
        bool all = true;
        for (int a = 0; a < nvf && all; ++a) all = on_side(cv[ck.lf[lf][a]], s);
        if (!all) continue;
        for (int a = 0; a < nvf; ++a) side_faces[s].push_back(cv[ck.lf[lf][a]]);
        break;
      }
    }
  }
  g.patch_type.assign(6, 0);
  g.patch_ntri.assign(6, 0);
  g.patch_nquad.assign(6, 0);
  g.patch_name = {"x-", "x+", "y-", "y+", "z-", "z+"};
  for (int s = 0; s < 6; ++s) {
    std::vector<int32_t>& dst = kind == RFLU_MESH_HEX ? g.bquad2v : g.btri2v;
    dst.insert(dst.end(), side_faces[s].begin(), side_faces[s].end());
    (kind == RFLU_MESH_HEX ? g.patch_nquad : g.patch_ntri)[s] =
        (int32_t)side_faces[s].size() / (kind == RFLU_MESH_HEX ? 4 : 3);
  }
  g.title = "box";
  return 0;
}

}  // namespace

// The region of a grid.  Faces are created in the order cells (tetrahedra, hexahedra,
// prisms, pyramids, as the reference numbers them globally) and their local faces
// are visited; the first visiting cell is c1 and the face keeps that cell's local
// vertex cycle, which points out of c1 (modflu/RFLU_ModFaceList.F90:784-1000).
// Boundary faces are the patches' faces, patch by patch, triangles then
// quadrilaterals, with the grid's vertex cycles and the cell found by matching
// vertex sets (RFLU_ModFaceList.F90:1025-1085).
int rflu::build_region(rflu_mesh& m, const rflu_grid& g, const int32_t* bc_type,
                       const int32_t* bc_opt, const double* bc_pout, int want_stencils) {
  const int V = g.n_verts;
  if (V < 4 || g.xyz.size() != (size_t)3 * V) { g_err = "grid: bad vertex count"; return 1; }
  const std::vector<int32_t>* lists[4] = {&g.tet2v, &g.hex2v, &g.pri2v, &g.pyr2v};
  int64_t ncell64 = 0, nkind[4];
  int kinds_present = 0, only = -1;
  for (int t = 0; t < 4; ++t) {
    if (lists[t]->size() % KINDS[t].nv) { g_err = "grid: ragged cell list"; return 1; }
    nkind[t] = (int64_t)lists[t]->size() / KINDS[t].nv;
    ncell64 += nkind[t];
    if (nkind[t]) { ++kinds_present; only = t; }
  }
  if (ncell64 < 1) { g_err = "grid: no cells"; return 1; }
  if (ncell64 > 400000000LL) { g_err = "mesh too large"; return 1; }
  const int P = (int)g.patch_type.size();
  if (P < 1 || (int)g.patch_ntri.size() != P || (int)g.patch_nquad.size() != P) {
    g_err = "grid: no patches";
    return 1;
  }
  {
    int64_t nt = 0, nq = 0;
    for (int p = 0; p < P; ++p) {
      if (g.patch_ntri[p] < 0 || g.patch_nquad[p] < 0) { g_err = "grid: bad patch"; return 1; }
      nt += g.patch_ntri[p];
      nq += g.patch_nquad[p];
    }
    if ((size_t)(3 * nt) != g.btri2v.size() || (size_t)(4 * nq) != g.bquad2v.size()) {
      g_err = "grid: patch face counts do not match the face lists";
      return 1;
    }
  }
  for (int p = 0; p < P; ++p)
    if (bc_type[p] != RFLU_BC_OUTFLOW && bc_type[p] != RFLU_BC_SLIPWALL &&
        bc_type[p] != RFLU_BC_INFLOW_VELTEMP && bc_type[p] != RFLU_BC_INFLOW_TOTANG &&
        bc_type[p] != RFLU_BC_FARFIELD && bc_type[p] != RFLU_BC_INJECTION &&
        bc_type[p] != RFLU_BC_VIRTUAL) {
      g_err = "patch " + std::to_string(p + 1) + " has boundary type " +
              std::to_string(bc_type[p]) +
              "; built types: inflow (10, 11), outflow (20), slip wall (60), farfield (80), "
              "injection (90); symmetry (100) and periodic (110) through rflu_mesh_build_sype";
      return 4;
    }
  for (const std::vector<int32_t>* l : lists)
    for (int32_t v : *l)
      if (v < 0 || v >= V) { g_err = "grid: cell vertex out of range"; return 1; }
  for (int32_t v : g.btri2v)
    if (v < 0 || v >= V) { g_err = "grid: patch vertex out of range"; return 1; }
  for (int32_t v : g.bquad2v)
    if (v < 0 || v >= V) { g_err = "grid: patch vertex out of range"; return 1; }

  m.n_verts = V;
  m.xyz = g.xyz;
  m.n_cells = (int)ncell64;
  m.n_tets = (int)nkind[0]; m.n_hexs = (int)nkind[1];
  m.n_pris = (int)nkind[2]; m.n_pyrs = (int)nkind[3];
  m.nvpc = kinds_present == 1 ? KINDS[only].nv : 0;
  m.n_patches = P;
  m.c2v_ptr.assign((size_t)m.n_cells + 1, 0);
  m.c2v.clear();
  m.c2v.reserve(g.tet2v.size() + g.hex2v.size() + g.pri2v.size() + g.pyr2v.size());
  std::vector<int8_t> ckind(m.n_cells);
  {
    int c = 0;
    for (int t = 0; t < 4; ++t)
      for (int64_t i = 0; i < nkind[t]; ++i, ++c) {
        ckind[c] = (int8_t)t;
        m.c2v.insert(m.c2v.end(), lists[t]->begin() + KINDS[t].nv * i,
                     lists[t]->begin() + KINDS[t].nv * (i + 1));
        m.c2v_ptr[c + 1] = m.c2v_ptr[c] + KINDS[t].nv;
      }
  }

  struct Pending { int32_t c1, c2, v[4], bface; };
  std::vector<Pending> faces;
  faces.reserve((size_t)m.n_cells * 3 + 16);
  std::unordered_map<FaceKey, int32_t, FaceKeyHash> map;
  map.reserve(faces.capacity() * 2);
  auto make_key = [](const int32_t* fv, int nvf) {
    FaceKey key;
    for (int a = 0; a < 4; ++a) key.v[a] = a < nvf ? fv[a] : -1;
    std::sort(key.v, key.v + nvf);
    return key;
  };
  for (int c = 0; c < m.n_cells; ++c) {
    const CellKind& ck = KINDS[ckind[c]];
    const int32_t* cv = &m.c2v[m.c2v_ptr[c]];
    for (int lf = 0; lf < ck.nf; ++lf) {
      const int nvf = ck.lf[lf][3] < 0 ? 3 : 4;
      int32_t fv[4] = {-1, -1, -1, -1};
      for (int a = 0; a < nvf; ++a) fv[a] = cv[ck.lf[lf][a]];
      FaceKey key = make_key(fv, nvf);
      auto it = map.find(key);
      if (it != map.end()) {
        Pending& p = faces[it->second];
        if (p.c2 >= 0) { g_err = "non-manifold face"; return 2; }
        p.c2 = c;
        map.erase(it);
        continue;
      }
      Pending p;
      p.c1 = c; p.c2 = -1; p.bface = -1;
      std::memcpy(p.v, fv, sizeof(fv));
      map.emplace(key, (int32_t)faces.size());
      faces.push_back(p);
    }
  }
  // boundary faces, patch by patch
  m.bpatch_ptr.assign(P + 1, 0);
  m.bf2c.clear(); m.bf2v.clear();
  {
    size_t it_tri = 0, it_quad = 0;
    for (int p = 0; p < P; ++p) {
      for (int pass = 0; pass < 2; ++pass) {
        const int nvf = pass == 0 ? 3 : 4;
        const int cnt = pass == 0 ? g.patch_ntri[p] : g.patch_nquad[p];
        for (int k = 0; k < cnt; ++k) {
          const int32_t* src = pass == 0 ? &g.btri2v[3 * it_tri++] : &g.bquad2v[4 * it_quad++];
          int32_t fv[4] = {-1, -1, -1, -1};
          for (int a = 0; a < nvf; ++a) fv[a] = src[a];
          auto f = map.find(make_key(fv, nvf));
          if (f == map.end() || faces[f->second].bface >= 0) {
            g_err = "patch " + std::to_string(p + 1) + " face " + std::to_string(k + 1) +
                    (f == map.end() ? " is not a boundary face of any cell"
                                    : " is on the boundary twice");
            return 3;
          }
          faces[f->second].bface = (int32_t)m.bf2c.size();
          m.bf2c.push_back(faces[f->second].c1);
          for (int a = 0; a < 4; ++a) m.bf2v.push_back(fv[a]);
        }
      }
      m.bpatch_ptr[p + 1] = (int32_t)m.bf2c.size();
    }
  }
  if (m.bf2c.size() != map.size()) {
    g_err = "grid: " + std::to_string(map.size() - m.bf2c.size()) +
            " cell faces have one cell and lie on no patch";
    return 3;
  }
  m.f2c.clear(); m.f2v.clear();
  for (const Pending& p : faces) {
    if (p.c2 < 0) continue;
    m.f2c.push_back(p.c1); m.f2c.push_back(p.c2);
    for (int a = 0; a < 4; ++a) m.f2v.push_back(p.v[a]);
  }
  std::vector<Pending>().swap(faces);
  m.n_faces = (int)(m.f2c.size() / 2);
  m.n_bfaces = (int)m.bf2c.size();
  m.patch_bc.assign(bc_type, bc_type + P);
  m.patch_opt.assign(bc_opt, bc_opt + P);
  m.patch_pout.assign(bc_pout, bc_pout + P);

  // Geometry (modflu/RFLU_ModGeometry.F90:307-665): approximate centroids, face
  // vectors/centroids, volumes and centroids from the divergence theorem.
  const int N = m.n_cells, F = m.n_faces, B = m.n_bfaces;
  std::vector<double> app((size_t)3 * N, 0.0);
  for (int c = 0; c < N; ++c) {
    const int nv = m.c2v_ptr[c + 1] - m.c2v_ptr[c];
    const double term = 1.0 / nv;
    double s[3] = {0, 0, 0};
    for (int a = m.c2v_ptr[c]; a < m.c2v_ptr[c + 1]; ++a)
      for (int d = 0; d < 3; ++d) s[d] += m.xyz[3 * m.c2v[a] + d];
    for (int d = 0; d < 3; ++d) app[3 * c + d] = term * s[d];
  }
  m.vol.assign(N, 0.0);
  m.cofg.assign((size_t)3 * N, 0.0);
  m.fn.resize((size_t)4 * F);
  m.fc.resize((size_t)3 * F);
  for (int f = 0; f < F; ++f) {
    double v[3], c[3];
    face_geom(m, &m.f2v[4 * (size_t)f], v, c);
    double vm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double term = 1.0 / vm;
    m.fn[4 * f] = v[0] * term; m.fn[4 * f + 1] = v[1] * term;
    m.fn[4 * f + 2] = v[2] * term; m.fn[4 * f + 3] = vm;
    m.fc[3 * f] = c[0]; m.fc[3 * f + 1] = c[1]; m.fc[3 * f + 2] = c[2];
    int c1 = m.f2c[2 * f], c2 = m.f2c[2 * f + 1];
    double t = (c[0] - app[3 * c1]) * v[0] + (c[1] - app[3 * c1 + 1]) * v[1] +
               (c[2] - app[3 * c1 + 2]) * v[2];
    m.vol[c1] = m.vol[c1] + t;
    t = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
    for (int d = 0; d < 3; ++d) m.cofg[3 * c1 + d] = m.cofg[3 * c1 + d] + t * c[d];
    t = (c[0] - app[3 * c2]) * v[0] + (c[1] - app[3 * c2 + 1]) * v[1] +
        (c[2] - app[3 * c2 + 2]) * v[2];
    m.vol[c2] = m.vol[c2] - t;
    t = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
    for (int d = 0; d < 3; ++d) m.cofg[3 * c2 + d] = m.cofg[3 * c2 + d] - t * c[d];
  }
  m.bfn.resize((size_t)4 * B);
  m.bfc.resize((size_t)3 * B);
  for (int b = 0; b < B; ++b) {
    double v[3], c[3];
    face_geom(m, &m.bf2v[4 * (size_t)b], v, c);
    double vm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double term = 1.0 / vm;
    m.bfn[4 * b] = v[0] * term; m.bfn[4 * b + 1] = v[1] * term;
    m.bfn[4 * b + 2] = v[2] * term; m.bfn[4 * b + 3] = vm;
    m.bfc[3 * b] = c[0]; m.bfc[3 * b + 1] = c[1]; m.bfc[3 * b + 2] = c[2];
    int c1 = m.bf2c[b];
    double t = (c[0] - app[3 * c1]) * v[0] + (c[1] - app[3 * c1 + 1]) * v[1] +
               (c[2] - app[3 * c1 + 2]) * v[2];
    if (!(t > 0.0)) {   // a patch face's vertex cycle must point out of its cell
      int p = (int)(std::upper_bound(m.bpatch_ptr.begin(), m.bpatch_ptr.end(), b) -
                    m.bpatch_ptr.begin()) - 1;
      g_err = "patch " + std::to_string(p + 1) + " face " +
              std::to_string(b - m.bpatch_ptr[p] + 1) + " points into its cell";
      return 5;
    }
    m.vol[c1] = m.vol[c1] + t;
    t = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
    for (int d = 0; d < 3; ++d) m.cofg[3 * c1 + d] = m.cofg[3 * c1 + d] + t * c[d];
  }
  {
    const double THRD = 1.0 / 3.0;
    for (int c = 0; c < N; ++c) {
      m.vol[c] = THRD * m.vol[c];
      double term = 1.0 / (4.0 * m.vol[c]);
      for (int d = 0; d < 3; ++d) m.cofg[3 * c + d] = term * m.cofg[3 * c + d];
      if (!(m.vol[c] > 0.0)) { g_err = "non-positive cell volume"; return 5; }
    }
  }

  // Cell -> face lists in summation order: interior faces ascending, then boundary
  // faces patch by patch (the order of the scatter loops in
  // modflu/RFLU_ModHLLCFlux.F90:2207-2425).
  m.c2f_ptr.assign(N + 1, 0);
  for (int f = 0; f < F; ++f) { m.c2f_ptr[m.f2c[2 * f] + 1]++; m.c2f_ptr[m.f2c[2 * f + 1] + 1]++; }
  for (int b = 0; b < B; ++b) m.c2f_ptr[m.bf2c[b] + 1]++;
  for (int c = 0; c < N; ++c) m.c2f_ptr[c + 1] += m.c2f_ptr[c];
  m.c2f_idx.assign(m.c2f_ptr[N], 0);
  {
    std::vector<int32_t> fill(m.c2f_ptr.begin(), m.c2f_ptr.end() - 1);
    for (int f = 0; f < F; ++f) {
      m.c2f_idx[fill[m.f2c[2 * f]]++] = f + 1;
      m.c2f_idx[fill[m.f2c[2 * f + 1]]++] = -(f + 1);
    }
    for (int b = 0; b < B; ++b) m.c2f_idx[fill[m.bf2c[b]]++] = F + b + 1;
  }

  if (want_stencils) {
    build_stencils(m);
    build_bface_stencils(m);
  }
  return 0;
}

namespace {

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return 6;
  }
}

}  // namespace

extern "C" {

int rflu_grid_box(int kind, int nx, int ny, int nz, double lx, double ly, double lz, double x0,
                  double y0, double z0, double jitter, uint64_t seed, rflu_grid** out) {
  if (!out) { g_err = "null argument"; return 1; }
  *out = nullptr;
  rflu_grid* g = new (std::nothrow) rflu_grid();
  if (!g) { g_err = "out of host memory"; return 6; }
  int rc = guarded([&] { return grid_box(*g, kind, nx, ny, nz, lx, ly, lz, x0, y0, z0, jitter, seed); });
  if (rc) { delete g; return rc; }
  *out = g;
  return 0;
}

int rflu_mesh_build(const rflu_grid* g, const int32_t* bc_type, const int32_t* bc_opt,
                    const double* bc_pout, int build_stencils_flag, rflu_mesh** out) {
  if (!g || !out || !bc_type || !bc_opt || !bc_pout) { g_err = "null argument"; return 1; }
  *out = nullptr;
  rflu_mesh* m = new (std::nothrow) rflu_mesh();
  if (!m) { g_err = "out of host memory"; return 6; }
  int rc = guarded([&] {
    return rflu::build_region(*m, *g, bc_type, bc_opt, bc_pout, build_stencils_flag);
  });
  if (rc) { delete m; return rc; }
  m->kind = RFLU_MESH_GRID;
  *out = m;
  return 0;
}

int rflu_mesh_box(int kind, int nx, int ny, int nz, double lx, double ly, double lz, double x0,
                  double y0, double z0, double jitter, uint64_t seed, const int32_t bc_type[6],
                  const int32_t bc_opt[6], const double bc_pout[6], int build_stencils_flag,
                  rflu_mesh** out) {
  if (!out || !bc_type || !bc_opt || !bc_pout) { g_err = "null argument"; return 1; }
  *out = nullptr;
  rflu_mesh* m = new (std::nothrow) rflu_mesh();
  if (!m) { g_err = "out of host memory"; return 6; }
  int rc = guarded([&] {
    rflu_grid g;
    int r = grid_box(g, kind, nx, ny, nz, lx, ly, lz, x0, y0, z0, jitter, seed);
    return r ? r : rflu::build_region(*m, g, bc_type, bc_opt, bc_pout, build_stencils_flag);
  });
  if (rc) { delete m; return rc; }
  m->kind = kind; m->nx = nx; m->ny = ny; m->nz = nz;
  *out = m;
  return 0;
}

void rflu_mesh_destroy(rflu_mesh* m) { delete m; }

int rflu_mesh_sizes(const rflu_mesh* m, int64_t out[9]) {
  if (!m || !out) { g_err = "null argument"; return 1; }
  out[0] = m->n_cells; out[1] = m->n_faces; out[2] = m->n_bfaces; out[3] = m->n_patches;
  out[4] = m->n_verts; out[5] = (int64_t)m->c2cs_idx.size(); out[6] = m->nvpc;
  out[7] = m->stencil_max; out[8] = m->stencil_min;
  return 0;
}

int rflu_mesh_array(const rflu_mesh* m, const char* name, const void** ptr, int64_t* count,
                    int* dtype) {
  if (!m || !name || !ptr || !count || !dtype) { g_err = "null argument"; return 1; }
  std::string n(name);
  auto I = [&](const std::vector<int32_t>& v) { *ptr = v.data(); *count = (int64_t)v.size(); *dtype = 0; return 0; };
  auto D = [&](const std::vector<double>& v) { *ptr = v.data(); *count = (int64_t)v.size(); *dtype = 1; return 0; };
  if (n == "f2c") return I(m->f2c);
  if (n == "f2v") return I(m->f2v);
  if (n == "bf2c") return I(m->bf2c);
  if (n == "bf2v") return I(m->bf2v);
  if (n == "bpatch_ptr") return I(m->bpatch_ptr);
  if (n == "patch_bc") return I(m->patch_bc);
  if (n == "patch_opt") return I(m->patch_opt);
  if (n == "c2v") return I(m->c2v);
  if (n == "c2v_ptr") return I(m->c2v_ptr);
  if (n == "c2cs_ptr") return I(m->c2cs_ptr);
  if (n == "c2cs_idx") return I(m->c2cs_idx);
  if (n == "c2cs_bptr") return I(m->c2cs_bptr);
  if (n == "c2cs_bface") return I(m->c2cs_bface);
  if (n == "c2f_ptr") return I(m->c2f_ptr);
  if (n == "c2f_idx") return I(m->c2f_idx);
  if (n == "xyz") return D(m->xyz);
  if (n == "fn") return D(m->fn);
  if (n == "fc") return D(m->fc);
  if (n == "bfn") return D(m->bfn);
  if (n == "bfc") return D(m->bfc);
  if (n == "patch_pout") return D(m->patch_pout);
  if (n == "vol") return D(m->vol);
  if (n == "cofg") return D(m->cofg);
  if (n == "xyz_moms") return D(m->xyz_moms);
  if (n == "virt_src") return I(m->virt_src);
  if (n == "virt_xf") return I(m->virt_xf);
  if (n == "virt_patch") return I(m->virt_patch);
  if (n == "patch_map") return I(m->patch_map);
  if (n == "xforms") return D(m->xforms);
  g_err = "unknown array name: " + n;
  return 2;
}

const char* rflu_mesh_last_error(void) { return g_err.c_str(); }

}  // extern "C"
