// Host-side region builder: box meshes of hexahedra or Kuhn tetrahedra, face list,
// geometry, cell-to-cell stencils and their R-factors.  See include/rflu_mesh.h.
//
// Everything here runs once per region on the host, as the reference's
// RFLU_InitFlowSolver does (rocflu/RFLU_InitFlowSolver.F90:495-946).
#include "rflu_mesh.h"

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

struct rflu_mesh {
  int kind = 0, nx = 0, ny = 0, nz = 0;
  int nvpc = 8;  // vertices per cell
  int nfpc = 6;  // faces per cell
  int n_verts = 0, n_cells = 0, n_faces = 0, n_bfaces = 0, n_patches = 6;
  std::vector<double> xyz;
  std::vector<int32_t> c2v;
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
};

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

// Local faces of a hexahedron, corner index = bx + 2*by + 4*bz, given as perimeter
// cycles; orientation is fixed afterwards from the undisplaced geometry.
const int HEX_FACES[6][4] = {{0, 2, 6, 4}, {1, 3, 7, 5}, {0, 1, 5, 4},
                             {2, 3, 7, 6}, {0, 1, 3, 2}, {4, 5, 7, 6}};
const int TET_FACES[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

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
  const int N = m.n_cells, nv = m.nvpc;
  // vertex -> cells, ascending cell order
  std::vector<int32_t> v2c_ptr(m.n_verts + 1, 0), v2c;
  for (int c = 0; c < N; ++c)
    for (int k = 0; k < nv; ++k) v2c_ptr[m.c2v[nv * c + k] + 1]++;
  for (int v = 0; v < m.n_verts; ++v) v2c_ptr[v + 1] += v2c_ptr[v];
  v2c.resize(v2c_ptr[m.n_verts]);
  {
    std::vector<int32_t> fill(v2c_ptr.begin(), v2c_ptr.end() - 1);
    for (int c = 0; c < N; ++c)
      for (int k = 0; k < nv; ++k) v2c[fill[m.c2v[nv * c + k]]++] = c;
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
    tmp.clear();
    for (int k = 0; k < nv; ++k) {
      int v = m.c2v[nv * icg + k];
      for (int j = v2c_ptr[v]; j < v2c_ptr[v + 1]; ++j) tmp.push_back(v2c[j]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    tmp.erase(std::remove(tmp.begin(), tmp.end(), icg), tmp.end());
    frac.clear();
    fid.clear();
    int vs[8];
    for (int k = 0; k < nv; ++k) vs[k] = m.c2v[nv * icg + k];
    std::sort(vs, vs + nv);
    for (int c2 : tmp) {
      int vs2[8];
      for (int k = 0; k < nv; ++k) vs2[k] = m.c2v[nv * c2 + k];
      std::sort(vs2, vs2 + nv);
      int common = 0;
      for (int a = 0, b = 0; a < nv && b < nv;) {
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
        for (int k = 0; k < nv && !full; ++k) {
          int v = m.c2v[nv * c + k];
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

int build_box(rflu_mesh& m, int kind, int nx, int ny, int nz, double lx, double ly, double lz,
              double x0, double y0, double z0, double jitter, uint64_t seed,
              const int32_t* bc_type, const int32_t* bc_opt, const double* bc_pout,
              int want_stencils) {
  if (nx < 1 || ny < 1 || nz < 1) { g_err = "box dimensions must be >= 1"; return 1; }
  if (kind != RFLU_MESH_HEX && kind != RFLU_MESH_TET6) { g_err = "unknown cell kind"; return 1; }
  if (!(jitter >= 0.0 && jitter < 0.25)) { g_err = "jitter must be in [0, 0.25)"; return 1; }
  const int64_t ncell64 = (int64_t)nx * ny * nz * (kind == RFLU_MESH_TET6 ? 6 : 1);
  if (ncell64 > 400000000LL) { g_err = "mesh too large"; return 1; }
  m.kind = kind; m.nx = nx; m.ny = ny; m.nz = nz;
  m.nvpc = kind == RFLU_MESH_HEX ? 8 : 4;
  m.nfpc = kind == RFLU_MESH_HEX ? 6 : 4;
  const int NX = nx + 1, NY = ny + 1, NZ = nz + 1;
  m.n_verts = NX * NY * NZ;
  const double dx = lx / nx, dy = ly / ny, dz = lz / nz;

  // Undisplaced node positions (used to orient faces), then displaced ones.
  std::vector<double> xyz0((size_t)3 * m.n_verts);
  m.xyz.resize((size_t)3 * m.n_verts);
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
        m.xyz[3 * v] = px; m.xyz[3 * v + 1] = py; m.xyz[3 * v + 2] = pz;
      }

  // Cells
  const int nsub = kind == RFLU_MESH_HEX ? 1 : 6;
  m.n_cells = (int)ncell64;
  m.c2v.resize((size_t)m.nvpc * m.n_cells);
  // Kuhn split: the tetrahedron of axis order (a,b,c) is 0 -> e_a -> e_a+e_b -> 7.
  const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
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
          for (int b = 0; b < 8; ++b) m.c2v[8 * box + b] = corner[b];
        } else {
          for (int t = 0; t < 6; ++t) {
            int b1 = 1 << perms[t][0];
            int b2 = b1 | (1 << perms[t][1]);
            int32_t* tv = &m.c2v[4 * (6 * box + t)];
            tv[0] = corner[0]; tv[1] = corner[b1]; tv[2] = corner[b2]; tv[3] = corner[7];
          }
        }
      }
  (void)nsub;

  // Face list: faces are created in the order cells and their local faces are
  // visited; the first visiting cell is c1 and the vertex cycle is oriented so that
  // the face vector points out of c1.
  struct Pending { int32_t c1, c2, v[4], patch; };
  std::vector<Pending> faces;
  faces.reserve((size_t)m.n_cells * (kind == RFLU_MESH_HEX ? 3 : 2) + 16);
  std::unordered_map<FaceKey, int32_t, FaceKeyHash> map;
  map.reserve(faces.capacity() * 2);
  const int nfl = m.nfpc, nvf = kind == RFLU_MESH_HEX ? 4 : 3;
  for (int c = 0; c < m.n_cells; ++c) {
    const int32_t* cv = &m.c2v[(size_t)m.nvpc * c];
    double cc[3] = {0, 0, 0};
    for (int a = 0; a < m.nvpc; ++a)
      for (int d = 0; d < 3; ++d) cc[d] += xyz0[3 * cv[a] + d] / m.nvpc;
    for (int lf = 0; lf < nfl; ++lf) {
      int32_t fv[4] = {-1, -1, -1, -1};
      for (int a = 0; a < nvf; ++a)
        fv[a] = kind == RFLU_MESH_HEX ? cv[HEX_FACES[lf][a]] : cv[TET_FACES[lf][a]];
      FaceKey key;
      std::memcpy(key.v, fv, sizeof(fv));
      std::sort(key.v, key.v + nvf);
      auto it = map.find(key);
      if (it != map.end()) {
        Pending& p = faces[it->second];
        if (p.c2 >= 0) { g_err = "non-manifold face"; return 2; }
        p.c2 = c;
        map.erase(it);
        continue;
      }
      // orient outward from c using the undisplaced nodes
      double vec[3], cen[3];
      {
        const double* p1 = &xyz0[3 * fv[0]];
        const double* p2 = &xyz0[3 * fv[1]];
        const double* p3 = &xyz0[3 * fv[2]];
        if (nvf == 3) { face_vector_tria(p1, p2, p3, vec); face_centroid_tria(p1, p2, p3, cen); }
        else {
          const double* p4 = &xyz0[3 * fv[3]];
          face_vector_quad(p1, p2, p3, p4, vec);
          face_centroid_quad(p1, p2, p3, p4, cen);
        }
      }
      double dot = 0;
      for (int d = 0; d < 3; ++d) dot += vec[d] * (cen[d] - cc[d]);
      if (dot < 0) std::reverse(fv, fv + nvf);
      Pending p;
      p.c1 = c; p.c2 = -1; p.patch = -1;
      std::memcpy(p.v, fv, sizeof(fv));
      map.emplace(key, (int32_t)faces.size());
      faces.push_back(p);
    }
  }
  // Split into interior faces and boundary faces per patch.
  auto vijk = [&](int v, int& i, int& j, int& k) {
    i = v % NX; j = (v / NX) % NY; k = v / (NX * NY);
  };
  std::vector<std::vector<int32_t>> pfaces(6);
  m.f2c.clear(); m.f2v.clear();
  for (size_t f = 0; f < faces.size(); ++f) {
    Pending& p = faces[f];
    if (p.c2 >= 0) {
      m.f2c.push_back(p.c1); m.f2c.push_back(p.c2);
      for (int a = 0; a < 4; ++a) m.f2v.push_back(p.v[a]);
      continue;
    }
    int side = -1;
    int cnt[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < nvf; ++a) {
      int i, j, k;
      vijk(p.v[a], i, j, k);
      cnt[0] += i == 0; cnt[1] += i == nx; cnt[2] += j == 0;
      cnt[3] += j == ny; cnt[4] += k == 0; cnt[5] += k == nz;
    }
    for (int s = 0; s < 6; ++s)
      if (cnt[s] == nvf) { side = s; break; }
    if (side < 0) { g_err = "boundary face not on a box side"; return 3; }
    pfaces[side].push_back((int32_t)f);
  }
  m.n_faces = (int)(m.f2c.size() / 2);
  m.bpatch_ptr.assign(7, 0);
  m.bf2c.clear(); m.bf2v.clear();
  for (int s = 0; s < 6; ++s) {
    for (int32_t f : pfaces[s]) {
      m.bf2c.push_back(faces[f].c1);
      for (int a = 0; a < 4; ++a) m.bf2v.push_back(faces[f].v[a]);
    }
    m.bpatch_ptr[s + 1] = (int32_t)m.bf2c.size();
  }
  m.n_bfaces = (int)m.bf2c.size();
  m.patch_bc.assign(bc_type, bc_type + 6);
  m.patch_opt.assign(bc_opt, bc_opt + 6);
  m.patch_pout.assign(bc_pout, bc_pout + 6);
  for (int s = 0; s < 6; ++s)
    if (bc_type[s] != RFLU_BC_OUTFLOW && bc_type[s] != RFLU_BC_SLIPWALL &&
        bc_type[s] != RFLU_BC_INFLOW_VELTEMP && bc_type[s] != RFLU_BC_INFLOW_TOTANG &&
        bc_type[s] != RFLU_BC_FARFIELD && bc_type[s] != RFLU_BC_INJECTION) {
      g_err = "patch types: inflow (10, 11), outflow (20), slip wall (60), farfield (80), "
              "injection (90)";
      return 4;
    }

  // Geometry (modflu/RFLU_ModGeometry.F90:307-665): approximate centroids, face
  // vectors/centroids, volumes and centroids from the divergence theorem.
  const int N = m.n_cells, F = m.n_faces, B = m.n_bfaces;
  std::vector<double> app((size_t)3 * N, 0.0);
  {
    const double term = 1.0 / m.nvpc;
    for (int c = 0; c < N; ++c) {
      double s[3] = {0, 0, 0};
      for (int a = 0; a < m.nvpc; ++a)
        for (int d = 0; d < 3; ++d) s[d] += m.xyz[3 * m.c2v[(size_t)m.nvpc * c + a] + d];
      for (int d = 0; d < 3; ++d) app[3 * c + d] = term * s[d];
    }
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

}  // namespace

extern "C" {

int rflu_mesh_box(int kind, int nx, int ny, int nz, double lx, double ly, double lz, double x0,
                  double y0, double z0, double jitter, uint64_t seed, const int32_t bc_type[6],
                  const int32_t bc_opt[6], const double bc_pout[6], int build_stencils_flag,
                  rflu_mesh** out) {
  if (!out || !bc_type || !bc_opt || !bc_pout) { g_err = "null argument"; return 1; }
  *out = nullptr;
  rflu_mesh* m = new (std::nothrow) rflu_mesh();
  if (!m) { g_err = "out of host memory"; return 6; }
  int rc;
  try {
    rc = build_box(*m, kind, nx, ny, nz, lx, ly, lz, x0, y0, z0, jitter, seed, bc_type, bc_opt,
                   bc_pout, build_stencils_flag);
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    rc = 6;
  }
  if (rc) { delete m; return rc; }
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
  g_err = "unknown array name: " + n;
  return 2;
}

const char* rflu_mesh_last_error(void) { return g_err.c_str(); }

}  // extern "C"
