// Symmetry and periodic patches as virtual cells.
//
// The reference turns a symmetry or periodic patch into layers of virtual cells when it
// partitions the grid (utilities/part/rflupart.F90:361-372 calls
// RFLU_SYPE_AddVirtualCells, modflu/RFLU_ModSymmetryPeriodic.F90:148-335): for each
// such patch in patch order, the cells near the source patch (the patch itself for a
// symmetry patch, its related patch for a periodic one) -- the cells touching its
// vertices and two further layers of their neighbours
// (RFLU_SYPE_AddVirtualCellsInv2, :471-787) -- are copied across the patch, mirrored
// in its plane or moved by the period, and added as virtual cells
// (RFLU_SYPE_AugmentCellLists, :813-1421).  Virtual cells that earlier patches made
// are copied too, which fills the edges and corners where patches meet; a periodic
// patch does not copy the virtual cells of its own related patch.  The faces the
// patch held become faces between actual and virtual cells.  At run time a virtual
// cell takes the state and the final gradients of the actual cell it mirrors, through
// the transform of the copies that made it (the border exchange of
// libfloflu/RungeKuttaMP.F90:229-236 and 440-459 with RFLU_RELP_TransformWrapper).
//
// Here the construction works on the grid, before the region is built: copied cells
// get new vertices except where a copy lands on an existing vertex (the patch plane,
// and where copies of different patches meet), the copies follow the actual cells in
// the cell numbering, the outer faces of the virtual layers form one extra patch, and
// the region (faces, geometry, stencils) is then built from the augmented grid by
// rflu::build_region.  Layers are vertex-neighbour layers (the reference's stencils
// of its first-order basic stencil: every vertex neighbour up to 40).
#include "rflu_mesh.h"
#include "mesh_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// Coincident-vertex lookup: buckets of side h, each query checks the 27 around it.
struct VertexIndex {
  double h = 1.0;
  std::unordered_map<uint64_t, std::vector<int32_t>> buckets;
  static uint64_t key(int64_t i, int64_t j, int64_t k) {
    return ((uint64_t)i * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)j * 0xC2B2AE3D27D4EB4FULL) ^
           ((uint64_t)k * 0x165667B19E3779F9ULL);
  }
  int64_t cell(double x) const { return (int64_t)std::floor(x / h); }
  void add(const double* x, int32_t v) {
    buckets[key(cell(x[0]), cell(x[1]), cell(x[2]))].push_back(v);
  }
  int32_t find(const double* x, const std::vector<double>& X, double tol) const {
    const int64_t i = cell(x[0]), j = cell(x[1]), k = cell(x[2]);
    for (int64_t a = i - 1; a <= i + 1; ++a)
      for (int64_t b = j - 1; b <= j + 1; ++b)
        for (int64_t c = k - 1; c <= k + 1; ++c) {
          auto it = buckets.find(key(a, b, c));
          if (it == buckets.end()) continue;
          for (int32_t v : it->second)
            if (std::fabs(X[3 * (size_t)v] - x[0]) <= tol &&
                std::fabs(X[3 * (size_t)v + 1] - x[1]) <= tol &&
                std::fabs(X[3 * (size_t)v + 2] - x[2]) <= tol)
              return v;
        }
    return -1;
  }
};

// Vertex orders that undo the handedness a reflection gives a cell, per kind (the
// local faces of the copy then point out of it again).
const int MIRROR_ORDER[4][8] = {{0, 2, 1, 3},
                                {0, 3, 2, 1, 4, 7, 6, 5},
                                {0, 2, 1, 3, 5, 4},
                                {0, 3, 2, 1, 4}};

using Mat = std::array<double, 9>;

Mat matmul(const Mat& a, const Mat& b) {
  Mat c{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
      c[3 * i + j] = s;
    }
  return c;
}

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
FaceKey make_key(const int32_t* fv, int nvf) {
  FaceKey key;
  for (int a = 0; a < 4; ++a) key.v[a] = a < nvf ? fv[a] : -1;
  std::sort(key.v, key.v + nvf);
  return key;
}

int fail(int rc, const std::string& msg) {
  rflu::set_error(msg);
  return rc;
}

int build_sype(rflu_mesh& m, const rflu_grid& g, const int32_t* bc_type, const int32_t* bc_opt,
               const double* bc_pout, const int32_t* related, const int32_t* axis,
               const double* angle, int want_stencils) {
  const std::vector<int32_t>* lists[4] = {&g.tet2v, &g.hex2v, &g.pri2v, &g.pyr2v};
  int kind = -1, nkinds = 0;
  for (int t = 0; t < 4; ++t)
    if (!lists[t]->empty()) { kind = t; ++nkinds; }
  if (nkinds != 1) return fail(1, "symmetry/periodic patches need a grid of one cell kind");
  const rflu::CellKind& ck = rflu::KINDS[kind];
  const int nv = ck.nv;
  std::vector<int32_t> c2v(*lists[kind]);
  const int n_act = (int)(c2v.size() / nv);
  const int P = (int)g.patch_type.size();
  std::vector<double> X(g.xyz);
  const int V0 = g.n_verts;
  if (V0 < 4 || X.size() != (size_t)3 * V0) return fail(1, "grid: bad vertex count");
  for (int32_t v : c2v)
    if (v < 0 || v >= V0) return fail(1, "grid: cell vertex out of range");

  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int v = 0; v < V0; ++v)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], X[3 * (size_t)v + d]);
      hi[d] = std::max(hi[d], X[3 * (size_t)v + d]);
    }
  const double diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                (hi[2] - lo[2]) * (hi[2] - lo[2]));
  const double tol = 1e-9 * diag;
  VertexIndex idx;
  idx.h = 2.0 * tol;
  for (int v = 0; v < V0; ++v) idx.add(&X[3 * (size_t)v], v);

  // faces and vertices of each patch
  std::vector<int> tri0(P + 1, 0), quad0(P + 1, 0);
  for (int p = 0; p < P; ++p) {
    tri0[p + 1] = tri0[p] + g.patch_ntri[p];
    quad0[p + 1] = quad0[p] + g.patch_nquad[p];
  }
  if ((size_t)3 * tri0[P] != g.btri2v.size() || (size_t)4 * quad0[P] != g.bquad2v.size())
    return fail(1, "grid: patch face counts do not match the face lists");
  std::vector<std::vector<int32_t>> pverts(P);
  for (int p = 0; p < P; ++p) {
    auto& pv = pverts[p];
    pv.insert(pv.end(), g.btri2v.begin() + 3 * (size_t)tri0[p], g.btri2v.begin() + 3 * (size_t)tri0[p + 1]);
    pv.insert(pv.end(), g.bquad2v.begin() + 4 * (size_t)quad0[p], g.bquad2v.begin() + 4 * (size_t)quad0[p + 1]);
    std::sort(pv.begin(), pv.end());
    pv.erase(std::unique(pv.begin(), pv.end()), pv.end());
  }

  std::vector<int32_t> src(n_act), xf(n_act, 0), vpatch(n_act, -1);
  for (int c = 0; c < n_act; ++c) src[c] = c;
  std::vector<Mat> mats = {Mat{1, 0, 0, 0, 1, 0, 0, 0, 1}};
  std::vector<std::vector<int32_t>> v2c(V0);
  for (int c = 0; c < n_act; ++c)
    for (int a = 0; a < nv; ++a) v2c[c2v[(size_t)nv * c + a]].push_back(c);

  bool any = false;
  for (int p = 0; p < P; ++p) {
    const int t = bc_type[p];
    if (t != RFLU_BC_SYMMETRY && t != RFLU_BC_PERIODIC) continue;
    any = true;
    const std::string pname = "patch " + std::to_string(p + 1);
    if (pverts[p].empty()) return fail(1, pname + " has no faces");
    Mat M{1, 0, 0, 0, 1, 0, 0, 0, 1};
    double T[3] = {0.0, 0.0, 0.0};
    int srcp = p;
    bool mirror = false;
    if (t == RFLU_BC_PERIODIC) {
      const int q = related ? related[p] - 1 : -1;
      if (q < 0 || q >= P || q == p || bc_type[q] != RFLU_BC_PERIODIC || related[q] - 1 != p)
        return fail(1, pname + ": a periodic patch needs a periodic related patch that "
                               "relates back to it");
      const int a = axis ? axis[p] - 1 : -1;
      if (a < 0 || a > 2) return fail(1, pname + ": periodic AXIS must be 1, 2 or 3");
      if (angle && angle[p] != 0.0)
        return fail(1, pname + ": rotational periodicity (ANGLE /= 0) is not built");
      auto plane = [&](int pp, double& c) {
        c = X[3 * (size_t)pverts[pp][0] + a];
        for (int32_t v : pverts[pp])
          if (std::fabs(X[3 * (size_t)v + a] - c) > tol) return false;
        return true;
      };
      double cp, cq;
      if (!plane(p, cp) || !plane(q, cq))
        return fail(1, pname + ": periodic patches must be planes normal to their AXIS");
      T[a] = cp - cq;
      srcp = q;
    } else {
      // the plane of the patch: normal of its first face, every vertex on it
      const int32_t* f = g.patch_ntri[p] ? &g.btri2v[3 * (size_t)tri0[p]] : &g.bquad2v[4 * (size_t)quad0[p]];
      const double* x0 = &X[3 * (size_t)f[0]];
      const double* x1 = &X[3 * (size_t)f[1]];
      const double* x2 = &X[3 * (size_t)f[2]];
      double e1[3], e2[3], n[3];
      for (int d = 0; d < 3; ++d) { e1[d] = x1[d] - x0[d]; e2[d] = x2[d] - x0[d]; }
      n[0] = e1[1] * e2[2] - e1[2] * e2[1];
      n[1] = e1[2] * e2[0] - e1[0] * e2[2];
      n[2] = e1[0] * e2[1] - e1[1] * e2[0];
      const double nm = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (!(nm > 0.0)) return fail(1, pname + ": degenerate face");
      for (double& c : n) c /= nm;
      // an axis-aligned plane keeps an exact mirror
      for (int d = 0; d < 3; ++d)
        if (std::fabs(std::fabs(n[d]) - 1.0) <= 1e-12) {
          for (int e = 0; e < 3; ++e) n[e] = e == d ? (n[d] > 0 ? 1.0 : -1.0) : 0.0;
          break;
        }
      double pd = n[0] * x0[0] + n[1] * x0[1] + n[2] * x0[2];
      for (int32_t v : pverts[p]) {
        const double* x = &X[3 * (size_t)v];
        if (std::fabs(n[0] * x[0] + n[1] * x[1] + n[2] * x[2] - pd) > tol)
          return fail(1, pname + ": a symmetry patch must be planar");
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[3 * i + j] = (i == j ? 1.0 : 0.0) - 2.0 * n[i] * n[j];
      for (int d = 0; d < 3; ++d) T[d] = 2.0 * pd * n[d];
      mirror = true;
    }

    // cells near the source patch: touching its vertices, then two vertex layers
    const int ncells = (int)src.size();
    auto eligible = [&](int c) {
      return !(t == RFLU_BC_PERIODIC && vpatch[c] == srcp);
    };
    std::vector<char> in(ncells, 0);
    std::vector<int32_t> layer, all;
    for (int32_t v : pverts[srcp])
      for (int32_t c : v2c[v])
        if (eligible(c) && !in[c]) { in[c] = 1; layer.push_back(c); }
    all = layer;
    for (int l = 2; l <= 3; ++l) {
      std::vector<int32_t> next;
      for (int32_t c : layer)
        for (int a = 0; a < nv; ++a)
          for (int32_t c2 : v2c[c2v[(size_t)nv * c + a]])
            if (eligible(c2) && !in[c2]) { in[c2] = 1; next.push_back(c2); }
      all.insert(all.end(), next.begin(), next.end());
      layer.swap(next);
    }
    std::sort(all.begin(), all.end());

    const int first_new = ncells;
    for (int32_t s : all) {
      int32_t vv[8];
      for (int a = 0; a < nv; ++a) {
        const double* x = &X[3 * (size_t)c2v[(size_t)nv * s + a]];
        double y[3];
        for (int i = 0; i < 3; ++i)
          y[i] = M[3 * i] * x[0] + M[3 * i + 1] * x[1] + M[3 * i + 2] * x[2] + T[i];
        int32_t w = idx.find(y, X, tol);
        if (w < 0) {
          w = (int32_t)(X.size() / 3);
          X.insert(X.end(), y, y + 3);
          idx.add(y, w);
          v2c.emplace_back();
        }
        vv[a] = w;
      }
      for (int a = 0; a < nv; ++a)
        c2v.push_back(mirror ? vv[MIRROR_ORDER[kind][a]] : vv[a]);
      src.push_back(src[s]);
      const Mat comp = matmul(M, mats[xf[s]]);
      int k = 0;
      for (; k < (int)mats.size(); ++k) {
        bool same = true;
        for (int e = 0; e < 9 && same; ++e) same = std::fabs(mats[k][e] - comp[e]) <= 1e-12;
        if (same) break;
      }
      if (k == (int)mats.size()) mats.push_back(comp);
      xf.push_back(k);
      vpatch.push_back(p);
    }
    for (int c = first_new; c < (int)src.size(); ++c)
      for (int a = 0; a < nv; ++a) v2c[c2v[(size_t)nv * c + a]].push_back(c);
  }
  if (!any) return fail(1, "no symmetry or periodic patch");
  const int N = (int)src.size();
  if ((int64_t)N > 400000000LL) return fail(1, "mesh too large");

  // the augmented grid: kept patches in grid order, then the outer faces of the
  // virtual layers
  rflu_grid g2;
  g2.title = g.title;
  g2.n_verts = (int)(X.size() / 3);
  g2.xyz = X;
  std::vector<int32_t>* out_lists[4] = {&g2.tet2v, &g2.hex2v, &g2.pri2v, &g2.pyr2v};
  *out_lists[kind] = c2v;
  std::vector<int32_t> bct, bco, pmap;
  std::vector<double> bcp;
  std::unordered_map<FaceKey, int, FaceKeyHash> kept, sype_faces;
  for (int p = 0; p < P; ++p) {
    const bool sy = bc_type[p] == RFLU_BC_SYMMETRY || bc_type[p] == RFLU_BC_PERIODIC;
    for (int k = tri0[p]; k < tri0[p + 1]; ++k)
      (sy ? sype_faces : kept)[make_key(&g.btri2v[3 * (size_t)k], 3)] = p;
    for (int k = quad0[p]; k < quad0[p + 1]; ++k)
      (sy ? sype_faces : kept)[make_key(&g.bquad2v[4 * (size_t)k], 4)] = p;
    if (sy) continue;
    g2.patch_type.push_back(g.patch_type[p]);
    g2.patch_ntri.push_back(g.patch_ntri[p]);
    g2.patch_nquad.push_back(g.patch_nquad[p]);
    g2.patch_name.push_back(p < (int)g.patch_name.size() ? g.patch_name[p] : "");
    g2.btri2v.insert(g2.btri2v.end(), g.btri2v.begin() + 3 * (size_t)tri0[p], g.btri2v.begin() + 3 * (size_t)tri0[p + 1]);
    g2.bquad2v.insert(g2.bquad2v.end(), g.bquad2v.begin() + 4 * (size_t)quad0[p], g.bquad2v.begin() + 4 * (size_t)quad0[p + 1]);
    bct.push_back(bc_type[p]);
    bco.push_back(bc_opt[p]);
    bcp.push_back(bc_pout[p]);
    pmap.push_back(p);
  }
  struct Once { int count; int32_t cyc[4]; int nvf; };
  std::unordered_map<FaceKey, Once, FaceKeyHash> faces;
  std::vector<FaceKey> order;
  faces.reserve((size_t)N * ck.nf);
  for (int c = 0; c < N; ++c)
    for (int lf = 0; lf < ck.nf; ++lf) {
      const int nvf = ck.lf[lf][3] < 0 ? 3 : 4;
      int32_t fv[4] = {-1, -1, -1, -1};
      for (int a = 0; a < nvf; ++a) fv[a] = c2v[(size_t)nv * c + ck.lf[lf][a]];
      FaceKey key = make_key(fv, nvf);
      auto it = faces.find(key);
      if (it != faces.end()) {
        if (++it->second.count > 2) return fail(2, "non-manifold face among the virtual cells");
        continue;
      }
      Once o{1, {fv[0], fv[1], fv[2], fv[3]}, nvf};
      faces.emplace(key, o);
      order.push_back(key);
    }
  int vtri = 0, vquad = 0;
  std::vector<int32_t> vt, vq;
  for (const FaceKey& key : order) {
    const Once& o = faces[key];
    if (o.count != 1 || kept.count(key)) continue;
    auto s = sype_faces.find(key);
    if (s != sype_faces.end())
      return fail(3, "patch " + std::to_string(s->second + 1) +
                     " keeps a face without a virtual cell behind it");
    if (o.nvf == 3) { vt.insert(vt.end(), o.cyc, o.cyc + 3); ++vtri; }
    else { vq.insert(vq.end(), o.cyc, o.cyc + 4); ++vquad; }
  }
  g2.patch_type.push_back(RFLU_BC_VIRTUAL);
  g2.patch_ntri.push_back(vtri);
  g2.patch_nquad.push_back(vquad);
  g2.patch_name.push_back("virtual");
  g2.btri2v.insert(g2.btri2v.end(), vt.begin(), vt.end());
  g2.bquad2v.insert(g2.bquad2v.end(), vq.begin(), vq.end());
  bct.push_back(RFLU_BC_VIRTUAL);
  bco.push_back(0);
  bcp.push_back(0.0);
  pmap.push_back(-1);

  int rc = rflu::build_region(m, g2, bct.data(), bco.data(), bcp.data(), want_stencils);
  if (rc) return rc;
  // the outer faces of the virtual layers reach virtual cells only, whose residuals
  // are not used; they are slip walls for the flux kernels (the wall-face lists of the
  // stencils were built without them)
  m.patch_bc.back() = RFLU_BC_SLIPWALL;
  m.virt_src.assign(src.begin() + n_act, src.end());
  m.virt_xf.assign(xf.begin() + n_act, xf.end());
  m.virt_patch.assign(vpatch.begin() + n_act, vpatch.end());
  m.patch_map = pmap;
  m.xforms.clear();
  for (const Mat& a : mats) m.xforms.insert(m.xforms.end(), a.begin(), a.end());
  return 0;
}

}  // namespace

extern "C" int rflu_mesh_build_sype(const rflu_grid* g, const int32_t* bc_type,
                                    const int32_t* bc_opt, const double* bc_pout,
                                    const int32_t* related, const int32_t* axis,
                                    const double* angle, int build_stencils, rflu_mesh** out) {
  if (!g || !out || !bc_type || !bc_opt || !bc_pout) return fail(1, "null argument");
  *out = nullptr;
  rflu_mesh* m = new (std::nothrow) rflu_mesh();
  if (!m) return fail(6, "out of host memory");
  int rc;
  try {
    rc = build_sype(*m, *g, bc_type, bc_opt, bc_pout, related, axis, angle, build_stencils);
  } catch (const std::bad_alloc&) {
    rc = fail(6, "out of host memory");
  }
  if (rc) { delete m; return rc; }
  m->kind = RFLU_MESH_GRID;
  *out = m;
  return 0;
}
