// CENTAUR ASCII grid files (.hyb.asc): reader and writer of rflu_grid.
//
// The layout is the one RFLU_ReadGridCENTAURASCII reads
// (utilities/part/RFLU_ModCENTAUR.F90:776-1315), all records fixed-width:
//   title (A80); nVert (I16); x, y, z blocks of nVert reals, 5 per line, 16 or 24
//   characters each (the width is taken from the first coordinate line); nVert
//   dummy integers (10I16); for tetrahedra, hexahedra, prisms and pyramids in turn:
//   the count (I16) and, when nonzero, one block of counts per local vertex (10I16)
//   and a dummy block; nPatches (I16); when nonzero: three blocks of nPatches
//   integers (patch code, last triangle, last quadrilateral: cumulative), one name
//   per line (A80), nBTris (I16) with three vertex blocks, nBQuads (I16) with four.
// Vertex numbers are 1-based in the file.  Patches without faces are dropped and
// the rest renumbered, as RFLU_ConvCENTAUR2ROCFLU does (:375-669).
#include "rflu_mesh.h"
#include "mesh_internal.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace {

struct Reader {
  std::ifstream in;
  std::string path, line;
  long lineno = 0;
  std::string err;

  bool next() {
    if (!std::getline(in, line)) {
      err = path + ": unexpected end of file after line " + std::to_string(lineno);
      return false;
    }
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
  }
  bool fail(const std::string& what) {
    err = path + ":" + std::to_string(lineno) + ": " + what;
    return false;
  }
  // the k-th fixed-width field of the current line, blanks removed (Fortran's
  // list of blank-padded records with blanks ignored)
  std::string field(int k, int w) const {
    std::string f;
    for (size_t i = (size_t)k * w; i < (size_t)(k + 1) * w && i < line.size(); ++i)
      if (line[i] != ' ' && line[i] != '\t') f.push_back(line[i]);
    return f;
  }
  bool int_field(int k, int w, int64_t& v) {
    std::string f = field(k, w);
    if (f.empty()) { v = 0; return true; }
    char* end = nullptr;
    errno = 0;
    long long x = std::strtoll(f.c_str(), &end, 10);
    if (*end || errno) return fail("not an integer field: '" + f + "'");
    v = x;
    return true;
  }
  // Ew.d input: a field with a decimal point carries its own digits; without one the
  // last d mantissa digits are decimals
  bool real_field(int k, int w, int d, double& v) {
    std::string f = field(k, w);
    if (f.empty()) { v = 0.0; return true; }
    for (char& c : f)
      if (c == 'D' || c == 'd') c = 'E';
    if (f.find('.') == std::string::npos) {
      size_t e = f.find_first_of("Ee");
      if (e == std::string::npos) e = f.find_first_of("+-", 1);
      std::string mant = f.substr(0, e), ex = e == std::string::npos ? "" : f.substr(e);
      std::string sign;
      if (!mant.empty() && (mant[0] == '+' || mant[0] == '-')) { sign = mant.substr(0, 1); mant = mant.substr(1); }
      while ((int)mant.size() <= d) mant = "0" + mant;
      mant.insert(mant.size() - d, ".");
      if (!ex.empty() && ex[0] != 'E' && ex[0] != 'e') ex = "E" + ex;
      f = sign + mant + ex;
    }
    char* end = nullptr;
    v = std::strtod(f.c_str(), &end);
    if (*end) return fail("not a real field: '" + f + "'");
    return true;
  }
  bool one_int(int64_t& v) { return next() && int_field(0, 16, v); }
  bool int_block(int64_t n, std::vector<int32_t>* out) {
    for (int64_t i = 0; i < n; i += 10) {
      if (!next()) return false;
      for (int k = 0; k < 10 && i + k < n; ++k) {
        int64_t v;
        if (!int_field(k, 16, v)) return false;
        if (out) out->push_back((int32_t)v);
      }
    }
    return true;
  }
};

// cells of one kind: nv blocks of n (vertex l of every cell), then a dummy block
bool read_cells(Reader& r, int nv, std::vector<int32_t>& x2v, int nvert) {
  int64_t n;
  if (!r.one_int(n)) return false;
  if (n < 0 || n > 400000000) return r.fail("bad cell count");
  if (n == 0) return true;
  std::vector<int32_t> blk;
  blk.reserve((size_t)n * nv);
  for (int l = 0; l < nv; ++l)
    if (!r.int_block(n, &blk)) return false;
  if (!r.int_block(n, nullptr)) return false;
  x2v.resize((size_t)n * nv);
  for (int l = 0; l < nv; ++l)
    for (int64_t c = 0; c < n; ++c) {
      int32_t v = blk[(size_t)l * n + c];
      if (v < 1 || v > nvert) return r.fail("cell vertex out of range");
      x2v[(size_t)c * nv + l] = v - 1;
    }
  return true;
}

bool read_faces(Reader& r, int nv, int64_t n, std::vector<int32_t>& f2v, int nvert) {
  std::vector<int32_t> blk;
  blk.reserve((size_t)n * nv);
  for (int l = 0; l < nv; ++l)
    if (!r.int_block(n, &blk)) return false;
  f2v.resize((size_t)n * nv);
  for (int l = 0; l < nv; ++l)
    for (int64_t c = 0; c < n; ++c) {
      int32_t v = blk[(size_t)l * n + c];
      if (v < 1 || v > nvert) return r.fail("patch vertex out of range");
      f2v[(size_t)c * nv + l] = v - 1;
    }
  return true;
}

int read_centaur(const char* path, rflu_grid& g) {
  Reader r;
  r.path = path;
  r.in.open(path);
  if (!r.in) { rflu::set_error(std::string("cannot open ") + path); return 7; }
  auto failed = [&]() { rflu::set_error(r.err); return 8; };
  if (!r.next()) return failed();
  g.title = r.line.substr(0, 80);
  int64_t nv;
  if (!r.one_int(nv)) return failed();
  if (nv < 1 || nv > 2000000000) { r.fail("bad vertex count"); return failed(); }
  g.n_verts = (int)nv;
  // field width from the first coordinate line
  std::streampos mark = r.in.tellg();
  long mark_line = r.lineno;
  if (!r.next()) return failed();
  size_t len = r.line.find_last_not_of(" \t");
  len = len == std::string::npos ? 0 : len + 1;
  const int per_line = (int)std::min<int64_t>(5, nv);
  const int w = (int)(len / per_line);
  if ((int)len != per_line * w || (w != 16 && w != 24)) {
    r.fail("coordinates must be 5 values per line, 16 or 24 characters each; the first "
           "line has " + std::to_string(len) + " characters");
    return failed();
  }
  r.in.seekg(mark);
  r.lineno = mark_line;
  const int d = w == 16 ? 9 : 16;
  g.xyz.assign((size_t)3 * nv, 0.0);
  for (int i = 0; i < 3; ++i)
    for (int64_t j = 0; j < nv; j += 5) {
      if (!r.next()) return failed();
      for (int k = 0; k < 5 && j + k < nv; ++k) {
        double x;
        if (!r.real_field(k, w, d, x)) return failed();
        g.xyz[3 * (size_t)(j + k) + i] = x;
      }
    }
  if (!r.int_block(nv, nullptr)) return failed();
  if (!read_cells(r, 4, g.tet2v, g.n_verts) || !read_cells(r, 8, g.hex2v, g.n_verts) ||
      !read_cells(r, 6, g.pri2v, g.n_verts) || !read_cells(r, 5, g.pyr2v, g.n_verts))
    return failed();
  int64_t np;
  if (!r.one_int(np)) return failed();
  if (np < 0 || np > 100000) { r.fail("bad patch count"); return failed(); }
  if (np == 0) return 0;
  std::vector<int32_t> info;
  for (int i = 0; i < 3; ++i)
    if (!r.int_block(np, &info)) return failed();
  std::vector<std::string> names((size_t)np);
  for (int64_t p = 0; p < np; ++p) {
    if (!r.next()) return failed();
    std::string n = r.line.substr(0, 80);
    size_t a = n.find_first_not_of(' '), b = n.find_last_not_of(' ');
    names[p] = a == std::string::npos ? "" : n.substr(a, b - a + 1);
  }
  int64_t ntri, nquad;
  std::vector<int32_t> tri, quad;
  if (!r.one_int(ntri)) return failed();
  if (ntri < 0) { r.fail("bad triangle count"); return failed(); }
  if (ntri > 0 && !read_faces(r, 3, ntri, tri, g.n_verts)) return failed();
  if (!r.one_int(nquad)) return failed();
  if (nquad < 0) { r.fail("bad quadrilateral count"); return failed(); }
  if (nquad > 0 && !read_faces(r, 4, nquad, quad, g.n_verts)) return failed();
  int32_t tprev = 0, qprev = 0;
  for (int64_t p = 0; p < np; ++p) {
    int32_t tend = info[np + p], qend = info[2 * np + p];
    if (tend < tprev || qend < qprev || tend > ntri || qend > nquad) {
      r.err = std::string(path) + ": patch " + std::to_string(p + 1) + " face ranges are not "
              "ascending within the face counts";
      return failed();
    }
    if (tend == tprev && qend == qprev) {
      std::fprintf(stderr, "%s: patch %lld is empty and is dropped\n", path, (long long)p + 1);
      continue;
    }
    g.patch_type.push_back(info[p]);
    g.patch_ntri.push_back(tend - tprev);
    g.patch_nquad.push_back(qend - qprev);
    g.patch_name.push_back(names[p]);
    g.btri2v.insert(g.btri2v.end(), tri.begin() + 3 * (size_t)tprev, tri.begin() + 3 * (size_t)tend);
    g.bquad2v.insert(g.bquad2v.end(), quad.begin() + 4 * (size_t)qprev,
                     quad.begin() + 4 * (size_t)qend);
    tprev = tend;
    qprev = qend;
  }
  if (tprev != ntri || qprev != nquad) {
    r.err = std::string(path) + ": the patches do not cover every boundary face";
    return failed();
  }
  return 0;
}

// Fortran-style writer of fixed-width records
struct Writer {
  FILE* f = nullptr;
  bool ok = true;
  void ints(const std::vector<int32_t>& v, size_t n, size_t stride, size_t off, int add) {
    for (size_t i = 0; i < n; ++i) {
      ok = ok && std::fprintf(f, "%16d", v[i * stride + off] + add) == 16;
      if (i % 10 == 9 || i + 1 == n) ok = ok && std::fputc('\n', f) != EOF;
    }
  }
  void dummy(size_t n) {
    for (size_t i = 0; i < n; ++i) {
      ok = ok && std::fprintf(f, "%16d", 0) == 16;
      if (i % 10 == 9 || i + 1 == n) ok = ok && std::fputc('\n', f) != EOF;
    }
  }
  void one(int64_t v) { ok = ok && std::fprintf(f, "%16lld\n", (long long)v) == 17; }
};

int write_centaur(const rflu_grid& g, const char* path, int width) {
  if (width != 16 && width != 24) { rflu::set_error("field width must be 16 or 24"); return 1; }
  // 24-character fields carry 17 significant digits, enough to read every double back
  const char* fmt = width == 16 ? "%16.9E" : "%24.16E";
  for (double x : g.xyz)
    if (!std::isfinite(x) || (x != 0.0 && std::fabs(x) < 1e-99) || std::fabs(x) >= 1e100) {
      rflu::set_error("coordinate outside the range of a two-digit exponent");
      return 1;
    }
  Writer w;
  w.f = std::fopen(path, "w");
  if (!w.f) { rflu::set_error(std::string("cannot create ") + path); return 7; }
  w.ok = std::fprintf(w.f, "%.80s\n", g.title.c_str()) > 0;
  const size_t V = (size_t)g.n_verts;
  w.one((int64_t)V);
  for (int i = 0; i < 3; ++i)
    for (size_t j = 0; j < V; ++j) {
      w.ok = w.ok && std::fprintf(w.f, fmt, g.xyz[3 * j + i]) == width;
      if (j % 5 == 4 || j + 1 == V) w.ok = w.ok && std::fputc('\n', w.f) != EOF;
    }
  w.dummy(V);
  const std::vector<int32_t>* lists[4] = {&g.tet2v, &g.hex2v, &g.pri2v, &g.pyr2v};
  const int nvk[4] = {4, 8, 6, 5};
  for (int t = 0; t < 4; ++t) {
    size_t n = lists[t]->size() / nvk[t];
    w.one((int64_t)n);
    if (!n) continue;
    for (int l = 0; l < nvk[t]; ++l) w.ints(*lists[t], n, nvk[t], l, 1);
    w.dummy(n);
  }
  const size_t P = g.patch_type.size();
  w.one((int64_t)P);
  if (P) {
    std::vector<int32_t> tend(P), qend(P);
    int32_t ts = 0, qs = 0;
    for (size_t p = 0; p < P; ++p) {
      ts += g.patch_ntri[p]; qs += g.patch_nquad[p];
      tend[p] = ts; qend[p] = qs;
    }
    w.ints(g.patch_type, P, 1, 0, 0);
    w.ints(tend, P, 1, 0, 0);
    w.ints(qend, P, 1, 0, 0);
    for (size_t p = 0; p < P; ++p) {
      std::string n = p < g.patch_name.size() ? g.patch_name[p] : "";
      if (n.empty()) n = "patch " + std::to_string(p + 1);
      w.ok = w.ok && std::fprintf(w.f, "%.80s\n", n.c_str()) > 0;
    }
    size_t nt = g.btri2v.size() / 3, nq = g.bquad2v.size() / 4;
    w.one((int64_t)nt);
    for (int l = 0; l < 3 && nt; ++l) w.ints(g.btri2v, nt, 3, l, 1);
    w.one((int64_t)nq);
    for (int l = 0; l < 4 && nq; ++l) w.ints(g.bquad2v, nq, 4, l, 1);
  }
  bool closed = std::fclose(w.f) == 0;
  if (!w.ok || !closed) { rflu::set_error(std::string("write failed: ") + path); return 7; }
  return 0;
}

}  // namespace

extern "C" {

int rflu_grid_read_centaur(const char* path, rflu_grid** out) {
  if (!path || !out) { rflu::set_error("null argument"); return 1; }
  *out = nullptr;
  rflu_grid* g = new (std::nothrow) rflu_grid();
  if (!g) { rflu::set_error("out of host memory"); return 6; }
  int rc;
  try {
    rc = read_centaur(path, *g);
  } catch (const std::bad_alloc&) {
    rflu::set_error("out of host memory");
    rc = 6;
  }
  if (rc) { delete g; return rc; }
  *out = g;
  return 0;
}

int rflu_grid_write_centaur(const rflu_grid* g, const char* path, int field_width) {
  if (!g || !path) { rflu::set_error("null argument"); return 1; }
  return write_centaur(*g, path, field_width);
}

void rflu_grid_destroy(rflu_grid* g) { delete g; }

int rflu_grid_sizes(const rflu_grid* g, int64_t out[8]) {
  if (!g || !out) { rflu::set_error("null argument"); return 1; }
  out[0] = g->n_verts;
  out[1] = (int64_t)g->tet2v.size() / 4; out[2] = (int64_t)g->hex2v.size() / 8;
  out[3] = (int64_t)g->pri2v.size() / 6; out[4] = (int64_t)g->pyr2v.size() / 5;
  out[5] = (int64_t)g->patch_type.size();
  out[6] = (int64_t)g->btri2v.size() / 3; out[7] = (int64_t)g->bquad2v.size() / 4;
  return 0;
}

int rflu_grid_array(const rflu_grid* g, const char* name, const void** ptr, int64_t* count,
                    int* dtype) {
  if (!g || !name || !ptr || !count || !dtype) { rflu::set_error("null argument"); return 1; }
  std::string n(name);
  auto I = [&](const std::vector<int32_t>& v) { *ptr = v.data(); *count = (int64_t)v.size(); *dtype = 0; return 0; };
  if (n == "xyz") { *ptr = g->xyz.data(); *count = (int64_t)g->xyz.size(); *dtype = 1; return 0; }
  if (n == "tet2v") return I(g->tet2v);
  if (n == "hex2v") return I(g->hex2v);
  if (n == "pri2v") return I(g->pri2v);
  if (n == "pyr2v") return I(g->pyr2v);
  if (n == "patch_type") return I(g->patch_type);
  if (n == "patch_ntri") return I(g->patch_ntri);
  if (n == "patch_nquad") return I(g->patch_nquad);
  if (n == "btri2v") return I(g->btri2v);
  if (n == "bquad2v") return I(g->bquad2v);
  rflu::set_error("unknown grid array name: " + n);
  return 2;
}

int rflu_grid_create(int n_verts, const double* xyz, int64_t n_tets, const int32_t* tet2v,
                     int64_t n_hexs, const int32_t* hex2v, int64_t n_pris, const int32_t* pri2v,
                     int64_t n_pyrs, const int32_t* pyr2v, int n_patches,
                     const int32_t* patch_type, const int32_t* patch_ntri,
                     const int32_t* patch_nquad, const int32_t* btri2v, const int32_t* bquad2v,
                     rflu_grid** out) {
  if (!out || n_verts < 1 || !xyz || n_patches < 0 || n_tets < 0 || n_hexs < 0 || n_pris < 0 ||
      n_pyrs < 0 || (n_patches && (!patch_type || !patch_ntri || !patch_nquad))) {
    rflu::set_error("bad argument");
    return 1;
  }
  *out = nullptr;
  rflu_grid* g = new (std::nothrow) rflu_grid();
  if (!g) { rflu::set_error("out of host memory"); return 6; }
  try {
    g->title = "grid";
    g->n_verts = n_verts;
    g->xyz.assign(xyz, xyz + 3 * (size_t)n_verts);
    auto take = [](std::vector<int32_t>& v, const int32_t* src, int64_t n, int nv) {
      if (n) v.assign(src, src + (size_t)n * nv);
    };
    if ((n_tets && !tet2v) || (n_hexs && !hex2v) || (n_pris && !pri2v) || (n_pyrs && !pyr2v)) {
      delete g;
      rflu::set_error("null cell list");
      return 1;
    }
    take(g->tet2v, tet2v, n_tets, 4);
    take(g->hex2v, hex2v, n_hexs, 8);
    take(g->pri2v, pri2v, n_pris, 6);
    take(g->pyr2v, pyr2v, n_pyrs, 5);
    int64_t nt = 0, nq = 0;
    for (int p = 0; p < n_patches; ++p) {
      g->patch_type.push_back(patch_type[p]);
      g->patch_ntri.push_back(patch_ntri[p]);
      g->patch_nquad.push_back(patch_nquad[p]);
      g->patch_name.push_back("patch " + std::to_string(p + 1));
      nt += patch_ntri[p];
      nq += patch_nquad[p];
    }
    if ((nt && !btri2v) || (nq && !bquad2v) || nt < 0 || nq < 0) {
      delete g;
      rflu::set_error("bad patch face lists");
      return 1;
    }
    take(g->btri2v, btri2v, nt, 3);
    take(g->bquad2v, bquad2v, nq, 4);
  } catch (const std::bad_alloc&) {
    delete g;
    rflu::set_error("out of host memory");
    return 6;
  }
  *out = g;
  return 0;
}

}  // extern "C"
