// Fluid dumps: the reference's restart files restart_data/fluid_<step>_<region>.dat.
//
// Stream layout (modflu/RFLU_ModRestartData.F90:7-33; writer RFLU_WriteRestartData
// :136-215, reader RFLU_ReadRestartData :693-872), raw bytes in native order, no
// record markers:
//   char[8] "RFLUDUMP", int32 version (5; 4 is still read), int64 step, double time,
//   int32 region, nCells, nCv, form (1 pure, 2 volume-weighted), withPicl,
//   double cv(iCv, 1:nCells) one variable after the other,
//   [PICL fields when withPicl = 1],
//   int32 number of NSCBC patches, then per patch int32 global patch, nBFaces and
//   double face states cv(iCv, 1:nBFaces) one variable after the other.
// A dump is written as <name>.tmp and renamed when complete (RFLU_FinishRestartFile
// :904-918).  The gas-only path writes the pure form, no PICL fields and no NSCBC
// patches, and reads what the reference's reader accepts for such a region.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace rflu_dump {

constexpr int32_t VERSION = 5;
constexpr int32_t FORM_PURE = 1;
constexpr int32_t FORM_VOLUME_WEIGHTED = 2;

enum Status { OK = 0, FILE_ERR = 1, FORMAT_ERR = 2 };

struct Header {
  int32_t version = VERSION;
  int64_t step = 0;
  double time = 0.0;
  int32_t region = 0, n_cells = 0, n_cv = 0, form = FORM_PURE, with_picl = 0;
};

struct File {
  FILE* f = nullptr;
  explicit File(FILE* f_) : f(f_) {}
  ~File() { if (f) std::fclose(f); }
  bool read(void* p, size_t n) { return std::fread(p, 1, n, f) == n; }
  bool write(const void* p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
  template <class T> bool get(T& v) { return read(&v, sizeof(T)); }
  template <class T> bool put(const T& v) { return write(&v, sizeof(T)); }
};

// restart_data/fluid_<step, 8 digits>_<region, 5 digits>.dat under out_dir
// (FluidDumpName, :258-266; out_dir ends in '/' like global%outDir, or is empty).
inline std::string name(const std::string& out_dir, int64_t step, int32_t region) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "restart_data/fluid_%08lld_%05d.dat", (long long)step, (int)region);
  return out_dir + buf;
}

// cv_var[iv][c]: variable iv of cell c (the file's order).  Returns OK, or FILE_ERR
// with msg set; a failed write leaves no file under the final name.
inline Status write(const std::string& path, const Header& h, const double* const* cv_var,
                    std::string& msg) {
  const std::string tmp = path + ".tmp";
  {
    File f(std::fopen(tmp.c_str(), "wb"));
    if (!f.f) { msg = "cannot open " + tmp; return FILE_ERR; }
    bool ok = f.write("RFLUDUMP", 8) && f.put(h.version) && f.put(h.step) && f.put(h.time) &&
              f.put(h.region) && f.put(h.n_cells) && f.put(h.n_cv) && f.put(h.form) &&
              f.put(h.with_picl);
    for (int iv = 0; ok && iv < h.n_cv; ++iv)
      ok = f.write(cv_var[iv], sizeof(double) * (size_t)h.n_cells);
    const int32_t n_nscbc = 0;
    ok = ok && f.put(n_nscbc);
    ok = ok && std::fflush(f.f) == 0;
    int rc = std::fclose(f.f);
    f.f = nullptr;
    if (!ok || rc != 0) {
      std::remove(tmp.c_str());
      msg = "cannot write " + tmp;
      return FILE_ERR;
    }
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    msg = "cannot rename " + tmp + " to " + path;
    return FILE_ERR;
  }
  return OK;
}

// Header only (RFLU_ChooseStartStep / RFLU_FluidDumpTime read the first fields).
inline Status read_header(File& f, const std::string& path, Header& h, std::string& msg) {
  char magic[8];
  if (!(f.read(magic, 8) && f.get(h.version) && f.get(h.step) && f.get(h.time) &&
        f.get(h.region) && f.get(h.n_cells) && f.get(h.n_cv) && f.get(h.form) &&
        f.get(h.with_picl)) ||
      std::memcmp(magic, "RFLUDUMP", 8) != 0 || (h.version != VERSION && h.version != 4)) {
    msg = path + ": not a fluid dump of this version (rerun rfluinit for step 0)";
    return FORMAT_ERR;
  }
  return OK;
}

// Reads a dump into cv_var[iv][c] with the reader's checks for a gas-only region of
// n_cells cells and n_cv variables: sizes must match, the form must be the pure one
// (a gas-only run expects it at every step), PICL fields are read past as the
// reader does (version 4: piclVF of every cell and three gradient fields; 5: the
// nonzero cells and their values, then a flag and the gradient fields), and the
// region must have face states for every NSCBC patch of the dump (it has none).
inline Status read(const std::string& path, int32_t n_cells, int32_t n_cv, Header& h,
                   double* const* cv_var, std::string& msg) {
  File f(std::fopen(path.c_str(), "rb"));
  if (!f.f) { msg = "cannot open " + path; return FILE_ERR; }
  Status s = read_header(f, path, h, msg);
  if (s != OK) return s;
  if (h.n_cells != n_cells || h.n_cv != n_cv) {
    msg = path + ": cells or variables differ from the grid of this region";
    return FORMAT_ERR;
  }
  if (h.form != FORM_PURE) {
    msg = path + ": written by a run with particles on/off differently from this one";
    return FORMAT_ERR;
  }
  const std::string rd = path + ": read error";
  for (int iv = 0; iv < n_cv; ++iv)
    if (!f.read(cv_var[iv], sizeof(double) * (size_t)n_cells)) { msg = rd; return FILE_ERR; }
  if (h.with_picl == 1) {
    std::vector<double> skip((size_t)n_cells);
    int n_grad = 0;
    if (h.version == 4) {
      if (!f.read(skip.data(), sizeof(double) * skip.size())) { msg = rd; return FILE_ERR; }
      n_grad = 3;
    } else {
      int32_t nnz = 0, with_grad = 0;
      if (!f.get(nnz) || nnz < 0 || nnz > n_cells) { msg = rd; return FILE_ERR; }
      std::vector<int32_t> cells((size_t)nnz);
      if (!f.read(cells.data(), sizeof(int32_t) * cells.size()) ||
          !f.read(skip.data(), sizeof(double) * (size_t)nnz)) { msg = rd; return FILE_ERR; }
      for (int32_t c : cells)
        if (c < 1 || c > n_cells) { msg = rd; return FILE_ERR; }
      if (!f.get(with_grad)) { msg = rd; return FILE_ERR; }
      n_grad = 3 * with_grad;
    }
    for (int i = 0; i < n_grad; ++i)
      if (!f.read(skip.data(), sizeof(double) * skip.size())) { msg = rd; return FILE_ERR; }
  }
  int32_t n_nscbc = 0;
  if (!f.get(n_nscbc)) { msg = rd; return FILE_ERR; }
  if (n_nscbc != 0) {
    msg = path + ": NSCBC patch not found in this region";
    return FORMAT_ERR;
  }
  return OK;
}

}  // namespace rflu_dump
