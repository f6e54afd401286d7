// The reference's .inp and .bc user input files (include/rflu_input.h).
#include "rflu_input.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

// NINT: nearest integer, halves away from zero
int32_t nint(double x) { return (int32_t)std::lround(x); }
// INT(ABS(x)+0.5)
int32_t iabs_round(double x) { return (int32_t)(std::fabs(x) + 0.5); }

struct Lines {
  std::ifstream in;
  std::string path, line;
  long lineno = 0;
  bool next() {
    if (!std::getline(in, line)) return false;
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
  }
  std::string where() const { return path + ":" + std::to_string(lineno); }
};

// TRIM: trailing blanks removed (tabs stay, as in Fortran)
std::string trim_right(const std::string& s) {
  size_t e = s.find_last_not_of(' ');
  return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

// LineHasKey (libfloflu/ReadInputUtil.F90:686-700)
bool line_has_key(const std::string& line, const char* key) {
  size_t nc = std::strlen(key);
  if (line.compare(0, nc, key) != 0) return false;
  return line.size() == nc || line[nc] == ' ' || line[nc] == '\t';
}

// First list-directed item after column `from`, read as a real.
bool list_real(const std::string& line, size_t from, double& v) {
  size_t i = from;
  while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == ',')) ++i;
  size_t j = i;
  while (j < line.size() && line[j] != ' ' && line[j] != '\t' && line[j] != ',' && line[j] != '/' &&
         line[j] != '!')
    ++j;
  std::string tok = line.substr(i, j - i);
  if (tok.empty()) return false;
  for (char& c : tok)
    if (c == 'D' || c == 'd') c = 'E';
  char* end = nullptr;
  v = std::strtod(tok.c_str(), &end);
  return *end == 0;
}

// A section: "KEY value" lines up to and including the next line starting with '#'.
// first_wins: ReadSection (the first occurrence of a key counts); otherwise
// ReadPatchSection (the last).  PATCH and NAME lines are handed to `special`.
struct Section {
  std::vector<const char*> keys;
  std::vector<double> vals;
  std::vector<bool> defined;
  explicit Section(std::initializer_list<const char*> k)
      : keys(k), vals(k.size(), 0.0), defined(k.size(), false) {}
  bool has(int i) const { return defined[i - 1]; }
  double v(int i) const { return vals[i - 1]; }
};

template <class Special>
bool read_section(Lines& L, Section& s, bool first_wins, Special special) {
  while (L.next()) {
    const std::string& line = L.line;
    if (!line.empty() && line[0] == '#') return true;
    int r = special(line);
    if (r < 0) return false;
    if (r > 0) continue;
    for (size_t k = 0; k < s.keys.size(); ++k) {
      if (first_wins && s.defined[k]) continue;
      if (!line_has_key(line, s.keys[k])) continue;
      double x;
      if (!list_real(line, std::strlen(s.keys[k]), x)) {
        g_err = L.where() + ": no value for " + s.keys[k];
        return false;
      }
      s.vals[k] = x;
      s.defined[k] = true;
      if (first_wins) break;
    }
  }
  g_err = L.path + ": section without its closing '#' line";
  return false;
}

bool read_plain(Lines& L, Section& s) {
  return read_section(L, s, true, [](const std::string&) { return 0; });
}

void inp_defaults(rflu_inp& p) {
  std::memset(&p, 0, sizeof(p));
  p.gas_model = 1;
  p.cfl = 3.0; p.discr = 1; p.order = 1; p.eps_entr = 0.05; p.diss_fact = 1.0;
  p.reconst = 2; p.dimens = 3;
  p.creconst_cells_weight = 1.0; p.creconst_faces_weight = 1.0;
  p.toler_ict = 1.0e-11;
  p.stencil_dimens_cells = p.stencil_dimens_faces = p.stencil_dimens_bfaces = 3;
  p.order_bfaces = 1;
  p.sigma = 1.0; p.kp = 0.25; p.ku = 0.75; p.minf = 0.10;
  p.flow_type = 0; p.solver_type = 1; p.rk_scheme = 1; p.start_step = 0;
  p.max_iter = 10000; p.write_iter = 1000; p.print_iter = 1;
  p.dt_imposed = 1.0e-5; p.dt_min_limit = 1.0e-10; p.max_time = 1.0e-4;
  p.print_time = 1.0e-12; p.write_time = 5.0e-5; p.res_tol = 1.0e-5;
  p.ref_velocity = 100.0; p.ref_pressure = 1.0e+5; p.ref_density = 1.2;
  p.ref_cp = 1004.5; p.ref_gamma = 1.4; p.ref_length = 1.0;
}

int format_code(double v) {
  if (v > 0.9 && v < 1.1) return 1;
  if (v > 1.9 && v < 2.1) return 2;
  if (v > 2.9 && v < 3.1) return 3;
  if (v > 3.9) return 4;
  return 0;
}

bool inp_numerics(Lines& L, rflu_inp& p) {
  Section s({"CFL", "DISCR", "ORDER", "ENTROPY", "DISSFACT", "RECONST", "DIMENS", "CRECONSTC",
             "CRECONSTF", "CRECONSTCW", "CRECONSTFW", "TOLERICT", "SDIMENSC", "SDIMENSF",
             "SDIMENSBF", "ORDERBF", "AXIFLAG", "SIGMA", "KP", "KU", "MINF"});
  if (!read_plain(L, s)) return false;
  if (s.has(1)) p.cfl = std::fabs(s.v(1));
  if (s.has(2)) p.discr = nint(s.v(2));
  if (s.has(3)) p.order = nint(s.v(3));
  if (s.has(4)) p.eps_entr = std::fabs(s.v(4));
  if (s.has(5)) p.diss_fact = std::fabs(s.v(5));
  if (s.has(6)) p.reconst = nint(s.v(6));
  if (s.has(7)) p.dimens = nint(s.v(7));
  if (s.has(8)) p.creconst_cells = nint(s.v(8));
  if (s.has(9)) p.creconst_faces = nint(s.v(9));
  if (s.has(10)) p.creconst_cells_weight = std::fabs(s.v(10));
  if (s.has(11)) p.creconst_faces_weight = std::fabs(s.v(11));
  if (s.has(12)) p.toler_ict = std::fabs(s.v(12));
  if (s.has(13)) p.stencil_dimens_cells = (int32_t)std::fabs(s.v(13));
  if (s.has(14)) p.stencil_dimens_faces = (int32_t)std::fabs(s.v(14));
  if (s.has(15)) p.stencil_dimens_bfaces = (int32_t)std::fabs(s.v(15));
  if (s.has(16)) p.order_bfaces = (int32_t)std::fabs(s.v(16));
  if (s.has(17) && s.v(17) == 1.0) p.axi_flag = 1;
  p.sigma = s.has(18) ? std::fabs(s.v(18)) : 1.0;
  p.kp = s.has(19) ? std::fabs(s.v(19)) : 0.25;
  p.ku = s.has(20) ? std::fabs(s.v(20)) : 0.75;
  p.minf = s.has(21) ? std::fabs(s.v(21)) : 0.10;
  return true;
}

bool inp_timestep(Lines& L, rflu_inp& p) {
  Section s({"FLOWTYPE", "TIMESTEP", "MAXTIME", "WRITIME", "PRNTIME", "MAXITER", "RESTOL",
             "WRIITER", "PRNITER", "START_STEP", "STARTITER", "SOLVERTYPE", "ORDER",
             "MAXSUBITER", "TOLSUBITER", "PREDICTSOL", "DTMINLIMIT", "RKSCHEME"});
  if (!read_plain(L, s)) return false;
  if (s.has(1)) p.flow_type = s.v(1) > 0.9 ? 1 : 0;
  if (s.has(2)) p.dt_imposed = std::fabs(s.v(2));
  if (s.has(3)) p.max_time = std::fabs(s.v(3));
  if (s.has(4)) p.write_time = std::fabs(s.v(4));
  p.print_time = s.has(5) ? std::fabs(s.v(5)) : p.write_time;
  if (s.has(6)) p.max_iter = iabs_round(s.v(6));
  if (s.has(7)) p.res_tol = std::fabs(s.v(7));
  if (s.has(8)) p.write_iter = std::max(1, iabs_round(s.v(8)));
  if (s.has(9)) p.print_iter = std::max(1, iabs_round(s.v(9)));
  if (s.has(12)) {
    int t = nint(s.v(12));
    p.solver_type = t == 1 ? 1 : t == 2 ? 2 : 3;
  }
  if (s.has(17)) p.dt_min_limit = std::fabs(s.v(17));
  if (s.has(18)) p.rk_scheme = iabs_round(s.v(18));
  if (s.has(10)) p.start_step = nint(s.v(10));
  return true;
}

bool inp_reference(Lines& L, rflu_inp& p) {
  Section s({"ABSVEL", "PRESS", "DENS", "CP", "GAMMA", "LENGTH"});
  if (!read_plain(L, s)) return false;
  if (s.has(1)) p.ref_velocity = std::fabs(s.v(1));
  if (s.has(2)) p.ref_pressure = std::fabs(s.v(2));
  if (s.has(3)) p.ref_density = std::fabs(s.v(3));
  if (s.has(4)) p.ref_cp = std::fabs(s.v(4));
  if (s.has(5)) p.ref_gamma = std::fabs(s.v(5));
  if (s.has(6)) p.ref_length = std::fabs(s.v(6));
  return true;
}

bool inp_initflow(Lines& L, rflu_inp& p) {
  static std::vector<std::string> names;
  if (names.empty()) {
    names = {"FLAG", "VELX", "VELY", "VELZ", "PRESS", "DENS", "TEMP"};
    for (int i = 1; i <= 19; ++i) names.push_back("IVAL" + std::to_string(i));
    for (int i = 1; i <= 25; ++i) names.push_back("RVAL" + std::to_string(i));
  }
  Section s({});
  for (const std::string& n : names) {
    s.keys.push_back(n.c_str());
    s.vals.push_back(0.0);
    s.defined.push_back(false);
  }
  if (!read_plain(L, s)) return false;
  if (s.has(1)) p.init_flag = nint(s.v(1));
  for (int d = 0; d < 3; ++d)
    if (s.has(2 + d)) p.init_vel[d] = s.v(2 + d);
  if (s.has(5)) p.init_press = s.v(5);
  if (s.has(6)) p.init_dens = s.v(6);
  if (s.has(7)) p.init_temp = s.v(7);
  for (int i = 0; i < 19; ++i)
    if (s.has(8 + i)) p.init_ival[i] = (int32_t)s.v(8 + i);
  for (int i = 0; i < 25; ++i)
    if (s.has(27 + i)) p.init_rval[i] = s.v(27 + i);
  return true;
}

int read_inp(const char* path, rflu_inp& p) {
  inp_defaults(p);
  Lines L;
  L.path = path;
  L.in.open(path);
  if (!L.in) { g_err = std::string("cannot open ") + path; return 7; }
  while (L.next()) {
    const std::string head = trim_right(L.line);
    bool ok = true;
    if (head == "# FORMATS") {
      Section s({"GRID", "SOLUTION", "GRIDSRC"});
      ok = read_plain(L, s);
      if (s.has(1)) p.grid_format = format_code(s.v(1));
      if (s.has(2)) p.solution_format = format_code(s.v(2));
      if (s.has(3)) p.grid_source = nint(s.v(3));
      p.sections |= RFLU_INP_FORMATS;
    } else if (head == "# FLOWMODEL") {
      Section s({"MODEL", "MOVEGRID"});
      ok = read_plain(L, s);
      if (s.has(1)) p.flow_model = nint(s.v(1));
      if (s.has(2) && (nint(s.v(2)) == 0 || nint(s.v(2)) == 1)) p.move_grid = nint(s.v(2));
      p.sections |= RFLU_INP_FLOWMODEL;
    } else if (head == "# MIXTURE") {
      Section s({"FROZENFLAG", "GASMODEL"});
      ok = read_plain(L, s);
      if (s.has(2)) p.gas_model = nint(s.v(2));
      p.sections |= RFLU_INP_MIXTURE;
    } else if (head == "# NUMERICS") {
      ok = inp_numerics(L, p);
      p.sections |= RFLU_INP_NUMERICS;
    } else if (head == "# TIMESTEP") {
      ok = inp_timestep(L, p);
      p.sections |= RFLU_INP_TIMESTEP;
    } else if (head == "# REFERENCE") {
      ok = inp_reference(L, p);
      p.sections |= RFLU_INP_REFERENCE;
    } else if (head == "# INITFLOW") {
      ok = inp_initflow(L, p);
      p.sections |= RFLU_INP_INITFLOW;
    }
    if (!ok) return 8;
  }
  return 0;
}

// ---------------------------------------------------------------------------------
// .bc

struct PatchRange { int beg = 1, end = 999999; std::string name = "None"; };

// ReadPatchSection's PATCH and NAME lines (libfloflu/ReadInputUtil.F90:454-534)
int patch_line(const Lines& L, const std::string& line, PatchRange& r) {
  if (line.compare(0, 5, "PATCH") == 0) {
    double a, b;
    size_t i = 5;
    if (!list_real(line, i, a)) { g_err = L.where() + ": PATCH needs two numbers"; return -1; }
    // the second item follows the first
    while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == ',')) ++i;
    while (i < line.size() && line[i] != ' ' && line[i] != '\t' && line[i] != ',') ++i;
    if (!list_real(line, i, b)) { g_err = L.where() + ": PATCH needs two numbers"; return -1; }
    r.beg = (int)a; r.end = (int)b;
    if (r.beg <= 0) r.beg = 1;
    if (r.end <= 0) r.end = 999999;
    if (r.end < r.beg) r.end = r.beg;
    return 1;
  }
  if (line.compare(0, 4, "NAME") == 0) {
    size_t i = 4;
    while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) ++i;
    size_t j = i;
    while (j < line.size() && line[j] != ' ' && line[j] != '\t' && line[j] != ',') ++j;
    r.name = line.substr(i, j - i);
    return 1;
  }
  if (line.compare(0, 7, "DISTRIB") == 0 || line.compare(0, 4, "FILE") == 0) return 1;
  return 0;
}

int space_order(const Section& s, int key, int order_bfaces) {
  if (!s.has(key)) return order_bfaces;
  return nint(s.v(key)) == 2 ? 2 : 1;
}

int read_bc(const char* path, int n_patches, int order_bfaces, rflu_bc_patch* out) {
  for (int p = 0; p < n_patches; ++p) {
    std::memset(&out[p], 0, sizeof(rflu_bc_patch));
    out[p].order = order_bfaces;
  }
  Lines L;
  L.path = path;
  L.in.open(path);
  if (!L.in) { g_err = std::string("cannot open ") + path; return 7; }
  const double pi = 4.0 * std::atan(1.0);
  const double deg2rad = 1.0 / (180.0 / pi);
  while (L.next()) {
    const std::string head = trim_right(L.line);
    if (head == "# END") break;
    int type = 0;
    std::unique_ptr<Section> sp;
    if (head == "# BC_SLIPW") {
      type = RFLU_BCT_SLIPWALL;
      sp.reset(new Section({"COUPLED", "MVPATCH", "SMGRID", "MOVEDIR", "STATS", "CRECONST", "KIND",
                            "THRUSTFLAG", "ORDER", "PARTWALL"}));
    } else if (head == "# BC_OUTFLOW") {
      type = RFLU_BCT_OUTFLOW;
      sp.reset(new Section({"TYPE", "PRESS", "MVPATCH", "SMGRID", "ORDER", "STATS", "MOVEDIR",
                            "COUPLED", "KIND", "REFLECT", "NSCBCK", "THRUSTFLAG"}));
    } else if (head == "# BC_INFLOW_VELTEMP") {
      type = RFLU_BCT_INFLOW_VELTEMP;
      sp.reset(new Section({"TYPE", "VELX", "VELY", "VELZ", "TEMP", "PRESS", "MVPATCH", "SMGRID",
                            "MOVEDIR", "COUPLED", "KIND", "REFLECT", "THRUSTFLAG", "ORDER"}));
    } else if (head == "# BC_INFLOW" || head == "# BC_INFLOW_TOTANG") {
      type = RFLU_BCT_INFLOW_TOTANG;
      sp.reset(new Section({"TYPE", "FIXED", "PTOT", "TTOT", "BETAH", "BETAV", "MACH", "MVPATCH",
                            "SMGRID", "MOVEDIR", "COUPLED", "KIND", "REFLECT", "THRUSTFLAG",
                            "ORDER"}));
    } else if (head == "# BC_FARF") {
      type = RFLU_BCT_FARFIELD;
      sp.reset(new Section({"CORR", "MACH", "ATTACK", "SLIP", "PRESS", "TEMP", "MVPATCH", "SMGRID",
                            "ORDER", "MOVEDIR", "COUPLED", "KIND", "THRUSTFLAG"}));
    } else if (head == "# BC_INJECT") {
      type = RFLU_BCT_INJECTION;
      sp.reset(new Section({"MFRATE", "TEMP", "COUPLED", "BFLAG", "MVPATCH", "SMGRID", "MOVEDIR",
                            "STATS", "CRECONST", "KIND", "THRUSTFLAG", "ORDER"}));
    } else if (head == "# BC_PERIODIC") {
      type = RFLU_BCT_PERIODIC;
      sp.reset(new Section({"RELPATCH", "ANGLE", "AXIS", "MVPATCH", "SMGRID"}));
    } else if (head == "# BC_SYMMETRY") {
      type = RFLU_BCT_SYMMETRY;
      sp.reset(new Section({"MVPATCH", "SMGRID"}));
    } else if (head == "# BC_NOSLIP" || head == "# BC_NOSLIP_HFLUX") {
      type = RFLU_BCT_NOSLIPWALL_HFLUX;
      sp.reset(new Section({"HFLUX", "ORDER"}));
    } else if (head == "# BC_NOSLIP_TEMP") {
      type = RFLU_BCT_NOSLIPWALL_TEMP;
      sp.reset(new Section({"TEMP", "ORDER"}));
    } else if (head == "# BC_VIRTUAL") {
      type = RFLU_BCT_VIRTUAL;
      sp.reset(new Section({"MVPATCH", "SMGRID"}));
    } else {
      continue;
    }
    Section& s = *sp;
    PatchRange r;
    const long head_line = L.lineno;
    if (!read_section(L, s, false, [&](const std::string& line) { return patch_line(L, line, r); }))
      return 8;
    const std::string where = L.path + ":" + std::to_string(head_line);
    if (r.end > n_patches && r.end != 999999) {
      g_err = where + ": patch range ends at " + std::to_string(r.end) + " but the grid has " +
              std::to_string(n_patches) + " patches";
      return 8;
    }
    auto missing = [&](const char* what) {
      g_err = where + ": " + head.substr(2) + " section without " + what;
      return 8;
    };
    for (int ip = r.beg; ip <= std::min(r.end, n_patches); ++ip) {
      rflu_bc_patch& b = out[ip - 1];
      std::memset(&b, 0, sizeof(b));
      b.bc_type = type;
      std::strncpy(b.name, r.name.c_str(), sizeof(b.name) - 1);
      b.order = order_bfaces;
      switch (type) {
        case RFLU_BCT_SLIPWALL:
          b.order = space_order(s, 9, order_bfaces);
          b.creconst = s.has(6) && nint(s.v(6)) == 1 ? 1 : 0;
          break;
        case RFLU_BCT_OUTFLOW: {
          if (!s.has(1) || nint(s.v(1)) < 0 || nint(s.v(1)) > 2) return missing("TYPE 0, 1 or 2");
          b.opt = nint(s.v(1));
          if (b.opt != 0) {
            if (!s.has(2)) return missing("PRESS");
            b.pout = s.v(2);
          }
          b.order = space_order(s, 5, order_bfaces);
          break;
        }
        case RFLU_BCT_INFLOW_VELTEMP: {
          if (!s.has(1) || nint(s.v(1)) < 0 || nint(s.v(1)) > 2) return missing("TYPE 0, 1 or 2");
          b.opt = nint(s.v(1));
          const int ndata = b.opt != 1 ? 5 : 4;
          for (int i = 0; i < ndata; ++i)
            if (!s.has(2 + i)) return missing("VELX, VELY, VELZ, TEMP (and PRESS unless TYPE 1)");
          for (int i = 0; i < 4; ++i) b.vals[i] = s.v(2 + i);
          if (b.opt != 1) b.vals[4] = s.v(6);
          b.order = space_order(s, 14, order_bfaces);
          break;
        }
        case RFLU_BCT_INFLOW_TOTANG: {
          if (!s.has(1) || nint(s.v(1)) < 0 || nint(s.v(1)) > 2) return missing("TYPE 0, 1 or 2");
          b.opt = nint(s.v(1));
          b.opt2 = s.has(2) && nint(s.v(2)) >= 0 && nint(s.v(2)) <= 1 ? nint(s.v(2)) : 0;
          const int ndata = b.opt != 1 ? 5 : 4;
          for (int i = 0; i < ndata; ++i)
            if (!s.has(3 + i)) return missing("PTOT, TTOT, BETAH, BETAV (and MACH unless TYPE 1)");
          b.vals[0] = s.v(3);
          b.vals[1] = s.v(4);
          b.vals[2] = s.v(5) * deg2rad;
          b.vals[3] = s.v(6) * deg2rad;
          if (b.opt != 1) b.vals[4] = s.v(7);
          b.order = space_order(s, 15, order_bfaces);
          break;
        }
        case RFLU_BCT_FARFIELD: {
          if (!s.has(1) || (nint(s.v(1)) != 0 && nint(s.v(1)) != 1)) return missing("CORR 0 or 1");
          b.opt2 = nint(s.v(1));
          for (int i = 0; i < 5; ++i)
            if (!s.has(2 + i)) return missing("MACH, ATTACK, SLIP, PRESS, TEMP");
          b.vals[0] = s.v(2);
          b.vals[1] = s.v(3) * deg2rad;
          b.vals[2] = s.v(4) * deg2rad;
          b.vals[3] = s.v(5);
          b.vals[4] = s.v(6);
          b.order = space_order(s, 9, order_bfaces);
          break;
        }
        case RFLU_BCT_INJECTION:
          if (!s.has(1) || !s.has(2)) return missing("MFRATE, TEMP");
          b.vals[0] = s.v(1);
          b.vals[1] = s.v(2);
          b.creconst = s.has(9) && nint(s.v(9)) == 1 ? 1 : 0;
          b.order = space_order(s, 12, order_bfaces);
          break;
        case RFLU_BCT_PERIODIC:
          if (!s.has(1) || !s.has(2) || !s.has(3)) return missing("RELPATCH, ANGLE, AXIS");
          b.related = nint(s.v(1));
          b.angle = s.v(2) * deg2rad;
          b.axis = nint(s.v(3));
          break;
        default:
          break;
      }
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int rflu_read_inp(const char* path, rflu_inp* out) {
  if (!path || !out) { g_err = "null argument"; return 1; }
  try {
    return read_inp(path, *out);
  } catch (const std::exception& e) {
    g_err = e.what();
    return 8;
  }
}

int rflu_read_bc(const char* path, int n_patches, int order_bfaces, rflu_bc_patch* out) {
  if (!path || !out || n_patches < 1) { g_err = "bad argument"; return 1; }
  try {
    return read_bc(path, n_patches, order_bfaces, out);
  } catch (const std::exception& e) {
    g_err = e.what();
    return 8;
  }
}

const char* rflu_input_last_error(void) { return g_err.c_str(); }

}  // extern "C"
