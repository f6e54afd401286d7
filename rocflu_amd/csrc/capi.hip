// C-ABI of the explicit residual/update path (include/rocflu_amd.h): region
// registry, upload of the t_region arrays into the device layout, and the launch
// sequences of RungeKuttaMP's callees.
#include "rocflu_amd.h"
#include "kernels.hpp"
#include "adjoint.hpp"
#include "dump.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

using namespace rflu;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(RFLU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

enum KernelId { KID_FACE = 0, KID_UPDATE = 1, KID_GRAD = 2, KID_BLEND = 3, KID_DT = 4, KID_BFACE = 5, KID_N = 6 };

// One border of a region (t_border): the actual cells it sends to a neighbouring
// region and the virtual cells it receives from it, as internal cell numbers.
enum { FBUF_F64 = 0, FBUF_MIX = 1 };

struct Border {
  int nbr = 0;          // neighbour's handle (the region itself for symmetry/periodic)
  int nbr_border = 0;   // the neighbour's border that matches this one
  int nsend = 0, nrecv = 0;
  int *send = nullptr, *recv = nullptr;
  bool turn = false;    // data received through this border is turned by xf
  Xform xf;
};

struct Region {
  int N = 0, F = 0, B = 0, P = 0;
  int nact = 0;                // actual cells; [nact, N) are virtual cells
  int base = 0;                // the descriptor's index base (border cell lists)
  std::vector<Border> borders;
  int order = 1, discr = 3, reconst = 0, rk = 2, prec = 0;
  int fbuf = 0;                // flux buffer: FBUF_F64 or FBUF_MIX (FluxMix)
  double cfl = 0.5, dt_imposed = 0, dt_min_limit = 0;
  double end_time = HUGE_VAL;  // global%maxTime of the fused loop
  Gas gas{};                   // case 0 (the only case of a single region)
  bool has_flowthrough = false;
  // batched cases (K == 1: a single region; per-case tables still hold its values)
  int K = 1;
  std::vector<int> cell0_h, bface0_h;
  std::vector<CaseParams> par_h;
  unsigned short* cid = nullptr;
  Gas* gas_tab = nullptr;
  CaseParams* par = nullptr;
  int *cell0 = nullptr, *bface0 = nullptr;
  unsigned long long* dtmin = nullptr;
  int *pbc = nullptr, *popt = nullptr, *popt2 = nullptr, *porder2 = nullptr;
  double *ppout = nullptr, *pvals = nullptr;
  std::vector<double> pvals_h;
  std::vector<int> bptr_h;     // first boundary face of every patch (all cases), P+1
  int* bptr = nullptr;
  // constrained gradients (CRECONSTC): (cell, variable) entries and their rows
  int con_n = 0, con_R = 0;
  int *con_ecell = nullptr, *con_evar = nullptr, *con_rptr = nullptr, *con_rcell = nullptr;
  double *con_rval = nullptr, *con_rw = nullptr;
  // derivative work buffers (allocated by the first tangent/adjoint call)
  double *adj_g = nullptr, *adj_gb = nullptr, *adj_gv = nullptr, *adj_pv = nullptr;
  double *adj_dflux = nullptr, *adj_dvals = nullptr;
  // second order: primitive-state tangent/adjoint [5][N], gradient values/tangents and
  // adjoints [15][N], limiter selections [5][N], the transposed stencil (CSR)
  double *adj_p = nullptr, *adj_ga = nullptr, *adj_gb2 = nullptr, *adj_gc = nullptr, *adj_gr = nullptr;
  double *adj_selv = nullptr;
  int *adj_seln = nullptr, *adj_tptr = nullptr, *adj_tidx = nullptr;
  double *adj_wpre = nullptr, *adj_wsum = nullptr;   // WENO-XYZ blend and normalisers
  // constrained gradients: entries of every cell and the rows that read every cell (CSR)
  int *adj_cptr = nullptr, *adj_cent = nullptr, *adj_ctptr = nullptr, *adj_ctrow = nullptr,
      *adj_ctent = nullptr;
  hipStream_t stream = nullptr;
  // device buffers
  int2* f2c = nullptr;
  double4 *fn = nullptr, *fc = nullptr, *cell = nullptr, *bfn = nullptr, *bfc = nullptr;
  int *bf2c = nullptr, *bpatch = nullptr, *c2f_ptr = nullptr, *c2f_idx = nullptr;
  int *sten_ptr = nullptr, *sten_idx = nullptr;
  double *moms = nullptr, *epsq = nullptr;
  int* vnb = nullptr;          // limiter neighbours [vq][N]
  double* vgeo = nullptr;      // limiter cofg - fc [3*vq][N]
  int vq = 0;
  double* srad = nullptr;      // face spectral radii [F+B] of the time step
  double *cv = nullptr, *cv_old = nullptr, *rhs_sum = nullptr, *dv = nullptr, *prim = nullptr;
  void *grad = nullptr, *grad2 = nullptr, *flux = nullptr;
  double* rhs = nullptr;
  double* dt_part = nullptr;
  Scalars* sc = nullptr;
  // internal cell numbering: device cell i holds the caller's cell new2old[i]
  std::vector<int> new2old;
  // stencil tiles staged in LDS by the gradient kernel
  int *tile_uptr = nullptr, *tile_uidx = nullptr, *tile_lptr = nullptr;
  unsigned short* tile_loc = nullptr;
  unsigned short* tile_vloc = nullptr;
  size_t tile_loc_len = 0;    // entries of the [member][lane] slot table
  void* gw = nullptr;         // GRAD_WEIGHTS 1: stencil weights [3][tile_loc_len];
                              // 2: inverse distances [tile_loc_len]
  int gw_kind = 0;
  int n_tiles = 0, tile_cap = 0;
  int grad_tile = 256;        // tuning: cells per LDS tile (128, 256 or 512)
  // face tiles of the second-order interior flux (FACE_TILE; 0 = one thread per face)
  int face_tile = 0, n_face_tiles = 0, face_cap = 0, face_fcap = 0;
  bool update_tile = false;   // tuning: the update reads its tile's face fluxes from LDS
  int *ft_fptr = nullptr, *ft_eptr = nullptr, *ft_eidx = nullptr;
  unsigned* ft_fslot = nullptr;
  bool use_lds = false;       // tiles were built (every tile fits in LDS)
  int grad_lds = 1;           // tuning: stage stencils in LDS when tiles exist (2: the
                              // mixed mode's FP32 records)
  int grad_unroll = 4;        // tuning: stencil members whose weights are formed together
  int face_threads = 256;     // tuning: threads per face-tile block (256 or 512)
  bool grad_fast = false;     // tuning: moment form of the gradient (LDS tiles; not bit-identical)
  // per-callee path: the update asked for by rflu_gpu_rk_update (stage, dt) and the
  // dependent variables of rflu_gpu_set_dependent_vars, not launched yet; the next
  // rflu_gpu_after_update launches them with its checks as one fused update, any other
  // entry point launches them separately first (flush_pending): the same bits
  int pend_stage = 0;
  double pend_dt = 0.0;
  bool pend_dv = false;
  int nxcd = 8;               // tuning: XCD-aware block order (MI355X: 8 XCDs); 1 = off
  int n_dt_parts = 0;
  int face_block = 256, cell_block = 256, grad_block = 256;
  hipGraphExec_t graph = nullptr;
  bool graph_dirty = true;
  // kernel timing
  bool profiling = false;
  // PERSISTENT: one launch per time step for a batch (k_step_persistent); its grid
  // barrier words (count, generation, error) and the blocks that fit the device at once
  bool persistent = false;
  unsigned* pbar = nullptr;
  int pers_blocks = 0;
  size_t pers_shm = 0;
  int pers_key = -1;
  struct Ev { int kid; hipEvent_t a, b; };
  std::vector<Ev> events;
  double kern_ms[KID_N] = {};
  long long kern_n[KID_N] = {};

  Mesh mesh() const {
    Mesh m;
    m.N = N; m.F = F; m.B = B; m.nact = nact;
    m.f2c = f2c; m.fn = fn; m.fc = fc; m.cell = cell;
    m.bf2c = bf2c; m.bfn = bfn; m.bfc = bfc; m.bpatch = bpatch;
    m.c2f_ptr = c2f_ptr; m.c2f_idx = c2f_idx;
    m.sten_ptr = sten_ptr; m.sten_idx = sten_idx; m.moms = moms; m.epsq = epsq;
    m.vq = vq; m.vnb = vnb; m.vgeo = vgeo; m.nxcd = nxcd;
    return m;
  }
  State state() const {
    State s;
    s.cv = cv; s.cv_old = cv_old; s.rhs_sum = rhs_sum; s.dv = dv; s.prim = prim;
    return s;
  }
  const void* grad_final() const { return reconst == RFLU_RECONST_WENO_XYZ ? grad2 : grad; }
  Cases cases() const {
    Cases c;
    c.K = K; c.cid = K > 1 ? cid : nullptr; c.gas = gas_tab; c.sc = sc;
    c.cell0 = cell0; c.bface0 = bface0; c.par = par; c.dtmin = dtmin;
    return c;
  }
  Patches patches() const {
    Patches p;
    p.bc = pbc; p.opt = popt; p.opt2 = popt2; p.order2 = porder2; p.pout = ppout; p.vals = pvals;
    return p;
  }
  // scalars that gate a whole launch: the single case's, none for a batch
  const Scalars* gate_sc() const { return K == 1 ? sc : nullptr; }
  Constr constr() const {
    Constr c;
    c.n = con_n; c.ecell = con_ecell; c.evar = con_evar; c.rptr = con_rptr;
    c.rcell = con_rcell; c.rval = con_rval; c.rw = con_rw; c.R = con_R;
    return c;
  }
  ConstrD constr_d() const {
    ConstrD d;
    d.C = constr();
    d.cptr = con_n > 0 ? adj_cptr : nullptr;
    d.cent = adj_cent; d.trptr = adj_ctptr; d.trow = adj_ctrow; d.tent = adj_ctent;
    return d;
  }
  CaseParams case0() const {
    CaseParams c;
    c.cfl = cfl; c.dt_imposed = dt_imposed; c.dt_min_limit = dt_min_limit; c.end_time = end_time;
    return c;
  }
};

std::mutex g_mu;
std::vector<Region*> g_regions;

Region* get(int32_t h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (h < 1 || h > (int32_t)g_regions.size()) return nullptr;
  return g_regions[h - 1];
}

template <typename T>
int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(RFLU_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}
template <typename T>
int dput(T* d, const std::vector<T>& h, hipStream_t s) {
  if (h.empty()) return 0;
  HIP_TRY(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return 0;
}

void free_region(Region* R) {
  if (!R) return;
  if (R->stream) (void)hipStreamSynchronize(R->stream);
  if (R->graph) (void)hipGraphExecDestroy(R->graph);
  for (auto& e : R->events) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  void* bufs[] = {R->f2c, R->fn, R->fc, R->cell, R->bfn, R->bfc, R->bf2c, R->bpatch, R->c2f_ptr,
                  R->c2f_idx, R->sten_ptr, R->sten_idx, R->moms, R->epsq, R->cv, R->cv_old,
                  R->rhs_sum, R->dv, R->prim, R->grad, R->grad2, R->flux, R->rhs, R->dt_part,
                  R->sc, R->tile_uptr, R->tile_uidx, R->tile_lptr, R->tile_loc, R->tile_vloc,
                  R->ft_fptr, R->ft_eptr, R->ft_eidx, R->ft_fslot,
                  R->vnb, R->vgeo, R->srad, R->gw, R->cid, R->gas_tab, R->par, R->cell0,
                  R->bface0, R->dtmin, R->pbc, R->popt, R->popt2, R->porder2, R->ppout, R->pvals,
                  R->bptr, R->adj_g, R->adj_gb, R->adj_gv, R->adj_pv, R->adj_dflux,
                  R->adj_dvals, R->adj_p, R->adj_ga, R->adj_gb2, R->adj_gc, R->adj_gr, R->adj_selv,
                  R->adj_seln, R->adj_tptr, R->adj_tidx, R->adj_wpre, R->adj_wsum,
                  R->adj_cptr, R->adj_cent, R->adj_ctptr, R->adj_ctrow, R->adj_ctent,
                  R->con_ecell, R->con_evar, R->con_rptr, R->con_rcell,
                  R->con_rval, R->con_rw, R->pbar};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (Border& b : R->borders) {
    if (b.send) (void)hipFree(b.send);
    if (b.recv) (void)hipFree(b.recv);
  }
  if (R->stream) (void)hipStreamDestroy(R->stream);
  delete R;
}

inline unsigned nblk(int n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- renumbering -------------------------------------------------------------------

inline uint64_t spread3(uint64_t x) {
  x &= 0x1fffffULL;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

// Cells of one case [c0, c1) sorted along a Morton curve of their centroids (ties
// keep the caller's order), so that consecutive cells -- one wavefront, one LDS tile
// -- are compact in space.  Cases keep their ranges.
void morton_order(const std::vector<double4>& cell, int c0, int c1, std::vector<int>& n2o) {
  const int N = c1 - c0;
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int i = c0; i < c1; ++i) {
    const double4& c = cell[i];
    lo[0] = std::min(lo[0], c.x); hi[0] = std::max(hi[0], c.x);
    lo[1] = std::min(lo[1], c.y); hi[1] = std::max(hi[1], c.y);
    lo[2] = std::min(lo[2], c.z); hi[2] = std::max(hi[2], c.z);
  }
  std::vector<std::pair<uint64_t, int>> key(N);
  for (int i = 0; i < N; ++i) {
    const double p[3] = {cell[c0 + i].x, cell[c0 + i].y, cell[c0 + i].z};
    uint64_t q[3];
    for (int d = 0; d < 3; ++d) {
      double ext = hi[d] - lo[d];
      double s = ext > 0 ? (p[d] - lo[d]) / ext : 0.0;
      q[d] = (uint64_t)(s * 2097151.0);
    }
    key[i] = {spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2), c0 + i};
  }
  std::sort(key.begin(), key.end());
  for (int i = 0; i < N; ++i) n2o[c0 + i] = key[i].second;
}

void renumber(Region* R, std::vector<double4>& hcell, std::vector<int2>& hf2c,
              std::vector<double4>& hfn, std::vector<double4>& hfc, std::vector<int>& hbf2c,
              std::vector<int>& hc2f_ptr, std::vector<int>& hc2f_idx, std::vector<int>& hsten_ptr,
              std::vector<int>& hsten_idx, std::vector<double>& hmoms, std::vector<double>& hepsq) {
  const int N = R->N, F = R->F;
  std::vector<int> n2o(N), o2n(N);
  if (R->nact < N) {
    // virtual cells stay after the actual ones
    morton_order(hcell, 0, R->nact, n2o);
    morton_order(hcell, R->nact, N, n2o);
  } else {
    for (int c = 0; c < R->K; ++c) morton_order(hcell, R->cell0_h[c], R->cell0_h[c + 1], n2o);
  }
  for (int i = 0; i < N; ++i) o2n[n2o[i]] = i;
  {
    std::vector<double4> c2(N);
    for (int i = 0; i < N; ++i) c2[i] = hcell[n2o[i]];
    hcell.swap(c2);
  }
  // faces ordered by their lower new cell index
  std::vector<int> fn2o(F), fo2n(F);
  {
    std::vector<std::pair<int64_t, int>> key(F);
    for (int f = 0; f < F; ++f) {
      int a = o2n[hf2c[f].x], b = o2n[hf2c[f].y];
      key[f] = {((int64_t)std::min(a, b) << 32) | (int64_t)std::max(a, b), f};
    }
    std::sort(key.begin(), key.end());
    for (int k = 0; k < F; ++k) { fn2o[k] = key[k].second; fo2n[key[k].second] = k; }
    std::vector<int2> f2(F);
    std::vector<double4> n2(F), c2(F);
    for (int k = 0; k < F; ++k) {
      int2 o = hf2c[fn2o[k]];
      f2[k] = make_int2(o2n[o.x], o2n[o.y]);
      n2[k] = hfn[fn2o[k]];
      c2[k] = hfc[fn2o[k]];
    }
    hf2c.swap(f2); hfn.swap(n2); hfc.swap(c2);
  }
  for (int& c : hbf2c) c = o2n[c];
  {
    std::vector<int> p2(N + 1, 0), i2(hc2f_idx.size());
    for (int i = 0; i < N; ++i) p2[i + 1] = p2[i] + (hc2f_ptr[n2o[i] + 1] - hc2f_ptr[n2o[i]]);
    for (int i = 0; i < N; ++i) {
      int o = n2o[i];
      for (int e = hc2f_ptr[o], k = p2[i]; e < hc2f_ptr[o + 1]; ++e, ++k) {
        int s = hc2f_idx[e];
        if (s > F) i2[k] = s;                          // boundary face: unchanged
        else if (s > 0) i2[k] = fo2n[s - 1] + 1;
        else i2[k] = -(fo2n[-s - 1] + 1);
      }
    }
    hc2f_ptr.swap(p2); hc2f_idx.swap(i2);
  }
  if (R->order == 2) {
    std::vector<int> p2(N + 1, 0), i2(hsten_idx.size());
    for (int i = 0; i < N; ++i) p2[i + 1] = p2[i] + (hsten_ptr[n2o[i] + 1] - hsten_ptr[n2o[i]]);
    for (int i = 0; i < N; ++i) {
      int o = n2o[i];
      for (int e = hsten_ptr[o], k = p2[i]; e < hsten_ptr[o + 1]; ++e, ++k) i2[k] = o2n[hsten_idx[e]];
    }
    hsten_ptr.swap(p2); hsten_idx.swap(i2);
    std::vector<double> m2(hmoms.size()), e2(N);
    for (int k = 0; k < 6; ++k)
      for (int i = 0; i < N; ++i) m2[(size_t)k * N + i] = hmoms[(size_t)k * N + n2o[i]];
    for (int i = 0; i < N; ++i) e2[i] = hepsq[n2o[i]];
    hmoms.swap(m2); hepsq.swap(e2);
  }
  R->new2old.swap(n2o);
}

// LDS tiles of the gradient kernel: tile t covers cells [t*TS, t*TS+TS) (TS =
// R->grad_tile, one thread per cell); its union of stencil members (and, with the
// limiter, face neighbours) is staged in LDS; loc[lptr[t] + s*TS + lane] is the LDS
// slot of the lane's s-th stencil member (member order unchanged) and
// vloc[(t*vq + q)*TS + lane] that of its q-th limiter neighbour.  A tile's records
// (8 doubles each when the weights are formed in the kernel) must fit the 160 KiB of
// LDS of a gfx950 CU.
constexpr int GRAD_TILE_CAP = 2560;

void build_tiles(Region* R, const std::vector<int>& sp, const std::vector<int>& si,
                 const std::vector<int>& vnb, bool venkat, std::vector<int>& uptr,
                 std::vector<int>& uidx, std::vector<int>& lptr, std::vector<unsigned short>& loc,
                 std::vector<unsigned short>& vloc) {
  const int N = R->N, TS = R->grad_tile, VQ = R->vq;
  const int nt = (N + TS - 1) / TS;
  auto give_up = [&]() {
    R->use_lds = false;
    uptr.clear(); uidx.clear(); lptr.clear(); loc.clear(); vloc.clear();
  };
  // tiles are independent: unions and slot tables in parallel, then concatenated
  std::vector<std::vector<int>> tu(nt);
  std::vector<std::vector<unsigned short>> tl(nt);
  vloc.assign(venkat ? (size_t)nt * VQ * TS : 1, (unsigned short)0xFFFF);
  std::vector<char> too_big(nt, 0);
  auto work = [&](int t0, int t1) {
    std::vector<int> u;
    for (int t = t0; t < t1; ++t) {
      int c0 = t * TS, c1 = std::min(N, c0 + TS);
      u.clear();
      int smax = 0;
      for (int i = c0; i < c1; ++i) {
        smax = std::max(smax, sp[i + 1] - sp[i]);
        for (int e = sp[i]; e < sp[i + 1]; ++e) u.push_back(si[e]);
        if (venkat)
          for (int q = 0; q < VQ && vnb[(size_t)q * N + i] >= 0; ++q) u.push_back(vnb[(size_t)q * N + i]);
      }
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      if ((int)u.size() > GRAD_TILE_CAP) {   // does not fit: gradient kernel reads global memory
        too_big[t] = 1;
        continue;
      }
      auto slot = [&](int j) { return (unsigned short)(std::lower_bound(u.begin(), u.end(), j) - u.begin()); };
      std::vector<unsigned short>& L = tl[t];
      L.assign((size_t)smax * TS, (unsigned short)0);
      for (int i = c0; i < c1; ++i) {
        int lane = i - c0;
        for (int e = sp[i], s = 0; e < sp[i + 1]; ++e, ++s) L[(size_t)s * TS + lane] = slot(si[e]);
        if (venkat)
          for (int q = 0; q < VQ && vnb[(size_t)q * N + i] >= 0; ++q)
            vloc[((size_t)t * VQ + q) * TS + lane] = slot(vnb[(size_t)q * N + i]);
      }
      tu[t] = u;
    }
  };
  int nth = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  nth = std::max(1, std::min(nth, nt / 64));
  if (nth <= 1) {
    work(0, nt);
  } else {
    std::vector<std::thread> pool;
    for (int k = 0; k < nth; ++k)
      pool.emplace_back(work, (int)((long long)nt * k / nth), (int)((long long)nt * (k + 1) / nth));
    for (auto& th : pool) th.join();
  }
  for (int t = 0; t < nt; ++t)
    if (too_big[t]) {
      give_up();
      return;
    }
  uptr.assign(nt + 1, 0);
  lptr.assign(nt + 1, 0);
  size_t nu = 0, nl = 0;
  int cap = 0;
  for (int t = 0; t < nt; ++t) {
    nu += tu[t].size();
    nl += tl[t].size();
    cap = std::max(cap, (int)tu[t].size());
  }
  if (nu > (size_t)0x7fffffff || nl > (size_t)0x7fffffff) {
    give_up();
    return;
  }
  uidx.clear();
  uidx.reserve(nu);
  loc.clear();
  loc.reserve(nl);
  for (int t = 0; t < nt; ++t) {
    uidx.insert(uidx.end(), tu[t].begin(), tu[t].end());
    loc.insert(loc.end(), tl[t].begin(), tl[t].end());
    uptr[t + 1] = (int)uidx.size();
    lptr[t + 1] = (int)loc.size();
    std::vector<int>().swap(tu[t]);
    std::vector<unsigned short>().swap(tl[t]);
  }
  // rows of zeros after the last tile: the kernels' prefetch loads read up to a group
  // (at most 8 members) past a lane's last member, or row 0 whatever its member count,
  // so every such read stays in bounds
  loc.resize(loc.size() + 8 * 512, (unsigned short)0);
  R->n_tiles = nt;
  R->tile_cap = cap;
  R->use_lds = true;
}

// Limiter neighbourhood of every cell from its cell-to-face list: the cell across
// each interior face (in list order) and cofg - fc of that face.
void build_limiter_geometry(Region* R, const std::vector<int>& c2f_ptr,
                            const std::vector<int>& c2f_idx, const std::vector<int2>& f2c,
                            const std::vector<double4>& fc, const std::vector<double4>& cell,
                            std::vector<int>& vnb, std::vector<double>& vgeo) {
  const int N = R->N, F = R->F;
  int vq = 1;
  for (int i = 0; i < N; ++i) {
    int n = 0;
    for (int e = c2f_ptr[i]; e < c2f_ptr[i + 1]; ++e) {
      int s = c2f_idx[e];
      if ((s > 0 ? s : -s) <= F) ++n;
    }
    vq = std::max(vq, n);
  }
  R->vq = vq;
  vnb.assign((size_t)vq * N, -1);
  vgeo.assign((size_t)3 * vq * N, 0.0);
  for (int i = 0; i < N; ++i) {
    int q = 0;
    for (int e = c2f_ptr[i]; e < c2f_ptr[i + 1]; ++e) {
      int s = c2f_idx[e];
      int f = (s > 0 ? s : -s) - 1;
      if (f >= F) continue;
      vnb[(size_t)q * N + i] = s > 0 ? f2c[f].y : f2c[f].x;
      vgeo[(size_t)(3 * q) * N + i] = cell[i].x - fc[f].x;
      vgeo[(size_t)(3 * q + 1) * N + i] = cell[i].y - fc[f].y;
      vgeo[(size_t)(3 * q + 2) * N + i] = cell[i].z - fc[f].z;
      ++q;
    }
  }
}

// ---- kernel dispatch ------------------------------------------------------------

// LDS bytes of a face tile's records (cofg, primitive state, [sound speed,] gradient).
template <typename GT>
size_t face_tile_lds(int cap, bool srad) {
  return (size_t)cap * ((srad ? 9 : 8) * sizeof(double) + 15 * sizeof(GT));
}

template <typename T, typename GT, typename FT, bool SR>
int launch_fluxes_t(Region* R) {
  Mesh M = R->mesh();
  State S = R->state();
  const GT* g = (const GT*)R->grad_final();
  FT* fl = (FT*)R->flux;
  dim3 fb(R->face_block);
  if (R->F > 0 && R->order == 2 && R->face_tile > 0) {
    FaceTiles FTl;
    FTl.fptr = R->ft_fptr; FTl.eptr = R->ft_eptr; FTl.eidx = R->ft_eidx; FTl.fslot = R->ft_fslot;
    FTl.ts = R->face_tile; FTl.cap = R->face_cap; FTl.fcap = R->face_fcap;
    const size_t shm = face_tile_lds<GT>(R->face_cap, SR);
#define RFLU_FACE_TILE(D, NT)                                                                    \
  hipLaunchKernelGGL((k_face_flux_tile<D, T, GT, FT, SR, NT>), dim3(R->n_face_tiles), dim3(NT),    \
                     shm, R->stream, M, S, R->gas, R->cases(), FTl, g, fl, R->srad, R->gate_sc())
    if (R->face_threads == 512) {
      if (R->discr == 1) RFLU_FACE_TILE(1, 512);
      else if (R->discr == 3) RFLU_FACE_TILE(3, 512);
      else RFLU_FACE_TILE(4, 512);
    } else {
      if (R->discr == 1) RFLU_FACE_TILE(1, 256);
      else if (R->discr == 3) RFLU_FACE_TILE(3, 256);
      else RFLU_FACE_TILE(4, 256);
    }
#undef RFLU_FACE_TILE
    return 0;
  }
  if (R->F > 0) {
#define RFLU_FACE(D, O)                                                                            \
  hipLaunchKernelGGL((k_face_flux<D, O, T, GT, FT, SR>), dim3(nblk(R->F, R->face_block)), fb, 0, \
                     R->stream, M, S, R->gas, R->cases(), g, fl, R->srad, R->gate_sc())
    if (R->order == 1) {
      if (R->discr == 1) RFLU_FACE(1, 1);
      else if (R->discr == 3) RFLU_FACE(3, 1);
      else RFLU_FACE(4, 1);
    } else {
      if (R->discr == 1) RFLU_FACE(1, 2);
      else if (R->discr == 3) RFLU_FACE(3, 2);
      else RFLU_FACE(4, 2);
    }
#undef RFLU_FACE
  }
  return 0;
}

template <typename GT, typename FT, bool SR>
int launch_bfaces_t(Region* R) {
  if (R->B == 0) return 0;
  Mesh M = R->mesh();
  State S = R->state();
  const GT* g = (const GT*)R->grad_final();
  if (R->order == 1)
    hipLaunchKernelGGL((k_bface_flux<1, GT, FT, SR>), dim3(nblk(R->B, 256)), dim3(256), 0, R->stream,
                       M, S, R->gas, R->cases(), R->patches(), g, (FT*)R->flux, R->srad, R->gate_sc());
  else
    hipLaunchKernelGGL((k_bface_flux<2, GT, FT, SR>), dim3(nblk(R->B, 256)), dim3(256), 0, R->stream,
                       M, S, R->gas, R->cases(), R->patches(), g, (FT*)R->flux, R->srad, R->gate_sc());
  return 0;
}

Tiles region_tiles(const Region* R) {
  Tiles T;
  T.uptr = R->tile_uptr; T.uidx = R->tile_uidx; T.lptr = R->tile_lptr; T.loc = R->tile_loc;
  T.vloc = R->tile_vloc;
  T.ts = R->grad_tile; T.cap = R->tile_cap;
  return T;
}

// The formed-weight tile kernel with the LDS stride capc (0: the runtime one).
template <typename GT, bool V, int U, int TS, bool TW>
void launch_gradient_lds(Region* R, const Mesh& M, const State& S, const Tiles& T, int capc,
                         size_t shm) {
  const GT* tw = TW ? (const GT*)R->gw : nullptr;
#define RFLU_LDS(C)                                                                           \
  hipLaunchKernelGGL((k_gradient_lds<GT, V, U, TS, C, TW>), dim3(R->n_tiles), dim3(TS), shm, \
                     R->stream, M, S, T, tw, (GT*)R->grad, R->gate_sc())
  switch (capc) {
    case 1024: RFLU_LDS(1024); break;
    case 1280: RFLU_LDS(1280); break;
    case 1536: RFLU_LDS(1536); break;
    case 2048: RFLU_LDS(2048); break;
    default: RFLU_LDS(0); break;
  }
#undef RFLU_LDS
}

template <typename GT, bool V, int U, int TS>
void launch_gradient_tiles(Region* R, const Mesh& M, const State& S) {
  Tiles T = region_tiles(R);
  if (R->gw && R->gw_kind == 1 && !R->grad_fast) {
    size_t shm = (size_t)5 * R->tile_cap * sizeof(double);
    hipLaunchKernelGGL((k_gradient_ldsw<GT, V, 4 * U, TS>), dim3(R->n_tiles), dim3(TS), shm,
                       R->stream, M, S, T, R->tile_loc_len, (const GT*)R->gw, (GT*)R->grad, R->gate_sc());
  } else {
    // a compile-time LDS stride when it keeps the workgroups per CU (160 KiB of LDS)
    const int cap = R->tile_cap;
    auto per_cu = [](int c) { return (160 * 1024) / (8 * c * (int)sizeof(double)); };
    int capc = 0;
    for (int c : {1024, 1280, 1536, 2048})
      if (cap <= c) {
        if (per_cu(c) == per_cu(cap)) capc = c;
        break;
      }
    const size_t shm = (size_t)8 * (capc ? capc : cap) * sizeof(double);
    if (R->grad_fast) {
      switch (capc) {
#define RFLU_MOM(C)                                                                          \
  hipLaunchKernelGGL((k_gradient_mom<GT, V, U, TS, C>), dim3(R->n_tiles), dim3(TS), shm, R->stream, \
                     M, S, T, (GT*)R->grad, R->gate_sc())
        case 1024: RFLU_MOM(1024); break;
        case 1280: RFLU_MOM(1280); break;
        case 1536: RFLU_MOM(1536); break;
        case 2048: RFLU_MOM(2048); break;
        default: RFLU_MOM(0); break;
#undef RFLU_MOM
      }
      return;
    }
    // stored inverse distances (GRAD_WEIGHTS 2) with groups of 4 members
    if constexpr (U == 4) {
      if (R->gw && R->gw_kind == 2) {
        launch_gradient_lds<GT, V, U, TS, true>(R, M, S, T, capc, shm);
        return;
      }
    }
    launch_gradient_lds<GT, V, U, TS, false>(R, M, S, T, capc, shm);
  }
}

template <typename GT, bool V, int U>
void launch_gradient_kernel(Region* R, const Mesh& M, const State& S) {
  if (R->use_lds && R->grad_lds) {
    switch (R->grad_tile) {
      case 128: launch_gradient_tiles<GT, V, U, 128>(R, M, S); break;
      case 512: launch_gradient_tiles<GT, V, U, 512>(R, M, S); break;
      default: launch_gradient_tiles<GT, V, U, 256>(R, M, S); break;
    }
  } else {
    hipLaunchKernelGGL((k_gradient<GT, V, U>), dim3(nblk(R->N, R->grad_block)), dim3(R->grad_block),
                       0, R->stream, M, S, (GT*)R->grad, R->gate_sc());
  }
}

// The mixed mode's FP32-record tiles (GRAD_LDS 2).
template <bool V, int U>
void launch_gradient_lds32(Region* R, const Mesh& M, const State& S) {
  Tiles T = region_tiles(R);
  const size_t shm = (size_t)2 * R->tile_cap * sizeof(float4);
#define RFLU_L32(TS)                                                                           \
  hipLaunchKernelGGL((k_gradient_lds32<V, U, TS>), dim3(R->n_tiles), dim3(TS), shm, R->stream, M, \
                     S, T, (float*)R->grad, R->gate_sc())
  switch (R->grad_tile) {
    case 128: RFLU_L32(128); break;
    case 512: RFLU_L32(512); break;
    default: RFLU_L32(256); break;
  }
#undef RFLU_L32
}

template <typename GT>
int launch_gradients_t(Region* R) {
  Mesh M = R->mesh();
  State S = R->state();
  // with constrained cells the limiter runs after them, as a pass of its own
  bool v = R->reconst == RFLU_RECONST_LIM_VENKAT && R->con_n == 0;
  if constexpr (sizeof(GT) == 4) {
    if (R->grad_lds == 2 && R->use_lds && !R->grad_fast && !R->gw) {
      if (R->grad_unroll == 8) v ? launch_gradient_lds32<true, 8>(R, M, S) : launch_gradient_lds32<false, 8>(R, M, S);
      else v ? launch_gradient_lds32<true, 4>(R, M, S) : launch_gradient_lds32<false, 4>(R, M, S);
      return 0;
    }
  }
  // groups of 8 members: formed weights over 256-cell LDS tiles only (else groups of 4)
  if (R->grad_unroll == 8 && R->use_lds && R->grad_lds && R->grad_tile == 256 &&
      !R->grad_fast && !R->gw) {
    Tiles T = region_tiles(R);
    const int cap = R->tile_cap;
    auto per_cu = [](int c) { return (160 * 1024) / (8 * c * (int)sizeof(double)); };
    int capc = 0;
    for (int c : {1024, 1280, 1536, 2048})
      if (cap <= c) {
        if (per_cu(c) == per_cu(cap)) capc = c;
        break;
      }
    const size_t shm = (size_t)8 * (capc ? capc : cap) * sizeof(double);
    if (v) launch_gradient_lds<GT, true, 8, 256, false>(R, M, S, T, capc, shm);
    else launch_gradient_lds<GT, false, 8, 256, false>(R, M, S, T, capc, shm);
    return 0;
  }
  switch (R->grad_unroll) {
    case 1: v ? launch_gradient_kernel<GT, true, 1>(R, M, S) : launch_gradient_kernel<GT, false, 1>(R, M, S); break;
    case 2: v ? launch_gradient_kernel<GT, true, 2>(R, M, S) : launch_gradient_kernel<GT, false, 2>(R, M, S); break;
    default: v ? launch_gradient_kernel<GT, true, 4>(R, M, S) : launch_gradient_kernel<GT, false, 4>(R, M, S); break;
  }
  return 0;
}

template <typename GT>
int launch_blend_t(Region* R) {
  if (R->reconst != RFLU_RECONST_WENO_XYZ) return 0;
  Mesh M = R->mesh();
  State S = R->state();
  const size_t shm = (size_t)8 * R->tile_cap * sizeof(GT);
  if (R->use_lds && R->grad_lds && shm <= 160 * 1024) {
    // over the gradient's stencil tiles, 8 components of every union record in LDS
    Tiles T = region_tiles(R);
#define RFLU_WENO(TS)                                                                          \
  hipLaunchKernelGGL((k_weno_clip_lds<GT, TS>), dim3(R->n_tiles), dim3(TS), shm, R->stream, M, \
                     S, T, (const GT*)R->grad, (GT*)R->grad2, R->gate_sc())
    switch (R->grad_tile) {
      case 128: RFLU_WENO(128); break;
      case 512: RFLU_WENO(512); break;
      default: RFLU_WENO(256); break;
    }
#undef RFLU_WENO
    return 0;
  }
  hipLaunchKernelGGL((k_weno_clip<GT>), dim3(nblk(R->N, R->grad_block)), dim3(R->grad_block), 0,
                     R->stream, M, S, (const GT*)R->grad, (GT*)R->grad2, R->gate_sc());
  return 0;
}

template <typename FT, bool FUSED, bool PRIM>
int launch_update_rk(Region* R, int stage, int last, double dt_in) {
  Mesh M = R->mesh();
  State S = R->state();
  dim3 cb(R->cell_block), cg(nblk(R->N, R->cell_block));
  const FT* fl = (const FT*)R->flux;
  const Cases C = R->cases();
  if (R->update_tile && R->face_tile > 0) {
    FaceTiles FTl;
    FTl.fptr = R->ft_fptr; FTl.eptr = R->ft_eptr; FTl.eidx = R->ft_eidx; FTl.fslot = R->ft_fslot;
    FTl.ts = R->face_tile; FTl.cap = R->face_cap; FTl.fcap = R->face_fcap;
    const size_t shm = (size_t)5 * R->face_fcap * sizeof(typename FluxIO<FT>::val_t);
    dim3 tb(R->face_tile), tg(R->n_face_tiles);
    switch (R->rk) {
      case 1: hipLaunchKernelGGL((k_update_tile<1, FT, FUSED, PRIM>), tg, tb, shm, R->stream, M, S, R->gas, C, FTl, fl, R->sc, stage, last, dt_in); break;
      case 2: hipLaunchKernelGGL((k_update_tile<2, FT, FUSED, PRIM>), tg, tb, shm, R->stream, M, S, R->gas, C, FTl, fl, R->sc, stage, last, dt_in); break;
      case 3: hipLaunchKernelGGL((k_update_tile<3, FT, FUSED, PRIM>), tg, tb, shm, R->stream, M, S, R->gas, C, FTl, fl, R->sc, stage, last, dt_in); break;
      default: hipLaunchKernelGGL((k_update_tile<4, FT, FUSED, PRIM>), tg, tb, shm, R->stream, M, S, R->gas, C, FTl, fl, R->sc, stage, last, dt_in); break;
    }
    return 0;
  }
  switch (R->rk) {
    case 1: hipLaunchKernelGGL((k_update<1, FT, FUSED, PRIM>), cg, cb, 0, R->stream, M, S, R->gas, C, fl, R->sc, stage, last, dt_in); break;
    case 2: hipLaunchKernelGGL((k_update<2, FT, FUSED, PRIM>), cg, cb, 0, R->stream, M, S, R->gas, C, fl, R->sc, stage, last, dt_in); break;
    case 3: hipLaunchKernelGGL((k_update<3, FT, FUSED, PRIM>), cg, cb, 0, R->stream, M, S, R->gas, C, fl, R->sc, stage, last, dt_in); break;
    default: hipLaunchKernelGGL((k_update<4, FT, FUSED, PRIM>), cg, cb, 0, R->stream, M, S, R->gas, C, fl, R->sc, stage, last, dt_in); break;
  }
  return 0;
}

// Event bracketing for per-kernel timing.
struct Timed {
  Region* R;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(Region* r, int k) : R(r), kid(k) {
    if (R->profiling) {
      (void)hipEventCreateWithFlags(&a, hipEventDisableSystemFence);
      (void)hipEventCreateWithFlags(&b, hipEventDisableSystemFence);
      (void)hipEventRecord(a, R->stream);
    }
  }
  ~Timed() {
    if (R->profiling) {
      (void)hipEventRecord(b, R->stream);
      R->events.push_back({kid, a, b});
    }
  }
};

void collect_events(Region* R) {
  for (auto& e : R->events) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      R->kern_ms[e.kid] += ms;
      R->kern_n[e.kid] += 1;
    }
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  R->events.clear();
}

// FP32 gradient storage/arithmetic (the mixed mode; FP32_GRAD is its older name).
inline bool grad32(const Region* R) { return R->prec == RFLU_PREC_MIXED || R->prec == RFLU_PREC_FP32_GRAD; }
// FP32 interior-flux arithmetic (the FP32_FLUX diagnostic split only).
inline bool flux32(const Region* R) { return R->prec == RFLU_PREC_FP32_FLUX; }

// Calls fn with a null pointer of the region's flux-buffer type (double or FluxMix),
// so one generic lambda launches the matching kernel instantiation.
template <typename Fn>
int with_fbuf(const Region* R, Fn&& fn) {
  if (R->fbuf == FBUF_MIX) return fn((FluxMix*)nullptr);
  return fn((double*)nullptr);
}

int launch_gradients(Region* R) {
  if (R->order != 2) return 0;
  int rc;
  {
    Timed t(R, KID_GRAD);
    rc = grad32(R) ? launch_gradients_t<float>(R) : launch_gradients_t<double>(R);
  }
  if (rc) return rc;
  if (R->con_n > 0) {
    Timed t(R, KID_GRAD);
    Mesh M = R->mesh();
    State S = R->state();
    Constr C = R->constr();
    dim3 g(nblk(R->con_n, 256)), b(256);
    if (grad32(R)) hipLaunchKernelGGL((k_grad_constr<float>), g, b, 0, R->stream, M, S, C, (float*)R->grad, R->gate_sc());
    else hipLaunchKernelGGL((k_grad_constr<double>), g, b, 0, R->stream, M, S, C, (double*)R->grad, R->gate_sc());
    if (R->reconst == RFLU_RECONST_LIM_VENKAT) {
      if (!R->vnb || !R->vgeo) return fail(RFLU_ERR_ARG, "limiter neighbour lists missing");
      dim3 gc(nblk(R->N, R->grad_block)), bc(R->grad_block);
      if (grad32(R)) hipLaunchKernelGGL((k_venkat_apply<float>), gc, bc, 0, R->stream, M, S, (float*)R->grad, R->gate_sc());
      else hipLaunchKernelGGL((k_venkat_apply<double>), gc, bc, 0, R->stream, M, S, (double*)R->grad, R->gate_sc());
    }
  }
  if (R->reconst == RFLU_RECONST_LIM_BARTHJESP) {
    if (!R->vnb || !R->vgeo) return fail(RFLU_ERR_ARG, "limiter neighbour lists missing");
    Timed t(R, KID_BLEND);
    Mesh M = R->mesh();
    State S = R->state();
    dim3 gc(nblk(R->N, R->grad_block)), bc(R->grad_block);
    if (grad32(R)) hipLaunchKernelGGL((k_barth_jesp<float>), gc, bc, 0, R->stream, M, S, (float*)R->grad, R->gate_sc());
    else hipLaunchKernelGGL((k_barth_jesp<double>), gc, bc, 0, R->stream, M, S, (double*)R->grad, R->gate_sc());
  }
  if (R->reconst == RFLU_RECONST_WENO_XYZ) {
    Timed t(R, KID_BLEND);
    rc = grad32(R) ? launch_blend_t<float>(R) : launch_blend_t<double>(R);
  }
  return rc;
}

template <bool SR>
int launch_fluxes_p(Region* R) {
  int rc;
  {
    Timed t(R, KID_FACE);
    switch (R->prec) {
      case RFLU_PREC_FP64: rc = launch_fluxes_t<double, double, double, SR>(R); break;
      case RFLU_PREC_FP32_FLUX: rc = launch_fluxes_t<float, double, FluxMix, SR>(R); break;
      default: rc = launch_fluxes_t<double, float, double, SR>(R); break;   // mixed
    }
  }
  if (rc) return rc;
  {
    Timed t(R, KID_BFACE);
    switch (R->prec) {
      case RFLU_PREC_FP64: rc = launch_bfaces_t<double, double, SR>(R); break;
      case RFLU_PREC_FP32_FLUX: rc = launch_bfaces_t<double, FluxMix, SR>(R); break;
      default: rc = launch_bfaces_t<float, double, SR>(R); break;   // mixed
    }
  }
  return rc;
}

// Stage 1's mass in/out over flow-through patches from the boundary mass fluxes
// stored at flux[F + b].
void launch_mass_inout(Region* R) {
  if (!R->has_flowthrough) return;
  Mesh M = R->mesh();
  with_fbuf(R, [&](auto* tag) {
    using FT = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL((k_mass_inout<FT>), dim3(R->K), dim3(256), 0, R->stream, M, R->cases(),
                       R->patches(), (const FT*)R->flux);
    return 0;
  });
}

// srad: also store the face spectral radii of the time step (first stage of a fused
// step, whose face pass sees the step-start state; the fused step counts the mass
// after its time step, which may end a case).
int launch_fluxes(Region* R, int stage, bool srad = false) {
  int rc = srad ? launch_fluxes_p<true>(R) : launch_fluxes_p<false>(R);
  if (rc) return rc;
  if (stage == 1 && !srad) launch_mass_inout(R);
  return 0;
}

int launch_update(Region* R, int stage, bool fused, double dt_in = -1.0, bool advance = true) {
  Timed t(R, KID_UPDATE);
  // the last stage of a fused step advances the device clock (not for the host's stages)
  int last = advance && stage == (R->rk == 1 ? 4 : R->rk == 3 ? 1 : 3);
  bool prim = R->order == 2;
  return with_fbuf(R, [&](auto* tag) {
    using FT = std::remove_pointer_t<decltype(tag)>;
    if (fused) return prim ? launch_update_rk<FT, true, true>(R, stage, last, dt_in)
                           : launch_update_rk<FT, true, false>(R, stage, last, dt_in);
    return launch_update_rk<FT, false, false>(R, stage, last, dt_in);
  });
}

// Minimum local time step from the face spectral radii already in R->srad, then
// RFLU_MinimumTimeStep's rules.
int launch_dt_reduce(Region* R, double time_in, double end_time) {
  Mesh M = R->mesh();
  if (R->K > 1) {
    hipLaunchKernelGGL(k_dt_gather_cases, dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M,
                       R->cases(), (const double*)R->srad);
    hipLaunchKernelGGL(k_dt_finalize_cases, dim3(nblk(R->K, 64)), dim3(64), 0, R->stream, R->cases());
    return 0;
  }
  CaseParams cp = R->case0();
  cp.end_time = end_time;
  // at most 2048 blocks, each over a grid-strided share of the cells: the single-block
  // minimum then reads 8 partial minima per thread
  const int nparts = std::min(R->n_dt_parts, 2048);
  hipLaunchKernelGGL(k_dt_gather, dim3(nparts), dim3(256), 0, R->stream, M,
                     (const double*)R->srad, R->dt_part, R->sc);
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, R->stream, (const double*)R->dt_part,
                     nparts, R->sc, cp, time_in);
  return 0;
}

// Time step as a pass of its own (the unfused entry point).
int launch_time_step(Region* R, double time_in = -1.0, double end_time = HUGE_VAL) {
  Timed t(R, KID_DT);
  Mesh M = R->mesh();
  State S = R->state();
  const int nf = R->F + R->B;
  if (nf > 0)
    hipLaunchKernelGGL(k_face_srad, dim3(nblk(nf, 256)), dim3(256), 0, R->stream, M, S, R->gas,
                       R->cases(), R->patches(), R->srad, R->gate_sc());
  return launch_dt_reduce(R, time_in, end_time);
}

int nstages(int rk) { return rk == 1 ? 4 : rk == 2 ? 3 : rk == 3 ? 1 : 3; }

// Whether the persistent step can run this region as it is configured (empty if so).
std::string persistent_unsupported(const Region* R) {
  if (R->K < 2) return "PERSISTENT needs a batch of cases";
  if (R->order != 2 || R->rk != RFLU_RK3_SSP) return "PERSISTENT needs order 2 and SSP-RK3";
  if (R->reconst != RFLU_RECONST_NONE && R->reconst != RFLU_RECONST_LIM_VENKAT)
    return "PERSISTENT supports no modifier or the Venkatakrishnan limiter";
  if (R->con_n > 0) return "PERSISTENT does not run constrained gradients";
  if (R->prec == RFLU_PREC_FP32_FLUX) return "PERSISTENT runs FP64 or mixed precision";
  if (!R->use_lds || R->grad_lds != 1 || R->grad_tile != 256)
    return "PERSISTENT needs 256-cell gradient tiles in LDS (GRAD_TILE 256, GRAD_LDS 1)";
  if (R->face_tile <= 0 || 256 % R->face_tile) return "PERSISTENT needs face tiles (FACE_TILE)";
  return "";
}

template <typename GT, int D, bool V, bool FAST>
int launch_persistent_t(Region* R, const PersistentPlan& pl, const GridBar& gb, bool dry) {
  auto kern = k_step_persistent<GT, D, V, FAST>;
  const int key = (int)sizeof(GT) * 1000 + D * 100 + (V ? 10 : 0) + (FAST ? 1 : 0);
  if (R->pers_key != key || R->pers_blocks <= 0) {
    int per_cu = 0, dev = 0, ncu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern),
                                                         256, R->pers_shm));
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (per_cu < 1) return fail(RFLU_ERR_ARG, "PERSISTENT: the step's LDS does not fit a CU");
    R->pers_blocks = per_cu * ncu;
    R->pers_key = key;
  }
  if (dry) return 0;
  Mesh M = R->mesh();
  State S = R->state();
  FaceTiles FTl;
  FTl.fptr = R->ft_fptr; FTl.eptr = R->ft_eptr; FTl.eidx = R->ft_eidx; FTl.fslot = R->ft_fslot;
  FTl.ts = R->face_tile; FTl.cap = R->face_cap; FTl.fcap = R->face_fcap;
  hipLaunchKernelGGL(k_bar_reset, dim3(1), dim3(1), 0, R->stream, gb);
  hipLaunchKernelGGL(kern, dim3(R->pers_blocks), dim3(256), R->pers_shm, R->stream, M, S, R->gas,
                     R->cases(), R->patches(), region_tiles(R), FTl, (GT*)R->grad,
                     (double*)R->flux, R->srad, pl, gb);
  return 0;
}

// One time step of a batch as one persistent launch (the phases of launch_step).  dry:
// only size the grid (before a graph capture, which should see launches only).
int launch_persistent(Region* R, bool dry = false) {
  std::string why = persistent_unsupported(R);
  if (!why.empty()) return fail(RFLU_ERR_ARG, why);
  PersistentPlan pl;
  pl.n_grad_tiles = R->n_tiles;
  pl.n_face_tiles = R->n_face_tiles;
  pl.upd_per_iter = 256 / R->face_tile;
  pl.upd_lds_stride = (int)(5 * (size_t)R->face_fcap * sizeof(double));
  pl.n_cell_blocks = (int)nblk(R->N, 256);
  pl.flowthrough = R->has_flowthrough ? 1 : 0;
  const bool g32 = grad32(R);
  size_t shm = (size_t)8 * R->tile_cap * sizeof(double);
  shm = std::max(shm, g32 ? face_tile_lds<float>(R->face_cap, true)
                          : face_tile_lds<double>(R->face_cap, true));
  shm = std::max(shm, (size_t)pl.upd_per_iter * pl.upd_lds_stride);
  if (shm != R->pers_shm) {
    R->pers_shm = shm;
    R->pers_blocks = 0;
  }
  GridBar gb;
  gb.count = R->pbar;
  gb.gen = R->pbar + 1;
  gb.err = (int*)(R->pbar + 2);
  const bool v = R->reconst == RFLU_RECONST_LIM_VENKAT;
#define RFLU_PERS(GT, D)                                                             \
  (v ? (R->grad_fast ? launch_persistent_t<GT, D, true, true>(R, pl, gb, dry)            \
                     : launch_persistent_t<GT, D, true, false>(R, pl, gb, dry))          \
     : (R->grad_fast ? launch_persistent_t<GT, D, false, true>(R, pl, gb, dry)           \
                     : launch_persistent_t<GT, D, false, false>(R, pl, gb, dry)))
  int rc;
  if (g32) rc = R->discr == 1 ? RFLU_PERS(float, 1) : R->discr == 3 ? RFLU_PERS(float, 3) : RFLU_PERS(float, 4);
  else rc = R->discr == 1 ? RFLU_PERS(double, 1) : R->discr == 3 ? RFLU_PERS(double, 3) : RFLU_PERS(double, 4);
#undef RFLU_PERS
  return rc;
}

// After persistent steps: whether a grid barrier timed out (a read of the error word).
int persistent_check(Region* R) {
  if (!R->persistent || !R->pbar) return 0;
  int err = 0;
  HIP_TRY(hipMemcpyAsync(&err, R->pbar + 2, sizeof(int), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  if (err) return fail(RFLU_ERR_HIP, "persistent step: a grid barrier wait timed out");
  return 0;
}

// One fused time step.  The time step only enters at the first update, so it is
// reduced after the first stage's face pass, which also stores the face spectral
// radii of the step-start state.
int launch_step(Region* R) {
  if (R->persistent) return launch_persistent(R);
  int rc;
  for (int s = 1; s <= nstages(R->rk); ++s) {
    if ((rc = launch_gradients(R))) return rc;
    if ((rc = launch_fluxes(R, s, s == 1))) return rc;
    if (s == 1) {
      {
        Timed t(R, KID_DT);
        if ((rc = launch_dt_reduce(R, -1.0, R->end_time))) return rc;
      }
      launch_mass_inout(R);
    }
    if ((rc = launch_update(R, s, true))) return rc;
  }
  return 0;
}

int read_cases(Region* R, std::vector<Scalars>& v) {
  v.resize(R->K);
  HIP_TRY(hipMemcpyAsync(v.data(), R->sc, R->K * sizeof(Scalars), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  return 0;
}

// The region's scalars; for a batch, counters and mass summed over the cases, the
// earliest time, smallest dt and fewest steps, finished when every case is.
int read_scalars(Region* R, Scalars* h) {
  std::vector<Scalars> v;
  int rc = read_cases(R, v);
  if (rc) return rc;
  *h = v[0];
  for (int c = 1; c < R->K; ++c) {
    const Scalars& s = v[c];
    h->time = std::min(h->time, s.time);
    h->dt = std::min(h->dt, s.dt);
    h->dt_local_min = std::min(h->dt_local_min, s.dt_local_min);
    h->mass_in += s.mass_in;
    h->mass_out += s.mass_out;
    h->n_invalid += s.n_invalid;
    h->n_negative += s.n_negative;
    h->n_bounds_fix += s.n_bounds_fix;
    h->n_pos_fix += s.n_pos_fix;
    h->steps = std::min(h->steps, s.steps);
    h->abort |= s.abort;
    h->fail |= s.fail;
    h->last_step &= s.last_step;
    h->finished &= s.finished;
  }
  return 0;
}

int status_from(const Scalars& s) {
  if (s.n_invalid) return fail(RFLU_ERR_INVALID_VALUE, "invalid (NaN) values detected");
  if (s.n_negative) return fail(RFLU_ERR_NEGATIVE_POSDEF, "negative positive-definite variables detected");
  return RFLU_OK;
}

// Pseudo-inverse of an m x 3 matrix a (column-major) by one-sided Jacobi SVD, with the
// reference's truncation: singular values are sorted descending and the first one
// below 1% of its predecessor (or s1 <= 100 eps) drops with all later ones
// (libflu/RFLU_InvertMatrixSVD.F90).  ainv is 3 x m, column-major.  Returns the
// number of dropped singular values.
int pinv_svd3(int m, const double* a, double* ainv) {
  std::vector<double> u(a, a + 3 * (size_t)m);
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < m; ++i) {
          al += u[p * m + i] * u[p * m + i];
          be += u[q * m + i] * u[q * m + i];
          ga += u[p * m + i] * u[q * m + i];
        }
        if (ga == 0.0) continue;
        off = std::max(off, std::fabs(ga) / std::sqrt(al * be));
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < m; ++i) {
          double up = u[p * m + i], uq = u[q * m + i];
          u[p * m + i] = c * up - s * uq;
          u[q * m + i] = s * up + c * uq;
        }
        for (int i = 0; i < 3; ++i) {
          double vp = v[i][p], vq = v[i][q];
          v[i][p] = c * vp - s * vq;
          v[i][q] = s * vp + c * vq;
        }
      }
    if (off < 1e-15) break;
  }
  double sv[3];
  int ord[3] = {0, 1, 2};
  for (int j = 0; j < 3; ++j) {
    double n = 0;
    for (int i = 0; i < m; ++i) n += u[j * m + i] * u[j * m + i];
    sv[j] = std::sqrt(n);
  }
  std::sort(ord, ord + 3, [&](int x, int y) { return sv[x] > sv[y]; });
  double sinv[3] = {0, 0, 0};
  int dropped = 0;
  if (sv[ord[0]] > 100.0 * 2.220446049250313e-16) sinv[0] = 1.0 / sv[ord[0]];
  else dropped = 1;
  for (int k = 1; k < 3 && !dropped; ++k) {
    if (sv[ord[k]] > 0.01 * sv[ord[k - 1]]) sinv[k] = 1.0 / sv[ord[k]];
    else dropped = 3 - k;
  }
  // A+ = sum_k v_k (1/s_k) u_k^T with u_k = U[:, k] / s_k
  for (int c = 0; c < m; ++c)
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) {
        int j = ord[k];
        if (sinv[k] == 0.0) continue;
        acc += v[i][j] * sinv[k] * sinv[k] * u[j * m + c];
      }
      ainv[c * 3 + i] = acc;
    }
  return dropped;
}

// FloatEqual (modfloflu/ModTools.F90:129-148).
bool float_equal(double a, double b, double tol) {
  return std::fabs(a - b) <= (1.0 + 0.5 * std::fabs(a + b)) * tol;
}

// Rows of the constrained gradients of one region (RFLU_ComputeGradCellsConstr,
// modflu/RFLU_ModDifferentiationCells.F90:1180-1554): for every cell with wall faces in
// its stencil and every velocity component with constraint rows there (slip walls
// whose normal is that axis, RFLU_GetConstrType, RFLU_ModConstraintUtils.F90:187-262,
// value 0, :470-478), the column-scaled least-squares matrix of the members and the
// constraint rows (weight cReconstCellsWeight), its pseudo-inverse, and the per-row
// weights term*aInv(:, row).  Cell numbers are the concatenated caller numbering.
struct ConstrHost {
  std::vector<int> ecell, evar, rptr{0}, rcell;
  std::vector<double> rval, rw[3];
};

int build_constrained(const rflu_region_desc* d, int c0, int p0, const std::vector<int>& bptr,
                      const std::vector<int>& sten_ptr, const std::vector<int>& sten_idx,
                      const std::vector<double4>& cell, const std::vector<double4>& bfn,
                      const std::vector<double4>& bfc, const std::vector<int>& pbc,
                      ConstrHost& H) {
  const int n = d->n_cells, base = d->index_base;
  const double cwt = d->creconst_weight;
  std::vector<double> a, ai, term;
  std::vector<int> rc;
  if (d->c2cs_bptr[0] != 0) return fail(RFLU_ERR_ARG, "c2cs_bptr[0] must be 0");
  for (int lc = 0; lc < n; ++lc) {
    const int b0 = d->c2cs_bptr[lc], nb = d->c2cs_bptr[lc + 1] - b0;
    if (nb < 0) return fail(RFLU_ERR_ARG, "c2cs_bptr not monotone");
    if (nb == 0) continue;
    const int icg = c0 + lc;
    const double4 xc = cell[icg];
    const int s0 = sten_ptr[icg], ns = sten_ptr[icg + 1] - s0;
    for (int iv = 1; iv <= 3; ++iv) {
      std::vector<int> rows;   // global boundary faces of the constraint rows
      for (int k = 0; k < nb; ++k) {
        const int p = p0 + d->c2cs_bface[2 * (size_t)(b0 + k)] - base;
        const int f = d->c2cs_bface[2 * (size_t)(b0 + k) + 1] - base;
        if (p < p0 || p >= p0 + d->n_patches) return fail(RFLU_ERR_ARG, "c2cs_bface patch out of range");
        const int b = bptr[p] + f;
        if (f < 0 || b >= bptr[p + 1]) return fail(RFLU_ERR_ARG, "c2cs_bface face out of range");
        const double nn = iv == 1 ? bfn[b].x : iv == 2 ? bfn[b].y : bfn[b].z;
        if (pbc[p] == BC_SLIPWALL && float_equal(std::fabs(nn), 1.0, 1.0E-6)) rows.push_back(b);
      }
      if (rows.empty()) continue;
      const int m = ns + (int)rows.size();
      a.assign(3 * (size_t)m, 0.0);
      ai.assign(3 * (size_t)m, 0.0);
      term.assign(m, 0.0);
      rc.assign(m, -1);
      for (int s = 0; s < ns; ++s) {
        const int j = sten_idx[s0 + s];
        const double dx = cell[j].x - xc.x, dy = cell[j].y - xc.y, dz = cell[j].z - xc.z;
        term[s] = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
        a[s] = term[s] * dx;
        a[m + s] = term[s] * dy;
        a[2 * (size_t)m + s] = term[s] * dz;
        rc[s] = j;
      }
      for (size_t k = 0; k < rows.size(); ++k) {
        const int r = ns + (int)k;
        const double4 xf = bfc[rows[k]];
        const double dx = xf.x - xc.x, dy = xf.y - xc.y, dz = xf.z - xc.z;
        term[r] = cwt / std::sqrt(dx * dx + dy * dy + dz * dz);
        a[r] = term[r] * dx;
        a[m + r] = term[r] * dy;
        a[2 * (size_t)m + r] = term[r] * dz;
      }
      double colmax[3];
      for (int c = 0; c < 3; ++c) {
        colmax[c] = -HUGE_VAL;
        for (int i = 0; i < m; ++i) colmax[c] = std::max(colmax[c], std::fabs(a[c * (size_t)m + i]));
        for (int i = 0; i < m; ++i) a[c * (size_t)m + i] = a[c * (size_t)m + i] / colmax[c];
      }
      // the reference stops on a singular matrix (RFLU_ModDifferentiationCells.F90:1432-1436)
      if (pinv_svd3(m, a.data(), ai.data()) != 0)
        return fail(RFLU_ERR_ARG, "singular constrained-gradient matrix at cell " + std::to_string(lc));
      H.ecell.push_back(icg);
      H.evar.push_back(iv);
      for (int i = 0; i < m; ++i) {
        H.rcell.push_back(rc[i]);
        H.rval.push_back(0.0);   // slip wall: the wall-normal velocity is zero
        for (int c = 0; c < 3; ++c) H.rw[c].push_back(term[i] * (ai[3 * (size_t)i + c] / colmax[c]));
      }
      H.rptr.push_back((int)H.rcell.size());
    }
  }
  return 0;
}

// Validates one descriptor (one case); ref, when given, is the batch's first case,
// whose scheme every case must share.
int check_desc(const rflu_region_desc* d, const rflu_region_desc* ref) {
  const int N = d->n_cells, F = d->n_faces, P = d->n_patches;
  const int base = d->index_base;
  if (N <= 0 || F < 0 || P < 0 || d->n_virtual < 0) return fail(RFLU_ERR_ARG, "bad region sizes");
  if ((long long)N + d->n_virtual > 0x7fffffff) return fail(RFLU_ERR_ARG, "cell count must fit int32");
  if (base != 0 && base != 1) return fail(RFLU_ERR_ARG, "index_base must be 0 or 1");
  if (!d->f2c || !d->fn || !d->cofg || !d->vol || (P > 0 && (!d->bpatch_ptr || !d->patch_bc)))
    return fail(RFLU_ERR_ARG, "missing grid arrays");
  if (d->order != 1 && d->order != 2) return fail(RFLU_ERR_ARG, "order must be 1 or 2");
  if (d->discr != 1 && d->discr != 3 && d->discr != 4) return fail(RFLU_ERR_ARG, "discr must be 1, 3 or 4");
  if (d->rk < 1 || d->rk > 4) return fail(RFLU_ERR_ARG, "rk must be 1..4");
  if (d->precision < 0 || d->precision > 3) return fail(RFLU_ERR_ARG, "precision must be 0..3");
  if (d->order == 2) {
    if (!d->fc || !d->c2cs_ptr || !d->c2cs_idx || !d->xyz_moms)
      return fail(RFLU_ERR_ARG, "order 2 needs fc, stencils and xyz_moms");
    if (d->reconst != 0 && d->reconst != 2 && d->reconst != 10 && d->reconst != 11)
      return fail(RFLU_ERR_ARG, "reconst must be 0, 2, 10 or 11");
  }
  if (!(d->gamma > 1.0) || !(d->cp > 0.0)) return fail(RFLU_ERR_ARG, "bad gas constants");
  const int B = P > 0 ? d->bpatch_ptr[P] : 0;
  if (B > 0 && (!d->bf2c || !d->bfn || !d->bfc)) return fail(RFLU_ERR_ARG, "missing boundary arrays");
  if (d->bpatch_ptr && d->bpatch_ptr[0] != 0) return fail(RFLU_ERR_ARG, "bpatch_ptr[0] must be 0");
  for (int p = 0; p < P; ++p) {
    const int bc = d->patch_bc[p];
    const int opt = d->patch_opt ? d->patch_opt[p] : 0;
    const int opt2 = d->patch_opt2 ? d->patch_opt2[p] : 0;
    if (bc != BC_OUTFLOW && bc != BC_SLIPWALL && bc != BC_INFLOW_VELTEMP &&
        bc != BC_INFLOW_TOTANG && bc != BC_FARFIELD && bc != BC_INJECTION)
      return fail(RFLU_ERR_ARG, "patch types: inflow (10, 11), outflow (20), slip wall (60), "
                                "farfield (80), injection (90)");
    if (bc == BC_SLIPWALL) continue;
    if (bc != BC_INJECTION && bc != BC_FARFIELD && (opt < 0 || opt > 2))
      return fail(RFLU_ERR_ARG, "inflow/outflow option must be 0, 1 or 2");
    if (bc == BC_OUTFLOW) continue;
    if (!d->patch_vals) return fail(RFLU_ERR_ARG, "inflow, farfield and injection patches need patch_vals");
    const double* v = d->patch_vals + 5 * (size_t)p;
    if (bc == BC_INFLOW_VELTEMP &&
        (!(v[3] > 0.0) || (opt != 1 && !(v[4] > 0.0))))
      return fail(RFLU_ERR_ARG, "inflow temperature (and non-subsonic pressure) must be > 0");
    if (bc == BC_INFLOW_TOTANG &&
        (!(v[0] > 0.0) || !(v[1] > 0.0) || (opt != 1 && !(v[4] > 0.0)) || opt2 < 0 || opt2 > 1))
      return fail(RFLU_ERR_ARG, "total-condition inflow needs ptot, ttot (and a non-subsonic "
                                "Mach number) > 0 and a fixed-angle switch of 0 or 1");
    if (bc == BC_FARFIELD && (!(v[0] >= 0.0) || !(v[3] > 0.0) || !(v[4] > 0.0)))
      return fail(RFLU_ERR_ARG, "farfield Mach number must be >= 0, pressure and temperature > 0");
    if (bc == BC_FARFIELD && opt2 != 0)
      return fail(RFLU_ERR_ARG, "the farfield vortex correction (BCOPT_CORR_YES) is not supported");
    if (bc == BC_INJECTION && (!(v[0] == v[0]) || (v[0] > 0.0 && !(v[1] > 0.0))))
      return fail(RFLU_ERR_ARG, "injection temperature must be > 0 where the mass flux is");
    if (bc == BC_INJECTION && d->order == 2 && d->creconst)
      return fail(RFLU_ERR_ARG, "constrained gradients at injection patches are not supported");
  }
  if (ref && (d->n_virtual > 0 || ref->n_virtual > 0))
    return fail(RFLU_ERR_ARG, "a batch of cases takes no virtual cells");
  if (ref && (d->order != ref->order || d->discr != ref->discr || d->rk != ref->rk ||
              d->precision != ref->precision || (d->order == 2 && d->reconst != ref->reconst)))
    return fail(RFLU_ERR_ARG, "every case of a batch needs the same discr, order, reconst, rk and precision");
  return 0;
}

// Gas constants of one case (libflu/RFLU_SetGasVars.F90:172-178 for gamma').
Gas gas_of(const rflu_region_desc* d) {
  Gas G;
  G.g = d->gamma; G.cp = d->cp; G.gc = d->cp - d->cp / d->gamma;
  double m = 8314.3 / G.gc;          // MixtPerf_M_R
  G.mm = m;
  G.rgasp = 8314.3 / m;              // MixtPerf_R_M
  G.gp = d->cp / (d->cp - G.rgasp);  // MixtPerf_G_CpR
  G.eps_entr = d->eps_entr; G.diss_fact = d->diss_fact;
  return G;
}

// Builds a region from K case descriptors (K == 1: one t_region).  The cases' cells,
// faces, boundary faces and patches are concatenated in case order; every per-case
// list keeps its own order.
int create_cases(int32_t K, const rflu_region_desc* descs, int32_t* handle) {
  if (!descs || !handle || K < 1 || K > 65535) return fail(RFLU_ERR_ARG, "bad arguments");
  *handle = 0;
  int rc = 0;
  std::vector<long long> co(K + 1, 0), fo(K + 1, 0), bo(K + 1, 0), po(K + 1, 0), so(K + 1, 0);
  for (int c = 0; c < K; ++c) {
    const rflu_region_desc* d = descs + c;
    if ((rc = check_desc(d, c ? descs : nullptr))) return rc;
    co[c + 1] = co[c] + d->n_cells + d->n_virtual;
    fo[c + 1] = fo[c] + d->n_faces;
    bo[c + 1] = bo[c] + (d->n_patches > 0 ? d->bpatch_ptr[d->n_patches] : 0);
    po[c + 1] = po[c] + d->n_patches;
    so[c + 1] = so[c] + (d->order == 2 ? d->c2cs_ptr[d->n_cells + d->n_virtual] : 0);
  }
  if (co[K] > 0x7fffffff || fo[K] + bo[K] > 0x7fffffff || so[K] > 0x7fffffff)
    return fail(RFLU_ERR_ARG, "cell, face and stencil counts must fit int32");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
    return fail(RFLU_ERR_HIP, "no GPU visible");
  const int N = (int)co[K], F = (int)fo[K], B = (int)bo[K], P = (int)po[K];
  const rflu_region_desc* d0 = descs;

  Region* R = new Region();
  R->N = N; R->F = F; R->B = B; R->P = P; R->K = K;
  R->nact = K == 1 ? d0->n_cells : N;
  R->base = d0->index_base;
  R->order = d0->order; R->discr = d0->discr; R->reconst = d0->order == 2 ? d0->reconst : 0;
  R->rk = d0->rk; R->prec = d0->precision;
  R->cfl = d0->cfl; R->dt_imposed = d0->dt_imposed; R->dt_min_limit = d0->dt_min_limit;
  R->gas = gas_of(d0);
  std::vector<Gas> hgas(K);
  R->par_h.resize(K);
  R->cell0_h.assign(co.begin(), co.end());
  R->bface0_h.assign(bo.begin(), bo.end());
  std::vector<int> hpbc(P), hpopt(P), hpopt2(P), hporder2(P);
  std::vector<double> hppout(P), hpvals((size_t)5 * P, 0.0);
  R->bptr_h.assign(P + 1, 0);
  for (int c = 0; c < K; ++c) {
    const rflu_region_desc* d = descs + c;
    hgas[c] = gas_of(d);
    CaseParams& cp = R->par_h[c];
    cp.cfl = d->cfl; cp.dt_imposed = d->dt_imposed; cp.dt_min_limit = d->dt_min_limit;
    cp.end_time = HUGE_VAL;
    for (int p = 0; p < d->n_patches; ++p) {
      const size_t q = (size_t)po[c] + p;
      hpbc[q] = d->patch_bc[p];
      hpopt[q] = d->patch_opt ? d->patch_opt[p] : 0;
      hpopt2[q] = d->patch_opt2 ? d->patch_opt2[p] : 0;
      hporder2[q] = d->patch_order ? (d->patch_order[p] == 2) : 0;
      hppout[q] = d->patch_pout ? d->patch_pout[p] : 0.0;
      if (d->patch_vals)
        for (int k = 0; k < 5; ++k) hpvals[5 * q + k] = d->patch_vals[5 * (size_t)p + k];
      if (d->patch_bc[p] != BC_SLIPWALL) R->has_flowthrough = true;   // mass crosses the patch
      R->bptr_h[q] = (int)bo[c] + d->bpatch_ptr[p];
      R->bptr_h[q + 1] = (int)bo[c] + d->bpatch_ptr[p + 1];
    }
  }

  // ---- host-side packing ----
  std::vector<int2> hf2c(F);
  std::vector<double4> hfn(F), hfc(F), hcell(N), hbfn(B), hbfc(B);
  std::vector<int> hbf2c(B), hbpatch(B);
  std::vector<int> hsten_ptr, hsten_idx;
  std::vector<double> hmoms, hepsq;
  if (R->order == 2) {
    hsten_ptr.assign(N + 1, 0);
    hsten_idx.resize(so[K]);
    hmoms.resize((size_t)6 * N);
    hepsq.resize(N);
  }
  for (int k = 0; k < K; ++k) {
    const rflu_region_desc* d = descs + k;
    // n counts actual and virtual cells; boundary faces belong to actual cells, or (the
    // outer faces of symmetry/periodic virtual layers) to virtual ones
    const int n = d->n_cells + d->n_virtual, nf = d->n_faces;
    const int np = d->n_patches, base = d->index_base;
    const int c0 = (int)co[k], f0 = (int)fo[k], b0 = (int)bo[k], p0 = (int)po[k];
    const int nb = (int)(bo[k + 1] - bo[k]);
    for (int f = 0; f < nf; ++f) {
      int c1 = d->f2c[2 * (size_t)f] - base, c2 = d->f2c[2 * (size_t)f + 1] - base;
      if (c1 < 0 || c1 >= n || c2 < 0 || c2 >= n || c1 == c2) {
        free_region(R);
        return fail(RFLU_ERR_ARG, "f2c out of range at face " + std::to_string(f) + " of case " + std::to_string(k));
      }
      hf2c[f0 + f] = make_int2(c0 + c1, c0 + c2);
      hfn[f0 + f] = make_double4(d->fn[4 * (size_t)f], d->fn[4 * (size_t)f + 1], d->fn[4 * (size_t)f + 2], d->fn[4 * (size_t)f + 3]);
      if (d->fc) hfc[f0 + f] = make_double4(d->fc[3 * (size_t)f], d->fc[3 * (size_t)f + 1], d->fc[3 * (size_t)f + 2], 0.0);
      else hfc[f0 + f] = make_double4(0, 0, 0, 0);
    }
    for (int c = 0; c < n; ++c) {
      hcell[c0 + c] = make_double4(d->cofg[3 * (size_t)c], d->cofg[3 * (size_t)c + 1], d->cofg[3 * (size_t)c + 2], d->vol[c]);
      if (!(d->vol[c] > 0.0)) { free_region(R); return fail(RFLU_ERR_ARG, "non-positive volume"); }
    }
    for (int p = 0; p < np; ++p)
      for (int b = d->bpatch_ptr[p]; b < d->bpatch_ptr[p + 1]; ++b) {
        if (b < 0 || b >= nb) { free_region(R); return fail(RFLU_ERR_ARG, "bad bpatch_ptr"); }
        int c1 = d->bf2c[b] - base;
        if (c1 < 0 || c1 >= n) { free_region(R); return fail(RFLU_ERR_ARG, "bf2c out of range"); }
        hbf2c[b0 + b] = c0 + c1;
        hbpatch[b0 + b] = p0 + p;
        hbfn[b0 + b] = make_double4(d->bfn[4 * (size_t)b], d->bfn[4 * (size_t)b + 1], d->bfn[4 * (size_t)b + 2], d->bfn[4 * (size_t)b + 3]);
        hbfc[b0 + b] = make_double4(d->bfc[3 * (size_t)b], d->bfc[3 * (size_t)b + 1], d->bfc[3 * (size_t)b + 2], 0.0);
      }
    if (R->order == 2) {
      if (d->c2cs_ptr[0] != 0) { free_region(R); return fail(RFLU_ERR_ARG, "c2cs_ptr[0] must be 0"); }
      const int s0 = (int)so[k];
      for (int c = 0; c < n; ++c) {
        if (d->c2cs_ptr[c + 1] < d->c2cs_ptr[c]) { free_region(R); return fail(RFLU_ERR_ARG, "c2cs_ptr not monotone"); }
        hsten_ptr[c0 + c + 1] = s0 + d->c2cs_ptr[c + 1];
      }
      for (int s = 0; s < d->c2cs_ptr[n]; ++s) {
        int j = d->c2cs_idx[s] - base;
        if (j < 0 || j >= n) { free_region(R); return fail(RFLU_ERR_ARG, "c2cs_idx out of range"); }
        hsten_idx[(size_t)s0 + s] = c0 + j;
      }
      for (int c = 0; c < n; ++c)
        for (int q = 0; q < 6; ++q) hmoms[(size_t)q * N + c0 + c] = d->xyz_moms[10 * (size_t)c + q];
      // Venkatakrishnan eps^2 = (K*vol^(1/3))^3, K = 1 (modflu/RFLU_ModLimiters.F90:473-476)
      const double THRD = 1.0 / 3.0, KV = 1.0;
      for (int c = 0; c < n; ++c) {
        double ds = std::pow(d->vol[c], THRD);
        hepsq[c0 + c] = (KV * ds) * (KV * ds) * (KV * ds);
      }
    }
  }
  // cell -> face list in the reference's scatter order (interior faces in face order,
  // then boundary faces patch by patch)
  std::vector<int> hc2f_ptr(N + 1, 0), hc2f_idx;
  for (int f = 0; f < F; ++f) { hc2f_ptr[hf2c[f].x + 1]++; hc2f_ptr[hf2c[f].y + 1]++; }
  for (int b = 0; b < B; ++b) hc2f_ptr[hbf2c[b] + 1]++;
  for (int c = 0; c < N; ++c) hc2f_ptr[c + 1] += hc2f_ptr[c];
  hc2f_idx.resize(hc2f_ptr[N]);
  {
    std::vector<int> fill(hc2f_ptr.begin(), hc2f_ptr.end() - 1);
    for (int f = 0; f < F; ++f) {
      hc2f_idx[fill[hf2c[f].x]++] = f + 1;
      hc2f_idx[fill[hf2c[f].y]++] = -(f + 1);
    }
    for (int b = 0; b < B; ++b) hc2f_idx[fill[hbf2c[b]]++] = F + b + 1;
  }

  // ---- constrained-gradient rows (before renumbering; cell numbers remapped below) ----
  ConstrHost hcon;
  if (R->order == 2)
    for (int k = 0; k < K; ++k) {
      const rflu_region_desc* d = descs + k;
      if (!d->creconst) continue;
      if (!d->c2cs_bptr || !d->c2cs_bface) { free_region(R); return fail(RFLU_ERR_ARG, "creconst needs c2cs_bptr and c2cs_bface"); }
      if ((rc = build_constrained(d, (int)co[k], (int)po[k], R->bptr_h, hsten_ptr, hsten_idx, hcell,
                                  hbfn, hbfc, hpbc, hcon))) {
        free_region(R);
        return rc;
      }
    }

  // ---- internal renumbering (Morton order of cell centroids within each case) ----
  // Storage order only: every per-cell face list keeps the reference's summation
  // order (original face order) and every stencil keeps its member order, so the
  // arithmetic is the same as in the original numbering.
  renumber(R, hcell, hf2c, hfn, hfc, hbf2c, hc2f_ptr, hc2f_idx, hsten_ptr, hsten_idx, hmoms, hepsq);
  if (!hcon.ecell.empty()) {
    std::vector<int> o2n(N);
    for (int i = 0; i < N; ++i) o2n[R->new2old[i]] = i;
    for (int& c : hcon.ecell) c = o2n[c];
    for (int& c : hcon.rcell) if (c >= 0) c = o2n[c];
  }
  std::vector<unsigned short> hcid;
  if (K > 1) {
    hcid.resize(N);
    for (int k = 0; k < K; ++k)
      for (long long i = co[k]; i < co[k + 1]; ++i) hcid[i] = (unsigned short)k;
  }
  std::vector<int> htile_uptr, htile_uidx, htile_lptr;
  std::vector<unsigned short> htile_loc, htile_vloc;
  std::vector<int> hvnb;
  std::vector<double> hvgeo;
  const bool venkat = R->order == 2 && R->reconst == RFLU_RECONST_LIM_VENKAT;
  // face neighbours and cofg - fc of the limiters (Barth-Jespersen reads them in a pass
  // of its own, outside the gradient tiles)
  if (venkat || (R->order == 2 && R->reconst == RFLU_RECONST_LIM_BARTHJESP))
    build_limiter_geometry(R, hc2f_ptr, hc2f_idx, hf2c, hfc, hcell, hvnb, hvgeo);
  if (R->order == 2)
    build_tiles(R, hsten_ptr, hsten_idx, hvnb, venkat, htile_uptr, htile_uidx, htile_lptr,
                htile_loc, htile_vloc);

  // ---- device allocation ----
  if (hipStreamCreate(&R->stream) != hipSuccess) { free_region(R); return fail(RFLU_ERR_HIP, "hipStreamCreate"); }
  const size_t FB = (size_t)F + B;
  // flux buffer: FP64, except with FP32 interior fluxes (mass and energy FP32, momenta FP64)
  R->fbuf = flux32(R) ? FBUF_MIX : FBUF_F64;
  const size_t fbytes = R->fbuf == FBUF_MIX ? 32 : 40;
  const size_t gsz = grad32(R) ? sizeof(float) : sizeof(double);
  R->n_dt_parts = (int)nblk(N, 256);
  std::vector<int> hcell0(co.begin(), co.end()), hbface0(bo.begin(), bo.end());
  R->pvals_h = hpvals;
  if ((rc = dalloc(&R->f2c, F)) || (rc = dalloc(&R->fn, F)) || (rc = dalloc(&R->fc, F)) ||
      (rc = dalloc(&R->cell, N)) || (rc = dalloc(&R->bfn, B)) || (rc = dalloc(&R->bfc, B)) ||
      (rc = dalloc(&R->bf2c, B)) || (rc = dalloc(&R->bpatch, B)) ||
      (rc = dalloc(&R->c2f_ptr, N + 1)) || (rc = dalloc(&R->c2f_idx, hc2f_idx.size())) ||
      (rc = dalloc(&R->cv, 5 * (size_t)N)) || (rc = dalloc(&R->rhs_sum, 5 * (size_t)N)) ||
      (rc = dalloc(&R->dv, 3 * (size_t)N)) || (rc = dalloc(&R->rhs, 5 * (size_t)N)) ||
      (rc = dalloc(&R->dt_part, R->n_dt_parts)) || (rc = dalloc(&R->sc, K)) ||
      (rc = dalloc(&R->srad, FB)) || (rc = dalloc(&R->gas_tab, K)) || (rc = dalloc(&R->par, K)) ||
      (rc = dalloc(&R->cell0, K + 1)) || (rc = dalloc(&R->bface0, K + 1)) ||
      (rc = dalloc(&R->dtmin, (size_t)K * DTMIN_STRIDE)) || (rc = dalloc(&R->pbc, P)) || (rc = dalloc(&R->popt, P)) ||
      (rc = dalloc(&R->porder2, P)) || (rc = dalloc(&R->ppout, P)) || (rc = dalloc(&R->popt2, P)) ||
      (rc = dalloc(&R->pvals, (size_t)5 * P)) || (rc = dalloc(&R->bptr, P + 1)) ||
      (K > 1 && (rc = dalloc(&R->cid, N)))) {
    free_region(R);
    return rc;
  }
  if (R->rk == 1 || R->rk == 4) {
    if ((rc = dalloc(&R->cv_old, 5 * (size_t)N))) { free_region(R); return rc; }
  }
  {
    hipError_t e = hipMalloc(&R->flux, fbytes * FB + 16);
    if (e != hipSuccess) { free_region(R); return fail(RFLU_ERR_NOMEM, "hipMalloc flux"); }
  }
  if (R->order == 2) {
    if ((rc = dalloc(&R->prim, 5 * (size_t)N)) || (rc = dalloc(&R->sten_ptr, N + 1)) ||
        (rc = dalloc(&R->sten_idx, hsten_idx.size())) || (rc = dalloc(&R->moms, 6 * (size_t)N)) ||
        (rc = dalloc(&R->epsq, N))) {
      free_region(R);
      return rc;
    }
    hipError_t e = hipMalloc(&R->grad, 15 * (size_t)N * gsz);
    if (e == hipSuccess && R->reconst == RFLU_RECONST_WENO_XYZ) e = hipMalloc(&R->grad2, 15 * (size_t)N * gsz);
    if (e != hipSuccess) { free_region(R); return fail(RFLU_ERR_NOMEM, "hipMalloc grad"); }
  } else {
    // never read at order 1, but kernels take a valid pointer
    if ((rc = dalloc(&R->prim, 5))) { free_region(R); return rc; }
    hipError_t e = hipMalloc(&R->grad, 64);
    if (e != hipSuccess) { free_region(R); return fail(RFLU_ERR_NOMEM, "hipMalloc grad"); }
  }
  if ((rc = dput(R->f2c, hf2c, R->stream)) || (rc = dput(R->fn, hfn, R->stream)) ||
      (rc = dput(R->fc, hfc, R->stream)) || (rc = dput(R->cell, hcell, R->stream)) ||
      (rc = dput(R->bfn, hbfn, R->stream)) || (rc = dput(R->bfc, hbfc, R->stream)) ||
      (rc = dput(R->bf2c, hbf2c, R->stream)) || (rc = dput(R->bpatch, hbpatch, R->stream)) ||
      (rc = dput(R->c2f_ptr, hc2f_ptr, R->stream)) || (rc = dput(R->c2f_idx, hc2f_idx, R->stream)) ||
      (rc = dput(R->gas_tab, hgas, R->stream)) || (rc = dput(R->par, R->par_h, R->stream)) ||
      (rc = dput(R->cell0, hcell0, R->stream)) || (rc = dput(R->bface0, hbface0, R->stream)) ||
      (rc = dput(R->pbc, hpbc, R->stream)) || (rc = dput(R->popt, hpopt, R->stream)) ||
      (rc = dput(R->porder2, hporder2, R->stream)) || (rc = dput(R->ppout, hppout, R->stream)) ||
      (rc = dput(R->popt2, hpopt2, R->stream)) ||
      (rc = dput(R->pvals, hpvals, R->stream)) || (rc = dput(R->bptr, R->bptr_h, R->stream)) ||
      (K > 1 && (rc = dput(R->cid, hcid, R->stream)))) {
    free_region(R);
    return rc;
  }
  if (R->order == 2) {
    if ((rc = dput(R->sten_ptr, hsten_ptr, R->stream)) || (rc = dput(R->sten_idx, hsten_idx, R->stream)) ||
        (rc = dput(R->moms, hmoms, R->stream)) || (rc = dput(R->epsq, hepsq, R->stream))) {
      free_region(R);
      return rc;
    }
    if (!hvnb.empty()) {   // the limiters' neighbour lists (Venkatakrishnan, Barth-Jespersen)
      if ((rc = dalloc(&R->vnb, hvnb.size())) || (rc = dalloc(&R->vgeo, hvgeo.size())) ||
          (rc = dput(R->vnb, hvnb, R->stream)) || (rc = dput(R->vgeo, hvgeo, R->stream))) {
        free_region(R);
        return rc;
      }
    }
    if (R->use_lds) {
      if ((rc = dalloc(&R->tile_uptr, htile_uptr.size())) || (rc = dalloc(&R->tile_uidx, htile_uidx.size())) ||
          (rc = dalloc(&R->tile_lptr, htile_lptr.size())) || (rc = dalloc(&R->tile_loc, htile_loc.size())) ||
          (rc = dput(R->tile_uptr, htile_uptr, R->stream)) || (rc = dput(R->tile_uidx, htile_uidx, R->stream)) ||
          (rc = dput(R->tile_lptr, htile_lptr, R->stream)) || (rc = dput(R->tile_loc, htile_loc, R->stream)) ||
          (rc = dalloc(&R->tile_vloc, htile_vloc.size())) || (rc = dput(R->tile_vloc, htile_vloc, R->stream))) {
        free_region(R);
        return rc;
      }
      R->tile_loc_len = htile_loc.size();
    }
  }
  if (!hcon.ecell.empty()) {
    R->con_n = (int)hcon.ecell.size();
    R->con_R = (int)hcon.rcell.size();
    std::vector<double> rw;
    rw.reserve(3 * (size_t)R->con_R);
    for (int c = 0; c < 3; ++c) rw.insert(rw.end(), hcon.rw[c].begin(), hcon.rw[c].end());
    if ((rc = dalloc(&R->con_ecell, R->con_n)) || (rc = dalloc(&R->con_evar, R->con_n)) ||
        (rc = dalloc(&R->con_rptr, R->con_n + 1)) || (rc = dalloc(&R->con_rcell, R->con_R)) ||
        (rc = dalloc(&R->con_rval, R->con_R)) || (rc = dalloc(&R->con_rw, 3 * (size_t)R->con_R)) ||
        (rc = dput(R->con_ecell, hcon.ecell, R->stream)) || (rc = dput(R->con_evar, hcon.evar, R->stream)) ||
        (rc = dput(R->con_rptr, hcon.rptr, R->stream)) || (rc = dput(R->con_rcell, hcon.rcell, R->stream)) ||
        (rc = dput(R->con_rval, hcon.rval, R->stream)) || (rc = dput(R->con_rw, rw, R->stream))) {
      free_region(R);
      return rc;
    }
  }
  hipLaunchKernelGGL(k_fill_scalars, dim3(nblk(K, 64)), dim3(64), 0, R->stream, R->sc, 0, K, 0.0, R->dtmin);
  if (hipMemsetAsync(R->srad, 0, FB * sizeof(double), R->stream) != hipSuccess ||
      hipStreamSynchronize(R->stream) != hipSuccess) {
    free_region(R);
    return fail(RFLU_ERR_HIP, "region upload failed");
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_regions.push_back(R);
  *handle = (int32_t)g_regions.size();
  return RFLU_OK;
}

}  // namespace

// Dependent variables of every cell (and the primitive copy at order 2).
int launch_set_dv(Region* R) {
  Mesh M = R->mesh();
  State S = R->state();
  hipLaunchKernelGGL(k_set_dv, dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M, S, R->gas, R->cases());
  if (R->order == 2)
    hipLaunchKernelGGL(k_set_prim, dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M, S);
  HIP_TRY(hipGetLastError());
  return RFLU_OK;
}

// Launches what rflu_gpu_rk_update and rflu_gpu_set_dependent_vars left pending, as
// separate kernels (the state they would have produced, before any other use).
int flush_pending(Region* R) {
  if (R->pend_stage == 0) return 0;
  const int s = R->pend_stage;
  const bool dv = R->pend_dv;
  R->pend_stage = 0;
  R->pend_dv = false;
  int rc = launch_update(R, s, false, R->pend_dt);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return dv ? launch_set_dv(R) : 0;
}

extern "C" {

const char* rflu_gpu_last_error(void) { return g_err.c_str(); }

int rflu_gpu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int rflu_gpu_create_region(const rflu_region_desc* d, int32_t* handle) {
  return create_cases(1, d, handle);
}

int rflu_gpu_create_batch(int32_t n_cases, const rflu_region_desc* descs, int32_t* handle) {
  return create_cases(n_cases, descs, handle);
}

int rflu_gpu_destroy_region(int32_t h) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_regions[h - 1] = nullptr;
  }
  free_region(R);
  return RFLU_OK;
}

int rflu_gpu_upload_state(int32_t h, const double* cv, const double* dv, double time) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!cv) return fail(RFLU_ERR_ARG, "null cv");
  const int N = R->N;
  std::vector<double> hc((size_t)5 * N);
  for (int c = 0; c < N; ++c)
    for (int k = 0; k < 5; ++k) hc[(size_t)k * N + c] = cv[5 * (size_t)R->new2old[c] + k];
  HIP_TRY(hipMemcpyAsync(R->cv, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice, R->stream));
  if (dv) {
    std::vector<double> hd((size_t)3 * N);
    for (int c = 0; c < N; ++c)
      for (int k = 0; k < 3; ++k) hd[(size_t)k * N + c] = dv[3 * (size_t)R->new2old[c] + k];
    HIP_TRY(hipMemcpyAsync(R->dv, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice, R->stream));
    HIP_TRY(hipStreamSynchronize(R->stream));
  } else {
    Mesh M = R->mesh();
    State S = R->state();
    hipLaunchKernelGGL(k_set_dv, dim3(nblk(N, 256)), dim3(256), 0, R->stream, M, S, R->gas, R->cases());
  }
  if (R->order == 2) {
    Mesh M = R->mesh();
    State S = R->state();
    hipLaunchKernelGGL(k_set_prim, dim3(nblk(N, 256)), dim3(256), 0, R->stream, M, S);
  }
  hipLaunchKernelGGL(k_fill_scalars, dim3(nblk(R->K, 64)), dim3(64), 0, R->stream, R->sc, 0, R->K, time, R->dtmin);
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

int rflu_gpu_download_state(int32_t h, double* cv, double* dv) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  const int N = R->N;
  std::vector<double> hc((size_t)5 * N), hd((size_t)3 * N);
  HIP_TRY(hipMemcpyAsync(hc.data(), R->cv, hc.size() * sizeof(double), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipMemcpyAsync(hd.data(), R->dv, hd.size() * sizeof(double), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  if (cv)
    for (int c = 0; c < N; ++c)
      for (int k = 0; k < 5; ++k) cv[5 * (size_t)R->new2old[c] + k] = hc[(size_t)k * N + c];
  if (dv)
    for (int c = 0; c < N; ++c)
      for (int k = 0; k < 3; ++k) dv[3 * (size_t)R->new2old[c] + k] = hd[(size_t)k * N + c];
  return RFLU_OK;
}

int rflu_gpu_time_step(int32_t h, double current_time, double* dt_out) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (R->K > 1) return fail(RFLU_ERR_ARG, "a batch of cases advances through rflu_gpu_time_steps");
  if (!(current_time >= 0.0)) return fail(RFLU_ERR_ARG, "current_time must be >= 0");
  int rc = launch_time_step(R, current_time);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  if (dt_out) {
    // the step's one read of the device: it also reports a check that failed in an
    // earlier stage (the region's state stopped there)
    Scalars s;
    if ((rc = read_scalars(R, &s))) return rc;
    *dt_out = s.dt;
    if ((rc = status_from(s))) return rc;
  }
  return RFLU_OK;
}

int rflu_gpu_rk_init(int32_t h, int32_t istage) {
  // cvOld and rhsSum are kept by the update kernel: the stage-start state is the
  // current cv, and RK4 / SSP-RK3 save the step-start state at stage 1.  diss is zero
  // for inviscid runs, so rhs starts at zero inside the gather.
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (istage < 1 || istage > nstages(R->rk)) return fail(RFLU_ERR_ARG, "bad stage");
  return RFLU_OK;
}

int rflu_gpu_cell_gradients(int32_t h) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  int rc = launch_gradients(R);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return RFLU_OK;
}

int rflu_gpu_convective_fluxes(int32_t h, int32_t istage) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (istage < 1 || istage > nstages(R->rk)) return fail(RFLU_ERR_ARG, "bad stage");
  int rc = launch_fluxes(R, istage);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return RFLU_OK;
}

int rflu_gpu_rk_update(int32_t h, int32_t istage, double dt_min) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (R->K > 1) return fail(RFLU_ERR_ARG, "a batch of cases advances through rflu_gpu_time_steps");
  if (istage < 1 || istage > nstages(R->rk)) return fail(RFLU_ERR_ARG, "bad stage");
  if (!(dt_min > 0.0)) return fail(RFLU_ERR_ARG, "dt_min must be > 0");
  int rc = flush_pending(R);
  if (rc) return rc;
  R->pend_stage = istage;
  R->pend_dt = dt_min;
  R->pend_dv = false;
  return RFLU_OK;
}

int rflu_gpu_set_dependent_vars(int32_t h) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (R->pend_stage > 0 && !R->pend_dv) {
    R->pend_dv = true;
    return RFLU_OK;
  }
  int rc = flush_pending(R);
  if (rc) return rc;
  return launch_set_dv(R);
}

int rflu_gpu_after_update(int32_t h, int32_t istage) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (istage < 1 || istage > nstages(R->rk)) return fail(RFLU_ERR_ARG, "bad stage");
  if (R->pend_stage > 0 && R->pend_dv) {
    // update, dependent variables, checks and the primitive copy in one pass
    const int s = R->pend_stage;
    R->pend_stage = 0;
    R->pend_dv = false;
    int rc = launch_update(R, s, true, R->pend_dt, false);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return RFLU_OK;
  }
  int rc = flush_pending(R);
  if (rc) return rc;
  Mesh M = R->mesh();
  State S = R->state();
  if (R->order == 2)
    hipLaunchKernelGGL((k_after_update<true>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M, S, R->gas, R->cases(), R->sc);
  else
    hipLaunchKernelGGL((k_after_update<false>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M, S, R->gas, R->cases(), R->sc);
  HIP_TRY(hipGetLastError());
  return RFLU_OK;
}

int rflu_gpu_check_status(int32_t h) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  int rc = flush_pending(R);
  if (rc) return rc;
  Scalars s;
  if ((rc = read_scalars(R, &s))) return rc;
  return status_from(s);
}

int rflu_gpu_set_borders(int32_t h, int32_t n_borders, const int32_t* nbr_handle,
                         const int32_t* nbr_border, const int32_t* send_ptr,
                         const int32_t* send_cells, const int32_t* recv_ptr,
                         const int32_t* recv_cells) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (R->K != 1) return fail(RFLU_ERR_ARG, "borders belong to single regions");
  if (n_borders < 0 || (n_borders > 0 && (!nbr_handle || !nbr_border || !send_ptr || !recv_ptr)))
    return fail(RFLU_ERR_ARG, "bad border arguments");
  if (n_borders > 0 && (send_ptr[0] != 0 || recv_ptr[0] != 0))
    return fail(RFLU_ERR_ARG, "send_ptr[0] and recv_ptr[0] must be 0");
  std::vector<int> o2n(R->N);
  for (int i = 0; i < R->N; ++i) o2n[R->new2old[i]] = i;
  std::vector<Border> nb(n_borders);
  auto drop = [&nb]() {
    for (Border& b : nb) {
      if (b.send) (void)hipFree(b.send);
      if (b.recv) (void)hipFree(b.recv);
    }
  };
  for (int b = 0; b < n_borders; ++b) {
    Border& B = nb[b];
    B.nbr = nbr_handle[b];
    B.nbr_border = nbr_border[b] - R->base;
    B.nsend = send_ptr[b + 1] - send_ptr[b];
    B.nrecv = recv_ptr[b + 1] - recv_ptr[b];
    if (!get(B.nbr) || B.nbr_border < 0 || B.nsend < 0 || B.nrecv < 0 ||
        (B.nsend > 0 && !send_cells) || (B.nrecv > 0 && !recv_cells)) {
      drop();
      return fail(RFLU_ERR_ARG, "border " + std::to_string(b) + ": bad neighbour or lists");
    }
    std::vector<int> hs(B.nsend), hr(B.nrecv);
    for (int e = 0; e < B.nsend; ++e) {
      const int c = send_cells[send_ptr[b] + e] - R->base;
      if (c < 0 || c >= R->nact) { drop(); return fail(RFLU_ERR_ARG, "send cells must be actual cells"); }
      hs[e] = o2n[c];
    }
    for (int e = 0; e < B.nrecv; ++e) {
      const int c = recv_cells[recv_ptr[b] + e] - R->base;
      if (c < R->nact || c >= R->N) { drop(); return fail(RFLU_ERR_ARG, "receive cells must be virtual cells"); }
      hr[e] = o2n[c];
    }
    int rc;
    if ((rc = dalloc(&B.send, hs.size())) || (rc = dalloc(&B.recv, hr.size())) ||
        (rc = dput(B.send, hs, R->stream)) || (rc = dput(B.recv, hr, R->stream))) {
      drop();
      return rc;
    }
  }
  if (hipStreamSynchronize(R->stream) != hipSuccess) { drop(); return fail(RFLU_ERR_HIP, "border upload failed"); }
  for (Border& b : R->borders) {
    if (b.send) (void)hipFree(b.send);
    if (b.recv) (void)hipFree(b.recv);
  }
  R->borders.swap(nb);
  return RFLU_OK;
}

int rflu_gpu_set_border_transform(int32_t h, int32_t border, const double* m) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  const int b = border - R->base;
  if (b < 0 || b >= (int)R->borders.size()) return fail(RFLU_ERR_ARG, "no such border");
  Border& B = R->borders[b];
  if (!m) {
    B.turn = false;
    return RFLU_OK;
  }
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(m[k])) return fail(RFLU_ERR_ARG, "transform entries must be finite");
  for (int k = 0; k < 9; ++k) B.xf.m[k] = m[k];
  B.turn = true;
  return RFLU_OK;
}

int rflu_gpu_exchange(int32_t n_regions, const int32_t* handles, int32_t what) {
  if (n_regions < 1 || !handles) return fail(RFLU_ERR_ARG, "bad arguments");
  if (what != RFLU_XCHG_STATE && what != RFLU_XCHG_GRAD) return fail(RFLU_ERR_ARG, "what must be RFLU_XCHG_STATE or RFLU_XCHG_GRAD");
  std::vector<Region*> rs(n_regions);
  for (int i = 0; i < n_regions; ++i) {
    rs[i] = get(handles[i]);
    if (!rs[i]) return fail(RFLU_ERR_HANDLE, "bad region handle");
    if (int frc = flush_pending(rs[i])) return frc;
    if (rs[i]->K != 1) return fail(RFLU_ERR_ARG, "borders belong to single regions");
    if (what == RFLU_XCHG_GRAD && rs[i]->order != 2) return fail(RFLU_ERR_ARG, "gradients exist at order 2 only");
  }
  // check every pairing before anything moves
  for (int i = 0; i < n_regions; ++i)
    for (const Border& b : rs[i]->borders) {
      Region* D = get(b.nbr);
      if (!D || std::find(handles, handles + n_regions, b.nbr) == handles + n_regions)
        return fail(RFLU_ERR_ARG, "a border's neighbour is not in the exchange");
      if (b.nbr_border >= (int)D->borders.size() || D->borders[b.nbr_border].nbr != handles[i] ||
          D->borders[b.nbr_border].nrecv != b.nsend)
        return fail(RFLU_ERR_ARG, "border lists do not match (ERR_BUFFERDIM_MISMATCH)");
      if (what == RFLU_XCHG_GRAD && (grad32(D) != grad32(rs[i]) || D->reconst != rs[i]->reconst))
        return fail(RFLU_ERR_ARG, "neighbouring regions need the same precision and reconstruction");
    }
  // a barrier: every region's queued work ends before its cells are read or written
  for (Region* R : rs) HIP_TRY(hipStreamSynchronize(R->stream));
  for (int i = 0; i < n_regions; ++i) {
    Region* R = rs[i];
    for (const Border& b : R->borders) {
      if (b.nsend == 0) continue;
      Region* D = get(b.nbr);
      const int* didx = D->borders[b.nbr_border].recv;
      const dim3 g(nblk(b.nsend, 256)), bs(256);
      const Border& rb = D->borders[b.nbr_border];
      if (rb.turn) {
        // symmetry and periodic virtual cells: momentum and gradients turned
        // (RFLU_RELP_TransformWrapper)
        if (what == RFLU_XCHG_STATE)
          hipLaunchKernelGGL((k_halo_turn<double>), g, bs, 0, R->stream, (const double*)R->cv, R->N,
                             D->cv, D->N, 5, (const int*)b.send, didx, b.nsend, rb.xf);
        else if (grad32(R))
          hipLaunchKernelGGL((k_halo_turn<float>), g, bs, 0, R->stream, (const float*)R->grad_final(), R->N,
                             (float*)D->grad_final(), D->N, 15, (const int*)b.send, didx, b.nsend, rb.xf);
        else
          hipLaunchKernelGGL((k_halo_turn<double>), g, bs, 0, R->stream, (const double*)R->grad_final(), R->N,
                             (double*)D->grad_final(), D->N, 15, (const int*)b.send, didx, b.nsend, rb.xf);
      } else if (what == RFLU_XCHG_STATE)
        hipLaunchKernelGGL((k_halo_copy<double>), g, bs, 0, R->stream, (const double*)R->cv, R->N,
                           D->cv, D->N, 5, (const int*)b.send, didx, b.nsend);
      else if (grad32(R))
        hipLaunchKernelGGL((k_halo_copy<float>), g, bs, 0, R->stream, (const float*)R->grad_final(), R->N,
                           (float*)D->grad_final(), D->N, 15, (const int*)b.send, didx, b.nsend);
      else
        hipLaunchKernelGGL((k_halo_copy<double>), g, bs, 0, R->stream, (const double*)R->grad_final(), R->N,
                           (double*)D->grad_final(), D->N, 15, (const int*)b.send, didx, b.nsend);
    }
  }
  HIP_TRY(hipGetLastError());
  for (Region* R : rs) HIP_TRY(hipStreamSynchronize(R->stream));
  if (what == RFLU_XCHG_STATE)
    for (Region* R : rs) {
      if (R->nact == R->N) continue;
      Mesh M = R->mesh();
      State S = R->state();
      const dim3 g(nblk(R->N - R->nact, 256)), bs(256);
      if (R->order == 2)
        hipLaunchKernelGGL((k_set_vars_range<true>), g, bs, 0, R->stream, M, S, R->gas, R->nact, R->N);
      else
        hipLaunchKernelGGL((k_set_vars_range<false>), g, bs, 0, R->stream, M, S, R->gas, R->nact, R->N);
    }
  HIP_TRY(hipGetLastError());
  return RFLU_OK;
}

int rflu_gpu_residual(int32_t h, double* rhs) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!rhs) return fail(RFLU_ERR_ARG, "null rhs");
  int rc;
  if ((rc = launch_gradients(R)) || (rc = launch_fluxes(R, 0))) return rc;
  Mesh M = R->mesh();
  with_fbuf(R, [&](auto* tag) {
    using FT = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL((k_gather_rhs<FT>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M,
                       (const FT*)R->flux, R->rhs);
    return 0;
  });
  HIP_TRY(hipGetLastError());
  const int N = R->N;
  std::vector<double> hr((size_t)5 * N);
  HIP_TRY(hipMemcpyAsync(hr.data(), R->rhs, hr.size() * sizeof(double), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  for (int c = 0; c < N; ++c)
    for (int k = 0; k < 5; ++k) rhs[5 * (size_t)R->new2old[c] + k] = hr[(size_t)k * N + c];
  return RFLU_OK;
}

int rflu_gpu_time_steps(int32_t h, int64_t nsteps, int32_t use_graph) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (nsteps < 0) return fail(RFLU_ERR_ARG, "nsteps < 0");
  if (R->nact < R->N)
    return fail(RFLU_ERR_ARG, "a region with virtual cells advances through the per-callee "
                              "entry points and rflu_gpu_exchange");
  int rc;
  if (R->persistent && (rc = launch_persistent(R, true))) return rc;
  if (use_graph && !R->profiling) {
    if (R->graph_dirty || !R->graph) {
      if (R->graph) { (void)hipGraphExecDestroy(R->graph); R->graph = nullptr; }
      hipGraph_t g;
      HIP_TRY(hipStreamBeginCapture(R->stream, hipStreamCaptureModeThreadLocal));
      rc = launch_step(R);
      hipError_t ce = hipStreamEndCapture(R->stream, &g);
      if (rc) return rc;
      if (ce != hipSuccess) return fail(RFLU_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&R->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(RFLU_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
      R->graph_dirty = false;
    }
    for (int64_t s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(R->graph, R->stream));
  } else {
    for (int64_t s = 0; s < nsteps; ++s)
      if ((rc = launch_step(R))) return rc;
  }
  HIP_TRY(hipGetLastError());
  if ((rc = persistent_check(R))) return rc;
  Scalars sc;
  if ((rc = read_scalars(R, &sc))) return rc;
  if (R->profiling) collect_events(R);
  return status_from(sc);
}

int rflu_gpu_set_end_time(int32_t h, double end_time) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!(end_time > 0.0)) return fail(RFLU_ERR_ARG, "end_time must be > 0");
  R->end_time = end_time;
  for (CaseParams& cp : R->par_h) cp.end_time = end_time;
  int rc = dput(R->par, R->par_h, R->stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(R->stream));
  R->graph_dirty = true;
  return RFLU_OK;
}

int rflu_gpu_get_stats(int32_t h, rflu_gpu_stats* st) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!st) return fail(RFLU_ERR_ARG, "null stats");
  Scalars s;
  int rc = read_scalars(R, &s);
  if (rc) return rc;
  st->n_invalid = (int64_t)s.n_invalid;
  st->n_negative = (int64_t)s.n_negative;
  st->n_bounds_fix = (int64_t)s.n_bounds_fix;
  st->n_pos_fix = (int64_t)s.n_pos_fix;
  st->mass_in = s.mass_in;
  st->mass_out = s.mass_out;
  st->time = s.time;
  st->dt = s.dt;
  st->steps = (int64_t)s.steps;
  st->finished = s.finished;
  return RFLU_OK;
}

int rflu_gpu_get_case_stats(int32_t h, int32_t icase, rflu_gpu_stats* st) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!st || icase < 0 || icase >= R->K) return fail(RFLU_ERR_ARG, "bad case or null stats");
  std::vector<Scalars> v;
  int rc = read_cases(R, v);
  if (rc) return rc;
  const Scalars& s = v[icase];
  st->n_invalid = (int64_t)s.n_invalid;
  st->n_negative = (int64_t)s.n_negative;
  st->n_bounds_fix = (int64_t)s.n_bounds_fix;
  st->n_pos_fix = (int64_t)s.n_pos_fix;
  st->mass_in = s.mass_in;
  st->mass_out = s.mass_out;
  st->time = s.time;
  st->dt = s.dt;
  st->steps = (int64_t)s.steps;
  st->finished = s.finished;
  return RFLU_OK;
}

int rflu_gpu_reset_stats(int32_t h) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  hipLaunchKernelGGL(k_reset_counters, dim3(nblk(R->K, 64)), dim3(64), 0, R->stream, R->sc, R->K);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

int rflu_gpu_synchronize(int32_t h) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

int rflu_gpu_set_launch(int32_t h, int32_t face_block, int32_t cell_block, int32_t grad_block) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  auto ok = [](int b) { return b == 0 || (b >= 64 && b <= 1024 && b % 64 == 0); };
  if (!ok(face_block) || !ok(cell_block) || !ok(grad_block))
    return fail(RFLU_ERR_ARG, "block sizes must be multiples of 64 in [64, 1024]");
  if (face_block) R->face_block = face_block;
  if (cell_block) R->cell_block = cell_block;
  if (grad_block) R->grad_block = grad_block;
  R->graph_dirty = true;
  return RFLU_OK;
}

int rflu_gpu_time_kernel(int32_t h, int32_t which, int32_t reps, double* ms) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!ms || reps < 1 || which < 0 || which > 5) return fail(RFLU_ERR_ARG, "bad arguments");
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  bool prof = R->profiling;
  R->profiling = false;
  int rc = 0;
  HIP_TRY(hipEventRecord(a, R->stream));
  for (int r = 0; r < reps && !rc; ++r) {
    switch (which) {
      case 0: rc = launch_fluxes(R, 2); break;
      case 1: rc = launch_update(R, 1, false, 1.0e-30); break;
      case 2: rc = launch_gradients(R); break;
      case 3: rc = grad32(R) ? launch_blend_t<float>(R) : launch_blend_t<double>(R); break;
      case 4: rc = launch_time_step(R); break;
      default:
        if (!(rc = launch_gradients(R)) && !(rc = launch_fluxes(R, 2))) rc = launch_update(R, 1, true);
        break;
    }
  }
  R->profiling = prof;
  HIP_TRY(hipEventRecord(b, R->stream));
  HIP_TRY(hipEventSynchronize(b));
  float t = 0;
  HIP_TRY(hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  *ms = t / reps;
  return RFLU_OK;
}

}  // extern "C"

namespace {

// Stencil weights of the tiles' [member][lane] layout, formed on the device.
int form_weights(Region* R, int kind) {
  const size_t gsz = grad32(R) ? sizeof(float) : sizeof(double);
  const size_t L = R->tile_loc_len;
  hipError_t e = hipMalloc(&R->gw, (kind == 2 ? 1 : 3) * L * gsz + 16);
  if (e != hipSuccess) {
    R->gw = nullptr;
    R->gw_kind = 0;
    return fail(RFLU_ERR_NOMEM, std::string("hipMalloc stencil weights: ") + hipGetErrorString(e));
  }
  R->gw_kind = kind;
  Mesh M = R->mesh();
  Tiles T = region_tiles(R);
  if (kind == 2) {
    if (grad32(R))
      hipLaunchKernelGGL((k_grad_terms<float>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream,
                         M, T, (float*)R->gw);
    else
      hipLaunchKernelGGL((k_grad_terms<double>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream,
                         M, T, (double*)R->gw);
  } else if (grad32(R))
    hipLaunchKernelGGL((k_grad_weights<float>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream,
                       M, T, L, (float*)R->gw);
  else
    hipLaunchKernelGGL((k_grad_weights<double>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream,
                       M, T, L, (double*)R->gw);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(R->stream));
  return 0;
}

// Tiles of another size: the stencils and limiter neighbours come back from the
// device, the tables are rebuilt on the host and replace the old ones (and the stored
// weights, whose layout follows the tiles).  Tiles that would not fit LDS leave the
// region as it was.
// Face tiles of ts cells (0: none).  Needs the faces ordered by their lower cell, as
// the renumbering leaves them; a tile's records must fit 80 KiB of LDS so that
// several blocks share a CU.
void free_face_tiles(Region* R) {
  void* old[] = {R->ft_fptr, R->ft_eptr, R->ft_eidx, R->ft_fslot};
  for (void* b : old)
    if (b) (void)hipFree(b);
  R->ft_fptr = R->ft_eptr = R->ft_eidx = nullptr;
  R->ft_fslot = nullptr;
  R->face_tile = R->n_face_tiles = R->face_cap = R->face_fcap = 0;
}

int set_face_tiles(Region* R, int ts) {
  if (ts == 0) {
    free_face_tiles(R);
    return 0;
  }
  const int N = R->N, F = R->F;
  std::vector<int2> f2c(F);
  HIP_TRY(hipStreamSynchronize(R->stream));
  if (F > 0) HIP_TRY(hipMemcpy(f2c.data(), R->f2c, F * sizeof(int2), hipMemcpyDeviceToHost));
  const int nt = (N + ts - 1) / ts;
  std::vector<int> fptr(nt + 1, 0), eptr(nt + 1, 0), eidx;
  std::vector<unsigned> fslot(F);
  int prev = 0;
  for (int f = 0; f < F; ++f) {
    const int lo = std::min(f2c[f].x, f2c[f].y) / ts;
    if (lo < prev) return fail(RFLU_ERR_ARG, "faces are not ordered by their lower cell");
    prev = lo;
    fptr[lo + 1]++;
  }
  for (int t = 0; t < nt; ++t) fptr[t + 1] += fptr[t];
  int cap = 0, fcap = 0;
  std::vector<int> ext;
  for (int t = 0; t < nt; ++t) {
    fcap = std::max(fcap, fptr[t + 1] - fptr[t]);
    const int c0 = t * ts, nown = std::min(ts, N - c0);
    ext.clear();
    for (int f = fptr[t]; f < fptr[t + 1]; ++f) {
      const int hi = std::max(f2c[f].x, f2c[f].y);
      if (hi >= c0 + nown) ext.push_back(hi);
    }
    std::sort(ext.begin(), ext.end());
    ext.erase(std::unique(ext.begin(), ext.end()), ext.end());
    auto slot = [&](int c) {
      if (c >= c0 && c < c0 + nown) return c - c0;
      return nown + (int)(std::lower_bound(ext.begin(), ext.end(), c) - ext.begin());
    };
    for (int f = fptr[t]; f < fptr[t + 1]; ++f)
      fslot[f] = (unsigned)slot(f2c[f].x) | ((unsigned)slot(f2c[f].y) << 16);
    eidx.insert(eidx.end(), ext.begin(), ext.end());
    eptr[t + 1] = (int)eidx.size();
    cap = std::max(cap, nown + (int)ext.size());
  }
  const bool fp32 = R->prec == RFLU_PREC_MIXED || R->prec == RFLU_PREC_FP32_GRAD;
  const size_t bytes = fp32 ? face_tile_lds<float>(cap, true) : face_tile_lds<double>(cap, true);
  if (cap > 65535 || bytes > 80 * 1024)   // two blocks per CU at most 80 KiB each
    return fail(RFLU_ERR_ARG, "face tiles of that size do not fit LDS");
  free_face_tiles(R);
  int rc;
  if ((rc = dalloc(&R->ft_fptr, fptr.size())) || (rc = dalloc(&R->ft_eptr, eptr.size())) ||
      (rc = dalloc(&R->ft_eidx, std::max<size_t>(eidx.size(), 1))) ||
      (rc = dalloc(&R->ft_fslot, std::max<size_t>(fslot.size(), 1))) ||
      (rc = dput(R->ft_fptr, fptr, R->stream)) || (rc = dput(R->ft_eptr, eptr, R->stream)) ||
      (!eidx.empty() && (rc = dput(R->ft_eidx, eidx, R->stream))) ||
      (!fslot.empty() && (rc = dput(R->ft_fslot, fslot, R->stream)))) {
    free_face_tiles(R);
    return rc;
  }
  HIP_TRY(hipStreamSynchronize(R->stream));
  R->face_tile = ts;
  R->n_face_tiles = nt;
  R->face_cap = cap;
  R->face_fcap = fcap;
  return 0;
}

int rebuild_tiles(Region* R, int ts) {
  const int N = R->N;
  const bool venkat = R->reconst == RFLU_RECONST_LIM_VENKAT;
  std::vector<int> sp(N + 1), si, vnb;
  HIP_TRY(hipStreamSynchronize(R->stream));
  HIP_TRY(hipMemcpy(sp.data(), R->sten_ptr, (N + 1) * sizeof(int), hipMemcpyDeviceToHost));
  si.resize(sp[N]);
  if (!si.empty())
    HIP_TRY(hipMemcpy(si.data(), R->sten_idx, si.size() * sizeof(int), hipMemcpyDeviceToHost));
  if (venkat) {
    vnb.resize((size_t)R->vq * N);
    HIP_TRY(hipMemcpy(vnb.data(), R->vnb, vnb.size() * sizeof(int), hipMemcpyDeviceToHost));
  }
  const int old_ts = R->grad_tile, old_nt = R->n_tiles, old_cap = R->tile_cap;
  const bool old_lds = R->use_lds;
  std::vector<int> uptr, uidx, lptr;
  std::vector<unsigned short> loc, vloc;
  R->grad_tile = ts;
  build_tiles(R, sp, si, vnb, venkat, uptr, uidx, lptr, loc, vloc);
  if (!R->use_lds) {
    R->grad_tile = old_ts; R->n_tiles = old_nt; R->tile_cap = old_cap; R->use_lds = old_lds;
    return fail(RFLU_ERR_ARG, "stencil tiles of that size do not fit LDS");
  }
  int *d_uptr = nullptr, *d_uidx = nullptr, *d_lptr = nullptr;
  unsigned short *d_loc = nullptr, *d_vloc = nullptr;
  int rc;
  if ((rc = dalloc(&d_uptr, uptr.size())) || (rc = dalloc(&d_uidx, uidx.size())) ||
      (rc = dalloc(&d_lptr, lptr.size())) || (rc = dalloc(&d_loc, loc.size())) ||
      (rc = dalloc(&d_vloc, vloc.size())) || (rc = dput(d_uptr, uptr, R->stream)) ||
      (rc = dput(d_uidx, uidx, R->stream)) || (rc = dput(d_lptr, lptr, R->stream)) ||
      (rc = dput(d_loc, loc, R->stream)) || (rc = dput(d_vloc, vloc, R->stream)) ||
      hipStreamSynchronize(R->stream) != hipSuccess) {
    void* bufs[] = {d_uptr, d_uidx, d_lptr, d_loc, d_vloc};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    R->grad_tile = old_ts; R->n_tiles = old_nt; R->tile_cap = old_cap; R->use_lds = old_lds;
    return rc ? rc : fail(RFLU_ERR_HIP, "tile upload failed");
  }
  void* old[] = {R->tile_uptr, R->tile_uidx, R->tile_lptr, R->tile_loc, R->tile_vloc};
  for (void* b : old)
    if (b) (void)hipFree(b);
  R->tile_uptr = d_uptr; R->tile_uidx = d_uidx; R->tile_lptr = d_lptr;
  R->tile_loc = d_loc; R->tile_vloc = d_vloc;
  R->tile_loc_len = loc.size();
  if (R->gw) {
    HIP_TRY(hipFree(R->gw));
    R->gw = nullptr;
    if ((rc = form_weights(R, R->gw_kind))) return rc;
  }
  return 0;
}

}  // namespace

// The allocation-free tangent's work arrays live in one static array per GPU (the
// code object's own data segment, allocated when the library's kernels are loaded on
// that GPU and held until the process ends).  Host code addresses it only through the
// device address hipGetSymbolAddress returns on that GPU.
static constexpr size_t kScratchDoubles = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// The current GPU's address of `scratch` (looked up once per GPU, with it current).
static int scratch_address(double** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(RFLU_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(RFLU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}


extern "C" {

int rflu_gpu_set_option(int32_t h, int32_t option, int32_t value) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  switch (option) {
    case RFLU_OPT_GRAD_LDS:
      if (value < 0 || value > 2) return fail(RFLU_ERR_ARG, "GRAD_LDS is 0, 1 or 2");
      if (value == 2 && !grad32(R))
        return fail(RFLU_ERR_ARG, "GRAD_LDS 2 (FP32 records) needs FP32 gradients");
      R->grad_lds = value;
      break;
    case RFLU_OPT_GRAD_UNROLL:
      if (value != 1 && value != 2 && value != 4 && value != 8)
        return fail(RFLU_ERR_ARG, "GRAD_UNROLL is 1, 2, 4 or 8");
      R->grad_unroll = value;
      break;
    case RFLU_OPT_XCD:
      if (value < 1 || value > 64) return fail(RFLU_ERR_ARG, "XCD is 1..64");
      R->nxcd = value;
      break;
    case RFLU_OPT_GRAD_WEIGHTS: {
      if (value < 0 || value > 2) return fail(RFLU_ERR_ARG, "GRAD_WEIGHTS is 0, 1 or 2");
      if (R->gw && R->gw_kind == value) break;
      if (value != 0 && (R->order != 2 || !R->use_lds))
        return fail(RFLU_ERR_ARG, "GRAD_WEIGHTS needs a second-order region with LDS tiles");
      if (R->gw) {
        HIP_TRY(hipStreamSynchronize(R->stream));
        HIP_TRY(hipFree(R->gw));
        R->gw = nullptr;
        R->gw_kind = 0;
      }
      if (value == 0) break;
      int rc = form_weights(R, value);
      if (rc) return rc;
      break;
    }
    case RFLU_OPT_GRAD_FAST:
      if (value != 0 && value != 1) return fail(RFLU_ERR_ARG, "GRAD_FAST is 0 or 1");
      if (value == 1 && (R->order != 2 || !R->use_lds))
        return fail(RFLU_ERR_ARG, "GRAD_FAST needs a second-order region with LDS tiles");
      R->grad_fast = value == 1;
      break;
    case RFLU_OPT_UPDATE_TILE:
      if (value != 0 && value != 1) return fail(RFLU_ERR_ARG, "UPDATE_TILE is 0 or 1");
      R->update_tile = value == 1;
      break;
    case RFLU_OPT_FACE_TILE: {
      if (value != 0 && value != 64 && value != 128 && value != 256)
        return fail(RFLU_ERR_ARG, "FACE_TILE is 0, 64, 128 or 256");
      if (value != 0 && R->order != 2) return fail(RFLU_ERR_ARG, "FACE_TILE needs a second-order region");
      if (value == R->face_tile) break;
      int rc = set_face_tiles(R, value);
      if (rc) return rc;
      break;
    }
    case RFLU_OPT_FACE_THREADS:
      if (value != 256 && value != 512) return fail(RFLU_ERR_ARG, "FACE_THREADS is 256 or 512");
      R->face_threads = value;
      break;
    case RFLU_OPT_PERSISTENT: {
      if (value != 0 && value != 1) return fail(RFLU_ERR_ARG, "PERSISTENT is 0 or 1");
      if (value == 1) {
        std::string why = persistent_unsupported(R);
        if (!why.empty()) return fail(RFLU_ERR_ARG, why);
        if (!R->pbar) {
          if (int rc = dalloc(&R->pbar, 4)) return rc;
          HIP_TRY(hipMemset(R->pbar, 0, 4 * sizeof(unsigned)));
        }
        // a barrier wait that timed out stays reported until the option is set again
        HIP_TRY(hipStreamSynchronize(R->stream));
        HIP_TRY(hipMemset(R->pbar + 2, 0, sizeof(unsigned)));
      }
      R->persistent = value == 1;
      break;
    }
    case RFLU_OPT_GRAD_TILE: {
      if (value != 128 && value != 256 && value != 512)
        return fail(RFLU_ERR_ARG, "GRAD_TILE is 128, 256 or 512");
      if (R->order != 2) return fail(RFLU_ERR_ARG, "GRAD_TILE needs a second-order region");
      if (value == R->grad_tile) break;
      int rc = rebuild_tiles(R, value);
      if (rc) return rc;
      break;
    }
    default:
      return fail(RFLU_ERR_ARG, "unknown option");
  }
  R->graph_dirty = true;
  return RFLU_OK;
}

int rflu_gpu_get_option(int32_t h, int32_t option, int32_t* value) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (!value) return fail(RFLU_ERR_ARG, "null value");
  switch (option) {
    case RFLU_OPT_GRAD_LDS:
      // the moment form and stored weights stage FP64-layout records (GRAD_LDS 1)
      *value = !R->use_lds ? 0 : (R->grad_lds == 2 && (R->grad_fast || R->gw)) ? 1 : R->grad_lds;
      break;
    case RFLU_OPT_GRAD_UNROLL: *value = R->grad_unroll; break;
    case RFLU_OPT_TILE_CAP: *value = R->tile_cap; break;
    case RFLU_OPT_XCD: *value = R->nxcd; break;
    case RFLU_OPT_GRAD_WEIGHTS: *value = R->gw ? R->gw_kind : 0; break;
    case RFLU_OPT_GRAD_TILE: *value = R->grad_tile; break;
    case RFLU_OPT_FACE_TILE: *value = R->face_tile; break;
    case RFLU_OPT_UPDATE_TILE: *value = (R->update_tile && R->face_tile > 0) ? 1 : 0; break;
    case RFLU_OPT_GRAD_FAST: *value = (R->grad_fast && R->use_lds && R->grad_lds) ? 1 : 0; break;
    case RFLU_OPT_PERSISTENT: *value = R->persistent ? 1 : 0; break;
    case RFLU_OPT_FACE_THREADS: *value = R->face_threads; break;
    default: return fail(RFLU_ERR_ARG, "unknown option");
  }
  return RFLU_OK;
}

// Per-kernel timing of the launches made by rflu_gpu_time_steps (HIP events on the
// region's stream around each launch; graphs are not used while on).
// ---- differentiation path: device-resident states in the internal cell order ------

int rflu_gpu_cell_order(int32_t h, int32_t* new2old) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (!new2old) return fail(RFLU_ERR_ARG, "null array");
  for (int i = 0; i < R->N; ++i) new2old[i] = R->new2old[i];
  return RFLU_OK;
}

int rflu_gpu_cell_volumes(int32_t h, double* vol) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (!vol) return fail(RFLU_ERR_ARG, "null array");
  std::vector<double4> c(R->N);
  HIP_TRY(hipMemcpyAsync(c.data(), R->cell, c.size() * sizeof(double4), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  for (int i = 0; i < R->N; ++i) vol[i] = c[i].w;
  return RFLU_OK;
}

int rflu_gpu_patch_faces(int32_t h, int32_t patch, int32_t* cells, double* areas) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (patch < 0 || patch >= R->P || !cells || !areas) return fail(RFLU_ERR_ARG, "bad arguments");
  const int b0 = R->bptr_h[patch], nb = R->bptr_h[patch + 1] - b0;
  if (nb == 0) return RFLU_OK;
  std::vector<double4> n(nb);
  HIP_TRY(hipMemcpyAsync(cells, R->bf2c + b0, nb * sizeof(int32_t), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipMemcpyAsync(n.data(), R->bfn + b0, nb * sizeof(double4), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  for (int k = 0; k < nb; ++k) areas[k] = n[k].w;
  return RFLU_OK;
}

int rflu_gpu_set_patch_values(int32_t h, int32_t patch, const double* vals) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (patch < 0 || patch >= R->P || !vals) return fail(RFLU_ERR_ARG, "bad arguments");
  for (int k = 0; k < 5; ++k) R->pvals_h[5 * (size_t)patch + k] = vals[k];
  HIP_TRY(hipMemcpyAsync(R->pvals + 5 * (size_t)patch, R->pvals_h.data() + 5 * (size_t)patch,
                         5 * sizeof(double), hipMemcpyHostToDevice, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

int rflu_gpu_set_state_dev(int32_t h, const double* cv) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!cv) return fail(RFLU_ERR_ARG, "null state");
  HIP_TRY(hipMemcpyAsync(R->cv, cv, (size_t)5 * R->N * sizeof(double), hipMemcpyDeviceToDevice, R->stream));
  Mesh M = R->mesh();
  State S = R->state();
  hipLaunchKernelGGL(k_set_dv, dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M, S, R->gas, R->cases());
  if (R->order == 2)
    hipLaunchKernelGGL(k_set_prim, dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M, S);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

int rflu_gpu_get_state_dev(int32_t h, double* cv) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!cv) return fail(RFLU_ERR_ARG, "null state");
  HIP_TRY(hipMemcpyAsync(cv, R->cv, (size_t)5 * R->N * sizeof(double), hipMemcpyDeviceToDevice, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

int rflu_gpu_residual_dev(int32_t h, double* rhs) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!rhs) return fail(RFLU_ERR_ARG, "null rhs");
  int rc;
  if ((rc = launch_gradients(R)) || (rc = launch_fluxes(R, 0))) return rc;
  Mesh M = R->mesh();
  with_fbuf(R, [&](auto* tag) {
    using FT = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL((k_gather_rhs<FT>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M,
                       (const FT*)R->flux, rhs);
    return 0;
  });
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

namespace {
// Whether the region's residual has derivative kernels (no allocation).
static int derivative_check(Region* R) {
  if (R->K != 1) return fail(RFLU_ERR_ARG, "derivatives are for a single region");
  if (R->nact < R->N) return fail(RFLU_ERR_ARG, "derivatives are for regions without virtual cells");
  if (R->prec != RFLU_PREC_FP64) return fail(RFLU_ERR_ARG, "derivatives are FP64");
  if (R->order == 2 && R->reconst != RFLU_RECONST_NONE && R->reconst != RFLU_RECONST_LIM_VENKAT &&
      R->reconst != RFLU_RECONST_WENO_XYZ)
    return fail(RFLU_ERR_ARG, "second-order derivatives cover no modifier, WENO-XYZ or the "
                              "Venkatakrishnan limiter");
  return 0;
}

// derivative_check, then the region's derivative buffers (allocated on first use).
static int derivative_ready(Region* R) {
  int rc = derivative_check(R);
  if (rc) return rc;
  const size_t F = R->F, B = R->B, P = R->P, N = R->N;
  const bool o2 = R->order == 2;
  if (!R->adj_g && ((rc = dalloc(&R->adj_g, (o2 ? 20 : 10) * F)) ||
                    (rc = dalloc(&R->adj_gb, (o2 ? 10 : 5) * B)) ||
                    (rc = dalloc(&R->adj_gv, 5 * B)) || (rc = dalloc(&R->adj_pv, 5 * P)) ||
                    (rc = dalloc(&R->adj_dflux, 5 * (F + B))) || (rc = dalloc(&R->adj_dvals, 5 * P))))
    return rc;
  if (o2 && !R->adj_p) {
    if ((rc = dalloc(&R->adj_p, 5 * N)) || (rc = dalloc(&R->adj_ga, 15 * N)) ||
        (rc = dalloc(&R->adj_gb2, 15 * N)) || (rc = dalloc(&R->adj_gc, 15 * N)) ||
        (rc = dalloc(&R->adj_gr, 15 * N)) ||
        (rc = dalloc(&R->adj_selv, 5 * N)) || (rc = dalloc(&R->adj_seln, 5 * N)))
      return rc;
    // transposed stencil: for every cell j, the cells whose stencil holds j, ascending
    std::vector<int> sp(N + 1), si;
    HIP_TRY(hipStreamSynchronize(R->stream));
    HIP_TRY(hipMemcpy(sp.data(), R->sten_ptr, (N + 1) * sizeof(int), hipMemcpyDeviceToHost));
    si.resize(sp[N]);
    if (!si.empty())
      HIP_TRY(hipMemcpy(si.data(), R->sten_idx, si.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> tptr(N + 1, 0), tidx(si.size());
    for (int j : si) tptr[j + 1]++;
    for (size_t j = 0; j < N; ++j) tptr[j + 1] += tptr[j];
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (size_t i = 0; i < N; ++i)
      for (int s = sp[i]; s < sp[i + 1]; ++s) tidx[fill[si[s]]++] = (int)i;
    if ((rc = dalloc(&R->adj_tptr, tptr.size())) || (rc = dalloc(&R->adj_tidx, tidx.size())) ||
        (rc = dput(R->adj_tptr, tptr, R->stream)) || (rc = dput(R->adj_tidx, tidx, R->stream)))
      return rc;
    HIP_TRY(hipStreamSynchronize(R->stream));
  }
  if (o2 && R->reconst == RFLU_RECONST_WENO_XYZ && !R->adj_wpre) {
    if ((rc = dalloc(&R->adj_wpre, 15 * N)) || (rc = dalloc(&R->adj_wsum, 15 * N))) return rc;
  }
  if (o2 && R->con_n > 0 && !R->adj_cptr) {
    // the entries of every cell, and for every cell j the rows whose cell is j
    const int E = R->con_n, RR = R->con_R;
    std::vector<int> ec(E), rp(E + 1), rcl(RR);
    HIP_TRY(hipStreamSynchronize(R->stream));
    HIP_TRY(hipMemcpy(ec.data(), R->con_ecell, E * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rp.data(), R->con_rptr, (E + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rcl.data(), R->con_rcell, RR * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> cptr(N + 1, 0), cent(E), tptr(N + 1, 0);
    for (int e = 0; e < E; ++e) cptr[ec[e] + 1]++;
    for (int r = 0; r < RR; ++r)
      if (rcl[r] >= 0) tptr[rcl[r] + 1]++;
    for (size_t j = 0; j < N; ++j) {
      cptr[j + 1] += cptr[j];
      tptr[j + 1] += tptr[j];
    }
    std::vector<int> trow(tptr[N]), tent(tptr[N]);
    std::vector<int> cf(cptr.begin(), cptr.end() - 1), tf(tptr.begin(), tptr.end() - 1);
    for (int e = 0; e < E; ++e) {
      cent[cf[ec[e]]++] = e;
      for (int r = rp[e]; r < rp[e + 1]; ++r)
        if (rcl[r] >= 0) {
          const int q = tf[rcl[r]]++;
          trow[q] = r;
          tent[q] = e;
        }
    }
    if ((rc = dalloc(&R->adj_cptr, cptr.size())) || (rc = dalloc(&R->adj_cent, cent.size())) ||
        (rc = dalloc(&R->adj_ctptr, tptr.size())) || (rc = dalloc(&R->adj_ctrow, trow.size())) ||
        (rc = dalloc(&R->adj_ctent, tent.size())) || (rc = dput(R->adj_cptr, cptr, R->stream)) ||
        (rc = dput(R->adj_cent, cent, R->stream)) || (rc = dput(R->adj_ctptr, tptr, R->stream)) ||
        (rc = dput(R->adj_ctrow, trow, R->stream)) || (rc = dput(R->adj_ctent, tent, R->stream)))
      return rc;
    HIP_TRY(hipStreamSynchronize(R->stream));
  }
  return 0;
}
}  // namespace

namespace {
// Device work arrays of the tangent: per-face tangent fluxes 5(F+B), the patch values'
// tangents 5P (device copy, or null), and at order 2 the primitive tangents 5N, the
// limited gradient's values and tangents 15N each and, with WENO-XYZ, the blended ones.
struct JvpWork {
  double *dflux, *dp, *ga, *gb, *gc, *gr;
  const double* dvals;
};

// drhs = dR/dcv . dcv (+ dR/dvals . dvals) with the kernels of the tangent, on the
// region's stream; returns after the stream has finished.
static int jvp_launch(Region* R, const double* dcv, const JvpWork& w, double* drhs) {
  Mesh M = R->mesh();
  State S = R->state();
  if (R->order == 2) {
    // dprim, then the limited gradient's values and tangents, then the faces
    const dim3 gc(nblk(R->N, 256)), bs(256);
    hipLaunchKernelGGL(k_prim_jvp, gc, bs, 0, R->stream, M, S, R->gas, dcv, w.dp);
    if (R->reconst == RFLU_RECONST_LIM_VENKAT)
      hipLaunchKernelGGL(k_grad2_jvp<true>, gc, bs, 0, R->stream, M, R->constr_d(), S, (const double*)w.dp, w.ga, w.gb);
    else
      hipLaunchKernelGGL(k_grad2_jvp<false>, gc, bs, 0, R->stream, M, R->constr_d(), S, (const double*)w.dp, w.ga, w.gb);
    const double *gv = w.ga, *gd = w.gb;
    if (R->reconst == RFLU_RECONST_WENO_XYZ) {
      // blend and clip the raw gradient's values and tangents
      hipLaunchKernelGGL(k_weno_jvp, gc, bs, 0, R->stream, M, S, (const double*)w.ga,
                         (const double*)w.gb, w.gc, w.gr);
      gv = w.gc;
      gd = w.gr;
    }
    if (R->F > 0) {
      dim3 g(nblk(R->F, 256));
      const double* dp = w.dp;
      if (R->discr == 1) hipLaunchKernelGGL(k_face2_jvp<1>, g, bs, 0, R->stream, M, S, R->gas, dp, gv, gd, w.dflux);
      else if (R->discr == 3) hipLaunchKernelGGL(k_face2_jvp<3>, g, bs, 0, R->stream, M, S, R->gas, dp, gv, gd, w.dflux);
      else hipLaunchKernelGGL(k_face2_jvp<4>, g, bs, 0, R->stream, M, S, R->gas, dp, gv, gd, w.dflux);
    }
    if (R->B > 0)
      hipLaunchKernelGGL(k_bface2_jvp, dim3(nblk(R->B, 256)), bs, 0, R->stream, M, S, R->gas, R->patches(),
                         (const double*)w.dp, gv, gd, w.dvals, w.dflux);
    hipLaunchKernelGGL((k_gather_rhs<double>), gc, bs, 0, R->stream, M, (const double*)w.dflux, drhs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(R->stream));
    return RFLU_OK;
  }
  if (R->F > 0) {
    dim3 g(nblk(R->F, 256)), b(256);
    if (R->discr == 1) hipLaunchKernelGGL(k_face_jvp<1>, g, b, 0, R->stream, M, S, R->gas, dcv, w.dflux);
    else if (R->discr == 3) hipLaunchKernelGGL(k_face_jvp<3>, g, b, 0, R->stream, M, S, R->gas, dcv, w.dflux);
    else hipLaunchKernelGGL(k_face_jvp<4>, g, b, 0, R->stream, M, S, R->gas, dcv, w.dflux);
  }
  if (R->B > 0)
    hipLaunchKernelGGL(k_bface_jvp, dim3(nblk(R->B, 256)), dim3(256), 0, R->stream, M, S, R->gas,
                       R->patches(), dcv, w.dvals, w.dflux);
  hipLaunchKernelGGL((k_gather_rhs<double>), dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M,
                     (const double*)w.dflux, drhs);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

// Doubles of the tangent's work arrays for region R, each rounded up to 256 bytes.
static size_t jvp_work_doubles(const Region* R) {
  auto up = [](size_t n) { return (n + 31) & ~(size_t)31; };
  const size_t N = R->N;
  size_t n = up(5 * ((size_t)R->F + R->B)) + up(5 * (size_t)R->P);
  if (R->order == 2) {
    n += up(5 * N) + 2 * up(15 * N);
    if (R->reconst == RFLU_RECONST_WENO_XYZ) n += 2 * up(15 * N);
  }
  return n;
}
}  // namespace

int rflu_gpu_residual_jvp(int32_t h, const double* dcv, const double* dvals, double* drhs) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!dcv || !drhs) return fail(RFLU_ERR_ARG, "null array");
  int rc = derivative_ready(R);
  if (rc) return rc;
  if (dvals && R->P > 0)
    HIP_TRY(hipMemcpyAsync(R->adj_dvals, dvals, 5 * (size_t)R->P * sizeof(double), hipMemcpyHostToDevice, R->stream));
  JvpWork w{R->adj_dflux, R->adj_p, R->adj_ga, R->adj_gb2, R->adj_gc, R->adj_gr,
            dvals ? (const double*)R->adj_dvals : nullptr};
  return jvp_launch(R, dcv, w, drhs);
}

int rflu_gpu_residual_jvp_static(int32_t h, const double* dcv, const double* dvals, double* drhs) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!dcv || !drhs) return fail(RFLU_ERR_ARG, "null array");
  int rc = derivative_check(R);
  if (rc) return rc;
  const size_t need = jvp_work_doubles(R);
  if (need > kScratchDoubles)
    return fail(RFLU_ERR_ARG, "the allocation-free tangent needs " + std::to_string(need * 8) +
                                  " bytes of work arrays; its static scratch holds " +
                                  std::to_string(kScratchDoubles * 8) +
                                  " (use rflu_gpu_residual_jvp)");
  double* base = nullptr;
  if ((rc = scratch_address(&base))) return rc;
  auto up = [](size_t n) { return (n + 31) & ~(size_t)31; };
  const size_t N = R->N;
  double* q = base;
  auto take = [&](size_t n) { double* r = q; q += up(n); return r; };
  JvpWork w{};
  w.dflux = take(5 * ((size_t)R->F + R->B));
  double* dv = take(5 * (size_t)R->P);
  if (R->order == 2) {
    w.dp = take(5 * N);
    w.ga = take(15 * N);
    w.gb = take(15 * N);
    if (R->reconst == RFLU_RECONST_WENO_XYZ) {
      w.gc = take(15 * N);
      w.gr = take(15 * N);
    }
  }
  if (dvals && R->P > 0) {
    HIP_TRY(hipMemcpyAsync(dv, dvals, 5 * (size_t)R->P * sizeof(double), hipMemcpyHostToDevice, R->stream));
    w.dvals = dv;
  }
  return jvp_launch(R, dcv, w, drhs);
}

int rflu_gpu_residual_vjp(int32_t h, const double* mu, double* lam, double* dvals) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!mu || !lam) return fail(RFLU_ERR_ARG, "null array");
  int rc = derivative_ready(R);
  if (rc) return rc;
  Mesh M = R->mesh();
  State S = R->state();
  if (R->order == 2) {
    const dim3 gc(nblk(R->N, 256)), bs(256);
    const bool venkat = R->reconst == RFLU_RECONST_LIM_VENKAT;
    double *graw = R->adj_gb2, *glim = R->adj_ga, *aG = R->adj_gc, *aR = R->adj_gr;
    const bool weno = R->reconst == RFLU_RECONST_WENO_XYZ;
    // forward sweep: raw and limited (or blended and clipped) gradients
    if (venkat) hipLaunchKernelGGL(k_grad2_fwd<true>, gc, bs, 0, R->stream, M, R->constr_d(), S, graw, glim);
    else hipLaunchKernelGGL(k_grad2_fwd<false>, gc, bs, 0, R->stream, M, R->constr_d(), S, graw, glim);
    if (weno)
      hipLaunchKernelGGL(k_weno_fwd, gc, bs, 0, R->stream, M, S, (const double*)graw, glim,
                         R->adj_wpre, R->adj_wsum, R->adj_seln);
    if (R->F > 0) {
      dim3 g(nblk(R->F, 256));
      if (R->discr == 1) hipLaunchKernelGGL(k_face2_vjp<1>, g, bs, 0, R->stream, M, S, R->gas, (const double*)glim, mu, R->adj_g);
      else if (R->discr == 3) hipLaunchKernelGGL(k_face2_vjp<3>, g, bs, 0, R->stream, M, S, R->gas, (const double*)glim, mu, R->adj_g);
      else hipLaunchKernelGGL(k_face2_vjp<4>, g, bs, 0, R->stream, M, S, R->gas, (const double*)glim, mu, R->adj_g);
    }
    if (R->B > 0)
      hipLaunchKernelGGL(k_bface2_vjp, dim3(nblk(R->B, 256)), bs, 0, R->stream, M, S, R->gas, R->patches(),
                         (const double*)glim, mu, R->adj_gb, R->adj_gv);
    hipLaunchKernelGGL(k_cell2_vjp, gc, bs, 0, R->stream, M, (const double*)R->adj_g, (const double*)R->adj_gb,
                       R->adj_p, aG);
    if (venkat) {
      hipLaunchKernelGGL(k_venkat_vjp, gc, bs, 0, R->stream, M, S, (const double*)graw, (const double*)aG,
                         aR, R->adj_p, R->adj_seln, R->adj_selv);
      hipLaunchKernelGGL(k_venkat_nb_gather, gc, bs, 0, R->stream, M, (const int*)R->adj_seln,
                         (const double*)R->adj_selv, R->adj_p);
    } else if (weno) {
      hipLaunchKernelGGL(k_weno_vjp, gc, bs, 0, R->stream, M, (const int*)R->adj_tptr,
                         (const int*)R->adj_tidx, (const double*)graw, (const double*)R->adj_wpre,
                         (const double*)R->adj_wsum, (const int*)R->adj_seln, (const double*)aG, aR);
    } else {
      aR = aG;
    }
    if (R->con_n > 0) {
      hipLaunchKernelGGL(k_constr_vjp, gc, bs, 0, R->stream, M, R->constr_d(), (const double*)aR, R->adj_p);
      hipLaunchKernelGGL(k_constr_clear, dim3(nblk(R->con_n, 256)), bs, 0, R->stream, M, R->constr(), aR);
    }
    hipLaunchKernelGGL(k_grad_vjp, gc, bs, 0, R->stream, M, (const int*)R->adj_tptr, (const int*)R->adj_tidx,
                       (const double*)aR, R->adj_p);
    hipLaunchKernelGGL(k_prim_vjp, gc, bs, 0, R->stream, M, S, R->gas, (const double*)R->adj_p, lam);
    if (dvals && R->P > 0) {
      hipLaunchKernelGGL(k_patch_vals_sum, dim3(R->P), dim3(256), 0, R->stream, M, (const int*)R->bptr,
                         (const double*)R->adj_gv, R->adj_pv);
      HIP_TRY(hipMemcpyAsync(dvals, R->adj_pv, 5 * (size_t)R->P * sizeof(double), hipMemcpyDeviceToHost, R->stream));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(R->stream));
    return RFLU_OK;
  }
  if (R->F > 0) {
    dim3 g(nblk(R->F, 256)), b(256);
    if (R->discr == 1) hipLaunchKernelGGL(k_face_vjp<1>, g, b, 0, R->stream, M, S, R->gas, mu, R->adj_g);
    else if (R->discr == 3) hipLaunchKernelGGL(k_face_vjp<3>, g, b, 0, R->stream, M, S, R->gas, mu, R->adj_g);
    else hipLaunchKernelGGL(k_face_vjp<4>, g, b, 0, R->stream, M, S, R->gas, mu, R->adj_g);
  }
  if (R->B > 0)
    hipLaunchKernelGGL(k_bface_vjp, dim3(nblk(R->B, 256)), dim3(256), 0, R->stream, M, S, R->gas,
                       R->patches(), mu, R->adj_gb, R->adj_gv);
  hipLaunchKernelGGL(k_cell_vjp, dim3(nblk(R->N, 256)), dim3(256), 0, R->stream, M,
                     (const double*)R->adj_g, (const double*)R->adj_gb, lam);
  if (dvals && R->P > 0) {
    hipLaunchKernelGGL(k_patch_vals_sum, dim3(R->P), dim3(256), 0, R->stream, M, (const int*)R->bptr,
                       (const double*)R->adj_gv, R->adj_pv);
    HIP_TRY(hipMemcpyAsync(dvals, R->adj_pv, 5 * (size_t)R->P * sizeof(double), hipMemcpyDeviceToHost, R->stream));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(R->stream));
  return RFLU_OK;
}

int rflu_gpu_set_profiling(int32_t h, int32_t on) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  R->profiling = on != 0;
  for (int k = 0; k < KID_N; ++k) { R->kern_ms[k] = 0; R->kern_n[k] = 0; }
  return RFLU_OK;
}

int rflu_gpu_get_kernel_times(int32_t h, double* ms, int64_t* counts, int32_t n) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!ms || !counts || n < 1) return fail(RFLU_ERR_ARG, "bad arguments");
  for (int k = 0; k < n && k < KID_N; ++k) { ms[k] = R->kern_ms[k]; counts[k] = R->kern_n[k]; }
  return RFLU_OK;
}

int rflu_dump_name(char* buf, int32_t buflen, const char* out_dir, int64_t step, int32_t region) {
  if (!buf || buflen <= 0) return fail(RFLU_ERR_ARG, "null or empty name buffer");
  if (step < 0 || step > 99999999 || region < 0 || region > 99999)
    return fail(RFLU_ERR_ARG, "step or region does not fit the dump name's fields");
  std::string s = rflu_dump::name(out_dir ? out_dir : "", step, region);
  if ((int64_t)s.size() + 1 > buflen) return fail(RFLU_ERR_ARG, "name buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return RFLU_OK;
}

namespace {

int dump_status(rflu_dump::Status s, const std::string& msg) {
  if (s == rflu_dump::OK) return RFLU_OK;
  return fail(s == rflu_dump::FILE_ERR ? RFLU_ERR_FILE : RFLU_ERR_DUMP, msg);
}

}  // namespace

int rflu_dump_write(const char* path, int64_t step, double time, int32_t region,
                    int32_t n_cells, int32_t n_cv, const double* cv) {
  if (!path || (!cv && n_cells > 0) || n_cells < 0 || n_cv < 1)
    return fail(RFLU_ERR_ARG, "bad dump arguments");
  std::vector<std::vector<double>> var(n_cv, std::vector<double>((size_t)n_cells));
  std::vector<const double*> ptrs(n_cv);
  for (int iv = 0; iv < n_cv; ++iv) {
    for (int c = 0; c < n_cells; ++c) var[iv][c] = cv[(size_t)n_cv * c + iv];
    ptrs[iv] = var[iv].data();
  }
  rflu_dump::Header h;
  h.step = step; h.time = time; h.region = region; h.n_cells = n_cells; h.n_cv = n_cv;
  std::string msg;
  return dump_status(rflu_dump::write(path, h, ptrs.data(), msg), msg);
}

int rflu_dump_read(const char* path, int32_t n_cells, int32_t n_cv, int64_t* step,
                   double* time, int32_t* region, double* cv) {
  if (!path || !cv || n_cells < 0 || n_cv < 1) return fail(RFLU_ERR_ARG, "bad dump arguments");
  std::vector<std::vector<double>> var(n_cv, std::vector<double>((size_t)n_cells));
  std::vector<double*> ptrs(n_cv);
  for (int iv = 0; iv < n_cv; ++iv) ptrs[iv] = var[iv].data();
  rflu_dump::Header h;
  std::string msg;
  int rc = dump_status(rflu_dump::read(path, n_cells, n_cv, h, ptrs.data(), msg), msg);
  if (rc) return rc;
  for (int c = 0; c < n_cells; ++c)
    for (int iv = 0; iv < n_cv; ++iv) cv[(size_t)n_cv * c + iv] = var[iv][c];
  if (step) *step = h.step;
  if (time) *time = h.time;
  if (region) *region = h.region;
  return RFLU_OK;
}

int rflu_gpu_write_dump(int32_t h, int32_t icase, const char* path, int64_t step, double time,
                        int32_t region) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!path || icase < 0 || icase >= R->K) return fail(RFLU_ERR_ARG, "bad case or null path");
  // the actual cells, as RFLU_WriteRestartData writes cv(:,1:nCells) (virtual cells
  // come first-to-last after them in both numberings)
  const int N = R->N, c0 = R->cell0_h[icase];
  const int n = R->K == 1 ? R->nact : R->cell0_h[icase + 1] - c0;
  // device [5][N] in the region's cell order -> the case's cells in the caller's order
  std::vector<double> dev((size_t)5 * n);
  for (int iv = 0; iv < 5; ++iv)
    HIP_TRY(hipMemcpyAsync(dev.data() + (size_t)iv * n, R->cv + (size_t)iv * N + c0,
                           n * sizeof(double), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  std::vector<double> var((size_t)5 * n);
  for (int c = 0; c < n; ++c) {
    int o = R->new2old[c0 + c] - c0;
    for (int iv = 0; iv < 5; ++iv) var[(size_t)iv * n + o] = dev[(size_t)iv * n + c];
  }
  const double* ptrs[5];
  for (int iv = 0; iv < 5; ++iv) ptrs[iv] = var.data() + (size_t)iv * n;
  rflu_dump::Header hd;
  hd.step = step; hd.time = time; hd.region = region; hd.n_cells = n; hd.n_cv = 5;
  std::string msg;
  return dump_status(rflu_dump::write(path, hd, ptrs, msg), msg);
}

int rflu_gpu_read_dump(int32_t h, int32_t icase, const char* path, int64_t* step, double* time,
                       int32_t* region) {
  Region* R = get(h);
  if (!R) return fail(RFLU_ERR_HANDLE, "bad region handle");
  if (int frc = flush_pending(R)) return frc;
  if (!path || icase < 0 || icase >= R->K) return fail(RFLU_ERR_ARG, "bad case or null path");
  // the actual cells (RFLU_ReadRestartData reads cv(:,1:nCells)); the next exchange
  // sets the virtual ones
  const int N = R->N, c0 = R->cell0_h[icase];
  const int n = R->K == 1 ? R->nact : R->cell0_h[icase + 1] - c0;
  std::vector<double> var((size_t)5 * n);
  double* ptrs[5];
  for (int iv = 0; iv < 5; ++iv) ptrs[iv] = var.data() + (size_t)iv * n;
  rflu_dump::Header hd;
  std::string msg;
  int rc = dump_status(rflu_dump::read(path, n, 5, hd, ptrs, msg), msg);
  if (rc) return rc;
  std::vector<double> dev((size_t)5 * n);
  for (int c = 0; c < n; ++c) {
    int o = R->new2old[c0 + c] - c0;
    for (int iv = 0; iv < 5; ++iv) dev[(size_t)iv * n + c] = var[(size_t)iv * n + o];
  }
  HIP_TRY(hipStreamSynchronize(R->stream));
  for (int iv = 0; iv < 5; ++iv)
    HIP_TRY(hipMemcpyAsync(R->cv + (size_t)iv * N + c0, dev.data() + (size_t)iv * n,
                           n * sizeof(double), hipMemcpyHostToDevice, R->stream));
  // dv = f(cv) holds in every cell after an update, so recomputing it over the whole
  // region leaves the other cases unchanged
  Mesh M = R->mesh();
  State S = R->state();
  hipLaunchKernelGGL(k_set_dv, dim3(nblk(N, 256)), dim3(256), 0, R->stream, M, S, R->gas, R->cases());
  if (R->order == 2) hipLaunchKernelGGL(k_set_prim, dim3(nblk(N, 256)), dim3(256), 0, R->stream, M, S);
  hipLaunchKernelGGL(k_fill_scalars, dim3(1), dim3(64), 0, R->stream, R->sc, icase, icase + 1,
                     hd.time, R->dtmin);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(R->stream));
  if (step) *step = hd.step;
  if (time) *time = hd.time;
  if (region) *region = hd.region;
  return RFLU_OK;
}

}  // extern "C"
