// Device kernels of the explicit residual/update path for gfx950.
//
// Layout in HBM (all structure-of-arrays so a wavefront's 64 lanes read 64
// consecutive cells or faces): cv[5][N], dv[3][N], prim[5][N] = (rho,u,v,w,p),
// grad[15][N] with component index 3*var+axis, flux[5][F+B].  Face geometry is
// packed as double4 (nx,ny,nz,|A|) and double4 (xc,yc,zc,-); cells as double4
// (x,y,z,vol).  Interior faces come first in the flux buffer, then the boundary
// faces patch by patch.  The cell-to-face list c2f stores +(f+1) where the cell is
// f2c(1,f) or the face is a boundary face, and -(f+1) where it is f2c(2,f), in the
// reference's scatter order, so the per-cell gather sums in that order.
//
// Arithmetic follows the reference's expressions operand by operand; the library
// is compiled with FMA contraction off so FP64 results match the reference's
// gfortran build rounding for rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rflu {

constexpr int BC_INFLOW_TOTANG = 10;
constexpr int BC_INFLOW_VELTEMP = 11;
constexpr int BC_OUTFLOW = 20;
constexpr int BC_SLIPWALL = 60;
constexpr int BC_FARFIELD = 80;
constexpr int BC_INJECTION = 90;
constexpr int BCOPT_SUPERSONIC = 0;
constexpr int BCOPT_SUBSONIC = 1;
constexpr int BCOPT_MIXED = 2;

// Boundary patches of a region (all cases), indexed by bpatch[b].
struct Patches {
  const int* bc;
  const int* opt;       // BCSWI_INFLOW_TYPE / BCSWI_OUTFLOW_TYPE
  const int* opt2;      // BCSWI_INFLOW_FIXED of total-condition inflow patches
  const int* order2;
  const double* pout;
  const double* vals;   // (5, P) the patch's BCDAT values 1..5
};

// Inflow boundary state with prescribed velocity and temperature (BC_INFLOW_VELTEMP,
// rocflu/RFLU_CentralFirstPatch.F90:375-414, RFLU_CentralSecondPatch.F90:473-495):
// subsonic takes the interior density, otherwise the prescribed pressure fixes it
// (MixtPerf_D_PRT, MixtPerf_P_DRT, MixtPerf_Eo_DGPUVW).  b = (rho, rho u, rho v,
// rho w, rho E) of the boundary state; returns its pressure.
template <typename T, typename V>
__device__ __forceinline__ T inflow_veltemp(int bcOpt, const V* vals, T rl, double rgas, double g,
                                            T* b) {
  const T ur = T(vals[0]), vr = T(vals[1]), wr = T(vals[2]), tr = T(vals[3]);
  T pr, rr;
  if (bcOpt != 1) {
    pr = T(vals[4]);
    rr = pr / (T(rgas) * tr);
  } else {
    rr = rl;
    pr = rr * T(rgas) * tr;
  }
  b[0] = rr;
  b[1] = rr * ur;
  b[2] = rr * vr;
  b[3] = rr * wr;
  b[4] = rr * (pr / (rr * (T(g) - T(1.0))) + T(0.5) * (ur * ur + vr * vr + wr * wr));
  return pr;
}

struct Gas {
  double g, cp, gc;            // refGamma, refCp, R = cp - cp/g (flux functions)
  double gp, rgasp, mm;        // gamma', R' and M of the dependent-variable path
  double eps_entr, diss_fact;
};

// Device-resident scalars of a region.
struct Scalars {
  double time, dt;
  double dt_local_min;
  double mass_in, mass_out;
  unsigned long long n_invalid, n_negative, n_bounds_fix, n_pos_fix;
  unsigned long long steps;
  int abort;      // a check failed in an earlier launch: every launch returns at once
  int last_step;  // the end-time clamp shortened this step (RFLU_TimeStepping.F90:445-448)
  int finished;   // set by the next step's time-step kernel after the last step
  int fail;       // a check failed (set by the launch that checks; promoted to abort by
                  // the next launch that is not an update, so that the launch that finds
                  // the failure still updates and checks every cell)
};

// Whether the update kernels skip the case: a failure of an earlier launch, or done.
// They must not read `fail`, which their own launch may be setting.
__device__ __forceinline__ bool halted(const Scalars* sc) { return sc->abort || sc->finished; }
// Whether any other kernel skips it.
__device__ __forceinline__ bool stopped(const Scalars* sc) {
  return sc->abort || sc->fail || sc->finished;
}
// fail -> abort, by one thread of a launch that runs between two updates.
__device__ __forceinline__ void promote(Scalars* sc) {
  if (sc->fail && !sc->abort) sc->abort = 1;
}
// Early exit of a whole launch: the single-case scalars, or none (nullptr) for a batch
// of cases, whose kernels run every case and leave stopped cases to the update.
// Promotes a failure found by the previous update.
__device__ __forceinline__ bool gate(const Scalars* sc) {
  if (!sc) return false;
  if (sc->fail && !sc->abort && blockIdx.x == 0 && threadIdx.x == 0)
    promote(const_cast<Scalars*>(sc));
  return stopped(sc);
}

// Per-case time-step controls (RFLU_MinimumTimeStep inputs and global%maxTime).
struct CaseParams {
  double cfl, dt_imposed, dt_min_limit, end_time;
};

// Slots of the per-case minimum time step are this many 8-byte words apart.
constexpr int DTMIN_STRIDE = 32;

// Cases of a batched region: K independent cases whose cells, faces and boundary
// faces occupy consecutive ranges.  With cid == nullptr the region is one case and
// the kernels use the by-value Gas and the scalars they are given.
struct Cases {
  int K;
  const unsigned short* cid;   // [N] case of each cell, or nullptr
  const Gas* gas;              // [K]
  Scalars* sc;                 // [K]
  const int* cell0;            // [K+1] first cell of each case
  const int* bface0;           // [K+1] first boundary face of each case
  const CaseParams* par;       // [K]
  unsigned long long* dtmin;   // [K * DTMIN_STRIDE] minimum local dt of the step (bits of a
                               // positive double) at c * DTMIN_STRIDE
};

struct Mesh {
  int N, F, B;
  int nact;                // actual cells [0, nact); virtual cells [nact, N) follow
  const int2* f2c;
  const double4* fn;
  const double4* fc;
  const double4* cell;     // cofg + vol
  const int* bf2c;
  const double4* bfn;
  const double4* bfc;
  const int* bpatch;
  const int* c2f_ptr;
  const int* c2f_idx;
  const int* sten_ptr;
  const int* sten_idx;
  const double* moms;      // [6][N]: r11 r12 r22 r13 r23 r33
  const double* epsq;      // [N] Venkatakrishnan eps^2 = (vol^(1/3))^3
  // Limiter neighbourhood: for the q-th interior face of cell i (cell-to-face order),
  // vnb[q*N+i] is the cell across it (-1 past the last one) and vgeo[(3q+d)*N+i] the
  // component d of cofg(i) - fc(face).
  int vq;                  // slots per cell (largest interior-face count)
  const int* vnb;
  const double* vgeo;
  int nxcd;                // blocks are spread over this many XCDs (1: no remap)
};

// Logical block index for an XCD-aware launch.  The dispatcher hands consecutive
// workgroups to the XCDs in turn; remapping gives each XCD one contiguous range of
// cells or faces, so neighbouring blocks share that XCD's L2.  A bijection on
// [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block(int nxcd) {
  const int b = (int)blockIdx.x, n = (int)gridDim.x;
  if (nxcd <= 1) return b;
  const int x = b % nxcd, l = b / nxcd, q = n / nxcd, r = n % nxcd;
  return x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

struct State {
  double* cv;      // [5][N]
  double* cv_old;  // [5][N]  step-start state (RK4, SSP-RK3)
  double* rhs_sum; // [5][N]
  double* dv;      // [3][N]
  double* prim;    // [5][N]
};

// Face-flux buffer [5][F+B] (interior faces, then boundary faces).  FT = double or
// float holds the five components at that width; FT = FluxMix (the mixed mode) holds
// mass and energy in FP32 and the three momentum components in FP64: F+B floats of
// mass, F+B floats of energy, then [3][F+B] doubles -- 32 bytes per face.  val_t is
// the register/LDS type of one value; the residual sums in FP64 whatever it is.
struct FluxMix {};
template <typename FT>
struct FluxIO {
  using val_t = FT;
  static __device__ __forceinline__ void put(FT* p, size_t FB, int k, size_t f, double v) {
    p[k * FB + f] = FT(v);
  }
  static __device__ __forceinline__ FT get(const FT* p, size_t FB, int k, size_t f) {
    return p[k * FB + f];
  }
};
template <>
struct FluxIO<FluxMix> {
  using val_t = double;
  static __device__ __forceinline__ void put(FluxMix* p, size_t FB, int k, size_t f, double v) {
    if (k == 0) ((float*)p)[f] = float(v);
    else if (k == 4) ((float*)p)[FB + f] = float(v);
    else ((double*)p)[(size_t)k * FB + f] = v;   // doubles 1..3 of [4][F+B] (floats fill 0)
  }
  static __device__ __forceinline__ double get(const FluxMix* p, size_t FB, int k, size_t f) {
    if (k == 0) return double(((const float*)p)[f]);
    if (k == 4) return double(((const float*)p)[FB + f]);
    return ((const double*)p)[(size_t)k * FB + f];
  }
};

template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ float tabs(float x) { return fabsf(x); }

// 1 / sqrt(x) with the two roundings of the reference (the correctly rounded square
// root, then the correctly rounded quotient).  For x in [2^-700, 2^700] it runs the
// compiler's own Newton-Raphson sequence for sqrt (v_rsq_f64, then FMA corrections)
// without its range scaling and special-value selects, which leave such arguments
// unchanged, and the last two FMA steps of its sequence for 1.0/y, seeded with the
// square root's own 1/(2 sqrt(x)) in place of v_rcp_f64 and a first step.  That start
// gives the correctly rounded quotient except where the square root's significand is
// all ones; those arguments, and those outside the range, take the library path.  The
// bits are those of 1.0 / sqrt(x) (checked on the GPU over 5.4e9 arguments, hard cases
// of both roundings included).  The stencil weights use it once per member.
// Whether inv_sqrt_fast gives the bits of 1.0 / sqrt(x).
__device__ __forceinline__ bool inv_sqrt_fast_ok(double x) { return x >= 0x1p-700 && x <= 0x1p700; }

// The sequences of inv_sqrt_cr without its range check (arguments in [2^-700, 2^700]).
// exact is false where the square root's significand is all ones: there the seeded
// quotient below can miss the correctly rounded one, and the caller takes the library.
__device__ __forceinline__ double inv_sqrt_fast(double x, bool& exact) {
  const double y = __builtin_amdgcn_rsq(x);
  const double s0 = x * y;
  const double h0 = y * 0.5;
  const double r0 = __builtin_fma(-h0, s0, 0.5);
  const double s1 = __builtin_fma(s0, r0, s0);
  const double d0 = __builtin_fma(-s1, s1, x);
  const double h1 = __builtin_fma(h0, r0, h0);
  const double s2 = __builtin_fma(d0, h1, s1);
  const double d1 = __builtin_fma(-s2, s2, x);
  const double q = __builtin_fma(d1, h1, s2);     // sqrt(x)
  exact = ((unsigned long long)__double_as_longlong(q) & 0xFFFFFFFFFFFFFull) != 0xFFFFFFFFFFFFFull;
  // 2 h1 is 1/sqrt(x) to about 2^-51: it starts the quotient's iteration where the
  // sequence for 1.0/q arrives after v_rcp_f64 and its first step
  double r = h1 + h1;
  const double e = __builtin_fma(-q, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double f = __builtin_fma(-q, r, 1.0);
  return __builtin_fma(f, r, r);                  // 1.0 / q
}

__device__ __forceinline__ double inv_sqrt_cr(double x) {
  if (!inv_sqrt_fast_ok(x)) return 1.0 / sqrt(x);
  bool exact;
  const double r = inv_sqrt_fast(x, exact);
  return exact ? r : 1.0 / sqrt(x);
}
// FP32 gradients (the mixed mode): the hardware reciprocal square root (within 1 ulp;
// the gradients are FP32 approximations of the FP64 ones either way).
__device__ __forceinline__ float inv_sqrt_cr(float x) { return __builtin_amdgcn_rsqf(x); }

// inv_sqrt_cr of n arguments at once: the fast sequences of all of them first as
// straight-line code (independent chains the scheduler can interleave; a branch per
// argument would serialize them), then the library path for any out of range.
template <int n>
__device__ __forceinline__ void inv_sqrt_cr_n(const double* x, double* r) {
  bool ok = true, ex[n];
#pragma unroll
  for (int k = 0; k < n; ++k) {
    r[k] = inv_sqrt_fast(x[k], ex[k]);
    ex[k] = ex[k] && inv_sqrt_fast_ok(x[k]);
    ok = ok && ex[k];
  }
  if (!ok) {
#pragma unroll
    for (int k = 0; k < n; ++k)
      if (!ex[k]) r[k] = 1.0 / sqrt(x[k]);
  }
}
template <int n>
__device__ __forceinline__ void inv_sqrt_cr_n(const float* x, float* r) {
#pragma unroll
  for (int k = 0; k < n; ++k) r[k] = inv_sqrt_cr(x[k]);
}

// ---------------------------------------------------------------------------------
// Flux functions (interior faces).  See the oracle for the line-by-line citations:
// HLLC modflu/RFLU_ModHLLCFlux.F90:1064-1117, Roe modflu/RFLU_ModRoeFlux.F90:
// 1838-1894 with RFLU_ModEntropyFixes.F90:95-106, AUSM+ RFLU_ModAUSMFlux.F90:282-367.
// ---------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void hllc(T g, T nx, T ny, T nz, T nm, T rl, T ul, T vl, T wl, T pl,
                                     T al, T el, T Hl, T rr, T ur, T vr, T wr, T pr, T ar, T er,
                                     T Hr, T* flx) {
  const T one = T(1), half = T(0.5), zero = T(0);
  T ql = ul * nx + vl * ny + wl * nz;
  T qr = ur * nx + vr * ny + wr * nz;
  T wt1 = tsqrt(rr / rl);
  T wt2 = one / (one + wt1);
  T us = wt2 * (ul + wt1 * ur);
  T vs = wt2 * (vl + wt1 * vr);
  T ws = wt2 * (wl + wt1 * wr);
  T qs = wt2 * (ql + wt1 * qr);
  T Hs = wt2 * (Hl + wt1 * Hr);
  T Vms = us * us + vs * vs + ws * ws;
  T as = tsqrt((g - one) * (Hs - half * Vms));
  T sl = tmin(ql - al, qs - as);
  T sr = tmax(qr + ar, qs + as);
  if (sl > zero) {
    flx[0] = ql * rl * nm;
    flx[1] = (ql * rl * ul + pl * nx) * nm;
    flx[2] = (ql * rl * vl + pl * ny) * nm;
    flx[3] = (ql * rl * wl + pl * nz) * nm;
    flx[4] = (ql * (rl * el + pl)) * nm;
  } else if (sr < zero) {
    flx[0] = qr * rr * nm;
    flx[1] = (qr * rr * ur + pr * nx) * nm;
    flx[2] = (qr * rr * vr + pr * ny) * nm;
    flx[3] = (qr * rr * wr + pr * nz) * nm;
    flx[4] = (qr * (rr * er + pr)) * nm;
  } else {
    T sm = (rr * qr * (sr - qr) - rl * ql * (sl - ql) - pr + pl) /
           (rr * (sr - qr) - rl * (sl - ql));
    T ps = pl + rl * (ql - sl) * (ql - sm);
    T term, rs, rus, rvs, rws, res;
    if (sm > zero) {
      term = one / (sl - sm);
      rs = term * (sl - ql) * rl;
      rus = term * ((sl - ql) * rl * ul + (ps - pl) * nx);
      rvs = term * ((sl - ql) * rl * vl + (ps - pl) * ny);
      rws = term * ((sl - ql) * rl * wl + (ps - pl) * nz);
      res = term * ((sl - ql) * rl * el + (ps * sm - pl * ql));
    } else {
      term = one / (sr - sm);
      rs = term * (sr - qr) * rr;
      rus = term * ((sr - qr) * rr * ur + (ps - pr) * nx);
      rvs = term * ((sr - qr) * rr * vr + (ps - pr) * ny);
      rws = term * ((sr - qr) * rr * wr + (ps - pr) * nz);
      res = term * ((sr - qr) * rr * er + (ps * sm - pr * qr));
    }
    flx[0] = sm * rs * nm;
    flx[1] = (sm * rus + ps * nx) * nm;
    flx[2] = (sm * rvs + ps * ny) * nm;
    flx[3] = (sm * rws + ps * nz) * nm;
    flx[4] = (sm * (res + ps)) * nm;
  }
}

template <typename T> __device__ __forceinline__ T ef_hh(T l, T d) {
  return l > d ? l : (l * l + d * d) / (T(2) * d);
}

template <typename T>
__device__ __forceinline__ void roe(T g, T epsentr, T dissFact, T nx, T ny, T nz, T nm, T rl,
                                    T ul, T vl, T wl, T pl, T Hl, T rr, T ur, T vr, T wr, T pr,
                                    T Hr, T* flx) {
  const T one = T(1), half = T(0.5), two = T(2);
  T ql = ul * nx + vl * ny + wl * nz;
  T qr = ur * nx + vr * ny + wr * nz;
  T rh = tsqrt(rl * rr);
  T wt = rl / (rl + rh);
  T uh = wt * ul + (one - wt) * ur;
  T vh = wt * vl + (one - wt) * vr;
  T wh = wt * wl + (one - wt) * wr;
  T Hh = wt * Hl + (one - wt) * Hr;
  T qh = uh * nx + vh * ny + wh * nz;
  T sh = half * (uh * uh + vh * vh + wh * wh);
  T ah = tsqrt((g - one) * (Hh - half * (two * sh)));
  T l1 = tabs(qh - ah);
  T l2 = tabs(qh);
  T l5 = tabs(qh + ah);
  T efc = epsentr * l5;
  l1 = ef_hh(l1, efc);
  l2 = ef_hh(l2, efc);
  l5 = ef_hh(l5, efc);
  T dr = rr - rl, du = ur - ul, de = vr - vl, dw = wr - wl, dp = pr - pl;
  T dq = du * nx + de * ny + dw * nz;
  T dv1 = (dp - rh * ah * dq) / (two * ah * ah);
  T dv5 = (dp + rh * ah * dq) / (two * ah * ah);
  T dv2 = dr - dp / (ah * ah);
  T t1 = l1 * dv1, t2 = l2 * dv2, t3 = l2 * rh, t5 = l5 * dv5;
  T hd = half * dissFact;
  T fc1 = half * (ql * rl + qr * rr) * nm;
  T fc2 = half * (ql * rl * ul + pl * nx + qr * rr * ur + pr * nx) * nm;
  T fc3 = half * (ql * rl * vl + pl * ny + qr * rr * vr + pr * ny) * nm;
  T fc4 = half * (ql * rl * wl + pl * nz + qr * rr * wr + pr * nz) * nm;
  T fc5 = half * (ql * rl * Hl + qr * rr * Hr) * nm;
  T fd1 = hd * (t1 + t2 + t5) * nm;
  T fd2 = hd * (t1 * (uh - nx * ah) + t2 * uh + t3 * (du - nx * dq) + t5 * (uh + nx * ah)) * nm;
  T fd3 = hd * (t1 * (vh - ny * ah) + t2 * vh + t3 * (de - ny * dq) + t5 * (vh + ny * ah)) * nm;
  T fd4 = hd * (t1 * (wh - nz * ah) + t2 * wh + t3 * (dw - nz * dq) + t5 * (wh + nz * ah)) * nm;
  T fd5 = hd * (t1 * (Hh - qh * ah) + t2 * sh + t3 * (uh * du + vh * de + wh * dw - qh * dq) +
                t5 * (Hh + qh * ah)) * nm;
  flx[0] = fc1 - fd1;
  flx[1] = fc2 - fd2;
  flx[2] = fc3 - fd3;
  flx[3] = fc4 - fd4;
  flx[4] = fc5 - fd5;
}

template <typename T>
__device__ __forceinline__ void ausm(T nx, T ny, T nz, T nm, T rl, T ul, T vl, T wl, T pl, T Hl,
                                     T al, T rr, T ur, T vr, T wr, T pr, T Hr, T ar, T* flx) {
  const T one = T(1), half = T(0.5), q4 = T(0.25), e8 = T(0.125), c3 = T(0.1875), two = T(2);
  T ql = ul * nx + vl * ny + wl * nz;
  T qr = ur * nx + vr * ny + wr * nz;
  T af = half * (al + ar);
  T ml = ql / af, mla = tabs(ml);
  T mr = qr / af, mra = tabs(mr);
  T mlp, wtl, mrm, wtr;
  if (mla <= one) {
    mlp = q4 * (ml + one) * (ml + one) + e8 * (ml * ml - one) * (ml * ml - one);
    wtl = q4 * (ml + one) * (ml + one) * (two - ml) + c3 * ml * (ml * ml - one) * (ml * ml - one);
  } else {
    mlp = half * (ml + mla);
    wtl = half * (one + ml / mla);
  }
  if (mra <= one) {
    mrm = -q4 * (mr - one) * (mr - one) - e8 * (mr * mr - one) * (mr * mr - one);
    wtr = q4 * (mr - one) * (mr - one) * (two + mr) - c3 * mr * (mr * mr - one) * (mr * mr - one);
  } else {
    mrm = half * (mr - mra);
    wtr = half * (one - mr / mra);
  }
  T mf = mlp + mrm;
  T mfa = tabs(mf);
  T mfp = half * (mf + mfa);
  T mfm = half * (mf - mfa);
  T pf = wtl * pl + wtr * pr;
  flx[0] = (af * (mfp * rl + mfm * rr)) * nm;
  flx[1] = (af * (mfp * rl * ul + mfm * rr * ur) + pf * nx) * nm;
  flx[2] = (af * (mfp * rl * vl + mfm * rr * vr) + pf * ny) * nm;
  flx[3] = (af * (mfp * rl * wl + mfm * rr * wr) + pf * nz) * nm;
  flx[4] = (af * (mfp * rl * Hl + mfm * rr * Hr)) * nm;
}

// ---------------------------------------------------------------------------------
// Mixed precision: the same three fluxes in FP32 with their pressure terms in FP64.
// A face's momentum flux carries p*n*|A| with p of order one, while the momenta
// normal to the main flow stay orders of magnitude smaller; rounding p*n*|A| to FP32
// (6e-8 of p|A| per face) would set their error.  So the face pressures PL, PR arrive
// in FP64 (the cell's FP64 pressure plus its FP32 reconstruction increment) and the
// pressure parts of the momentum fluxes -- p*n*|A| of the upwind and central forms,
// HLLC's star pressure, AUSM+'s split pressure with its weights -- are formed in FP64;
// the mass flux, the convective and dissipative parts and the energy flux stay FP32.
// Pressure differences enter the FP32 parts as FP64 differences rounded once.  The
// momentum components leave in FP64.  Mathematically these are hllc/roe/ausm above.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void hllc_mp(float g, double nxd, double nyd, double nzd, double nmd,
                                        float rl, float ul, float vl, float wl, double PL, float al,
                                        float el, float Hl, float rr, float ur, float vr, float wr,
                                        double PR, float ar, float er, float Hr, double* flx) {
  const float one = 1.f, half = 0.5f, zero = 0.f;
  const float nx = float(nxd), ny = float(nyd), nz = float(nzd), nm = float(nmd);
  const float pl = float(PL), pr = float(PR);
  float ql = ul * nx + vl * ny + wl * nz;
  float qr = ur * nx + vr * ny + wr * nz;
  float wt1 = sqrtf(rr / rl);
  float wt2 = one / (one + wt1);
  float us = wt2 * (ul + wt1 * ur);
  float vs = wt2 * (vl + wt1 * vr);
  float ws = wt2 * (wl + wt1 * wr);
  float qs = wt2 * (ql + wt1 * qr);
  float Hs = wt2 * (Hl + wt1 * Hr);
  float Vms = us * us + vs * vs + ws * ws;
  float as = sqrtf((g - one) * (Hs - half * Vms));
  float sl = tmin(ql - al, qs - as);
  float sr = tmax(qr + ar, qs + as);
  if (sl > zero) {
    flx[0] = double(ql * rl * nm);
    flx[1] = (double(ql * rl * ul) + PL * nxd) * nmd;
    flx[2] = (double(ql * rl * vl) + PL * nyd) * nmd;
    flx[3] = (double(ql * rl * wl) + PL * nzd) * nmd;
    flx[4] = double((ql * (rl * el + pl)) * nm);
  } else if (sr < zero) {
    flx[0] = double(qr * rr * nm);
    flx[1] = (double(qr * rr * ur) + PR * nxd) * nmd;
    flx[2] = (double(qr * rr * vr) + PR * nyd) * nmd;
    flx[3] = (double(qr * rr * wr) + PR * nzd) * nmd;
    flx[4] = double((qr * (rr * er + pr)) * nm);
  } else {
    float sm = (rr * qr * (sr - qr) - rl * ql * (sl - ql) + float(PL - PR)) /
               (rr * (sr - qr) - rl * (sl - ql));
    const float dl = rl * (ql - sl) * (ql - sm);   // ps - pl
    const double PS = PL + double(dl);
    const float ps = float(PS);
    float term, rs, rus, rvs, rws, res;
    if (sm > zero) {
      term = one / (sl - sm);
      rs = term * (sl - ql) * rl;
      rus = term * ((sl - ql) * rl * ul + dl * nx);
      rvs = term * ((sl - ql) * rl * vl + dl * ny);
      rws = term * ((sl - ql) * rl * wl + dl * nz);
      res = term * ((sl - ql) * rl * el + (ps * sm - pl * ql));
    } else {
      const float dr = float(PS - PR);             // ps - pr
      term = one / (sr - sm);
      rs = term * (sr - qr) * rr;
      rus = term * ((sr - qr) * rr * ur + dr * nx);
      rvs = term * ((sr - qr) * rr * vr + dr * ny);
      rws = term * ((sr - qr) * rr * wr + dr * nz);
      res = term * ((sr - qr) * rr * er + (ps * sm - pr * qr));
    }
    flx[0] = double(sm * rs * nm);
    flx[1] = (double(sm * rus) + PS * nxd) * nmd;
    flx[2] = (double(sm * rvs) + PS * nyd) * nmd;
    flx[3] = (double(sm * rws) + PS * nzd) * nmd;
    flx[4] = double((sm * (res + ps)) * nm);
  }
}

__device__ __forceinline__ void roe_mp(float g, float epsentr, float dissFact, double nxd,
                                       double nyd, double nzd, double nmd, float rl, float ul,
                                       float vl, float wl, double PL, float Hl, float rr, float ur,
                                       float vr, float wr, double PR, float Hr, double* flx) {
  const float one = 1.f, half = 0.5f, two = 2.f;
  const float nx = float(nxd), ny = float(nyd), nz = float(nzd), nm = float(nmd);
  float ql = ul * nx + vl * ny + wl * nz;
  float qr = ur * nx + vr * ny + wr * nz;
  float rh = sqrtf(rl * rr);
  float wt = rl / (rl + rh);
  float uh = wt * ul + (one - wt) * ur;
  float vh = wt * vl + (one - wt) * vr;
  float wh = wt * wl + (one - wt) * wr;
  float Hh = wt * Hl + (one - wt) * Hr;
  float qh = uh * nx + vh * ny + wh * nz;
  float sh = half * (uh * uh + vh * vh + wh * wh);
  float ah = sqrtf((g - one) * (Hh - half * (two * sh)));
  float l1 = fabsf(qh - ah);
  float l2 = fabsf(qh);
  float l5 = fabsf(qh + ah);
  float efc = epsentr * l5;
  l1 = ef_hh(l1, efc);
  l2 = ef_hh(l2, efc);
  l5 = ef_hh(l5, efc);
  float dr = rr - rl, du = ur - ul, de = vr - vl, dw = wr - wl, dp = float(PR - PL);
  float dq = du * nx + de * ny + dw * nz;
  float dv1 = (dp - rh * ah * dq) / (two * ah * ah);
  float dv5 = (dp + rh * ah * dq) / (two * ah * ah);
  float dv2 = dr - dp / (ah * ah);
  float t1 = l1 * dv1, t2 = l2 * dv2, t3 = l2 * rh, t5 = l5 * dv5;
  float hd = half * dissFact;
  const double PC = 0.5 * (PL + PR);
  float fc1 = half * (ql * rl + qr * rr) * nm;
  double fc2 = (double(half * (ql * rl * ul + qr * rr * ur)) + PC * nxd) * nmd;
  double fc3 = (double(half * (ql * rl * vl + qr * rr * vr)) + PC * nyd) * nmd;
  double fc4 = (double(half * (ql * rl * wl + qr * rr * wr)) + PC * nzd) * nmd;
  float fc5 = half * (ql * rl * Hl + qr * rr * Hr) * nm;
  float fd1 = hd * (t1 + t2 + t5) * nm;
  float fd2 = hd * (t1 * (uh - nx * ah) + t2 * uh + t3 * (du - nx * dq) + t5 * (uh + nx * ah)) * nm;
  float fd3 = hd * (t1 * (vh - ny * ah) + t2 * vh + t3 * (de - ny * dq) + t5 * (vh + ny * ah)) * nm;
  float fd4 = hd * (t1 * (wh - nz * ah) + t2 * wh + t3 * (dw - nz * dq) + t5 * (wh + nz * ah)) * nm;
  float fd5 = hd * (t1 * (Hh - qh * ah) + t2 * sh + t3 * (uh * du + vh * de + wh * dw - qh * dq) +
                    t5 * (Hh + qh * ah)) * nm;
  flx[0] = double(fc1 - fd1);
  flx[1] = fc2 - double(fd2);
  flx[2] = fc3 - double(fd3);
  flx[3] = fc4 - double(fd4);
  flx[4] = double(fc5 - fd5);
}

// AUSM+ pressure weights P+(M) and P-(M) of one side (alpha = 3/16), in FP64.
__device__ __forceinline__ double ausm_pweight(double m, bool sub, double side) {
  if (sub)
    return 0.25 * (m + side) * (m + side) * (2.0 - side * m) +
           side * 0.1875 * m * (m * m - 1.0) * (m * m - 1.0);
  return 0.5 * (1.0 + side * m / fabs(m));
}

__device__ __forceinline__ void ausm_mp(double nxd, double nyd, double nzd, double nmd, float rl,
                                        float ul, float vl, float wl, double PL, float Hl, float al,
                                        float rr, float ur, float vr, float wr, double PR, float Hr,
                                        float ar, double* flx) {
  const float one = 1.f, half = 0.5f, q4 = 0.25f, e8 = 0.125f;
  const float nx = float(nxd), ny = float(nyd), nz = float(nzd), nm = float(nmd);
  float ql = ul * nx + vl * ny + wl * nz;
  float qr = ur * nx + vr * ny + wr * nz;
  float af = half * (al + ar);
  float ml = ql / af, mla = fabsf(ml);
  float mr = qr / af, mra = fabsf(mr);
  const bool subl = mla <= one, subr = mra <= one;
  float mlp = subl ? q4 * (ml + one) * (ml + one) + e8 * (ml * ml - one) * (ml * ml - one)
                   : half * (ml + mla);
  float mrm = subr ? -q4 * (mr - one) * (mr - one) - e8 * (mr * mr - one) * (mr * mr - one)
                   : half * (mr - mra);
  const double wtl = ausm_pweight(double(ml), subl, 1.0);
  const double wtr = ausm_pweight(double(mr), subr, -1.0);
  float mf = mlp + mrm;
  float mfa = fabsf(mf);
  float mfp = half * (mf + mfa);
  float mfm = half * (mf - mfa);
  const double PF = wtl * PL + wtr * PR;
  flx[0] = double((af * (mfp * rl + mfm * rr)) * nm);
  flx[1] = (double(af * (mfp * rl * ul + mfm * rr * ur)) + PF * nxd) * nmd;
  flx[2] = (double(af * (mfp * rl * vl + mfm * rr * vr)) + PF * nyd) * nmd;
  flx[3] = (double(af * (mfp * rl * wl + mfm * rr * wr)) + PF * nzd) * nmd;
  flx[4] = double((af * (mfp * rl * Hl + mfm * rr * Hr)) * nm);
}

// Face state of cell c for a second-order flux: the flux loop reads cv after
// CellGradientsMP's cons->prim->cons round trip, so momentum is rho*(u) with
// u = (1/rho)*m, and velocities are that momentum times 1/rho again.
template <typename T, typename GV>
__device__ __forceinline__ void recon_values(double rd, double pu, double pv, double pw, double pd,
                                             double cx, double cy, double cz, GV gv, double xf,
                                             double yf, double zf, T& r, T& u, T& v, T& w, T& p,
                                             T& ir, double& P) {
  double ird = 1.0 / rd;
  double ud = (rd * pu) * ird;
  double vd = (rd * pv) * ird;
  double wd = (rd * pw) * ird;
  T dx = T(xf - cx), dy = T(yf - cy), dz = T(zf - cz);
  r = T(rd) + T(gv(0)) * dx + T(gv(1)) * dy + T(gv(2)) * dz;
  u = T(ud) + T(gv(3)) * dx + T(gv(4)) * dy + T(gv(5)) * dz;
  v = T(vd) + T(gv(6)) * dx + T(gv(7)) * dy + T(gv(8)) * dz;
  w = T(wd) + T(gv(9)) * dx + T(gv(10)) * dy + T(gv(11)) * dz;
  if constexpr (sizeof(T) == 4) {
    // mixed precision: the face pressure in FP64, the cell's value plus the FP32 increment
    P = pd + double(T(gv(12)) * dx + T(gv(13)) * dy + T(gv(14)) * dz);
    p = T(P);
  } else {
    p = T(pd) + T(gv(12)) * dx + T(gv(13)) * dy + T(gv(14)) * dz;
    P = double(p);
  }
  ir = T(ird);
}

template <typename T, typename GT>
__device__ __forceinline__ void recon_state(const State& S, const Mesh& M, const GT* grad, int c,
                                            double xf, double yf, double zf, T& r, T& u, T& v,
                                            T& w, T& p, T& ir, double& P) {
  const int N = M.N;
  const double4 cc = M.cell[c];
  const GT* g = grad + c;
  recon_values<T>(S.prim[c], S.prim[N + c], S.prim[2 * N + c], S.prim[3 * N + c],
                  S.prim[4 * N + c], cc.x, cc.y, cc.z,
                  [&](int k) { return g[(size_t)k * N]; }, xf, yf, zf, r, u, v, w, p, ir, P);
}

// Spectral radius (|v.n| + c)|A| of every face, interior faces first, then boundary
// faces (rocflu/RFLU_TimeStepInviscid.F90:165-304); velocities are m*(1/rho).
__device__ __forceinline__ double srad_interior(double u1, double v1, double w1, double a1,
                                                double u2, double v2, double w2, double a2,
                                                double4 n) {
  double vc = 0.5 * ((u1 + u2) * n.x + (v1 + v2) * n.y + (w1 + w2) * n.z);
  double as = 0.5 * (a1 + a2);
  return (fabs(vc) + as) * n.w;
}
__device__ __forceinline__ double srad_boundary(double u1, double v1, double w1, double a1,
                                                double4 n) {
  double vc = u1 * n.x + v1 * n.y + w1 * n.z;
  return (fabs(vc) + a1) * n.w;
}

// Velocity of cell c from cv (order 1) or from the primitive copy (order 2), which
// holds (1/rho)*m with the same rounding.
template <int ORDER>
__device__ __forceinline__ void cell_velocity(const State& S, int N, int c, double& u, double& v,
                                              double& w) {
  if (ORDER == 2) {
    u = S.prim[(size_t)N + c]; v = S.prim[2 * (size_t)N + c]; w = S.prim[3 * (size_t)N + c];
  } else {
    double ir = 1.0 / S.cv[c];
    u = S.cv[(size_t)N + c] * ir; v = S.cv[2 * (size_t)N + c] * ir; w = S.cv[3 * (size_t)N + c] * ir;
  }
}

template <int ORDER>
__device__ __forceinline__ double face_srad(const Mesh& M, const State& S, int f) {
  const int N = M.N;
  double u1, v1, w1, u2, v2, w2;
  if (f < M.F) {
    int2 cc = M.f2c[f];
    cell_velocity<ORDER>(S, N, cc.x, u1, v1, w1);
    cell_velocity<ORDER>(S, N, cc.y, u2, v2, w2);
    return srad_interior(u1, v1, w1, S.dv[2 * (size_t)N + cc.x], u2, v2, w2,
                         S.dv[2 * (size_t)N + cc.y], M.fn[f]);
  }
  int b = f - M.F;
  int c = M.bf2c[b];
  cell_velocity<ORDER>(S, N, c, u1, v1, w1);
  return srad_boundary(u1, v1, w1, S.dv[2 * (size_t)N + c], M.bfn[b]);
}

// Second-order interior flux from the two reconstructed face states (the DISCR
// dispatch of rocflu/RFLU_ComputeFluxInv.F90:123-140).
template <int DISCR, typename T>
__device__ __forceinline__ void face_flux2(const Gas& G, T nx, T ny, T nz, T nm, T rl, T ul, T vl,
                                           T wl, T pl, T irl, T rr, T ur, T vr, T wr, T pr,
                                           T irr, T* flx) {
  const T g = T(G.g), cp = T(G.cp), gc = T(G.gc), one = T(1), half = T(0.5);
  if (DISCR == 3) {
    // RFLU_ModHLLCFlux.F90:2262-2267: E of the face state, H with the cell 1/rho.
    T el = pl / (rl * (g - one)) + half * (ul * ul + vl * vl + wl * wl);
    T al = tsqrt(g * pl / rl);
    T Hl = el + irl * pl;
    T er = pr / (rr * (g - one)) + half * (ur * ur + vr * vr + wr * wr);
    T ar = tsqrt(g * pr / rr);
    T Hr = er + irr * pr;
    hllc<T>(g, nx, ny, nz, nm, rl, ul, vl, wl, pl, al, el, Hl, rr, ur, vr, wr, pr, ar, er, Hr,
            flx);
  } else if (DISCR == 1) {
    T tl = pl / (rl * gc), tr = pr / (rr * gc);
    T Hl = cp * tl + half * (ul * ul + vl * vl + wl * wl);
    T Hr = cp * tr + half * (ur * ur + vr * vr + wr * wr);
    roe<T>(g, T(G.eps_entr), T(G.diss_fact), nx, ny, nz, nm, rl, ul, vl, wl, pl, Hl, rr, ur, vr,
           wr, pr, Hr, flx);
  } else {
    T tl = pl / (rl * gc), tr = pr / (rr * gc);
    T al = tsqrt(g * gc * tl), ar = tsqrt(g * gc * tr);
    T Hl = cp * tl + half * (ul * ul + vl * vl + wl * wl);
    T Hr = cp * tr + half * (ur * ur + vr * vr + wr * wr);
    ausm<T>(nx, ny, nz, nm, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, flx);
  }
}

// face_flux2 in mixed precision: FP32 states, FP64 face pressures and geometry.
template <int DISCR>
__device__ __forceinline__ void face_flux2_mp(const Gas& G, double4 n, float rl, float ul,
                                              float vl, float wl, double PL, float irl, float rr,
                                              float ur, float vr, float wr, double PR, float irr,
                                              double* flx) {
  const float g = float(G.g), cp = float(G.cp), gc = float(G.gc), one = 1.f, half = 0.5f;
  const float pl = float(PL), pr = float(PR);
  if (DISCR == 3) {
    float el = pl / (rl * (g - one)) + half * (ul * ul + vl * vl + wl * wl);
    float al = sqrtf(g * pl / rl);
    float Hl = el + irl * pl;
    float er = pr / (rr * (g - one)) + half * (ur * ur + vr * vr + wr * wr);
    float ar = sqrtf(g * pr / rr);
    float Hr = er + irr * pr;
    hllc_mp(g, n.x, n.y, n.z, n.w, rl, ul, vl, wl, PL, al, el, Hl, rr, ur, vr, wr, PR, ar, er, Hr,
            flx);
  } else if (DISCR == 1) {
    float tl = pl / (rl * gc), tr = pr / (rr * gc);
    float Hl = cp * tl + half * (ul * ul + vl * vl + wl * wl);
    float Hr = cp * tr + half * (ur * ur + vr * vr + wr * wr);
    roe_mp(g, float(G.eps_entr), float(G.diss_fact), n.x, n.y, n.z, n.w, rl, ul, vl, wl, PL, Hl,
           rr, ur, vr, wr, PR, Hr, flx);
  } else {
    float tl = pl / (rl * gc), tr = pr / (rr * gc);
    float al = sqrtf(g * gc * tl), ar = sqrtf(g * gc * tr);
    float Hl = cp * tl + half * (ul * ul + vl * vl + wl * wl);
    float Hr = cp * tr + half * (ur * ur + vr * vr + wr * wr);
    ausm_mp(n.x, n.y, n.z, n.w, rl, ul, vl, wl, PL, Hl, al, rr, ur, vr, wr, PR, Hr, ar, flx);
  }
}

// Flux of interior face f (computed with f2c(1,f) = cc.x as the left state).
template <int DISCR, int ORDER, typename T, typename GT>
__device__ __forceinline__ void interior_face_flux(const Mesh& M, const State& S, const Gas& G,
                                                   const GT* grad, int f, int2 cc,
                                                   double* flx) {
  const int N = M.N;
  int c1 = cc.x, c2 = cc.y;
  double4 n = M.fn[f];
  T nx = T(n.x), ny = T(n.y), nz = T(n.z), nm = T(n.w);
  constexpr bool MP = sizeof(T) == 4;   // FP32 arithmetic, FP64 pressure terms
  if (ORDER == 1) {
    const T g = T(G.g);
    (void)g;
    double rl = S.cv[c1], rr = S.cv[c2];
    double irl = 1.0 / rl, irr = 1.0 / rr;
    double ul = S.cv[N + c1] * irl, vl = S.cv[2 * N + c1] * irl, wl = S.cv[3 * N + c1] * irl;
    double ur = S.cv[N + c2] * irr, vr = S.cv[2 * N + c2] * irr, wr = S.cv[3 * N + c2] * irr;
    double pl = S.dv[c1], pr = S.dv[c2];
    if (DISCR == 3) {
      double al = S.dv[2 * N + c1], ar = S.dv[2 * N + c2];
      double el = S.cv[4 * N + c1] * irl, er = S.cv[4 * N + c2] * irr;
      double Hl = el + irl * pl, Hr = er + irr * pr;
      if constexpr (MP)
        hllc_mp(g, n.x, n.y, n.z, n.w, T(rl), T(ul), T(vl), T(wl), pl, T(al), T(el), T(Hl), T(rr),
                T(ur), T(vr), T(wr), pr, T(ar), T(er), T(Hr), flx);
      else
        hllc<T>(g, nx, ny, nz, nm, T(rl), T(ul), T(vl), T(wl), T(pl), T(al), T(el), T(Hl), T(rr),
                T(ur), T(vr), T(wr), T(pr), T(ar), T(er), T(Hr), flx);
    } else if (DISCR == 1) {
      double tl = S.dv[N + c1], tr = S.dv[N + c2];
      double Hl = G.cp * tl + 0.5 * (ul * ul + vl * vl + wl * wl);
      double Hr = G.cp * tr + 0.5 * (ur * ur + vr * vr + wr * wr);
      if constexpr (MP)
        roe_mp(g, T(G.eps_entr), T(G.diss_fact), n.x, n.y, n.z, n.w, T(rl), T(ul), T(vl), T(wl),
               pl, T(Hl), T(rr), T(ur), T(vr), T(wr), pr, T(Hr), flx);
      else
        roe<T>(g, T(G.eps_entr), T(G.diss_fact), nx, ny, nz, nm, T(rl), T(ul), T(vl), T(wl), T(pl),
               T(Hl), T(rr), T(ur), T(vr), T(wr), T(pr), T(Hr), flx);
    } else {
      double al = S.dv[2 * N + c1], ar = S.dv[2 * N + c2];
      double Hl = S.cv[4 * N + c1] * irl + pl / rl;
      double Hr = S.cv[4 * N + c2] * irr + pr / rr;
      if constexpr (MP)
        ausm_mp(n.x, n.y, n.z, n.w, T(rl), T(ul), T(vl), T(wl), pl, T(Hl), T(al), T(rr), T(ur),
                T(vr), T(wr), pr, T(Hr), T(ar), flx);
      else
        ausm<T>(nx, ny, nz, nm, T(rl), T(ul), T(vl), T(wl), T(pl), T(Hl), T(al), T(rr), T(ur),
                T(vr), T(wr), T(pr), T(Hr), T(ar), flx);
    }
  } else {
    double4 xf = M.fc[f];
    T rl, ul, vl, wl, pl, irl, rr, ur, vr, wr, pr, irr;
    double PL, PR;
    recon_state<T, GT>(S, M, grad, c1, xf.x, xf.y, xf.z, rl, ul, vl, wl, pl, irl, PL);
    recon_state<T, GT>(S, M, grad, c2, xf.x, xf.y, xf.z, rr, ur, vr, wr, pr, irr, PR);
    if constexpr (MP)
      face_flux2_mp<DISCR>(G, n, rl, ul, vl, wl, PL, irl, rr, ur, vr, wr, PR, irr, flx);
    else
      face_flux2<DISCR, T>(G, nx, ny, nz, nm, rl, ul, vl, wl, pl, irl, rr, ur, vr, wr, pr, irr, flx);
  }
}

// SRAD: also store the face's spectral radius for the time step (the first stage of
// a fused step sees the step-start state, which is what the time step reads).
template <int DISCR, int ORDER, typename T, typename GT, typename FT, bool SRAD>
__global__ void k_face_flux(Mesh M, State S, Gas G, Cases C, const GT* __restrict__ grad,
                            FT* __restrict__ flux, double* __restrict__ srad,
                            const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  int f = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (f >= M.F) return;
  // both cells in one 8-byte load, before anything waits on either
  const int2 cc = M.f2c[f];
  if (SRAD) srad[f] = face_srad<ORDER>(M, S, f);
  double flx[5];
  if (C.cid) G = C.gas[C.cid[cc.x]];
  interior_face_flux<DISCR, ORDER, T, GT>(M, S, G, grad, f, cc, flx);
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) FluxIO<FT>::put(flux, FB, k, f, flx[k]);
}

// Face tiles of the second-order interior flux.  Faces are ordered by their lower cell,
// so the faces whose lower cell lies in tile t = [t*ts, t*ts + ts) are the contiguous
// range [fptr[t], fptr[t+1]).  A tile's records are its own cells (slots 0..ts-1,
// loaded with consecutive addresses) followed by the higher cells outside the tile
// that its faces reach (eidx[eptr[t]..eptr[t+1]), ascending); fslot[f] holds the
// slots of f2c(1,f) and f2c(2,f) in its low and high 16 bits.
struct FaceTiles {
  const int* fptr;
  const int* eptr;
  const int* eidx;
  const unsigned* fslot;
  int ts;
  int cap;    // records of the largest tile
  int fcap;   // faces of the largest tile
};

// Second-order interior fluxes of one face tile per block: the tile's records (cofg,
// primitive state, gradient; with SRAD also the sound speed) are staged in LDS with
// coalesced loads, then each thread evaluates faces of the tile from LDS with the
// arithmetic of k_face_flux, so the results are the same bits.  The scattered
// per-face gathers of k_face_flux are what bounds it (its address and data paths are
// busy nearly all the time); here only the records of cells outside the tile are
// gathered.
template <int DISCR, typename T, typename GT, typename FT, bool SRAD>
__device__ __forceinline__ void face_tile_body(Mesh M, State S, Gas G, Cases C,
                                                        FaceTiles FTl, const GT* __restrict__ grad,
                                                        FT* __restrict__ flux,
                                                        double* __restrict__ srad, int t, double* lds) {
  const int N = M.N, cap = FTl.cap;
  const int c0 = t * FTl.ts;
  const int nown = min(FTl.ts, N - c0);
  const int e0 = FTl.eptr[t], nrec = nown + (FTl.eptr[t + 1] - e0);
  constexpr int ND = SRAD ? 9 : 8;   // x y z rho u v w p (c)
  GT* lg = (GT*)(lds + (size_t)ND * cap);
  // two records per thread and pass, every load of both issued before the first store
  for (int k0 = threadIdx.x; k0 < nrec; k0 += 2 * blockDim.x) {
    int kr[2], cr[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      kr[r] = k0 + r * (int)blockDim.x;
      const int k = kr[r];
      cr[r] = k < nown ? c0 + k : k < nrec ? FTl.eidx[e0 + k - nown] : c0;
    }
    double4 cc[2];
    double pv[2][5], a[2];
    GT gv[2][15];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int c = cr[r];
      cc[r] = M.cell[c];
#pragma unroll
      for (int q = 0; q < 5; ++q) pv[r][q] = S.prim[(size_t)q * N + c];
#pragma unroll
      for (int q = 0; q < 15; ++q) gv[r][q] = grad[(size_t)q * N + c];
      a[r] = SRAD ? S.dv[2 * (size_t)N + c] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int k = kr[r];
      if (k >= nrec) continue;
      lds[k] = cc[r].x;
      lds[cap + k] = cc[r].y;
      lds[2 * cap + k] = cc[r].z;
#pragma unroll
      for (int q = 0; q < 5; ++q) lds[(size_t)(3 + q) * cap + k] = pv[r][q];
      if (SRAD) lds[(size_t)8 * cap + k] = a[r];
#pragma unroll
      for (int q = 0; q < 15; ++q) lg[(size_t)q * cap + k] = gv[r][q];
    }
  }
  __syncthreads();
  const size_t FB = (size_t)M.F + M.B;
  const int f1 = FTl.fptr[t + 1];
  for (int f = FTl.fptr[t] + (int)threadIdx.x; f < f1; f += blockDim.x) {
    const unsigned sl = FTl.fslot[f];
    const int s1 = (int)(sl & 0xFFFFu), s2 = (int)(sl >> 16);
    const double4 n = M.fn[f];
    const double4 xf = M.fc[f];
    Gas Gf = G;
    if (C.cid) Gf = C.gas[C.cid[M.f2c[f].x]];
    if (SRAD)
      srad[f] = srad_interior(lds[4 * cap + s1], lds[5 * cap + s1], lds[6 * cap + s1],
                              lds[8 * cap + s1], lds[4 * cap + s2], lds[5 * cap + s2],
                              lds[6 * cap + s2], lds[8 * cap + s2], n);
    T rl, ul, vl, wl, pl, irl, rr, ur, vr, wr, pr, irr;
    double PL, PR;
    recon_values<T>(lds[3 * cap + s1], lds[4 * cap + s1], lds[5 * cap + s1], lds[6 * cap + s1],
                    lds[7 * cap + s1], lds[s1], lds[cap + s1], lds[2 * cap + s1],
                    [&](int q) { return lg[(size_t)q * cap + s1]; }, xf.x, xf.y, xf.z, rl, ul,
                    vl, wl, pl, irl, PL);
    recon_values<T>(lds[3 * cap + s2], lds[4 * cap + s2], lds[5 * cap + s2], lds[6 * cap + s2],
                    lds[7 * cap + s2], lds[s2], lds[cap + s2], lds[2 * cap + s2],
                    [&](int q) { return lg[(size_t)q * cap + s2]; }, xf.x, xf.y, xf.z, rr, ur,
                    vr, wr, pr, irr, PR);
    double flx[5];
    if constexpr (sizeof(T) == 4)
      face_flux2_mp<DISCR>(Gf, n, rl, ul, vl, wl, PL, irl, rr, ur, vr, wr, PR, irr, flx);
    else
      face_flux2<DISCR, T>(Gf, T(n.x), T(n.y), T(n.z), T(n.w), rl, ul, vl, wl, pl, irl, rr, ur,
                           vr, wr, pr, irr, flx);
#pragma unroll
    for (int k = 0; k < 5; ++k) FluxIO<FT>::put(flux, FB, k, f, flx[k]);
  }
}

// NT threads per block: 256, or 512 (twice the resident waves for tiles whose records
// hold a CU to two blocks).
template <int DISCR, typename T, typename GT, typename FT, bool SRAD, int NT>
__global__ __launch_bounds__(NT) void k_face_flux_tile(Mesh M, State S, Gas G, Cases C,
                                                        FaceTiles FTl, const GT* __restrict__ grad,
                                                        FT* __restrict__ flux,
                                                        double* __restrict__ srad,
                                                        const Scalars* __restrict__ sc) {
  extern __shared__ double lds[];
  if (gate(sc)) return;  // uniform over the grid: no lane reaches the barrier alone
  face_tile_body<DISCR, T, GT, FT, SRAD>(M, S, G, C, FTl, grad, flux, srad, xcd_block(M.nxcd), lds);
}


// Outflow boundary state (libfloflu/BcondOutflowPerf.F90:111-157), FP64.
__device__ __forceinline__ void bcond_outflow(int bcOpt, double pout, double sxn, double syn,
                                              double szn, double gamma, double rho, double rhou,
                                              double rhov, double rhow, double rhoe,
                                              double press, double* b) {
  double u = rhou / rho, v = rhov / rho, w = rhow / rho;
  double csound = sqrt(gamma * press / rho);
  double mach = sqrt(u * u + v * v + w * w) / csound;
  if (mach < 1.0 && (bcOpt == BCOPT_SUBSONIC || bcOpt == BCOPT_MIXED)) {
    double rrhoc = 1.0 / (rho * csound);
    double deltp = press - pout;
    double rb = rho - deltp / (csound * csound);
    double ub = u + sxn * deltp * rrhoc;
    double vb = v + syn * deltp * rrhoc;
    double wb = w + szn * deltp * rrhoc;
    double vnd = ub * sxn + vb * syn + wb * szn;
    if (vnd < 0.0) {
      ub = copysign(1.0, u) * fmax(fabs(ub), fabs(u));
      vb = copysign(1.0, v) * fmax(fabs(vb), fabs(v));
      wb = copysign(1.0, w) * fmax(fabs(wb), fabs(w));
    }
    b[0] = rb;
    b[1] = rb * ub;
    b[2] = rb * vb;
    b[3] = rb * wb;
    b[4] = rb * (pout / (rb * (gamma - 1.0)) + 0.5 * (ub * ub + vb * vb + wb * wb));
  } else {
    b[0] = rho; b[1] = rhou; b[2] = rhov; b[3] = rhow; b[4] = rhoe;
  }
}

// Elementary functions of the boundary states for any arithmetic T (double here; the
// derivative kernels add dual-number overloads).
__device__ __forceinline__ double tcos(double x) { return cos(x); }
__device__ __forceinline__ double tsin(double x) { return sin(x); }
__device__ __forceinline__ double tpow(double x, double e) { return pow(x, e); }

// Total-condition inflow state (BC_INFLOW_TOTANG, libfloflu/BcondInflowPerf.F90:95-223):
// vals = (ptot, ttot, betah, betav, Mach).  Subsonic faces keep the interior Riemann
// invariant a - (g-1)/2 q.n and solve the total-enthalpy quadratic for the boundary
// sound speed; supersonic faces, and mixed-option faces with |q.n| >= a, take the
// state of the given Mach number.  A subsonic-option face with |q.n| >= a has no
// state in the reference (its outputs stay unset); it is NaN here, which the
// validity check reports.  b = (rho, rho u, rho v, rho w, rho E); returns p.
template <typename T>
__device__ __forceinline__ T bcond_inflow_totang(int bcOptType, int bcOptFixed, const T* vals,
                                                 double sxn, double syn, double szn, const Gas& G,
                                                 T rl, T rul, T rvl, T rwl, T* b) {
  const double rgas = G.rgasp, g = G.gp;
  const T ptot = vals[0], ttot = vals[1], betah = vals[2], betav = vals[3];
  const T mach = bcOptType != BCOPT_SUBSONIC ? vals[4] : T(0.0);
  T pr, tr, ar;
  if (bcOptType == BCOPT_SUBSONIC || bcOptType == BCOPT_MIXED) {
    const double gm1 = g - 1.0;
    const T ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
    const T a02 = T(g * rgas) * ttot;
    const T al = tsqrt(a02 - T(0.5 * (g - 1.0)) * (ul * ul + vl * vl + wl * wl));
    const T ql = ul * T(sxn) + vl * T(syn) + wl * T(szn);
    if (tabs(ql) < al) {
      const T sl = tsqrt(ul * ul + vl * vl + wl * wl);
      const T eta = (bcOptFixed == 0 && sl > T(1.0E-6)) ? ql / sl : T(-1.0);
      const T Rm = al - T(0.5 * gm1) * ql;
      const T disc = T(0.5 * gm1) * (eta * eta) *
                     (a02 / (Rm * Rm) * (T(1.0) + T(0.5 * gm1) * (eta * eta)) - T(1.0));
      T sr;
      if (disc < T(0.0)) {
        ar = tsqrt(a02);
        tr = ttot;
        pr = ptot;
        sr = T(0.0);
      } else {
        ar = Rm / (T(1.0) + T(0.5 * gm1) * eta * eta) * (T(1.0) + tsqrt(disc));
        tr = ar * ar / T(g * rgas);
        pr = ptot * tpow(tr / ttot, g / (g - 1.0));
        const T c2 = ar * ar;
        sr = a02 > c2 ? tsqrt(T(2.0 / (g - 1.0)) * (a02 - c2)) : T(0.0);
      }
      const T rr = pr / (T(rgas) * tr);
      const T ur = sr * tcos(betah) * tcos(betav);
      const T vr = sr * tsin(betah);
      const T wr = sr * tcos(betah) * tsin(betav);
      b[0] = rr;
      b[1] = rr * ur;
      b[2] = rr * vr;
      b[3] = rr * wr;
      b[4] = rr * (pr / (rr * T(g - 1.0)) + T(0.5) * sr * sr);
      return pr;
    }
    if (bcOptType != BCOPT_MIXED) {
      const T nan = T(__builtin_nan(""));
      b[0] = b[1] = b[2] = b[3] = b[4] = nan;
      return nan;
    }
  }
  pr = ptot / tpow(T(1.0) + T(0.5 * (g - 1.0)) * mach * mach, g / (g - 1.0));
  tr = ttot / (T(1.0) + T(0.5 * (g - 1.0)) * mach * mach);
  const T rr = pr / (T(rgas) * tr);
  ar = tsqrt(T(g * rgas) * tr);
  const T ur = mach * ar * tcos(betah) * tcos(betav);
  const T vr = mach * ar * tsin(betah);
  const T wr = mach * ar * tcos(betah) * tsin(betav);
  b[0] = rr;
  b[1] = rr * ur;
  b[2] = rr * vr;
  b[3] = rr * wr;
  b[4] = rr * (pr / (rr * T(g - 1.0)) + T(0.5) * (ur * ur + vr * vr + wr * wr));
  return pr;
}

// Farfield state (BC_FARFIELD, modflu/RFLU_ModRindStates.F90:248-420, no vortex
// correction): vals = (Mach, attack, sideslip, p, T).  The interior pressure is
// formed from rho E.  Returns the boundary pressure.
template <typename T>
__device__ __forceinline__ T rind_farfield(const Gas& G, const T* vals, double nx, double ny,
                                           double nz, T rl, T rul, T rvl, T rwl, T rel, T* b) {
  const double rgas = G.rgasp, g = G.gp;
  const T machInf = vals[0], alphaInf = vals[1], betaInf = vals[2], pInf = vals[3],
          tInf = vals[4];
  const T ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
  const T ql = ul * T(nx) + vl * T(ny) + wl * T(nz);
  const T el = rel / rl;
  const T sl2 = ul * ul + vl * vl + wl * wl;
  const T pl = T(g - 1.0) * rl * (el - T(0.5) * sl2);
  const T ri = pInf / (T(rgas) * tInf);
  const T pi = pInf;
  const T qi = machInf * tsqrt(T(g * rgas) * tInf);
  const T ui = qi * tcos(alphaInf) * tcos(betaInf);
  const T vi = qi * tsin(alphaInf) * tcos(betaInf);
  const T wi = qi * tsin(betaInf);
  if (machInf < T(1.0)) {
    const T al = tsqrt(T(g) * pl / rl);
    T pb, rb, ub, vb, wb;
    if (ql < T(0.0)) {
      pb = T(0.5) * (pi + pl - rl * al * ((ui - ul) * T(nx) + (vi - vl) * T(ny) + (wi - wl) * T(nz)));
      rb = ri - (pi - pb) / (al * al);
      ub = ui - T(nx) * (pi - pb) / (rl * al);
      vb = vi - T(ny) * (pi - pb) / (rl * al);
      wb = wi - T(nz) * (pi - pb) / (rl * al);
    } else {
      pb = pi;
      rb = rl - (pl - pi) / (al * al);
      ub = ul + T(nx) * (pl - pi) / (rl * al);
      vb = vl + T(ny) * (pl - pi) / (rl * al);
      wb = wl + T(nz) * (pl - pi) / (rl * al);
    }
    b[0] = rb;
    b[1] = rb * ub;
    b[2] = rb * vb;
    b[3] = rb * wb;
    b[4] = rb * (pb / (rb * T(g - 1.0)) + T(0.5) * (ub * ub + vb * vb + wb * wb));
    return pb;
  }
  if (ql < T(0.0)) {
    b[0] = ri;
    b[1] = ri * ui;
    b[2] = ri * vi;
    b[3] = ri * wi;
    b[4] = ri * (pi / (ri * T(g - 1.0)) + T(0.5) * qi * qi);
    return pi;
  }
  b[0] = rl; b[1] = rul; b[2] = rvl; b[3] = rwl; b[4] = rel;
  return pl;
}

// Injected gas at a burning face (modflu/RFLU_ModRindStates.F90:459-501): density from
// the interior pressure at the injection temperature, velocity -minj/rho along n.
template <typename T>
__device__ __forceinline__ void rind_inject(const Gas& G, double nx, double ny, double nz, T minj,
                                            T tinj, T pl, double fs, T& rl, T& ul, T& vl, T& wl,
                                            T& Hl) {
  rl = pl / (T(G.rgasp) * tinj);
  const T ql = -minj / rl + T(fs);
  ul = ql * T(nx);
  vl = ql * T(ny);
  wl = ql * T(nz);
  Hl = T(G.cp) * tinj + T(0.5) * (ul * ul + vl * vl + wl * wl);
}

// Wall pressure of a non-burning injection face (modflu/RFLU_ModRindStates.F90:535-584):
// isentropic expansion when the gas recedes from the wall, the piston-shock relation
// when it approaches.
template <typename T>
__device__ __forceinline__ T rind_slipwall(const Gas& G, double nx, double ny, double nz, T rl,
                                           T rul, T rvl, T rwl, double fs, T pl) {
  const double g = G.gp;
  const T irl = T(1.0) / rl;
  const T ul = irl * rul, vl = irl * rvl, wl = irl * rwl;
  const T ql = ul * T(nx) + vl * T(ny) + wl * T(nz) - T(fs);
  const T al = tsqrt(T(g) * pl / rl);
  if (ql < T(0.0)) {
    const T term = T(1.0) + T(0.5 * (g - 1.0)) * ql / al;
    return pl * tpow(term, 2.0 * g / (g - 1.0));
  }
  const double term = (g + 1.0) / 4.0;
  return pl + T(term) * rl * ql * (ql + tsqrt(al * al + T(term * term) * ql * ql) / T(term));
}

// Flux of an injection face (the reference's form for both orders): the injected
// mass, momentum and enthalpy with the interior pressure when burning (minj > 0), a
// wall with the corrected pressure otherwise.
template <typename T>
__device__ __forceinline__ void injection_flux(const Gas& G, const T* v, double nx, double ny,
                                               double nz, double nm, T rl, T rul, T rvl, T rwl,
                                               T pl, T* flx) {
  const double fs = 0.0;
  const T minj = v[0];
  if (minj > T(0.0)) {
    T ri, ui, vi, wi, Hi;
    rind_inject<T>(G, nx, ny, nz, minj, v[1], pl, fs, ri, ui, vi, wi, Hi);
    flx[0] = -minj * T(nm);
    flx[1] = (-minj * ui + pl * T(nx)) * T(nm);
    flx[2] = (-minj * vi + pl * T(ny)) * T(nm);
    flx[3] = (-minj * wi + pl * T(nz)) * T(nm);
    flx[4] = (-minj * Hi + pl * T(fs)) * T(nm);
  } else {
    const T pw = rind_slipwall<T>(G, nx, ny, nz, rl, rul, rvl, rwl, fs, pl);
    flx[0] = T(0.0);
    flx[1] = pw * T(nx) * T(nm);
    flx[2] = pw * T(ny) * T(nm);
    flx[3] = pw * T(nz) * T(nm);
    flx[4] = pw * T(fs) * T(nm);
  }
}

// Flux of a boundary face of patch p from its interior state, by patch type (the
// cases of rocflu/RFLU_CentralFirstPatch.F90 and RFLU_CentralSecondPatch.F90): slip
// wall (0, p n |A|, 0); injection: injected mass, momentum and enthalpy with the
// interior pressure, or a wall with the corrected pressure when minj <= 0; the
// others the central average of the interior and boundary states.
__device__ __forceinline__ void patch_face_flux(const Gas& G, const Patches& P, int p, double nx,
                                                double ny, double nz, double nm, double rl,
                                                double rul, double rvl, double rwl, double rel,
                                                double pl, double* flx) {
  const int bc = P.bc[p];
  if (bc == BC_SLIPWALL) {
    flx[0] = 0.0;
    flx[1] = pl * nx * nm;
    flx[2] = pl * ny * nm;
    flx[3] = pl * nz * nm;
    flx[4] = 0.0;
    return;
  }
  const double* v = P.vals + 5 * (size_t)p;
  if (bc == BC_INJECTION) {
    injection_flux<double>(G, v, nx, ny, nz, nm, rl, rul, rvl, rwl, pl, flx);
    return;
  }
  const int opt = P.opt[p];
  double pr, bs[5];
  if (bc == BC_INFLOW_VELTEMP) {
    pr = inflow_veltemp<double>(opt, v, rl, G.rgasp, G.gp, bs);
  } else if (bc == BC_INFLOW_TOTANG) {
    pr = bcond_inflow_totang<double>(opt, P.opt2[p], v, nx, ny, nz, G, rl, rul, rvl, rwl, bs);
  } else if (bc == BC_FARFIELD) {
    pr = rind_farfield<double>(G, v, nx, ny, nz, rl, rul, rvl, rwl, rel, bs);
  } else {
    pr = opt != BCOPT_SUPERSONIC ? P.pout[p] : pl;
    bcond_outflow(opt, pr, nx, ny, nz, G.gp, rl, rul, rvl, rwl, rel, pl, bs);
  }
  const double ql = (rul * nx + rvl * ny + rwl * nz) / rl;
  const double qr = (bs[1] * nx + bs[2] * ny + bs[3] * nz) / bs[0];
  flx[0] = 0.5 * (ql * rl + qr * bs[0]) * nm;
  flx[1] = 0.5 * (ql * rul + pl * nx + qr * bs[1] + pr * nx) * nm;
  flx[2] = 0.5 * (ql * rvl + pl * ny + qr * bs[2] + pr * ny) * nm;
  flx[3] = 0.5 * (ql * rwl + pl * nz + qr * bs[3] + pr * nz) * nm;
  flx[4] = 0.5 * (ql * (rel + pl) + qr * (bs[4] + pr)) * nm;
}

// Spectral radius of boundary face b.  Injection faces
// (rocflu/RFLU_TimeStepInviscid.F90:214-260) count the injected gas's normal velocity
// and its sound speed at the injection temperature when burning, and the cell's sound
// speed alone otherwise; every other patch the one-sided cell term.
template <int ORDER>
__device__ __forceinline__ double bface_srad(const Mesh& M, const State& S, const Gas& G,
                                             const Patches& P, int b) {
  const int p = M.bpatch[b];
  if (P.bc[p] == BC_INJECTION) {
    const double* v = P.vals + 5 * (size_t)p;
    const double4 n = M.bfn[b];
    const int c = M.bf2c[b];
    const double minj = v[0];
    double vc, a1;
    if (minj > 0.0) {
      double rl, ul, vl, wl, Hl;
      rind_inject<double>(G, n.x, n.y, n.z, minj, v[1], S.dv[c], 0.0, rl, ul, vl, wl, Hl);
      vc = ul * n.x + vl * n.y + wl * n.z;
      a1 = sqrt(G.gp * G.rgasp * v[1]);
    } else {
      vc = 0.0;
      a1 = S.dv[2 * (size_t)M.N + c];
    }
    return (fabs(vc) + a1) * n.w;
  }
  return face_srad<ORDER>(M, S, M.F + b);
}

// Boundary faces (rocflu/RFLU_CentralFirstPatch.F90:571-906,
// rocflu/RFLU_CentralSecondPatch.F90:634-1024).  Slip walls store a zero mass flux.
// Boundary fluxes stay FP64 in every precision mode (they are few).
// Flux of boundary face b (patch fluxes, FP64).
template <int ORDER, typename GT>
__device__ __forceinline__ void boundary_face_flux(const Mesh& M, const State& S, const Gas& G,
                                                   const Patches& P, const GT* grad, int b,
                                                   double* flx) {
  const int N = M.N;
  int c1 = M.bf2c[b];
  int p = M.bpatch[b];
  double4 n = M.bfn[b];
  double nx = n.x, ny = n.y, nz = n.z, nm = n.w;
  double rl, rul, rvl, rwl, rel, pl;
  bool o2 = ORDER == 2 && P.order2[p];
  if (!o2) {
    if (ORDER == 2) {
      // First-order patch in a second-order run: cv as left by CellGradientsMP,
      // momentum rho*((1/rho)*m) and slot 5 holding p
      // (modflu/RFLU_ModConvertCv.F90:560-574).
      rl = S.prim[c1];
      rul = rl * S.prim[N + c1];
      rvl = rl * S.prim[2 * N + c1];
      rwl = rl * S.prim[3 * N + c1];
      rel = S.prim[4 * N + c1];
      pl = S.prim[4 * N + c1];
    } else {
      rl = S.cv[c1];
      rul = S.cv[N + c1];
      rvl = S.cv[2 * N + c1];
      rwl = S.cv[3 * N + c1];
      rel = S.cv[4 * N + c1];
      pl = S.dv[c1];
    }
  } else {
    double4 xf = M.bfc[b];
    double4 cc = M.cell[c1];
    double dx = xf.x - cc.x, dy = xf.y - cc.y, dz = xf.z - cc.z;
    double r0 = S.prim[c1];
    double ir = 1.0 / r0;
    double ul = (r0 * S.prim[N + c1]) * ir;
    double vl = (r0 * S.prim[2 * N + c1]) * ir;
    double wl = (r0 * S.prim[3 * N + c1]) * ir;
    double p0 = S.prim[4 * N + c1];
    const GT* g = grad + c1;
    rl = r0 + double(g[0]) * dx + double(g[(size_t)N]) * dy + double(g[2 * (size_t)N]) * dz;
    ul = ul + double(g[3 * (size_t)N]) * dx + double(g[4 * (size_t)N]) * dy + double(g[5 * (size_t)N]) * dz;
    vl = vl + double(g[6 * (size_t)N]) * dx + double(g[7 * (size_t)N]) * dy + double(g[8 * (size_t)N]) * dz;
    wl = wl + double(g[9 * (size_t)N]) * dx + double(g[10 * (size_t)N]) * dy + double(g[11 * (size_t)N]) * dz;
    pl = p0 + double(g[12 * (size_t)N]) * dx + double(g[13 * (size_t)N]) * dy + double(g[14 * (size_t)N]) * dz;
    rul = rl * ul;
    rvl = rl * vl;
    rwl = rl * wl;
    rel = rl * (pl / (rl * (G.gp - 1.0)) + 0.5 * (ul * ul + vl * vl + wl * wl));
  }
  patch_face_flux(G, P, p, nx, ny, nz, nm, rl, rul, rvl, rwl, rel, pl, flx);
}

template <int ORDER, typename GT, typename FT, bool SRAD>
__device__ __forceinline__ void bface_body(const Mesh& M, const State& S, Gas G, const Cases& C,
                                           const Patches& P, const GT* __restrict__ grad,
                                           FT* __restrict__ flux, double* __restrict__ srad,
                                           int b) {
  if (C.cid) {
    const int c = C.cid[M.bf2c[b]];
    G = C.gas[c];
    if (b == C.bface0[c]) promote(C.sc + c);   // once per case and stage
  }
  if (SRAD) srad[M.F + b] = bface_srad<ORDER>(M, S, G, P, b);
  double flx[5];
  boundary_face_flux<ORDER, GT>(M, S, G, P, grad, b, flx);
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) FluxIO<FT>::put(flux, FB, k, M.F + b, flx[k]);
}

template <int ORDER, typename GT, typename FT, bool SRAD>
__global__ void k_bface_flux(Mesh M, State S, Gas G, Cases C, Patches P,
                             const GT* __restrict__ grad, FT* __restrict__ flux,
                             double* __restrict__ srad, const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= M.B) return;
  bface_body<ORDER, GT, FT, SRAD>(M, S, G, C, P, grad, flux, srad, b);
}

// Dependent variables p, T, c (libflu/RFLU_SetDependentVars.F90:255-284).
__device__ __forceinline__ void dep_vars(const Gas& G, double r, double mx, double my, double mz,
                                         double E, double& p, double& t, double& a) {
  double ir = 1.0 / r;
  double Eo = ir * E;
  double Vm2 = ir * ir * (mx * mx + my * my + mz * mz);
  p = (G.gp - 1.0) * r * (Eo - 0.5 * Vm2);
  t = p / (r * G.rgasp);
  a = sqrt(G.gp * G.rgasp * t);
}

// Gathers the residual of cell i over its cell-to-face list.  Lists of up to 8 faces
// (tetrahedra 4, prisms and pyramids 5, hexahedra 6) load every face id and then every
// flux value before the first addition, so a cell waits for memory twice instead of
// twice per face; the additions keep the list order.
template <typename FT>
__device__ __forceinline__ void gather_residual(const Mesh& M, const FT* __restrict__ flux, int i,
                                                double* acc) {
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) acc[k] = -0.0;
  int e0 = M.c2f_ptr[i], e1 = M.c2f_ptr[i + 1];
  const int cnt = e1 - e0;
  if (cnt <= 8) {
    int sg[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sg[q] = q < cnt ? M.c2f_idx[e0 + q] : 1;
    typename FluxIO<FT>::val_t v[8][5];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const size_t f = (size_t)((sg[q] > 0 ? sg[q] : -sg[q]) - 1);
#pragma unroll
      for (int k = 0; k < 5; ++k) v[q][k] = q < cnt ? FluxIO<FT>::get(flux, FB, k, f) : 0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (q >= cnt) break;
      if (sg[q] > 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] = acc[k] + double(v[q][k]);
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] = acc[k] - double(v[q][k]);
      }
    }
    return;
  }
  for (int e = e0; e < e1; ++e) {
    int s = M.c2f_idx[e];
    if (s > 0) {
      size_t f = (size_t)(s - 1);
#pragma unroll
      for (int k = 0; k < 5; ++k) acc[k] = acc[k] + double(FluxIO<FT>::get(flux, FB, k, f));
    } else {
      size_t f = (size_t)(-s - 1);
#pragma unroll
      for (int k = 0; k < 5; ++k) acc[k] = acc[k] - double(FluxIO<FT>::get(flux, FB, k, f));
    }
  }
}

// gather_residual for a cell of face tile t: the fluxes of the tile's own faces
// [f0, f0+nf) come from their LDS copy lf ([5][fcap]), the others from memory; the
// values and the order of the additions are those of gather_residual.
template <typename FT>
__device__ __forceinline__ void gather_residual_tile(const Mesh& M, const FT* __restrict__ flux,
                                                     int i, int f0, int nf,
                                                     const typename FluxIO<FT>::val_t* lf,
                                                     int fcap, double* acc) {
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) acc[k] = -0.0;
  int e0 = M.c2f_ptr[i], e1 = M.c2f_ptr[i + 1];
  for (int e = e0; e < e1; e += 8) {
    const int cnt = min(8, e1 - e);
    int sg[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sg[q] = q < cnt ? M.c2f_idx[e + q] : 1;
    typename FluxIO<FT>::val_t v[8][5];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f = (sg[q] > 0 ? sg[q] : -sg[q]) - 1;
      const unsigned l = (unsigned)(f - f0);
      if (q < cnt && l < (unsigned)nf) {
#pragma unroll
        for (int k = 0; k < 5; ++k) v[q][k] = lf[k * fcap + (int)l];
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) v[q][k] = q < cnt ? FluxIO<FT>::get(flux, FB, k, (size_t)f) : 0;
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (q >= cnt) break;
      if (sg[q] > 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] = acc[k] + double(v[q][k]);
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] = acc[k] - double(v[q][k]);
      }
    }
  }
}

// Wavefront-level OR/ADD helpers for the rare after-update events.
__device__ __forceinline__ void count_event(unsigned long long* ctr) { atomicAdd(ctr, 1ull); }

// After-update checks of one cell (libfloflu/AfterUpdateMP.F90:134-136).  dv must be
// the dependent variables of the updated cv.  Returns the final p for prim.
__device__ __forceinline__ void after_update_cell(const Gas& G, double* q, double& p, double& t,
                                                  double& a, Scalars* sc) {
  // RFLU_CheckValidity.F90:141-222
  {
    double rrho = 1.0 / q[0];
    double u = rrho * q[1], v = rrho * q[2], w = rrho * q[3];
    if (isnan(q[0]) || isnan(u) || isnan(v) || isnan(w) || isnan(p) || isnan(t)) {
      count_event(&sc->n_invalid);
      sc->fail = 1;
    }
  }
  // RFLU_EnforceBounds.F90:169-231 (p recomputed with its own rounding, dv untouched)
  {
    double r = q[0], ir = 1.0 / r;
    double u = ir * q[1], v = ir * q[2], w = ir * q[3];
    double Eo = ir * q[4];
    double Vm2 = u * u + v * v + w * w;
    double pe = (G.gp - 1.0) * r * (Eo - 0.5 * Vm2);
    if (pe <= 0.0 && fabs(pe) <= 1.0E-12) {
      count_event(&sc->n_bounds_fix);
      pe = 1.0E-12;
      q[4] = r * (pe / (r * (G.gp - 1.0)) + 0.5 * (u * u + v * v + w * w));
    }
  }
  // RFLU_CheckPositivity.F90:147-253
  {
    double rho = q[0], rrho = 1.0 / rho;
    double u = rrho * q[1], v = rrho * q[2], w = rrho * q[3];
    double Vm2 = u * u + v * v + w * w;
    if (p <= 0.0) {
      count_event(&sc->n_pos_fix);
      double Eo = 1.0e4 + 0.5 * Vm2;
      q[4] = rho * Eo;
      dep_vars(G, q[0], q[1], q[2], q[3], q[4], p, t, a);
    }
    if (rho <= 0.0 || p <= 0.0) {
      count_event(&sc->n_negative);
      sc->fail = 1;
    }
  }
}

// RK update (libfloflu/RkUpdateGeneric.F90:255-325) of one cell given its residual.
// q holds the stage-start state on entry and the updated state on return.
template <int RK>
__device__ __forceinline__ void rk_update_cell(const State& S, int N, int i, int stage, double dt,
                                               double vol, const double* acc, double* q) {
  if (RK == 1) {
    const double ark[4] = {0.5, 0.5, 1.0, 1.0 / 6.0};
    const double grk[4] = {1.0, 2.0, 2.0, 1.0};
    double fac = ark[stage - 1] * dt;
    double adtv = fac / vol;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      size_t o = (size_t)k * N + i;
      double old = stage == 1 ? q[k] : S.cv_old[o];
      if (stage == 1) {
        S.cv_old[o] = q[k];
        q[k] = old - adtv * acc[k];
        S.rhs_sum[o] = acc[k];
      } else if (stage == 4) {
        q[k] = old - adtv * (acc[k] + S.rhs_sum[o]);
      } else {
        q[k] = old - adtv * acc[k];
        S.rhs_sum[o] = S.rhs_sum[o] + grk[stage - 1] * acc[k];
      }
    }
  } else if (RK == 2) {
    const double ark[3] = {8.0 / 15.0, 5.0 / 12.0, 0.75};
    const double grk[3] = {0.0, 17.0 / 25.0, 5.0 / 9.0};
    double fac = ark[stage - 1] * dt;
    double adtv = fac / vol;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      size_t o = (size_t)k * N + i;
      if (stage == 1) {
        q[k] = q[k] - adtv * acc[k];
        S.rhs_sum[o] = acc[k];
      } else if (stage == 2) {
        q[k] = q[k] - adtv * (acc[k] - grk[1] * S.rhs_sum[o]);
        S.rhs_sum[o] = acc[k];
      } else {
        q[k] = q[k] - adtv * (acc[k] - grk[2] * S.rhs_sum[o]);
      }
    }
  } else if (RK == 3) {
    double fac = 1.0 * dt;
    double adtv = fac / vol;
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k] = q[k] - adtv * acc[k];
  } else {
    // Shu-Osher SSP-RK3 (no counterpart in the reference)
    double adtv = dt / vol;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      size_t o = (size_t)k * N + i;
      double us = q[k] - adtv * acc[k];
      if (stage == 1) {
        S.cv_old[o] = q[k];
        q[k] = us;
      } else if (stage == 2) {
        q[k] = 0.75 * S.cv_old[o] + 0.25 * us;
      } else {
        q[k] = (1.0 / 3.0) * S.cv_old[o] + (2.0 / 3.0) * us;
      }
    }
  }
}

// Residual gather + RK update (+ dependent variables + after-update checks when
// FUSED) for every cell.  Stage-start state is the current cv, so Wray and Euler
// need no separate cvOld copy.
template <int RK, typename FT, bool FUSED, bool WRITE_PRIM>
__global__ void k_update(Mesh M, State S, Gas G, Cases C, const FT* __restrict__ flux,
                         Scalars* sc, int stage, int last_stage, double dt_in) {
  if (!C.cid && halted(sc)) return;
  const int N = M.N;
  int i = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (i >= M.nact) return;   // virtual cells take their state from the halo exchange
  int first = 0;
  if (C.cid) {
    const int c = C.cid[i];
    sc = C.sc + c;
    if (halted(sc)) return;   // this case is done or failed: its state stays
    G = C.gas[c];
    first = C.cell0[c];
  }
  double dt = dt_in > 0.0 ? dt_in : sc->dt;
  if (last_stage && FUSED && i == first) {
    sc->time = sc->time + dt;
    sc->steps = sc->steps + 1;
  }
  double acc[5];
  gather_residual<FT>(M, flux, i, acc);
  double q[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k] = S.cv[(size_t)k * N + i];
  rk_update_cell<RK>(S, N, i, stage, dt, M.cell[i].w, acc, q);
  if (FUSED) {
    double p, t, a;
    dep_vars(G, q[0], q[1], q[2], q[3], q[4], p, t, a);
    after_update_cell(G, q, p, t, a, sc);
    S.dv[i] = p;
    S.dv[(size_t)N + i] = t;
    S.dv[2 * (size_t)N + i] = a;
    if (WRITE_PRIM) {
      double ir = 1.0 / q[0];
      S.prim[i] = q[0];
      S.prim[(size_t)N + i] = ir * q[1];
      S.prim[2 * (size_t)N + i] = ir * q[2];
      S.prim[3 * (size_t)N + i] = ir * q[3];
      S.prim[4 * (size_t)N + i] = p;
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) S.cv[(size_t)k * N + i] = q[k];
}

// k_update over face tiles: a block of ts threads updates the cells of tile t and
// first copies the fluxes of the tile's own faces (one contiguous range) into LDS with
// consecutive loads, so that only the faces a cell shares with an earlier tile, and
// its boundary faces, are gathered from memory.  Same arithmetic, same bits.
// Tile t's part of k_update_tile for the lanes [0, nl) of a block (lane = this thread's):
// copies the tile's own face fluxes into lf, waits for the whole block (every thread of
// the block calls it once, valid or not), then updates the tile's cells.
template <int RK, typename FT, bool FUSED, bool WRITE_PRIM>
__device__ __forceinline__ void update_tile_body(const Mesh& M, const State& S, Gas G,
                                                 const Cases& C, const FaceTiles& FTl,
                                                 const FT* __restrict__ flux, Scalars* sc,
                                                 int stage, int last_stage, double dt_in, int t,
                                                 int lane, int nl, bool valid, void* lds) {
  const int N = M.N;
  const int f0 = valid ? FTl.fptr[t] : 0, nf = valid ? FTl.fptr[t + 1] - f0 : 0;
  const int fcap = FTl.fcap;
  using VT = typename FluxIO<FT>::val_t;
  VT* lf = (VT*)lds;
  const size_t FB = (size_t)M.F + M.B;
  for (int k = lane; k < nf; k += nl) {
    VT v[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) v[q] = FluxIO<FT>::get(flux, FB, q, (size_t)(f0 + k));
#pragma unroll
    for (int q = 0; q < 5; ++q) lf[q * fcap + k] = v[q];
  }
  __syncthreads();
  if (!valid) return;
  const int i = t * FTl.ts + lane;
  if (i >= M.nact) return;
  int first = 0;
  if (C.cid) {
    const int c = C.cid[i];
    sc = C.sc + c;
    if (halted(sc)) return;
    G = C.gas[c];
    first = C.cell0[c];
  }
  double dt = dt_in > 0.0 ? dt_in : sc->dt;
  if (last_stage && FUSED && i == first) {
    sc->time = sc->time + dt;
    sc->steps = sc->steps + 1;
  }
  double acc[5];
  gather_residual_tile<FT>(M, flux, i, f0, nf, lf, fcap, acc);
  double q[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k] = S.cv[(size_t)k * N + i];
  rk_update_cell<RK>(S, N, i, stage, dt, M.cell[i].w, acc, q);
  if (FUSED) {
    double p, tt, a;
    dep_vars(G, q[0], q[1], q[2], q[3], q[4], p, tt, a);
    after_update_cell(G, q, p, tt, a, sc);
    S.dv[i] = p;
    S.dv[(size_t)N + i] = tt;
    S.dv[2 * (size_t)N + i] = a;
    if (WRITE_PRIM) {
      double ir = 1.0 / q[0];
      S.prim[i] = q[0];
      S.prim[(size_t)N + i] = ir * q[1];
      S.prim[2 * (size_t)N + i] = ir * q[2];
      S.prim[3 * (size_t)N + i] = ir * q[3];
      S.prim[4 * (size_t)N + i] = p;
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) S.cv[(size_t)k * N + i] = q[k];
}

template <int RK, typename FT, bool FUSED, bool WRITE_PRIM>
__global__ void k_update_tile(Mesh M, State S, Gas G, Cases C, FaceTiles FTl,
                              const FT* __restrict__ flux, Scalars* sc, int stage,
                              int last_stage, double dt_in) {
  extern __shared__ double lds[];
  if (!C.cid && halted(sc)) return;  // uniform over the grid
  update_tile_body<RK, FT, FUSED, WRITE_PRIM>(M, S, G, C, FTl, flux, sc, stage, last_stage, dt_in,
                                              xcd_block(M.nxcd), (int)threadIdx.x,
                                              (int)blockDim.x, true, lds);
}

// Dependent variables (and the primitive copy used by second-order kernels).
__global__ void k_set_dv(Mesh M, State S, Gas G, Cases C) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (C.cid) G = C.gas[C.cid[i]];
  double p, t, a;
  double r = S.cv[i];
  dep_vars(G, r, S.cv[(size_t)N + i], S.cv[2 * (size_t)N + i], S.cv[3 * (size_t)N + i],
           S.cv[4 * (size_t)N + i], p, t, a);
  S.dv[i] = p;
  S.dv[(size_t)N + i] = t;
  S.dv[2 * (size_t)N + i] = a;
}

// Primitive copy (rho, u, v, w, p) from cv and dv (CellGradientsMP's conversion,
// modflu/RFLU_ModConvertCv.F90:217-227).
__global__ void k_set_prim(Mesh M, State S) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double r = S.cv[i];
  double ir = 1.0 / r;
  S.prim[i] = r;
  S.prim[(size_t)N + i] = ir * S.cv[(size_t)N + i];
  S.prim[2 * (size_t)N + i] = ir * S.cv[2 * (size_t)N + i];
  S.prim[3 * (size_t)N + i] = ir * S.cv[3 * (size_t)N + i];
  S.prim[4 * (size_t)N + i] = S.dv[i];
}

// Dependent variables and, at order 2, the primitive copy of the cells [i0, i1) --
// RFLU_SetVarsContWrapper over the virtual cells after a halo exchange
// (libfloflu/RungeKuttaMP.F90:451-452), with k_set_dv's and k_set_prim's arithmetic.
template <bool PRIM>
__global__ void k_set_vars_range(Mesh M, State S, Gas G, int i0, int i1) {
  const int N = M.N;
  int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= i1) return;
  double p, t, a;
  double r = S.cv[i];
  dep_vars(G, r, S.cv[(size_t)N + i], S.cv[2 * (size_t)N + i], S.cv[3 * (size_t)N + i],
           S.cv[4 * (size_t)N + i], p, t, a);
  S.dv[i] = p;
  S.dv[(size_t)N + i] = t;
  S.dv[2 * (size_t)N + i] = a;
  if (PRIM) {
    double ir = 1.0 / r;
    S.prim[i] = r;
    S.prim[(size_t)N + i] = ir * S.cv[(size_t)N + i];
    S.prim[2 * (size_t)N + i] = ir * S.cv[2 * (size_t)N + i];
    S.prim[3 * (size_t)N + i] = ir * S.cv[3 * (size_t)N + i];
    S.prim[4 * (size_t)N + i] = p;
  }
}

// Halo copy between two regions on one device (RFLU_MPI_CopyCellData,
// modflu/RFLU_ModMPI.F90:332-390): the ncomp components ([k][ns] layout) of the
// sending region's cells sidx[e] to the receiving region's cells didx[e] ([k][nd]).
template <typename T>
__global__ void k_halo_copy(const T* __restrict__ src, int ns, T* __restrict__ dst, int nd,
                            int ncomp, const int* __restrict__ sidx, const int* __restrict__ didx,
                            int n) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int a = sidx[e], b = didx[e];
  for (int k = 0; k < ncomp; ++k) dst[(size_t)k * nd + b] = src[(size_t)k * ns + a];
}

// A 3x3 transform, row-major (the turn of a symmetry or periodic copy).
struct Xform {
  double m[9];
};

// The halo copy of symmetry/periodic virtual cells: state (ncomp 5: the momentum
// vector turned, y_i = sum_k m_ik x_k) or gradients (ncomp 15, component 3v+d: the
// density and pressure gradients turned as vectors, the velocity gradient G_ij =
// d u_i / d x_j as M (G M^T), sums in index order) (RFLU_RELP_TransformWrapper and
// RFLU_RELP_TransformGradWrapper, modflu/RFLU_ModRelatedPatches.F90:380-435, 673-735).
template <typename T>
__global__ void k_halo_turn(const T* __restrict__ src, int ns, T* __restrict__ dst, int nd,
                            int ncomp, const int* __restrict__ sidx,
                            const int* __restrict__ didx, int n, Xform X) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int a = sidx[e], b = didx[e];
  const double* m = X.m;
  auto in = [&](int k) { return (double)src[(size_t)k * ns + a]; };
  auto out = [&](int k, double v) { dst[(size_t)k * nd + b] = (T)v; };
  if (ncomp == 5) {
    out(0, in(0));
    const double x0 = in(1), x1 = in(2), x2 = in(3);
    for (int i = 0; i < 3; ++i) out(1 + i, m[3 * i] * x0 + m[3 * i + 1] * x1 + m[3 * i + 2] * x2);
    out(4, in(4));
    return;
  }
  for (int v = 0; v < 5; v += 4) {
    const double x0 = in(3 * v), x1 = in(3 * v + 1), x2 = in(3 * v + 2);
    for (int i = 0; i < 3; ++i)
      out(3 * v + i, m[3 * i] * x0 + m[3 * i + 1] * x1 + m[3 * i + 2] * x2);
  }
  double G[9], K[9];
  for (int k = 0; k < 9; ++k) G[k] = in(3 + k);   // rows u, v, w; columns x, y, z
  // M (G M^T), the reference's order
  for (int i = 0; i < 3; ++i)
    for (int l = 0; l < 3; ++l)
      K[3 * i + l] = G[3 * i] * m[3 * l] + G[3 * i + 1] * m[3 * l + 1] + G[3 * i + 2] * m[3 * l + 2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out(3 + 3 * i + j, m[3 * i] * K[j] + m[3 * i + 1] * K[3 + j] + m[3 * i + 2] * K[6 + j]);
}

// After-update checks as a separate pass (the unfused entry points).
template <bool WRITE_PRIM>
__global__ void k_after_update(Mesh M, State S, Gas G, Cases C, Scalars* sc) {
  const int N = M.N;
  if (!C.cid && halted(sc)) return;   // stopped at an earlier stage's failure
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.nact) return;   // the checks cover actual cells (AfterUpdateMP over nCells)
  if (C.cid) {
    G = C.gas[C.cid[i]];
    sc = C.sc + C.cid[i];
    if (halted(sc)) return;
  }
  double q[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k] = S.cv[(size_t)k * N + i];
  double p = S.dv[i], t = S.dv[(size_t)N + i], a = S.dv[2 * (size_t)N + i];
  after_update_cell(G, q, p, t, a, sc);
  S.cv[4 * (size_t)N + i] = q[4];
  S.dv[i] = p;
  S.dv[(size_t)N + i] = t;
  S.dv[2 * (size_t)N + i] = a;
  if (WRITE_PRIM) S.prim[4 * (size_t)N + i] = p;
}

// Residual gather alone (for rflu_gpu_residual).
template <typename FT>
__global__ void k_gather_rhs(Mesh M, const FT* __restrict__ flux, double* __restrict__ rhs) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double acc[5];
  gather_residual<FT>(M, flux, i, acc);
#pragma unroll
  for (int k = 0; k < 5; ++k) rhs[(size_t)k * N + i] = acc[k];
}

// ---------------------------------------------------------------------------------
// Cell gradients: inverse-distance weighted least squares with R-factor weights
// recomputed per stencil member (RFLU_ComputeGradCellsFast_3D,
// modflu/RFLU_ModDifferentiationCells.F90:931-1144), optionally followed in the
// same pass by the Venkatakrishnan limiter (modflu/RFLU_ModLimiters.F90:407-522,
// 1100-1106), which needs only this cell's gradient and its face neighbours' values.
// ---------------------------------------------------------------------------------
template <typename T>
struct GradCoef {
  T c12, c13, c23, c1111, c2222, c3333;
};

template <typename T>
__device__ __forceinline__ GradCoef<T> grad_coef(const Mesh& M, int i) {
  const int N = M.N;
  double r11 = M.moms[i], r12 = M.moms[(size_t)N + i], r22 = M.moms[2 * (size_t)N + i];
  double r13 = M.moms[3 * (size_t)N + i], r23 = M.moms[4 * (size_t)N + i],
         r33 = M.moms[5 * (size_t)N + i];
  double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
  double c12 = -c11 * r12;
  double c13 = -(c11 * r13 + c12 * c22 * r23);
  double c23 = -c22 * r23;
  GradCoef<T> c;
  c.c12 = T(c12); c.c13 = T(c13); c.c23 = T(c23);
  c.c1111 = T(c11 * c11); c.c2222 = T(c22 * c22); c.c3333 = T(c33 * c33);
  return c;
}

// Inverse distance of one stencil member, the reference's 1/SQRT(dx^2 + dy^2 + dz^2).
template <typename T>
__device__ __forceinline__ T grad_term(double ddx, double ddy, double ddz) {
  T dx = T(ddx), dy = T(ddy), dz = T(ddz);
  return inv_sqrt_cr(dx * dx + dy * dy + dz * dz);
}

// Weights of one stencil member (xj - xi) with its inverse distance term.
template <typename T>
__device__ __forceinline__ void grad_weights_t(const GradCoef<T>& c, T term, double ddx,
                                               double ddy, double ddz, T* w) {
  T dx = T(ddx), dy = T(ddy), dz = T(ddz);
  dx = term * dx;
  dy = term * dy;
  dz = term * dz;
  T term1 = c.c1111 * (dx);
  T term2 = c.c2222 * (dy + c.c12 * dx);
  T term3 = c.c3333 * (dz + c.c23 * dy + c.c13 * dx);
  w[0] = term * (term1 + c.c12 * term2 + c.c13 * term3);
  w[1] = term * (term2 + c.c23 * term3);
  w[2] = term * (term3);
}

// Weights of one stencil member (xj - xi) and its contribution (var_j - var_i).
template <typename T>
__device__ __forceinline__ void grad_weights(const GradCoef<T>& c, double ddx, double ddy,
                                             double ddz, T* w) {
  grad_weights_t<T>(c, grad_term<T>(ddx, ddy, ddz), ddx, ddy, ddz, w);
}

template <typename T>
__device__ __forceinline__ void grad_accum(const T* w, const double* dvar, T* g) {
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    T dVar = T(dvar[k]);
    g[3 * k + 0] = g[3 * k + 0] + w[0] * dVar;
    g[3 * k + 1] = g[3 * k + 1] + w[1] * dVar;
    g[3 * k + 2] = g[3 * k + 2] + w[2] * dVar;
  }
}

template <typename T>
__device__ __forceinline__ void grad_member(const GradCoef<T>& c, double ddx, double ddy,
                                            double ddz, const double* dvar, T* g) {
  T w[3];
  grad_weights<T>(c, ddx, ddy, ddz, w);
  grad_accum<T>(w, dvar, g);
}

// Venkatakrishnan limiter of cell i in FP64: varMin/varMax over the interior-face
// neighbours, then the face-wise limiter with dx = cofg - fc (the reference's sign).
// idx(q) gives the handle of the q-th neighbour (negative past the last one) and
// vals(h, v) loads that neighbour's (rho,u,v,w,p).  MIN/MAX reductions do not depend
// on the face order, so the precomputed neighbour slots give the reference's values
// exactly.  With FP64 gradients and at most 8 neighbours (every element type) all
// handles and face offsets are requested before the first use (measured faster); with
// FP32 gradients the extra registers cost more than the waits they save.
template <typename T>
__device__ __forceinline__ void venkat_face(double d2, double dm, double epsq, double& lim) {
  const double TINY = 1.0E-12;
  double numer = (dm * dm + epsq) * d2 + 2 * d2 * d2 * dm;
  double denom = d2 * (dm * dm + 2 * d2 * d2 + dm * d2 + epsq);
  double term = numer / (denom + TINY);
  if (d2 != 0.0) lim = term < lim ? term : lim;
}

template <typename T, typename Idx, typename Vals>
__device__ __forceinline__ void venkat_limit(const Mesh& M, int i, const double* vi, T* g,
                                             Idx idx, Vals vals) {
  const int N = M.N;
  double vmn[5], vmx[5], lim[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) { vmn[k] = vi[k]; vmx[k] = vi[k]; lim[k] = 1.0; }
  if (sizeof(T) == 8 && M.vq <= 8) {
    int h[8];
    double gx[8], gy[8], gz[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) h[q] = q < M.vq ? idx(q) : -1;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool on = q < M.vq;
      gx[q] = on ? M.vgeo[(size_t)(3 * q) * N + i] : 0.0;
      gy[q] = on ? M.vgeo[(size_t)(3 * q + 1) * N + i] : 0.0;
      gz[q] = on ? M.vgeo[(size_t)(3 * q + 2) * N + i] : 0.0;
    }
    const double epsq = M.epsq[i];
    int nv = 8;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (h[q] < 0 && q < nv) nv = q;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (q >= nv) break;
      double vo[5];
      vals(h[q], vo);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        vmn[k] = vmn[k] < vo[k] ? vmn[k] : vo[k];
        vmx[k] = vmx[k] > vo[k] ? vmx[k] : vo[k];
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (q >= nv) break;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        double d2 = double(g[3 * k]) * gx[q] + double(g[3 * k + 1]) * gy[q] + double(g[3 * k + 2]) * gz[q];
        // the reference's two branches (d2 > 0 with varMax, d2 < 0 with varMin) share one
        // expression; selecting the operand keeps the lanes of a wavefront together
        double dm = d2 > 0.0 ? vmx[k] - vi[k] : vmn[k] - vi[k];
        venkat_face<T>(d2, dm, epsq, lim[k]);
      }
    }
  } else {
    int nv = 0;
    for (int q = 0; q < M.vq; ++q) {
      int hq = idx(q);
      if (hq < 0) break;
      ++nv;
      double vo[5];
      vals(hq, vo);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        vmn[k] = vmn[k] < vo[k] ? vmn[k] : vo[k];
        vmx[k] = vmx[k] > vo[k] ? vmx[k] : vo[k];
      }
    }
    const double epsq = M.epsq[i];
    for (int q = 0; q < nv; ++q) {
      double dx = M.vgeo[(size_t)(3 * q) * N + i];
      double dy = M.vgeo[(size_t)(3 * q + 1) * N + i];
      double dz = M.vgeo[(size_t)(3 * q + 2) * N + i];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        double d2 = double(g[3 * k]) * dx + double(g[3 * k + 1]) * dy + double(g[3 * k + 2]) * dz;
        double dm = d2 > 0.0 ? vmx[k] - vi[k] : vmn[k] - vi[k];
        venkat_face<T>(d2, dm, epsq, lim[k]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) g[3 * k + d] = T(lim[k] * double(g[3 * k + d]));
}

template <typename GT, bool VENKAT, int U>
__global__ void k_gradient(Mesh M, State S, GT* __restrict__ grad, const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  const int N = M.N;
  int i = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  GradCoef<GT> c = grad_coef<GT>(M, i);
  double4 ci = M.cell[i];
  double vi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) vi[k] = S.prim[(size_t)k * N + i];
  GT g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = GT(0);
  int s0 = M.sten_ptr[i], s1 = M.sten_ptr[i + 1];
  int s = s0;
  for (; s + U <= s1; s += U) {
    GT w[U][3];
    double dvar[U][5];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int j = M.sten_idx[s + u];
      double4 cj = M.cell[j];
      grad_weights<GT>(c, cj.x - ci.x, cj.y - ci.y, cj.z - ci.z, w[u]);
#pragma unroll
      for (int k = 0; k < 5; ++k) dvar[u][k] = S.prim[(size_t)k * N + j] - vi[k];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) grad_accum<GT>(w[u], dvar[u], g);
  }
  for (; s < s1; ++s) {
    int j = M.sten_idx[s];
    double4 cj = M.cell[j];
    double dvar[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) dvar[k] = S.prim[(size_t)k * N + j] - vi[k];
    grad_member<GT>(c, cj.x - ci.x, cj.y - ci.y, cj.z - ci.z, dvar, g);
  }
  if (VENKAT)
    venkat_limit<GT>(M, i, vi, g, [&](int q) { return M.vnb[(size_t)q * N + i]; },
                     [&](int o, double* vo) {
#pragma unroll
                       for (int k = 0; k < 5; ++k) vo[k] = S.prim[(size_t)k * N + o];
                     });
#pragma unroll
  for (int k = 0; k < 15; ++k) grad[(size_t)k * N + i] = g[k];
}

// Stencil tiles: the union of the stencil members and face neighbours of TS (128, 256
// or 512) consecutive cells is staged in LDS once per tile (centroid and primitive state, 8
// doubles per member), and each lane then walks its own stencil, in the reference's
// member order, reading LDS slots through a coalesced [member][lane] table of 16-bit
// slot numbers; vloc[(q*TS) + lane] (per tile) holds the slot of the lane's q-th
// limiter neighbour (0xFFFF past the last).
struct Tiles {
  const int* uptr;
  const int* uidx;
  const int* lptr;
  const unsigned short* loc;
  const unsigned short* vloc;
  int ts;
  int cap;
};

// LDS records of the formed-weight gradient tiles (centroid and primitive state): four
// arrays of 16-byte pairs (x,y) (z,rho) (u,v) (w,p), each read with one 128-bit LDS
// load (against eight 64-bit arrays: same FP64 time, 17% less with FP32 gradients,
// whose member loop waits on LDS more).
__device__ __forceinline__ void rec_store(double* lds, int cap, int k, double4 c, const double* v) {
  double2* lp = (double2*)lds;
  lp[k] = make_double2(c.x, c.y);
  lp[cap + k] = make_double2(c.z, v[0]);
  lp[2 * cap + k] = make_double2(v[1], v[2]);
  lp[3 * cap + k] = make_double2(v[3], v[4]);
}
__device__ __forceinline__ void rec_load(const double* lds, int cap, int k, double& x, double& y,
                                         double& z, double* v) {
  const double2* lp = (const double2*)lds;
  const double2 a = lp[k], b = lp[cap + k], c = lp[2 * cap + k], d = lp[3 * cap + k];
  x = a.x; y = a.y; z = b.x;
  v[0] = b.y; v[1] = c.x; v[2] = c.y; v[3] = d.x; v[4] = d.y;
}
__device__ __forceinline__ void rec_vals(const double* lds, int cap, int k, double* v) {
  const double2* lp = (const double2*)lds;
  const double2 b = lp[cap + k], c = lp[2 * cap + k], d = lp[3 * cap + k];
  v[0] = b.y; v[1] = c.x; v[2] = c.y; v[3] = d.x; v[4] = d.y;
}

// Stages a tile's union records in LDS: each thread issues the loads of up to four
// records before writing any of them, so the gathers overlap.  Layout: with COORDS the
// pair arrays of rec_store, otherwise rho, u, v, w, p, each an array of cap doubles.
template <bool COORDS>
__device__ __forceinline__ void stage_tile(const Mesh& M, const State& S, const int* uidx, int nu,
                                           int cap, double* lds) {
  const int N = M.N;
  double* lv = lds + (COORDS ? 3 * cap : 0);
  for (int k0 = threadIdx.x; k0 < nu; k0 += 4 * blockDim.x) {
    int jj[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int k = k0 + r * (int)blockDim.x;
      jj[r] = k < nu ? uidx[k] : -1;
    }
    double4 cc[4];
    double pp[4][5];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = jj[r] >= 0 ? jj[r] : 0;
      if (COORDS) cc[r] = M.cell[j];
#pragma unroll
      for (int q = 0; q < 5; ++q) pp[r][q] = S.prim[(size_t)q * N + j];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (jj[r] < 0) continue;
      const int k = k0 + r * (int)blockDim.x;
      if (COORDS) {
        rec_store(lds, cap, k, cc[r], pp[r]);
      } else {
#pragma unroll
        for (int q = 0; q < 5; ++q) lv[q * cap + k] = pp[r][q];
      }
    }
  }
}

// CAPC > 0: the LDS arrays' stride is that compile-time constant (>= T.cap), so the
// arrays' offsets become immediates of the LDS loads instead of per-member additions.
// TW: each member's inverse distance comes from tw ([member][lane], formed once per
// region by k_grad_terms with the same arithmetic) instead of its square root and
// quotient, loaded one group ahead like the slot numbers; same bits.
template <typename GT, bool VENKAT, int U, int TS, int CAPC, bool TW>
__device__ __forceinline__ void grad_lds_body(Mesh M, State S, Tiles T,
                                                      const GT* __restrict__ tw,
                                                      GT* __restrict__ grad, int t, double* lds) {
  const int N = M.N;
  const int cap = CAPC > 0 ? CAPC : T.cap;
  // the cell's own data and first slot numbers are requested before the staging
  const int i = t * T.ts + (int)threadIdx.x;
  const bool live = i < N;
  const int ic = live ? i : N - 1;
  GradCoef<GT> c = grad_coef<GT>(M, ic);
  const double4 ci = M.cell[ic];
  double vi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) vi[k] = S.prim[(size_t)k * N + ic];
  const int cnt = live ? M.sten_ptr[ic + 1] - M.sten_ptr[ic] : 0;
  const size_t lbase = (size_t)T.lptr[t] + threadIdx.x;
  const unsigned short* L = T.loc + lbase;
  const GT* TWp = TW ? tw + lbase : nullptr;
  // Slot numbers (and stored terms) are loaded one group of U members ahead of their use.
  // Slot rows are read past the lane's last member too (the next tile's rows, or the
  // padding rows after the last tile) and those values are never used, so the loads
  // take a running row pointer with immediate offsets; stored terms keep row 0 there.
  int nk[U];
  GT nt[U];
  const unsigned short* Ln = L + (size_t)U * TS;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (TW) {
      const size_t ro = (size_t)(u < cnt ? u : 0) * TS;
      const int kv = (int)L[ro];
      nk[u] = u < cnt ? kv : 0;
      const GT tv = TWp[ro];
      nt[u] = u < cnt ? tv : GT(1);
    } else {
      nk[u] = (int)L[u * TS];
    }
  }
  const int u0 = T.uptr[t];
  stage_tile<true>(M, S, T.uidx + u0, T.uptr[t + 1] - u0, cap, lds);
  __syncthreads();
  if (!live) return;
  GT g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = GT(0);
  int s = 0;
  // U members' weights are formed together (independent sqrt/divide chains), then
  // accumulated one member after the other in stencil order.
  for (; s + U <= cnt; s += U) {
    int kk[U];
    GT tt[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      kk[u] = nk[u];
      if (TW) {
        const int sn = s + U + u;
        const size_t ro = (size_t)(sn < cnt ? sn : 0) * TS;
        const int kv = (int)L[ro];
        nk[u] = sn < cnt ? kv : 0;
        tt[u] = nt[u];
        const GT tv = TWp[ro];
        nt[u] = sn < cnt ? tv : GT(1);
      } else {
        nk[u] = (int)Ln[u * TS];
      }
    }
    Ln += (size_t)U * TS;
    GT w[U][3];
    double dvar[U][5];
    double ddx[U], ddy[U], ddz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int k = kk[u];
      double xj, yj, zj, vj[5];
      rec_load(lds, cap, k, xj, yj, zj, vj);
      ddx[u] = xj - ci.x;
      ddy[u] = yj - ci.y;
      ddz[u] = zj - ci.z;
#pragma unroll
      for (int q = 0; q < 5; ++q) dvar[u][q] = vj[q] - vi[q];
    }
    if (!TW) {
      // the group's inverse distances as independent chains (grad_term's arithmetic)
      GT d2[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const GT dx = GT(ddx[u]), dy = GT(ddy[u]), dz = GT(ddz[u]);
        d2[u] = dx * dx + dy * dy + dz * dz;
      }
      inv_sqrt_cr_n<U>(d2, tt);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) grad_weights_t<GT>(c, tt[u], ddx[u], ddy[u], ddz[u], w[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) grad_accum<GT>(w[u], dvar[u], g);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (s + u < cnt) {
      int k = nk[u];
      double dvar[5], xj, yj, zj, vj[5];
      rec_load(lds, cap, k, xj, yj, zj, vj);
#pragma unroll
      for (int q = 0; q < 5; ++q) dvar[q] = vj[q] - vi[q];
      const double dx = xj - ci.x, dy = yj - ci.y, dz = zj - ci.z;
      GT w[3];
      if (TW) grad_weights_t<GT>(c, nt[u], dx, dy, dz, w);
      else grad_weights<GT>(c, dx, dy, dz, w);
      grad_accum<GT>(w, dvar, g);
    }
  }
  if (VENKAT) {
    const unsigned short* V = T.vloc + (size_t)t * M.vq * T.ts + threadIdx.x;
    venkat_limit<GT>(M, i, vi, g,
                     [&](int q) {
                       int k = V[(size_t)q * T.ts];
                       return k == 0xFFFF ? -1 : k;
                     },
                     [&](int k, double* vo) { rec_vals(lds, cap, k, vo); });
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) grad[(size_t)k * N + i] = g[k];
}

template <typename GT, bool VENKAT, int U, int TS, int CAPC, bool TW>
__global__ __launch_bounds__(TS) void k_gradient_lds(Mesh M, State S, Tiles T,
                                                      const GT* __restrict__ tw,
                                                      GT* __restrict__ grad,
                                                      const Scalars* __restrict__ sc) {
  extern __shared__ double lds[];
  if (gate(sc)) return;  // uniform over the grid: no lane reaches the barrier alone
  grad_lds_body<GT, VENKAT, U, TS, CAPC, TW>(M, S, T, tw, grad, xcd_block(M.nxcd), lds);
}


// Gradient of the mixed mode over the same tiles with FP32 records: each union record is
// staged as eight floats -- its centroid and primitive state less those of the tile's
// first cell, rounded once from FP64 -- so the member differences the loop forms keep
// the accuracy of those deviations; half the LDS of the FP64 records (512-cell tiles fit
// twice per CU) and FP32 arithmetic throughout, the inverse distance from the hardware
// reciprocal square root.  Not bit-identical to the FP64-record kernel (the mixed
// mode's gradients are FP32 either way).  The Venkatakrishnan limiter stays FP64 and
// reads its face neighbours' values from memory.
template <bool VENKAT, int U, int TS>
__global__ __launch_bounds__(TS) void k_gradient_lds32(Mesh M, State S, Tiles T,
                                                        float* __restrict__ grad,
                                                        const Scalars* __restrict__ sc) {
  extern __shared__ double lds_raw[];
  if (gate(sc)) return;  // uniform over the grid: no lane reaches the barrier alone
  const int N = M.N, cap = T.cap;
  float4* la = (float4*)lds_raw;   // x, y, z, rho (less the reference cell's)
  float4* lb = la + cap;           // u, v, w, p
  const int t = xcd_block(M.nxcd);
  const int r0 = t * TS;           // the tile's first cell: the reference
  const int i = r0 + (int)threadIdx.x;
  const bool live = i < N;
  const int ic = live ? i : N - 1;
  const double4 cr = M.cell[r0];
  double vr[5], vi[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    vr[q] = S.prim[(size_t)q * N + r0];
    vi[q] = S.prim[(size_t)q * N + ic];
  }
  GradCoef<float> c = grad_coef<float>(M, ic);
  const double4 ci = M.cell[ic];
  const float xi = float(ci.x - cr.x), yi = float(ci.y - cr.y), zi = float(ci.z - cr.z);
  float vd[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) vd[q] = float(vi[q] - vr[q]);
  const int cnt = live ? M.sten_ptr[ic + 1] - M.sten_ptr[ic] : 0;
  const unsigned short* L = T.loc + T.lptr[t] + threadIdx.x;
  // slot rows one group ahead, past the lane's last member too (never used; padded)
  int nk[U];
#pragma unroll
  for (int u = 0; u < U; ++u) nk[u] = (int)L[u * TS];
  const unsigned short* Ln = L + (size_t)U * TS;
  const int u0 = T.uptr[t], nu = T.uptr[t + 1] - u0;
  for (int k0 = threadIdx.x; k0 < nu; k0 += 2 * TS) {
    int jj[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int k = k0 + r * TS;
      jj[r] = k < nu ? T.uidx[u0 + k] : r0;
    }
    double4 cc[2];
    double pp[2][5];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      cc[r] = M.cell[jj[r]];
#pragma unroll
      for (int q = 0; q < 5; ++q) pp[r][q] = S.prim[(size_t)q * N + jj[r]];
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int k = k0 + r * TS;
      if (k >= nu) continue;
      la[k] = make_float4(float(cc[r].x - cr.x), float(cc[r].y - cr.y), float(cc[r].z - cr.z),
                          float(pp[r][0] - vr[0]));
      lb[k] = make_float4(float(pp[r][1] - vr[1]), float(pp[r][2] - vr[2]),
                          float(pp[r][3] - vr[3]), float(pp[r][4] - vr[4]));
    }
  }
  __syncthreads();
  if (!live) return;
  float g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = 0.f;
  auto member = [&](int k, float* gg) {
    const float4 a = la[k], b = lb[k];
    const float dx = a.x - xi, dy = a.y - yi, dz = a.z - zi;
    const float term = inv_sqrt_cr(dx * dx + dy * dy + dz * dz);
    float w[3];
    grad_weights_t<float>(c, term, dx, dy, dz, w);
    const double dvar[5] = {double(a.w - vd[0]), double(b.x - vd[1]), double(b.y - vd[2]),
                            double(b.z - vd[3]), double(b.w - vd[4])};
    grad_accum<float>(w, dvar, gg);
  };
  int s = 0;
  for (; s + U <= cnt; s += U) {
    int kk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      kk[u] = nk[u];
      nk[u] = (int)Ln[u * TS];
    }
    Ln += (size_t)U * TS;
#pragma unroll
    for (int u = 0; u < U; ++u) member(kk[u], g);
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (s + u < cnt) member(nk[u], g);
  if (VENKAT)
    venkat_limit<float>(M, i, vi, g, [&](int q) { return M.vnb[(size_t)q * N + i]; },
                        [&](int o, double* vo) {
#pragma unroll
                          for (int k = 0; k < 5; ++k) vo[k] = S.prim[(size_t)k * N + o];
                        });
#pragma unroll
  for (int k = 0; k < 15; ++k) grad[(size_t)k * N + i] = g[k];
}

// Moment form of the same least-squares gradient.  The member loop's weights are
// w_j = term^2 (A d_j) with term = 1/|d_j|, d_j = x_j - x_i and A the cell's symmetric
// matrix built from the R-factor coefficients (A = (sum_j d_j d_j^T / |d_j|^2)^-1), so
// the gradient of variable v is A (sum_j (v_j - v_i) d_j / |d_j|^2).  Per member this
// is one reciprocal of |d|^2 (hardware estimate and two Newton steps) and 21 fused
// multiply-adds; A is applied once per cell.  Mathematically equal to the member loop,
// not bit for bit (other rounding, fused operations), so it is a tuning option outside
// the bit-identical set.
template <typename T>
struct MomMat {
  T a00, a01, a02, a11, a12, a22;
};

template <typename T>
__device__ __forceinline__ MomMat<T> mom_mat(const Mesh& M, int i) {
  const GradCoef<double> c = grad_coef<double>(M, i);
  MomMat<T> a;
  a.a00 = T(c.c1111 + c.c12 * c.c12 * c.c2222 + c.c13 * c.c13 * c.c3333);
  a.a01 = T(c.c12 * c.c2222 + c.c13 * c.c23 * c.c3333);
  a.a02 = T(c.c13 * c.c3333);
  a.a11 = T(c.c2222 + c.c23 * c.c23 * c.c3333);
  a.a12 = T(c.c23 * c.c3333);
  a.a22 = T(c.c3333);
  return a;
}

__device__ __forceinline__ double recip_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float recip_nr(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double tfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float tfma(float a, float b, float c) { return fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ void mom_member(double ddx, double ddy, double ddz, const double* dvar,
                                           T* acc) {
  const T dx = T(ddx), dy = T(ddy), dz = T(ddz);
  const T inv = recip_nr(tfma(dz, dz, tfma(dy, dy, dx * dx)));
  const T ex = dx * inv, ey = dy * inv, ez = dz * inv;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const T dv = T(dvar[k]);
    acc[3 * k] = tfma(dv, ex, acc[3 * k]);
    acc[3 * k + 1] = tfma(dv, ey, acc[3 * k + 1]);
    acc[3 * k + 2] = tfma(dv, ez, acc[3 * k + 2]);
  }
}

template <typename T>
__device__ __forceinline__ void mom_apply(const MomMat<T>& a, T* g) {
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const T x = g[3 * k], y = g[3 * k + 1], z = g[3 * k + 2];
    g[3 * k] = tfma(a.a02, z, tfma(a.a01, y, a.a00 * x));
    g[3 * k + 1] = tfma(a.a12, z, tfma(a.a11, y, a.a01 * x));
    g[3 * k + 2] = tfma(a.a22, z, tfma(a.a12, y, a.a02 * x));
  }
}

// The LDS-tile gradient (stage_tile, slot table, Venkatakrishnan limiter as in
// k_gradient_lds) with the moment form of the member loop; U members per group.
template <typename GT, bool VENKAT, int U, int TS, int CAPC>
__device__ __forceinline__ void grad_mom_body(Mesh M, State S, Tiles T,
                                                      GT* __restrict__ grad, int t, double* lds) {
  const int N = M.N;
  const int cap = CAPC > 0 ? CAPC : T.cap;
  const int i = t * T.ts + (int)threadIdx.x;
  const bool live = i < N;
  const int ic = live ? i : N - 1;
  const double4 ci = M.cell[ic];
  double vi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) vi[k] = S.prim[(size_t)k * N + ic];
  const int cnt = live ? M.sten_ptr[ic + 1] - M.sten_ptr[ic] : 0;
  const unsigned short* L = T.loc + T.lptr[t] + threadIdx.x;
  // slot rows one group ahead, past the lane's last member too (never used; padded)
  int nk[U];
#pragma unroll
  for (int u = 0; u < U; ++u) nk[u] = (int)L[u * TS];
  const unsigned short* Ln = L + (size_t)U * TS;
  const int u0 = T.uptr[t];
  stage_tile<true>(M, S, T.uidx + u0, T.uptr[t + 1] - u0, cap, lds);
  __syncthreads();
  if (!live) return;
  GT g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = GT(0);
  int s = 0;
  for (; s + U <= cnt; s += U) {
    int kk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      kk[u] = nk[u];
      nk[u] = (int)Ln[u * TS];
    }
    Ln += (size_t)U * TS;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kk[u];
      double dvar[5];
      {
        double xj, yj, zj, vj[5];
        rec_load(lds, cap, k, xj, yj, zj, vj);
#pragma unroll
        for (int q = 0; q < 5; ++q) dvar[q] = vj[q] - vi[q];
        mom_member<GT>(xj - ci.x, yj - ci.y, zj - ci.z, dvar, g);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (s + u < cnt) {
      const int k = nk[u];
      double dvar[5];
      {
        double xj, yj, zj, vj[5];
        rec_load(lds, cap, k, xj, yj, zj, vj);
#pragma unroll
        for (int q = 0; q < 5; ++q) dvar[q] = vj[q] - vi[q];
        mom_member<GT>(xj - ci.x, yj - ci.y, zj - ci.z, dvar, g);
      }
    }
  }
  mom_apply<GT>(mom_mat<GT>(M, i), g);
  if (VENKAT) {
    const unsigned short* V = T.vloc + (size_t)t * M.vq * T.ts + threadIdx.x;
    venkat_limit<GT>(M, i, vi, g,
                     [&](int q) {
                       int k = V[(size_t)q * T.ts];
                       return k == 0xFFFF ? -1 : k;
                     },
                     [&](int k, double* vo) { rec_vals(lds, cap, k, vo); });
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) grad[(size_t)k * N + i] = g[k];
}

template <typename GT, bool VENKAT, int U, int TS, int CAPC>
__global__ __launch_bounds__(TS) void k_gradient_mom(Mesh M, State S, Tiles T,
                                                      GT* __restrict__ grad,
                                                      const Scalars* __restrict__ sc) {
  extern __shared__ double lds[];
  if (gate(sc)) return;  // uniform over the grid: no lane reaches the barrier alone
  grad_mom_body<GT, VENKAT, U, TS, CAPC>(M, S, T, grad, xcd_block(M.nxcd), lds);
}


// Stencil weights of every (cell, member) pair, formed once per region with the same
// arithmetic as the member loop (the weights depend on the grid only).  Stored in
// the tiles' [member][lane] order: w[d*L + lptr[t] + s*TS + lane] for component d.
template <typename GT>
__global__ void k_grad_weights(Mesh M, Tiles T, size_t L, GT* __restrict__ w) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int t = i / T.ts, lane = i - t * T.ts;
  GradCoef<GT> c = grad_coef<GT>(M, i);
  double4 ci = M.cell[i];
  const size_t base = (size_t)T.lptr[t] + lane;
  const int s0 = M.sten_ptr[i], s1 = M.sten_ptr[i + 1];
  for (int s = s0; s < s1; ++s) {
    double4 cj = M.cell[M.sten_idx[s]];
    GT ww[3];
    grad_weights<GT>(c, cj.x - ci.x, cj.y - ci.y, cj.z - ci.z, ww);
    size_t o = base + (size_t)(s - s0) * T.ts;
    w[o] = ww[0];
    w[L + o] = ww[1];
    w[2 * L + o] = ww[2];
  }
}

// Inverse distances of every (cell, member) pair in [member][lane] order of the tiles
// (the stored terms of k_gradient_lds<..., TW = true>).
template <typename GT>
__global__ void k_grad_terms(Mesh M, Tiles T, GT* __restrict__ tw) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int t = i / T.ts, lane = i - t * T.ts;
  double4 ci = M.cell[i];
  const size_t base = (size_t)T.lptr[t] + lane;
  const int s0 = M.sten_ptr[i], s1 = M.sten_ptr[i + 1];
  for (int s = s0; s < s1; ++s) {
    double4 cj = M.cell[M.sten_idx[s]];
    tw[base + (size_t)(s - s0) * T.ts] = grad_term<GT>(cj.x - ci.x, cj.y - ci.y, cj.z - ci.z);
  }
}

// Gradient with precomputed weights: only the members' primitive state is staged
// in LDS (5 doubles per record), weights stream from HBM in [member][lane] order
// through a ring of P members loaded ahead of their use (nontemporal loads: read once
// per pass, they should not evict the state the tiles gather), and the accumulation
// runs in stencil order as in the member loop above.
template <typename GT, bool VENKAT, int P, int TS>
__global__ __launch_bounds__(TS) void k_gradient_ldsw(Mesh M, State S, Tiles T, size_t L,
                                                       const GT* __restrict__ w,
                                                       GT* __restrict__ grad,
                                                       const Scalars* __restrict__ sc) {
  extern __shared__ double lds[];
  if (gate(sc)) return;  // uniform over the grid: no lane reaches the barrier alone
  const int N = M.N;
  const int cap = T.cap;
  const double* lv = lds;  // 5 arrays: rho, u, v, w, p
  const int t = xcd_block(M.nxcd);
  const int i = t * T.ts + (int)threadIdx.x;
  const bool live = i < N;
  const int ic = live ? i : N - 1;
  double vi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) vi[k] = S.prim[(size_t)k * N + ic];
  const int cnt = live ? M.sten_ptr[ic + 1] - M.sten_ptr[ic] : 0;
  const size_t base = (size_t)T.lptr[t] + threadIdx.x;
  const unsigned short* Lp = T.loc + base;
  const GT* W = w + base;
  int rk[P];
  GT rw[P][3];
  auto fetch = [&](int m, int sn) {
    bool ok = sn < cnt;
    size_t o = (size_t)(ok ? sn : 0) * T.ts;
    rk[m] = ok ? (int)Lp[o] : 0;
    rw[m][0] = ok ? __builtin_nontemporal_load(W + o) : GT(0);
    rw[m][1] = ok ? __builtin_nontemporal_load(W + L + o) : GT(0);
    rw[m][2] = ok ? __builtin_nontemporal_load(W + 2 * L + o) : GT(0);
  };
#pragma unroll
  for (int m = 0; m < P; ++m) fetch(m, m);
  const int u0 = T.uptr[t];
  stage_tile<false>(M, S, T.uidx + u0, T.uptr[t + 1] - u0, cap, lds);
  __syncthreads();
  if (!live) return;
  GT g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = GT(0);
  int s = 0;
  for (; s + P <= cnt; s += P) {
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int k = rk[m];
      GT ww[3] = {rw[m][0], rw[m][1], rw[m][2]};
      fetch(m, s + P + m);
      double dvar[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) dvar[q] = lv[q * cap + k] - vi[q];
      grad_accum<GT>(ww, dvar, g);
    }
  }
#pragma unroll
  for (int m = 0; m < P; ++m) {
    if (s + m < cnt) {
      double dvar[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) dvar[q] = lv[q * cap + rk[m]] - vi[q];
      grad_accum<GT>(rw[m], dvar, g);
    }
  }
  if (VENKAT) {
    const unsigned short* V = T.vloc + (size_t)t * M.vq * T.ts + threadIdx.x;
    venkat_limit<GT>(M, i, vi, g,
                     [&](int q) {
                       int k = V[(size_t)q * T.ts];
                       return k == 0xFFFF ? -1 : k;
                     },
                     [&](int k, double* vo) {
#pragma unroll
                       for (int r = 0; r < 5; ++r) vo[r] = lv[r * cap + k];
                     });
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) grad[(size_t)k * N + i] = g[k];
}

// Constrained gradients at walls (RFLU_ComputeGradCellsConstr,
// modflu/RFLU_ModDifferentiationCells.F90:1180-1554): entry e is one (cell, variable)
// with constraint rows; its rows [rptr[e], rptr[e+1]) hold the stencil members (cell
// index) and then the constraint rows (-1, wall value), each with the 3 weights
// term*aInv(:,row) formed at region creation from the column-scaled pseudo-inverse.
// The result replaces that variable's gradient, accumulated rows in order, before the
// WENO blend or the limiter.
struct Constr {
  int n;
  const int* ecell;
  const int* evar;
  const int* rptr;
  const int* rcell;
  const double* rval;
  const double* rw;   // [3][R]
  int R;
};

template <typename GT>
__global__ void k_grad_constr(Mesh M, State S, Constr C, GT* __restrict__ grad,
                              const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= C.n) return;
  const int N = M.N;
  const int i = C.ecell[e], v = C.evar[e];
  const double vi = S.prim[(size_t)v * N + i];
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int r = C.rptr[e]; r < C.rptr[e + 1]; ++r) {
    const int j = C.rcell[r];
    const double dVar = (j >= 0 ? S.prim[(size_t)v * N + j] : C.rval[r]) - vi;
    gx = gx + C.rw[r] * dVar;
    gy = gy + C.rw[(size_t)C.R + r] * dVar;
    gz = gz + C.rw[2 * (size_t)C.R + r] * dVar;
  }
  grad[(size_t)(3 * v) * N + i] = GT(gx);
  grad[(size_t)(3 * v + 1) * N + i] = GT(gy);
  grad[(size_t)(3 * v + 2) * N + i] = GT(gz);
}

// The Venkatakrishnan limiter as a pass of its own (when constrained gradients sit
// between the least-squares pass and the limiter).
template <typename GT>
__global__ void k_venkat_apply(Mesh M, State S, GT* __restrict__ grad,
                               const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  const int N = M.N;
  int i = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double vi[5];
  GT g[15];
#pragma unroll
  for (int k = 0; k < 5; ++k) vi[k] = S.prim[(size_t)k * N + i];
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = grad[(size_t)k * N + i];
  venkat_limit<GT>(M, i, vi, g, [&](int q) { return M.vnb[(size_t)q * N + i]; },
                   [&](int o, double* vo) {
#pragma unroll
                     for (int k = 0; k < 5; ++k) vo[k] = S.prim[(size_t)k * N + o];
                   });
#pragma unroll
  for (int k = 0; k < 15; ++k) grad[(size_t)k * N + i] = g[k];
}

// Eigen-decomposition of a symmetric 3x3 matrix m (column-major) by cyclic Jacobi
// rotations, as RFLU_SymmetricEigen3 (modflu/RFLU_ModLimiters.F90:791-886): sweeps
// until the off-diagonal squares fall to 1e-24 of the total (at most 50), then values
// ascending with their vectors (columns of vec, vec[i + 3 j]).
__device__ __forceinline__ void sym_eigen3(const double* m, double* vec, double* val) {
  double r[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) r[k] = m[k];
#pragma unroll
  for (int k = 0; k < 9; ++k) vec[k] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
  double total = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) total = total + r[k] * r[k];
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = r[3] * r[3] + r[6] * r[6] + r[7] * r[7];
    if (off <= 1.0E-24 * total) break;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int ip = pq < 2 ? 0 : 1, iq = pq == 0 ? 1 : 2, ir = 3 - ip - iq;
      const double apq = r[ip + 3 * iq];
      if (apq == 0.0) continue;
      const double theta = (r[iq + 3 * iq] - r[ip + 3 * ip]) / (2.0 * apq);
      const double t = fabs(theta) > 1.0E+150
                           ? 0.5 / theta
                           : copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0);
      const double sn = t * c;
      const double arp = r[ir + 3 * ip], arq = r[ir + 3 * iq];
      r[ip + 3 * ip] = r[ip + 3 * ip] - t * r[ip + 3 * iq];
      r[iq + 3 * iq] = r[iq + 3 * iq] + t * r[ip + 3 * iq];
      r[ip + 3 * iq] = 0.0;
      r[iq + 3 * ip] = 0.0;
      r[ir + 3 * ip] = c * arp - sn * arq;
      r[ip + 3 * ir] = r[ir + 3 * ip];
      r[ir + 3 * iq] = sn * arp + c * arq;
      r[iq + 3 * ir] = r[ir + 3 * iq];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double vp = vec[i + 3 * ip], vq = vec[i + 3 * iq];
        vec[i + 3 * ip] = c * vp - sn * vq;
        vec[i + 3 * iq] = sn * vp + c * vq;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) val[i] = r[i + 3 * i];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    int imin = i;
    for (int k = i + 1; k < 3; ++k)
      if (val[k] < val[imin]) imin = k;  // the first of equal minima
    if (imin != i) {
      const double h = val[i];
      val[i] = val[imin];
      val[imin] = h;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double cc = vec[k + 3 * i];
        vec[k + 3 * i] = vec[k + 3 * imin];
        vec[k + 3 * imin] = cc;
      }
    }
  }
}

// RECONST 10 for one cell, after the gradient pass (and the constrained rows):
// RFLU_ComputeLimiterBarthJesp (modflu/RFLU_ModLimiters.F90:132-322) for rho and p,
// RFLU_LimitGradVelocityBarthJesp (:588-764) for the velocity gradient G (limited along
// the principal axes of its symmetric part, against the neighbours' velocity
// differences along those axes), and RFLU_LimitGradCells (:1100-1106).  Face
// neighbours and cofg - fc come from the limiter lists (vnb, vgeo); min/max reductions
// do not depend on the face order, and the face neighbour's velocity difference
// u_o - u_i has the bits of the reference's side-2 projection negated.
template <typename GT>
__global__ void k_barth_jesp(Mesh M, State S, GT* __restrict__ grad,
                             const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  const int N = M.N;
  const int i = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double vi[5], g[15];
#pragma unroll
  for (int k = 0; k < 5; ++k) vi[k] = S.prim[(size_t)k * N + i];
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = double(grad[(size_t)k * N + i]);
  // rho and p: neighbour extrema, then min over faces of min(1, Delta/Delta_2)
  double mn0 = vi[0], mx0 = vi[0], mn4 = vi[4], mx4 = vi[4];
  int nq = 0;
  for (int q = 0; q < M.vq; ++q) {
    const int o = M.vnb[(size_t)q * N + i];
    if (o < 0) break;
    ++nq;
    const double r = S.prim[o], p = S.prim[4 * (size_t)N + o];
    mn0 = mn0 < r ? mn0 : r;
    mx0 = mx0 > r ? mx0 : r;
    mn4 = mn4 < p ? mn4 : p;
    mx4 = mx4 > p ? mx4 : p;
  }
  double l0 = 1.0, l4 = 1.0;
  // velocity: principal axes of the symmetric part of G (G(i,k) = g[3 (1+i) + k])
  double Sm[9], A[9], ev[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) Sm[a + 3 * b] = 0.5 * (g[3 * (1 + a) + b] + g[3 * (1 + b) + a]);
  sym_eigen3(Sm, A, ev);
  double dmn[3] = {0.0, 0.0, 0.0}, dmx[3] = {0.0, 0.0, 0.0}, al[3] = {1.0, 1.0, 1.0};
  for (int q = 0; q < nq; ++q) {
    const int o = M.vnb[(size_t)q * N + i];
    const double dx = M.vgeo[(size_t)(3 * q) * N + i];
    const double dy = M.vgeo[(size_t)(3 * q + 1) * N + i];
    const double dz = M.vgeo[(size_t)(3 * q + 2) * N + i];
    double d2 = g[0] * dx + g[1] * dy + g[2] * dz;
    if (d2 > 0.0) {
      const double t = 1.0 < (mx0 - vi[0]) / d2 ? 1.0 : (mx0 - vi[0]) / d2;
      l0 = t < l0 ? t : l0;
    } else if (d2 < 0.0) {
      const double t = 1.0 < (mn0 - vi[0]) / d2 ? 1.0 : (mn0 - vi[0]) / d2;
      l0 = t < l0 ? t : l0;
    }
    d2 = g[12] * dx + g[13] * dy + g[14] * dz;
    if (d2 > 0.0) {
      const double t = 1.0 < (mx4 - vi[4]) / d2 ? 1.0 : (mx4 - vi[4]) / d2;
      l4 = t < l4 ? t : l4;
    } else if (d2 < 0.0) {
      const double t = 1.0 < (mn4 - vi[4]) / d2 ? 1.0 : (mn4 - vi[4]) / d2;
      l4 = t < l4 ? t : l4;
    }
    const double du[3] = {S.prim[(size_t)N + o] - vi[1], S.prim[2 * (size_t)N + o] - vi[2],
                          S.prim[3 * (size_t)N + o] - vi[3]};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double pr = A[0 + 3 * j] * du[0] + A[1 + 3 * j] * du[1] + A[2 + 3 * j] * du[2];
      dmn[j] = dmn[j] < pr ? dmn[j] : pr;
      dmx[j] = dmx[j] > pr ? dmx[j] : pr;
    }
  }
  for (int q = 0; q < nq; ++q) {
    const double off[3] = {M.vgeo[(size_t)(3 * q) * N + i], M.vgeo[(size_t)(3 * q + 1) * N + i],
                           M.vgeo[(size_t)(3 * q + 2) * N + i]};
    double sh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
      sh[a] = g[3 * (1 + a)] * off[0] + g[3 * (1 + a) + 1] * off[1] + g[3 * (1 + a) + 2] * off[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double pr = A[0 + 3 * j] * sh[0] + A[1 + 3 * j] * sh[1] + A[2 + 3 * j] * sh[2];
      if (pr > 0.0) {
        const double t = dmx[j] / pr;
        al[j] = al[j] < t ? al[j] : t;
      } else if (pr < 0.0) {
        const double t = dmn[j] / pr;
        al[j] = al[j] < t ? al[j] : t;
      }
    }
  }
  double emax = fabs(ev[0]);
  emax = fabs(ev[1]) > emax ? fabs(ev[1]) : emax;
  emax = fabs(ev[2]) > emax ? fabs(ev[2]) : emax;
  const double escale = 1.0E-3 * emax;
  const bool e12 = ev[1] - ev[0] <= escale, e23 = ev[2] - ev[1] <= escale;
  if (e12 && e23) {
    double m = al[0] < al[1] ? al[0] : al[1];
    m = m < al[2] ? m : al[2];
    al[0] = al[1] = al[2] = m;
  } else if (e12) {
    const double m = al[0] < al[1] ? al[0] : al[1];
    al[0] = al[1] = m;
  } else if (e23) {
    const double m = al[1] < al[2] ? al[1] : al[2];
    al[1] = al[2] = m;
  }
  double rot[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int l = 0; l < 3; ++l)
      rot[j + 3 * l] = al[j] * (A[0 + 3 * j] * g[3 + l] + A[1 + 3 * j] * g[6 + l] +
                                A[2 + 3 * j] * g[9 + l]);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int l = 0; l < 3; ++l)
      g[3 * (1 + a) + l] = A[a + 3 * 0] * rot[0 + 3 * l] + A[a + 3 * 1] * rot[1 + 3 * l] +
                           A[a + 3 * 2] * rot[2 + 3 * l];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g[d] = l0 * g[d];
    g[12 + d] = l4 * g[12 + d];
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) grad[(size_t)k * N + i] = GT(g[k]);
}

// Positivity clip of rho and p (modflu/RFLU_ModLimiters.F90:1199-1267): a variable's
// gradient is zeroed when its extrapolation to any face of the cell is not positive.
template <typename T>
__device__ __forceinline__ void weno_clip_cell(const Mesh& M, const State& S, int i, T* g) {
  const int N = M.N;
  double4 ci = M.cell[i];
  const int pos[2] = {0, 4};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int iv = pos[q];
    double varc = S.prim[(size_t)iv * N + i];
    int e0 = M.c2f_ptr[i], e1 = M.c2f_ptr[i + 1];
    for (int e = e0; e < e1; ++e) {
      int sgn = M.c2f_idx[e];
      int f = (sgn > 0 ? sgn : -sgn) - 1;
      double4 xf = f < M.F ? M.fc[f] : M.bfc[f - M.F];
      double dx = xf.x - ci.x, dy = xf.y - ci.y, dz = xf.z - ci.z;
      double varf = varc + double(g[3 * iv]) * dx + double(g[3 * iv + 1]) * dy +
                    double(g[3 * iv + 2]) * dz;
      if (varf <= 0.0) {
        g[3 * iv] = T(0);
        g[3 * iv + 1] = T(0);
        g[3 * iv + 2] = T(0);
        break;
      }
    }
  }
}

// WENO-XYZ blend (modflu/RFLU_ModWENO.F90:1437-1508) followed by the positivity clip
// of rho and p (modflu/RFLU_ModLimiters.F90:1199-1267).
template <typename GT>
__global__ void k_weno_clip(Mesh M, State S, const GT* __restrict__ gin, GT* __restrict__ gout,
                            const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  const int N = M.N;
  int i = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  typedef GT T;
  const T eps = T(1.0E-15);
  int s0 = M.sten_ptr[i], s1 = M.sten_ptr[i + 1];
  // members outermost: each member's index is loaded once and its 15 values are
  // requested together; every component still sums its members in stencil order
  T sumw[15], loc[15], g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    T term = gin[(size_t)k * N + i];
    sumw[k] = T(1) / (term * term + eps);
    loc[k] = sumw[k] * term;
  }
  int jn = s0 < s1 ? M.sten_idx[s0] : 0;
  for (int s = s0; s < s1; ++s) {
    const int j = jn;
    if (s + 1 < s1) jn = M.sten_idx[s + 1];
    T t[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) t[k] = gin[(size_t)k * N + j];
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      T w = T(1) / (t[k] * t[k] + eps);
      sumw[k] = sumw[k] + w;
      loc[k] = loc[k] + w * t[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = loc[k] / sumw[k];
  weno_clip_cell<T>(M, S, i, g);
#pragma unroll
  for (int k = 0; k < 15; ++k) gout[(size_t)k * N + i] = g[k];
}


// The blend over the gradient kernel's stencil tiles: the raw gradients of a tile's
// union records are staged in LDS, components 0-7 and then 8-14 (8 values per record
// fit the gradient's LDS budget in FP64), and every lane blends its own stencil in
// member order through the tile's slot table -- the same sums in the same order.
template <typename GT, int TS>
__global__ __launch_bounds__(TS) void k_weno_clip_lds(Mesh M, State S, Tiles T,
                                                      const GT* __restrict__ gin,
                                                      GT* __restrict__ gout,
                                                      const Scalars* __restrict__ sc) {
  extern __shared__ double lds_raw[];
  if (gate(sc)) return;  // uniform over the grid: no lane reaches the barrier alone
  GT* lds = (GT*)lds_raw;
  const int N = M.N, cap = T.cap;
  const int t = xcd_block(M.nxcd);
  const int i = t * TS + (int)threadIdx.x;
  const bool live = i < N;
  const int ic = live ? i : N - 1;
  const int cnt = live ? M.sten_ptr[ic + 1] - M.sten_ptr[ic] : 0;
  const unsigned short* L = T.loc + T.lptr[t] + threadIdx.x;
  const int u0 = T.uptr[t], nu = T.uptr[t + 1] - u0;
  const GT eps = GT(1.0E-15);
  GT g[15];
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    constexpr int NQ = 8;
    const int k0 = half ? 8 : 0, nq = half ? 7 : 8;
    if (half) __syncthreads();   // the first half's reads are done
    for (int k = threadIdx.x; k < nu; k += TS) {
      const int j = T.uidx[u0 + k];
      GT v[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) v[q] = q < nq ? gin[(size_t)(k0 + q) * N + j] : GT(0);
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (q < nq) lds[q * cap + k] = v[q];
    }
    __syncthreads();
    if (live) {
      GT sumw[NQ], loc[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (q >= nq) break;
        GT term = gin[(size_t)(k0 + q) * N + i];
        sumw[q] = GT(1) / (term * term + eps);
        loc[q] = sumw[q] * term;
      }
      // the next member's slot read one row ahead (the row after the last member is the
      // next tile's or the padding: read, never used)
      int kn = (int)L[0];
      for (int s = 0; s < cnt; ++s) {
        const int k = kn;
        kn = (int)L[(size_t)(s + 1) * TS];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (q >= nq) break;
          const GT tq = lds[q * cap + k];
          GT w = GT(1) / (tq * tq + eps);
          sumw[q] = sumw[q] + w;
          loc[q] = loc[q] + w * tq;
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (q < nq) g[k0 + q] = loc[q] / sumw[q];
    }
  }
  if (!live) return;
  weno_clip_cell<GT>(M, S, i, g);
#pragma unroll
  for (int k = 0; k < 15; ++k) gout[(size_t)k * N + i] = g[k];
}

// ---------------------------------------------------------------------------------
// Time step: per-cell spectral-radius sums gathered over the cell-to-face list in
// the reference's order (rocflu/RFLU_TimeStepInviscid.F90:165-304), block minimum,
// then one block applies RFLU_MinimumTimeStep's rules (:126-245).
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  return v;
}

// Face spectral radii as a pass of their own (the unfused time-step entry point).
__global__ void k_face_srad(Mesh M, State S, Gas G, Cases C, Patches P,
                            double* __restrict__ srad, const Scalars* __restrict__ sc) {
  if (gate(sc)) return;
  int f = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (f >= M.F + M.B) return;
  if (f < M.F) {
    srad[f] = face_srad<1>(M, S, f);
    return;
  }
  if (C.cid) G = C.gas[C.cid[M.bf2c[f - M.F]]];
  srad[f] = bface_srad<1>(M, S, G, P, f - M.F);
}

// Sum of the spectral radii of cell i's faces in cell-to-face order (all loads of a
// list of up to 8 faces issued before the first addition).
__device__ __forceinline__ double cell_srad_sum(const Mesh& M, const double* __restrict__ srad,
                                                int i) {
  const int e0 = M.c2f_ptr[i], e1 = M.c2f_ptr[i + 1];
  const int cnt = e1 - e0;
  double sum = 0.0;
  if (cnt <= 8) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int sgn = q < cnt ? M.c2f_idx[e0 + q] : 1;
      v[q] = q < cnt ? srad[(sgn > 0 ? sgn : -sgn) - 1] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < cnt) sum = sum + v[q];
    return sum;
  }
  for (int e = e0; e < e1; ++e) {
    int sgn = M.c2f_idx[e];
    sum = sum + srad[(sgn > 0 ? sgn : -sgn) - 1];
  }
  return sum;
}

// Local time step vol/sum(srad) per cell, summed over the cell-to-face list in the
// reference's scatter order, block minimum into part[].
__global__ void k_dt_gather(Mesh M, const double* __restrict__ srad, double* __restrict__ part,
                            const Scalars* __restrict__ sc) {
  __shared__ double red[16];
  const int blk = xcd_block(M.nxcd);
  double dtc = HUGE_VAL;
  if (!stopped(sc)) {
    // the minimum does not depend on the order: each block takes a grid-strided share
    // (a lane's first cell is taken as it is, so a NaN stays one as with a cell per lane)
    const int stride = (int)(gridDim.x * blockDim.x);
    bool first = true;
    for (int i = blk * (int)blockDim.x + (int)threadIdx.x; i < M.nact; i += stride) {
      const double d = M.cell[i].w / cell_srad_sum(M, srad, i);
      dtc = first ? d : fmin(dtc, d);
      first = false;
    }
  }
  dtc = wave_min(dtc);
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = dtc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmin(m, red[w]);
    part[blk] = m;
  }
}

// RFLU_MinimumTimeStep's rules (rocflu/RFLU_MinimumTimeStep.F90:126-245) for one case
// given its minimum local dt, then the end-time clamp of RFLU_TimeStepping
// (:445-448).  time_in >= 0 sets global%currentTime from the host first (the unfused
// entry point).
__device__ __forceinline__ void finalize_dt(Scalars* sc, double mm, const CaseParams& cp,
                                            double time_in) {
  sc->dt_local_min = mm;
  if (time_in >= 0.0) sc->time = time_in;
  double wI = 0.0, wC = 0.0;
  double d = cp.cfl * mm;
  if (sc->time <= 1.0E-08) wI = 1.0;
  else wC = 1.0;
  d = wI / cp.dt_imposed + wC / d;
  d = 1.0 / d;
  double dtMin = HUGE_VAL;
  if (d < dtMin) dtMin = d;
  if (dtMin <= 0.0) dtMin = 6.0E-12;
  dtMin = dtMin > cp.dt_min_limit ? dtMin : cp.dt_min_limit;
  if (sc->last_step) sc->finished = 1;
  if (!stopped(sc)) {
    if (sc->time + dtMin > cp.end_time) {
      dtMin = cp.end_time - sc->time;
      sc->last_step = 1;
    }
    sc->dt = dtMin;
  }
}

__global__ void k_dt_finalize(const double* __restrict__ part, int nparts, Scalars* sc,
                              CaseParams cp, double time_in) {
  __shared__ double red[16];
  double m = HUGE_VAL;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) m = fmin(m, part[k]);
  m = wave_min(m);
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) mm = fmin(mm, red[w]);
    finalize_dt(sc, mm, cp, time_in);
  }
}

// Batched cases: per-case minimum of vol/sum(srad) by an order-independent atomic
// minimum on the bits of the (positive) local time steps.  Wavefronts whose lanes all
// belong to one case (all but the few straddling a case boundary) reduce first, then
// the block merges its wavefronts, so a block issues one atomic per case it covers.
// Each case's slot sits on its own 256-byte line (DTMIN_STRIDE doubles apart).
__device__ __forceinline__ void dt_gather_cases_body(const Mesh& M, const Cases& C,
                                                     const double* __restrict__ srad, int blk) {
  __shared__ int wc[16];
  __shared__ double wm[16];
  const int N = M.N;
  int i = blk * blockDim.x + threadIdx.x;
  int c = -1;
  double dtc = HUGE_VAL;
  if (i < N) {
    c = C.cid[i];
    if (!stopped(C.sc + c)) dtc = M.cell[i].w / cell_srad_sum(M, srad, i);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c0 = __shfl(c, 0, 64);
  if (__all(c == c0)) {
    dtc = wave_min(dtc);
    if (lane == 0) {
      wc[wv] = c0;
      wm[wv] = dtc;
    }
  } else {
    if (c >= 0 && dtc < HUGE_VAL)
      atomicMin(C.dtmin + (size_t)c * DTMIN_STRIDE, (unsigned long long)__double_as_longlong(dtc));
    if (lane == 0) wc[wv] = -1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    int cc = -1;
    double m = HUGE_VAL;
    for (int w = 0; w <= nw; ++w) {
      const int cw = w < nw ? wc[w] : -1;
      if (cw != cc) {
        if (cc >= 0 && m < HUGE_VAL)
          atomicMin(C.dtmin + (size_t)cc * DTMIN_STRIDE, (unsigned long long)__double_as_longlong(m));
        cc = cw;
        m = HUGE_VAL;
      }
      if (cw >= 0) m = fmin(m, wm[w]);
    }
  }
}

__global__ void k_dt_gather_cases(Mesh M, Cases C, const double* __restrict__ srad) {
  dt_gather_cases_body(M, C, srad, xcd_block(M.nxcd));
}

__device__ __forceinline__ void dt_finalize_case(const Cases& C, int c) {
  unsigned long long bits = C.dtmin[(size_t)c * DTMIN_STRIDE];
  C.dtmin[(size_t)c * DTMIN_STRIDE] = 0x7FF0000000000000ULL;   // +inf for the next step
  finalize_dt(C.sc + c, __longlong_as_double((long long)bits), C.par[c], -1.0);
}

__global__ void k_dt_finalize_cases(Cases C) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C.K) return;
  dt_finalize_case(C, c);
}

// Outflow mass bookkeeping of stage 1 (rocflu/RFLU_CentralFirstPatch.F90:697-700):
// block c sums case c's outflow faces' mass fluxes in face order per thread stride.
// The fused step launches it after its time step is set, so a case that has finished
// (or failed) counts no mass for a step it does not take.
template <typename FT>
__device__ __forceinline__ void mass_inout_body(const Mesh& M, const Cases& C, const Patches& P,
                                                const FT* __restrict__ flux, int c) {
  __shared__ double sin_[256], sout[256];
  Scalars* sc = C.sc + c;
  double a = 0.0, b = 0.0;
  const bool live = !stopped(sc);
  if (live) {
    for (int k = C.bface0[c] + threadIdx.x; k < C.bface0[c + 1]; k += blockDim.x) {
      int p = M.bpatch[k];
      if (P.bc[p] == BC_SLIPWALL) continue;   // inflow and outflow faces carry mass
      double m = double(FluxIO<FT>::get(flux, (size_t)M.F + M.B, 0, (size_t)M.F + k));
      a = a - fmin(m, 0.0);
      b = b + fmax(m, 0.0);
    }
  }
  sin_[threadIdx.x] = a;
  sout[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sin_[threadIdx.x] += sin_[threadIdx.x + s];
      sout[threadIdx.x] += sout[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && live) {
    sc->mass_in = sc->mass_in + sin_[0];
    sc->mass_out = sc->mass_out + sout[0];
  }
}

template <typename FT>
__global__ void k_mass_inout(Mesh M, Cases C, Patches P, const FT* __restrict__ flux) {
  mass_inout_body<FT>(M, C, P, flux, (int)blockIdx.x);
}

// Clock and counters of cases [c0, c1) as after an upload at `time`.
__global__ void k_fill_scalars(Scalars* scs, int c0, int c1, double time,
                               unsigned long long* dtmin) {
  int c = c0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= c1) return;
  Scalars* sc = scs + c;
  sc->time = time;
  sc->dt = 0.0;
  sc->dt_local_min = 0.0;
  sc->mass_in = 0.0;
  sc->mass_out = 0.0;
  sc->n_invalid = 0;
  sc->n_negative = 0;
  sc->n_bounds_fix = 0;
  sc->n_pos_fix = 0;
  sc->steps = 0;
  sc->abort = 0;
  sc->last_step = 0;
  sc->finished = 0;
  sc->fail = 0;
  dtmin[(size_t)c * DTMIN_STRIDE] = 0x7FF0000000000000ULL;
}

__global__ void k_reset_counters(Scalars* scs, int K) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= K) return;
  Scalars* sc = scs + c;
  sc->mass_in = 0.0;
  sc->mass_out = 0.0;
  sc->n_invalid = 0;
  sc->n_negative = 0;
  sc->n_bounds_fix = 0;
  sc->n_pos_fix = 0;
  sc->abort = 0;
  sc->last_step = 0;
  sc->finished = 0;
  sc->fail = 0;
}

// ---------------------------------------------------------------------------------
// Persistent step of a batch: one launch advances every case by one time step (all RK
// stages).  Its blocks stay resident and take work items of every case in turn --
// gradient tiles, face tiles and boundary faces, the per-case time step, update tiles
// -- with a grid-wide barrier between the phases where the separate launches of
// launch_step have their kernel boundaries, so every value is the one those launches
// compute (same bodies, same order within each item).  The grid is sized by the host
// to the blocks that fit the device at once; every barrier wait is bounded in time and
// a timeout sets `err` (the host reports it) and lets every block run to its end.
// ---------------------------------------------------------------------------------
struct GridBar {
  unsigned* count;   // arrivals at the current barrier
  unsigned* gen;     // barriers completed
  int* err;          // a wait timed out
};

// Arrivals back to zero before a launch (the error word stays set once a wait timed out).
__global__ void k_bar_reset(GridBar gb) { *gb.count = 0u; }

// Bounded grid barrier: block 0..n-1 arrive, the last one opens the next generation.
// Returns false when a wait timed out (now or at an earlier barrier).
__device__ __forceinline__ bool grid_sync(const GridBar& gb) {
  __shared__ int ok;
  // every wave's own stores (not only thread 0's) complete and released at device scope
  // before its block arrives (a release fence: write-back, no invalidation); thread 0
  // polls with relaxed loads and acquires once after the wait, which covers the block
  // (the CU's caches)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g0 = __hip_atomic_load(gb.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned n = gridDim.x;
    if (atomicAdd(gb.count, 1u) == n - 1) {
      __hip_atomic_store(gb.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(gb.gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      // the constant-rate wall clock (100 MHz on MI355X): give up after ~5 s, or after
      // 2^26 polls whatever the clock says
      const unsigned long long t0 = (unsigned long long)wall_clock64();
      for (unsigned polls = 0;
           __hip_atomic_load(gb.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0; ++polls) {
        if (__hip_atomic_load(gb.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if ((unsigned long long)wall_clock64() - t0 > 500000000ull || polls > (1u << 26)) {
          atomicExch(gb.err, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();
    ok = __hip_atomic_load(gb.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
  }
  __syncthreads();
  return ok != 0;
}

struct PersistentPlan {
  int n_grad_tiles;    // 256-cell gradient tiles
  int n_face_tiles;    // face tiles of FTl.ts cells (the update runs over the same tiles)
  int upd_per_iter;    // update tiles per 256-thread block iteration (256 / FTl.ts)
  int upd_lds_stride;  // bytes of LDS per update tile
  int n_cell_blocks;   // 256-cell blocks of the time-step gather
  int flowthrough;     // cases have inflow/outflow patches: stage 1 counts their mass
};

template <typename GT, int DISCR, bool VENKAT, bool FAST>
__global__ __launch_bounds__(256) void k_step_persistent(Mesh M, State S, Gas G, Cases C,
                                                         Patches P, Tiles T, FaceTiles FTl,
                                                         GT* __restrict__ grad,
                                                         double* __restrict__ flux,
                                                         double* __restrict__ srad,
                                                         PersistentPlan pl, GridBar gb) {
  extern __shared__ double lds[];
  const int nb = (int)gridDim.x, b0 = (int)blockIdx.x;
  for (int stage = 1; stage <= 3; ++stage) {
    for (int t = b0; t < pl.n_grad_tiles; t += nb) {
      if (FAST) grad_mom_body<GT, VENKAT, 4, 256, 0>(M, S, T, grad, t, lds);
      else grad_lds_body<GT, VENKAT, 4, 256, 0, false>(M, S, T, nullptr, grad, t, lds);
      __syncthreads();
    }
    if (!grid_sync(gb)) return;
    for (int t = b0; t < pl.n_face_tiles; t += nb) {
      if (stage == 1) face_tile_body<DISCR, double, GT, double, true>(M, S, G, C, FTl, grad, flux, srad, t, lds);
      else face_tile_body<DISCR, double, GT, double, false>(M, S, G, C, FTl, grad, flux, srad, t, lds);
      __syncthreads();
    }
    for (int b = b0 * 256 + (int)threadIdx.x; b < M.B; b += nb * 256) {
      if (stage == 1) bface_body<2, GT, double, true>(M, S, G, C, P, grad, flux, srad, b);
      else bface_body<2, GT, double, false>(M, S, G, C, P, grad, flux, srad, b);
    }
    if (!grid_sync(gb)) return;
    if (stage == 1) {
      for (int k = b0; k < pl.n_cell_blocks; k += nb) {
        dt_gather_cases_body(M, C, srad, k);
        __syncthreads();
      }
      if (!grid_sync(gb)) return;
      for (int c = b0 * 256 + (int)threadIdx.x; c < C.K; c += nb * 256) dt_finalize_case(C, c);
      if (!grid_sync(gb)) return;
      if (pl.flowthrough)
        for (int c = b0; c < C.K; c += nb) {
          mass_inout_body<double>(M, C, P, flux, c);
          __syncthreads();
        }
    }
    const int per = pl.upd_per_iter, ts = FTl.ts;
    const int part = (int)threadIdx.x / ts, lane = (int)threadIdx.x - part * ts;
    for (int t2 = b0; t2 * per < pl.n_face_tiles; t2 += nb) {
      const int t = t2 * per + part;
      update_tile_body<4, double, true, true>(M, S, G, C, FTl, flux, C.sc, stage, stage == 3,
                                              -1.0, t, lane, ts, t < pl.n_face_tiles,
                                              (char*)lds + (size_t)part * pl.upd_lds_stride);
      __syncthreads();
    }
    if (stage < 3 && !grid_sync(gb)) return;
  }
}

}  // namespace rflu
