// Tangent (forward) and adjoint (reverse) derivatives of the residual
// R(cv) = sum over a cell's faces of +-F, for verification of the explicit path's
// discrete adjoint.  First order: below.  Second order (WLSQ gradients, optional
// Venkatakrishnan limiter, linear reconstruction): after the first-order kernels.
//
// Each face flux is re-evaluated in forward-mode dual numbers (value, derivative)
// from the conservative states of its cells -- the dependent variables p, T, c of
// libflu/RFLU_SetDependentVars.F90:255-284 included, since dv is a function of cv
// after every update -- through the same templated flux functions (HLLC, Roe,
// AUSM+) and boundary states as the primal kernels.  One seeded direction gives a
// Jacobian-vector product; the transposed product is assembled column by column
// (the 10 cell-state inputs of an interior face, the 5 of a boundary face and the
// 5 inflow values of an inflow face), contracted with the face's residual adjoint,
// and gathered per cell over the cell-to-face list in a fixed order.
#pragma once

#include "kernels.hpp"

namespace rflu {

// Dual number a + b eps (eps^2 = 0).
struct D1 {
  double v, d;
  __device__ __forceinline__ D1() : v(0.0), d(0.0) {}
  __device__ __forceinline__ D1(double x) : v(x), d(0.0) {}
  __device__ __forceinline__ D1(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return D1(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return D1(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ D1 operator-(D1 a) { return D1(-a.v, -a.d); }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return D1(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ D1 operator/(D1 a, D1 b) {
  double q = a.v / b.v;
  return D1(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ bool operator<(D1 a, D1 b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(D1 a, D1 b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(D1 a, D1 b) { return a.v <= b.v; }
__device__ __forceinline__ D1 tsqrt(D1 a) {
  double s = sqrt(a.v);
  return D1(s, a.d / (2.0 * s));
}
__device__ __forceinline__ D1 tabs(D1 a) { return a.v < 0.0 ? -a : a; }
__device__ __forceinline__ double dval(double a) { return a; }
__device__ __forceinline__ double dval(D1 a) { return a.v; }
__device__ __forceinline__ D1 tcos(D1 a) { return D1(cos(a.v), -sin(a.v) * a.d); }
__device__ __forceinline__ D1 tsin(D1 a) { return D1(sin(a.v), cos(a.v) * a.d); }
__device__ __forceinline__ D1 tpow(D1 a, double e) {
  const double p = pow(a.v, e);
  return D1(p, e * pow(a.v, e - 1.0) * a.d);
}

// Boundary state and pressure of the patches whose state comes from the patch values
// (total-condition inflow, farfield), or false for the other types.
template <typename T>
__device__ __forceinline__ bool rind_state_t(const Gas& G, int bc, int opt, int opt2, const T* vals,
                                             const T* q, double nx, double ny, double nz, T* bs,
                                             T& prb) {
  if (bc == BC_INFLOW_TOTANG) {
    prb = bcond_inflow_totang<T>(opt, opt2, vals, nx, ny, nz, G, q[0], q[1], q[2], q[3], bs);
    return true;
  }
  if (bc == BC_FARFIELD) {
    prb = rind_farfield<T>(G, vals, nx, ny, nz, q[0], q[1], q[2], q[3], q[4], bs);
    return true;
  }
  return false;
}

// Patches whose flux depends on their values (the value adjoints are formed there).
__device__ __forceinline__ bool has_value_inputs(int bc) {
  return bc == BC_INFLOW_VELTEMP || bc == BC_INFLOW_TOTANG || bc == BC_FARFIELD ||
         bc == BC_INJECTION;
}

// Dependent variables p, T, c of a conservative state (RFLU_SetDependentVars).
template <typename T>
__device__ __forceinline__ void dep_vars_t(const Gas& G, const T* q, T& p, T& t, T& a) {
  T ir = T(1.0) / q[0];
  T Eo = ir * q[4];
  T Vm2 = ir * ir * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  p = T(G.gp - 1.0) * q[0] * (Eo - T(0.5) * Vm2);
  t = p / (q[0] * T(G.rgasp));
  a = tsqrt(T(G.gp * G.rgasp) * t);
}

// First-order interior flux from the two cells' conservative states, as
// interior_face_flux<DISCR, 1> forms it from cv and dv.
template <int DISCR, typename T>
__device__ __forceinline__ void flux1_t(const Gas& G, const T* L, const T* Rt, double4 n, T* flx) {
  T pl, tl, al, pr, tr, ar;
  dep_vars_t<T>(G, L, pl, tl, al);
  dep_vars_t<T>(G, Rt, pr, tr, ar);
  const T nx = T(n.x), ny = T(n.y), nz = T(n.z), nm = T(n.w);
  T rl = L[0], rr = Rt[0];
  T irl = T(1.0) / rl, irr = T(1.0) / rr;
  T ul = L[1] * irl, vl = L[2] * irl, wl = L[3] * irl;
  T ur = Rt[1] * irr, vr = Rt[2] * irr, wr = Rt[3] * irr;
  if (DISCR == 3) {
    T el = L[4] * irl, er = Rt[4] * irr;
    T Hl = el + irl * pl, Hr = er + irr * pr;
    hllc<T>(T(G.g), nx, ny, nz, nm, rl, ul, vl, wl, pl, al, el, Hl, rr, ur, vr, wr, pr, ar, er,
            Hr, flx);
  } else if (DISCR == 1) {
    T Hl = T(G.cp) * tl + T(0.5) * (ul * ul + vl * vl + wl * wl);
    T Hr = T(G.cp) * tr + T(0.5) * (ur * ur + vr * vr + wr * wr);
    roe<T>(T(G.g), T(G.eps_entr), T(G.diss_fact), nx, ny, nz, nm, rl, ul, vl, wl, pl, Hl, rr, ur,
           vr, wr, pr, Hr, flx);
  } else {
    T Hl = L[4] * irl + pl / rl;
    T Hr = Rt[4] * irr + pr / rr;
    ausm<T>(nx, ny, nz, nm, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, flx);
  }
}

// Outflow boundary state (libfloflu/BcondOutflowPerf.F90:111-157), generic arithmetic.
template <typename T>
__device__ __forceinline__ void bcond_outflow_t(int bcOpt, T pout, double sxn, double syn,
                                                double szn, double gamma, const T* q, T press,
                                                T* b) {
  T rho = q[0];
  T u = q[1] / rho, v = q[2] / rho, w = q[3] / rho;
  T csound = tsqrt(T(gamma) * press / rho);
  T mach = tsqrt(u * u + v * v + w * w) / csound;
  if (mach < T(1.0) && (bcOpt == BCOPT_SUBSONIC || bcOpt == BCOPT_MIXED)) {
    T rrhoc = T(1.0) / (rho * csound);
    T deltp = press - pout;
    T rb = rho - deltp / (csound * csound);
    T ub = u + T(sxn) * deltp * rrhoc;
    T vb = v + T(syn) * deltp * rrhoc;
    T wb = w + T(szn) * deltp * rrhoc;
    T vnd = ub * T(sxn) + vb * T(syn) + wb * T(szn);
    if (vnd < T(0.0)) {
      // sign(u)*max(|ub|, |u|), componentwise
      T au = tabs(ub) < tabs(u) ? tabs(u) : tabs(ub);
      T av = tabs(vb) < tabs(v) ? tabs(v) : tabs(vb);
      T aw = tabs(wb) < tabs(w) ? tabs(w) : tabs(wb);
      ub = u < T(0.0) ? -au : au;
      vb = v < T(0.0) ? -av : av;
      wb = w < T(0.0) ? -aw : aw;
    }
    b[0] = rb;
    b[1] = rb * ub;
    b[2] = rb * vb;
    b[3] = rb * wb;
    b[4] = rb * (pout / (rb * T(gamma - 1.0)) + T(0.5) * (ub * ub + vb * vb + wb * wb));
  } else {
    for (int k = 0; k < 5; ++k) b[k] = q[k];
  }
}

// First-order boundary-face flux (k_bface_flux at order 1) from the cell's
// conservative state and the patch values.
template <typename T>
__device__ __forceinline__ void bflux1_t(const Gas& G, int bc, int opt, int opt2, double pout,
                                         const T* vals, const T* q, double4 n, T* flx) {
  T pl, tl, al;
  dep_vars_t<T>(G, q, pl, tl, al);
  const double nx = n.x, ny = n.y, nz = n.z, nm = n.w;
  if (bc == BC_SLIPWALL) {
    flx[0] = T(0.0);
    flx[1] = pl * T(nx) * T(nm);
    flx[2] = pl * T(ny) * T(nm);
    flx[3] = pl * T(nz) * T(nm);
    flx[4] = T(0.0);
    return;
  }
  if (bc == BC_INJECTION) {
    injection_flux<T>(G, vals, nx, ny, nz, nm, q[0], q[1], q[2], q[3], pl, flx);
    return;
  }
  T bs[5], pr;
  if (rind_state_t<T>(G, bc, opt, opt2, vals, q, nx, ny, nz, bs, pr)) {
  } else if (bc == BC_INFLOW_VELTEMP) {
    pr = inflow_veltemp<T, T>(opt, vals, q[0], G.rgasp, G.gp, bs);
  } else {
    pr = opt != BCOPT_SUPERSONIC ? T(pout) : pl;
    bcond_outflow_t<T>(opt, pr, nx, ny, nz, G.gp, q, pl, bs);
  }
  T ql = (q[1] * T(nx) + q[2] * T(ny) + q[3] * T(nz)) / q[0];
  T qr = (bs[1] * T(nx) + bs[2] * T(ny) + bs[3] * T(nz)) / bs[0];
  flx[0] = T(0.5) * (ql * q[0] + qr * bs[0]) * T(nm);
  flx[1] = T(0.5) * (ql * q[1] + pl * T(nx) + qr * bs[1] + pr * T(nx)) * T(nm);
  flx[2] = T(0.5) * (ql * q[2] + pl * T(ny) + qr * bs[2] + pr * T(ny)) * T(nm);
  flx[3] = T(0.5) * (ql * q[3] + pl * T(nz) + qr * bs[3] + pr * T(nz)) * T(nm);
  flx[4] = T(0.5) * (ql * (q[4] + pl) + qr * (bs[4] + pr)) * T(nm);
}

__device__ __forceinline__ void load_cv(const State& S, int N, int c, double* q) {
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k] = S.cv[(size_t)k * N + c];
}

// ---- tangent: dF of every face along (dcv, dvals), into dflux[5][F+B] -------------
template <int DISCR>
__global__ void k_face_jvp(Mesh M, State S, Gas G, const double* __restrict__ dcv,
                           double* __restrict__ dflux) {
  int f = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (f >= M.F) return;
  const int N = M.N;
  int2 cc = M.f2c[f];
  double ql[5], qr[5];
  load_cv(S, N, cc.x, ql);
  load_cv(S, N, cc.y, qr);
  D1 L[5], R[5], flx[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    L[k] = D1(ql[k], dcv[(size_t)k * N + cc.x]);
    R[k] = D1(qr[k], dcv[(size_t)k * N + cc.y]);
  }
  flux1_t<DISCR, D1>(G, L, R, M.fn[f], flx);
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) dflux[k * FB + f] = flx[k].d;
}

__global__ void k_bface_jvp(Mesh M, State S, Gas G, Patches P, const double* __restrict__ dcv,
                            const double* __restrict__ dvals, double* __restrict__ dflux) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= M.B) return;
  const int N = M.N;
  const int c = M.bf2c[b], p = M.bpatch[b];
  double q0[5];
  load_cv(S, N, c, q0);
  D1 q[5], vals[5], flx[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    q[k] = D1(q0[k], dcv[(size_t)k * N + c]);
    vals[k] = D1(P.vals[5 * (size_t)p + k], dvals ? dvals[5 * (size_t)p + k] : 0.0);
  }
  bflux1_t<D1>(G, P.bc[p], P.opt[p], P.opt2[p], P.pout[p], vals, q, M.bfn[b], flx);
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) dflux[k * FB + M.F + b] = flx[k].d;
}

// ---- adjoint ----------------------------------------------------------------------
// Interior face f: with w = mu(c1) - mu(c2) (R(c1) += F, R(c2) -= F), the cell-state
// adjoints (dF/dU_c1)^T w and (dF/dU_c2)^T w go to g[j][f], j < 5 for c1, 5..9 c2.
template <int DISCR>
__global__ void k_face_vjp(Mesh M, State S, Gas G, const double* __restrict__ mu,
                           double* __restrict__ g) {
  int f = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (f >= M.F) return;
  const int N = M.N;
  const size_t F = (size_t)M.F;
  int2 cc = M.f2c[f];
  double ql[5], qr[5], w[5];
  load_cv(S, N, cc.x, ql);
  load_cv(S, N, cc.y, qr);
#pragma unroll
  for (int k = 0; k < 5; ++k) w[k] = mu[(size_t)k * N + cc.x] - mu[(size_t)k * N + cc.y];
  const double4 n = M.fn[f];
  for (int j = 0; j < 10; ++j) {
    D1 L[5], R[5], flx[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      L[k] = D1(ql[k], j == k ? 1.0 : 0.0);
      R[k] = D1(qr[k], j == k + 5 ? 1.0 : 0.0);
    }
    flux1_t<DISCR, D1>(G, L, R, n, flx);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) s = s + w[k] * flx[k].d;
    g[(size_t)j * F + f] = s;
  }
}

// Boundary face b of cell c: (dF/dU_c)^T mu(c) into gb[j][b] and, for an inflow
// patch, (dF/dvals)^T mu(c) into gv[j][b] (zero elsewhere).
__global__ void k_bface_vjp(Mesh M, State S, Gas G, Patches P, const double* __restrict__ mu,
                            double* __restrict__ gb, double* __restrict__ gv) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= M.B) return;
  const int N = M.N;
  const size_t B = (size_t)M.B;
  const int c = M.bf2c[b], p = M.bpatch[b];
  const int bc = P.bc[p];
  double q0[5], w[5];
  load_cv(S, N, c, q0);
#pragma unroll
  for (int k = 0; k < 5; ++k) w[k] = mu[(size_t)k * N + c];
  const double4 n = M.bfn[b];
  const int nseed = has_value_inputs(bc) ? 10 : 5;
  for (int j = 0; j < 10; ++j) {
    if (j >= nseed) {
      gv[(size_t)(j - 5) * B + b] = 0.0;
      continue;
    }
    D1 q[5], vals[5], flx[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      q[k] = D1(q0[k], j == k ? 1.0 : 0.0);
      vals[k] = D1(P.vals[5 * (size_t)p + k], j == k + 5 ? 1.0 : 0.0);
    }
    bflux1_t<D1>(G, bc, P.opt[p], P.opt2[p], P.pout[p], vals, q, n, flx);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) s = s + w[k] * flx[k].d;
    if (j < 5) gb[(size_t)j * B + b] = s;
    else gv[(size_t)(j - 5) * B + b] = s;
  }
}

// lambda(c) = sum over the cell's faces of its share, in cell-to-face order.
__global__ void k_cell_vjp(Mesh M, const double* __restrict__ g, const double* __restrict__ gb,
                           double* __restrict__ lam) {
  const int N = M.N;
  int i = xcd_block(M.nxcd) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const size_t F = (size_t)M.F, B = (size_t)M.B;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = M.c2f_ptr[i]; e < M.c2f_ptr[i + 1]; ++e) {
    int s = M.c2f_idx[e];
    if (s > M.F) {
      size_t b = (size_t)(s - 1 - M.F);
#pragma unroll
      for (int k = 0; k < 5; ++k) acc[k] = acc[k] + gb[k * B + b];
    } else if (s > 0) {
      size_t f = (size_t)(s - 1);
#pragma unroll
      for (int k = 0; k < 5; ++k) acc[k] = acc[k] + g[k * F + f];
    } else {
      size_t f = (size_t)(-s - 1);
#pragma unroll
      for (int k = 0; k < 5; ++k) acc[k] = acc[k] + g[(k + 5) * F + f];
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) lam[(size_t)k * N + i] = acc[k];
}

// Per-patch sums of the inflow-value adjoints, one block per patch, fixed order.
__global__ void k_patch_vals_sum(Mesh M, const int* __restrict__ bpatch_ptr,
                                 const double* __restrict__ gv, double* __restrict__ out) {
  __shared__ double red[5][256];
  const int p = blockIdx.x;
  const size_t B = (size_t)M.B;
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = bpatch_ptr[p] + threadIdx.x; b < bpatch_ptr[p + 1]; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < 5; ++k) a[k] = a[k] + gv[k * B + b];
#pragma unroll
  for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 5; ++k) out[5 * p + k] = red[k][0];
}

// ===================================================================================
// Second order.  R(cv) = sum_f +-Phi(recon(prim_c1, g'_c1), recon(prim_c2, g'_c2)) with
//   prim = P(cv) = (rho, m/rho, p(cv))            (CellGradientsMP's conversion)
//   g    = G prim, g_i = sum_s w_is (prim_js - prim_i)   (linear in prim)
//   g'   = lim(prim, g) g                          (Venkatakrishnan, per variable)
// The tangent runs these maps in dual numbers.  The adjoint transposes them one by one,
// each as a gather in a fixed order: face Jacobians seeded in dual numbers (the 5
// primitive inputs of each side through the whole reconstruction and the 5
// reconstructed values, whose adjoint times dx gives the gradient adjoint), the
// limiter at its active face (the minimising face and the neighbour that sets the
// min/max), the gradient through the transposed stencil, and P per cell.
// ===================================================================================

// Primitive state (rho, (1/rho) m, p) of a conservative state, as k_set_prim forms it
// from cv and dv.
template <typename T>
__device__ __forceinline__ void prim_t(const Gas& G, const T* q, T* pr) {
  T p, t, a;
  dep_vars_t<T>(G, q, p, t, a);
  T ir = T(1.0) / q[0];
  pr[0] = q[0];
  pr[1] = ir * q[1];
  pr[2] = ir * q[2];
  pr[3] = ir * q[3];
  pr[4] = p;
}

// Face state of a cell from its primitive state and (limited) gradient at (xf - cofg),
// recon_state's arithmetic; ir is the cell's 1/rho.
template <typename T>
__device__ __forceinline__ void recon_t(const T* pr, const T* g, double dxd, double dyd,
                                        double dzd, T* q, T& ir) {
  T rd = pr[0];
  T ird = T(1.0) / rd;
  T ud = (rd * pr[1]) * ird, vd = (rd * pr[2]) * ird, wd = (rd * pr[3]) * ird;
  T dx = T(dxd), dy = T(dyd), dz = T(dzd);
  q[0] = rd + g[0] * dx + g[1] * dy + g[2] * dz;
  q[1] = ud + g[3] * dx + g[4] * dy + g[5] * dz;
  q[2] = vd + g[6] * dx + g[7] * dy + g[8] * dz;
  q[3] = wd + g[9] * dx + g[10] * dy + g[11] * dz;
  q[4] = pr[4] + g[12] * dx + g[13] * dy + g[14] * dz;
  ir = ird;
}

// Second-order interior flux from the two face states (interior_face_flux<DISCR, 2>).
template <int DISCR, typename T>
__device__ __forceinline__ void flux2_t(const Gas& G, const T* L, T irl, const T* Rt, T irr,
                                        double4 n, T* flx) {
  const T nx = T(n.x), ny = T(n.y), nz = T(n.z), nm = T(n.w);
  const T g = T(G.g), cp = T(G.cp), gc = T(G.gc), one = T(1.0), half = T(0.5);
  T rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  T rr = Rt[0], ur = Rt[1], vr = Rt[2], wr = Rt[3], pr = Rt[4];
  if (DISCR == 3) {
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

// Boundary flux of a second-order run (boundary_face_flux<2>) from the cell's primitive
// state pr and limited gradient g: a second-order patch reconstructs at the face
// (recon_t's values; E from the face state), a first-order patch reads the primitive
// copy with p in the energy slot.
template <typename T>
__device__ __forceinline__ void bflux2_t(const Gas& G, int bc, int opt, int opt2, int order2,
                                         double pout, const T* vals, const T* pr, const T* g,
                                         double dxd, double dyd, double dzd, double4 n, T* flx) {
  const double nx = n.x, ny = n.y, nz = n.z, nm = n.w;
  T q[5], pl;
  if (!order2) {
    q[0] = pr[0];
    q[1] = pr[0] * pr[1];
    q[2] = pr[0] * pr[2];
    q[3] = pr[0] * pr[3];
    q[4] = pr[4];
    pl = pr[4];
  } else {
    T f[5], ir;
    recon_t<T>(pr, g, dxd, dyd, dzd, f, ir);
    T rl = f[0], ul = f[1], vl = f[2], wl = f[3];
    pl = f[4];
    q[0] = rl;
    q[1] = rl * ul;
    q[2] = rl * vl;
    q[3] = rl * wl;
    q[4] = rl * (pl / (rl * T(G.gp - 1.0)) + T(0.5) * (ul * ul + vl * vl + wl * wl));
  }
  if (bc == BC_SLIPWALL) {
    flx[0] = T(0.0);
    flx[1] = pl * T(nx) * T(nm);
    flx[2] = pl * T(ny) * T(nm);
    flx[3] = pl * T(nz) * T(nm);
    flx[4] = T(0.0);
    return;
  }
  if (bc == BC_INJECTION) {
    injection_flux<T>(G, vals, nx, ny, nz, nm, q[0], q[1], q[2], q[3], pl, flx);
    return;
  }
  T bs[5], prb;
  if (rind_state_t<T>(G, bc, opt, opt2, vals, q, nx, ny, nz, bs, prb)) {
  } else if (bc == BC_INFLOW_VELTEMP) {
    prb = inflow_veltemp<T, T>(opt, vals, q[0], G.rgasp, G.gp, bs);
  } else {
    prb = opt != BCOPT_SUPERSONIC ? T(pout) : pl;
    bcond_outflow_t<T>(opt, prb, nx, ny, nz, G.gp, q, pl, bs);
  }
  T ql = (q[1] * T(nx) + q[2] * T(ny) + q[3] * T(nz)) / q[0];
  T qr = (bs[1] * T(nx) + bs[2] * T(ny) + bs[3] * T(nz)) / bs[0];
  flx[0] = T(0.5) * (ql * q[0] + qr * bs[0]) * T(nm);
  flx[1] = T(0.5) * (ql * q[1] + pl * T(nx) + qr * bs[1] + prb * T(nx)) * T(nm);
  flx[2] = T(0.5) * (ql * q[2] + pl * T(ny) + qr * bs[2] + prb * T(ny)) * T(nm);
  flx[3] = T(0.5) * (ql * q[3] + pl * T(nz) + qr * bs[3] + prb * T(nz)) * T(nm);
  flx[4] = T(0.5) * (ql * (q[4] + pl) + qr * (bs[4] + prb)) * T(nm);
}

// Venkatakrishnan function of one face side: numer / (denom + TINY).
template <typename T>
__device__ __forceinline__ T venkat_phi(T d2, T dm, double epsq) {
  const T e = T(epsq);
  T numer = (dm * dm + e) * d2 + T(2.0) * d2 * d2 * dm;
  T denom = d2 * (dm * dm + T(2.0) * d2 * d2 + dm * d2 + e);
  return numer / (denom + T(1.0E-12));
}

// Wall-constrained velocity gradients (CRECONSTC 1) in the derivative kernels: the
// entries and rows that k_grad_constr applies, the entries of every cell (cptr/cent,
// ascending entry) and, for every cell j, the rows that read j's value (trptr over
// trow/tent: row and its entry, ascending row).  cptr is null without constraints.
struct ConstrD {
  Constr C;
  const int* cptr;
  const int* cent;
  const int* trptr;
  const int* trow;
  const int* tent;
};

// Overwrites the raw gradient of cell i with its constrained rows (k_grad_constr's
// arithmetic and order: g_d = sum_r rw_d(r) (value(r) - v_i), value(r) the row cell's
// primitive or the row's wall value).
template <typename T, typename Load>
__device__ __forceinline__ void constrain_t(const ConstrD& CD, int i, Load pr, const T* vi,
                                            T* graw) {
  const Constr& C = CD.C;
  for (int q = CD.cptr[i]; q < CD.cptr[i + 1]; ++q) {
    const int e = CD.cent[q], v = C.evar[e];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
      if (k != v) continue;
      T g0 = T(0.0), g1 = T(0.0), g2 = T(0.0);
      for (int r = C.rptr[e]; r < C.rptr[e + 1]; ++r) {
        const int j = C.rcell[r];
        T vj[5];
        if (j >= 0) pr(j, vj);
        else vj[k] = T(C.rval[r]);
        const T dVar = vj[k] - vi[k];
        g0 = g0 + T(C.rw[r]) * dVar;
        g1 = g1 + T(C.rw[(size_t)C.R + r]) * dVar;
        g2 = g2 + T(C.rw[2 * (size_t)C.R + r]) * dVar;
      }
      graw[3 * k] = g0;
      graw[3 * k + 1] = g1;
      graw[3 * k + 2] = g2;
    }
  }
}

// Gradient and limited gradient of cell i in arithmetic T from the primitive states
// (pr(j) loads cell j's 5 values), venkat_limit's selections: min/max keep the earlier
// value on ties, the limiter the smaller term, faces with d2 = 0 skipped.  Constrained
// entries replace their components between the least-squares pass and the limiter.
template <typename T, bool VENKAT, typename Load>
__device__ __forceinline__ void grad_limited_t(const Mesh& M, const ConstrD& CD, int i, Load pr,
                                               T* graw, T* glim) {
  const int N = M.N;
  GradCoef<double> c = grad_coef<double>(M, i);
  const double4 ci = M.cell[i];
  T vi[5];
  pr(i, vi);
#pragma unroll
  for (int k = 0; k < 15; ++k) graw[k] = T(0.0);
  for (int s = M.sten_ptr[i]; s < M.sten_ptr[i + 1]; ++s) {
    const int j = M.sten_idx[s];
    const double4 cj = M.cell[j];
    double w[3];
    grad_weights<double>(c, cj.x - ci.x, cj.y - ci.y, cj.z - ci.z, w);
    T vj[5];
    pr(j, vj);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      T dVar = vj[k] - vi[k];
      graw[3 * k + 0] = graw[3 * k + 0] + T(w[0]) * dVar;
      graw[3 * k + 1] = graw[3 * k + 1] + T(w[1]) * dVar;
      graw[3 * k + 2] = graw[3 * k + 2] + T(w[2]) * dVar;
    }
  }
  if (CD.cptr) constrain_t<T>(CD, i, pr, vi, graw);
#pragma unroll
  for (int k = 0; k < 15; ++k) glim[k] = graw[k];
  if (!VENKAT) return;
  T vmn[5], vmx[5], lim[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) { vmn[k] = vi[k]; vmx[k] = vi[k]; lim[k] = T(1.0); }
  int nv = 0;
  for (int q = 0; q < M.vq; ++q) {
    int o = M.vnb[(size_t)q * N + i];
    if (o < 0) break;
    ++nv;
    T vo[5];
    pr(o, vo);
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
      T d2 = graw[3 * k] * T(dx) + graw[3 * k + 1] * T(dy) + graw[3 * k + 2] * T(dz);
      T dm = d2 > T(0.0) ? vmx[k] - vi[k] : vmn[k] - vi[k];
      T term = venkat_phi<T>(d2, dm, epsq);
      if (d2 < T(0.0) || d2 > T(0.0)) lim[k] = term < lim[k] ? term : lim[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) glim[3 * k + d] = lim[k] * graw[3 * k + d];
}

__device__ __forceinline__ void load_prim(const State& S, int N, int c, double* p) {
#pragma unroll
  for (int k = 0; k < 5; ++k) p[k] = S.prim[(size_t)k * N + c];
}

// ---- second-order tangent ------------------------------------------------------------
// dprim = dP(cv) dcv
__global__ void k_prim_jvp(Mesh M, State S, Gas G, const double* __restrict__ dcv,
                           double* __restrict__ dprim) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double q0[5];
  load_cv(S, N, i, q0);
  D1 q[5], pr[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k] = D1(q0[k], dcv[(size_t)k * N + i]);
  prim_t<D1>(G, q, pr);
#pragma unroll
  for (int k = 0; k < 5; ++k) dprim[(size_t)k * N + i] = pr[k].d;
}

// Limited gradient (value and tangent) of every cell: gv, gd [15][N].
template <bool VENKAT>
__global__ void k_grad2_jvp(Mesh M, ConstrD CD, State S, const double* __restrict__ dprim,
                            double* __restrict__ gv, double* __restrict__ gd) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  D1 graw[15], glim[15];
  grad_limited_t<D1, VENKAT>(M, CD, i, [&](int c, D1* v) {
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = D1(S.prim[(size_t)k * N + c], dprim[(size_t)k * N + c]);
  }, graw, glim);
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    gv[(size_t)k * N + i] = glim[k].v;
    gd[(size_t)k * N + i] = glim[k].d;
  }
}

__device__ __forceinline__ void load_side(const State& S, int N, int c, const double* dprim,
                                          const double* gv, const double* gd, D1* pr, D1* g) {
#pragma unroll
  for (int k = 0; k < 5; ++k) pr[k] = D1(S.prim[(size_t)k * N + c], dprim[(size_t)k * N + c]);
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = D1(gv[(size_t)k * N + c], gd[(size_t)k * N + c]);
}

template <int DISCR>
__global__ void k_face2_jvp(Mesh M, State S, Gas G, const double* __restrict__ dprim,
                            const double* __restrict__ gv, const double* __restrict__ gd,
                            double* __restrict__ dflux) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= M.F) return;
  const int N = M.N;
  const int2 cc = M.f2c[f];
  const double4 xf = M.fc[f], c1 = M.cell[cc.x], c2 = M.cell[cc.y];
  D1 pl[5], gl[15], pr[5], gr[15], L[5], R[5], irl, irr, flx[5];
  load_side(S, N, cc.x, dprim, gv, gd, pl, gl);
  load_side(S, N, cc.y, dprim, gv, gd, pr, gr);
  recon_t<D1>(pl, gl, xf.x - c1.x, xf.y - c1.y, xf.z - c1.z, L, irl);
  recon_t<D1>(pr, gr, xf.x - c2.x, xf.y - c2.y, xf.z - c2.z, R, irr);
  flux2_t<DISCR, D1>(G, L, irl, R, irr, M.fn[f], flx);
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) dflux[k * FB + f] = flx[k].d;
}

__global__ void k_bface2_jvp(Mesh M, State S, Gas G, Patches P, const double* __restrict__ dprim,
                             const double* __restrict__ gv, const double* __restrict__ gd,
                             const double* __restrict__ dvals, double* __restrict__ dflux) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= M.B) return;
  const int N = M.N;
  const int c = M.bf2c[b], p = M.bpatch[b];
  const double4 xf = M.bfc[b], cc = M.cell[c];
  D1 pr[5], g[15], vals[5], flx[5];
  load_side(S, N, c, dprim, gv, gd, pr, g);
#pragma unroll
  for (int k = 0; k < 5; ++k)
    vals[k] = D1(P.vals[5 * (size_t)p + k], dvals ? dvals[5 * (size_t)p + k] : 0.0);
  bflux2_t<D1>(G, P.bc[p], P.opt[p], P.opt2[p], P.order2[p], P.pout[p], vals, pr, g, xf.x - cc.x,
               xf.y - cc.y, xf.z - cc.z, M.bfn[b], flx);
  const size_t FB = (size_t)M.F + M.B;
#pragma unroll
  for (int k = 0; k < 5; ++k) dflux[k * FB + M.F + b] = flx[k].d;
}

// ---- second-order adjoint ------------------------------------------------------------
// Values of the limited gradient and of the raw one (the adjoint's forward sweep).
template <bool VENKAT>
__global__ void k_grad2_fwd(Mesh M, ConstrD CD, State S, double* __restrict__ graw,
                            double* __restrict__ glim) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double gr[15], gl[15];
  grad_limited_t<double, VENKAT>(M, CD, i, [&](int c, double* v) { load_prim(S, N, c, v); }, gr, gl);
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    graw[(size_t)k * N + i] = gr[k];
    glim[(size_t)k * N + i] = gl[k];
  }
}

// ---- WENO-XYZ and the positivity clip (RECONST 2) ------------------------------------
// Per component, g'_i = L_i / S_i with L_i = sum w_j g_j, S_i = sum w_j over the cell
// itself and its stencil members, w = 1/(g^2 + 1e-15) (k_weno_clip's order), then the
// clip zeroes the three components of rho or p when an extrapolation to one of the
// cell's faces is not positive.  The clip is a selection: its derivative keeps or
// zeroes those components as the values decide.  The adjoint of the blend is a gather
// over the cell itself and then the transposed stencil, with
// dg'_i/dg_j = (w_j + (g_j - g'_i) dw_j) / S_i, dw_j = -2 g_j w_j^2.

// Clip decisions of cell i (keep[0] rho, keep[1] p) from its blended gradient g.
template <typename T>
__device__ __forceinline__ void weno_clip_keep(const Mesh& M, const State& S, int i, const T* g,
                                               bool* keep) {
  const int N = M.N;
  const double4 ci = M.cell[i];
  const int pos[2] = {0, 4};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int iv = pos[q];
    const double varc = S.prim[(size_t)iv * N + i];
    keep[q] = true;
    for (int e = M.c2f_ptr[i]; e < M.c2f_ptr[i + 1]; ++e) {
      const int sgn = M.c2f_idx[e];
      const int f = (sgn > 0 ? sgn : -sgn) - 1;
      const double4 xf = f < M.F ? M.fc[f] : M.bfc[f - M.F];
      const double dx = xf.x - ci.x, dy = xf.y - ci.y, dz = xf.z - ci.z;
      const double varf = varc + dval(g[3 * iv]) * dx + dval(g[3 * iv + 1]) * dy +
                          dval(g[3 * iv + 2]) * dz;
      if (varf <= 0.0) {
        keep[q] = false;
        break;
      }
    }
  }
}

// Tangent: blended and clipped gradient values gv and tangents gd from the raw
// gradient's values graw and tangents draw.
__global__ void k_weno_jvp(Mesh M, State S, const double* __restrict__ graw,
                           const double* __restrict__ draw, double* __restrict__ gv,
                           double* __restrict__ gd) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const D1 eps(1.0E-15);
  const int s0 = M.sten_ptr[i], s1 = M.sten_ptr[i + 1];
  D1 g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    const D1 term(graw[(size_t)k * N + i], draw[(size_t)k * N + i]);
    D1 sumw = D1(1.0) / (term * term + eps);
    D1 loc = sumw * term;
    for (int s = s0; s < s1; ++s) {
      const int j = M.sten_idx[s];
      const D1 t(graw[(size_t)k * N + j], draw[(size_t)k * N + j]);
      const D1 w = D1(1.0) / (t * t + eps);
      sumw = sumw + w;
      loc = loc + w * t;
    }
    g[k] = loc / sumw;
  }
  bool keep[2];
  weno_clip_keep<D1>(M, S, i, g, keep);
#pragma unroll
  for (int q = 0; q < 2; ++q)
    if (!keep[q])
      for (int d = 0; d < 3; ++d) g[3 * (q ? 4 : 0) + d] = D1(0.0);
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    gv[(size_t)k * N + i] = g[k].v;
    gd[(size_t)k * N + i] = g[k].d;
  }
}

// Adjoint forward sweep: the clipped blend gfin (what the fluxes read), the blend before
// the clip gpre, the normalisers sw and the clip decisions keep[q*N + i].
__global__ void k_weno_fwd(Mesh M, State S, const double* __restrict__ graw,
                           double* __restrict__ gfin, double* __restrict__ gpre,
                           double* __restrict__ sw, int* __restrict__ keepo) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double eps = 1.0E-15;
  const int s0 = M.sten_ptr[i], s1 = M.sten_ptr[i + 1];
  double g[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    const double term = graw[(size_t)k * N + i];
    double sumw = 1.0 / (term * term + eps);
    double loc = sumw * term;
    for (int s = s0; s < s1; ++s) {
      const double t = graw[(size_t)k * N + M.sten_idx[s]];
      const double w = 1.0 / (t * t + eps);
      sumw = sumw + w;
      loc = loc + w * t;
    }
    g[k] = loc / sumw;
    gpre[(size_t)k * N + i] = g[k];
    sw[(size_t)k * N + i] = sumw;
  }
  bool keep[2];
  weno_clip_keep<double>(M, S, i, g, keep);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    keepo[(size_t)q * N + i] = keep[q] ? 1 : 0;
    if (!keep[q])
      for (int d = 0; d < 3; ++d) g[3 * (q ? 4 : 0) + d] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) gfin[(size_t)k * N + i] = g[k];
}

// Adjoint: aR_j = sum over i in {j} + T(j) of dg'_i/dg_j keep_i aG_i, per component.
__global__ void k_weno_vjp(Mesh M, const int* __restrict__ tptr, const int* __restrict__ tidx,
                           const double* __restrict__ graw, const double* __restrict__ gpre,
                           const double* __restrict__ sw, const int* __restrict__ keep,
                           const double* __restrict__ aG, double* __restrict__ aR) {
  const int N = M.N;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double eps = 1.0E-15;
  const int t0 = tptr[j], t1 = tptr[j + 1];
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    const int v = k / 3;
    const int q = v == 0 ? 0 : v == 4 ? 1 : -1;
    const double g = graw[(size_t)k * N + j];
    const double w = 1.0 / (g * g + eps);
    const double dw = -2.0 * g * w * w;
    double acc = 0.0;
    for (int t = t0 - 1; t < t1; ++t) {
      const int i = t < t0 ? j : tidx[t];
      if (q >= 0 && !keep[(size_t)q * N + i]) continue;
      acc = acc + (w + (g - gpre[(size_t)k * N + i]) * dw) / sw[(size_t)k * N + i] *
                      aG[(size_t)k * N + i];
    }
    aR[(size_t)k * N + j] = acc;
  }
}

// Interior face f, w = mu(c1) - mu(c2): gF[j][f] for j = 0..19, side s = j / 10 (c1, c2):
// j%10 < 5 the adjoint of the side's primitive state through the whole face state,
// j%10 >= 5 that of its reconstructed value (times dx it is the gradient adjoint).
template <int DISCR>
__global__ void k_face2_vjp(Mesh M, State S, Gas G, const double* __restrict__ glim,
                            const double* __restrict__ mu, double* __restrict__ gF) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= M.F) return;
  const int N = M.N;
  const size_t F = (size_t)M.F;
  const int2 cc = M.f2c[f];
  const double4 xf = M.fc[f], c1 = M.cell[cc.x], c2 = M.cell[cc.y], n = M.fn[f];
  double p1[5], p2[5], g1[15], g2[15], w[5];
  load_prim(S, N, cc.x, p1);
  load_prim(S, N, cc.y, p2);
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    g1[k] = glim[(size_t)k * N + cc.x];
    g2[k] = glim[(size_t)k * N + cc.y];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) w[k] = mu[(size_t)k * N + cc.x] - mu[(size_t)k * N + cc.y];
  for (int j = 0; j < 20; ++j) {
    const int side = j / 10, r = j % 10;
    D1 pl[5], pr[5], gl[15], gr[15], L[5], R[5], irl, irr, flx[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      pl[k] = D1(p1[k], side == 0 && r == k ? 1.0 : 0.0);
      pr[k] = D1(p2[k], side == 1 && r == k ? 1.0 : 0.0);
    }
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      gl[k] = D1(g1[k]);
      gr[k] = D1(g2[k]);
    }
    recon_t<D1>(pl, gl, xf.x - c1.x, xf.y - c1.y, xf.z - c1.z, L, irl);
    recon_t<D1>(pr, gr, xf.x - c2.x, xf.y - c2.y, xf.z - c2.z, R, irr);
    if (r >= 5) {
      if (side == 0) L[r - 5].d = 1.0;
      else R[r - 5].d = 1.0;
    }
    flux2_t<DISCR, D1>(G, L, irl, R, irr, n, flx);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) s = s + w[k] * flx[k].d;
    gF[(size_t)j * F + f] = s;
  }
}

// Boundary face b of cell c: gB[j][b] (j < 5 primitive state, 5..9 reconstructed values
// of a second-order patch) and, for an inflow patch, gV[j][b] (the patch values).
__global__ void k_bface2_vjp(Mesh M, State S, Gas G, Patches P, const double* __restrict__ glim,
                             const double* __restrict__ mu, double* __restrict__ gB,
                             double* __restrict__ gV) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= M.B) return;
  const int N = M.N;
  const size_t B = (size_t)M.B;
  const int c = M.bf2c[b], p = M.bpatch[b];
  const int bc = P.bc[p], o2 = P.order2[p];
  const double4 xf = M.bfc[b], cc = M.cell[c], n = M.bfn[b];
  double p0[5], g0[15], w[5];
  load_prim(S, N, c, p0);
#pragma unroll
  for (int k = 0; k < 15; ++k) g0[k] = glim[(size_t)k * N + c];
#pragma unroll
  for (int k = 0; k < 5; ++k) w[k] = mu[(size_t)k * N + c];
  for (int j = 0; j < 15; ++j) {
    const bool vseed = j >= 10;
    if ((j >= 5 && j < 10 && !o2) || (vseed && !has_value_inputs(bc))) {
      if (j < 10) gB[(size_t)j * B + b] = 0.0;
      else gV[(size_t)(j - 10) * B + b] = 0.0;
      continue;
    }
    D1 pr[5], g[15], vals[5], flx[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      pr[k] = D1(p0[k], j == k ? 1.0 : 0.0);
      vals[k] = D1(P.vals[5 * (size_t)p + k], j == k + 10 ? 1.0 : 0.0);
    }
#pragma unroll
    for (int k = 0; k < 15; ++k) g[k] = D1(g0[k]);
    if (j >= 5 && j < 10) {
      // seed the reconstructed value: d(face value)/d(g . dx) = 1 through g's component
      // along the face offset, scaled back below
      const int v = j - 5;
      const double d[3] = {xf.x - cc.x, xf.y - cc.y, xf.z - cc.z};
      const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
#pragma unroll
      for (int a = 0; a < 3; ++a) g[3 * v + a].d = d[a] / dd;
    }
    bflux2_t<D1>(G, bc, P.opt[p], P.opt2[p], o2, P.pout[p], vals, pr, g, xf.x - cc.x, xf.y - cc.y,
                 xf.z - cc.z, n, flx);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) s = s + w[k] * flx[k].d;
    if (j < 10) gB[(size_t)j * B + b] = s;
    else gV[(size_t)(j - 10) * B + b] = s;
  }
}

// Per cell, over its faces in cell-to-face order: the primitive-state adjoint aP [5][N]
// and the limited-gradient adjoint aG [15][N] (reconstructed-value adjoint times the
// face offset xf - cofg).
__global__ void k_cell2_vjp(Mesh M, const double* __restrict__ gF, const double* __restrict__ gB,
                            double* __restrict__ aP, double* __restrict__ aG) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const size_t F = (size_t)M.F, B = (size_t)M.B;
  const double4 ci = M.cell[i];
  double ap[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, ag[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) ag[k] = 0.0;
  for (int e = M.c2f_ptr[i]; e < M.c2f_ptr[i + 1]; ++e) {
    const int s = M.c2f_idx[e];
    const double* base;
    size_t stride, idx;
    double4 xf;
    if (s > M.F) {
      idx = (size_t)(s - 1 - M.F);
      base = gB;
      stride = B;
      xf = M.bfc[idx];
    } else {
      const size_t f = (size_t)((s > 0 ? s : -s) - 1);
      idx = f;
      base = gF + (s > 0 ? 0 : 10 * F);
      stride = F;
      xf = M.fc[f];
    }
    const double d[3] = {xf.x - ci.x, xf.y - ci.y, xf.z - ci.z};
#pragma unroll
    for (int k = 0; k < 5; ++k) ap[k] = ap[k] + base[(size_t)k * stride + idx];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double aq = base[(size_t)(5 + k) * stride + idx];
#pragma unroll
      for (int a = 0; a < 3; ++a) ag[3 * k + a] = ag[3 * k + a] + aq * d[a];
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) aP[(size_t)k * N + i] = ap[k];
#pragma unroll
  for (int k = 0; k < 15; ++k) aG[(size_t)k * N + i] = ag[k];
}

// Limiter adjoint of cell i: aG (limited) -> aR (raw gradient) and the primitive-state
// adjoint of the limiter's min/max: -a_dm to the cell itself (added to aP) and +a_dm to
// the neighbour that set the min/max (selN/selV [5][N], -1 when the cell itself did).
__global__ void k_venkat_vjp(Mesh M, State S, const double* __restrict__ graw,
                             const double* __restrict__ aG, double* __restrict__ aR,
                             double* __restrict__ aP, int* __restrict__ selN,
                             double* __restrict__ selV) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double vi[5], g[15];
  load_prim(S, N, i, vi);
#pragma unroll
  for (int k = 0; k < 15; ++k) g[k] = graw[(size_t)k * N + i];
  double vmn[5], vmx[5];
  int amn[5], amx[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) { vmn[k] = vi[k]; vmx[k] = vi[k]; amn[k] = -1; amx[k] = -1; }
  int nv = 0;
  for (int q = 0; q < M.vq; ++q) {
    int o = M.vnb[(size_t)q * N + i];
    if (o < 0) break;
    ++nv;
    double vo[5];
    load_prim(S, N, o, vo);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (!(vmn[k] < vo[k])) { vmn[k] = vo[k]; amn[k] = o; }
      if (!(vmx[k] > vo[k])) { vmx[k] = vo[k]; amx[k] = o; }
    }
  }
  const double epsq = M.epsq[i];
  for (int k = 0; k < 5; ++k) {
    double lim = 1.0;
    int qs = -1;
    for (int q = 0; q < nv; ++q) {
      double dx = M.vgeo[(size_t)(3 * q) * N + i];
      double dy = M.vgeo[(size_t)(3 * q + 1) * N + i];
      double dz = M.vgeo[(size_t)(3 * q + 2) * N + i];
      double d2 = g[3 * k] * dx + g[3 * k + 1] * dy + g[3 * k + 2] * dz;
      double dm = d2 > 0.0 ? vmx[k] - vi[k] : vmn[k] - vi[k];
      double term = venkat_phi<double>(d2, dm, epsq);
      if (d2 < 0.0 || d2 > 0.0) {
        if (term < lim) { lim = term; qs = q; }
      }
    }
    double sadj = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) sadj = sadj + aG[(size_t)(3 * k + a) * N + i] * g[3 * k + a];
    double ar[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) ar[a] = lim * aG[(size_t)(3 * k + a) * N + i];
    int sn = -1;
    double sv = 0.0;
    if (qs >= 0) {
      double dx[3] = {M.vgeo[(size_t)(3 * qs) * N + i], M.vgeo[(size_t)(3 * qs + 1) * N + i],
                      M.vgeo[(size_t)(3 * qs + 2) * N + i]};
      double d2 = g[3 * k] * dx[0] + g[3 * k + 1] * dx[1] + g[3 * k + 2] * dx[2];
      const bool up = d2 > 0.0;
      double dm = up ? vmx[k] - vi[k] : vmn[k] - vi[k];
      double pd2 = venkat_phi<D1>(D1(d2, 1.0), D1(dm), epsq).d;
      double pdm = venkat_phi<D1>(D1(d2), D1(dm, 1.0), epsq).d;
      double a_d2 = sadj * pd2, a_dm = sadj * pdm;
#pragma unroll
      for (int a = 0; a < 3; ++a) ar[a] = ar[a] + a_d2 * dx[a];
      const int from = up ? amx[k] : amn[k];
      if (from >= 0) {
        aP[(size_t)k * N + i] = aP[(size_t)k * N + i] - a_dm;
        sn = from;
        sv = a_dm;
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) aR[(size_t)(3 * k + a) * N + i] = ar[a];
    selN[(size_t)k * N + i] = sn;
    selV[(size_t)k * N + i] = sv;
  }
}

// aP(j) += the limiter adjoints of the face neighbours whose min/max j set (the
// face-neighbour relation is symmetric, so j's own list finds them).
__global__ void k_venkat_nb_gather(Mesh M, const int* __restrict__ selN,
                                   const double* __restrict__ selV, double* __restrict__ aP) {
  const int N = M.N;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < M.vq; ++q) {
    int i = M.vnb[(size_t)q * N + j];
    if (i < 0) break;
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if (selN[(size_t)k * N + i] == j) acc[k] = acc[k] + selV[(size_t)k * N + i];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) aP[(size_t)k * N + j] = aP[(size_t)k * N + j] + acc[k];
}

// Constrained-gradient adjoint, before k_grad_vjp: aP(v, j) += sum over the rows r
// that read cell j (entry (i, v)) of rw(:, r) . aR(3v..3v+2, i), minus, for j's own
// entries, sum over all their rows of rw(:, r) . aR(3v..3v+2, j) (each row subtracts
// the cell's own value).  k_constr_clear then zeroes the constrained components of aR,
// whose least-squares values the constraints replaced.
__global__ void k_constr_vjp(Mesh M, ConstrD CD, const double* __restrict__ aR,
                             double* __restrict__ aP) {
  const int N = M.N;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const Constr& C = CD.C;
  const size_t R = C.R;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int q = CD.trptr[j]; q < CD.trptr[j + 1]; ++q) {
    const int r = CD.trow[q], e = CD.tent[q];
    const int i = C.ecell[e], v = C.evar[e];
    const double s = C.rw[r] * aR[(size_t)(3 * v) * N + i] +
                     C.rw[R + r] * aR[(size_t)(3 * v + 1) * N + i] +
                     C.rw[2 * R + r] * aR[(size_t)(3 * v + 2) * N + i];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (k == v - 1) acc[k] = acc[k] + s;
  }
  for (int q = CD.cptr[j]; q < CD.cptr[j + 1]; ++q) {
    const int e = CD.cent[q], v = C.evar[e];
    const double a0 = aR[(size_t)(3 * v) * N + j], a1 = aR[(size_t)(3 * v + 1) * N + j],
                 a2 = aR[(size_t)(3 * v + 2) * N + j];
    double s = 0.0;
    for (int r = C.rptr[e]; r < C.rptr[e + 1]; ++r)
      s = s + (C.rw[r] * a0 + C.rw[R + r] * a1 + C.rw[2 * R + r] * a2);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (k == v - 1) acc[k] = acc[k] - s;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) aP[(size_t)(k + 1) * N + j] = aP[(size_t)(k + 1) * N + j] + acc[k];
}

__global__ void k_constr_clear(Mesh M, Constr C, double* __restrict__ aR) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= C.n) return;
  const int N = M.N, i = C.ecell[e], v = C.evar[e];
#pragma unroll
  for (int d = 0; d < 3; ++d) aR[(size_t)(3 * v + d) * N + i] = 0.0;
}

// Gradient adjoint: aP(j) += sum over the cells i whose stencil holds j (transposed
// stencil tptr/tidx) of w_ij . aR(i), minus sum over j's own members of w_js . aR(j).
__global__ void k_grad_vjp(Mesh M, const int* __restrict__ tptr, const int* __restrict__ tidx,
                           const double* __restrict__ aR, double* __restrict__ aP) {
  const int N = M.N;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double4 cj = M.cell[j];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = tptr[j]; e < tptr[j + 1]; ++e) {
    const int i = tidx[e];
    const double4 ci = M.cell[i];
    GradCoef<double> c = grad_coef<double>(M, i);
    double w[3];
    grad_weights<double>(c, cj.x - ci.x, cj.y - ci.y, cj.z - ci.z, w);
#pragma unroll
    for (int k = 0; k < 5; ++k)
      acc[k] = acc[k] + (w[0] * aR[(size_t)(3 * k) * N + i] + w[1] * aR[(size_t)(3 * k + 1) * N + i] +
                         w[2] * aR[(size_t)(3 * k + 2) * N + i]);
  }
  GradCoef<double> c = grad_coef<double>(M, j);
  double ws[3] = {0.0, 0.0, 0.0};
  for (int s = M.sten_ptr[j]; s < M.sten_ptr[j + 1]; ++s) {
    const double4 cs = M.cell[M.sten_idx[s]];
    double w[3];
    grad_weights<double>(c, cs.x - cj.x, cs.y - cj.y, cs.z - cj.z, w);
    ws[0] = ws[0] + w[0];
    ws[1] = ws[1] + w[1];
    ws[2] = ws[2] + w[2];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k)
    aP[(size_t)k * N + j] = aP[(size_t)k * N + j] + acc[k] -
                            (ws[0] * aR[(size_t)(3 * k) * N + j] + ws[1] * aR[(size_t)(3 * k + 1) * N + j] +
                             ws[2] * aR[(size_t)(3 * k + 2) * N + j]);
}

// lam = dP(cv)^T aP, per cell.
__global__ void k_prim_vjp(Mesh M, State S, Gas G, const double* __restrict__ aP,
                           double* __restrict__ lam) {
  const int N = M.N;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double q0[5], a[5];
  load_cv(S, N, i, q0);
#pragma unroll
  for (int k = 0; k < 5; ++k) a[k] = aP[(size_t)k * N + i];
  for (int j = 0; j < 5; ++j) {
    D1 q[5], pr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k] = D1(q0[k], j == k ? 1.0 : 0.0);
    prim_t<D1>(G, q, pr);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) s = s + a[k] * pr[k].d;
    lam[(size_t)j * N + i] = s;
  }
}

}  // namespace rflu
