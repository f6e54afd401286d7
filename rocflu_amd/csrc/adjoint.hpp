// Tangent (forward) and adjoint (reverse) derivatives of the first-order residual
// R(cv) = sum over a cell's faces of +-F, for verification of the explicit path's
// discrete adjoint.
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
__device__ __forceinline__ void bflux1_t(const Gas& G, int bc, int opt, double pout, const T* vals,
                                         const T* q, double4 n, T* flx) {
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
  T bs[5], pr;
  if (bc == BC_INFLOW_VELTEMP) {
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
  bflux1_t<D1>(G, P.bc[p], P.opt[p], P.pout[p], vals, q, M.bfn[b], flx);
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
  const int nseed = bc == BC_INFLOW_VELTEMP ? 10 : 5;
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
    bflux1_t<D1>(G, bc, P.opt[p], P.pout[p], vals, q, n, flx);
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

}  // namespace rflu
