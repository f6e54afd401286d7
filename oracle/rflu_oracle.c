/*
 * rflu_oracle.c -- CPU restatement of RocfluMP's explicit residual/update path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the GPU path; see rflu_oracle.h).
 *
 * FP64 throughout, built with -ffp-contract=off so every expression rounds as the
 * reference's gfortran build does (no FMA contraction, CMakeLists.txt:44,
 * machines/dane.sh:6).  Loops follow the reference's order: residuals are scattered
 * face by face, interior faces first and then patch by patch, which fixes the
 * per-cell summation order the GPU gather reproduces.  Expressions keep the
 * reference's operand order and parenthesisation.
 *
 * Threads (orc_set_threads, for timing the CPU path on several cores): per-cell loops
 * run under OpenMP, and the three face scatters (residual, Venkatakrishnan min/max and
 * limiter, spectral-radius sums) become per-cell gathers over each cell's faces in
 * ascending face order (interior faces, then boundary faces in patch order), which adds
 * the same terms in the same order as the scatter; min and max do not depend on order.
 * With one thread (the default) every loop is the sequential restatement.
 */
#include "rflu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void orc_set_threads(int n) {
#ifdef _OPENMP
  g_threads = n < 1 ? 1 : n;
#else
  (void)n;
  g_threads = 1;
#endif
}

int orc_get_threads(void) { return g_threads; }

#define ORC_PAR _Pragma("omp parallel for schedule(static) num_threads(g_threads) if(g_threads > 1)")

#define BC_INFLOW_TOTANG 10
#define BC_INFLOW_VELTEMP 11
#define BC_OUTFLOW 20
#define BC_SLIPWALL 60
#define BC_FARFIELD 80
#define BC_INJECTION 90
#define BCOPT_SUPERSONIC 0
#define BCOPT_SUBSONIC 1
#define BCOPT_MIXED 2

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* Gas constants of the dependent-variable path: gv(CP) = refCp and gv(MOL) = M(R)
 * (libflu/RFLU_SetGasVars.F90:172-178), read back as R' = 8314.3/M and
 * gamma' = cp/(cp - R') (libfloflu/MixtPerf_R.F90:93, MixtPerf_G.F90:78). */
static void gas_prime(const orc_region* r, double* mm, double* rgas, double* gam) {
  double gc0 = r->cp - r->cp / r->gamma;     /* MixtPerf_R_CpG */
  double m = 8314.3 / gc0;                   /* MixtPerf_M_R  */
  double gc = 8314.3 / m;                    /* MixtPerf_R_M  */
  *mm = m;
  *rgas = gc;
  *gam = r->cp / (r->cp - gc);               /* MixtPerf_G_CpR */
}

/* libflu/RFLU_SetDependentVars.F90:255-284 (TCPERF, explicit solver). */
void orc_set_dependent_vars(const orc_region* r, const double* cv, double* dv, int beg, int end) {
  double mm, gc, g;
  gas_prime(r, &mm, &gc, &g);
#pragma omp parallel for schedule(static) num_threads(g_threads) if(g_threads > 1 && end - beg > 4096)
  for (int icg = beg; icg < end; ++icg) {
    const double* c = cv + 5 * (size_t)icg;
    double rr = c[0];
    double ir = 1.0 / rr;
    double Eo = ir * c[4];
    double Vm2 = ir * ir * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]);
    double p = (g - 1.0) * rr * (Eo - 0.5 * Vm2);  /* MixtPerf_P_DEoGVm2 */
    double t = p / (rr * gc);                       /* MixtPerf_T_DPR */
    dv[3 * (size_t)icg + 0] = p;
    dv[3 * (size_t)icg + 1] = t;
    dv[3 * (size_t)icg + 2] = sqrt(g * gc * t);     /* MixtPerf_C_GRT */
  }
}

/* ---------------------------------------------------------------------------------
 * Face flux functions.  Inputs are the left/right states the reference computes in
 * its face loop; outputs flx(1:5) (already multiplied by |A|).
 * ------------------------------------------------------------------------------- */

/* HLLC (Batten): modflu/RFLU_ModHLLCFlux.F90:1064-1117 (order 1), 2310-2363 (order 2). */
static void hllc_core(double g, double nx, double ny, double nz, double nm, double fs,
                      double rl, double ul, double vl, double wl, double pl, double al,
                      double el, double Hl, double rr, double ur, double vr, double wr,
                      double pr, double ar, double er, double Hr, double* flx) {
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double qr = ur * nx + vr * ny + wr * nz - fs;
  double wt1 = sqrt(rr / rl);
  double wt2 = 1.0 / (1.0 + wt1);
  double us = wt2 * (ul + wt1 * ur);
  double vs = wt2 * (vl + wt1 * vr);
  double ws = wt2 * (wl + wt1 * wr);
  double qs = wt2 * (ql + wt1 * qr);
  double Hs = wt2 * (Hl + wt1 * Hr);
  double Vms = us * us + vs * vs + ws * ws;
  double as = sqrt((g - 1.0) * (Hs - 0.5 * Vms)); /* MixtPerf_C_GHoVm2 */
  double sl = dmin(ql - al, qs - as);
  double sr = dmax(qr + ar, qs + as);
  if (sl > 0.0) {
    flx[0] = ql * rl * nm;
    flx[1] = (ql * rl * ul + pl * nx) * nm;
    flx[2] = (ql * rl * vl + pl * ny) * nm;
    flx[3] = (ql * rl * wl + pl * nz) * nm;
    flx[4] = (ql * (rl * el + pl) + pl * fs) * nm;
  } else if (sr < 0.0) {
    flx[0] = qr * rr * nm;
    flx[1] = (qr * rr * ur + pr * nx) * nm;
    flx[2] = (qr * rr * vr + pr * ny) * nm;
    flx[3] = (qr * rr * wr + pr * nz) * nm;
    flx[4] = (qr * (rr * er + pr) + pr * fs) * nm;
  } else {
    double sm = (rr * qr * (sr - qr) - rl * ql * (sl - ql) - pr + pl) /
                (rr * (sr - qr) - rl * (sl - ql));
    double ps = pl + rl * (ql - sl) * (ql - sm);
    double term, rs, rus, rvs, rws, res;
    if (sm > 0.0) {
      term = 1.0 / (sl - sm);
      rs = term * (sl - ql) * rl;
      rus = term * ((sl - ql) * rl * ul + (ps - pl) * nx);
      rvs = term * ((sl - ql) * rl * vl + (ps - pl) * ny);
      rws = term * ((sl - ql) * rl * wl + (ps - pl) * nz);
      res = term * ((sl - ql) * rl * el + (ps * sm - pl * ql));
    } else {
      term = 1.0 / (sr - sm);
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
    flx[4] = (sm * (res + ps) - ps * fs) * nm;
  }
}

/* Harten-Hyman entropy fix, modflu/RFLU_ModEntropyFixes.F90:95-106. */
static inline double ef_hh(double l, double d) { return l > d ? l : (l * l + d * d) / (2.0 * d); }

/* Roe FDS: modflu/RFLU_ModRoeFlux.F90:1838-1894 (order 1), 2762-2818 (order 2). */
static void roe_core(double g, double epsentr, double dissFact, double nx, double ny, double nz,
                     double nm, double fs, double rl, double ul, double vl, double wl, double pl,
                     double Hl, double rr, double ur, double vr, double wr, double pr, double Hr,
                     double* flx) {
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double qr = ur * nx + vr * ny + wr * nz - fs;
  double rh = sqrt(rl * rr);
  double wt = rl / (rl + rh);
  double uh = wt * ul + (1.0 - wt) * ur;
  double vh = wt * vl + (1.0 - wt) * vr;
  double wh = wt * wl + (1.0 - wt) * wr;
  double Hh = wt * Hl + (1.0 - wt) * Hr;
  double qh = uh * nx + vh * ny + wh * nz - fs;
  double sh = 0.5 * (uh * uh + vh * vh + wh * wh);
  double ah = sqrt((g - 1.0) * (Hh - 0.5 * (2.0 * sh)));
  double l1 = fabs(qh - ah);
  double l2 = fabs(qh);
  double l5 = fabs(qh + ah);
  double efc = epsentr * l5;
  l1 = ef_hh(l1, efc);
  l2 = ef_hh(l2, efc);
  l5 = ef_hh(l5, efc);
  double dr = rr - rl;
  double du = ur - ul;
  double de = vr - vl;
  double dw = wr - wl;
  double dp = pr - pl;
  double dq = du * nx + de * ny + dw * nz;
  double dv1 = (dp - rh * ah * dq) / (2.0 * ah * ah);
  double dv5 = (dp + rh * ah * dq) / (2.0 * ah * ah);
  double dv2 = dr - dp / (ah * ah);
  double t1 = l1 * dv1;
  double t2 = l2 * dv2;
  double t3 = l2 * rh;
  double t5 = l5 * dv5;
  double fc1 = 0.5 * (ql * rl + qr * rr) * nm;
  double fc2 = 0.5 * (ql * rl * ul + pl * nx + qr * rr * ur + pr * nx) * nm;
  double fc3 = 0.5 * (ql * rl * vl + pl * ny + qr * rr * vr + pr * ny) * nm;
  double fc4 = 0.5 * (ql * rl * wl + pl * nz + qr * rr * wr + pr * nz) * nm;
  double fc5 = 0.5 * (ql * rl * Hl + pl * fs + qr * rr * Hr + pr * fs) * nm;
  double fd1 = 0.5 * dissFact * (t1 + t2 + t5) * nm;
  double fd2 = 0.5 * dissFact * (t1 * (uh - nx * ah) + t2 * uh + t3 * (du - nx * dq) + t5 * (uh + nx * ah)) * nm;
  double fd3 = 0.5 * dissFact * (t1 * (vh - ny * ah) + t2 * vh + t3 * (de - ny * dq) + t5 * (vh + ny * ah)) * nm;
  double fd4 = 0.5 * dissFact * (t1 * (wh - nz * ah) + t2 * wh + t3 * (dw - nz * dq) + t5 * (wh + nz * ah)) * nm;
  double fd5 = 0.5 * dissFact * (t1 * (Hh - qh * ah) + t2 * sh + t3 * (uh * du + vh * de + wh * dw - qh * dq) + t5 * (Hh + qh * ah)) * nm;
  flx[0] = fc1 - fd1;
  flx[1] = fc2 - fd2;
  flx[2] = fc3 - fd3;
  flx[3] = fc4 - fd4;
  flx[4] = fc5 - fd5;
}

/* AUSM+ flux function, modflu/RFLU_ModAUSMFlux.F90:282-367. */
static void ausm_core(double nx, double ny, double nz, double nm, double fs, double rl, double ul,
                      double vl, double wl, double pl, double Hl, double al, double rr, double ur,
                      double vr, double wr, double pr, double Hr, double ar, double* flx) {
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double qr = ur * nx + vr * ny + wr * nz - fs;
  double af = 0.5 * (al + ar);
  double ml = ql / af, mla = fabs(ml);
  double mr = qr / af, mra = fabs(mr);
  double mlp, wtl, mrm, wtr;
  if (mla <= 1.0) {
    mlp = 0.25 * (ml + 1.0) * (ml + 1.0) + 0.125 * (ml * ml - 1.0) * (ml * ml - 1.0);
    wtl = 0.25 * (ml + 1.0) * (ml + 1.0) * (2.0 - ml) + 0.1875 * ml * (ml * ml - 1.0) * (ml * ml - 1.0);
  } else {
    mlp = 0.5 * (ml + mla);
    wtl = 0.5 * (1.0 + ml / mla);
  }
  if (mra <= 1.0) {
    mrm = -0.25 * (mr - 1.0) * (mr - 1.0) - 0.125 * (mr * mr - 1.0) * (mr * mr - 1.0);
    wtr = 0.25 * (mr - 1.0) * (mr - 1.0) * (2.0 + mr) - 0.1875 * mr * (mr * mr - 1.0) * (mr * mr - 1.0);
  } else {
    mrm = 0.5 * (mr - mra);
    wtr = 0.5 * (1.0 - mr / mra);
  }
  double mf = mlp + mrm;
  double mfa = fabs(mf);
  double mfp = 0.5 * (mf + mfa);
  double mfm = 0.5 * (mf - mfa);
  double pf = wtl * pl + wtr * pr;
  flx[0] = (af * (mfp * rl + mfm * rr)) * nm;
  flx[1] = (af * (mfp * rl * ul + mfm * rr * ur) + pf * nx) * nm;
  flx[2] = (af * (mfp * rl * vl + mfm * rr * vr) + pf * ny) * nm;
  flx[3] = (af * (mfp * rl * wl + mfm * rr * wr) + pf * nz) * nm;
  flx[4] = (af * (mfp * rl * Hl + mfm * rr * Hr) + pf * fs) * nm;
}

/* Outflow boundary state, libfloflu/BcondOutflowPerf.F90:111-157. */
static void bcond_outflow_perf(int bcOpt, double pout, double sxn, double syn, double szn,
                               double cpgas, double mol, double rho, double rhou, double rhov,
                               double rhow, double rhoe, double press, double* rhob,
                               double* rhoub, double* rhovb, double* rhowb, double* rhoeb) {
  double rgas = 8314.3 / mol;
  double gamma = cpgas / (cpgas - rgas);
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
      ub = copysign(1.0, u) * dmax(fabs(ub), fabs(u));
      vb = copysign(1.0, v) * dmax(fabs(vb), fabs(v));
      wb = copysign(1.0, w) * dmax(fabs(wb), fabs(w));
    }
    *rhob = rb;
    *rhoub = rb * ub;
    *rhovb = rb * vb;
    *rhowb = rb * wb;
    *rhoeb = rb * (pout / (rb * (gamma - 1.0)) + 0.5 * (ub * ub + vb * vb + wb * wb));
  } else {
    *rhob = rho; *rhoub = rhou; *rhovb = rhov; *rhowb = rhow; *rhoeb = rhoe;
  }
}

/* Inflow boundary state with prescribed velocity and temperature (BC_INFLOW_VELTEMP,
 * rocflu/RFLU_CentralFirstPatch.F90:375-414, RFLU_CentralSecondPatch.F90:473-495):
 * subsonic takes the interior density, otherwise the prescribed pressure fixes it
 * (MixtPerf_D_PRT, MixtPerf_P_DRT, MixtPerf_Eo_DGPUVW). */
static void inflow_veltemp(int bcOpt, const double* vals, double rl, double rgas, double g,
                           double* rr, double* rur, double* rvr, double* rwr, double* rer,
                           double* pr) {
  double ur = vals[0], vr = vals[1], wr = vals[2], tr = vals[3];
  if (bcOpt != BCOPT_SUBSONIC) {
    *pr = vals[4];
    *rr = *pr / (rgas * tr);
  } else {
    *rr = rl;
    *pr = *rr * rgas * tr;
  }
  *rer = *rr * (*pr / (*rr * (g - 1.0)) + 0.5 * (ur * ur + vr * vr + wr * wr));
  *rur = *rr * ur;
  *rvr = *rr * vr;
  *rwr = *rr * wr;
}

/* Inflow boundary state from total pressure, total temperature and flow angles
 * (BC_INFLOW_TOTANG), libfloflu/BcondInflowPerf.F90:95-223: subsonic faces keep the
 * interior Riemann invariant Rm = a - (g-1)/2 q.n and solve the total-enthalpy
 * quadratic for the boundary sound speed; supersonic (and mixed faces with
 * |q.n| >= a) take the state of the given Mach number.  A subsonic-option face with
 * |q.n| >= a leaves the reference's outputs unset; here they are NaN, which the
 * validity check then reports. */
static void bcond_inflow_perf(int bcOptType, int bcOptFixed, double ptot, double ttot,
                              double betah, double betav, double mach, double sxn, double syn,
                              double szn, double cpgas, double mm, double rl, double rul,
                              double rvl, double rwl, double* rr, double* rur, double* rvr,
                              double* rwr, double* rer, double* pr) {
  double rgas = 8314.3 / mm;                   /* MixtPerf_R_M */
  double g = cpgas / (cpgas - rgas);           /* MixtPerf_G_CpR */
  double tr, ar, ur, vr, wr;
  if (bcOptType == BCOPT_SUBSONIC || bcOptType == BCOPT_MIXED) {
    double gm1 = g - 1.0;
    double ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
    double a02 = g * rgas * ttot;                                    /* MixtPerf_C2_GRT */
    double al = sqrt(a02 - 0.5 * (g - 1.0) * (ul * ul + vl * vl + wl * wl)); /* C_Co2GUVW */
    double ql = ul * sxn + vl * syn + wl * szn;
    if (fabs(ql) < al) {
      double sl = sqrt(ul * ul + vl * vl + wl * wl);
      double eta;
      if (bcOptFixed == 0)
        eta = sl > 1.0E-6 ? ql / sl : -1.0;
      else
        eta = -1.0;
      double Rm = al - 0.5 * gm1 * ql;
      double disc = 0.5 * gm1 * (eta * eta) *
                    (a02 / (Rm * Rm) * (1.0 + 0.5 * gm1 * (eta * eta)) - 1.0);
      double sr;
      if (disc < 0.0) {
        ar = sqrt(a02);
        tr = ttot;
        *pr = ptot;
        sr = 0.0;
      } else {
        ar = Rm / (1.0 + 0.5 * gm1 * eta * eta) * (1.0 + sqrt(disc));
        tr = ar * ar / (g * rgas);                                   /* MixtPerf_T_CGR */
        *pr = ptot * pow(tr / ttot, g / (g - 1.0));                  /* MixtPerf_P_GPoTTo */
        double c2 = ar * ar;                                         /* MixtPerf_Vm_C2Co2G */
        sr = a02 > c2 ? sqrt(2.0 / (g - 1.0) * (a02 - c2)) : 0.0;
      }
      *rr = *pr / (rgas * tr);                                       /* MixtPerf_D_PRT */
      ur = sr * cos(betah) * cos(betav);
      vr = sr * sin(betah);
      wr = sr * cos(betah) * sin(betav);
      *rer = *rr * (*pr / (*rr * (g - 1.0)) + 0.5 * sr * sr);        /* MixtPerf_Eo_DGPVm */
      *rur = *rr * ur;
      *rvr = *rr * vr;
      *rwr = *rr * wr;
      return;
    }
    if (bcOptType != BCOPT_MIXED) {
      *rr = *rur = *rvr = *rwr = *rer = *pr = NAN;
      return;
    }
  }
  *pr = ptot / pow(1.0 + 0.5 * (g - 1.0) * mach * mach, g / (g - 1.0));  /* MixtPerf_P_GMaPo */
  tr = ttot / (1.0 + 0.5 * (g - 1.0) * mach * mach);                    /* MixtPerf_T_GMaTo */
  *rr = *pr / (rgas * tr);
  ar = sqrt(g * rgas * tr);                                             /* MixtPerf_C_GRT */
  ur = mach * ar * cos(betah) * cos(betav);
  vr = mach * ar * sin(betah);
  wr = mach * ar * cos(betah) * sin(betav);
  *rer = *rr * (*pr / (*rr * (g - 1.0)) + 0.5 * (ur * ur + vr * vr + wr * wr));
  *rur = *rr * ur;
  *rvr = *rr * vr;
  *rwr = *rr * wr;
}

/* Farfield boundary state (BC_FARFIELD), modflu/RFLU_ModRindStates.F90:248-420 without
 * the point-vortex correction (BCOPT_CORR_NO): free stream from Mach number, angles of
 * attack and sideslip, pressure and temperature; subsonic faces blend it with the
 * interior state along the acoustic characteristics, supersonic faces take one or the
 * other by the sign of the interior normal velocity.  The interior pressure is formed
 * from rho E (MixtPerf_P_DEoGVm2). */
static void rind_farfield_perf(double cpgas, double mm, double nx, double ny, double nz,
                               double machInf, double pInf, double tInf, double alphaInf,
                               double betaInf, double rl, double rul, double rvl, double rwl,
                               double rel, double* rr, double* rur, double* rvr, double* rwr,
                               double* rer, double* pr) {
  double rgas = 8314.3 / mm;
  double g = cpgas / (cpgas - rgas);
  double ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
  double ql = ul * nx + vl * ny + wl * nz;
  double el = rel / rl;
  double sl2 = ul * ul + vl * vl + wl * wl;
  double pl = (g - 1.0) * rl * (el - 0.5 * sl2);
  double ri = pInf / (rgas * tInf);
  double pi = pInf;
  double qi = machInf * sqrt(g * rgas * tInf);
  double ui = qi * cos(alphaInf) * cos(betaInf);
  double vi = qi * sin(alphaInf) * cos(betaInf);
  double wi = qi * sin(betaInf);
  if (machInf < 1.0) {
    double al = sqrt(g * pl / rl);
    double pb, rb, ub, vb, wb;
    if (ql < 0.0) {
      pb = 0.5 * (pi + pl - rl * al * ((ui - ul) * nx + (vi - vl) * ny + (wi - wl) * nz));
      rb = ri - (pi - pb) / (al * al);
      ub = ui - nx * (pi - pb) / (rl * al);
      vb = vi - ny * (pi - pb) / (rl * al);
      wb = wi - nz * (pi - pb) / (rl * al);
    } else {
      pb = pi;
      rb = rl - (pl - pi) / (al * al);
      ub = ul + nx * (pl - pi) / (rl * al);
      vb = vl + ny * (pl - pi) / (rl * al);
      wb = wl + nz * (pl - pi) / (rl * al);
    }
    *rr = rb;
    *rur = rb * ub;
    *rvr = rb * vb;
    *rwr = rb * wb;
    *rer = rb * (pb / (rb * (g - 1.0)) + 0.5 * (ub * ub + vb * vb + wb * wb));
    *pr = pb;
  } else if (ql < 0.0) {
    *rr = ri;
    *rur = ri * ui;
    *rvr = ri * vi;
    *rwr = ri * wi;
    *rer = ri * (pi / (ri * (g - 1.0)) + 0.5 * qi * qi);
    *pr = pi;
  } else {
    *rr = rl; *rur = rul; *rvr = rvl; *rwr = rwl; *rer = rel;
    *pr = pl;
  }
}

/* Injected gas at a burning surface (BC_INJECTION, mass flux minj > 0),
 * modflu/RFLU_ModRindStates.F90:459-501: density from the interior pressure and the
 * injection temperature, velocity -minj/rho along the normal, total enthalpy. */
static void rind_inject_perf(double cpgas, double mm, double nx, double ny, double nz,
                             double minj, double tinj, double pl, double fs, double* rl,
                             double* ul, double* vl, double* wl, double* Hl) {
  double rgas = 8314.3 / mm;
  *rl = pl / (rgas * tinj);
  double ql = -minj / *rl + fs;
  *ul = ql * nx;
  *vl = ql * ny;
  *wl = ql * nz;
  *Hl = cpgas * tinj + 0.5 * (*ul * *ul + *vl * *vl + *wl * *wl);   /* MixtPerf_Ho_CpTUVW */
}

/* Wall pressure of a non-burning injection face, modflu/RFLU_ModRindStates.F90:535-584:
 * isentropic expansion for a receding normal velocity, the piston-shock relation for
 * an approaching one. */
static void rind_slipwall_perf(double cpgas, double mm, double nx, double ny, double nz,
                               double rl, double rul, double rvl, double rwl, double fs,
                               double* pl) {
  double rgas = 8314.3 / mm;
  double g = cpgas / (cpgas - rgas);
  double irl = 1.0 / rl;
  double ul = irl * rul, vl = irl * rvl, wl = irl * rwl;
  double ql = ul * nx + vl * ny + wl * nz - fs;
  double al = sqrt(g * *pl / rl);
  if (ql < 0.0) {
    double term = 1.0 + 0.5 * (g - 1.0) * ql / al;
    *pl = *pl * pow(term, 2.0 * g / (g - 1.0));
  } else {
    double term = (g + 1.0) / 4.0;
    *pl = *pl + term * rl * ql * (ql + sqrt(al * al + term * term * ql * ql) / term);
  }
}

/* Central flux of an outflow face from the interior and boundary states,
 * rocflu/RFLU_CentralFirstPatch.F90:676-686 (same form in CentralSecondPatch). */
static void outflow_flux(double nx, double ny, double nz, double nm, double fs, double rl,
                         double rul, double rvl, double rwl, double rel, double pl, double rr,
                         double rur, double rvr, double rwr, double rer, double pr, double* flx) {
  double ql = (rul * nx + rvl * ny + rwl * nz) / rl - fs;
  double qr = (rur * nx + rvr * ny + rwr * nz) / rr - fs;
  flx[0] = 0.5 * (ql * rl + qr * rr) * nm;
  flx[1] = 0.5 * (ql * rul + pl * nx + qr * rur + pr * nx) * nm;
  flx[2] = 0.5 * (ql * rvl + pl * ny + qr * rvr + pr * ny) * nm;
  flx[3] = 0.5 * (ql * rwl + pl * nz + qr * rwr + pr * nz) * nm;
  flx[4] = 0.5 * (ql * (rel + pl) + pl * fs + qr * (rer + pr) + pr * fs) * nm;
}

/* Left/right states of an interior face for the flux functions.  Order 2 adds the
 * linear reconstruction at the face centroid (e.g. RFLU_ModHLLCFlux.F90:2231-2262). */
typedef struct { double r, u, v, w, p, a, e, H, ir; } face_state;

static void reconstruct(const orc_region* R, const double* cv, const double* dv,
                        const double* grad, int c, const double* xf, face_state* s) {
  const double* q = cv + 5 * (size_t)c;
  double r = q[0];
  double ir = 1.0 / r;
  double u = q[1] * ir, v = q[2] * ir, w = q[3] * ir;
  double p = dv[3 * (size_t)c];
  const double* gr = grad + 15 * (size_t)c;   /* grad(3,5,N) */
  double dx = xf[0] - R->cofg[3 * (size_t)c];
  double dy = xf[1] - R->cofg[3 * (size_t)c + 1];
  double dz = xf[2] - R->cofg[3 * (size_t)c + 2];
  s->r = r + gr[0] * dx + gr[1] * dy + gr[2] * dz;
  s->u = u + gr[3] * dx + gr[4] * dy + gr[5] * dz;
  s->v = v + gr[6] * dx + gr[7] * dy + gr[8] * dz;
  s->w = w + gr[9] * dx + gr[10] * dy + gr[11] * dz;
  s->p = p + gr[12] * dx + gr[13] * dy + gr[14] * dz;
  s->ir = ir;
}

/* Interior face flux, dispatched on DISCR and order (rocflu/RFLU_ComputeFluxInv.F90:123-140). */
static void interior_flux(const orc_region* R, const double* cv, const double* dv,
                          const double* grad, int f, double* flx) {
  int c1 = R->f2c[2 * (size_t)f], c2 = R->f2c[2 * (size_t)f + 1];
  const double* n = R->fn + 4 * (size_t)f;
  double nx = n[0], ny = n[1], nz = n[2], nm = n[3];
  const double fs = 0.0; /* static grid: gs(indGs*ifg) with indGs = 0 */
  const double g = R->gamma, cp = R->cp;
  const double gc = cp - cp / g;  /* MixtPerf_R_CpG */
  if (R->order == 1) {
    const double* L = cv + 5 * (size_t)c1;
    const double* Rr = cv + 5 * (size_t)c2;
    double rl = L[0], irl = 1.0 / rl;
    double ul = L[1] * irl, vl = L[2] * irl, wl = L[3] * irl;
    double pl = dv[3 * (size_t)c1], tl = dv[3 * (size_t)c1 + 1], al = dv[3 * (size_t)c1 + 2];
    double rr = Rr[0], irr = 1.0 / rr;
    double ur = Rr[1] * irr, vr = Rr[2] * irr, wr = Rr[3] * irr;
    double pr = dv[3 * (size_t)c2], tr = dv[3 * (size_t)c2 + 1], ar = dv[3 * (size_t)c2 + 2];
    if (R->discr == 3) {
      double el = L[4] * irl, er = Rr[4] * irr;
      double Hl = el + irl * pl, Hr = er + irr * pr;
      hllc_core(g, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, al, el, Hl, rr, ur, vr, wr, pr, ar,
                er, Hr, flx);
    } else if (R->discr == 1) {
      double Hl = cp * tl + 0.5 * (ul * ul + vl * vl + wl * wl);  /* MixtPerf_Ho_CpTUVW */
      double Hr = cp * tr + 0.5 * (ur * ur + vr * vr + wr * wr);
      roe_core(g, R->eps_entr, R->diss_fact, nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, rr, ur,
               vr, wr, pr, Hr, flx);
    } else {
      /* RFLU_ModAUSMFlux.F90:668,713: H = E/rho + p/rho */
      double Hl = L[4] * irl + pl / rl, Hr = Rr[4] * irr + pr / rr;
      ausm_core(nx, ny, nz, nm, fs, rl, ul, vl, wl, pl, Hl, al, rr, ur, vr, wr, pr, Hr, ar, flx);
    }
    return;
  }
  const double* xf = R->fc + 3 * (size_t)f;
  face_state sl, sr;
  reconstruct(R, cv, dv, grad, c1, xf, &sl);
  reconstruct(R, cv, dv, grad, c2, xf, &sr);
  if (R->discr == 3) {
    /* RFLU_ModHLLCFlux.F90:2262-2267: E from the face state, H mixes the cell 1/rho. */
    double el = sl.p / (sl.r * (g - 1.0)) + 0.5 * (sl.u * sl.u + sl.v * sl.v + sl.w * sl.w);
    double al = sqrt(g * sl.p / sl.r);
    double Hl = el + sl.ir * sl.p;
    double er = sr.p / (sr.r * (g - 1.0)) + 0.5 * (sr.u * sr.u + sr.v * sr.v + sr.w * sr.w);
    double ar = sqrt(g * sr.p / sr.r);
    double Hr = er + sr.ir * sr.p;
    hllc_core(g, nx, ny, nz, nm, fs, sl.r, sl.u, sl.v, sl.w, sl.p, al, el, Hl, sr.r, sr.u, sr.v,
              sr.w, sr.p, ar, er, Hr, flx);
  } else if (R->discr == 1) {
    double tl = sl.p / (sl.r * gc), tr = sr.p / (sr.r * gc);
    double Hl = cp * tl + 0.5 * (sl.u * sl.u + sl.v * sl.v + sl.w * sl.w);
    double Hr = cp * tr + 0.5 * (sr.u * sr.u + sr.v * sr.v + sr.w * sr.w);
    roe_core(g, R->eps_entr, R->diss_fact, nx, ny, nz, nm, fs, sl.r, sl.u, sl.v, sl.w, sl.p, Hl,
             sr.r, sr.u, sr.v, sr.w, sr.p, Hr, flx);
  } else {
    double tl = sl.p / (sl.r * gc), tr = sr.p / (sr.r * gc);
    double al = sqrt(g * gc * tl), ar = sqrt(g * gc * tr);
    double Hl = cp * tl + 0.5 * (sl.u * sl.u + sl.v * sl.v + sl.w * sl.w);
    double Hr = cp * tr + 0.5 * (sr.u * sr.u + sr.v * sr.v + sr.w * sr.w);
    ausm_core(nx, ny, nz, nm, fs, sl.r, sl.u, sl.v, sl.w, sl.p, Hl, al, sr.r, sr.u, sr.v, sr.w,
              sr.p, Hr, ar, flx);
  }
}

/* Flux of a boundary face from its interior state (rl, rul, rvl, rwl, rel, pl) by
 * patch type; the same in both orders (rocflu/RFLU_CentralFirstPatch.F90,
 * RFLU_CentralSecondPatch.F90, case by case):
 *   slip wall   (0, p n |A|, 0)                                   First :785-906
 *   inflow 10   BcondInflowPerf state, central average            First :228-374
 *   inflow 11   prescribed velocity/temperature state, central    First :375-570
 *   outflow 20  BcondOutflowPerf state, central average           First :571-784
 *   farfield 80 RFLU_SetRindStateFarfieldPerf state, central      First :948-1042
 *   injection 90  minj > 0: injected mass, momentum and enthalpy with the interior
 *               pressure; minj <= 0: a wall with the corrected pressure  First :1043-1127
 * Returns 1 where the face carries mass (every type but the slip wall). */
static int patch_flux(const orc_region* R, int patch, const double* n, double rl, double rul,
                      double rvl, double rwl, double rel, double pl, double* flx) {
  double nx = n[0], ny = n[1], nz = n[2], nm = n[3];
  const double fs = 0.0;
  double mm, rgas, gam;
  gas_prime(R, &mm, &rgas, &gam);
  int bc = R->patch_bc[patch];
  int opt = R->patch_opt[patch];
  int opt2 = R->patch_opt2 ? R->patch_opt2[patch] : 0;
  const double* vals = R->patch_vals ? R->patch_vals + 5 * (size_t)patch : NULL;
  double rr, rur, rvr, rwr, rer, pr;
  switch (bc) {
    case BC_SLIPWALL:
      flx[0] = 0.0;
      flx[1] = pl * nx * nm;
      flx[2] = pl * ny * nm;
      flx[3] = pl * nz * nm;
      flx[4] = pl * fs * nm;
      return 0;
    case BC_INJECTION: {
      double minj = vals[0];
      if (minj > 0.0) {
        double ri, ui, vi, wi, Hi;
        rind_inject_perf(R->cp, mm, nx, ny, nz, minj, vals[1], pl, fs, &ri, &ui, &vi, &wi, &Hi);
        flx[0] = -minj * nm;
        flx[1] = (-minj * ui + pl * nx) * nm;
        flx[2] = (-minj * vi + pl * ny) * nm;
        flx[3] = (-minj * wi + pl * nz) * nm;
        flx[4] = (-minj * Hi + pl * fs) * nm;
      } else {
        rind_slipwall_perf(R->cp, mm, nx, ny, nz, rl, rul, rvl, rwl, fs, &pl);
        flx[0] = 0.0;
        flx[1] = pl * nx * nm;
        flx[2] = pl * ny * nm;
        flx[3] = pl * nz * nm;
        flx[4] = pl * fs * nm;
      }
      return 1;
    }
    case BC_INFLOW_VELTEMP:
      inflow_veltemp(opt, vals, rl, rgas, gam, &rr, &rur, &rvr, &rwr, &rer, &pr);
      break;
    case BC_INFLOW_TOTANG:
      bcond_inflow_perf(opt, opt2, vals[0], vals[1], vals[2], vals[3],
                        opt != BCOPT_SUBSONIC ? vals[4] : 0.0, nx, ny, nz, R->cp, mm, rl, rul,
                        rvl, rwl, &rr, &rur, &rvr, &rwr, &rer, &pr);
      break;
    case BC_FARFIELD:
      rind_farfield_perf(R->cp, mm, nx, ny, nz, vals[0], vals[3], vals[4], vals[1], vals[2], rl,
                         rul, rvl, rwl, rel, &rr, &rur, &rvr, &rwr, &rer, &pr);
      break;
    default:  /* BC_OUTFLOW */
      pr = opt != BCOPT_SUPERSONIC ? R->patch_pout[patch] : pl;
      bcond_outflow_perf(opt, pr, nx, ny, nz, R->cp, mm, rl, rul, rvl, rwl, rel, pl, &rr, &rur,
                         &rvr, &rwr, &rer);
      break;
  }
  outflow_flux(nx, ny, nz, nm, fs, rl, rul, rvl, rwl, rel, pl, rr, rur, rvr, rwr, rer, pr, flx);
  return 1;
}

/* Test entry points: the boundary state of patch `patch` for interior state
 * q = (rho, rho u, rho v, rho w, rho E) and pressure pl at a face with n = (nx, ny, nz,
 * |A|), and that face's flux.  out = (rho, rho u, rho v, rho w, rho E, p) of the
 * boundary state; injection faces give (rho, u, v, w, H, -) of the injected gas when
 * burning and (-, -, -, -, -, p_wall) otherwise; outflow faces give the pressure the
 * face flux uses (the prescribed one unless the option is supersonic, also where the
 * state is copied, rocflu/RFLU_CentralFirstPatch.F90:595-599). */
void orc_rind_state(const orc_region* R, int patch, const double* n, const double* q, double pl,
                    double* out) {
  double mm, rgas, gam;
  gas_prime(R, &mm, &rgas, &gam);
  int bc = R->patch_bc[patch], opt = R->patch_opt[patch];
  int opt2 = R->patch_opt2 ? R->patch_opt2[patch] : 0;
  const double* v = R->patch_vals + 5 * (size_t)patch;
  for (int k = 0; k < 6; ++k) out[k] = NAN;
  switch (bc) {
    case BC_INFLOW_TOTANG:
      bcond_inflow_perf(opt, opt2, v[0], v[1], v[2], v[3], opt != BCOPT_SUBSONIC ? v[4] : 0.0,
                        n[0], n[1], n[2], R->cp, mm, q[0], q[1], q[2], q[3], &out[0], &out[1],
                        &out[2], &out[3], &out[4], &out[5]);
      break;
    case BC_FARFIELD:
      rind_farfield_perf(R->cp, mm, n[0], n[1], n[2], v[0], v[3], v[4], v[1], v[2], q[0], q[1],
                         q[2], q[3], q[4], &out[0], &out[1], &out[2], &out[3], &out[4], &out[5]);
      break;
    case BC_OUTFLOW:
      out[5] = opt != BCOPT_SUPERSONIC ? R->patch_pout[patch] : pl;
      bcond_outflow_perf(opt, out[5], n[0], n[1], n[2], R->cp, mm, q[0], q[1], q[2], q[3], q[4],
                         pl, &out[0], &out[1], &out[2], &out[3], &out[4]);
      break;
    case BC_INJECTION:
      if (v[0] > 0.0) {
        rind_inject_perf(R->cp, mm, n[0], n[1], n[2], v[0], v[1], pl, 0.0, &out[0], &out[1],
                         &out[2], &out[3], &out[4]);
      } else {
        out[5] = pl;
        rind_slipwall_perf(R->cp, mm, n[0], n[1], n[2], q[0], q[1], q[2], q[3], 0.0, &out[5]);
      }
      break;
    default:
      break;
  }
}

/* Test entry point: the flux of interior face f alone, from cv, dv and (order 2) the
 * gradients grad(3,5,N), as ConvectiveFluxesMP evaluates it before the scatter. */
void orc_interior_flux(const orc_region* R, const double* cv, const double* dv,
                       const double* grad, int f, double* flx) {
  interior_flux(R, cv, dv, grad, f, flx);
}

void orc_patch_flux(const orc_region* R, int patch, const double* n, const double* q, double pl,
                    double* flx) {
  (void)patch_flux(R, patch, n, q[0], q[1], q[2], q[3], q[4], pl, flx);
}

/* Boundary-face flux of patch face b (rocflu/RFLU_CentralFirstPatch.F90,
 * rocflu/RFLU_CentralSecondPatch.F90).  Boundary patches run at first order
 * unless the patch asks for second order; this restatement uses the region's
 * patch's order (patch_order[p] == 2 selects second order; the reference default,
 * spaceOrderBFaces = 1, is at libflu/RFLU_InitInputValues.F90:195 and the per-patch
 * ORDER key at modflu/RFLU_ModReadBcInputFile.F90:301-309).  A first-order patch
 * in a second-order run reads cv after CellGradientsMP's conversion, whose slot 5
 * holds p (see orc_cell_gradients), exactly as the reference's patches do.  Second
 * order reconstructs rho, u, v, w, p at the face centroid and forms rho E from them
 * (RFLU_CentralSecondPatch.F90:258-296). */
static void boundary_flux(const orc_region* R, const double* cv, const double* dv,
                          const double* grad, int patch, int b, int patch_order2, double* flx,
                          int* has_mass) {
  int c1 = R->bf2c[b];
  const double* n = R->bfn + 4 * (size_t)b;
  const double* q = cv + 5 * (size_t)c1;
  if (!patch_order2) {
    *has_mass = patch_flux(R, patch, n, q[0], q[1], q[2], q[3], q[4], dv[3 * (size_t)c1], flx);
    return;
  }
  double mm, rgas, gam;
  gas_prime(R, &mm, &rgas, &gam);
  double rl = q[0], irl = 1.0 / rl;
  const double* xf = R->bfc + 3 * (size_t)b;
  double dx = xf[0] - R->cofg[3 * (size_t)c1];
  double dy = xf[1] - R->cofg[3 * (size_t)c1 + 1];
  double dz = xf[2] - R->cofg[3 * (size_t)c1 + 2];
  double ul = q[1] * irl, vl = q[2] * irl, wl = q[3] * irl;
  double pl = dv[3 * (size_t)c1];
  const double* gr = grad + 15 * (size_t)c1;
  rl = rl + gr[0] * dx + gr[1] * dy + gr[2] * dz;
  ul = ul + gr[3] * dx + gr[4] * dy + gr[5] * dz;
  vl = vl + gr[6] * dx + gr[7] * dy + gr[8] * dz;
  wl = wl + gr[9] * dx + gr[10] * dy + gr[11] * dz;
  pl = pl + gr[12] * dx + gr[13] * dy + gr[14] * dz;
  double rel = rl * (pl / (rl * (gam - 1.0)) + 0.5 * (ul * ul + vl * vl + wl * wl));
  *has_mass = patch_flux(R, patch, n, rl, rl * ul, rl * vl, rl * wl, rel, pl, flx);
}

/* Faces of cell icg in the scatter's order: keys f (interior face f) and F + b
 * (boundary face b), ascending; sign[k] = +1 where the cell is c1 of an interior face,
 * -1 where it is c2, 0 for a boundary face.  Returns the count (at most 16). */
static int cell_faces_sorted(const orc_region* R, int icg, int* key, int* sign) {
  int n = 0;
  const int F = R->n_faces;
  for (int k = R->c2f_ptr[icg]; k < R->c2f_ptr[icg + 1] && n < 16; ++k) {
    int e = R->c2f_idx[k], kk, sg;
    if (e > F) { kk = e - 1; sg = 0; }
    else if (e > 0) { kk = e - 1; sg = 1; }
    else { kk = -e - 1; sg = -1; }
    int j = n++;
    while (j > 0 && key[j - 1] > kk) { key[j] = key[j - 1]; sign[j] = sign[j - 1]; --j; }
    key[j] = kk;
    sign[j] = sg;
  }
  return n;
}

static int patch_of(const orc_region* R, int b) {
  int p = 0;
  while (b >= R->bpatch_ptr[p + 1]) ++p;
  return p;
}

/* orc_fluxes with threads: face fluxes into buffers, then each cell adds its faces'
 * terms in the scatter's order; mass in/out summed over boundary faces in order. */
static void orc_fluxes_gather(const orc_region* R, const double* cv, const double* dv,
                              const double* grad, double* rhs, orc_stats* st, int stage) {
  const int F = R->n_faces, B = R->bpatch_ptr[R->n_patches], N = R->n_cells;
  double* ff = (double*)malloc(5 * ((size_t)F + (size_t)B) * sizeof(double));
  double* bf = ff + 5 * (size_t)F;
  ORC_PAR
  for (int f = 0; f < F; ++f) interior_flux(R, cv, dv, grad, f, ff + 5 * (size_t)f);
  ORC_PAR
  for (int b = 0; b < B; ++b) {
    int p = patch_of(R, b), has_mass;
    int o2 = R->order == 2 && R->patch_order && R->patch_order[p] == 2;
    boundary_flux(R, cv, dv, grad, p, b, o2, bf + 5 * (size_t)b, &has_mass);
  }
  ORC_PAR
  for (int c = 0; c < N; ++c) {
    int key[16], sign[16];
    int n = cell_faces_sorted(R, c, key, sign);
    double* r = rhs + 5 * (size_t)c;
    for (int k = 0; k < 5; ++k) r[k] = -0.0;
    for (int i = 0; i < n; ++i) {
      if (sign[i] > 0) {
        for (int k = 0; k < 5; ++k) r[k] = r[k] + ff[5 * (size_t)key[i] + k];
      } else if (sign[i] < 0) {
        for (int k = 0; k < 5; ++k) r[k] = r[k] - ff[5 * (size_t)key[i] + k];
      } else {
        int b = key[i] - F;
        const double* x = bf + 5 * (size_t)b;
        int k0 = R->patch_bc[patch_of(R, b)] == BC_SLIPWALL ? 1 : 0;
        for (int k = k0; k < 5; ++k) r[k] = r[k] + x[k];
      }
    }
  }
  if (st && stage == 1)
    for (int p = 0; p < R->n_patches; ++p) {
      if (R->patch_bc[p] == BC_SLIPWALL) continue;
      for (int b = R->bpatch_ptr[p]; b < R->bpatch_ptr[p + 1]; ++b) {
        st->mass_in = st->mass_in - dmin(bf[5 * (size_t)b], 0.0);
        st->mass_out = st->mass_out + dmax(bf[5 * (size_t)b], 0.0);
      }
    }
  free(ff);
}

/* ConvectiveFluxesMP -> RFLU_ComputeFluxInv: rhs = -diss (= 0 here), then the interior
 * face scatter in ascending face order and the patch scatter
 * (libfloflu/ConvectiveFluxes.F90:133-159, modflu/RFLU_ModHLLCFlux.F90:2207-2425). */
void orc_fluxes(const orc_region* R, const double* cv, const double* dv, const double* grad,
                double* rhs, orc_stats* st, int stage) {
  if (g_threads > 1) {
    orc_fluxes_gather(R, cv, dv, grad, rhs, st, stage);
    return;
  }
  size_t N = (size_t)R->n_cells;
  for (size_t i = 0; i < 5 * N; ++i) rhs[i] = -0.0;
  double flx[5];
  for (int f = 0; f < R->n_faces; ++f) {
    interior_flux(R, cv, dv, grad, f, flx);
    int c1 = R->f2c[2 * (size_t)f], c2 = R->f2c[2 * (size_t)f + 1];
    for (int k = 0; k < 5; ++k) rhs[5 * (size_t)c1 + k] = rhs[5 * (size_t)c1 + k] + flx[k];
    for (int k = 0; k < 5; ++k) rhs[5 * (size_t)c2 + k] = rhs[5 * (size_t)c2 + k] - flx[k];
  }
  for (int p = 0; p < R->n_patches; ++p) {
    int o2 = R->order == 2 && R->patch_order && R->patch_order[p] == 2;
    for (int b = R->bpatch_ptr[p]; b < R->bpatch_ptr[p + 1]; ++b) {
      int has_mass;
      boundary_flux(R, cv, dv, grad, p, b, o2, flx, &has_mass);
      int c1 = R->bf2c[b];
      if (R->patch_bc[p] == BC_SLIPWALL) {
        for (int k = 1; k < 5; ++k) rhs[5 * (size_t)c1 + k] = rhs[5 * (size_t)c1 + k] + flx[k];
      } else {
        for (int k = 0; k < 5; ++k) rhs[5 * (size_t)c1 + k] = rhs[5 * (size_t)c1 + k] + flx[k];
        if (st && stage == 1) {
          st->mass_in = st->mass_in - dmin(flx[0], 0.0);
          st->mass_out = st->mass_out + dmax(flx[0], 0.0);
        }
      }
    }
  }
}

/* RFLU_ComputeGradCellsFast_3D, modflu/RFLU_ModDifferentiationCells.F90:931-1144.
 * var (5,N) holds (rho,u,v,w,p); grad (3,5,N). */
void orc_gradients(const orc_region* R, const double* var, double* grad) {
  ORC_PAR
  for (int icg = 0; icg < R->n_cells; ++icg) {
    const double* mo = R->xyz_moms + 10 * (size_t)icg;
    double r11 = mo[0], r12 = mo[1], r22 = mo[2], r13 = mo[3], r23 = mo[4], r33 = mo[5];
    double c11 = 1.0 / r11, c22 = 1.0 / r22, c33 = 1.0 / r33;
    double c12 = -c11 * r12;
    double c13 = -(c11 * r13 + c12 * c22 * r23);
    double c23 = -c22 * r23;
    double* g = grad + 15 * (size_t)icg;
    for (int k = 0; k < 15; ++k) g[k] = 0.0;
    double xc = R->cofg[3 * (size_t)icg], yc = R->cofg[3 * (size_t)icg + 1],
           zc = R->cofg[3 * (size_t)icg + 2];
    double c1111 = c11 * c11, c2222 = c22 * c22, c3333 = c33 * c33;
    const double* vi = var + 5 * (size_t)icg;
    for (int s = R->c2cs_ptr[icg]; s < R->c2cs_ptr[icg + 1]; ++s) {
      int j = R->c2cs_idx[s];
      double dx = R->cofg[3 * (size_t)j] - xc;
      double dy = R->cofg[3 * (size_t)j + 1] - yc;
      double dz = R->cofg[3 * (size_t)j + 2] - zc;
      double term = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      dx = term * dx;
      dy = term * dy;
      dz = term * dz;
      double term1 = c1111 * (dx);
      double term2 = c2222 * (dy + c12 * dx);
      double term3 = c3333 * (dz + c23 * dy + c13 * dx);
      double wtx = term * (term1 + c12 * term2 + c13 * term3);
      double wty = term * (term2 + c23 * term3);
      double wtz = term * (term3);
      const double* vj = var + 5 * (size_t)j;
      for (int iv = 0; iv < 5; ++iv) {
        double dVar = vj[iv] - vi[iv];
        g[3 * iv + 0] = g[3 * iv + 0] + wtx * dVar;
        g[3 * iv + 1] = g[3 * iv + 1] + wty * dVar;
        g[3 * iv + 2] = g[3 * iv + 2] + wtz * dVar;
      }
    }
  }
}

/* Least-squares pseudo-inverse of an m x 3 matrix a (column-major) by Householder QR:
 * ainv (3 x m, column-major) = R^-1 Q^T.  The reference inverts with LAPACK's SVD
 * (libflu/RFLU_InvertMatrixSVD.F90, which drops singular values below 1% of the
 * previous); for the full-rank matrices of the constrained gradient both give the
 * same pseudo-inverse to round-off.  Returns 1 if a is rank deficient. */
static int pinv_qr3(int m, const double* a, double* ainv) {
  double* q = (double*)malloc((size_t)m * 3 * sizeof(double));
  double* v = (double*)malloc((size_t)m * 3 * sizeof(double));
  double beta[3], r[3][3];
  memcpy(q, a, (size_t)m * 3 * sizeof(double));
  for (int k = 0; k < 3; ++k) {
    double nrm = 0.0;
    for (int i = k; i < m; ++i) nrm += q[k * m + i] * q[k * m + i];
    nrm = sqrt(nrm);
    double alpha = q[k * m + k] > 0 ? -nrm : nrm;
    for (int i = 0; i < m; ++i) v[k * m + i] = i < k ? 0.0 : q[k * m + i];
    v[k * m + k] -= alpha;
    double vv = 0.0;
    for (int i = k; i < m; ++i) vv += v[k * m + i] * v[k * m + i];
    beta[k] = vv > 0.0 ? 2.0 / vv : 0.0;
    for (int j = k; j < 3; ++j) {
      double d = 0.0;
      for (int i = k; i < m; ++i) d += v[k * m + i] * q[j * m + i];
      for (int i = k; i < m; ++i) q[j * m + i] -= beta[k] * d * v[k * m + i];
    }
  }
  int bad = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = j >= i ? q[j * m + i] : 0.0;
  for (int i = 0; i < 3; ++i)
    if (fabs(r[i][i]) < 1e-12 * fabs(r[0][0])) bad = 1;
  /* columns of Q^T: apply the reflectors to unit vectors e_c, keep the first 3 rows,
   * then back-substitute R x = (Q^T e_c)[0..2] */
  double* e = (double*)malloc((size_t)m * sizeof(double));
  for (int c = 0; c < m; ++c) {
    for (int i = 0; i < m; ++i) e[i] = i == c ? 1.0 : 0.0;
    for (int k = 0; k < 3; ++k) {
      double d = 0.0;
      for (int i = k; i < m; ++i) d += v[k * m + i] * e[i];
      for (int i = k; i < m; ++i) e[i] -= beta[k] * d * v[k * m + i];
    }
    double x[3];
    for (int i = 2; i >= 0; --i) {
      double t = e[i];
      for (int j = i + 1; j < 3; ++j) t -= r[i][j] * x[j];
      x[i] = t / r[i][i];
    }
    for (int i = 0; i < 3; ++i) ainv[c * 3 + i] = x[i];
  }
  free(e); free(v); free(q);
  return bad;
}

/* FloatEqual (modfloflu/ModTools.F90:129-148). */
static int float_equal(double a, double b, double tol) {
  return fabs(a - b) <= (1.0 + 0.5 * fabs(a + b)) * tol;
}

/* RFLU_ComputeGradCellsConstr, modflu/RFLU_ModDifferentiationCells.F90:1180-1554, for
 * the patches of this restatement: slip walls constrain the velocity component along
 * an axis-aligned wall normal to zero (RFLU_GetConstrType/Value,
 * modflu/RFLU_ModConstraintUtils.F90:122-334, 366-745).  Cells with wall faces in
 * their stencil get, per constrained variable, the weighted least-squares gradient
 * of the members plus the constraint rows (weight cReconstCellsWeight), solved with
 * column scaling and a pseudo-inverse, accumulated members first.  var (5,N) is
 * (rho,u,v,w,p); only u, v, w can carry constraints here. */
void orc_gradients_constr(const orc_region* R, const double* var, double* grad) {
  if (!R->creconst || !R->c2cs_bptr) return;
  const double cwt = R->creconst_weight;
  ORC_PAR
  for (int icg = 0; icg < R->n_cells; ++icg) {
    int b0 = R->c2cs_bptr[icg], nb = R->c2cs_bptr[icg + 1] - b0;
    if (nb == 0) continue;
    const double* xc = R->cofg + 3 * (size_t)icg;
    int s0 = R->c2cs_ptr[icg], ns = R->c2cs_ptr[icg + 1] - s0;
    for (int iv = 1; iv <= 3; ++iv) {
      int axis = iv - 1;
      int nc = 0;
      int is_c[9];
      for (int k = 0; k < nb; ++k) {
        int p = R->c2cs_bface[2 * (size_t)(b0 + k)], f = R->c2cs_bface[2 * (size_t)(b0 + k) + 1];
        int b = R->bpatch_ptr[p] + f;
        is_c[k] = R->patch_bc[p] == BC_SLIPWALL &&
                  float_equal(fabs(R->bfn[4 * (size_t)b + axis]), 1.0, 1.0E-6);
        nc += is_c[k];
      }
      if (nc == 0) continue;
      int m = ns + nc;
      double* a = (double*)malloc((size_t)m * 3 * sizeof(double));
      double* ai = (double*)malloc((size_t)m * 3 * sizeof(double));
      double* term = (double*)malloc((size_t)m * sizeof(double));
      double* dvar = (double*)malloc((size_t)m * sizeof(double));
      for (int s = 0; s < ns; ++s) {
        int j = R->c2cs_idx[s0 + s];
        double dx = R->cofg[3 * (size_t)j] - xc[0];
        double dy = R->cofg[3 * (size_t)j + 1] - xc[1];
        double dz = R->cofg[3 * (size_t)j + 2] - xc[2];
        term[s] = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
        a[0 * m + s] = term[s] * dx;
        a[1 * m + s] = term[s] * dy;
        a[2 * m + s] = term[s] * dz;
        dvar[s] = var[5 * (size_t)j + iv] - var[5 * (size_t)icg + iv];
      }
      int row = ns;
      for (int k = 0; k < nb; ++k) {
        if (!is_c[k]) continue;
        int p = R->c2cs_bface[2 * (size_t)(b0 + k)], f = R->c2cs_bface[2 * (size_t)(b0 + k) + 1];
        const double* xf = R->bfc + 3 * (size_t)(R->bpatch_ptr[p] + f);
        double dx = xf[0] - xc[0], dy = xf[1] - xc[1], dz = xf[2] - xc[2];
        term[row] = cwt / sqrt(dx * dx + dy * dy + dz * dz);
        a[0 * m + row] = term[row] * dx;
        a[1 * m + row] = term[row] * dy;
        a[2 * m + row] = term[row] * dz;
        dvar[row] = 0.0 - var[5 * (size_t)icg + iv];   /* slip wall: u.n = 0 */
        ++row;
      }
      double colmax[3];
      for (int c = 0; c < 3; ++c) {
        colmax[c] = -HUGE_VAL;
        for (int i = 0; i < m; ++i) colmax[c] = dmax(colmax[c], fabs(a[c * m + i]));
        for (int i = 0; i < m; ++i) a[c * m + i] = a[c * m + i] / colmax[c];
      }
      pinv_qr3(m, a, ai);
      double g[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < m; ++i)
        for (int c = 0; c < 3; ++c) g[c] = g[c] + term[i] * (ai[3 * i + c] / colmax[c]) * dvar[i];
      for (int c = 0; c < 3; ++c) grad[15 * (size_t)icg + 3 * iv + c] = g[c];
      free(a); free(ai); free(term); free(dvar);
    }
  }
}

/* RFLU_WENOGradCellsXYZFast_3D, modflu/RFLU_ModWENO.F90:1437-1508: per component, blend
 * the own and member gradients with weights 1/(g^2 + 1e-15), own value first. */
void orc_weno_xyz(const orc_region* R, double* grad) {
  size_t N = (size_t)R->n_cells;
  double* eno = (double*)malloc(15 * N * sizeof(double));
  ORC_PAR
  for (size_t icg = 0; icg < N; ++icg) {
    for (int k = 0; k < 15; ++k) {
      double term = grad[15 * icg + k];
      double sumw = 1.0 / (term * term + 1.0E-15);
      double loc = sumw * term;
      for (int s = R->c2cs_ptr[icg]; s < R->c2cs_ptr[icg + 1]; ++s) {
        double t = grad[15 * (size_t)R->c2cs_idx[s] + k];
        double w = 1.0 / (t * t + 1.0E-15);
        sumw = sumw + w;
        loc = loc + w * t;
      }
      eno[15 * icg + k] = loc / sumw;
    }
  }
  memcpy(grad, eno, 15 * N * sizeof(double));
  free(eno);
}

/* RFLU_LimitGradCellsSimple, modflu/RFLU_ModLimiters.F90:1199-1267: zero the gradient of
 * a positive variable (rho, p) when its extrapolation to any face centroid is <= 0. */
void orc_limit_simple(const orc_region* R, const double* var, double* grad) {
  static const int pos_vars[2] = {0, 4};
  ORC_PAR
  for (int icg = 0; icg < R->n_cells; ++icg) {
    double xc = R->cofg[3 * (size_t)icg], yc = R->cofg[3 * (size_t)icg + 1],
           zc = R->cofg[3 * (size_t)icg + 2];
    for (int q = 0; q < 2; ++q) {
      int iv = pos_vars[q];
      double* g = grad + 15 * (size_t)icg + 3 * iv;
      double varc = var[5 * (size_t)icg + iv];
      for (int k = R->c2f_ptr[icg]; k < R->c2f_ptr[icg + 1]; ++k) {
        int e = R->c2f_idx[k];
        const double* xf;
        if (e > R->n_faces || e < -R->n_faces) xf = R->bfc + 3 * (size_t)(e - R->n_faces - 1);
        else xf = R->fc + 3 * (size_t)((e > 0 ? e : -e) - 1);
        double dx = xf[0] - xc, dy = xf[1] - yc, dz = xf[2] - zc;
        double varf = varc + g[0] * dx + g[1] * dy + g[2] * dz;
        if (varf <= 0.0) {
          g[0] = 0.0; g[1] = 0.0; g[2] = 0.0;
          break;
        }
      }
    }
  }
}

/* One cell's side of the Venkatakrishnan limiter at face centroid xf (the loop body of
 * RFLU_ComputeLimiterVenkat for that side). */
static void venkat_side(const orc_region* R, int ci, const double* xf, const double* var,
                        const double* grad, const double* vmin, const double* vmax, double* lim) {
  const double THRD = 1.0 / 3.0, TINY = 1.0E-12, K = 1.0;
  double dx = R->cofg[3 * (size_t)ci] - xf[0];
  double dy = R->cofg[3 * (size_t)ci + 1] - xf[1];
  double dz = R->cofg[3 * (size_t)ci + 2] - xf[2];
  double ds = pow(R->vol[ci], THRD);
  double epsq = (K * ds) * (K * ds) * (K * ds);
  for (int iv = 0; iv < 5; ++iv) {
    double vi = var[5 * (size_t)ci + iv];
    double dmn = vmin[5 * (size_t)ci + iv] - vi;
    double dmx = vmax[5 * (size_t)ci + iv] - vi;
    const double* g = grad + 15 * (size_t)ci + 3 * iv;
    double d2 = g[0] * dx + g[1] * dy + g[2] * dz;
    double* l = lim + 5 * (size_t)ci + iv;
    if (d2 > 0.0) {
      double numer = (dmx * dmx + epsq) * d2 + 2 * d2 * d2 * dmx;
      double denom = d2 * (dmx * dmx + 2 * d2 * d2 + dmx * d2 + epsq);
      double term = numer / (denom + TINY);
      *l = dmin(term, *l);
    } else if (d2 < 0.0) {
      double numer = (dmn * dmn + epsq) * d2 + 2 * d2 * d2 * dmn;
      double denom = d2 * (dmn * dmn + 2 * d2 * d2 + dmn * d2 + epsq);
      double term = numer / (denom + TINY);
      *l = dmin(term, *l);
    }
  }
}

/* orc_venkat's face loops with threads: each cell takes min and max over its interior
 * faces' other cells and the limiter minimum over its own sides. */
static void venkat_gather(const orc_region* R, const double* var, double* grad, double* vmin,
                          double* vmax, double* lim) {
  const int F = R->n_faces;
  ORC_PAR
  for (int c = 0; c < R->n_cells; ++c) {
    for (int k = R->c2f_ptr[c]; k < R->c2f_ptr[c + 1]; ++k) {
      int e = R->c2f_idx[k];
      if (e > F || e < -F) continue;
      int f = (e > 0 ? e : -e) - 1;
      int o = e > 0 ? R->f2c[2 * (size_t)f + 1] : R->f2c[2 * (size_t)f];
      for (int iv = 0; iv < 5; ++iv) {
        double vo = var[5 * (size_t)o + iv];
        vmin[5 * (size_t)c + iv] = dmin(vmin[5 * (size_t)c + iv], vo);
        vmax[5 * (size_t)c + iv] = dmax(vmax[5 * (size_t)c + iv], vo);
      }
    }
  }
  ORC_PAR
  for (int c = 0; c < R->n_cells; ++c) {
    for (int k = R->c2f_ptr[c]; k < R->c2f_ptr[c + 1]; ++k) {
      int e = R->c2f_idx[k];
      if (e > F || e < -F) continue;
      int f = (e > 0 ? e : -e) - 1;
      venkat_side(R, c, R->fc + 3 * (size_t)f, var, grad, vmin, vmax, lim);
    }
    for (int iv = 0; iv < 5; ++iv)
      for (int d = 0; d < 3; ++d)
        grad[15 * (size_t)c + 3 * iv + d] = lim[5 * (size_t)c + iv] * grad[15 * (size_t)c + 3 * iv + d];
  }
}

/* RFLU_ComputeLimiterVenkat + RFLU_LimitGradCells,
 * modflu/RFLU_ModLimiters.F90:407-522,1100-1106.  Note dx1 = cofg(c1) - fc (the
 * reference's sign) and eps^2 = (K*vol^(1/3))^3 with K = 1. */
void orc_venkat(const orc_region* R, const double* var, double* grad) {
  const double THRD = 1.0 / 3.0, TINY = 1.0E-12, K = 1.0;
  size_t N = (size_t)R->n_cells;
  double* vmin = (double*)malloc(5 * N * sizeof(double));
  double* vmax = (double*)malloc(5 * N * sizeof(double));
  double* lim = (double*)malloc(5 * N * sizeof(double));
  for (size_t i = 0; i < 5 * N; ++i) { vmin[i] = var[i]; vmax[i] = var[i]; lim[i] = 1.0; }
  if (g_threads > 1) {
    venkat_gather(R, var, grad, vmin, vmax, lim);
    free(vmin); free(vmax); free(lim);
    return;
  }
  for (int f = 0; f < R->n_faces; ++f) {
    int c1 = R->f2c[2 * (size_t)f], c2 = R->f2c[2 * (size_t)f + 1];
    for (int iv = 0; iv < 5; ++iv) {
      double v1 = var[5 * (size_t)c1 + iv], v2 = var[5 * (size_t)c2 + iv];
      vmin[5 * (size_t)c1 + iv] = dmin(vmin[5 * (size_t)c1 + iv], v2);
      vmin[5 * (size_t)c2 + iv] = dmin(vmin[5 * (size_t)c2 + iv], v1);
      vmax[5 * (size_t)c1 + iv] = dmax(vmax[5 * (size_t)c1 + iv], v2);
      vmax[5 * (size_t)c2 + iv] = dmax(vmax[5 * (size_t)c2 + iv], v1);
    }
  }
  for (int f = 0; f < R->n_faces; ++f) {
    int c[2] = {R->f2c[2 * (size_t)f], R->f2c[2 * (size_t)f + 1]};
    const double* xf = R->fc + 3 * (size_t)f;
    for (int side = 0; side < 2; ++side) {
      int ci = c[side];
      double dx = R->cofg[3 * (size_t)ci] - xf[0];
      double dy = R->cofg[3 * (size_t)ci + 1] - xf[1];
      double dz = R->cofg[3 * (size_t)ci + 2] - xf[2];
      double ds = pow(R->vol[ci], THRD);
      double epsq = (K * ds) * (K * ds) * (K * ds);
      for (int iv = 0; iv < 5; ++iv) {
        double vi = var[5 * (size_t)ci + iv];
        double dmn = vmin[5 * (size_t)ci + iv] - vi;
        double dmx = vmax[5 * (size_t)ci + iv] - vi;
        const double* g = grad + 15 * (size_t)ci + 3 * iv;
        double d2 = g[0] * dx + g[1] * dy + g[2] * dz;
        double* l = lim + 5 * (size_t)ci + iv;
        if (d2 > 0.0) {
          double numer = (dmx * dmx + epsq) * d2 + 2 * d2 * d2 * dmx;
          double denom = d2 * (dmx * dmx + 2 * d2 * d2 + dmx * d2 + epsq);
          double term = numer / (denom + TINY);
          *l = dmin(term, *l);
        } else if (d2 < 0.0) {
          double numer = (dmn * dmn + epsq) * d2 + 2 * d2 * d2 * dmn;
          double denom = d2 * (dmn * dmn + 2 * d2 * d2 + dmn * d2 + epsq);
          double term = numer / (denom + TINY);
          *l = dmin(term, *l);
        }
      }
    }
  }
  for (size_t i = 0; i < N; ++i)
    for (int iv = 0; iv < 5; ++iv)
      for (int d = 0; d < 3; ++d)
        grad[15 * i + 3 * iv + d] = lim[5 * i + iv] * grad[15 * i + 3 * iv + d];
  free(vmin); free(vmax); free(lim);
}

/* RFLU_SymmetricEigen3, modflu/RFLU_ModLimiters.F90:791-886: cyclic Jacobi rotations
 * of a symmetric 3x3 matrix m (column-major, m[i + 3 j]) until the off-diagonal squares
 * fall to 1e-24 of the total (at most 50 sweeps); values ascending, vectors as columns
 * (vec[i + 3 j] is component i of vector j). */
void orc_sym_eigen3(const double* m, double* vec, double* val) {
  double r[9];
  for (int k = 0; k < 9; ++k) r[k] = m[k];
  for (int k = 0; k < 9; ++k) vec[k] = 0.0;
  vec[0] = vec[4] = vec[8] = 1.0;
  double total = 0.0;
  for (int k = 0; k < 9; ++k) total = total + r[k] * r[k];
  for (int sweep = 1; sweep <= 50; ++sweep) {
    double off = r[3] * r[3] + r[6] * r[6] + r[7] * r[7];
    if (off <= 1.0E-24 * total) break;
    for (int ip = 0; ip < 2; ++ip) {
      for (int iq = ip + 1; iq < 3; ++iq) {
        double apq = r[ip + 3 * iq];
        if (apq == 0.0) continue;
        double theta = (r[iq + 3 * iq] - r[ip + 3 * ip]) / (2.0 * apq);
        double t;
        if (fabs(theta) > 1.0E+150)
          t = 0.5 / theta;
        else
          t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0);
        double sn = t * c;
        int ir = 3 - ip - iq;
        double arp = r[ir + 3 * ip], arq = r[ir + 3 * iq];
        r[ip + 3 * ip] = r[ip + 3 * ip] - t * r[ip + 3 * iq];
        r[iq + 3 * iq] = r[iq + 3 * iq] + t * r[ip + 3 * iq];
        r[ip + 3 * iq] = 0.0;
        r[iq + 3 * ip] = 0.0;
        r[ir + 3 * ip] = c * arp - sn * arq;
        r[ip + 3 * ir] = r[ir + 3 * ip];
        r[ir + 3 * iq] = sn * arp + c * arq;
        r[iq + 3 * ir] = r[ir + 3 * iq];
        for (int i = 0; i < 3; ++i) {
          double vp = vec[i + 3 * ip], vq = vec[i + 3 * iq];
          vec[i + 3 * ip] = c * vp - sn * vq;
          vec[i + 3 * iq] = sn * vp + c * vq;
        }
      }
    }
  }
  for (int i = 0; i < 3; ++i) val[i] = r[i + 3 * i];
  for (int i = 0; i < 2; ++i) {
    int imin = i;
    for (int k = i + 1; k < 3; ++k)
      if (val[k] < val[imin]) imin = k;  /* MINLOC: the first of equal minima */
    if (imin != i) {
      double h = val[i];
      val[i] = val[imin];
      val[imin] = h;
      for (int k = 0; k < 3; ++k) {
        double cc = vec[k + 3 * i];
        vec[k + 3 * i] = vec[k + 3 * imin];
        vec[k + 3 * imin] = cc;
      }
    }
  }
}

/* RECONST 10: RFLU_ComputeLimiterBarthJesp (modflu/RFLU_ModLimiters.F90:132-322) for rho
 * and p (the velocity skipped, iVelVar given: libfloflu/CellGradientsMP.F90:229-238),
 * then RFLU_LimitGradVelocityBarthJesp (:588-764), which limits the velocity gradient
 * along the principal axes of its symmetric part, and RFLU_LimitGradCells (:1100-1106).
 * var (5,N) holds (rho,u,v,w,p), grad (3,5,N).  dx = cofg - fc, the reference's sign;
 * Fortran MATMUL and DOT_PRODUCT sums run in index order. */
void orc_barth_jesp(const orc_region* R, const double* var, double* grad) {
  const size_t N = (size_t)R->n_cells;
  double* vmin = (double*)malloc(5 * N * sizeof(double));
  double* vmax = (double*)malloc(5 * N * sizeof(double));
  double* lim = (double*)malloc(5 * N * sizeof(double));
  double* axes = (double*)malloc(9 * N * sizeof(double));
  double* eig = (double*)malloc(3 * N * sizeof(double));
  double* alpha = (double*)malloc(3 * N * sizeof(double));
  double* dmn = (double*)malloc(3 * N * sizeof(double));
  double* dmx = (double*)malloc(3 * N * sizeof(double));
  static const int lv[2] = {0, 4};
  for (size_t i = 0; i < 5 * N; ++i) { vmin[i] = var[i]; vmax[i] = var[i]; lim[i] = 1.0; }
  for (int f = 0; f < R->n_faces; ++f) {
    int c1 = R->f2c[2 * (size_t)f], c2 = R->f2c[2 * (size_t)f + 1];
    for (int a = 0; a < 2; ++a) {
      int iv = lv[a];
      double v1 = var[5 * (size_t)c1 + iv], v2 = var[5 * (size_t)c2 + iv];
      vmin[5 * (size_t)c1 + iv] = dmin(vmin[5 * (size_t)c1 + iv], v2);
      vmin[5 * (size_t)c2 + iv] = dmin(vmin[5 * (size_t)c2 + iv], v1);
      vmax[5 * (size_t)c1 + iv] = dmax(vmax[5 * (size_t)c1 + iv], v2);
      vmax[5 * (size_t)c2 + iv] = dmax(vmax[5 * (size_t)c2 + iv], v1);
    }
  }
  for (int f = 0; f < R->n_faces; ++f) {
    const double* xf = R->fc + 3 * (size_t)f;
    for (int side = 0; side < 2; ++side) {
      int ci = R->f2c[2 * (size_t)f + side];
      double dx = R->cofg[3 * (size_t)ci] - xf[0];
      double dy = R->cofg[3 * (size_t)ci + 1] - xf[1];
      double dz = R->cofg[3 * (size_t)ci + 2] - xf[2];
      for (int a = 0; a < 2; ++a) {
        int iv = lv[a];
        const double* g = grad + 15 * (size_t)ci + 3 * iv;
        double d2 = g[0] * dx + g[1] * dy + g[2] * dz;
        double vi = var[5 * (size_t)ci + iv];
        double* l = lim + 5 * (size_t)ci + iv;
        if (d2 > 0.0) {
          double term = dmin(1.0, (vmax[5 * (size_t)ci + iv] - vi) / d2);
          *l = dmin(term, *l);
        } else if (d2 < 0.0) {
          double term = dmin(1.0, (vmin[5 * (size_t)ci + iv] - vi) / d2);
          *l = dmin(term, *l);
        }
      }
    }
  }
  /* velocity: principal axes of the symmetric part of the velocity gradient G, with
   * G(i,k) = d u_i / d x_k = grad(k, 1+i, c) */
  for (size_t c = 0; c < N; ++c) {
    const double* G = grad + 15 * c + 3;
    double S[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) S[i + 3 * j] = 0.5 * (G[3 * i + j] + G[3 * j + i]);
    orc_sym_eigen3(S, axes + 9 * c, eig + 3 * c);
    for (int j = 0; j < 3; ++j) { alpha[3 * c + j] = 1.0; dmn[3 * c + j] = 0.0; dmx[3 * c + j] = 0.0; }
  }
  for (int f = 0; f < R->n_faces; ++f) {
    int cc[2] = {R->f2c[2 * (size_t)f], R->f2c[2 * (size_t)f + 1]};
    double du[3];
    for (int i = 0; i < 3; ++i) du[i] = var[5 * (size_t)cc[1] + 1 + i] - var[5 * (size_t)cc[0] + 1 + i];
    for (int side = 0; side < 2; ++side) {
      size_t c = (size_t)cc[side];
      const double* A = axes + 9 * c;
      for (int j = 0; j < 3; ++j) {
        double proj = A[0 + 3 * j] * du[0] + A[1 + 3 * j] * du[1] + A[2 + 3 * j] * du[2];
        if (side == 1) proj = -proj;
        dmn[3 * c + j] = dmin(dmn[3 * c + j], proj);
        dmx[3 * c + j] = dmax(dmx[3 * c + j], proj);
      }
    }
  }
  for (int f = 0; f < R->n_faces; ++f) {
    const double* xf = R->fc + 3 * (size_t)f;
    for (int side = 0; side < 2; ++side) {
      size_t c = (size_t)R->f2c[2 * (size_t)f + side];
      double off[3] = {R->cofg[3 * c] - xf[0], R->cofg[3 * c + 1] - xf[1], R->cofg[3 * c + 2] - xf[2]};
      const double* G = grad + 15 * c + 3;
      double sh[3];
      for (int i = 0; i < 3; ++i) sh[i] = G[3 * i] * off[0] + G[3 * i + 1] * off[1] + G[3 * i + 2] * off[2];
      const double* A = axes + 9 * c;
      for (int j = 0; j < 3; ++j) {
        double proj = A[0 + 3 * j] * sh[0] + A[1 + 3 * j] * sh[1] + A[2 + 3 * j] * sh[2];
        if (proj > 0.0)
          alpha[3 * c + j] = dmin(alpha[3 * c + j], dmx[3 * c + j] / proj);
        else if (proj < 0.0)
          alpha[3 * c + j] = dmin(alpha[3 * c + j], dmn[3 * c + j] / proj);
      }
    }
  }
  for (size_t c = 0; c < N; ++c) {
    double* al = alpha + 3 * c;
    const double* ev = eig + 3 * c;
    double mx = fabs(ev[0]);
    if (fabs(ev[1]) > mx) mx = fabs(ev[1]);
    if (fabs(ev[2]) > mx) mx = fabs(ev[2]);
    double scale = 1.0E-3 * mx;
    int e12 = ev[1] - ev[0] <= scale, e23 = ev[2] - ev[1] <= scale;
    if (e12 && e23) {
      double m = dmin(dmin(al[0], al[1]), al[2]);
      al[0] = al[1] = al[2] = m;
    } else if (e12) {
      double m = dmin(al[0], al[1]);
      al[0] = al[1] = m;
    } else if (e23) {
      double m = dmin(al[1], al[2]);
      al[1] = al[2] = m;
    }
    /* rotated = A^T G, rows scaled by alpha, G = A rotated */
    double* G = grad + 15 * c + 3;
    const double* A = axes + 9 * c;
    double rot[9];
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l)
        rot[j + 3 * l] = A[0 + 3 * j] * G[3 * 0 + l] + A[1 + 3 * j] * G[3 * 1 + l] + A[2 + 3 * j] * G[3 * 2 + l];
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l) rot[j + 3 * l] = al[j] * rot[j + 3 * l];
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < 3; ++l)
        G[3 * i + l] = A[i + 3 * 0] * rot[0 + 3 * l] + A[i + 3 * 1] * rot[1 + 3 * l] + A[i + 3 * 2] * rot[2 + 3 * l];
  }
  /* RFLU_LimitGradCells: rho and p by their limiters (the velocity's are 1) */
  for (size_t i = 0; i < N; ++i)
    for (int a = 0; a < 2; ++a)
      for (int d = 0; d < 3; ++d)
        grad[15 * i + 3 * lv[a] + d] = lim[5 * i + lv[a]] * grad[15 * i + 3 * lv[a] + d];
  free(vmin); free(vmax); free(lim); free(axes); free(eig); free(alpha); free(dmn); free(dmx);
}

/* CellGradientsMP, libfloflu/CellGradientsMP.F90:160-296: cons -> (rho,u,v,w,p) in
 * place, gradients, reconstruction modifier, and back to conservative form.  The
 * back-conversion of the default build leaves p in slot 5
 * (modflu/RFLU_ModConvertCv.F90:560-574), which this restatement keeps. */
void orc_cell_gradients(const orc_region* R, double* cv, const double* dv, double* grad) {
  size_t N = (size_t)R->n_cells;
  ORC_PAR
  for (size_t i = 0; i < N; ++i) {
    double ir = 1.0 / cv[5 * i];
    cv[5 * i + 1] = ir * cv[5 * i + 1];
    cv[5 * i + 2] = ir * cv[5 * i + 2];
    cv[5 * i + 3] = ir * cv[5 * i + 3];
    cv[5 * i + 4] = dv[3 * i];
  }
  orc_gradients(R, cv, grad);
  orc_gradients_constr(R, cv, grad);
  if (R->reconst == 2) {
    orc_weno_xyz(R, grad);
    orc_limit_simple(R, cv, grad);
  } else if (R->reconst == 11) {
    orc_venkat(R, cv, grad);
  } else if (R->reconst == 10) {
    orc_barth_jesp(R, cv, grad);
  }
  ORC_PAR
  for (size_t i = 0; i < N; ++i) {
    double r = cv[5 * i];
    cv[5 * i + 1] = r * cv[5 * i + 1];
    cv[5 * i + 2] = r * cv[5 * i + 2];
    cv[5 * i + 3] = r * cv[5 * i + 3];
  }
}

static int n_actual(const orc_region* R) { return R->n_act > 0 ? R->n_act : R->n_cells; }

/* Symmetry/periodic virtual cells: the copy of RungeKuttaMP.F90:440-459 (state, then
 * RFLU_SetVarsContWrapper over the virtual cells) and :229-236 (final gradients), each
 * turned as modflu/RFLU_ModRelatedPatches.F90:380-435 and 673-735 turn them: momentum
 * m (rho u), density and pressure gradients m g, velocity gradient m (G m^T). */
void orc_fill_virtual_state(const orc_region* R, double* cv, double* dv) {
  const int na = n_actual(R);
  if (na == R->n_cells) return;
  for (int c = na; c < R->n_cells; ++c) {
    const int s = R->virt_src[c - na];
    const double* m = R->xforms + 9 * (size_t)R->virt_xf[c - na];
    const double* x = cv + 5 * (size_t)s;
    double* y = cv + 5 * (size_t)c;
    y[0] = x[0];
    for (int i = 0; i < 3; ++i) y[1 + i] = m[3 * i] * x[1] + m[3 * i + 1] * x[2] + m[3 * i + 2] * x[3];
    y[4] = x[4];
  }
  orc_set_dependent_vars(R, cv, dv, na, R->n_cells);
}

void orc_fill_virtual_grad(const orc_region* R, double* grad) {
  const int na = n_actual(R);
  if (na == R->n_cells) return;
  for (int c = na; c < R->n_cells; ++c) {
    const int s = R->virt_src[c - na];
    const double* m = R->xforms + 9 * (size_t)R->virt_xf[c - na];
    const double* x = grad + 15 * (size_t)s;
    double* y = grad + 15 * (size_t)c;
    for (int v = 0; v < 5; v += 4)
      for (int i = 0; i < 3; ++i)
        y[3 * v + i] = m[3 * i] * x[3 * v] + m[3 * i + 1] * x[3 * v + 1] + m[3 * i + 2] * x[3 * v + 2];
    double G[9], K[9];
    for (int k = 0; k < 9; ++k) G[k] = x[3 + k];
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < 3; ++l)
        K[3 * i + l] = G[3 * i] * m[3 * l] + G[3 * i + 1] * m[3 * l + 1] + G[3 * i + 2] * m[3 * l + 2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        y[3 + 3 * i + j] = m[3 * i] * K[j] + m[3 * i + 1] * K[3 + j] + m[3 * i + 2] * K[6 + j];
  }
}

int orc_nstages(int rk) { return rk == 1 ? 4 : rk == 2 ? 3 : rk == 3 ? 1 : rk == 4 ? 3 : 0; }

/* RkUpdateGeneric, libfloflu/RkUpdateGeneric.F90:255-325 (static grid); ark/grk from
 * libflu/RFLU_DerivedInputValues.F90:347-396.  rk = 4 is the Shu-Osher SSP-RK3, which the
 * reference does not have: u1 = u0 - dt/V R0, u2 = 3/4 u0 + 1/4 (u1 - dt/V R1),
 * u3 = 1/3 u0 + 2/3 (u2 - dt/V R2), with cvOld = u0 kept from stage 1. */
void orc_rk_update(const orc_region* R, int stage, double dt, double* cv, const double* cvOld,
                   const double* rhs, double* rhsSum) {
  static const double ark4[4] = {0.5, 0.5, 1.0, 1.0 / 6.0};
  static const double grk4[4] = {1.0, 2.0, 2.0, 1.0};
  static const double ark3[3] = {8.0 / 15.0, 5.0 / 12.0, 0.75};
  static const double grk3[3] = {0.0, 17.0 / 25.0, 5.0 / 9.0};
  size_t N = (size_t)R->n_cells;
  int s = stage - 1;
  if (R->rk == 1) {
    double fac = ark4[s] * dt;
    ORC_PAR
    for (size_t ic = 0; ic < N; ++ic) {
      double adtv = fac / R->vol[ic];
      for (int iv = 0; iv < 5; ++iv) {
        size_t k = 5 * ic + iv;
        if (stage == 1) {
          cv[k] = cvOld[k] - adtv * rhs[k];
          rhsSum[k] = rhs[k];
        } else if (stage == 4) {
          cv[k] = cvOld[k] - adtv * (rhs[k] + rhsSum[k]);
        } else {
          cv[k] = cvOld[k] - adtv * rhs[k];
          rhsSum[k] = rhsSum[k] + grk4[s] * rhs[k];
        }
      }
    }
  } else if (R->rk == 2) {
    double fac = ark3[s] * dt;
    ORC_PAR
    for (size_t ic = 0; ic < N; ++ic) {
      double adtv = fac / R->vol[ic];
      for (int iv = 0; iv < 5; ++iv) {
        size_t k = 5 * ic + iv;
        if (stage == 1) {
          cv[k] = cvOld[k] - adtv * rhs[k];
          rhsSum[k] = rhs[k];
        } else if (stage == 2) {
          cv[k] = cvOld[k] - adtv * (rhs[k] - grk3[s] * rhsSum[k]);
          rhsSum[k] = rhs[k];
        } else {
          cv[k] = cvOld[k] - adtv * (rhs[k] - grk3[s] * rhsSum[k]);
        }
      }
    }
  } else if (R->rk == 3) {
    double fac = 1.0 * dt;
    ORC_PAR
    for (size_t ic = 0; ic < N; ++ic) {
      double adtv = fac / R->vol[ic];
      for (int iv = 0; iv < 5; ++iv) {
        size_t k = 5 * ic + iv;
        cv[k] = cvOld[k] - adtv * rhs[k];
      }
    }
  } else {
    ORC_PAR
    for (size_t ic = 0; ic < N; ++ic) {
      double adtv = dt / R->vol[ic];
      for (int iv = 0; iv < 5; ++iv) {
        size_t k = 5 * ic + iv;
        /* rhsSum carries the stage's starting state u_{s-1} (set by the driver). */
        double us = rhsSum[k] - adtv * rhs[k];
        if (stage == 1) cv[k] = us;
        else if (stage == 2) cv[k] = 0.75 * cvOld[k] + 0.25 * us;
        else cv[k] = (1.0 / 3.0) * cvOld[k] + (2.0 / 3.0) * us;
      }
    }
  }
}

/* AfterUpdateMP, libfloflu/AfterUpdateMP.F90:134-136: CheckValidity
 * (libflu/RFLU_CheckValidity.F90:141-222), EnforceBounds (libflu/RFLU_EnforceBounds.F90:
 * 169-231) and CheckPositivity (libflu/RFLU_CheckPositivity.F90:147-253). */
int orc_after_update(const orc_region* R, double* cv, double* dv, orc_stats* st) {
  size_t N = (size_t)n_actual(R);   /* AfterUpdateMP covers the actual cells */
  double mm, rgas, gamma;
  gas_prime(R, &mm, &rgas, &gamma);
  int any_nan = 0;
#pragma omp parallel for schedule(static) num_threads(g_threads) if(g_threads > 1) reduction(|:any_nan)
  for (size_t i = 0; i < N; ++i) {
    double rho = cv[5 * i], rrho = 1.0 / rho;
    double u = rrho * cv[5 * i + 1], v = rrho * cv[5 * i + 2], w = rrho * cv[5 * i + 3];
    double p = dv[3 * i], t = dv[3 * i + 1];
    if (isnan(rho) || isnan(u) || isnan(v) || isnan(w) || isnan(p) || isnan(t)) any_nan = 1;
  }
  if (any_nan) return ORC_ERR_INVALID_VALUE;
  const double pTol = 1.0E-12;
  int64_t nfix = 0;
#pragma omp parallel for schedule(static) num_threads(g_threads) if(g_threads > 1) reduction(+:nfix)
  for (size_t i = 0; i < N; ++i) {
    double r = cv[5 * i], ir = 1.0 / r;
    double u = ir * cv[5 * i + 1], v = ir * cv[5 * i + 2], w = ir * cv[5 * i + 3];
    double Eo = ir * cv[5 * i + 4];
    double Vm2 = u * u + v * v + w * w;
    double p = (gamma - 1.0) * r * (Eo - 0.5 * Vm2);
    if (p <= 0.0 && fabs(p) <= pTol) {
      nfix++;
      p = pTol;
      cv[5 * i + 4] = r * (p / (r * (gamma - 1.0)) + 0.5 * (u * u + v * v + w * w));
    }
  }
  if (st) st->n_bounds_fix += nfix;
  /* RFLU_CheckPositivity.F90:147-253: every cell is floored and counted, and the error
   * stop comes after the loop */
  int64_t nlocs = 0;
  for (size_t i = 0; i < N; ++i) {
    double rho = cv[5 * i], rrho = 1.0 / rho;
    double u = rrho * cv[5 * i + 1], v = rrho * cv[5 * i + 2], w = rrho * cv[5 * i + 3];
    double Vm2 = u * u + v * v + w * w;
    double p = dv[3 * i];
    if (p <= 0.0) {
      if (st) st->n_pos_fix++;
      double Eo = 1.0e4 + 0.5 * Vm2;
      cv[5 * i + 4] = rho * Eo;
      orc_set_dependent_vars(R, cv, dv, (int)i, (int)i + 1);
      p = dv[3 * i];
    }
    if (rho <= 0.0 || p <= 0.0) nlocs++;
  }
  return nlocs > 0 ? ORC_ERR_NEGATIVE_POSDEF : ORC_OK;
}

/* RFLU_TimeStepInviscid, rocflu/RFLU_TimeStepInviscid.F90:153-334: spectral-radius
 * sums over interior faces and boundary faces, dt = vol/sum, minimum over cells. */
static double face_srad(const orc_region* R, const double* cv, const double* dv, int f) {
  const double fsu = 0.0;
  int c1 = R->f2c[2 * (size_t)f], c2 = R->f2c[2 * (size_t)f + 1];
  const double* n = R->fn + 4 * (size_t)f;
  double ir1 = 1.0 / cv[5 * (size_t)c1], ir2 = 1.0 / cv[5 * (size_t)c2];
  double u1 = cv[5 * (size_t)c1 + 1] * ir1, v1 = cv[5 * (size_t)c1 + 2] * ir1,
         w1 = cv[5 * (size_t)c1 + 3] * ir1, a1 = dv[3 * (size_t)c1 + 2];
  double u2 = cv[5 * (size_t)c2 + 1] * ir2, v2 = cv[5 * (size_t)c2 + 2] * ir2,
         w2 = cv[5 * (size_t)c2 + 3] * ir2, a2 = dv[3 * (size_t)c2 + 2];
  double vc = 0.5 * ((u1 + u2) * n[0] + (v1 + v2) * n[1] + (w1 + w2) * n[2]);
  double as = 0.5 * (a1 + a2);
  return (fabs(vc - fsu) + as) * n[3];
}

static double bface_srad(const orc_region* R, const double* cv, const double* dv, int b) {
  const double fsu = 0.0;
  int c1 = R->bf2c[b];
  const double* n = R->bfn + 4 * (size_t)b;
  int p = patch_of(R, b);
  if (R->patch_bc[p] == BC_INJECTION) {
    /* rocflu/RFLU_TimeStepInviscid.F90:214-260: the injected state's normal velocity
     * and the sound speed at the injection temperature; a non-burning face counts its
     * sound speed only. */
    const double* vals = R->patch_vals + 5 * (size_t)p;
    double minj = vals[0], vc, a1;
    if (minj > 0.0) {
      double mm, rgas, gam, rl, ul, vl, wl, Hl;
      gas_prime(R, &mm, &rgas, &gam);
      rind_inject_perf(R->cp, mm, n[0], n[1], n[2], minj, vals[1], dv[3 * (size_t)c1], fsu, &rl,
                       &ul, &vl, &wl, &Hl);
      vc = ul * n[0] + vl * n[1] + wl * n[2];
      a1 = sqrt(gam * rgas * vals[1]);
    } else {
      vc = fsu;
      a1 = dv[3 * (size_t)c1 + 2];
    }
    return (fabs(vc - fsu) + a1) * n[3];
  }
  double ir1 = 1.0 / cv[5 * (size_t)c1];
  double u1 = cv[5 * (size_t)c1 + 1] * ir1, v1 = cv[5 * (size_t)c1 + 2] * ir1,
         w1 = cv[5 * (size_t)c1 + 3] * ir1, a1 = dv[3 * (size_t)c1 + 2];
  double vc = u1 * n[0] + v1 * n[1] + w1 * n[2];
  return (fabs(vc - fsu) + a1) * n[3];
}

/* orc_local_dt with threads: per-face spectral radii, then each cell sums its faces'
 * in the scatter's order. */
static double local_dt_gather(const orc_region* R, const double* cv, const double* dv) {
  const int F = R->n_faces;
  double* sr = (double*)malloc((size_t)F * sizeof(double));
  ORC_PAR
  for (int f = 0; f < F; ++f) sr[f] = face_srad(R, cv, dv, f);
  double mn = HUGE_VAL;
  const int na = n_actual(R);   /* RFLU_MinimumTimeStep: the actual cells */
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(min:mn)
  for (int c = 0; c < na; ++c) {
    int key[16], sign[16];
    int n = cell_faces_sorted(R, c, key, sign);
    double s = 0.0;
    for (int i = 0; i < n; ++i)
      s = s + (key[i] < F ? sr[key[i]] : bface_srad(R, cv, dv, key[i] - F));
    double d = R->vol[c] / s;
    if (d < mn) mn = d;
  }
  free(sr);
  return mn;
}

double orc_local_dt(const orc_region* R, const double* cv, const double* dv) {
  if (g_threads > 1) return local_dt_gather(R, cv, dv);
  size_t N = (size_t)R->n_cells;
  double* dt = (double*)malloc(N * sizeof(double));
  for (size_t i = 0; i < N; ++i) dt[i] = 0.0;
  const double fsu = 0.0;
  for (int f = 0; f < R->n_faces; ++f) {
    int c1 = R->f2c[2 * (size_t)f], c2 = R->f2c[2 * (size_t)f + 1];
    const double* n = R->fn + 4 * (size_t)f;
    double ir1 = 1.0 / cv[5 * (size_t)c1], ir2 = 1.0 / cv[5 * (size_t)c2];
    double u1 = cv[5 * (size_t)c1 + 1] * ir1, v1 = cv[5 * (size_t)c1 + 2] * ir1,
           w1 = cv[5 * (size_t)c1 + 3] * ir1, a1 = dv[3 * (size_t)c1 + 2];
    double u2 = cv[5 * (size_t)c2 + 1] * ir2, v2 = cv[5 * (size_t)c2 + 2] * ir2,
           w2 = cv[5 * (size_t)c2 + 3] * ir2, a2 = dv[3 * (size_t)c2 + 2];
    double vc = 0.5 * ((u1 + u2) * n[0] + (v1 + v2) * n[1] + (w1 + w2) * n[2]);
    double as = 0.5 * (a1 + a2);
    double srad = (fabs(vc - fsu) + as) * n[3];
    dt[c1] = dt[c1] + srad;
    dt[c2] = dt[c2] + srad;
  }
  for (int p = 0; p < R->n_patches; ++p)
    for (int b = R->bpatch_ptr[p]; b < R->bpatch_ptr[p + 1]; ++b) {
      int c1 = R->bf2c[b];
      dt[c1] = dt[c1] + bface_srad(R, cv, dv, b);
    }
  double mn = HUGE_VAL;
  for (size_t i = 0; i < (size_t)n_actual(R); ++i) {
    dt[i] = R->vol[i] / dt[i];
    if (dt[i] < mn) mn = dt[i];
  }
  free(dt);
  return mn;
}

/* RFLU_MinimumTimeStep, rocflu/RFLU_MinimumTimeStep.F90:126-245: CFL scaling, the
 * imposed step while currentTime <= 1e-8 (formed as 1/(1/dtImposed + 0/dt)), the
 * non-positive guard and the DTMINLIMIT floor. */
double orc_min_timestep(const orc_region* R, double dt_local_min, double current_time) {
  double wI = 0.0, wC = 0.0;
  double d = R->cfl * dt_local_min;
  if (current_time <= 1.0E-08) wI = 1.0;
  else wC = 1.0;
  d = wI / R->dt_imposed + wC / d;
  d = 1.0 / d;
  double dtMin = HUGE_VAL;
  if (d < dtMin) dtMin = d;
  if (dtMin <= 0.0) dtMin = 6.0E-12;
  dtMin = dmax(dtMin, R->dt_min_limit);
  return dtMin;
}

/* One RK stage of RungeKuttaMP (libfloflu/RungeKuttaMP.F90:148-508) for one region. */
static int orc_stage(const orc_region* R, int stage, double dt, double* cv, double* dv,
                     double* cvOld, double* rhs, double* rhsSum, double* grad, orc_stats* st) {
  size_t N = (size_t)R->n_cells;
  /* RKInitMP -> RkInitGeneric.F90:128-160 */
  if (R->rk == 1) {
    if (stage == 1) memcpy(cvOld, cv, 5 * N * sizeof(double));
  } else if (R->rk == 4) {
    if (stage == 1) memcpy(cvOld, cv, 5 * N * sizeof(double));
    memcpy(rhsSum, cv, 5 * N * sizeof(double)); /* stage start state */
  } else {
    memcpy(cvOld, cv, 5 * N * sizeof(double));
  }
  if (R->order == 2) {
    orc_cell_gradients(R, cv, dv, grad);
    orc_fill_virtual_grad(R, grad);
  }
  orc_fluxes(R, cv, dv, grad, rhs, st, stage);
  orc_rk_update(R, stage, dt, cv, cvOld, rhs, rhsSum);
  orc_set_dependent_vars(R, cv, dv, 0, R->n_cells);
  int rc = orc_after_update(R, cv, dv, st);
  if (rc == ORC_OK) orc_fill_virtual_state(R, cv, dv);
  return rc;
}

int orc_run(const orc_region* R, double* cv, double* dv, int64_t nsteps, orc_stats* st) {
  orc_fill_virtual_state(R, cv, dv);
  size_t N = (size_t)R->n_cells;
  int ns = orc_nstages(R->rk);
  if (ns == 0 || (R->order != 1 && R->order != 2)) return ORC_ERR_ARG;
  double* cvOld = (double*)malloc(5 * N * sizeof(double));
  double* rhs = (double*)malloc(5 * N * sizeof(double));
  double* rhsSum = (double*)calloc(5 * N, sizeof(double));
  double* grad = (double*)calloc(15 * N, sizeof(double));
  int rc = ORC_OK;
  for (int64_t step = 0; step < nsteps && rc == ORC_OK; ++step) {
    double dtl = orc_local_dt(R, cv, dv);
    double dt = orc_min_timestep(R, dtl, st->time);
    for (int s = 1; s <= ns && rc == ORC_OK; ++s)
      rc = orc_stage(R, s, dt, cv, dv, cvOld, rhs, rhsSum, grad, st);
    st->time = st->time + dt;
    st->dt_last = dt;
    st->steps++;
  }
  free(cvOld); free(rhs); free(rhsSum); free(grad);
  return rc;
}

/* RFLU_TimeStepping's unsteady loop up to MAXTIME (rocflu/RFLU_TimeStepping.F90:
 * 429-448,628-640): dt, the end-time clamp dt = endTime - currentTime, the stages, and
 * the time advance. */
int orc_run_until(const orc_region* R, double* cv, double* dv, double end_time,
                  int64_t max_steps, orc_stats* st) {
  orc_fill_virtual_state(R, cv, dv);
  size_t N = (size_t)R->n_cells;
  int ns = orc_nstages(R->rk);
  if (ns == 0 || (R->order != 1 && R->order != 2)) return ORC_ERR_ARG;
  double* cvOld = (double*)malloc(5 * N * sizeof(double));
  double* rhs = (double*)malloc(5 * N * sizeof(double));
  double* rhsSum = (double*)calloc(5 * N, sizeof(double));
  double* grad = (double*)calloc(15 * N, sizeof(double));
  int rc = ORC_OK, finished = 0;
  while (!finished && st->steps < max_steps && rc == ORC_OK) {
    double dtl = orc_local_dt(R, cv, dv);
    double dt = orc_min_timestep(R, dtl, st->time);
    if (st->time + dt > end_time) {
      dt = end_time - st->time;
      finished = 1;
    }
    for (int s = 1; s <= ns && rc == ORC_OK; ++s)
      rc = orc_stage(R, s, dt, cv, dv, cvOld, rhs, rhsSum, grad, st);
    st->time = st->time + dt;
    st->dt_last = dt;
    st->steps++;
  }
  free(cvOld); free(rhs); free(rhsSum); free(grad);
  return rc;
}

void orc_residual(const orc_region* R, const double* cv_in, const double* dv, double* rhs) {
  size_t N = (size_t)R->n_cells;
  double* cv = (double*)malloc(5 * N * sizeof(double));
  double* grad = (double*)calloc(15 * N, sizeof(double));
  memcpy(cv, cv_in, 5 * N * sizeof(double));
  if (R->order == 2) orc_cell_gradients(R, cv, dv, grad);
  orc_fluxes(R, cv, dv, grad, rhs, NULL, 0);
  free(cv); free(grad);
}
