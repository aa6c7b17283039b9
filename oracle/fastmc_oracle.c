/*
 * FastMC trial-move energy path: CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's Fortran
 * algorithm in plain C so that tests can check the HIP product against it.
 * Only tests/, the smoke check and bench.py's cpu_baseline leg may load it,
 * and only as the checker.  The product never links or calls it.
 *
 * Each function cites the reference routine (path:line relative to the
 * reference checkout) whose semantics it follows.  Operation order inside
 * each formula follows the Fortran left-to-right evaluation so that the
 * oracle agrees with a Fortran build to the last few ulps.
 *
 * Pinning (tests/test_oracle_*.py): known answers held by the reference's
 * example run (framework E_recip+self = -11190.606165 kcal/mol; final guest
 * energy -731.5072668178669 kcal/mol of the REVIVE configuration), the
 * published RANMAR known-answer sequence for duni, closed forms for the
 * tables and tail terms, and invariants (insert-then-delete, S(k) after
 * commit == full recompute, brute-force neighbour sets).
 */
#include "fastmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Fortran nint: round half away from zero */
static inline double f_nint(double x) { return round(x); }

/* ------------------------------------------------------------------ */
/* cell utilities                                                       */
/* ------------------------------------------------------------------ */

/* utility_pack.f:1351-1415 (dcell) */
void orc_dcell(const double a[9], double b[10]) {
  b[0] = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  b[1] = sqrt(a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
  b[2] = sqrt(a[6] * a[6] + a[7] * a[7] + a[8] * a[8]);
  b[3] = (a[0] * a[3] + a[1] * a[4] + a[2] * a[5]) / (b[0] * b[1]);
  b[4] = (a[0] * a[6] + a[1] * a[7] + a[2] * a[8]) / (b[0] * b[2]);
  b[5] = (a[3] * a[6] + a[4] * a[7] + a[5] * a[8]) / (b[1] * b[2]);
  double axb1 = a[1] * a[5] - a[2] * a[4];
  double axb2 = a[2] * a[3] - a[0] * a[5];
  double axb3 = a[0] * a[4] - a[1] * a[3];
  double bxc1 = a[4] * a[8] - a[5] * a[7];
  double bxc2 = a[5] * a[6] - a[3] * a[8];
  double bxc3 = a[3] * a[7] - a[4] * a[6];
  double cxa1 = a[7] * a[2] - a[1] * a[8];
  double cxa2 = a[0] * a[8] - a[2] * a[6];
  double cxa3 = a[1] * a[6] - a[0] * a[7];
  b[9] = fabs(a[0] * bxc1 + a[1] * bxc2 + a[2] * bxc3);
  b[6] = b[9] / sqrt(bxc1 * bxc1 + bxc2 * bxc2 + bxc3 * bxc3);
  b[7] = b[9] / sqrt(cxa1 * cxa1 + cxa2 * cxa2 + cxa3 * cxa3);
  b[8] = b[9] / sqrt(axb1 * axb1 + axb2 * axb2 + axb3 * axb3);
}

/* utility_pack.f:1417-1467 (invert) */
void orc_invert(const double a[9], double b[9], double *det) {
  b[0] = a[4] * a[8] - a[5] * a[7];
  b[1] = a[2] * a[7] - a[1] * a[8];
  b[2] = a[1] * a[5] - a[2] * a[4];
  b[3] = a[5] * a[6] - a[3] * a[8];
  b[4] = a[0] * a[8] - a[2] * a[6];
  b[5] = a[2] * a[3] - a[0] * a[5];
  b[6] = a[3] * a[7] - a[4] * a[6];
  b[7] = a[1] * a[6] - a[0] * a[7];
  b[8] = a[0] * a[4] - a[1] * a[3];
  double d = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
  double r = 0.0;
  if (fabs(d) > 0.0) r = 1.0 / d;
  for (int i = 0; i < 9; i++) b[i] = r * b[i];
  *det = d;
}

static double f_sign(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }

/* utility_pack.f:1138-1349 (images), imcon 1..7 */
void orc_images(int imcon, int n, const double cell[9], double *x, double *y,
                double *z) {
  const double rt2 = 1.41421356623, rt3 = 1.7320508075;
  if (imcon == 1) {
    double aaa = 1.0 / cell[0];
    for (int i = 0; i < n; i++) {
      x[i] = x[i] - cell[0] * f_nint(aaa * x[i]);
      y[i] = y[i] - cell[0] * f_nint(aaa * y[i]);
      z[i] = z[i] - cell[0] * f_nint(aaa * z[i]);
    }
  } else if (imcon == 2) {
    double aaa = 1.0 / cell[0], bbb = 1.0 / cell[4], ccc = 1.0 / cell[8];
    for (int i = 0; i < n; i++) {
      x[i] = x[i] - cell[0] * f_nint(aaa * x[i]);
      y[i] = y[i] - cell[4] * f_nint(bbb * y[i]);
      z[i] = z[i] - cell[8] * f_nint(ccc * z[i]);
    }
  } else if (imcon == 3) {
    double rc[9], det;
    orc_invert(cell, rc, &det);
    for (int i = 0; i < n; i++) {
      double ssx = (rc[0] * x[i] + rc[3] * y[i] + rc[6] * z[i]);
      double ssy = (rc[1] * x[i] + rc[4] * y[i] + rc[7] * z[i]);
      double ssz = (rc[2] * x[i] + rc[5] * y[i] + rc[8] * z[i]);
      double xss = ssx - f_nint(ssx);
      double yss = ssy - f_nint(ssy);
      double zss = ssz - f_nint(ssz);
      x[i] = (cell[0] * xss + cell[3] * yss + cell[6] * zss);
      y[i] = (cell[1] * xss + cell[4] * yss + cell[7] * zss);
      z[i] = (cell[2] * xss + cell[5] * yss + cell[8] * zss);
    }
  } else if (imcon == 4) {
    double aaa = 1.0 / cell[0];
    for (int i = 0; i < n; i++) {
      x[i] = x[i] - cell[0] * f_nint(aaa * x[i]);
      y[i] = y[i] - cell[0] * f_nint(aaa * y[i]);
      z[i] = z[i] - cell[0] * f_nint(aaa * z[i]);
      if ((fabs(x[i]) + fabs(y[i]) + fabs(z[i])) >= (0.75 * cell[0])) {
        x[i] = x[i] - 0.5 * f_sign(cell[0], x[i]);
        y[i] = y[i] - 0.5 * f_sign(cell[0], y[i]);
        z[i] = z[i] - 0.5 * f_sign(cell[0], z[i]);
      }
    }
  } else if (imcon == 5) {
    double aaa = 1.0 / cell[0], bbb = 1.0 / cell[8];
    for (int i = 0; i < n; i++) {
      x[i] = x[i] - cell[0] * f_nint(aaa * x[i]);
      y[i] = y[i] - cell[0] * f_nint(aaa * y[i]);
      z[i] = z[i] - cell[8] * f_nint(bbb * z[i]);
      if ((fabs(x[i]) + fabs(y[i]) + fabs(rt2 * z[i])) >= cell[0]) {
        x[i] = x[i] - 0.5 * f_sign(cell[0], x[i]);
        y[i] = y[i] - 0.5 * f_sign(cell[0], y[i]);
        z[i] = z[i] - 0.5 * f_sign(cell[8], z[i]);
      }
    }
  } else if (imcon == 6) {
    double det = cell[0] * cell[4] - cell[1] * cell[3];
    det = 1.0 / det;
    double r1 = det * cell[4], r2 = -det * cell[1], r4 = -det * cell[3],
           r5 = det * cell[0];
    for (int i = 0; i < n; i++) {
      double ssx = r1 * x[i] + r4 * y[i];
      double ssy = r2 * x[i] + r5 * y[i];
      double xss = ssx - f_nint(ssx);
      double yss = ssy - f_nint(ssy);
      x[i] = cell[0] * xss + cell[3] * yss;
      y[i] = cell[1] * xss + cell[4] * yss;
    }
  } else if (imcon == 7) {
    double aaa = cell[0] / (rt3 * 2.0), bbb = cell[0] / rt3,
           ccc = rt3 / cell[0], ddd = 1.0 / cell[8];
    for (int i = 0; i < n; i++) {
      y[i] = y[i] - bbb * f_nint(ccc * y[i]);
      z[i] = z[i] - cell[8] * f_nint(ddd * z[i]);
      if ((fabs(y[i]) + fabs(rt3 * x[i])) >= bbb) {
        x[i] = x[i] - rt3 * f_sign(aaa, x[i]);
        y[i] = y[i] - f_sign(aaa, y[i]);
      }
    }
  }
}

/* utility_pack.f:429-441 / readinputs.f:463-485 (volm per imcon) */
double orc_volume(int imcon, const double cell[9]) {
  double p[10];
  orc_dcell(cell, p);
  if (imcon == 0) return 0.0;
  if (imcon == 4 || imcon == 5 || imcon == 7) return 0.5 * p[9];
  return p[9];
}

/* utility_pack.f:380-392 */
void orc_ewald_precision(double eps, double rcut, int imcon,
                         const double cell[9], double *alpha, int kmax[3]) {
  double p[10];
  orc_dcell(cell, p);
  eps = fmin(fabs(eps), 0.5);
  double tol = sqrt(fabs(log(eps * rcut)));
  double a = sqrt(fabs(log(eps * rcut * tol))) / rcut;
  double t2 = 2.0 * tol * a;
  double tol1 = sqrt(-log(eps * rcut * (t2 * t2)));
  double fac = 1.0;
  if (imcon == 4 || imcon == 5 || imcon == 7) fac = pow(2.0, 1.0 / 3.0);
  for (int d = 0; d < 3; d++)
    kmax[d] = (int)f_nint(0.25 + fac * p[d] * a * tol1 / ORC_PI);
  *alpha = a;
}

/* ------------------------------------------------------------------ */
/* tables                                                               */
/* ------------------------------------------------------------------ */

/* utility_pack.f:725 (mxgrid=max(1000,int(rvdw/0.01+0.5)+4)) */
int orc_mxgrid(double rvdw) {
  int m = (int)(rvdw / 0.01 + 0.5) + 4;
  return m > 1000 ? m : 1000;
}

/* ewald.f:1-30 (erfcgen); A&S 7.1.26 constants at ewald.f:9-10 */
double orc_erfcgen(double alpha, double rcut, int mxgrid, double *erc) {
  const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
  const double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
  double drewd = rcut / (double)(mxgrid - 4);
  for (int i = 1; i <= mxgrid; i++) {
    double rrr = (double)i * drewd;
    double tt = 1.0 / (1.0 + pp * alpha * rrr);
    double ar = alpha * rrr;
    double exp1 = exp(-(ar * ar));
    erc[i] = tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * exp1 /
             rrr;
  }
  erc[0] = erc[1];
  return drewd;
}

static double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

/* vdw_terms.f:189-396 (forgen) for keys 1..9; table ntpvdw is all zeros
 * (vdw_terms.f:160-182) */
double orc_forgen(int ntpvdw, const int *keys, const double *prm, double rvdw,
                  int mxgrid, double *vvv) {
  double dlrpot = rvdw / (double)(mxgrid - 4);
  int W = mxgrid + 1;
  for (int t = 0; t <= ntpvdw; t++)
    for (int i = 0; i <= mxgrid; i++) vvv[t * W + i] = 0.0;
  for (int t = 0; t < ntpvdw; t++) {
    const double *p = prm + 5 * t;
    double *v = vvv + t * W;
    int key = keys[t];
    for (int i = 1; i <= mxgrid; i++) {
      double rrr = (double)i * dlrpot;
      double val = 0.0;
      if (key == 1) {
        double r6 = ipow(rrr, 6);
        val = (p[0] / r6 - p[1]) / r6;
      } else if (key == 2) {
        double sr6 = ipow(p[1] / rrr, 6);
        val = 4.0 * p[0] * sr6 * (sr6 - 1.0);
      } else if (key == 3) {
        double ann = fmax(p[1], p[2]), amm = fmin(p[1], p[2]);
        val = p[0] / (ann - amm) *
              (amm * pow(p[3] / rrr, ann) - ann * pow(p[3] / rrr, amm));
      } else if (key == 4) {
        val = p[0] * exp(-rrr / p[1]) - p[2] / ipow(rrr, 6);
      } else if (key == 5) {
        val = p[0] * exp(p[1] * (p[2] - rrr)) - p[3] / ipow(rrr, 6) -
              p[4] / ipow(rrr, 8);
      } else if (key == 6) {
        val = p[0] / ipow(rrr, 12) - p[1] / ipow(rrr, 10);
      } else if (key == 7) {
        double eps = p[0], ann = p[1], amm = p[2], rr0 = p[3], rrc = p[4];
        if (rrc < 1e-6) rrc = rvdw;
        double gam = rrc / rr0;
        double bet = gam * pow((pow(gam, amm + 1.0) - 1.0) /
                                   (pow(gam, ann + 1.0) - 1.0),
                               1.0 / (ann - amm));
        eps = -eps * (ann - amm) /
              (amm * pow(bet, ann) * (1.0 + (ann / gam - ann - 1.0) / pow(gam, ann)) -
               ann * pow(bet, amm) * (1.0 + (amm / gam - amm - 1.0) / pow(gam, amm)));
        if (rrr > rrc) {
          val = 0.0;
        } else {
          val = eps / (ann - amm) *
                (amm * pow(bet, ann) * (pow(rr0 / rrr, ann) - pow(1.0 / gam, ann)) -
                 ann * pow(bet, amm) * (pow(rr0 / rrr, amm) - pow(1.0 / gam, amm)) +
                 ann * amm * ((rrr / (gam * rr0) - 1.0) *
                              (pow(bet / gam, ann) - pow(bet / gam, amm))));
        }
      } else if (key == 8) {
        double e1 = 1.0 - exp(-p[2] * (rrr - p[1]));
        val = p[0] * (e1 * e1 - 1.0);
      } else if (key == 9) {
        double eps = p[0], sig = p[1], rr0 = p[2];
        double ddd = sig * pow(2.0, 1.0 / 6.0);
        double r = rrr - rr0;
        if (r > ddd) {
          val = 0.0;
        } else if (r > 0.0) {
          double sr6 = ipow(sig / r, 6);
          val = 4.0 * eps * sr6 * (sr6 - 1.0) + eps;
        } else {
          val = 0.0; /* reference leaves this entry unset */
        }
      }
      v[i] = val;
    }
    v[0] = v[1];
  }
  return dlrpot;
}

/* ------------------------------------------------------------------ */
/* reciprocal space                                                     */
/* ------------------------------------------------------------------ */

/* the spherical cut of ewald.f:77-84 */
static double orc_rcpct2(const double cell[9], const int kmax[3],
                         double rcell[9]) {
  double det, buf[10];
  orc_invert(cell, rcell, &det);
  orc_dcell(rcell, buf);
  double rcpcut = fmin(fmin((double)kmax[0] * buf[6], (double)kmax[1] * buf[7]),
                       (double)kmax[2] * buf[8]);
  rcpcut = rcpcut * 1.05 * (2.0 * ORC_PI);
  return rcpcut * rcpcut;
}

int orc_kvec_count(const double cell[9], const int kmax[3]) {
  double rcell[9];
  double rcpct2 = orc_rcpct2(cell, kmax, rcell);
  const double twopi = 2.0 * ORC_PI;
  int kkk = 0, mmin = 0, nmin = 1;
  for (int ll = 0; ll <= kmax[0]; ll++) {
    double tmp = twopi * (double)ll;
    double rkx1 = tmp * rcell[0], rky1 = tmp * rcell[3], rkz1 = tmp * rcell[6];
    for (int mm = mmin; mm <= kmax[1]; mm++) {
      tmp = twopi * (double)mm;
      double rkx2 = rkx1 + tmp * rcell[1], rky2 = rky1 + tmp * rcell[4],
             rkz2 = rkz1 + tmp * rcell[7];
      for (int nn = nmin; nn <= kmax[2]; nn++) {
        tmp = twopi * (double)nn;
        double rkx3 = rkx2 + tmp * rcell[2], rky3 = rky2 + tmp * rcell[5],
               rkz3 = rkz2 + tmp * rcell[8];
        double rksq = rkx3 * rkx3 + rky3 * rky3 + rkz3 * rkz3;
        if (rksq <= rcpct2) kkk++;
      }
      nmin = -kmax[2];
    }
    mmin = -kmax[1];
  }
  return kkk;
}

/* Shared k-loop of ewald.f:97-331 (ewald1_guest) and :450-616 (ewald1).
 * Computes per k-vector the structure factor of the n atoms, and hands
 * (kkk, ckcs, ckss, akk) to the callback-free consumer through arrays. */
typedef struct {
  double *elc0, *els0, *elc1, *els1; /* x: running power + base */
  double *emc, *ems;                 /* y: [i*(kmax2+1)+l] */
  double *enc, *ens;                 /* z: [i*(kmax3+1)+l] */
  double *clm, *slm;
} orc_phase_ws;

static void orc_phase_alloc(orc_phase_ws *w, int n, const int kmax[3]) {
  w->elc0 = malloc(sizeof(double) * n);
  w->els0 = malloc(sizeof(double) * n);
  w->elc1 = malloc(sizeof(double) * n);
  w->els1 = malloc(sizeof(double) * n);
  w->emc = malloc(sizeof(double) * n * (kmax[1] + 1));
  w->ems = malloc(sizeof(double) * n * (kmax[1] + 1));
  w->enc = malloc(sizeof(double) * n * (kmax[2] + 1));
  w->ens = malloc(sizeof(double) * n * (kmax[2] + 1));
  w->clm = malloc(sizeof(double) * n);
  w->slm = malloc(sizeof(double) * n);
}
static void orc_phase_free(orc_phase_ws *w) {
  free(w->elc0); free(w->els0); free(w->elc1); free(w->els1);
  free(w->emc); free(w->ems); free(w->enc); free(w->ens);
  free(w->clm); free(w->slm);
}

/* returns number of k-vectors; fills ckcs/ckss/akk per k (length Nk) */
static int orc_kloop(int imcon, const double cell[9], double alpha,
                     const int kmax[3], int n, const double *x,
                     const double *y, const double *z, const double *q,
                     double *out_c, double *out_s, double *out_a) {
  const double twopi = 2.0 * ORC_PI;
  double rcell[9];
  double rcpct2 = orc_rcpct2(cell, kmax, rcell);
  double ralph = -0.25 / (alpha * alpha);
  int lconsw = !(imcon == 4 || imcon == 5 || imcon == 7);
  const int K2 = kmax[1] + 1, K3 = kmax[2] + 1;
  orc_phase_ws w;
  orc_phase_alloc(&w, n > 0 ? n : 1, kmax);
  for (int i = 0; i < n; i++) {
    w.elc0[i] = 1.0; w.els0[i] = 0.0;
    w.emc[i * K2] = 1.0; w.ems[i * K2] = 0.0;
    w.enc[i * K3] = 1.0; w.ens[i * K3] = 0.0;
    double ssx = rcell[0] * x[i] + rcell[3] * y[i] + rcell[6] * z[i];
    double ssy = rcell[1] * x[i] + rcell[4] * y[i] + rcell[7] * z[i];
    double ssz = rcell[2] * x[i] + rcell[5] * y[i] + rcell[8] * z[i];
    w.elc1[i] = cos(twopi * ssx);
    w.emc[i * K2 + 1] = cos(twopi * ssy);
    w.enc[i * K3 + 1] = cos(twopi * ssz);
    w.els1[i] = sin(twopi * ssx);
    w.ems[i * K2 + 1] = sin(twopi * ssy);
    w.ens[i * K3 + 1] = sin(twopi * ssz);
  }
  for (int l = 2; l <= kmax[1]; l++)
    for (int i = 0; i < n; i++) {
      double *c = w.emc + i * K2, *s = w.ems + i * K2;
      c[l] = c[l - 1] * c[1] - s[l - 1] * s[1];
      s[l] = s[l - 1] * c[1] + c[l - 1] * s[1];
    }
  for (int l = 2; l <= kmax[2]; l++)
    for (int i = 0; i < n; i++) {
      double *c = w.enc + i * K3, *s = w.ens + i * K3;
      c[l] = c[l - 1] * c[1] - s[l - 1] * s[1];
      s[l] = s[l - 1] * c[1] + c[l - 1] * s[1];
    }
  int kkk = 0, mmin = 0, nmin = 1;
  for (int ll = 0; ll <= kmax[0]; ll++) {
    double tmp = twopi * (double)ll;
    double rkx1 = tmp * rcell[0], rky1 = tmp * rcell[3], rkz1 = tmp * rcell[6];
    if (ll == 1) {
      for (int i = 0; i < n; i++) { w.elc0[i] = w.elc1[i]; w.els0[i] = w.els1[i]; }
    } else if (ll > 1) {
      for (int i = 0; i < n; i++) {
        double cs = w.elc0[i];
        w.elc0[i] = cs * w.elc1[i] - w.els0[i] * w.els1[i];
        w.els0[i] = w.els0[i] * w.elc1[i] + cs * w.els1[i];
      }
    }
    for (int mm = mmin; mm <= kmax[1]; mm++) {
      int m = abs(mm);
      tmp = twopi * (double)mm;
      double rkx2 = rkx1 + tmp * rcell[1], rky2 = rky1 + tmp * rcell[4],
             rkz2 = rkz1 + tmp * rcell[7];
      if (mm >= 0) {
        for (int i = 0; i < n; i++) {
          double ec = w.emc[i * K2 + m], es = w.ems[i * K2 + m];
          w.clm[i] = w.elc0[i] * ec - w.els0[i] * es;
          w.slm[i] = w.els0[i] * ec + es * w.elc0[i];
        }
      } else {
        for (int i = 0; i < n; i++) {
          double ec = w.emc[i * K2 + m], es = w.ems[i * K2 + m];
          w.clm[i] = w.elc0[i] * ec + w.els0[i] * es;
          w.slm[i] = w.els0[i] * ec - es * w.elc0[i];
        }
      }
      for (int nn = nmin; nn <= kmax[2]; nn++) {
        int nabs = abs(nn);
        tmp = twopi * (double)nn;
        double rkx3 = rkx2 + tmp * rcell[2], rky3 = rky2 + tmp * rcell[5],
               rkz3 = rkz2 + tmp * rcell[8];
        double rksq = rkx3 * rkx3 + rky3 * rky3 + rkz3 * rkz3;
        if (rksq <= rcpct2) {
          double ckcs = 0.0, ckss = 0.0;
          for (int i = 0; i < n; i++) {
            double tchge = q[i], tclm = w.clm[i], tslm = w.slm[i];
            double tenc = w.enc[i * K3 + nabs], tens = w.ens[i * K3 + nabs];
            double ckc, cks;
            if (nn >= 0) {
              ckc = tchge * (tclm * tenc - tslm * tens);
              cks = tchge * (tslm * tenc + tclm * tens);
            } else {
              ckc = tchge * (tclm * tenc + tslm * tens);
              cks = tchge * (tslm * tenc - tclm * tens);
            }
            ckcs = ckcs + ckc;
            ckss = ckss + cks;
          }
          double rrksq = 1.0 / rksq;
          double akk = lconsw ? exp(ralph * rksq) * rrksq
                              : 4.0 * exp(ralph * rksq) * rrksq;
          out_c[kkk] = ckcs;
          out_s[kkk] = ckss;
          out_a[kkk] = akk;
          kkk++;
        }
      }
      nmin = -kmax[2];
    }
    mmin = -kmax[1];
  }
  orc_phase_free(&w);
  return kkk;
}

/* ewald.f:368-630 (ewald1) */
double orc_ewald1(int imcon, const double cell[9], double alpha,
                  const int kmax[3], double epsq, int n, const double *x,
                  const double *y, const double *z, const double *q,
                  double *ckc, double *cks, int *newld, double *engsic_out) {
  const double twopi = 2.0 * ORC_PI;
  double volm = orc_volume(imcon, cell);
  double rvolm = twopi / volm;
  int lconsw = 1;
  if (imcon == 4 || imcon == 5 || imcon == 7) { rvolm = 0.5 * rvolm; lconsw = 0; }
  double engsic = 0.0;
  for (int i = 0; i < n; i++) engsic = engsic + q[i] * q[i];
  engsic = -ORC_R4PIE0 / epsq * alpha * engsic / ORC_SQRPI;
  int nk = orc_kvec_count(cell, kmax);
  double *akk = malloc(sizeof(double) * (nk > 0 ? nk : 1));
  int got = orc_kloop(imcon, cell, alpha, kmax, n, x, y, z, q, ckc, cks, akk);
  double engcpe = 0.0;
  for (int k = 0; k < got; k++)
    engcpe = engcpe + akk[k] * (ckc[k] * ckc[k] + cks[k] * cks[k]);
  free(akk);
  if (newld) *newld = got;
  if (engsic_out) *engsic_out = engsic;
  if (lconsw) return 2.0 * rvolm * ORC_R4PIE0 * engcpe / epsq + engsic;
  return rvolm * ORC_R4PIE0 * engcpe / epsq + engsic;
}

/* ewald.f:32-365 (ewald1_guest) */
double orc_ewald1_guest(int imcon, const double cell[9], double alpha,
                        const int kmax[3], double epsq, int natms,
                        const double *x, const double *y, const double *z,
                        const double *q, int kind, int pass,
                        const double *ckcsum, const double *ckssum,
                        double *ckcsnew, double *ckssnew) {
  const double twopi = 2.0 * ORC_PI;
  double volm = orc_volume(imcon, cell);
  double rvolm = twopi / volm;
  int lconsw = 1;
  if (imcon == 4 || imcon == 5 || imcon == 7) { rvolm = 0.5 * rvolm; lconsw = 0; }
  double engsic = 0.0;
  for (int i = 0; i < natms; i++) engsic = engsic + q[i] * q[i];
  engsic = -ORC_R4PIE0 / epsq * alpha * engsic / ORC_SQRPI;
  int nk = orc_kvec_count(cell, kmax);
  double *sc = malloc(sizeof(double) * (nk > 0 ? nk : 1) * 3);
  double *ss = sc + nk, *akk = sc + 2 * nk;
  orc_kloop(imcon, cell, alpha, kmax, natms, x, y, z, q, sc, ss, akk);
  double engcpe = 0.0;
  for (int k = 0; k < nk; k++) {
    double ckcs = sc[k], ckss = ss[k];
    double ckcold = ckcsum[k], cksold = ckssum[k];
    if (kind == 1) {
      ckcs = ckcs + ckcold;
      ckss = ckss + cksold;
      ckcsnew[k] = ckcs;
      ckssnew[k] = ckss;
    } else if (kind == 2) {
      ckcs = ckcold - ckcs;
      ckss = cksold - ckss;
      ckcsnew[k] = ckcs;
      ckssnew[k] = ckss;
    } else {
      if (pass == 1) {
        ckcsnew[k] = ckcold - ckcs;
        ckssnew[k] = cksold - ckss;
      } else {
        double ckcpass2 = ckcsnew[k], ckspass2 = ckssnew[k];
        ckcs = ckcpass2 + ckcs;
        ckss = ckspass2 + ckss;
        ckcsnew[k] = ckcs;
        ckssnew[k] = ckss;
      }
    }
    if (kind == 3) {
      if (pass == 2)
        engcpe = engcpe + akk[k] * (ckcs * ckcs + ckss * ckss -
                                    ckcold * ckcold - cksold * cksold);
    } else {
      engcpe = engcpe + akk[k] * (ckcs * ckcs + ckss * ckss -
                                  ckcold * ckcold - cksold * cksold);
    }
  }
  free(sc);
  double f = lconsw ? 2.0 * rvolm : rvolm;
  if (kind == 3) return f * ORC_R4PIE0 * engcpe / epsq;
  if (kind == 2) return f * ORC_R4PIE0 * engcpe / epsq - engsic;
  return f * ORC_R4PIE0 * engcpe / epsq + engsic;
}

/* ------------------------------------------------------------------ */
/* real space                                                           */
/* ------------------------------------------------------------------ */

/* ewald.f:632-694 (ewald2) */
double orc_ewald2(double chg, int n, const int *ilist, const double *rsq,
                  const double *atmcharge, const double *erc, double drewd,
                  double rcut, double epsq) {
  double rcsq = rcut * rcut;
  double rdrewd = 1.0 / drewd;
  double engcpe = 0.0;
  double chgea = chg / epsq * ORC_R4PIE0;
  for (int i = 0; i < n; i++) {
    int jatm = ilist[i];
    double chgprd = chgea * atmcharge[jatm];
    if (fabs(chgprd) > 1e-10) {
      double r2 = rsq[i];
      if (rcsq > r2) {
        double rrr = sqrt(r2);
        int ll = (int)(rrr * rdrewd);
        double ppp = rrr * rdrewd - (double)ll;
        double vk0 = erc[ll], vk1 = erc[ll + 1], vk2 = erc[ll + 2];
        double t1 = vk0 + (vk1 - vk0) * ppp;
        double t2 = vk1 + (vk2 - vk1) * (ppp - 1.0);
        double erfcr = (t1 + (t2 - t1) * ppp * 0.5) * chgprd;
        engcpe = engcpe + erfcr;
      }
    }
  }
  return engcpe;
}

/* vdw_terms.f:398-498 (srfrce) */
double orc_srfrce(int sittyp, int n, const int *ilist, const double *rsq,
                  const int *ltype, const int *lstvdw, const int *ltpvdw,
                  const double *vvv, int mxgrid, double rcut, double dlrpot) {
  double rcsq = rcut * rcut;
  double rdr = 1.0 / dlrpot;
  double engsrp = 0.0;
  double ai = (double)sittyp;
  const int W = mxgrid + 1;
  for (int m = 0; m < n; m++) {
    int jatm = ilist[m];
    double aj = (double)ltype[jatm];
    double ab;
    if (ai > aj)
      ab = ai * (ai - 1.0) * 0.5 + aj + 0.5;
    else
      ab = aj * (aj - 1.0) * 0.5 + ai + 0.5;
    int k = lstvdw[(int)ab - 1]; /* 1-based table index */
    const double *v = vvv + (k - 1) * W;
    if ((ltpvdw[k - 1] < 100) && (fabs(v[1]) > 1e-10)) {
      double r2 = rsq[m];
      if (rcsq > r2) {
        double rrr = sqrt(r2);
        int l = (int)(rrr * rdr);
        double ppp = rrr * rdr - (double)l;
        double vk0 = v[l], vk1 = v[l + 1], vk2 = v[l + 2];
        double t1 = vk0 + (vk1 - vk0) * ppp;
        double t2 = vk1 + (vk2 - vk1) * (ppp - 1.0);
        double omega = t1 + (t2 - t1) * ppp * 0.5;
        engsrp = omega + engsrp;
      }
    }
  }
  return engsrp;
}

/* ewald.f:697-758 (ewald3) */
double orc_ewald3(double chg, int n, const double *chgj, const double *dx,
                  const double *dy, const double *dz, double alpha,
                  double epsq) {
  const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
  const double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
  const double rr3 = 0.333333333333, r10 = 0.1, r42 = 0.02380952381,
               r216 = 4.62962962963e-3;
  double engcpe = 0.0;
  double chgea = chg / epsq * ORC_R4PIE0;
  for (int m = 0; m < n; m++) {
    double chgprd = chgea * chgj[m];
    double rsq = dx[m] * dx[m] + dy[m] * dy[m] + dz[m] * dz[m];
    double rrr = sqrt(rsq);
    double alpr = rrr * alpha;
    double alpr2 = alpr * alpr;
    double erfr;
    if (alpr < 1e-2) {
      erfr = 2.0 * chgprd * (alpha / ORC_SQRPI) *
             (1.0 + alpr2 * (-rr3 + alpr2 * (r10 + alpr2 * (-r42 + alpr2 * r216))));
    } else {
      double tt = 1.0 / (1.0 + pp * alpha * rrr);
      double ar = alpha * rrr;
      double exp1 = exp(-(ar * ar));
      erfr = (1.0 - tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * exp1) *
             chgprd / rrr;
    }
    engcpe = engcpe - erfr;
  }
  return engcpe;
}

/* gcmc.f:529-552 */
double orc_ewald3_guest(int imcon, const double cell[9], int natms,
                        const double *gx, const double *gy, const double *gz,
                        const double *q, double alpha, double epsq) {
  double sum = 0.0;
  double dx[64], dy[64], dz[64], cj[64];
  for (int i = 0; i < natms - 1; i++) {
    int ik = 0;
    for (int j = i + 1; j < natms; j++) {
      dx[ik] = gx[i] - gx[j];
      dy[ik] = gy[i] - gy[j];
      dz[ik] = gz[i] - gz[j];
      cj[ik] = q[j];
      ik++;
    }
    orc_images(imcon, ik, cell, dx, dy, dz);
    sum = sum + orc_ewald3(q[i], ik, cj, dx, dy, dz, alpha, epsq);
  }
  return sum;
}

/* ------------------------------------------------------------------ */
/* long-range tail                                                      */
/* ------------------------------------------------------------------ */

/* vdw_terms.f:500-682 (lrcorrect).  prm indexed by table (0-based), 5 each;
 * ltpvdw by table (0-based); lstvdw by loc8-1. */
double orc_lrcorrect(int imcon, int ntpatm, const int *numtyp,
                     const int *numfrz, const int *lstvdw, const int *ltpvdw,
                     const double *prm, double rcut, double volm,
                     double *steadd) {
  const double twopi = 2.0 * ORC_PI;
  double elrc = 0.0;
  int ivdw = 0;
  if (imcon == 0 || imcon == 6) return 0.0;
  for (int i = 1; i <= ntpatm; i++) {
    for (int j = 1; j <= i; j++) {
      double eadd = 0.0;
      int k = lstvdw[ivdw];
      int key = ltpvdw[k - 1];
      const double *p = prm + 5 * (k - 1);
      if (key == 0) {
        eadd = p[0];
      } else if (key == 1) {
        eadd = p[0] / (9.0 * ipow(rcut, 9)) - p[1] / (3.0 * ipow(rcut, 3));
      } else if (key == 2) {
        double eps = p[0], sig = p[1];
        eadd = 4.0 * eps * (ipow(sig, 12) / (9.0 * ipow(rcut, 9)) -
                            ipow(sig, 6) / (3.0 * ipow(rcut, 3)));
      } else if (key == 3) {
        double eps = p[0], ann = p[1], amm = p[2], rr0 = p[3];
        eadd = eps / (ann - amm) *
               (amm * pow(rr0, ann) / ((ann - 3.0) * pow(rcut, ann - 3.0)) -
                ann * pow(rr0, amm) / ((amm - 3.0) * pow(rcut, amm - 3.0)));
      } else if (key == 4) {
        eadd = -p[2] / (3.0 * ipow(rcut, 3));
      } else if (key == 5) {
        eadd = -p[3] / (3.0 * ipow(rcut, 3)) - p[4] / (5.0 * ipow(rcut, 5));
      } else if (key == 6) {
        eadd = p[0] / (9.0 * ipow(rcut, 9)) - p[1] / (7.0 * ipow(rcut, 7));
      }
      if (i != j) eadd = eadd * 2.0;
      steadd[ivdw] = eadd;
      double denprd = twopi *
                      ((double)numtyp[i - 1] * (double)numtyp[j - 1] -
                       (double)numfrz[i - 1] * (double)numfrz[j - 1]) /
                      (volm * volm);
      elrc = elrc + volm * denprd * eadd;
      ivdw++;
    }
  }
  return elrc;
}

/* vdw_terms.f:684-745 (gstlrcorrect) */
double orc_gstlrcorrect(int imcon, int ntpatm, const int *numtyp,
                        const int *numfrz, const double *steadd, double volm) {
  const double twopi = 2.0 * ORC_PI;
  double delrc = 0.0;
  int ivdw = 0;
  if (imcon == 0 || imcon == 6) return 0.0;
  for (int i = 1; i <= ntpatm; i++)
    for (int j = 1; j <= i; j++) {
      double eadd = steadd[ivdw++];
      double denprd = twopi *
                      ((double)numtyp[i - 1] * (double)numtyp[j - 1] -
                       (double)numfrz[i - 1] * (double)numfrz[j - 1]) /
                      (volm * volm);
      delrc = delrc + volm * denprd * eadd;
    }
  return delrc;
}

/* ------------------------------------------------------------------ */
/* per-move neighbour list                                              */
/* ------------------------------------------------------------------ */

/* utility_pack.f:3113-3177 (guestlistgen).  ind holds 0-based condensed
 * indices of the moving molecule's atoms, or -1. */
void orc_guestlistgen(int imcon, const double cell[9], double rcut,
                      double delr, int natms, const double *nx,
                      const double *ny, const double *nz, const int *ind,
                      int totatm, const double *xxx, const double *yyy,
                      const double *zzz, int *gstlentry, int *gstlist) {
  double rclim = (rcut + delr) * (rcut + delr);
  double *dx = malloc(sizeof(double) * (totatm > 0 ? totatm : 1) * 3);
  double *dy = dx + totatm, *dz = dy + totatm;
  int *jj = malloc(sizeof(int) * (totatm > 0 ? totatm : 1));
  for (int i = 0; i < natms; i++) {
    int ii = 0;
    gstlentry[i] = 0;
    for (int j = 0; j < totatm; j++) {
      int chk = 1;
      for (int t = 0; t < natms; t++)
        if (j == ind[t]) chk = 0;
      if (chk) {
        dx[ii] = nx[i] - xxx[j];
        dy[ii] = ny[i] - yyy[j];
        dz[ii] = nz[i] - zzz[j];
        jj[ii] = j;
        ii++;
      }
    }
    orc_images(imcon, ii, cell, dx, dy, dz);
    for (int t = 0; t < ii; t++) {
      double rsq;
      if (imcon == 6)
        rsq = dx[t] * dx[t] + dy[t] * dy[t];
      else
        rsq = dx[t] * dx[t] + dy[t] * dy[t] + dz[t] * dz[t];
      if (rsq <= rclim) {
        gstlist[(size_t)i * totatm + gstlentry[i]] = jj[t];
        gstlentry[i]++;
      }
    }
  }
  free(dx);
  free(jj);
}

/* ------------------------------------------------------------------ */
/* RNG                                                                  */
/* ------------------------------------------------------------------ */

/* gcmc.f:2760-2784 (duni set-up) */
void orc_duni_init(orc_duni *s, int i, int j, int k, int l) {
  s->ir = 97;
  s->jr = 33;
  for (int ii = 0; ii < 97; ii++) {
    float sv = 0.0f, t = 0.5f;
    for (int jj = 0; jj < 24; jj++) {
      int m = ((i * j) % 179) * k % 179;
      i = j;
      j = k;
      k = m;
      l = (53 * l + 1) % 169;
      if ((l * m) % 64 >= 32) sv = sv + t;
      t = 0.5f * t;
    }
    s->u[ii] = sv;
  }
  s->c = 362436.0f / 16777216.0f;
  s->cd = 7654321.0f / 16777216.0f;
  s->cm = 16777213.0f / 16777216.0f;
}

/* gcmc.f:2786-2799 (duni draw) */
double orc_duni_next(orc_duni *s) {
  float uni = s->u[s->ir - 1] - s->u[s->jr - 1];
  if (uni < 0.0f) uni = uni + 1.0f;
  s->u[s->ir - 1] = uni;
  s->ir = s->ir - 1;
  if (s->ir == 0) s->ir = 97;
  s->jr = s->jr - 1;
  if (s->jr == 0) s->jr = 97;
  s->c = s->c - s->cd;
  if (s->c < 0.0f) s->c = s->c + s->cm;
  uni = uni - s->c;
  if (uni < 0.0f) uni = uni + 1.0f;
  return (double)uni;
}

/* gcmc.f:2753-2758 */
void orc_duni_seed_from_clock(int seconds, int millis, int ijkl[4]) {
  ijkl[0] = (millis * 3) % 177 + 1;
  ijkl[1] = (seconds * 23) % 177 + 1;
  ijkl[2] = (millis * 5) % 177 + 1;
  ijkl[3] = (seconds * 3) % 168;
}

/* ------------------------------------------------------------------ */
/* acceptance                                                           */
/* ------------------------------------------------------------------ */

/* utility_pack.f:3049-3095 (energy_eval) */
int orc_energy_eval(double eng, double rande, double volm, double press,
                    int ngsts, double temp, double beta, int kind) {
  double test = 0.0;
  if (-1.0 * beta * eng > 700.0) {
    test = exp(700.0);
  } else if (kind == 3) {
    test = exp(-1.0 * beta * eng);
  } else if (kind == 1) {
    test = volm * press / ((double)(ngsts + 1) * ORC_BOLTZ * temp) *
           exp(-1.0 * beta * eng);
  } else if (kind == 2) {
    if (fabs(press) < 1e-7)
      test = 1.0;
    else
      test = (double)ngsts * ORC_BOLTZ * temp / (volm * press) *
             exp(-1.0 * beta * eng);
  }
  return rande < test;
}

/* ------------------------------------------------------------------ */
/* proposals                                                            */
/* ------------------------------------------------------------------ */

/* utility_pack.f:2870-2906 (rotation) */
void orc_rotation(double *x, double *y, double *z, int n, double q1,
                  double q2, double q3, double q4) {
  double Q[9];
  Q[0] = 2.0 * q2 * q2 + 2.0 * q1 * q1 - 1.0;
  Q[1] = 2.0 * q2 * q3 - 2.0 * q1 * q4;
  Q[2] = 2.0 * q2 * q4 + 2.0 * q1 * q3;
  Q[3] = 2.0 * q2 * q3 + 2.0 * q1 * q4;
  Q[4] = 2.0 * q1 * q1 + 2.0 * q3 * q3 - 1.0;
  Q[5] = 2.0 * q3 * q4 - 2.0 * q1 * q2;
  Q[6] = 2.0 * q2 * q4 - 2.0 * q1 * q3;
  Q[7] = 2.0 * q3 * q4 + 2.0 * q1 * q2;
  Q[8] = 2.0 * q4 * q4 + 2.0 * q1 * q1 - 1.0;
  for (int i = 0; i < n; i++) {
    double tx = Q[0] * x[i] + Q[1] * y[i] + Q[2] * z[i];
    double ty = Q[3] * x[i] + Q[4] * y[i] + Q[5] * z[i];
    double tz = Q[6] * x[i] + Q[7] * y[i] + Q[8] * z[i];
    x[i] = tx; y[i] = ty; z[i] = tz;
  }
}

/* quaternion from (theta, r2, r3, r4) as random_ins/random_rot build it */
static void orc_quat(double theta, double r2, double r3, double r4,
                     double q[4]) {
  double beta1 = 2.0 * r2 - 1.0, beta2 = 2.0 * r3 - 1.0, beta3 = 2.0 * r4 - 1.0;
  double sintheta = sin(theta / 2.0);
  double norm = sqrt(beta1 * beta1 + beta2 * beta2 + beta3 * beta3);
  double u1 = beta1 / norm, u2 = beta2 / norm, u3 = beta3 / norm;
  q[0] = cos(theta / 2.0);
  q[1] = sintheta * u1;
  q[2] = sintheta * u2;
  q[3] = sintheta * u3;
}

/* utility_pack.f:2545-2650 (random_ins) with cartesian (:3097-3111) */
void orc_random_ins(const double cell[9], int natms, const double *gx,
                    const double *gy, const double *gz, double ra, double rb,
                    double rc, double r1, double r2, double r3, double r4,
                    double *nx, double *ny, double *nz) {
  double xc = ra * cell[0] + rb * cell[3] + rc * cell[6];
  double yc = ra * cell[1] + rb * cell[4] + rc * cell[7];
  double zc = ra * cell[2] + rb * cell[5] + rc * cell[8];
  for (int i = 0; i < natms; i++) { nx[i] = gx[i]; ny[i] = gy[i]; nz[i] = gz[i]; }
  double theta = r1 * 2.0 * ORC_PI;
  double q[4];
  orc_quat(theta, r2, r3, r4, q);
  orc_rotation(nx, ny, nz, natms, q[0], q[1], q[2], q[3]);
  for (int i = 0; i < natms; i++) {
    nx[i] = nx[i] + xc;
    ny[i] = ny[i] + yc;
    nz[i] = nz[i] + zc;
  }
}

/* utility_pack.f:3002-3034 (com) */
static void orc_com(int natms, const double *mass, const double *x,
                    const double *y, const double *z, double *cx, double *cy,
                    double *cz) {
  double w = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  for (int i = 0; i < natms; i++) {
    double m = mass[i];
    w = w + m;
    sx = sx + x[i] * m;
    sy = sy + y[i] * m;
    sz = sz + z[i] * m;
  }
  *cx = sx / w; *cy = sy / w; *cz = sz / w;
}

/* utility_pack.f:2679-2763 (random_disp) + 2929-3000 (random_rot) */
void orc_random_disp(const double cell[9], const double rcell[9], int natms,
                     const double *mass, double *nx, double *ny, double *nz,
                     double delr, double rotangle, double ra, double rb,
                     double rc, double r1, double r2, double r3, double r4) {
  double halfdelr = delr * 0.5;
  ra = (2.0 * ra - 1.0) * halfdelr;
  rb = (2.0 * rb - 1.0) * halfdelr;
  rc = (2.0 * rc - 1.0) * halfdelr;
  for (int i = 0; i < natms; i++) {
    nx[i] = ra + nx[i];
    ny[i] = rb + ny[i];
    nz[i] = rc + nz[i];
  }
  double comx, comy, comz;
  orc_com(natms, mass, nx, ny, nz, &comx, &comy, &comz);
  double ssx = (rcell[0] * comx + rcell[3] * comy + rcell[6] * comz);
  double ssy = (rcell[1] * comx + rcell[4] * comy + rcell[7] * comz);
  double ssz = (rcell[2] * comx + rcell[5] * comy + rcell[8] * comz);
  double xss = ssx - floor(ssx), yss = ssy - floor(ssy), zss = ssz - floor(ssz);
  double x = (cell[0] * xss + cell[3] * yss + cell[6] * zss);
  double y = (cell[1] * xss + cell[4] * yss + cell[7] * zss);
  double z = (cell[2] * xss + cell[5] * yss + cell[8] * zss);
  double sx = x - comx, sy = y - comy, sz = z - comz;
  comx = x; comy = y; comz = z;
  for (int i = 0; i < natms; i++) {
    nx[i] = nx[i] + sx;
    ny[i] = ny[i] + sy;
    nz[i] = nz[i] + sz;
  }
  /* random_rot */
  for (int i = 0; i < natms; i++) {
    nx[i] = nx[i] - comx;
    ny[i] = ny[i] - comy;
    nz[i] = nz[i] - comz;
  }
  double theta = r1 * rotangle;
  double q[4];
  orc_quat(theta, r2, r3, r4, q);
  orc_rotation(nx, ny, nz, natms, q[0], q[1], q[2], q[3]);
  for (int i = 0; i < natms; i++) {
    nx[i] = nx[i] + comx;
    ny[i] = ny[i] + comy;
    nz[i] = nz[i] + comz;
  }
}

/* ------------------------------------------------------------------ */
/* system state                                                         */
/* ------------------------------------------------------------------ */

#define ORC_MAXG 8

typedef struct {
  int ns;
  double tx[64], ty[64], tz[64], q[64], mass[64];
  int type[64];
  double e3;
  int nmol, cap;
  double *x, *y, *z; /* [mol*ns + site] */
} orc_gtype;

struct orc_sys {
  int imcon;
  double cell[9], rcell[9], volm;
  double alpha, epsq;
  int kmax[3], nk;
  double rcut, delr, rvdw;
  int mxgrid;
  double drewd, dlrpot;
  double *erc, *vvv;
  int ntpatm, ntpvdw, ntab;
  int *lstvdw, *ltpvdw;
  double *prm;
  double *steadd;
  int nfw;
  double *fx, *fy, *fz, *fq;
  int *ftype, *ffrz;
  int ng;
  orc_gtype g[ORC_MAXG];
  /* condensed (guests first, then framework) */
  int totatm, cap_all;
  double *xxx, *yyy, *zzz, *chg;
  int *ltype, *lfrz;
  double *ckcsum, *ckssum, *ckcsnew, *ckssnew;
  int *gstlentry, *gstlist;
  double *rsqbuf, *dxb, *dyb, *dzb;
  int *ilb;
  int finalized;
  /* pending trial */
  int pend_kind, pend_g, pend_mol;
  double px[64], py[64], pz[64];
};

orc_sys *orc_sys_create(void) {
  orc_sys *s = calloc(1, sizeof(orc_sys));
  s->epsq = 1.0;
  return s;
}

void orc_sys_destroy(orc_sys *s) {
  if (!s) return;
  free(s->erc); free(s->vvv); free(s->lstvdw); free(s->ltpvdw); free(s->prm);
  free(s->steadd);
  free(s->fx); free(s->fy); free(s->fz); free(s->fq); free(s->ftype); free(s->ffrz);
  for (int t = 0; t < s->ng; t++) { free(s->g[t].x); free(s->g[t].y); free(s->g[t].z); }
  free(s->xxx); free(s->yyy); free(s->zzz); free(s->chg); free(s->ltype); free(s->lfrz);
  free(s->ckcsum); free(s->ckssum); free(s->ckcsnew); free(s->ckssnew);
  free(s->gstlentry); free(s->gstlist); free(s->rsqbuf); free(s->dxb);
  free(s->dyb); free(s->dzb); free(s->ilb);
  free(s);
}

int orc_sys_set_cell(orc_sys *s, int imcon, const double cell[9]) {
  double det;
  s->imcon = imcon;
  memcpy(s->cell, cell, sizeof(double) * 9);
  orc_invert(cell, s->rcell, &det);
  s->volm = orc_volume(imcon, cell);
  return 0;
}

int orc_sys_set_ewald(orc_sys *s, double alpha, int k1, int k2, int k3,
                      double epsq) {
  s->alpha = alpha;
  s->kmax[0] = k1; s->kmax[1] = k2; s->kmax[2] = k3;
  s->epsq = epsq;
  return 0;
}

int orc_sys_set_cutoffs(orc_sys *s, double rcut, double delr, double rvdw) {
  s->rcut = rcut; s->delr = delr; s->rvdw = rvdw;
  return 0;
}

/* vdw_terms.f:1-187 (define_van_der_waals) table bookkeeping */
int orc_sys_set_vdw(orc_sys *s, int ntpatm, int ntpvdw, const int *ta,
                    const int *tb, const int *keys, const double *prm) {
  s->ntpatm = ntpatm;
  s->ntpvdw = ntpvdw;
  s->ntab = ntpatm * (ntpatm + 1) / 2;
  int maxvdw = ntpvdw > s->ntab ? ntpvdw : s->ntab;
  maxvdw += 1;
  s->lstvdw = calloc(maxvdw, sizeof(int));
  s->ltpvdw = malloc(sizeof(int) * maxvdw);
  s->prm = calloc((size_t)maxvdw * 5, sizeof(double));
  for (int i = 0; i < maxvdw; i++) s->ltpvdw[i] = -1;
  for (int t = 0; t < ntpvdw; t++) {
    int a = ta[t], b = tb[t];
    int mx = a > b ? a : b, mn = a > b ? b : a;
    int keyvdw = (mx * (mx - 1)) / 2 + mn;
    s->lstvdw[keyvdw - 1] = t + 1;
    s->ltpvdw[t] = keys[t];
    for (int p = 0; p < 5; p++) s->prm[5 * t + p] = prm[5 * t + p];
  }
  if (ntpvdw < s->ntab)
    for (int i = 0; i < s->ntab; i++)
      if (s->lstvdw[i] == 0) s->lstvdw[i] = ntpvdw + 1;
  return 0;
}

int orc_sys_add_framework(orc_sys *s, int n, const double *x, const double *y,
                          const double *z, const double *q, const int *type,
                          const int *frozen) {
  int m = s->nfw + n;
  s->fx = realloc(s->fx, sizeof(double) * m);
  s->fy = realloc(s->fy, sizeof(double) * m);
  s->fz = realloc(s->fz, sizeof(double) * m);
  s->fq = realloc(s->fq, sizeof(double) * m);
  s->ftype = realloc(s->ftype, sizeof(int) * m);
  s->ffrz = realloc(s->ffrz, sizeof(int) * m);
  for (int i = 0; i < n; i++) {
    s->fx[s->nfw + i] = x[i]; s->fy[s->nfw + i] = y[i]; s->fz[s->nfw + i] = z[i];
    s->fq[s->nfw + i] = q[i]; s->ftype[s->nfw + i] = type[i];
    s->ffrz[s->nfw + i] = frozen ? frozen[i] : 1;
  }
  s->nfw = m;
  return 0;
}

int orc_sys_add_guest_type(orc_sys *s, int nsites, const double *x,
                           const double *y, const double *z, const double *q,
                           const int *type, const double *mass) {
  if (s->ng >= ORC_MAXG || nsites > 64) return -1;
  orc_gtype *g = &s->g[s->ng];
  memset(g, 0, sizeof(*g));
  g->ns = nsites;
  for (int i = 0; i < nsites; i++) {
    g->tx[i] = x[i]; g->ty[i] = y[i]; g->tz[i] = z[i];
    g->q[i] = q[i]; g->type[i] = type[i]; g->mass[i] = mass ? mass[i] : 1.0;
  }
  return s->ng++;
}

static void orc_gtype_reserve(orc_gtype *g, int nmol) {
  if (nmol <= g->cap) return;
  int c = g->cap ? g->cap : 16;
  while (c < nmol) c *= 2;
  g->x = realloc(g->x, sizeof(double) * c * g->ns);
  g->y = realloc(g->y, sizeof(double) * c * g->ns);
  g->z = realloc(g->z, sizeof(double) * c * g->ns);
  g->cap = c;
}

int orc_sys_add_guest_molecule(orc_sys *s, int gt, const double *x,
                               const double *y, const double *z) {
  orc_gtype *g = &s->g[gt];
  orc_gtype_reserve(g, g->nmol + 1);
  for (int i = 0; i < g->ns; i++) {
    g->x[g->nmol * g->ns + i] = x[i];
    g->y[g->nmol * g->ns + i] = y[i];
    g->z[g->nmol * g->ns + i] = z[i];
  }
  g->nmol++;
  return 0;
}

/* utility_pack.f:3325-3423 (condense): guests first, then framework */
static void orc_condense(orc_sys *s) {
  int tot = s->nfw;
  for (int t = 0; t < s->ng; t++) tot += s->g[t].nmol * s->g[t].ns;
  if (tot > s->cap_all) {
    int c = tot + 1024;
    s->xxx = realloc(s->xxx, sizeof(double) * c);
    s->yyy = realloc(s->yyy, sizeof(double) * c);
    s->zzz = realloc(s->zzz, sizeof(double) * c);
    s->chg = realloc(s->chg, sizeof(double) * c);
    s->ltype = realloc(s->ltype, sizeof(int) * c);
    s->lfrz = realloc(s->lfrz, sizeof(int) * c);
    s->gstlist = realloc(s->gstlist, sizeof(int) * (size_t)c * 64);
    s->rsqbuf = realloc(s->rsqbuf, sizeof(double) * c);
    s->dxb = realloc(s->dxb, sizeof(double) * c);
    s->dyb = realloc(s->dyb, sizeof(double) * c);
    s->dzb = realloc(s->dzb, sizeof(double) * c);
    s->ilb = realloc(s->ilb, sizeof(int) * c);
    s->cap_all = c;
  }
  int jj = 0;
  for (int t = 0; t < s->ng; t++) {
    orc_gtype *g = &s->g[t];
    for (int m = 0; m < g->nmol; m++)
      for (int i = 0; i < g->ns; i++) {
        s->xxx[jj] = g->x[m * g->ns + i];
        s->yyy[jj] = g->y[m * g->ns + i];
        s->zzz[jj] = g->z[m * g->ns + i];
        s->chg[jj] = g->q[i];
        s->ltype[jj] = g->type[i];
        s->lfrz[jj] = 0;
        jj++;
      }
  }
  for (int i = 0; i < s->nfw; i++) {
    s->xxx[jj] = s->fx[i]; s->yyy[jj] = s->fy[i]; s->zzz[jj] = s->fz[i];
    s->chg[jj] = s->fq[i]; s->ltype[jj] = s->ftype[i]; s->lfrz[jj] = s->ffrz[i];
    jj++;
  }
  s->totatm = tot;
}

int orc_sys_finalize(orc_sys *s) {
  s->mxgrid = orc_mxgrid(s->rvdw > 0 ? s->rvdw : s->rcut);
  s->erc = malloc(sizeof(double) * (s->mxgrid + 1));
  s->drewd = orc_erfcgen(s->alpha, s->rcut, s->mxgrid, s->erc);
  int maxvdw = s->ntpvdw > s->ntab ? s->ntpvdw : s->ntab;
  s->vvv = calloc((size_t)(maxvdw + 1) * (s->mxgrid + 1), sizeof(double));
  int *keys = malloc(sizeof(int) * (s->ntpvdw + 1));
  for (int t = 0; t < s->ntpvdw; t++) keys[t] = s->ltpvdw[t];
  s->dlrpot = orc_forgen(s->ntpvdw, keys, s->prm,
                         s->rvdw > 0 ? s->rvdw : s->rcut, s->mxgrid, s->vvv);
  free(keys);
  s->steadd = calloc(s->ntab + 1, sizeof(double));
  s->nk = orc_kvec_count(s->cell, s->kmax);
  s->ckcsum = calloc(s->nk + 1, sizeof(double));
  s->ckssum = calloc(s->nk + 1, sizeof(double));
  s->ckcsnew = calloc(s->nk + 1, sizeof(double));
  s->ckssnew = calloc(s->nk + 1, sizeof(double));
  s->gstlentry = calloc(64, sizeof(int));
  for (int t = 0; t < s->ng; t++) {
    orc_gtype *g = &s->g[t];
    g->e3 = orc_ewald3_guest(s->imcon, s->cell, g->ns, g->tx, g->ty, g->tz,
                             g->q, s->alpha, s->epsq);
  }
  orc_condense(s);
  s->finalized = 1;
  orc_sys_recip_full(s, 1);
  return 0;
}

int orc_sys_nk(const orc_sys *s) { return s->nk; }
void orc_sys_cell_copy(const orc_sys *s, int *imcon, double cell[9],
                       double *rcut) {
  *imcon = s->imcon;
  memcpy(cell, s->cell, sizeof(double) * 9);
  *rcut = s->rcut;
}
void orc_sys_grid_info(const orc_sys *s, int *mxgrid, double *drewd,
                       double *dlrpot) {
  *mxgrid = s->mxgrid;
  *drewd = s->drewd;
  *dlrpot = s->dlrpot;
}
int orc_sys_get_table(const orc_sys *s, int which, int t, double *out) {
  int W = s->mxgrid + 1;
  if (which == 0) {
    memcpy(out, s->erc, sizeof(double) * W);
    return 0;
  }
  int maxvdw = s->ntpvdw > s->ntab ? s->ntpvdw : s->ntab;
  if (t < 0 || t > maxvdw) return -1;
  memcpy(out, s->vvv + (size_t)t * W, sizeof(double) * W);
  return 0;
}
int orc_sys_nmol(const orc_sys *s, int gt) { return s->g[gt].nmol; }
int orc_sys_nsites(const orc_sys *s, int gt) { return s->g[gt].ns; }
double orc_sys_volume(const orc_sys *s) { return s->volm; }
double orc_sys_ewald3en(const orc_sys *s, int gt) { return s->g[gt].e3; }
void orc_sys_get_molecule(const orc_sys *s, int gt, int m, double *x,
                          double *y, double *z) {
  const orc_gtype *g = &s->g[gt];
  for (int i = 0; i < g->ns; i++) {
    x[i] = g->x[m * g->ns + i];
    y[i] = g->y[m * g->ns + i];
    z[i] = g->z[m * g->ns + i];
  }
}
void orc_sys_get_sk(const orc_sys *s, double *c, double *sn) {
  memcpy(c, s->ckcsum, sizeof(double) * s->nk);
  memcpy(sn, s->ckssum, sizeof(double) * s->nk);
}

double orc_sys_recip_full(orc_sys *s, int all) {
  int n = all ? s->totatm : s->nfw;
  int off = all ? 0 : s->totatm - s->nfw;
  double *tc = malloc(sizeof(double) * (s->nk + 1) * 2), *ts = tc + s->nk + 1;
  int newld;
  double e = orc_ewald1(s->imcon, s->cell, s->alpha, s->kmax, s->epsq, n,
                        s->xxx + off, s->yyy + off, s->zzz + off, s->chg + off,
                        tc, ts, &newld, NULL);
  if (all) {
    memcpy(s->ckcsum, tc, sizeof(double) * s->nk);
    memcpy(s->ckssum, ts, sizeof(double) * s->nk);
  }
  free(tc);
  return e;
}

static int orc_mol_offset(const orc_sys *s, int gt) {
  int off = 0;
  for (int t = 0; t < gt; t++) off += s->g[t].nmol * s->g[t].ns;
  return off;
}

/* real-space sum for one molecule position (gcmc.f:906-936 and siblings):
 * e2 and vdw accumulate site by site */
static void orc_realspace(orc_sys *s, int ns, const double *nx,
                          const double *ny, const double *nz, const int *ind,
                          const double *sq, const int *st, double sign,
                          double *e2, double *ev) {
  orc_guestlistgen(s->imcon, s->cell, s->rcut, s->delr, ns, nx, ny, nz, ind,
                   s->totatm, s->xxx, s->yyy, s->zzz, s->gstlentry,
                   s->gstlist);
  for (int i = 0; i < ns; i++) {
    int n = s->gstlentry[i];
    const int *lst = s->gstlist + (size_t)i * s->totatm;
    for (int j = 0; j < n; j++) {
      int jatm = lst[j];
      s->ilb[j] = jatm;
      s->dxb[j] = nx[i] - s->xxx[jatm];
      s->dyb[j] = ny[i] - s->yyy[jatm];
      s->dzb[j] = nz[i] - s->zzz[jatm];
    }
    orc_images(s->imcon, n, s->cell, s->dxb, s->dyb, s->dzb);
    for (int j = 0; j < n; j++)
      s->rsqbuf[j] = s->dxb[j] * s->dxb[j] + s->dyb[j] * s->dyb[j] +
                     s->dzb[j] * s->dzb[j];
    double a = orc_ewald2(sq[i], n, s->ilb, s->rsqbuf, s->chg, s->erc,
                          s->drewd, s->rcut, s->epsq);
    double b = orc_srfrce(st[i], n, s->ilb, s->rsqbuf, s->ltype, s->lstvdw,
                          s->ltpvdw, s->vvv, s->mxgrid, s->rcut, s->dlrpot);
    if (sign > 0) {
      *e2 = *e2 + a;
      *ev = *ev + b;
    } else {
      *e2 = *e2 - a;
      *ev = *ev - b;
    }
  }
}

/* gcmc.f:855-946 (insert), :1010-1112 (delete), :1176-1328 (displace):
 * energy assembly of one trial */
int orc_sys_trial(orc_sys *s, int kind, int gt, int imol, const double *x,
                  const double *y, const double *z, double out[3]) {
  if (gt < 0 || gt >= s->ng) return -1;
  orc_gtype *g = &s->g[gt];
  int ns = g->ns;
  int ind[64];
  double ox[64], oy[64], oz[64];
  double e1 = 0.0, e2 = 0.0, ev = 0.0;
  if (kind == 1) {
    for (int i = 0; i < ns; i++) ind[i] = -1;
    e1 = orc_ewald1_guest(s->imcon, s->cell, s->alpha, s->kmax, s->epsq, ns, x,
                          y, z, g->q, 1, 1, s->ckcsum, s->ckssum, s->ckcsnew,
                          s->ckssnew);
    orc_realspace(s, ns, x, y, z, ind, g->q, g->type, 1.0, &e2, &ev);
    for (int i = 0; i < ns; i++) { s->px[i] = x[i]; s->py[i] = y[i]; s->pz[i] = z[i]; }
  } else {
    if (imol < 0 || imol >= g->nmol) return -2;
    int off = orc_mol_offset(s, gt) + imol * ns;
    for (int i = 0; i < ns; i++) {
      ind[i] = off + i;
      ox[i] = g->x[imol * ns + i];
      oy[i] = g->y[imol * ns + i];
      oz[i] = g->z[imol * ns + i];
    }
    if (kind == 2) {
      e1 = orc_ewald1_guest(s->imcon, s->cell, s->alpha, s->kmax, s->epsq, ns,
                            ox, oy, oz, g->q, 2, 1, s->ckcsum, s->ckssum,
                            s->ckcsnew, s->ckssnew);
      orc_realspace(s, ns, ox, oy, oz, ind, g->q, g->type, 1.0, &e2, &ev);
    } else if (kind == 3) {
      orc_ewald1_guest(s->imcon, s->cell, s->alpha, s->kmax, s->epsq, ns, ox,
                       oy, oz, g->q, 3, 1, s->ckcsum, s->ckssum, s->ckcsnew,
                       s->ckssnew);
      orc_realspace(s, ns, ox, oy, oz, ind, g->q, g->type, -1.0, &e2, &ev);
      e1 = orc_ewald1_guest(s->imcon, s->cell, s->alpha, s->kmax, s->epsq, ns,
                            x, y, z, g->q, 3, 2, s->ckcsum, s->ckssum,
                            s->ckcsnew, s->ckssnew);
      orc_realspace(s, ns, x, y, z, ind, g->q, g->type, 1.0, &e2, &ev);
      for (int i = 0; i < ns; i++) { s->px[i] = x[i]; s->py[i] = y[i]; s->pz[i] = z[i]; }
    } else {
      return -3;
    }
  }
  s->pend_kind = kind;
  s->pend_g = gt;
  s->pend_mol = imol;
  out[0] = e1;
  out[1] = e2;
  out[2] = ev;
  return 0;
}

/* gcmc.f:965-981, 1134-1151, 1355-1368 (commit of an accepted move) */
int orc_sys_commit(orc_sys *s) {
  if (!s->pend_kind) return -1;
  orc_gtype *g = &s->g[s->pend_g];
  int ns = g->ns;
  if (s->pend_kind == 1) {
    orc_sys_add_guest_molecule(s, s->pend_g, s->px, s->py, s->pz);
  } else if (s->pend_kind == 2) {
    int m = s->pend_mol;
    for (int i = (m + 1) * ns; i < g->nmol * ns; i++) {
      g->x[i - ns] = g->x[i];
      g->y[i - ns] = g->y[i];
      g->z[i - ns] = g->z[i];
    }
    g->nmol--;
  } else {
    int m = s->pend_mol;
    for (int i = 0; i < ns; i++) {
      g->x[m * ns + i] = s->px[i];
      g->y[m * ns + i] = s->py[i];
      g->z[m * ns + i] = s->pz[i];
    }
  }
  for (int k = 0; k < s->nk; k++) {
    s->ckcsum[k] = s->ckcsnew[k];
    s->ckssum[k] = s->ckssnew[k];
  }
  orc_condense(s);
  s->pend_kind = 0;
  return 0;
}

static void orc_counts(const orc_sys *s, int *numtyp, int *numfrz) {
  for (int i = 0; i < s->ntpatm; i++) { numtyp[i] = 0; numfrz[i] = 0; }
  for (int i = 0; i < s->nfw; i++) {
    numtyp[s->ftype[i] - 1]++;
    if (s->ffrz[i]) numfrz[s->ftype[i] - 1]++;
  }
  for (int t = 0; t < s->ng; t++)
    for (int i = 0; i < s->g[t].ns; i++)
      numtyp[s->g[t].type[i] - 1] += s->g[t].nmol;
}

/* gcmc.f:891-901 / 1056-1069: delrc = gstlrcorrect(N+-1) - elrc */
double orc_sys_tail_delta(orc_sys *s, int gt, int sign, double elrc) {
  int numtyp[256], numfrz[256];
  orc_counts(s, numtyp, numfrz);
  /* steadd is filled by lrcorrect over the current counts */
  orc_lrcorrect(s->imcon, s->ntpatm, numtyp, numfrz, s->lstvdw, s->ltpvdw,
                s->prm, s->rcut, s->volm, s->steadd);
  for (int i = 0; i < s->g[gt].ns; i++) numtyp[s->g[gt].type[i] - 1] += sign;
  double t = orc_gstlrcorrect(s->imcon, s->ntpatm, numtyp, numfrz, s->steadd,
                              s->volm);
  return t - elrc;
}

/* gcmc.f:2518-2710 (single_point) */
void orc_sys_energy_full(orc_sys *s, double out[6]) {
  double efw = orc_sys_recip_full(s, 0);
  double *tc = malloc(sizeof(double) * (s->nk + 1) * 2), *ts = tc + s->nk + 1;
  double eall = orc_ewald1(s->imcon, s->cell, s->alpha, s->kmax, s->epsq,
                           s->totatm, s->xxx, s->yyy, s->zzz, s->chg, tc, ts,
                           NULL, NULL);
  free(tc);
  double e2 = 0.0, ev = 0.0, e3 = 0.0;
  int nguest_atoms = s->totatm - s->nfw;
  /* molecule id per condensed atom: guests get a molecule id, framework -1 */
  int *molid = malloc(sizeof(int) * (s->totatm > 0 ? s->totatm : 1));
  int jj = 0, mid = 0;
  for (int t = 0; t < s->ng; t++)
    for (int m = 0; m < s->g[t].nmol; m++, mid++)
      for (int i = 0; i < s->g[t].ns; i++) molid[jj++] = mid;
  for (; jj < s->totatm; jj++) molid[jj] = -1;
  double rcsq = s->rcut * s->rcut;
  double rclim = (s->rcut + s->delr) * (s->rcut + s->delr);
  if (nguest_atoms > 0) {
    for (int i = 0; i < s->totatm; i++) {
      int n = 0;
      for (int j = i + 1; j < s->totatm; j++) {
        if (s->lfrz[i] && s->lfrz[j]) continue;
        if (molid[i] >= 0 && molid[i] == molid[j]) continue;
        s->ilb[n] = j;
        s->dxb[n] = s->xxx[i] - s->xxx[j];
        s->dyb[n] = s->yyy[i] - s->yyy[j];
        s->dzb[n] = s->zzz[i] - s->zzz[j];
        n++;
      }
      orc_images(s->imcon, n, s->cell, s->dxb, s->dyb, s->dzb);
      int nn = 0;
      for (int t = 0; t < n; t++) {
        double r2 = s->dxb[t] * s->dxb[t] + s->dyb[t] * s->dyb[t] +
                    s->dzb[t] * s->dzb[t];
        if (r2 < rclim) {
          s->ilb[nn] = s->ilb[t];
          s->rsqbuf[nn] = r2;
          nn++;
        }
      }
      (void)rcsq;
      e2 += orc_ewald2(s->chg[i], nn, s->ilb, s->rsqbuf, s->chg, s->erc,
                       s->drewd, s->rcut, s->epsq);
      ev += orc_srfrce(s->ltype[i], nn, s->ilb, s->rsqbuf, s->ltype,
                       s->lstvdw, s->ltpvdw, s->vvv, s->mxgrid, s->rcut,
                       s->dlrpot);
    }
    /* exclusions: every intramolecular pair once (condense + ewald3) */
    jj = 0;
    for (int t = 0; t < s->ng; t++) {
      orc_gtype *g = &s->g[t];
      for (int m = 0; m < g->nmol; m++) {
        for (int i = 0; i < g->ns - 1; i++) {
          double dx[64], dy[64], dz[64], cj[64];
          int k = 0;
          for (int j = i + 1; j < g->ns; j++) {
            dx[k] = s->xxx[jj + i] - s->xxx[jj + j];
            dy[k] = s->yyy[jj + i] - s->yyy[jj + j];
            dz[k] = s->zzz[jj + i] - s->zzz[jj + j];
            cj[k] = g->q[j];
            k++;
          }
          orc_images(s->imcon, k, s->cell, dx, dy, dz);
          e3 += orc_ewald3(g->q[i], k, cj, dx, dy, dz, s->alpha, s->epsq);
        }
        jj += g->ns;
      }
    }
  }
  free(molid);
  int numtyp[256], numfrz[256];
  orc_counts(s, numtyp, numfrz);
  double elrc = orc_lrcorrect(s->imcon, s->ntpatm, numtyp, numfrz, s->lstvdw,
                              s->ltpvdw, s->prm, s->rcut, s->volm, s->steadd);
  out[0] = efw;
  out[1] = eall - efw;
  out[2] = e2;
  out[3] = ev;
  out[4] = e3;
  out[5] = elrc;
}

/* ------------------------------------------------------------------ */
/* chain: gcmc.f:589-1960 restated for insert/delete/displace           */
/* ------------------------------------------------------------------ */

struct orc_chain {
  orc_sys *s;
  orc_chain_params p;
  orc_duni rng;
  double cum[ORC_MAXG][9];
  double beta, statvolm;
  double elrc;
  double energy[ORC_MAXG];
  long disp_count[ORC_MAXG], accept_disp[ORC_MAXG];
  long tran_count[ORC_MAXG], accept_tran[ORC_MAXG];
  long rota_count, accept_rota;
  double delrdisp[ORC_MAXG], tran_delr[ORC_MAXG], rotangle, rota_rotangle;
  long gcmccount;
};

orc_chain *orc_chain_create(orc_sys *s, const orc_chain_params *p) {
  orc_chain *c = calloc(1, sizeof(orc_chain));
  c->s = s;
  c->p = *p;
  /* gcmc.f:274-331: normalise move frequencies */
  for (int g = 0; g < p->ntpguest; g++) {
    double f[9];
    double norm = 0.0;
    for (int i = 0; i < 9; i++) { f[i] = p->freq[g][i]; norm += f[i]; }
    if (norm == 0.0) {
      norm = 1.0;
      for (int i = 0; i < 9; i++) f[i] = 0.0;
      f[0] = 1.0 / 3.0; f[1] = 1.0 / 3.0; f[2] = 1.0 / 3.0;
    }
    c->cum[g][0] = f[0] / norm;
    for (int i = 1; i < 9; i++) c->cum[g][i] = c->cum[g][i - 1] + (f[i] / norm);
    c->delrdisp[g] = p->delr[g];
    c->tran_delr[g] = p->delr[g];
  }
  /* gcmc.f:334-336: the first duni call only sets the generator up, the
   * second one draws a number that is thrown away */
  orc_duni_init(&c->rng, p->seed[0], p->seed[1], p->seed[2], p->seed[3]);
  (void)orc_duni_next(&c->rng);
  c->rotangle = ORC_PI / 3.0;
  c->rota_rotangle = c->rotangle;
  c->beta = 1.0 / (ORC_KBOLTZ * p->temp);
  c->statvolm = s->volm * 1e-30;
  c->elrc = 0.0;
  return c;
}

void orc_chain_destroy(orc_chain *c) { free(c); }
double orc_chain_energy(const orc_chain *c, int g) { return c->energy[g]; }
void orc_chain_adapt_state(const orc_chain *c, int g, double out[4]) {
  out[0] = c->delrdisp[g];
  out[1] = c->rotangle;
  out[2] = c->tran_delr[g];
  out[3] = c->rota_rotangle;
}

static double orc_chain_draw(orc_chain *c) { return orc_duni_next(&c->rng); }

int orc_chain_run(orc_chain *c, long nsteps, orc_move_rec *rec) {
  orc_sys *s = c->s;
  const double delrmin = 0.1, minangle = ORC_PI / 18.0;
  for (long step = 0; step < nsteps; step++) {
    c->gcmccount++;
    int ig = 0;
    if (c->p.ntpguest > 1)
      ig = (int)floor(orc_chain_draw(c) * c->p.ntpguest);
    double randmov = orc_chain_draw(c);
    const double *cum = c->cum[ig];
    int insert = 0, del = 0, disp = 0, tran = 0, rota = 0, other = 0;
    if (randmov < cum[0]) insert = 1;
    else if (randmov < cum[1]) del = 1;
    else if (randmov < cum[2]) disp = 1;
    else if (randmov < cum[3]) return -10; /* jump: not on this path */
    else if (randmov < cum[4]) other = 1;  /* flex: no-op in the reference */
    else if (randmov < cum[5]) other = 1;  /* swap: no-op */
    else if (randmov < cum[6]) { disp = 1; tran = 1; }
    else if (randmov < cum[7]) { disp = 1; rota = 1; }
    else if (randmov < cum[8]) return -11; /* multi-jump */
    else disp = 1;
    orc_gtype *g = &s->g[ig];
    int natms = g->ns;
    int nmols = g->nmol;
    /* gcmc.f:794-834 adaptive step sizes */
    if (nmols >= 1 && c->disp_count[ig] >= 1) {
      if (c->disp_count[ig] % 100 == 0) {
        if (((double)c->accept_disp[ig] / (double)c->disp_count[ig]) >
            c->p.disp_ratio[ig])
          c->delrdisp[ig] = c->delrdisp[ig] * 1.05;
        else if (c->delrdisp[ig] > delrmin)
          c->delrdisp[ig] = c->delrdisp[ig] * 0.95;
      } else if (c->disp_count[ig] % 50 == 0) {
        if (((double)c->accept_disp[ig] / (double)c->disp_count[ig]) >
            c->p.disp_ratio[ig])
          c->rotangle = c->rotangle * 1.05;
        else if (c->rotangle > minangle)
          c->rotangle = c->rotangle * 0.95;
      }
    }
    if (nmols >= 1 && c->tran_count[ig] >= 1 && c->tran_count[ig] % 100 == 0) {
      if (((double)c->accept_tran[ig] / (double)c->tran_count[ig]) >
          c->p.tran_ratio[ig])
        c->tran_delr[ig] = c->tran_delr[ig] * 1.05;
      else if (c->tran_delr[ig] > delrmin)
        c->tran_delr[ig] = c->tran_delr[ig] * 0.95;
    }
    if (nmols >= 1 && c->rota_count >= 1 && c->rota_count % 100 == 0) {
      if (((double)c->accept_rota / (double)c->rota_count) > c->p.rota_ratio)
        c->rota_rotangle = c->rota_rotangle * 1.05;
      else if (c->rota_rotangle > minangle)
        c->rota_rotangle = c->rota_rotangle * 0.95;
    }
    if (nmols == 0) { insert = 1; del = 0; disp = 0; tran = 0; rota = 0; other = 0; }
    orc_move_rec r;
    memset(&r, 0, sizeof(r));
    r.iguest = ig;
    double out[3];
    double eu = c->p.engunit;
    if (insert) {
      double ra = orc_chain_draw(c), rb = orc_chain_draw(c), rc = orc_chain_draw(c);
      double r1 = orc_chain_draw(c), r2 = orc_chain_draw(c), r3 = orc_chain_draw(c),
             r4 = orc_chain_draw(c);
      double nx[64], ny[64], nz[64];
      orc_random_ins(s->cell, natms, g->tx, g->ty, g->tz, ra, rb, rc, r1, r2,
                     r3, r4, nx, ny, nz);
      orc_sys_trial(s, 1, ig, -1, nx, ny, nz, out);
      double delrc = orc_sys_tail_delta(s, ig, +1, c->elrc);
      double e3 = g->e3;
      double estep = (out[0] + out[1] + e3 + out[2] + delrc) / eu;
      double rande = orc_chain_draw(c);
      int acc = orc_energy_eval(estep, rande, c->statvolm, c->p.press[ig],
                                nmols, c->p.temp, c->beta, 1);
      r.kind = 1; r.estep = estep; r.e1 = out[0]; r.e2 = out[1]; r.vdw = out[2];
      r.e3 = e3; r.delrc = delrc; r.accepted = acc;
      if (acc) {
        orc_sys_commit(s);
        c->elrc = c->elrc + delrc;
        c->energy[ig] += estep;
      }
    } else if (del) {
      int randchoice = (int)floor(orc_chain_draw(c) * nmols);
      double nul[1] = {0};
      orc_sys_trial(s, 2, ig, randchoice, nul, nul, nul, out);
      double delrc = orc_sys_tail_delta(s, ig, -1, c->elrc);
      double e3 = g->e3;
      double estep = (out[0] - out[1] - e3 - out[2] + delrc) / eu;
      double rande = orc_chain_draw(c);
      int acc = orc_energy_eval(estep, rande, c->statvolm, c->p.press[ig],
                                nmols, c->p.temp, c->beta, 2);
      r.kind = 2; r.estep = estep; r.e1 = out[0]; r.e2 = out[1]; r.vdw = out[2];
      r.e3 = e3; r.delrc = delrc; r.accepted = acc;
      if (acc) {
        orc_sys_commit(s);
        c->elrc = c->elrc + delrc;
        c->energy[ig] += estep;
      }
    } else if (disp) {
      double ra, rb, rc, r1, r2, r3, r4, dd, da;
      if (tran) {
        c->tran_count[ig]++;
        ra = orc_chain_draw(c); rb = orc_chain_draw(c); rc = orc_chain_draw(c);
        r1 = r2 = r3 = r4 = 0.0;
        dd = c->tran_delr[ig]; da = 0.0;
      } else if (rota) {
        c->rota_count++;
        ra = rb = rc = 0.5;
        r1 = orc_chain_draw(c); r2 = orc_chain_draw(c); r3 = orc_chain_draw(c);
        r4 = orc_chain_draw(c);
        dd = 0.0; da = c->rota_rotangle;
      } else {
        c->disp_count[ig]++;
        ra = orc_chain_draw(c); rb = orc_chain_draw(c); rc = orc_chain_draw(c);
        r1 = orc_chain_draw(c); r2 = orc_chain_draw(c); r3 = orc_chain_draw(c);
        r4 = orc_chain_draw(c);
        dd = c->delrdisp[ig]; da = c->rotangle;
      }
      int randchoice = (int)floor(orc_chain_draw(c) * nmols);
      double nx[64], ny[64], nz[64];
      orc_sys_get_molecule(s, ig, randchoice, nx, ny, nz);
      orc_random_disp(s->cell, s->rcell, natms, g->mass, nx, ny, nz, dd, da,
                      ra, rb, rc, r1, r2, r3, r4);
      orc_sys_trial(s, 3, ig, randchoice, nx, ny, nz, out);
      double estep = (out[0] + out[1] + out[2]) / eu;
      int acc;
      if (estep < 0.0) {
        acc = 1;
      } else {
        double rande = orc_chain_draw(c);
        acc = orc_energy_eval(estep, rande, c->statvolm, c->p.press[ig], nmols,
                              c->p.temp, c->beta, 3);
      }
      r.kind = tran ? 4 : (rota ? 5 : 3);
      r.estep = estep; r.e1 = out[0]; r.e2 = out[1]; r.vdw = out[2];
      r.accepted = acc;
      if (acc) {
        if (tran) c->accept_tran[ig]++;
        else if (rota) c->accept_rota++;
        else c->accept_disp[ig]++;
        orc_sys_commit(s);
        c->energy[ig] += estep;
      }
    } else {
      (void)other;
      r.kind = 0;
    }
    r.nmol_after = g->nmol;
    if (rec) rec[step] = r;
  }
  return 0;
}
