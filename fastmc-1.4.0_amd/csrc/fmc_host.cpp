// Host-side set-up of the trial-move path (see fmc_host.h).
#include "fmc_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace fmc {

// Cell lengths, cosines, perpendicular widths, volume (utility_pack.f:1351).
void dcell(const double a[9], double b[10]) {
  const double* va = a;
  const double* vb = a + 3;
  const double* vc = a + 6;
  auto dot = [](const double* u, const double* v) {
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
  };
  b[0] = std::sqrt(dot(va, va));
  b[1] = std::sqrt(dot(vb, vb));
  b[2] = std::sqrt(dot(vc, vc));
  b[3] = dot(va, vb) / (b[0] * b[1]);
  b[4] = dot(va, vc) / (b[0] * b[2]);
  b[5] = dot(vb, vc) / (b[1] * b[2]);
  double axb[3] = {a[1] * a[5] - a[2] * a[4], a[2] * a[3] - a[0] * a[5],
                   a[0] * a[4] - a[1] * a[3]};
  double bxc[3] = {a[4] * a[8] - a[5] * a[7], a[5] * a[6] - a[3] * a[8],
                   a[3] * a[7] - a[4] * a[6]};
  double cxa[3] = {a[7] * a[2] - a[1] * a[8], a[0] * a[8] - a[2] * a[6],
                   a[1] * a[6] - a[0] * a[7]};
  b[9] = std::fabs(a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2]);
  b[6] = b[9] / std::sqrt(dot(bxc, bxc));
  b[7] = b[9] / std::sqrt(dot(cxa, cxa));
  b[8] = b[9] / std::sqrt(dot(axb, axb));
}

// 3x3 inverse by cofactors, column-major DL_POLY layout (utility_pack.f:1417).
void invert(const double a[9], double b[9], double* det) {
  double c[9];
  c[0] = a[4] * a[8] - a[5] * a[7];
  c[1] = a[2] * a[7] - a[1] * a[8];
  c[2] = a[1] * a[5] - a[2] * a[4];
  c[3] = a[5] * a[6] - a[3] * a[8];
  c[4] = a[0] * a[8] - a[2] * a[6];
  c[5] = a[2] * a[3] - a[0] * a[5];
  c[6] = a[3] * a[7] - a[4] * a[6];
  c[7] = a[1] * a[6] - a[0] * a[7];
  c[8] = a[0] * a[4] - a[1] * a[3];
  double d = a[0] * c[0] + a[3] * c[1] + a[6] * c[2];
  double r = std::fabs(d) > 0.0 ? 1.0 / d : 0.0;
  for (int i = 0; i < 9; ++i) b[i] = r * c[i];
  *det = d;
}

double volume(int imcon, const double cell[9]) {
  double p[10];
  dcell(cell, p);
  if (imcon == 0) return 0.0;
  if (imcon == 4 || imcon == 5 || imcon == 7) return 0.5 * p[9];
  return p[9];
}

// "ewald precision" -> alpha, kmax (utility_pack.f:380-392)
void ewald_precision(double eps, double rcut, int imcon, const double cell[9],
                     double* alpha, int kmax[3]) {
  double p[10];
  dcell(cell, p);
  eps = std::min(std::fabs(eps), 0.5);
  const double tol = std::sqrt(std::fabs(std::log(eps * rcut)));
  const double a = std::sqrt(std::fabs(std::log(eps * rcut * tol))) / rcut;
  const double t2 = 2.0 * tol * a;
  const double tol1 = std::sqrt(-std::log(eps * rcut * (t2 * t2)));
  const double fac =
      (imcon == 4 || imcon == 5 || imcon == 7) ? std::pow(2.0, 1.0 / 3.0) : 1.0;
  for (int d = 0; d < 3; ++d)
    kmax[d] = static_cast<int>(std::round(0.25 + fac * p[d] * a * tol1 / kPi));
  *alpha = a;
}

int mxgrid_for(double rvdw) {
  return std::max(1000, static_cast<int>(rvdw / 0.01 + 0.5) + 4);
}

// Half-space enumeration with the spherical cut of ewald.f:77-84,141-213.
std::vector<KVec> enumerate_kvecs(int imcon, const double cell[9],
                                  double alpha, const int kmax[3]) {
  double rc[9], det, w[10];
  invert(cell, rc, &det);
  dcell(rc, w);
  const double twopi = 2.0 * kPi;
  double cut = std::min({kmax[0] * w[6], kmax[1] * w[7], kmax[2] * w[8]});
  cut = cut * 1.05 * twopi;
  const double cut2 = cut * cut;
  const double ralph = -0.25 / (alpha * alpha);
  const bool lconsw = !(imcon == 4 || imcon == 5 || imcon == 7);
  std::vector<KVec> out;
  for (int l = 0; l <= kmax[0]; ++l) {
    const int m0 = (l == 0) ? 0 : -kmax[1];
    for (int m = m0; m <= kmax[1]; ++m) {
      const int n0 = (l == 0 && m == 0) ? 1 : -kmax[2];
      for (int n = n0; n <= kmax[2]; ++n) {
        const double tl = twopi * l, tm = twopi * m, tn = twopi * n;
        const double kx = tl * rc[0] + tm * rc[1] + tn * rc[2];
        const double ky = tl * rc[3] + tm * rc[4] + tn * rc[5];
        const double kz = tl * rc[6] + tm * rc[7] + tn * rc[8];
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 <= cut2) {
          double ak = std::exp(ralph * k2) * (1.0 / k2);
          if (!lconsw) ak *= 4.0;
          out.push_back({l, m, n, ak});
        }
      }
    }
  }
  return out;
}

namespace {

// A&S 7.1.26 erfc(alpha r)/r on r = i*drewd (ewald.f:1-30)
void build_erfc(double alpha, double rcut, int mxgrid, std::vector<double>* erc,
                double* drewd) {
  constexpr double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
  constexpr double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
  const double h = rcut / static_cast<double>(mxgrid - 4);
  erc->assign(mxgrid + 1, 0.0);
  for (int i = 1; i <= mxgrid; ++i) {
    const double r = i * h;
    const double tt = 1.0 / (1.0 + pp * alpha * r);
    const double ar = alpha * r;
    const double e = std::exp(-(ar * ar));
    (*erc)[i] = tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * e / r;
  }
  (*erc)[0] = (*erc)[1];
  *drewd = h;
}

double powi(double x, int n) {
  double r = 1.0;
  while (n-- > 0) r *= x;
  return r;
}

// One vdW table entry for key/prm at r (vdw_terms.f:225-392)
double vdw_value(int key, const double* p, double r, double rvdw) {
  switch (key) {
    case 1: {
      const double r6 = powi(r, 6);
      return (p[0] / r6 - p[1]) / r6;
    }
    case 2: {
      const double s6 = powi(p[1] / r, 6);
      return 4.0 * p[0] * s6 * (s6 - 1.0);
    }
    case 3: {
      const double ann = std::max(p[1], p[2]), amm = std::min(p[1], p[2]);
      return p[0] / (ann - amm) *
             (amm * std::pow(p[3] / r, ann) - ann * std::pow(p[3] / r, amm));
    }
    case 4:
      return p[0] * std::exp(-r / p[1]) - p[2] / powi(r, 6);
    case 5:
      return p[0] * std::exp(p[1] * (p[2] - r)) - p[3] / powi(r, 6) -
             p[4] / powi(r, 8);
    case 6:
      return p[0] / powi(r, 12) - p[1] / powi(r, 10);
    case 7: {
      double eps = p[0];
      const double ann = p[1], amm = p[2], r0 = p[3];
      double rc = p[4] < 1e-6 ? rvdw : p[4];
      const double gam = rc / r0;
      const double bet =
          gam * std::pow((std::pow(gam, amm + 1.0) - 1.0) /
                             (std::pow(gam, ann + 1.0) - 1.0),
                         1.0 / (ann - amm));
      eps = -eps * (ann - amm) /
            (amm * std::pow(bet, ann) *
                 (1.0 + (ann / gam - ann - 1.0) / std::pow(gam, ann)) -
             ann * std::pow(bet, amm) *
                 (1.0 + (amm / gam - amm - 1.0) / std::pow(gam, amm)));
      if (r > rc) return 0.0;
      return eps / (ann - amm) *
             (amm * std::pow(bet, ann) *
                  (std::pow(r0 / r, ann) - std::pow(1.0 / gam, ann)) -
              ann * std::pow(bet, amm) *
                  (std::pow(r0 / r, amm) - std::pow(1.0 / gam, amm)) +
              ann * amm *
                  ((r / (gam * r0) - 1.0) *
                   (std::pow(bet / gam, ann) - std::pow(bet / gam, amm))));
    }
    case 8: {
      const double e1 = 1.0 - std::exp(-p[2] * (r - p[1]));
      return p[0] * (e1 * e1 - 1.0);
    }
    case 9: {
      const double d = p[1] * std::pow(2.0, 1.0 / 6.0);
      const double rr = r - p[2];
      if (rr > d || rr <= 0.0) return 0.0;
      const double s6 = powi(p[1] / rr, 6);
      return 4.0 * p[0] * s6 * (s6 - 1.0) + p[0];
    }
    default:
      return 0.0;
  }
}

// lrcorrect's per-pair eadd (vdw_terms.f:555-650), doubled for i != j
double tail_eadd(int key, const double* p, double rc) {
  switch (key) {
    case 0:
      return p[0];
    case 1:
      return p[0] / (9.0 * powi(rc, 9)) - p[1] / (3.0 * powi(rc, 3));
    case 2:
      return 4.0 * p[0] *
             (powi(p[1], 12) / (9.0 * powi(rc, 9)) - powi(p[1], 6) / (3.0 * powi(rc, 3)));
    case 3: {
      const double ann = p[1], amm = p[2], r0 = p[3];
      return p[0] / (ann - amm) *
             (amm * std::pow(r0, ann) / ((ann - 3.0) * std::pow(rc, ann - 3.0)) -
              ann * std::pow(r0, amm) / ((amm - 3.0) * std::pow(rc, amm - 3.0)));
    }
    case 4:
      return -p[2] / (3.0 * powi(rc, 3));
    case 5:
      return -p[3] / (3.0 * powi(rc, 3)) - p[4] / (5.0 * powi(rc, 5));
    case 6:
      return p[0] / (9.0 * powi(rc, 9)) - p[1] / (7.0 * powi(rc, 7));
    default:
      return 0.0;
  }
}

double erf_as(double alpha, double r) {
  // 1 - A&S 7.1.26, with the small-argument series of ewald.f:741-751
  constexpr double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741;
  constexpr double a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
  const double tt = 1.0 / (1.0 + pp * alpha * r);
  const double ar = alpha * r;
  return 1.0 - tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) *
                   std::exp(-(ar * ar));
}

void images1(int imcon, const double cell[9], const double rc[9], double* x,
             double* y, double* z) {
  if (imcon == 1) {
    const double a = 1.0 / cell[0];
    *x -= cell[0] * std::round(a * *x);
    *y -= cell[0] * std::round(a * *y);
    *z -= cell[0] * std::round(a * *z);
  } else if (imcon == 2) {
    *x -= cell[0] * std::round(*x / cell[0]);
    *y -= cell[4] * std::round(*y / cell[4]);
    *z -= cell[8] * std::round(*z / cell[8]);
  } else if (imcon == 3) {
    const double sx = rc[0] * *x + rc[3] * *y + rc[6] * *z;
    const double sy = rc[1] * *x + rc[4] * *y + rc[7] * *z;
    const double sz = rc[2] * *x + rc[5] * *y + rc[8] * *z;
    const double ax = sx - std::round(sx), ay = sy - std::round(sy),
                 az = sz - std::round(sz);
    *x = cell[0] * ax + cell[3] * ay + cell[6] * az;
    *y = cell[1] * ax + cell[4] * ay + cell[7] * az;
    *z = cell[2] * ax + cell[5] * ay + cell[8] * az;
  }
}

}  // namespace

double ewald3_molecule(const SystemDesc& s, const GuestTemplate& g,
                       const double* x, const double* y, const double* z) {
  constexpr double r3 = 0.333333333333, r10 = 0.1, r42 = 0.02380952381,
                   r216 = 4.62962962963e-3;
  double sum = 0.0;
  for (int i = 0; i < g.ns - 1; ++i) {
    const double ce = g.q[i] / s.epsq * kR4pie0;
    for (int j = i + 1; j < g.ns; ++j) {
      double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
      images1(s.imcon, s.cell, s.rcell, &dx, &dy, &dz);
      const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double cp = ce * g.q[j];
      const double ar = r * s.alpha;
      double e;
      if (ar < 1e-2) {
        const double a2 = ar * ar;
        e = 2.0 * cp * (s.alpha / kSqrpi) *
            (1.0 + a2 * (-r3 + a2 * (r10 + a2 * (-r42 + a2 * r216))));
      } else {
        e = erf_as(s.alpha, r) * cp / r;
      }
      sum -= e;
    }
  }
  return sum;
}

int SystemDesc::finalize(std::string* err) {
  if (imcon < 1 || imcon > 3) {
    if (err) *err = "imcon must be 1, 2 or 3 on this path";
    return -3;
  }
  if (!have_ewald) {
    if (err) *err = "ewald parameters not set";
    return -2;
  }
  double det;
  invert(cell, rcell, &det);
  volm = volume(imcon, cell);
  double p[10];
  dcell(cell, p);
  const double width = std::min({p[6], p[7], p[8]}) / 2.0;
  if (rcut > width) {
    if (err) *err = "cutoff exceeds half the perpendicular cell width";
    return -7;
  }
  const double rv = rvdw > 0 ? rvdw : rcut;
  mxgrid = mxgrid_for(rv);
  build_erfc(alpha, rcut, mxgrid, &erc, &drewd);
  // vdW bookkeeping (vdw_terms.f:1-187)
  const int ntab_pairs = ntpatm * (ntpatm + 1) / 2;
  ntable = std::max(ntpvdw, ntab_pairs) + 1;
  lstvdw.assign(ntab_pairs, 0);
  ltpvdw.assign(ntable, -1);
  for (int t = 0; t < ntpvdw; ++t) {
    const int a = vdw_a[t], b = vdw_b[t];
    if (a < 1 || b < 1 || a > ntpatm || b > ntpatm) {
      if (err) *err = "vdw pair names an unknown atom type";
      return -1;
    }
    const int mx = std::max(a, b), mn = std::min(a, b);
    lstvdw[(mx * (mx - 1)) / 2 + mn - 1] = t + 1;
    ltpvdw[t] = vdw_key[t];
  }
  if (ntpvdw < ntab_pairs)
    for (int i = 0; i < ntab_pairs; ++i)
      if (lstvdw[i] == 0) lstvdw[i] = ntpvdw + 1;
  dlrpot = rv / static_cast<double>(mxgrid - 4);
  const int W = mxgrid + 1;
  vvv.assign(static_cast<size_t>(ntable) * W, 0.0);
  for (int t = 0; t < ntpvdw; ++t) {
    for (int i = 1; i <= mxgrid; ++i)
      vvv[static_cast<size_t>(t) * W + i] =
          vdw_value(vdw_key[t], &vdw_prm[5 * t], i * dlrpot, rv);
    vvv[static_cast<size_t>(t) * W] = vvv[static_cast<size_t>(t) * W + 1];
  }
  pairtab.assign(ntpatm * ntpatm, -1);
  for (int i = 1; i <= ntpatm; ++i)
    for (int j = 1; j <= ntpatm; ++j) {
      const int mx = std::max(i, j), mn = std::min(i, j);
      const int k = lstvdw[(mx * (mx - 1)) / 2 + mn - 1];  // 1-based table
      const bool use = ltpvdw[k - 1] < 100 &&
                       std::fabs(vvv[static_cast<size_t>(k - 1) * W + 1]) > 1e-10;
      pairtab[(i - 1) * ntpatm + (j - 1)] = use ? k - 1 : -1;
    }
  // tail terms per pair (vdw_terms.f:548-662)
  steadd.assign(ntab_pairs, 0.0);
  {
    int iv = 0;
    for (int i = 1; i <= ntpatm; ++i)
      for (int j = 1; j <= i; ++j, ++iv) {
        const int k = lstvdw[iv];
        const int key = ltpvdw[k - 1];
        double e = 0.0;
        if (key >= 0 && k - 1 < ntpvdw) e = tail_eadd(key, &vdw_prm[5 * (k - 1)], rcut);
        if (i != j) e *= 2.0;
        steadd[iv] = e;
      }
  }
  kvecs = enumerate_kvecs(imcon, cell, alpha, kmax);
  double sq = 0.0;
  for (double q : fq) sq += q * q;
  fw_engsic = -kR4pie0 / epsq * alpha * sq / kSqrpi;
  for (auto& g : guests) {
    double s2 = 0.0;
    for (int i = 0; i < g.ns; ++i) s2 += g.q[i] * g.q[i];
    g.engsic = -kR4pie0 / epsq * alpha * s2 / kSqrpi;
    g.e3 = ewald3_molecule(*this, g, g.x, g.y, g.z);
  }
  finalized = true;
  return 0;
}

void SystemDesc::counts(const std::vector<int>& nmol, std::vector<int>* numtyp,
                        std::vector<int>* numfrz) const {
  numtyp->assign(ntpatm, 0);
  numfrz->assign(ntpatm, 0);
  for (size_t i = 0; i < ftype.size(); ++i) {
    (*numtyp)[ftype[i] - 1]++;
    if (ffrz[i]) (*numfrz)[ftype[i] - 1]++;
  }
  for (size_t t = 0; t < guests.size(); ++t)
    for (int s = 0; s < guests[t].ns; ++s)
      (*numtyp)[guests[t].type[s] - 1] += nmol[t];
}

double SystemDesc::tail(const std::vector<int>& numtyp,
                        const std::vector<int>& numfrz) const {
  if (imcon == 0 || imcon == 6 || ntpvdw == 0) return 0.0;
  const double twopi = 2.0 * kPi;
  double e = 0.0;
  int iv = 0;
  for (int i = 1; i <= ntpatm; ++i)
    for (int j = 1; j <= i; ++j, ++iv) {
      const double den = twopi *
                         (static_cast<double>(numtyp[i - 1]) * numtyp[j - 1] -
                          static_cast<double>(numfrz[i - 1]) * numfrz[j - 1]) /
                         (volm * volm);
      e = e + volm * den * steadd[iv];
    }
  return e;
}

}  // namespace fmc
