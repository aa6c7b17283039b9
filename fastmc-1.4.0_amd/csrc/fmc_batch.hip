// Batched device-resident chains: many independent copies of the reference's
// move loop (gcmc.f:589-1960) advance inside one persistent launch, one
// chain per work-group.  Nothing returns to the host per move: duni, the
// insert/delete/displace proposals, the energy of the trial, the acceptance
// test and the commit all run on the GPU.
//
// Per chain, for the whole launch:
//   * S(k) (FP64) and a_k sit in registers, KPT k-vectors per thread
//     (k = tid + j*BT), so the reciprocal update reads no memory;
//   * the trial's structure-factor change dS(k) stays in registers and is
//     added to S(k) when the move is accepted;
//   * per-position tables q X[l] Y[m] and Z[n] of the phase factors
//     e^{i 2 pi l s} (ewald.f:97-137) sit in LDS;
//   * real space walks a per-framework candidate list: the framework is cut
//     into ~1.5 A cells and each cell lists every atom that can lie within
//     rcut of a point in it (a superset, so the pair set -- and therefore the
//     energy -- is exactly that of the reference's all-atom scan), plus all
//     guest atoms;
//   * the serial part runs wave-wide (RNG draws and proposal: every lane
//     the same scalar steps, lane i draw i, lane s site s) or on one lane
//     (acceptance and bookkeeping).
// Precision FP64 reproduces the oracle's accept/reject stream; MIXED keeps
// the per-position phase-factor products q X[l] Y[m], Z[n] and the trial's
// dS(k) in FP32 (a displacement's as the change of each site's term, formed
// in FP64 before rounding), with FP64 coordinates, S(k), real-space pair
// terms, energies and acceptance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "fastmc_amd.h"
#include "fmc_ctx.h"
#include "fmc_device.h"
#include "fmc_host.h"

namespace fmc {

constexpr int kBTab = 17;   // phase-table entries per dimension (kmax <= 16)
constexpr int kLogFields = 8;
constexpr int kWin = 5;          // statistics windows (gcmc.f:119)
constexpr int kWinQ = 5;         // sums per window: N, E, E*N, N^2, E^2
constexpr int kWinSnap = 2 * kWin - 1;   // running-sum snapshots that bound them
constexpr int kWinDoubles = kWinSnap * kWinQ * kMaxGuestTypes;
constexpr int kWinMax = 500000;  // samples beyond which the reference keeps
                                 // rolling averages instead (gcmc.f:1971)
constexpr int kLdsTypes = 32;   // atom types whose pair rows sit in LDS
constexpr int kSelPairs = kMaxGuestTypes * (kMaxGuestTypes - 1) / 2;
constexpr long long kSelWindow = 100000;   // readinputs.f:888-889

// The small per-system constants the serial lane and the per-move set-up
// read; the batch kernel keeps a copy in LDS for the whole launch.
struct SysQ {
  double cell[9], rcell[9];
  double epsq;
  double rvolm2;          // 2 * 2 pi / V
  double volm;
  int kmax[3];
  int gdim[3];
  int ngt;
  int ns[kMaxGuestTypes];
  double gq[kMaxGuestTypes][kMaxSites];
  int gtype[kMaxGuestTypes][kMaxSites];  // 0-based
  double gmass[kMaxGuestTypes][kMaxSites];
  double tx[kMaxGuestTypes][kMaxSites], ty[kMaxGuestTypes][kMaxSites],
      tz[kMaxGuestTypes][kMaxSites];
  double e3[kMaxGuestTypes];
  double engsic[kMaxGuestTypes];
  // tail(n) = t0 + sum_g tb[g] n_g + sum_gh tA[g][h] n_g n_h
  double t0, tb[kMaxGuestTypes], tA[kMaxGuestTypes][kMaxGuestTypes];
};

struct alignas(16) Rec {
  double x, y, z, q;
};

struct BSys {
  DevSys S;
  int nk;
  const int* kv;
  const double* ak;
  const double* sfw_re;   // framework S(k) (constant)
  const double* sfw_im;
  // candidate cells: cell c lists, in ascending atom order, every framework
  // atom that can lie within rcut of a point in it, as records
  // rec[c*gcap + i] (x, y, z, q) and rtype[c*gcap + i]; unused slots hold
  // NaN coordinates, which every pair test skips
  const Rec* grec;
  const int* grec_t;
  int gcap;
  SysQ q;
};

struct BChain {
  int sys;
  int ntpguest;
  int cap_atoms;
  // end-of-run windows (gcmc.f:1969-2031): window j sums the production
  // samples j*L ... j*L + L (L = win_len), the difference of the running
  // sums after j*L + L + 1 samples and after j*L samples.  The running sums
  // are copied to the snapshot area that follows S(k) in the sre buffer
  // (sre + nk: [snapshot][quantity][guest]) when the production count
  // reaches L, L+1, 2L, 2L+1, 3L, 3L+1, 4L, 4L+1, 5L+1.
  int win_len;    // 0: no windows (no move limit, or too many samples)
  double* sre;
  double* sim;
  double* dsre;   // dS scratch for the streaming (large-Nk) kernel
  double* dsim;
  double* gx[kMaxGuestTypes];
  double* gy[kMaxGuestTypes];
  double* gz[kMaxGuestTypes];
  double temp, beta, statvolm, engunit;
  double press[kMaxGuestTypes];
  double cum[kMaxGuestTypes][9];
  double disp_ratio[kMaxGuestTypes], tran_ratio[kMaxGuestTypes], rota_ratio;
  long long eq_steps;
  // RNG (gcmc.f:2712-2805)
  float u[97];
  float c, cd, cm;
  int ir, jr;
  // state
  int nmol[kMaxGuestTypes];
  double energy[kMaxGuestTypes];
  double elrc;
  double delrdisp[kMaxGuestTypes], tran_delr[kMaxGuestTypes], rotangle,
      rota_rotangle;
  int disp_count[kMaxGuestTypes], accept_disp[kMaxGuestTypes];
  int tran_count[kMaxGuestTypes], accept_tran[kMaxGuestTypes];
  int rota_count, accept_rota;
  long long steps, prod_steps;
  long long accepted[6], attempted[6];
  double sum_n[kMaxGuestTypes], sum_n2[kMaxGuestTypes], sum_e[kMaxGuestTypes];
  double sum_en[kMaxGuestTypes], sum_e2[kMaxGuestTypes];
  double* log;
  long long log_cap, log_n;
  int fault;   // 1: guest capacity reached (reference error 75)
  int win_snap;   // snapshots taken
  long long stop_at;   // the chain idles once steps reaches this
  long long recomp_every;   // > 0: FP64 recompute inside the launch every this many moves
  // FP64 full-energy bookkeeping: e_acc is the running sum of accepted
  // estep from the full energy at admission; e_full the last full
  // recompute (kcal/mol); drift = e_acc - e_full at that recompute
  double e_acc, e_full, drift, sk_dev;
  long long n_recompute;
  // selectivity (gcmc.f:1887-1916): windows of sel_nwind production
  // samples; pair p = i (i - 1) / 2 + j of guests j < i
  long long sel_nwind, sel_count;
  double sel_base[kMaxGuestTypes];   // sum_n when the last window closed
  double sel_mean[kSelPairs], sel_sec[kSelPairs], sel_var[kSelPairs];
};

// ---- RNG on one lane (gcmc.f:2786-2799) ----
struct Rng {
  float* u;
  float c, cd, cm;
  int ir, jr;
  __device__ double next() {
    float uni = u[ir - 1] - u[jr - 1];
    if (uni < 0.0f) uni = uni + 1.0f;
    u[ir - 1] = uni;
    ir = ir - 1;
    if (ir == 0) ir = 97;
    jr = jr - 1;
    if (jr == 0) jr = 97;
    c = c - cd;
    if (c < 0.0f) c = c + cm;
    uni = uni - c;
    if (uni < 0.0f) uni = uni + 1.0f;
    return static_cast<double>(uni);
  }
};

// Lane j's value of a wave-uniform computation (j uniform).
__device__ __forceinline__ float lane_f(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_d(double v, int j) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b), j));
  const unsigned hi =
      static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), j));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ int uni_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Draw i of the duni stream (gcmc.f:2786-2799) ahead of state R, on lane i
// of a wave.  Draw i reads u(ir_i), u(jr_i) and overwrites u(ir_i); both
// indices step down together from 97 and 33, so jr = ir + 33 (mod 97)
// throughout and draw i reads an entry that another draw overwrites only 33
// draws later: within kLook draws every read sees the table as the
// sequential draws would.  c, cd, cm and the table entries are multiples of
// 2^-24 below 1, so every float step is exact and c after i+1 steps is
// (c - (i+1) cd) mod cm in integers of 2^-24.
constexpr int kLook = 10;   // most draws one proposal takes
struct Ahead {
  int ir;    // u index of this lane's draw
  float w;   // table entry the draw leaves
  float c;   // c after the draw
  float d;   // the draw
};

__device__ __forceinline__ Ahead rng_ahead(const Rng& R, int lane) {
  Ahead A;
  const int i = lane < kLook ? lane : kLook - 1;
  int ir = R.ir - i;
  if (ir <= 0) ir += 97;
  int jr = R.jr - i;
  if (jr <= 0) jr += 97;
  A.ir = ir;
  float w = R.u[ir - 1] - R.u[jr - 1];
  if (w < 0.0f) w = w + 1.0f;
  A.w = w;
  const int cd = static_cast<int>(R.cd * 16777216.0f);
  const int cm = static_cast<int>(R.cm * 16777216.0f);
  int ci = (static_cast<int>(R.c * 16777216.0f) - (i + 1) * cd) % cm;
  if (ci < 0) ci += cm;
  A.c = static_cast<float>(ci) * (1.0f / 16777216.0f);
  float d = w - A.c;
  if (d < 0.0f) d = d + 1.0f;
  A.d = d;
  return A;
}

// the first n draws taken: their table entries written (one per lane) and
// the state after them, on every lane
__device__ __forceinline__ void rng_commit(Rng& R, const Ahead& A, int lane, int n) {
  if (lane < n) R.u[A.ir - 1] = A.w;
  if (n > 0) {
    R.c = lane_f(A.c, n - 1);
    R.ir = R.ir - n;
    if (R.ir <= 0) R.ir += 97;
    R.jr = R.jr - n;
    if (R.jr <= 0) R.jr += 97;
  }
}

// rotation matrix of rotation (utility_pack.f:2870-2906) from the
// quaternion built as random_ins / random_rot do
__device__ __forceinline__ void rot_matrix(double theta, double r2, double r3, double r4,
                                           double (&Q)[9]) {
  const double b1 = 2.0 * r2 - 1.0, b2 = 2.0 * r3 - 1.0, b3 = 2.0 * r4 - 1.0;
  double st, q1;
  sincos(theta / 2.0, &st, &q1);
  const double nrm = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
  const double q2 = st * (b1 / nrm), q3 = st * (b2 / nrm), q4 = st * (b3 / nrm);
  Q[0] = 2.0 * q2 * q2 + 2.0 * q1 * q1 - 1.0;
  Q[1] = 2.0 * q2 * q3 - 2.0 * q1 * q4;
  Q[2] = 2.0 * q2 * q4 + 2.0 * q1 * q3;
  Q[3] = 2.0 * q2 * q3 + 2.0 * q1 * q4;
  Q[4] = 2.0 * q1 * q1 + 2.0 * q3 * q3 - 1.0;
  Q[5] = 2.0 * q3 * q4 - 2.0 * q1 * q2;
  Q[6] = 2.0 * q2 * q4 - 2.0 * q1 * q3;
  Q[7] = 2.0 * q3 * q4 + 2.0 * q1 * q2;
  Q[8] = 2.0 * q4 * q4 + 2.0 * q1 * q1 - 1.0;
}

__device__ __forceinline__ void rot_apply(const double (&Q)[9], double& x, double& y,
                                          double& z) {
  const double tx = Q[0] * x + Q[1] * y + Q[2] * z;
  const double ty = Q[3] * x + Q[4] * y + Q[5] * z;
  const double tz = Q[6] * x + Q[7] * y + Q[8] * z;
  x = tx;
  y = ty;
  z = tz;
}

// energy_eval (utility_pack.f:3049-3095)
__device__ bool accept_test(double eng, double rande, double volm,
                            double press, int ngsts, double temp, double beta,
                            int kind) {
  double test = 0.0;
  if (-1.0 * beta * eng > 700.0) {
    test = exp(700.0);
  } else if (kind == FMC_DISPLACE) {
    test = exp(-1.0 * beta * eng);
  } else if (kind == FMC_INSERT) {
    test = volm * press / (static_cast<double>(ngsts + 1) * kBoltz * temp) *
           exp(-1.0 * beta * eng);
  } else if (kind == FMC_DELETE) {
    if (fabs(press) < 1e-7)
      test = 1.0;
    else
      test = static_cast<double>(ngsts) * kBoltz * temp / (volm * press) *
             exp(-1.0 * beta * eng);
  }
  return rande < test;
}

__device__ double tail_of(const SysQ& Z, const int* n) {
  double t = Z.t0;
  for (int g = 0; g < Z.ngt; ++g) t += Z.tb[g] * n[g];
  for (int g = 0; g < Z.ngt; ++g)
    for (int h = 0; h < Z.ngt; ++h) t += Z.tA[g][h] * n[g] * n[h];
  return t;
}

// The move as the serial lane hands it to the work-group.
struct MoveDesc {
  int kind;     // 0 no-op, FMC_INSERT/DELETE/DISPLACE
  int sub;      // 0 plain, 4 translate-only, 5 rotate-only
  int ig, imol, ns, slot;
  int accepted;
  double ox[kMaxSites], oy[kMaxSites], oz[kMaxSites];
  double nx[kMaxSites], ny[kMaxSites], nz[kMaxSites];
  // flattened real-space work: position p = (set, site) of the trial
  int np;                       // active positions (ns or 2 ns)
  int nguest;                   // guest atoms of every type
  int gpre[kMaxGuestTypes + 1]; // guest-type prefix over atoms
  int pset[2 * kMaxSites], psite[2 * kMaxSites];
  int plo[2 * kMaxSites];        // first candidate record of the cell
  // per position: coordinates, q/epsq*r4pie0, site type, sign of its terms
  double ppx[2 * kMaxSites], ppy[2 * kMaxSites], ppz[2 * kMaxSites];
  double pce[2 * kMaxSites];
  double pq[2 * kMaxSites];     // +q (new set) or -q (old set)
  int pts[2 * kMaxSites], psg[2 * kMaxSites];
  // pairtab row of each position's site type (ntpatm <= kLdsTypes)
  short ptab[2 * kMaxSites][kLdsTypes];
};

__device__ __forceinline__ int cell_of(const SysQ& Z, double x, double y,
                                       double z) {
  const double* rc = Z.rcell;
  int idx[3];
  for (int d = 0; d < 3; ++d) {
    double f = rc[d] * x + rc[3 + d] * y + rc[6 + d] * z;
    f -= floor(f);
    int i = static_cast<int>(f * Z.gdim[d]);
    idx[d] = i >= Z.gdim[d] ? Z.gdim[d] - 1 : (i < 0 ? 0 : i);
  }
  return (idx[0] * Z.gdim[1] + idx[1]) * Z.gdim[2] + idx[2];
}

template <typename T>
struct Cplx {
  T x, y;
};

__device__ __forceinline__ float fmat(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fmat(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

// dS(k) of the trial's positions: sum_p (q_p X_p[l] Y_p[m]) Z_p[n], with
// q X Y tabulated per (l, m) and Z per signed n (conjugates folded in); four
// multiply-adds per position, accumulated in T.  Both kernel forms call
// this one routine (contraction off, fused multiply-adds written out), so
// they round identically.
constexpr int kZRow = 2 * kBTab;   // Z table row: n = -16..16
template <typename T>
__device__ __forceinline__ void trial_ds(int np, const Cplx<T>* xy, int nlm, int lm,
                                         const Cplx<T>* zt, int zoff, T& dr, T& di) {
#pragma clang fp contract(off)
  T ar = 0, ai = 0;
  for (int p = 0; p < np; ++p) {
    const Cplx<T> a = xy[p * nlm + lm];
    const Cplx<T> c = zt[p * kZRow + zoff];
    ar = fmat(a.x, c.x, ar);
    ar = fmat(-a.y, c.y, ar);
    ai = fmat(a.y, c.x, ai);
    ai = fmat(a.x, c.y, ai);
  }
  dr = ar;
  di = ai;
}

// trial_ds for the k-vectors a thread owns, kpt of its KPT slots (kpt the
// same on every thread), positions outer: the LDS reads of one position for
// a group of four k issue together and wait once, rather than once per
// (k, position).  Each k accumulates over the positions in the same order
// as trial_ds, so the sums are the same.  A slot of a group beyond the
// thread's k-vectors carries index 0 and its sum is never used.
template <typename T, int KPT>
__device__ __forceinline__ void trial_ds_tile(int np, const Cplx<T>* xy, int nlm,
                                              const Cplx<T>* zt, int zoff, int kpt,
                                              const int (&kv)[KPT], T (&dr)[KPT],
                                              T (&di)[KPT]) {
#pragma clang fp contract(off)
  constexpr int G = 4;
  constexpr int NG = (KPT + G - 1) / G;
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    dr[j] = 0;
    di[j] = 0;
  }
  for (int p = 0; p < np; ++p) {
    const Cplx<T>* xp = xy + p * nlm;
    const Cplx<T>* zp = zt + p * kZRow + zoff;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      if (g * G < kpt) {
#pragma unroll
        for (int jj = 0; jj < G; ++jj) {
          const int j = g * G + jj;
          if (j >= KPT) break;
          const Cplx<T> a = xp[kv[j] & 0xffff];
          const Cplx<T> c = zp[static_cast<signed char>((kv[j] >> 16) & 0xff)];
          dr[j] = fmat(a.x, c.x, dr[j]);
          dr[j] = fmat(-a.y, c.y, dr[j]);
          di[j] = fmat(a.y, c.x, di[j]);
          di[j] = fmat(a.x, c.y, di[j]);
        }
      }
    }
  }
}

// a_k (|S + dS|^2 - |S|^2) = a_k (2 S.dS + |dS|^2): no cancellation of |S|^2
__device__ __forceinline__ double de_k(double ak, double s0, double s1, double dr, double di) {
#pragma clang fp contract(off)
  const double sd = fmat(s0, dr, s1 * di);
  const double dd = fmat(dr, dr, di * di);
  return ak * fmat(2.0, sd, dd);
}

// a / d for 0 <= a < 2^22 and 0 < d < 2^11 from rd = 1/d: the float
// quotient is within one of the true one, and one correction step fixes it
// (a runtime integer division costs tens of instructions)
__device__ __forceinline__ int div_rc(int a, int d, float rd) {
  int q = static_cast<int>(static_cast<float>(a) * rd);
  const int r = a - q * d;
  q += (r >= d) - (r < 0);
  return q;
}

__device__ __forceinline__ int ul(int v) { return static_cast<signed char>(v & 0xff); }
__device__ __forceinline__ int um(int v) { return static_cast<signed char>((v >> 8) & 0xff); }
__device__ __forceinline__ int un(int v) { return static_cast<signed char>((v >> 16) & 0xff); }

// ---------------------------------------------------------------------------
// The proposal: move selection, adaptive steps, draws and trial positions
// (gcmc.f:745-847, 855-877, 1010-1060, 1176-1248; utility_pack.f:2545-3047)
// ---------------------------------------------------------------------------
// The proposal as one wave computes it: every lane runs the same scalar
// steps on the same values (wave-uniform, so no lane waits on another), lane
// i holds draw i of the lookahead and lane s works on guest site s.  Lane 0
// writes the chain's counters and the descriptor's scalars.  The
// arithmetic and the draw order are those of the reference's sequential
// form; the centre of mass is summed over the sites in order.
__device__ void propose(const SysQ& Z, BChain& C, Rng& R, MoveDesc& D, int lane) {
  const double delrmin = 0.1, minangle = kPi / 18.0;
  const Ahead A = rng_ahead(R, lane);
  // draw 0 picks the guest type when there is more than one; draw g0 + j is
  // the j-th after it
  const int g0 = uni_i(C.ntpguest > 1 ? 1 : 0);
  int ig = 0;
  if (g0) ig = uni_i(static_cast<int>(floor(static_cast<double>(lane_f(A.d, 0)) * C.ntpguest)));
  const double randmov = lane_f(A.d, g0);
  const double* cum = C.cum[ig];
  int kind = 0, sub = 0;
  if (randmov < cum[0]) kind = FMC_INSERT;
  else if (randmov < cum[1]) kind = FMC_DELETE;
  else if (randmov < cum[2]) kind = FMC_DISPLACE;
  else if (randmov < cum[5]) kind = 0;        // jump (rejected at set-up), flex, swap
  else if (randmov < cum[6]) { kind = FMC_DISPLACE; sub = 4; }
  else if (randmov < cum[7]) { kind = FMC_DISPLACE; sub = 5; }
  else if (randmov < cum[8]) kind = 0;
  else kind = FMC_DISPLACE;
  const int ns = uni_i(Z.ns[ig]);
  const int nmols = uni_i(C.nmol[ig]);
  if (lane == 0) {
    if (nmols >= 1 && C.disp_count[ig] >= 1) {
      if (C.disp_count[ig] % 100 == 0) {
        if ((C.accept_disp[ig] / static_cast<double>(C.disp_count[ig])) > C.disp_ratio[ig])
          C.delrdisp[ig] = C.delrdisp[ig] * 1.05;
        else if (C.delrdisp[ig] > delrmin)
          C.delrdisp[ig] = C.delrdisp[ig] * 0.95;
      } else if (C.disp_count[ig] % 50 == 0) {
        if ((C.accept_disp[ig] / static_cast<double>(C.disp_count[ig])) > C.disp_ratio[ig])
          C.rotangle = C.rotangle * 1.05;
        else if (C.rotangle > minangle)
          C.rotangle = C.rotangle * 0.95;
      }
    }
    if (nmols >= 1 && C.tran_count[ig] >= 1 && C.tran_count[ig] % 100 == 0) {
      if ((C.accept_tran[ig] / static_cast<double>(C.tran_count[ig])) > C.tran_ratio[ig])
        C.tran_delr[ig] = C.tran_delr[ig] * 1.05;
      else if (C.tran_delr[ig] > delrmin)
        C.tran_delr[ig] = C.tran_delr[ig] * 0.95;
    }
    if (nmols >= 1 && C.rota_count >= 1 && C.rota_count % 100 == 0) {
      if ((C.accept_rota / static_cast<double>(C.rota_count)) > C.rota_ratio)
        C.rota_rotangle = C.rota_rotangle * 1.05;
      else if (C.rota_rotangle > minangle)
        C.rota_rotangle = C.rota_rotangle * 0.95;
    }
  }
  if (nmols == 0) { kind = FMC_INSERT; sub = 0; }
  kind = uni_i(kind);
  sub = uni_i(sub);
  // guest atoms of every type (the capacity check and the descriptor)
  int na = 0;
  for (int g = 0; g < Z.ngt; ++g) {
    if (lane == 0) D.gpre[g] = na;
    na += C.nmol[g] * Z.ns[g];
  }
  na = uni_i(na);
  if (lane == 0) {
    D.gpre[Z.ngt] = na;
    D.nguest = na;
    D.kind = kind;
    D.sub = sub;
    D.ig = ig;
    D.ns = ns;
    D.imol = -1;
    D.accepted = 0;
  }
  int taken = g0 + 1;   // draws taken (type, move)
  const bool site = lane < ns;
  if (kind == FMC_INSERT) {
    if (na + ns >= C.cap_atoms) {
      if (lane == 0) {
        C.fault = 1;
        D.kind = 0;
        D.np = 0;
      }
      rng_commit(R, A, lane, taken);
      return;
    }
    const double ra = lane_f(A.d, g0 + 1), rb = lane_f(A.d, g0 + 2), rc = lane_f(A.d, g0 + 3);
    const double r1 = lane_f(A.d, g0 + 4), r2 = lane_f(A.d, g0 + 5),
                 r3 = lane_f(A.d, g0 + 6), r4 = lane_f(A.d, g0 + 7);
    taken += 7;
    const double* cell = Z.cell;
    const double xc = ra * cell[0] + rb * cell[3] + rc * cell[6];
    const double yc = ra * cell[1] + rb * cell[4] + rc * cell[7];
    const double zc = ra * cell[2] + rb * cell[5] + rc * cell[8];
    double Q[9];
    rot_matrix(r1 * 2.0 * kPi, r2, r3, r4, Q);
    if (site) {
      double x = Z.tx[ig][lane], y = Z.ty[ig][lane], z = Z.tz[ig][lane];
      rot_apply(Q, x, y, z);
      D.nx[lane] = x + xc;
      D.ny[lane] = y + yc;
      D.nz[lane] = z + zc;
    }
    if (lane == 0) D.slot = nmols * ns;
  } else if (kind == FMC_DELETE) {
    const int imol = uni_i(static_cast<int>(floor(static_cast<double>(lane_f(A.d, g0 + 1)) * nmols)));
    taken += 1;
    if (lane == 0) D.imol = imol;
    if (site) {
      D.ox[lane] = gload(C.gx[ig] + imol * ns + lane);
      D.oy[lane] = gload(C.gy[ig] + imol * ns + lane);
      D.oz[lane] = gload(C.gz[ig] + imol * ns + lane);
    }
  } else if (kind == FMC_DISPLACE) {
    double ra, rb, rc, r1, r2, r3, r4, dd, da, rm;
    if (sub == 4) {
      if (lane == 0) C.tran_count[ig]++;
      ra = lane_f(A.d, g0 + 1); rb = lane_f(A.d, g0 + 2); rc = lane_f(A.d, g0 + 3);
      r1 = r2 = r3 = r4 = 0.0;
      dd = C.tran_delr[ig]; da = 0.0;
      rm = lane_f(A.d, g0 + 4);
      taken += 4;
    } else if (sub == 5) {
      if (lane == 0) C.rota_count++;
      ra = rb = rc = 0.5;
      r1 = lane_f(A.d, g0 + 1); r2 = lane_f(A.d, g0 + 2);
      r3 = lane_f(A.d, g0 + 3); r4 = lane_f(A.d, g0 + 4);
      dd = 0.0; da = C.rota_rotangle;
      rm = lane_f(A.d, g0 + 5);
      taken += 5;
    } else {
      if (lane == 0) C.disp_count[ig]++;
      ra = lane_f(A.d, g0 + 1); rb = lane_f(A.d, g0 + 2); rc = lane_f(A.d, g0 + 3);
      r1 = lane_f(A.d, g0 + 4); r2 = lane_f(A.d, g0 + 5);
      r3 = lane_f(A.d, g0 + 6); r4 = lane_f(A.d, g0 + 7);
      dd = C.delrdisp[ig]; da = C.rotangle;
      rm = lane_f(A.d, g0 + 8);
      taken += 8;
    }
    const int imol = uni_i(static_cast<int>(floor(rm * nmols)));
    if (lane == 0) D.imol = imol;
    double x = 0.0, y = 0.0, z = 0.0, m = 0.0;
    if (site) {
      x = gload(C.gx[ig] + imol * ns + lane);
      y = gload(C.gy[ig] + imol * ns + lane);
      z = gload(C.gz[ig] + imol * ns + lane);
      m = Z.gmass[ig][lane];
      D.ox[lane] = x;
      D.oy[lane] = y;
      D.oz[lane] = z;
    }
    const double h = dd * 0.5;
    const double ta = (2.0 * ra - 1.0) * h, tb = (2.0 * rb - 1.0) * h,
                 tc = (2.0 * rc - 1.0) * h;
    x = ta + x;
    y = tb + y;
    z = tc + z;
    double w = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int s = 0; s < ns; ++s) {
      const double ms = lane_d(m, s);
      w = w + ms;
      cx = cx + lane_d(x, s) * ms;
      cy = cy + lane_d(y, s) * ms;
      cz = cz + lane_d(z, s) * ms;
    }
    cx = cx / w; cy = cy / w; cz = cz / w;
    const double* rcl = Z.rcell;
    const double* cell = Z.cell;
    const double sx = (rcl[0] * cx + rcl[3] * cy + rcl[6] * cz);
    const double sy = (rcl[1] * cx + rcl[4] * cy + rcl[7] * cz);
    const double sz = (rcl[2] * cx + rcl[5] * cy + rcl[8] * cz);
    const double xs = sx - floor(sx), ys = sy - floor(sy), zs = sz - floor(sz);
    const double X = (cell[0] * xs + cell[3] * ys + cell[6] * zs);
    const double Yv = (cell[1] * xs + cell[4] * ys + cell[7] * zs);
    const double Zv = (cell[2] * xs + cell[5] * ys + cell[8] * zs);
    x = x + (X - cx);
    y = y + (Yv - cy);
    z = z + (Zv - cz);
    x = x - X;
    y = y - Yv;
    z = z - Zv;
    double Q[9];
    rot_matrix(r1 * da, r2, r3, r4, Q);
    rot_apply(Q, x, y, z);
    if (site) {
      D.nx[lane] = x + X;
      D.ny[lane] = y + Yv;
      D.nz[lane] = z + Zv;
    }
  }
  // trial positions: the old sites, then the new ones
  int np = 0;
  if (kind != 0) {
    const int nold = kind != FMC_INSERT ? ns : 0;
    np = nold + (kind != FMC_DELETE ? ns : 0);
    if (lane < np) {
      D.pset[lane] = lane < nold ? 0 : 1;
      D.psite[lane] = lane < nold ? lane : lane - nold;
    }
  }
  if (lane == 0) D.np = np;
  rng_commit(R, A, lane, taken);
}

// Production count at which snapshot k of the running sums is taken.
__host__ __device__ inline long long win_target(int k, int L) {
  return k < 2 * kWin - 2 ? static_cast<long long>(k / 2 + 1) * L + (k & 1)
                          : static_cast<long long>(kWin) * L + 1;
}

__device__ void window_snapshot(BChain& C, double* snaps) {
  while (C.win_snap < kWinSnap && C.prod_steps == win_target(C.win_snap, C.win_len)) {
    double* w = snaps + C.win_snap * kWinQ * kMaxGuestTypes;
    for (int g = 0; g < kMaxGuestTypes; ++g) {
      gstore(w + 0 * kMaxGuestTypes + g, C.sum_n[g]);
      gstore(w + 1 * kMaxGuestTypes + g, C.sum_e[g]);
      gstore(w + 2 * kMaxGuestTypes + g, C.sum_en[g]);
      gstore(w + 3 * kMaxGuestTypes + g, C.sum_n2[g]);
      gstore(w + 4 * kMaxGuestTypes + g, C.sum_e2[g]);
    }
    C.win_snap++;
  }
}

// A selectivity window closes (gcmc.f:1887-1916 with weight 1): its <N_g>
// are the production sums since the last close over the window's samples;
// every closed window counts, a pair whose <N_i> is zero is not updated.
__device__ __forceinline__ void selectivity_window(BChain& C) {
#pragma clang fp contract(off)
  C.sel_count++;
  const double av = static_cast<double>(C.sel_count);
  const double nw = static_cast<double>(C.sel_nwind);
  const double weight = 1.0;
  for (int i = 1; i < C.ntpguest; ++i) {
    const double aN = (C.sum_n[i] - C.sel_base[i]) / nw;
    if (aN == 0.0) continue;
    for (int j = 0; j < i; ++j) {
      const int p = i * (i - 1) / 2 + j;
      const double sij = ((C.sum_n[j] - C.sel_base[j]) / nw) / aN;
      const double delta = weight * (sij - C.sel_mean[p]);
      C.sel_mean[p] = C.sel_mean[p] + delta / av;
      C.sel_sec[p] = C.sel_sec[p] + weight * (sij * sij - C.sel_sec[p]) / (av - 1.0 + weight);
      const double w = av - 1.0 + weight;
      C.sel_var[p] = (w * C.sel_sec[p] - (w * C.sel_mean[p] * C.sel_mean[p])) / w;
    }
  }
  for (int g = 0; g < C.ntpguest; ++g) C.sel_base[g] = C.sum_n[g];
}

// Production statistics of one move (gcmc.f:1814-1848), after C.steps has
// counted it; true when it is a production sample.
__device__ __forceinline__ bool production_sample(BChain& C) {
  if (C.steps <= C.eq_steps) return false;
  C.prod_steps++;
  for (int g = 0; g < C.ntpguest; ++g) {
    const double n = C.nmol[g];
    const double e = C.energy[g];
    C.sum_n[g] += n;
    C.sum_n2[g] += n * n;
    C.sum_e[g] += e;
    C.sum_en[g] += e * n;
    C.sum_e2[g] += e * e;
  }
  if (C.ntpguest > 1 && C.prod_steps % C.sel_nwind == 0) selectivity_window(C);
  return true;
}

__device__ __forceinline__ void log_move(BChain& C, int slot, int ig, int acc, double estep,
                                         double e1, double e2, double ev) {
  if (C.log && C.log_n < C.log_cap) {
    double* r = C.log + kLogFields * C.log_n;
    r[0] = slot;
    r[1] = ig;
    r[2] = acc;
    r[3] = C.nmol[ig];
    r[4] = estep;
    r[5] = e1;
    r[6] = e2;
    r[7] = ev;
    C.log_n++;
  }
}

__device__ __forceinline__ int move_slot(const MoveDesc& D) {
  return D.kind == 0 ? 0 : (D.sub ? D.sub : D.kind);
}

// The decision alone (decide's arithmetic, gcmc.f:938-949, 1106-1116,
// 1326-1338), with the acceptance draw given: estep, the logged e1 and
// whether the move is accepted.  delrc is the tail change of an insert or
// delete.
__device__ __forceinline__ bool decision(const SysQ& Z, const BChain& C, const MoveDesc& D,
                                         double e1, double e2, double ev, double delrc,
                                         double rande, double& estep, double& e1out) {
  const int ig = D.ig;
  const int nmols = C.nmol[ig];
  bool acc = false;
  estep = 0.0;
  if (D.kind == FMC_INSERT || D.kind == FMC_DELETE) {
    if (D.kind == FMC_INSERT) {
      e1 = e1 + Z.engsic[ig];
      estep = (e1 + e2 + Z.e3[ig] + ev + delrc) / C.engunit;
    } else {
      e1 = e1 - Z.engsic[ig];
      estep = (e1 - e2 - Z.e3[ig] - ev + delrc) / C.engunit;
    }
    acc = accept_test(estep, rande, C.statvolm, C.press[ig], nmols, C.temp, C.beta, D.kind);
  } else if (D.kind == FMC_DISPLACE) {
    estep = (e1 + e2 + ev) / C.engunit;
    acc = estep < 0.0 ||
          accept_test(estep, rande, C.statvolm, C.press[ig], nmols, C.temp, C.beta,
                      FMC_DISPLACE);
  }
  e1out = e1;
  return acc;
}

// A bound on estep above which the move is certainly rejected for the
// acceptance draw rande: energy_eval accepts when rande < pref exp(-beta
// estep) (pref 1 for a displacement, the volume-pressure factor for an
// insert or delete), so estep > E* = -ln(rande / pref) / beta rejects.  The
// bound adds a margin of 1e-6 (1 + |E*|), far beyond the rounding of exp and
// log; a move at or below it is decided exactly.  Returns false when no
// bound applies (rande 0, a tiny or zero prefactor, the zero-pressure
// delete, a no-op).
__device__ bool reject_bound(const BChain& C, const MoveDesc& D, double rande, double* thr) {
  const int ig = D.ig;
  double pref;
  if (D.kind == FMC_DISPLACE) {
    pref = 1.0;
  } else if (D.kind == FMC_INSERT) {
    pref = C.statvolm * C.press[ig] / (static_cast<double>(C.nmol[ig] + 1) * kBoltz * C.temp);
  } else if (D.kind == FMC_DELETE) {
    if (fabs(C.press[ig]) < 1e-7) return false;
    pref = static_cast<double>(C.nmol[ig]) * kBoltz * C.temp / (C.statvolm * C.press[ig]);
  } else {
    return false;
  }
  if (!(rande > 0.0) || !(pref > 1e-300) || !(pref < 1e300)) return false;
  const double e = -log(rande / pref) / C.beta;
  *thr = e + 1e-6 * (1.0 + fabs(e));
  return true;
}

__device__ __forceinline__ double tail_change(const SysQ& Z, const BChain& C, const MoveDesc& D) {
  int n2[kMaxGuestTypes];
  for (int g = 0; g < Z.ngt; ++g) n2[g] = C.nmol[g];
  n2[D.ig] += (D.kind == FMC_INSERT) ? 1 : -1;
  return tail_of(Z, n2) - C.elrc;
}

// Acceptance and bookkeeping on the serial lane (gcmc.f:938-1002, 1106-1172,
// 1326-1378, 1814-1870).
__device__ void decide(const SysQ& Z, BChain& C, Rng& R, MoveDesc& D,
                       double e1, double e2, double ev, int nk) {
  const int ig = D.ig;
  const int nmols = C.nmol[ig];
  double estep = 0.0, delrc = 0.0;
  bool acc = false;
  const int slot = move_slot(D);
  if (D.kind == FMC_INSERT || D.kind == FMC_DELETE) {
    delrc = tail_change(Z, C, D);
    if (D.kind == FMC_INSERT) {
      e1 = e1 + Z.engsic[ig];
      estep = (e1 + e2 + Z.e3[ig] + ev + delrc) / C.engunit;
    } else {
      e1 = e1 - Z.engsic[ig];
      estep = (e1 - e2 - Z.e3[ig] - ev + delrc) / C.engunit;
    }
    const double rande = R.next();
    acc = accept_test(estep, rande, C.statvolm, C.press[ig], nmols, C.temp,
                      C.beta, D.kind);
    if (acc) {
      C.elrc = C.elrc + delrc;
      C.nmol[ig] = nmols + (D.kind == FMC_INSERT ? 1 : -1);
    }
  } else if (D.kind == FMC_DISPLACE) {
    estep = (e1 + e2 + ev) / C.engunit;
    if (estep < 0.0) {
      acc = true;
    } else {
      const double rande = R.next();
      acc = accept_test(estep, rande, C.statvolm, C.press[ig], nmols, C.temp,
                        C.beta, FMC_DISPLACE);
    }
    if (acc) {
      if (D.sub == 4) C.accept_tran[ig]++;
      else if (D.sub == 5) C.accept_rota++;
      else C.accept_disp[ig]++;
    }
  }
  if (acc) {
    C.energy[ig] += estep;
    C.e_acc += estep;
  }
  C.attempted[slot]++;
  if (acc) C.accepted[slot]++;
  D.accepted = acc ? 1 : 0;
  C.steps++;
  if (production_sample(C) && C.win_len > 0) window_snapshot(C, C.sre + nk);
  log_move(C, slot, ig, D.accepted, estep, e1, e2, ev);
}

// Exact S(k) of a chain: framework part plus every guest atom, FP64 with
// the phase reduced to [-1/2, 1/2] before sincospi (ewald.f:368-630 for the
// guest atoms).  One code path for the recompute kernel and the batch
// kernel's in-launch recompute, so both give the same bits.
__device__ __forceinline__ void exact_sk(const BSys& Y, const SysQ& Z, const BChain& C,
                                         int k, double& sx, double& sy) {
  const int v = Y.kv[k];
  const int l = ul(v), m = um(v), n = un(v);
  const double* rc = Z.rcell;
  sx = Y.sfw_re[k];
  sy = Y.sfw_im[k];
  for (int g = 0; g < Z.ngt; ++g) {
    const int nat = C.nmol[g] * Z.ns[g];
    for (int a = 0; a < nat; ++a) {
      const double x = C.gx[g][a], y = C.gy[g][a], z = C.gz[g][a];
      double ph = l * (rc[0] * x + rc[3] * y + rc[6] * z) +
                  m * (rc[1] * x + rc[4] * y + rc[7] * z) +
                  n * (rc[2] * x + rc[5] * y + rc[8] * z);
      ph -= rint(ph);
      double sv, cv;
      sincospi(2.0 * ph, &sv, &cv);
      const double q = Z.gq[g][a % Z.ns[g]];
      sx += q * cv;
      sy += q * sv;
    }
  }
}

// Real-space energy of all guests (single_point's pair part): guest atom i
// against its cell's framework candidates and every later guest atom of
// another molecule; threads tid, tid+nt, ... share the pairs of each i.
__device__ __forceinline__ void guest_pairs(const BSys& Y, const SysQ& Z, const BChain& C,
                                            int tid, int nt, double& e2, double& ev) {
  int gpre[kMaxGuestTypes + 1];
  int na = 0;
  for (int g = 0; g < kMaxGuestTypes; ++g) {
    gpre[g] = na;
    if (g < Z.ngt) na += C.nmol[g] * Z.ns[g];
  }
  gpre[kMaxGuestTypes] = na;
  for (int i = 0; i < na; ++i) {
    int g = 0;
    while (g + 1 < Z.ngt && gpre[g + 1] <= i) ++g;
    const int la = i - gpre[g];
    const int s = la % Z.ns[g];
    const int mol = la / Z.ns[g];
    const double xi = C.gx[g][la], yi = C.gy[g][la], zi = C.gz[g][la];
    const double ce = Z.gq[g][s] / Z.epsq * kR4pie0;
    const int ts = Z.gtype[g][s];
    const int c = cell_of(Z, xi, yi, zi);
    const int lo = c * Y.gcap;
    for (int t = tid; t < Y.gcap; t += nt) {
      const double* rr = &Y.grec[lo + t].x;
      pair_fp64(Y.S, ce, ts, xi - gload(rr), yi - gload(rr + 1), zi - gload(rr + 2),
                gload(rr + 3), gload(Y.grec_t + lo + t), e2, ev);
    }
    for (int j = i + 1 + tid; j < na; j += nt) {
      int h = 0;
      while (h + 1 < Z.ngt && gpre[h + 1] <= j) ++h;
      const int lb = j - gpre[h];
      if (h == g && lb / Z.ns[h] == mol) continue;
      const int sb = lb % Z.ns[h];
      pair_fp64(Y.S, ce, ts, xi - C.gx[h][lb], yi - C.gy[h][lb], zi - C.gz[h][lb],
                Z.gq[h][sb], Z.gtype[h][sb], e2, ev);
    }
  }
}

// Thread 0's part of a recompute: the FP64 totals into the chain's
// bookkeeping (see recompute_kernel).
__device__ __forceinline__ void recompute_finish(const SysQ& Z, BChain& C, double erec,
                                                 double e2, double ev, double dev, int base) {
  double e = Z.rvolm2 * kR4pie0 * erec / Z.epsq;
  for (int g = 0; g < Z.ngt; ++g) e += C.nmol[g] * (Z.engsic[g] + Z.e3[g]);
  e += e2 + ev;
  // the move loop's tail bookkeeping starts from elrc = 0 and adds
  // delrc = tail(N +/- 1) - elrc (gcmc.f:891-901): compare the running
  // sum with the same tail it carries, report the true one in e_full
  const double book = (e + C.elrc) / C.engunit;
  C.e_full = (e + tail_of(Z, C.nmol)) / C.engunit;
  if (base) {
    C.e_acc = book;
    for (int g = 0; g < Z.ngt; ++g) C.energy[g] = C.e_full;
  }
  C.drift = C.e_acc - book;
  C.sk_dev = dev;
  C.n_recompute++;
}

// The per-move tables of a trial's positions, one position per wave (waves
// w < nw of the caller's range take positions p = w, w + nw, ...): its
// phase factors e^{i 2 pi l f} for every dimension and l (ewald.f:97-137,
// one sincospi per lane, into the wave's own rows of tab), then, read back
// within the wave, its q X[l] Y[m] row and signed Z[n] row for the
// reciprocal sum, its pair-table row, and its coordinates, charge factors
// and candidate-list start for real space.  Positions are independent, so
// no work-group barrier is needed inside; the caller's next barrier
// publishes the tables.
//
// A displacement in DELTA form (mixed precision) sums, per site, the change
// of q e^{ik.r} rather than the new term minus the old one:
//   q (P_n Z_n - P_o Z_o) = q P_o (Z_n - Z_o) + q (P_n - P_o) Z_n,
// with P = X[l] Y[m].  The old position's rows are (q P_o, Z_n - Z_o), the
// new position's (q (P_n - P_o), Z_n).  The differences are formed from
// FP64 phase factors in FP64 and only then rounded to T, so their rounding
// error is relative to the change of the term, not to the term: the FP32
// difference of two nearly equal phase factors would lose the digits that
// the framework's large |S(k)| amplifies in 2 a_k S.dS.  Each position's
// wave computes the partner position's phase factors it needs into tabp.
template <typename T, bool DELTA>
__device__ __forceinline__ void build_tables(const SysQ& Z, const BSys& Y, MoveDesc& D,
                                             int wave, int nw, int lane,
                                             Cplx<double> (*tab)[3][kBTab],
                                             Cplx<double> (*tabp)[3][kBTab], Cplx<T>* xy_l,
                                             Cplx<T>* zt, int nlm) {
  const int np = D.np, ig = D.ig, kind = D.kind;
  const int ntp = Y.S.ntpatm;
  const int w2 = 2 * Z.kmax[1] + 1, k3 = Z.kmax[2], nz = 2 * k3 + 1;
  const float rw2 = 1.0f / w2;
  const bool delta = DELTA && kind == FMC_DISPLACE;
  for (int p = wave; p < np; p += nw) {
    const int set = D.pset[p], s = D.psite[p];
    const double x = set ? D.nx[s] : D.ox[s];
    const double y = set ? D.ny[s] : D.oy[s];
    const double z = set ? D.nz[s] : D.oz[s];
    const double q = Z.gq[ig][s];
    const double pq = (set || delta) ? q : -q;
    const int ts = Z.gtype[ig][s];
    // one sincospi per lane: first the position's own phase factors (kmax+1
    // per dimension, packed), then, in delta form, the partner position's
    // that the rows below need (the new Z for the old position, the old X
    // and Y for the new one); both fit one pass of the wave while the
    // entries number at most 64
    const int k0 = Z.kmax[0] + 1, k1 = Z.kmax[1] + 1, k2 = Z.kmax[2] + 1;
    const int nown = k0 + k1 + k2;
    const int d0 = set ? 0 : 2;   // the partner's first dimension
    const int npart = delta ? (set ? k0 + k1 : k2) : 0;
    for (int j = lane; j < nown + npart; j += kWave) {
      const bool own = j < nown;
      int d, l;
      if (own) {
        d = j < k0 ? 0 : (j < k0 + k1 ? 1 : 2);
        l = j - (d == 0 ? 0 : (d == 1 ? k0 : k0 + k1));
      } else {
        const int jj = j - nown;
        d = d0 + (d0 == 0 && jj >= k0 ? 1 : 0);
        l = jj - (d == 1 ? k0 : 0);
      }
      const double px = own ? x : (set ? D.ox[s] : D.nx[s]);
      const double py = own ? y : (set ? D.oy[s] : D.ny[s]);
      const double pz = own ? z : (set ? D.oz[s] : D.nz[s]);
      double f = Z.rcell[d] * px + Z.rcell[3 + d] * py + Z.rcell[6 + d] * pz;
      f = f - floor(f);
      double ph = l * f;
      ph = ph - rint(ph);
      double sv, cv;
      sincospi(2.0 * ph, &sv, &cv);
      if (!DELTA || own)
        tab[p][d][l] = {cv, sv};
      else
        tabp[p][d][l] = {cv, sv};
    }
    if (lane == 0) {
      D.plo[p] = cell_of(Z, x, y, z) * Y.gcap;
      D.ppx[p] = x;
      D.ppy[p] = y;
      D.ppz[p] = z;
      D.pce[p] = q / Z.epsq * kR4pie0;
      D.pq[p] = pq;
      D.pts[p] = ts;
      D.psg[p] = (kind == FMC_DISPLACE && set == 0) ? -1 : 1;
    }
    if (ntp <= kLdsTypes)
      for (int tj = lane; tj < ntp; tj += kWave)
        D.ptab[p][tj] = static_cast<short>(Y.S.pairtab[ts * ntp + tj]);
    // the phase factors of this position, written by this wave, read back
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // X[l] * Y[m] for every (l, m) (m < 0 conjugates Y); in delta form the
    // new position's row is X_n Y_n - X_o Y_o
    Cplx<T>* xr = xy_l + p * nlm;
    const bool dxy = delta && set;
    for (int r = lane; r < nlm; r += kWave) {
      const int l = div_rc(r, w2, rw2), m = r - l * w2 - Z.kmax[1];
      const int am = m < 0 ? -m : m;
      const Cplx<double> a = tab[p][0][l];
      const Cplx<double> b = tab[p][1][am];
      double ox, oy;
      if (m >= 0) {
        ox = a.x * b.x - a.y * b.y;
        oy = a.y * b.x + b.y * a.x;
      } else {
        ox = a.x * b.x + a.y * b.y;
        oy = a.y * b.x - b.y * a.x;
      }
      if (dxy) {
        const Cplx<double> c = tabp[p][0][l];
        const Cplx<double> e = tabp[p][1][am];
        if (m >= 0) {
          ox -= c.x * e.x - c.y * e.y;
          oy -= c.y * e.x + e.y * c.x;
        } else {
          ox -= c.x * e.x + c.y * e.y;
          oy -= c.y * e.x - e.y * c.x;
        }
      }
      xr[r] = {static_cast<T>(pq * ox), static_cast<T>(pq * oy)};
    }
    // signed Z[n]; in delta form the old position's row is Z_n - Z_o
    const bool dz = delta && !set;
    for (int i = lane; i < nz; i += kWave) {
      const int n = i - k3;
      const int an = n < 0 ? -n : n;
      Cplx<double> c = tab[p][2][an];
      if (dz) {
        const Cplx<double> e = tabp[p][2][an];
        c = {e.x - c.x, e.y - c.y};
      }
      zt[p * kZRow + i] = {static_cast<T>(c.x), static_cast<T>(n < 0 ? -c.y : c.y)};
    }
  }
}

// ---------------------------------------------------------------------------
// The persistent kernel
// ---------------------------------------------------------------------------
// STREAM: S(k), a_k and the k indices stay in global memory (L2-resident)
// and dS(k) goes to the chain's scratch; for Nk beyond what registers and
// LDS hold (kmax 16: Nk ~ 10^4).  Otherwise S(k) lives in registers, KPT
// per thread, a_k and k indices in LDS.  Both forms give a thread the same
// k = tid + j*BT (coalesced global and conflict-free LDS access) and reduce
// in the same order, so a chain's results do not depend on the form its
// launch used.  The register form keeps the trial's S + dS in registers
// too; the streaming form writes dS to the chain's scratch.
template <int BT, int KPT, bool MIXED, bool STREAM>
__global__ void __launch_bounds__(BT)
    batch_kernel(const BSys* __restrict__ systems, BChain* __restrict__ chains,
                 const int* __restrict__ ids, long long nmoves, int nkcap,
                 long long budget_ticks) {
  using T = typename std::conditional<MIXED, float, double>::type;
  __shared__ Cplx<double> tab[2 * kMaxSites][3][kBTab];   // per position p
  // a displacement's partner phase factors (mixed precision: delta form)
  __shared__ Cplx<double> tabp[MIXED ? 2 * kMaxSites : 1][3][kBTab];
  __shared__ Cplx<T> zt[2 * kMaxSites * kZRow];           // Z_p[n], signed n
  // Move descriptors and duni tables, two of each: while the waves work on
  // one move, the last wave proposes the next one on copies of the chain
  // state as if the current move were rejected (most are); the serial lane
  // adopts that proposal when the move is indeed rejected.
  // The chain state is double-buffered as well: the last wave also books
  // the current move as rejected in the other buffer, so that a rejection
  // costs the serial lane its decision and a buffer flip.
  __shared__ MoveDesc Dbuf[2];
  __shared__ float rng_u[2][97];
  __shared__ BChain Cbuf[2];
  __shared__ int csel, dsel, rsel, spec_hit;   // current buffers, adopted?
  __shared__ float cur_c, spec_c;       // duni state after this / the next proposal
  __shared__ int cur_ir, cur_jr, spec_ir, spec_jr;
  __shared__ double pre_rande, pre_delrc;   // this move's draw and tail change
  __shared__ double pre_thr;                // estep above which it is rejected
  __shared__ int pre_fast;                  // pre_thr applies
  __shared__ double red[4][BT / kWave];
  __shared__ int stop_flag;
  __shared__ int shift_lo, shift_hi;
  __shared__ SysQ Z;   // the system's small constants, for the whole launch
  BChain& C = Cbuf[0];   // the launch's first state (the loop rebinds it)

  const int tid = threadIdx.x;
  // The chain's mutable state and the system's small constants are copied
  // to LDS for the launch; tables and the framework are read from global
  // memory.
  BChain& Cg = chains[ids[blockIdx.x]];
  const BSys& Y = systems[Cg.sys];
  {
    const int* src = reinterpret_cast<const int*>(&Cg);
    int* dst = reinterpret_cast<int*>(&C);
    for (int i = tid; i < static_cast<int>(sizeof(BChain) / 4); i += BT) dst[i] = src[i];
  }
  {
    const int* src = reinterpret_cast<const int*>(&Y.q);
    int* dst = reinterpret_cast<int*>(&Z);
    for (int i = tid; i < static_cast<int>(sizeof(SysQ) / 4); i += BT) dst[i] = src[i];
  }
  __syncthreads();
  const int nk = Y.nk;
  // dynamic LDS: a_k [nkcap], packed k index [nkcap] (register form), then
  // the XY tables
  extern __shared__ double dyn_lds[];
  double* ak_l = dyn_lds;
  int* kv_l = reinterpret_cast<int*>(dyn_lds + (STREAM ? 0 : nkcap));
  // X[l]*Y[m] for every (l, m) of the half space, per trial position:
  // [p][l*(2*kmax2+1) + m + kmax2]
  const int nlm = (Z.kmax[0] + 1) * (2 * Z.kmax[1] + 1);
  Cplx<T>* xy_l =
      reinterpret_cast<Cplx<T>*>(STREAM ? reinterpret_cast<int*>(dyn_lds)
                                        : kv_l + ((nkcap + 3) & ~3));
  // this thread owns k = tid + j*BT, j < kpt = ceil(nk/BT); the index
  // (l, m) is precomputed into kv's upper bits
  const int kpt = (nk + BT - 1) / BT;
  if (!STREAM) {
    for (int k = tid; k < nk; k += BT) {
      ak_l[k] = Y.ak[k];
      const int v = Y.kv[k];
      const int lm = ul(v) * (2 * Z.kmax[1] + 1) + um(v) + Z.kmax[1];
      kv_l[k] = (v & 0xff0000) | lm;   // n in bits 16..23, (l,m) index below
    }
  }
  double sre[KPT], sim[KPT];
  T dre[KPT], dim[KPT];   // the trial's dS, as trial_ds summed it (exact)
  if (!STREAM) {
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const int k = tid + j * BT;
      const bool in = j < kpt && k < nk;
      sre[j] = in ? C.sre[k] : 0.0;
      sim[j] = in ? C.sim[k] : 0.0;
    }
  }
  Rng R = {};
  // time budget (constant-rate counter): thread 0 raises stop_flag after
  // the move that crosses it; the flag is read after a barrier
  const long long t_start = (tid == 0 && budget_ticks > 0)
                                ? static_cast<long long>(__builtin_amdgcn_s_memrealtime())
                                : 0;
  if (tid == 0) stop_flag = 0;
  __syncthreads();
  if (tid == 0) {
    for (int i = 0; i < 97; ++i) rng_u[0][i] = C.u[i];
    R.u = rng_u[0];
    csel = 0;
    dsel = 0;
    rsel = 0;
    spec_hit = 0;
    R.c = C.c; R.cd = C.cd; R.cm = C.cm; R.ir = C.ir; R.jr = C.jr;
  }
  __syncthreads();

  // moves left until the next in-launch recompute (the chain's cadence)

  for (long long mv = 0; mv < nmoves; ++mv) {
    BChain& C = Cbuf[csel];
    // C.steps and stop_flag were last written before the barrier ending the
    // previous move: every thread takes the same branch
    if (C.steps >= C.stop_at || stop_flag) break;
    // a prepared move is in place already (written before the last
    // barrier); only a fresh proposal needs one
    const bool fresh = !spec_hit;
    if (tid < kWave) {
      if (fresh) {
        // wave 0 proposes; its lanes take thread 0's duni state
        R.u = rng_u[rsel];
        R.c = lane_f(R.c, 0);
        R.ir = __builtin_amdgcn_readlane(R.ir, 0);
        R.jr = __builtin_amdgcn_readlane(R.jr, 0);
        R.cd = C.cd;
        R.cm = C.cm;
        propose(Z, C, R, Dbuf[dsel], tid);
      }
      if (tid == 0) {
        cur_c = R.c;
        cur_ir = R.ir;
        cur_jr = R.jr;
      }
    }
    if (fresh) __syncthreads();
    // dsel is written only by thread 0's decision, after the reduction barrier
    const int dcur = dsel;
    MoveDesc& D = Dbuf[dcur];
    const int kind = D.kind;
    const int ns = D.ns;
    const int ig = D.ig;
    double e1 = 0.0, e2 = 0.0, ev = 0.0;
    if (kind != 0) {
      const int np = D.np;
      // the positions' tables: built by every wave for a fresh proposal;
      // a prepared one had them built while the previous move was decided
      if (fresh) {
        build_tables<T, MIXED>(Z, Y, D, tid / kWave, BT / kWave, tid % kWave, tab, tabp, xy_l,
                               zt, nlm);
        __syncthreads();
      }
      const int ntp = Y.S.ntpatm;
      const bool ldsrow = ntp <= kLdsTypes;
      // ---- reciprocal space (ewald1_guest, ewald.f:141-319) ----
      // dS(k) = sum_new q e^{ik.r} - sum_old q e^{ik.r}
      if (STREAM) {
        const int w2 = 2 * Z.kmax[1] + 1;
        for (int k = tid; k < nk; k += BT) {
          const int v = Y.kv[k];
          const int lm = ul(v) * w2 + um(v) + Z.kmax[1];
          T dr32, di32;
          trial_ds<T>(np, xy_l, nlm, lm, zt, un(v) + Z.kmax[2], dr32, di32);
          const double dr = dr32, di = di32;
          const double s0 = gload(C.sre + k), s1 = gload(C.sim + k);
          gstore(C.dsre + k, dr);
          gstore(C.dsim + k, di);
          e1 += de_k(Y.ak[k], s0, s1, dr, di);
        }
      } else {
        // fully unrolled (predicated, no early exit) so that sre/sim/dre/dim
        // keep static register indices
        int kvj[KPT];
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
          const int k = tid + j * BT;
          kvj[j] = (j < kpt && k < nk) ? kv_l[k] : 0;
        }
        trial_ds_tile<T, KPT>(np, xy_l, nlm, zt, Z.kmax[2], __builtin_amdgcn_readfirstlane(kpt),
                              kvj, dre, dim);
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
          const int k = tid + j * BT;
          if (j < kpt && k < nk) {
            const double dr = dre[j], di = dim[j];
            e1 += de_k(ak_l[k], sre[j], sim[j], dr, di);
          }
        }
      }
      // ---- the next move's proposal, by the last wave, on copies of the
      //      chain state and duni table as this move leaves them if it is
      //      rejected (its acceptance draw taken, if it makes one) ----
      constexpr int RT = BT - kWave;   // threads on real space
      if (tid >= RT) {
        const int lane = tid - RT;
        BChain& Cn = Cbuf[csel ^ 1];
        const long long* src = reinterpret_cast<const long long*>(&C);
        long long* dst = reinterpret_cast<long long*>(&Cn);
        static_assert(sizeof(BChain) % 8 == 0, "BChain copies in 8-byte words");
        for (int i = lane; i < static_cast<int>(sizeof(BChain) / 8); i += kWave) dst[i] = src[i];
        for (int i = lane; i < 97; i += kWave) rng_u[rsel ^ 1][i] = rng_u[rsel][i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        {
          // every lane holds the same duni state
          Rng Rs;
          Rs.u = rng_u[rsel ^ 1];
          Rs.c = cur_c;
          Rs.cd = C.cd;
          Rs.cm = C.cm;
          Rs.ir = cur_ir;
          Rs.jr = cur_jr;
          // this move booked as rejected: its acceptance draw (every
          // rejected move with a parallel phase takes one), the attempt,
          // the step and the production sample
          (void)Rs.next();
          if (lane == 0) {
            Cn.attempted[move_slot(D)]++;
            Cn.steps++;
            (void)production_sample(Cn);
          }
          propose(Z, Cn, Rs, Dbuf[dsel ^ 1], lane);
          if (lane == 0) {
            spec_c = Rs.c;
            spec_ir = Rs.ir;
            spec_jr = Rs.jr;
          }
        }
      }
      // ---- real space: framework candidates of every position's cell and
      //      every guest atom, one flattened loop (ewald2 + srfrce), on all
      //      but the last wave.  It reads only what the set-up phases
      //      wrote, so no barrier separates it from the reciprocal part: a
      //      wave that finishes its k-vectors starts its pairs at once ----
      // the decision's inputs, by the first lane of the wave before it (it
      // finishes its real-space share ahead of the preparing wave): the
      // acceptance draw (read without taking it), the tail change, the
      // rejection bound
      if (tid == RT - kWave) {
        const float* u = rng_u[rsel];
        float uni = u[cur_ir - 1] - u[cur_jr - 1];
        if (uni < 0.0f) uni = uni + 1.0f;
        float c = cur_c - C.cd;
        if (c < 0.0f) c = c + C.cm;
        uni = uni - c;
        if (uni < 0.0f) uni = uni + 1.0f;
        const double rande = uni;
        pre_rande = rande;
        pre_delrc = (kind == FMC_INSERT || kind == FMC_DELETE) ? tail_change(Z, C, D) : 0.0;
        double thr = 0.0;
        pre_fast = reject_bound(C, D, rande, &thr) ? 1 : 0;
        pre_thr = thr;
      }
      if (tid < RT) {
        const int gcap = Y.gcap;
        const int nfwp = np * gcap;
        const int ng = D.nguest;
        const float rgcap = 1.0f / gcap, rng = 1.0f / max(ng, 1);
        const int total = nfwp + np * ng;
        const int own0 = (kind != FMC_INSERT) ? D.gpre[ig] + D.imol * ns : -1;
        double a2 = 0, av = 0;
        // kU items per thread per pass: their index loads, then their
        // coordinate loads, issue together (one memory latency per pass)
        constexpr int kU = 4;
        for (int base = tid; base < total; base += kU * RT) {
          int pp[kU], aa[kU];
#pragma unroll
          for (int u = 0; u < kU; ++u) {
            const int idx = base + u * RT;
            pp[u] = -1;
            aa[u] = 0;
            if (idx < nfwp) {
              const int p = div_rc(idx, gcap, rgcap);
              pp[u] = p;
              aa[u] = D.plo[p] + (idx - p * gcap);   // record slot
            } else if (idx < total) {
              const int r = idx - nfwp;
              const int p = div_rc(r, ng, rng);
              const int a = r - p * ng;
              if (!(own0 >= 0 && a >= own0 && a < own0 + ns)) {
                pp[u] = p;
                aa[u] = -1 - a;   // guest atom a
              }
            }
          }
          double ax[kU], ay[kU], az[kU], aq[kU];
          int at[kU];
#pragma unroll
          for (int u = 0; u < kU; ++u) {
            ax[u] = ay[u] = az[u] = aq[u] = 0.0;
            at[u] = 0;
            if (pp[u] < 0) continue;
            if (aa[u] >= 0) {
              const double* rr = &Y.grec[aa[u]].x;
              ax[u] = gload(rr);
              ay[u] = gload(rr + 1);
              az[u] = gload(rr + 2);
              aq[u] = gload(rr + 3);
              at[u] = gload(Y.grec_t + aa[u]);
            } else {
              const int a = -1 - aa[u];
              int g = 0;
              while (g + 1 < Z.ngt && D.gpre[g + 1] <= a) ++g;
              const int la = a - D.gpre[g];
              const int sa =
                  la - div_rc(la, Z.ns[g], __builtin_amdgcn_rcpf(static_cast<float>(Z.ns[g]))) *
                           Z.ns[g];
              ax[u] = gload(C.gx[g] + la);
              ay[u] = gload(C.gy[g] + la);
              az[u] = gload(C.gz[g] + la);
              aq[u] = Z.gq[g][sa];
              at[u] = Z.gtype[g][sa];
            }
          }
#pragma unroll
          for (int u = 0; u < kU; ++u) {
            if (pp[u] < 0) continue;
            const int p = pp[u];
            // FP64 in both precisions: the nearest pairs' terms reach tens
            // of kcal/mol and cancel against each other and against the
            // reciprocal part, so an FP32 rounding of each would dominate
            // the error of the move's energy change, and on this path FP64
            // pair arithmetic costs about what FP32 does
            double b2 = 0, bv = 0;
            const int vk =
                ldsrow ? D.ptab[p][at[u]] : gload(Y.S.pairtab + D.pts[p] * ntp + at[u]);
            pair_fp64_vk(Y.S, D.pce[p], vk, D.ppx[p] - ax[u], D.ppy[p] - ay[u],
                         D.ppz[p] - az[u], aq[u], b2, bv);
            if (D.psg[p] < 0) {
              a2 -= b2;
              av -= bv;
            } else {
              a2 += b2;
              av += bv;
            }
          }
        }
        e2 = a2;
        ev = av;
      }
    }
    // block reductions (fixed tree): DPP within each wave, then the waves'
    // totals in order
    e1 = wave_sum_dpp(e1);
    e2 = wave_sum_dpp(e2);
    ev = wave_sum_dpp(ev);
    if ((tid % kWave) == kWave - 1) {
      red[0][tid / kWave] = e1;
      red[1][tid / kWave] = e2;
      red[2][tid / kWave] = ev;
    }
    __syncthreads();
    // while thread 0 decides, waves 1.. build the tables of the prepared
    // next move (this move's are read no more); the next move uses them if
    // it adopts the prepared proposal, and rebuilds them otherwise
    if (kind != 0 && tid >= kWave) {
      MoveDesc& Dn = Dbuf[dcur ^ 1];
      if (Dn.kind != 0)
        build_tables<T, MIXED>(Z, Y, Dn, tid / kWave - 1, BT / kWave - 1, tid % kWave, tab,
                               tabp, xy_l, zt, nlm);
    }
    if (tid == 0) {
      double a = 0.0, b = 0.0, c = 0.0;
      for (int w = 0; w < BT / kWave; ++w) {
        a += red[0][w];
        b += red[1][w];
        c += red[2][w];
      }
      a = Z.rvolm2 * kR4pie0 * a / Z.epsq;
      if (budget_ticks > 0 &&
          static_cast<long long>(__builtin_amdgcn_s_memrealtime()) - t_start >= budget_ticks)
        stop_flag = 1;
      // A rejected move, when another move follows in this launch (a
      // launch never ends holding draws of a move it has not made), takes
      // the other buffers: the move booked as rejected and the next move
      // proposed.  Otherwise the serial lane books the move itself.  (A
      // no-op move has no parallel phase, hence nothing prepared.)
      spec_hit = 0;
      if (kind != 0 && !stop_flag && C.steps + 1 < C.stop_at && mv + 1 < nmoves) {
        double estep, e1x;
        bool rej;
        if (pre_fast) {
          // estep and the logged e1 as decide forms them; the bound first
          e1x = a;
          if (kind == FMC_INSERT) {
            e1x = e1x + Z.engsic[ig];
            estep = (e1x + b + Z.e3[ig] + c + pre_delrc) / C.engunit;
          } else if (kind == FMC_DELETE) {
            e1x = e1x - Z.engsic[ig];
            estep = (e1x - b - Z.e3[ig] - c + pre_delrc) / C.engunit;
          } else {
            estep = (e1x + b + c) / C.engunit;
          }
          rej = estep > pre_thr ||
                !decision(Z, C, D, a, b, c, pre_delrc, pre_rande, estep, e1x);
        } else {
          rej = !decision(Z, C, D, a, b, c, pre_delrc, pre_rande, estep, e1x);
        }
        if (rej && !(kind == FMC_DISPLACE && estep < 0.0)) {
          spec_hit = 1;
          D.accepted = 0;
          csel ^= 1;
          rsel ^= 1;
          dsel ^= 1;
          R.u = rng_u[rsel];
          R.c = spec_c;
          R.ir = spec_ir;
          R.jr = spec_jr;
          BChain& Cn = Cbuf[csel];
          if (Cn.win_len > 0 && Cn.steps > Cn.eq_steps) window_snapshot(Cn, Cn.sre + nk);
          log_move(Cn, move_slot(D), ig, 0, estep, e1x, b, c);
        }
      }
      if (!spec_hit) decide(Z, C, R, D, a, b, c, nk);
      shift_lo = 0;
      shift_hi = 0;
      if (D.accepted && kind == FMC_DELETE) {
        shift_lo = D.imol * ns;
        shift_hi = C.nmol[ig] * ns;   // already decremented
      }
    }
    __syncthreads();
    {
    // the chain state after this move (the other buffer if it was rejected)
    BChain& C = Cbuf[csel];
    if (D.accepted) {
      // S(k) <- S_new (ewald.f:260-293, gcmc.f:971-974)
      if (STREAM) {
        for (int k = tid; k < nk; k += BT) {
          gstore(C.sre + k, gload(C.sre + k) + gload(C.dsre + k));
          gstore(C.sim + k, gload(C.sim + k) + gload(C.dsim + k));
        }
      } else {
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
          const int k = tid + j * BT;
          if (j < kpt && k < nk) {
            sre[j] = sre[j] + static_cast<double>(dre[j]);
            sim[j] = sim[j] + static_cast<double>(dim[j]);
          }
        }
      }
      if (kind == FMC_INSERT) {
        if (tid < ns) {
          C.gx[ig][D.slot + tid] = D.nx[tid];
          C.gy[ig][D.slot + tid] = D.ny[tid];
          C.gz[ig][D.slot + tid] = D.nz[tid];
        }
      } else if (kind == FMC_DISPLACE) {
        if (tid < ns) {
          C.gx[ig][D.imol * ns + tid] = D.nx[tid];
          C.gy[ig][D.imol * ns + tid] = D.ny[tid];
          C.gz[ig][D.imol * ns + tid] = D.nz[tid];
        }
      } else if (kind == FMC_DELETE) {
        // successors shift down by one molecule (gcmc.f:1127-1131)
        for (int base = shift_lo; base < shift_hi; base += BT) {
          const int i = base + tid;
          double x = 0, y = 0, z = 0;
          if (i < shift_hi) {
            x = C.gx[ig][i + ns];
            y = C.gy[ig][i + ns];
            z = C.gz[ig][i + ns];
          }
          __syncthreads();
          if (i < shift_hi) {
            C.gx[ig][i] = x;
            C.gy[ig][i] = y;
            C.gz[ig][i] = z;
          }
          __syncthreads();
        }
      }
      __syncthreads();
    }
    // FP64 recompute every recomp_every moves of this chain, inside the
    // launch: S(k) <- exact, full energy, drift (as recompute_kernel; the
    // same exact_sk, so the same S bits).  Uniform branch (LDS values
    // written before the last barrier).
    if (C.recomp_every > 0 && C.steps % C.recomp_every == 0) {
      double erec = 0.0, dev = 0.0;
      if (STREAM) {
        for (int k = tid; k < nk; k += BT) {
          double sx, sy;
          exact_sk(Y, Z, C, k, sx, sy);
          const double fr = Y.sfw_re[k], fi = Y.sfw_im[k];
          dev = fmax(dev, fmax(fabs(sx - gload(C.sre + k)), fabs(sy - gload(C.sim + k))));
          gstore(C.sre + k, sx);
          gstore(C.sim + k, sy);
          erec += Y.ak[k] * ((sx - fr) * (sx + fr) + (sy - fi) * (sy + fi));
        }
      } else {
        for (int j = 0; j < kpt; ++j) {   // not unrolled: one copy of exact_sk
          const int k = tid + j * BT;
          if (k >= nk) break;
          double sx, sy;
          exact_sk(Y, Z, C, k, sx, sy);
          const double fr = Y.sfw_re[k], fi = Y.sfw_im[k];
#pragma unroll
          for (int jj = 0; jj < KPT; ++jj)
            if (jj == j) {
              dev = fmax(dev, fmax(fabs(sx - sre[jj]), fabs(sy - sim[jj])));
              sre[jj] = sx;
              sim[jj] = sy;
            }
          erec += ak_l[k] * ((sx - fr) * (sx + fr) + (sy - fi) * (sy + fi));
        }
      }
      double e2r = 0.0, evr = 0.0;
      guest_pairs(Y, Z, C, tid, BT, e2r, evr);
      erec = wave_sum_dpp(erec);
      e2r = wave_sum_dpp(e2r);
      evr = wave_sum_dpp(evr);
      for (int off = kWave / 2; off > 0; off >>= 1) dev = fmax(dev, __shfl_xor(dev, off, kWave));
      if ((tid % kWave) == kWave - 1) {
        red[0][tid / kWave] = erec;
        red[1][tid / kWave] = e2r;
        red[2][tid / kWave] = evr;
        red[3][tid / kWave] = dev;
      }
      __syncthreads();
      if (tid == 0) {
        double a = 0.0, b = 0.0, cc = 0.0, d = 0.0;
        for (int w = 0; w < BT / kWave; ++w) {
          a += red[0][w];
          b += red[1][w];
          cc += red[2][w];
          d = fmax(d, red[3][w]);
        }
        recompute_finish(Z, C, a, b, cc, d, 0);
      }
      __syncthreads();
    }
    }
  }
  // write the state back
  BChain& Cf = Cbuf[csel];
  if (!STREAM) {
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const int k = tid + j * BT;
      if (j < kpt && k < nk) {
        Cf.sre[k] = sre[j];
        Cf.sim[k] = sim[j];
      }
    }
  }
  if (tid == 0) {
    for (int i = 0; i < 97; ++i) Cf.u[i] = rng_u[rsel][i];
    Cf.c = R.c; Cf.ir = R.ir; Cf.jr = R.jr;
  }
  __syncthreads();
  {
    const int* src = reinterpret_cast<const int*>(&Cf);
    int* dst = reinterpret_cast<int*>(&Cg);
    for (int i = tid; i < static_cast<int>(sizeof(BChain) / 4); i += BT) dst[i] = src[i];
  }
}

// ---------------------------------------------------------------------------
// Candidate grid (built once per framework)
// ---------------------------------------------------------------------------
struct GridParams {
  DevSys S;
  int gdim[3];
  double reach2;   // (rcut + half the cell diagonal)^2
  int* count;
  Rec* rec;
  int* rtype;
  int cap;
};

__device__ __forceinline__ void grid_center(const GridParams& P, int c,
                                            double& x, double& y, double& z) {
  const int i2 = c % P.gdim[2];
  const int i1 = (c / P.gdim[2]) % P.gdim[1];
  const int i0 = c / (P.gdim[2] * P.gdim[1]);
  const double f0 = (i0 + 0.5) / P.gdim[0], f1 = (i1 + 0.5) / P.gdim[1],
               f2 = (i2 + 0.5) / P.gdim[2];
  x = P.S.cell[0] * f0 + P.S.cell[3] * f1 + P.S.cell[6] * f2;
  y = P.S.cell[1] * f0 + P.S.cell[4] * f1 + P.S.cell[7] * f2;
  z = P.S.cell[2] * f0 + P.S.cell[5] * f1 + P.S.cell[8] * f2;
}

// Is some periodic image of the separation within reach?  For imcon 1/2
// the per-axis nearest image is the shortest; for a parallelepiped the
// fractional-nint image of images (utility_pack.f:1221-1238) need not be, so
// its 26 neighbours are tried too (enough while reach < every
// perpendicular width: the shortest image then has fractional components
// within one of the nint image's).
__device__ __forceinline__ bool within_reach(const GridParams& P, double dx, double dy,
                                             double dz) {
  min_image(P.S, dx, dy, dz);
  if (dx * dx + dy * dy + dz * dz <= P.reach2) return true;
  if (P.S.imcon != 3) return false;
  const double* c = P.S.cell;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        const double x = dx + i * c[0] + j * c[3] + k * c[6];
        const double y = dy + i * c[1] + j * c[4] + k * c[7];
        const double z = dz + i * c[2] + j * c[5] + k * c[8];
        if (x * x + y * y + z * z <= P.reach2) return true;
      }
  return false;
}

__global__ void __launch_bounds__(256) grid_count_kernel(GridParams P) {
  __shared__ int cnt;
  const int c = blockIdx.x;
  double cx, cy, cz;
  grid_center(P, c, cx, cy, cz);
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int a = threadIdx.x; a < P.S.nfw; a += 256)
    if (within_reach(P, P.S.fx[a] - cx, P.S.fy[a] - cy, P.S.fz[a] - cz)) ++mine;
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) P.count[c] = cnt;
}

// ordered fill: candidates of a cell in increasing atom index, then NaN
// records up to the capacity
__global__ void __launch_bounds__(256) grid_fill_kernel(GridParams P) {
  __shared__ int scan[256];
  __shared__ int base;
  const int c = blockIdx.x;
  double cx, cy, cz;
  grid_center(P, c, cx, cy, cz);
  Rec* rec = P.rec + static_cast<size_t>(c) * P.cap;
  int* rt = P.rtype + static_cast<size_t>(c) * P.cap;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  for (int a0 = 0; a0 < P.S.nfw; a0 += 256) {
    const int a = a0 + threadIdx.x;
    int f = 0;
    if (a < P.S.nfw)
      f = within_reach(P, P.S.fx[a] - cx, P.S.fy[a] - cy, P.S.fz[a] - cz) ? 1 : 0;
    scan[threadIdx.x] = f;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const int v = threadIdx.x >= o ? scan[threadIdx.x - o] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    const int slot = base + scan[threadIdx.x] - 1;
    if (f && slot < P.cap) {
      rec[slot] = {P.S.fx[a], P.S.fy[a], P.S.fz[a], P.S.fq[a]};
      rt[slot] = P.S.ftype[a];
    }
    __syncthreads();
    if (threadIdx.x == 255) base += scan[255];
    __syncthreads();
  }
  const double nan = __builtin_nan("");
  for (int i = base + threadIdx.x; i < P.cap; i += 256) {
    rec[i] = {nan, nan, nan, 0.0};
    rt[i] = 0;
  }
}

// Periodic FP64 full recompute of a chain (no reference counterpart; the
// reference only accumulates estep, gcmc.f:981,1151,1368).  It replaces the
// running S(k) by S_fw(k) + the guests' exact sum (ewald.f:368-630 for the
// guest atoms) and evaluates the chain's whole guest energy as
// single_point does (gcmc.f:2518-2710): reciprocal (all - framework) incl.
// the guests' self terms, real space over every guest-framework pair and
// every guest-guest pair of different molecules (each once), the ewald3
// exclusions and the tail.  FP64 throughout, whatever the batch precision.
// base != 0 (admission): the chain's accumulators start from this energy,
// every guest type's energy(i) = spenergy as gcmc.f:555-560 sets it.
struct RecomputeParams {
  const BSys* sys;
  BChain* chains;
  const int* ids;
  double* maxdev;   // per work-group: largest |S_running - S_exact|
  int base;
  int one_id;       // >= 0: a single chain (admission) ...
  const BSys* one;  // ... whose system is this device copy
};

constexpr int kRecBT = 256;

__global__ void __launch_bounds__(kRecBT) recompute_kernel(RecomputeParams P) {
  __shared__ double red[4][kRecBT / kWave];
  const int tid = threadIdx.x;
  BChain& C = P.chains[P.one_id >= 0 ? P.one_id : P.ids[blockIdx.x]];
  const BSys& Y = P.one ? *P.one : P.sys[C.sys];
  double erec = 0.0, dev = 0.0;
  for (int k = tid; k < Y.nk; k += kRecBT) {
    double sx, sy;
    exact_sk(Y, Y.q, C, k, sx, sy);
    const double fr = Y.sfw_re[k], fi = Y.sfw_im[k];
    dev = fmax(dev, fmax(fabs(sx - C.sre[k]), fabs(sy - C.sim[k])));
    C.sre[k] = sx;
    C.sim[k] = sy;
    // |S|^2 - |S_fw|^2 = (S - S_fw).(S + S_fw)
    erec += Y.ak[k] * ((sx - fr) * (sx + fr) + (sy - fi) * (sy + fi));
  }
  double e2 = 0.0, ev = 0.0;
  guest_pairs(Y, Y.q, C, tid, kRecBT, e2, ev);
  erec = wave_sum(erec);
  e2 = wave_sum(e2);
  ev = wave_sum(ev);
  for (int off = kWave / 2; off > 0; off >>= 1) dev = fmax(dev, __shfl_xor(dev, off, kWave));
  if (tid % kWave == 0) {
    red[0][tid / kWave] = erec;
    red[1][tid / kWave] = e2;
    red[2][tid / kWave] = ev;
    red[3][tid / kWave] = dev;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0, cc = 0.0, d = 0.0;
    for (int w = 0; w < kRecBT / kWave; ++w) {
      a += red[0][w];
      b += red[1][w];
      cc += red[2][w];
      d = fmax(d, red[3][w]);
    }
    recompute_finish(Y.q, C, a, b, cc, d, P.base);
    P.maxdev[blockIdx.x] = d;
  }
}

// S(k) of n atoms (FP64, direct phases): the framework part used by resync
__global__ void __launch_bounds__(256) sfac_kernel(int nk, const int* kv,
                                                   const double* rcell, int n,
                                                   const double* x, const double* y,
                                                   const double* z, const double* q,
                                                   double* re, double* im) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nk) return;
  const int v = kv[k];
  const int l = ul(v), m = um(v), nn = un(v);
  double sx = 0.0, sy = 0.0;
  for (int a = 0; a < n; ++a) {
    double ph = l * (rcell[0] * x[a] + rcell[3] * y[a] + rcell[6] * z[a]) +
                m * (rcell[1] * x[a] + rcell[4] * y[a] + rcell[7] * z[a]) +
                nn * (rcell[2] * x[a] + rcell[5] * y[a] + rcell[8] * z[a]);
    ph -= rint(ph);
    double sv, cv;
    sincospi(2.0 * ph, &sv, &cv);
    sx += q[a] * cv;
    sy += q[a] * sv;
  }
  re[k] = sx;
  im[k] = sy;
}

}  // namespace fmc

// ===========================================================================
// Host side
// ===========================================================================
using namespace fmc;

#define BCHECK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      std::fprintf(stderr, "fastmc_amd batch: HIP error %s at %s:%d\n",  \
                   hipGetErrorString(e_), __FILE__, __LINE__);           \
      return FMC_ERR_HIP;                                                \
    }                                                                    \
  } while (0)

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <typename T>
int dput(std::vector<std::unique_ptr<DevBuf>>& keep, T** out, const T* h,
         size_t n) {
  auto b = std::make_unique<DevBuf>();
  const size_t bytes = std::max<size_t>(1, n) * sizeof(T);
  BCHECK(hipMalloc(&b->p, bytes));
  b->n = bytes;
  if (h && n) BCHECK(hipMemcpy(b->p, h, n * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<T*>(b->p);
  keep.push_back(std::move(b));
  return FMC_OK;
}

uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

uint64_t system_key(const SystemDesc& s) {
  uint64_t h = 1469598103934665603ull;
  h = fnv(h, &s.imcon, sizeof(int));
  h = fnv(h, s.cell, sizeof(s.cell));
  h = fnv(h, &s.alpha, sizeof(double));
  h = fnv(h, s.kmax, sizeof(s.kmax));
  h = fnv(h, &s.rcut, sizeof(double));
  h = fnv(h, &s.rvdw, sizeof(double));
  h = fnv(h, &s.epsq, sizeof(double));
  h = fnv(h, s.fx.data(), s.fx.size() * sizeof(double));
  h = fnv(h, s.fy.data(), s.fy.size() * sizeof(double));
  h = fnv(h, s.fz.data(), s.fz.size() * sizeof(double));
  h = fnv(h, s.fq.data(), s.fq.size() * sizeof(double));
  h = fnv(h, s.ftype.data(), s.ftype.size() * sizeof(int));
  h = fnv(h, s.vdw_prm.data(), s.vdw_prm.size() * sizeof(double));
  h = fnv(h, s.vdw_key.data(), s.vdw_key.size() * sizeof(int));
  h = fnv(h, s.vdw_a.data(), s.vdw_a.size() * sizeof(int));
  h = fnv(h, s.vdw_b.data(), s.vdw_b.size() * sizeof(int));
  // template fields one by one: the struct's padding and unused sites are
  // not part of the system
  for (const auto& g : s.guests) {
    const size_t n = static_cast<size_t>(g.ns);
    h = fnv(h, &g.ns, sizeof(int));
    for (const double* a : {g.x, g.y, g.z, g.q, g.mass}) h = fnv(h, a, n * sizeof(double));
    h = fnv(h, g.type, n * sizeof(int));
  }
  return h;
}

struct SysHost {
  uint64_t key;
  BSys dev;
  size_t bytes = 0;
  int refs = 0;     // chains using it; freed at zero
  int nlm = 0, npos = 0;
  BSys* d_self = nullptr;   // device copy of dev (admission recompute)
  std::vector<std::unique_ptr<DevBuf>> keep;
};

struct ChainHost {
  bool active = false;
  int sys = -1;
  std::vector<std::unique_ptr<DevBuf>> keep;
  size_t bytes = 0;
};

void duni_setup(int i, int j, int k, int l, BChain* C) {
  // gcmc.f:2760-2784, then one discarded draw (gcmc.f:335-336)
  C->ir = 97;
  C->jr = 33;
  for (int ii = 0; ii < 97; ++ii) {
    float s = 0.0f, t = 0.5f;
    for (int jj = 0; jj < 24; ++jj) {
      const int m = ((i * j) % 179) * k % 179;
      i = j;
      j = k;
      k = m;
      l = (53 * l + 1) % 169;
      if ((l * m) % 64 >= 32) s = s + t;
      t = 0.5f * t;
    }
    C->u[ii] = s;
  }
  C->c = 362436.0f / 16777216.0f;
  C->cd = 7654321.0f / 16777216.0f;
  C->cm = 16777213.0f / 16777216.0f;
  float uni = C->u[C->ir - 1] - C->u[C->jr - 1];
  if (uni < 0.0f) uni += 1.0f;
  C->u[C->ir - 1] = uni;
  C->ir = C->ir - 1 == 0 ? 97 : C->ir - 1;
  C->jr = C->jr - 1 == 0 ? 97 : C->jr - 1;
  C->c = C->c - C->cd;
  if (C->c < 0.0f) C->c += C->cm;
}

}  // namespace

struct fmc_batch {
  int max_chains = 0;
  int precision = FMC_PREC_FP64;
  int bt = 0, kpt = 0;   // block size; register tile (0: streaming form)
  bool tuned = false;    // false: autotune at the next launch
  bool force_stream = false;
  std::vector<std::unique_ptr<SysHost>> systems;
  std::vector<std::unique_ptr<ChainHost>> chains;
  BSys* d_sys = nullptr;
  int d_sys_cap = 0;
  BChain* d_chains = nullptr;
  int* d_ids = nullptr;
  double* d_dev = nullptr;
  double* d_dev1 = nullptr;   // admission recompute's output
  std::vector<BChain> h_chains;
  bool dirty_sys = true;
  hipStream_t stream = nullptr;
  hipStream_t setup = nullptr;   // table/grid building while a launch runs
  // One thread drives launches; another may admit chains meanwhile.  mu
  // guards the host-side tables; admit_mu serialises admissions (the set-up
  // stream); device waits happen outside both.
  std::mutex mu;
  std::mutex admit_mu;
  std::vector<char> reserved;   // slots being admitted
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // graph replays: one event pair around each replay of the move graph
  std::vector<std::pair<hipEvent_t, hipEvent_t>> cyc_ev;
  long long cyc_pending = 0;
  int timing_mode = 0;   // 1: ev0/ev1 bracket one kernel; 2: cyc_ev
  // captured move cycle: the batch kernel (graph) and the FP64 recompute
  // (graph_rec), and what they were captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph_rec = nullptr;
  hipGraphExec_t graph_rec_exec = nullptr;
  std::vector<int> graph_ids;
  long long graph_moves = -1;
  int graph_bt = 0, graph_kpt = -1;
  bool timing_pending = false;
  double kernel_ms = 0.0;
  long long launches = 0;
  int max_nk = 0;
  int max_nlm = 0;   // (kmax1+1)(2 kmax2+1) over the systems
  int max_pos = 0;   // trial positions: 2 x the largest guest
  std::vector<int> active_ids;
};

namespace {

using KernelFn = void (*)(const BSys*, BChain*, const int*, long long, int, long long);

struct KCfg {
  int bt, kpt;   // kpt 0: the streaming form
  bool mixed;
  KernelFn fn;
};

#define KC(BT_, KPT_)                                                      \
  {BT_, KPT_, false, batch_kernel<BT_, KPT_, false, false>},                \
      {BT_, KPT_, true, batch_kernel<BT_, KPT_, true, false>}
#define KS(BT_)                                                            \
  {BT_, 0, false, batch_kernel<BT_, 1, false, true>},                       \
      {BT_, 0, true, batch_kernel<BT_, 1, true, true>}

const KCfg kConfigs[] = {
    KC(256, 8), KC(256, 16), KC(256, 28), KC(512, 4), KC(512, 8), KC(512, 12),
    KC(512, 14),
    KS(256),    KS(512),
};

const KCfg* find_cfg(int bt, int kpt, bool mixed) {
  for (const auto& c : kConfigs)
    if (c.bt == bt && c.kpt == kpt && c.mixed == mixed) return &c;
  return nullptr;
}

// smallest template KPT for block size bt that covers nk
int kpt_for(int bt, int nk) {
  int best = 0;
  for (const auto& c : kConfigs)
    if (c.bt == bt && !c.mixed && c.kpt > 0 && c.bt * c.kpt >= nk &&
        (best == 0 || c.kpt < best))
      best = c.kpt;
  return best;
}

// Device data of a new system (tables, k-vectors, framework S(k), candidate
// records) built on the set-up stream; touches no shared batch state.
int build_system(fmc_batch* b, const fmc_ctx* ctx, uint64_t key,
                 std::unique_ptr<SysHost>* out) {
  const SystemDesc& s = ctx->sys;
  auto sh = std::make_unique<SysHost>();
  sh->key = key;
  BSys& Y = sh->dev;
  std::memset(&Y, 0, sizeof(Y));
  auto& K = sh->keep;
  int r;
  const int nfw = static_cast<int>(s.fx.size());
  DevSys& S = Y.S;
  S.imcon = s.imcon;
  std::memcpy(S.cell, s.cell, sizeof(S.cell));
  std::memcpy(S.rcell, s.rcell, sizeof(S.rcell));
  std::memcpy(Y.q.cell, s.cell, sizeof(Y.q.cell));
  std::memcpy(Y.q.rcell, s.rcell, sizeof(Y.q.rcell));
  Y.q.epsq = s.epsq;
  S.inv[0] = 1.0 / s.cell[0];
  S.inv[1] = 1.0 / s.cell[4];
  S.inv[2] = 1.0 / s.cell[8];
  S.rcut2 = s.rcut * s.rcut;
  S.rdrewd = 1.0 / s.drewd;
  S.rdlrpot = 1.0 / s.dlrpot;
  S.mxgrid = s.mxgrid;
  S.ntpatm = s.ntpatm;
  S.epsq = s.epsq;
  S.nfw = nfw;
  std::vector<int> ft0(nfw);
  for (int i = 0; i < nfw; ++i) ft0[i] = s.ftype[i] - 1;
  double *fx, *fy, *fz, *fq, *erc, *vvv, *ak;
  int *ftype, *pairtab, *kv;
  if ((r = dput(K, &fx, s.fx.data(), nfw)) || (r = dput(K, &fy, s.fy.data(), nfw)) ||
      (r = dput(K, &fz, s.fz.data(), nfw)) || (r = dput(K, &fq, s.fq.data(), nfw)) ||
      (r = dput(K, &ftype, ft0.data(), nfw)))
    return r;
  if ((r = dput(K, &erc, s.erc.data(), s.erc.size())) ||
      (r = dput(K, &vvv, s.vvv.data(), s.vvv.size())) ||
      (r = dput(K, &pairtab, s.pairtab.data(), s.pairtab.size())))
    return r;
  S.fx = fx; S.fy = fy; S.fz = fz; S.fq = fq; S.ftype = ftype;
  S.erc = erc; S.vvv = vvv; S.pairtab = pairtab;
  const int nk = static_cast<int>(s.kvecs.size());
  std::vector<int> kvh(nk);
  std::vector<double> akh(nk);
  for (int k = 0; k < nk; ++k) {
    kvh[k] = (s.kvecs[k].l & 0xff) | ((s.kvecs[k].m & 0xff) << 8) |
             ((s.kvecs[k].n & 0xff) << 16);
    akh[k] = s.kvecs[k].ak;
  }
  if ((r = dput(K, &kv, kvh.data(), nk)) || (r = dput(K, &ak, akh.data(), nk))) return r;
  Y.nk = nk;
  Y.kv = kv;
  Y.ak = ak;
  for (int d = 0; d < 3; ++d) Y.q.kmax[d] = s.kmax[d];
  if (s.kmax[0] >= kBTab || s.kmax[1] >= kBTab || s.kmax[2] >= kBTab)
    return FMC_ERR_UNSUPPORTED;
  Y.q.volm = s.volm;
  Y.q.rvolm2 = 2.0 * (2.0 * kPi / s.volm);
  // framework S(k)
  double *sre, *sim, *rc;
  if ((r = dput(K, &sre, static_cast<double*>(nullptr), nk)) ||
      (r = dput(K, &sim, static_cast<double*>(nullptr), nk)) ||
      (r = dput(K, &rc, s.rcell, 9)))
    return r;
  hipLaunchKernelGGL(sfac_kernel, dim3((nk + 255) / 256), dim3(256), 0, b->setup,
                     nk, kv, rc, nfw, fx, fy, fz, fq, sre, sim);
  BCHECK(hipGetLastError());
  Y.sfw_re = sre;
  Y.sfw_im = sim;
  // candidate grid: cells of >= 1.5 A along each perpendicular width, each
  // with a fixed-capacity list of framework atom records (one read per
  // candidate, no index indirection); the cell edge grows until the lists
  // fit in kGridBytes
  if (nfw >= (1 << 24)) return FMC_ERR_CAPACITY;
  double p[10];
  dcell(s.cell, p);
  const size_t kGridBytes = size_t(1) << 30;
  for (double edge = 1.5;; edge *= 1.25) {
    for (int d = 0; d < 3; ++d) Y.q.gdim[d] = std::max(1, static_cast<int>(p[6 + d] / edge));
    const int ncell = Y.q.gdim[0] * Y.q.gdim[1] * Y.q.gdim[2];
    double hd = 0.0;
    for (int sa = -1; sa <= 1; sa += 2)
      for (int sb = -1; sb <= 1; sb += 2) {
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = (s.cell[d] / Y.q.gdim[0] + sa * s.cell[3 + d] / Y.q.gdim[1] +
                  sb * s.cell[6 + d] / Y.q.gdim[2]) * 0.5;
        hd = std::max(hd, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      }
    GridParams G;
    G.S = S;
    for (int d = 0; d < 3; ++d) G.gdim[d] = Y.q.gdim[d];
    const double reach = s.rcut + hd + 1e-6;
    if (reach >= std::min(p[6], std::min(p[7], p[8]))) return FMC_ERR_CELL;
    G.reach2 = reach * reach;
    int* cnt;   // kept with the system (no free while a launch may run)
    if ((r = dput(K, &cnt, static_cast<int*>(nullptr), ncell))) return r;
    G.count = cnt;
    hipLaunchKernelGGL(grid_count_kernel, dim3(ncell), dim3(256), 0, b->setup, G);
    BCHECK(hipGetLastError());
    std::vector<int> hc(ncell);
    BCHECK(hipMemcpyAsync(hc.data(), cnt, ncell * sizeof(int), hipMemcpyDeviceToHost,
                          b->setup));
    BCHECK(hipStreamSynchronize(b->setup));
    const int cap = std::max(1, *std::max_element(hc.begin(), hc.end()));
    const size_t bytes = static_cast<size_t>(ncell) * cap * (sizeof(Rec) + sizeof(int));
    if (bytes > kGridBytes && ncell > 1) continue;
    Rec* rec;
    int* rt;
    if ((r = dput(K, &rec, static_cast<Rec*>(nullptr), static_cast<size_t>(ncell) * cap)) ||
        (r = dput(K, &rt, static_cast<int*>(nullptr), static_cast<size_t>(ncell) * cap)))
      return r;
    G.rec = rec;
    G.rtype = rt;
    G.cap = cap;
    hipLaunchKernelGGL(grid_fill_kernel, dim3(ncell), dim3(256), 0, b->setup, G);
    BCHECK(hipGetLastError());
    BCHECK(hipStreamSynchronize(b->setup));
    Y.grec = rec;
    Y.grec_t = rt;
    Y.gcap = cap;
    break;
  }
  // guests
  Y.q.ngt = static_cast<int>(s.guests.size());
  for (int g = 0; g < Y.q.ngt; ++g) {
    const GuestTemplate& t = s.guests[g];
    Y.q.ns[g] = t.ns;
    for (int i = 0; i < t.ns; ++i) {
      Y.q.gq[g][i] = t.q[i];
      Y.q.gtype[g][i] = t.type[i] - 1;
      Y.q.gmass[g][i] = t.mass[i];
      Y.q.tx[g][i] = t.x[i];
      Y.q.ty[g][i] = t.y[i];
      Y.q.tz[g][i] = t.z[i];
    }
    Y.q.e3[g] = t.e3;
    Y.q.engsic[g] = t.engsic;
  }
  // tail as a quadratic form in the molecule counts
  {
    const int T = s.ntpatm;
    std::vector<double> F(T, 0.0), Z(T, 0.0);
    std::vector<std::vector<double>> cg(Y.q.ngt, std::vector<double>(T, 0.0));
    for (int i = 0; i < nfw; ++i) {
      F[s.ftype[i] - 1] += 1.0;
      if (s.ffrz[i]) Z[s.ftype[i] - 1] += 1.0;
    }
    for (int g = 0; g < Y.q.ngt; ++g)
      for (int i = 0; i < s.guests[g].ns; ++i) cg[g][s.guests[g].type[i] - 1] += 1.0;
    double t0 = 0.0, tb[kMaxGuestTypes] = {0}, tA[kMaxGuestTypes][kMaxGuestTypes] = {{0}};
    const bool on = !(s.imcon == 0 || s.imcon == 6 || s.ntpvdw == 0);
    int iv = 0;
    for (int i = 0; i < T; ++i)
      for (int j = 0; j <= i; ++j, ++iv) {
        if (!on) continue;
        const double w = 2.0 * kPi * s.steadd[iv] / s.volm;
        t0 += w * (F[i] * F[j] - Z[i] * Z[j]);
        for (int g = 0; g < Y.q.ngt; ++g) {
          tb[g] += w * (F[i] * cg[g][j] + F[j] * cg[g][i]);
          for (int h = 0; h < Y.q.ngt; ++h) tA[g][h] += w * cg[g][i] * cg[h][j];
        }
      }
    Y.q.t0 = t0;
    for (int g = 0; g < Y.q.ngt; ++g) {
      Y.q.tb[g] = tb[g];
      for (int h = 0; h < Y.q.ngt; ++h) Y.q.tA[g][h] = tA[g][h];
    }
  }
  if ((r = dput(K, &sh->d_self, &Y, 1))) return r;
  for (auto& kb : K) sh->bytes += kb->n;
  sh->nlm = (s.kmax[0] + 1) * (2 * s.kmax[1] + 1);
  for (const auto& g : s.guests) sh->npos = std::max(sh->npos, 2 * g.ns);
  *out = std::move(sh);
  return FMC_OK;
}

// (batch lock held) index of a system, -1 if it is not there
int find_system(fmc_batch* b, uint64_t key) {
  for (size_t i = 0; i < b->systems.size(); ++i)
    if (b->systems[i] && b->systems[i]->key == key) return static_cast<int>(i);
  return -1;
}

// (batch lock held) takes a built system into the table
int insert_system(fmc_batch* b, std::unique_ptr<SysHost> sh) {
  b->dirty_sys = true;
  for (size_t i = 0; i < b->systems.size(); ++i)
    if (!b->systems[i]) {
      b->systems[i] = std::move(sh);
      return static_cast<int>(i);
    }
  b->systems.push_back(std::move(sh));
  return static_cast<int>(b->systems.size()) - 1;
}

// (batch lock held) one reference less; a system nobody uses is freed
void release_system(fmc_batch* b, int sys) {
  if (sys >= 0 && b->systems[sys] && --b->systems[sys]->refs <= 0) {
    b->systems[sys].reset();
    b->dirty_sys = true;
  }
}

void invalidate_graph(fmc_batch* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->graph_rec_exec) (void)hipGraphExecDestroy(b->graph_rec_exec);
  if (b->graph_rec) (void)hipGraphDestroy(b->graph_rec);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->graph_rec_exec = nullptr;
  b->graph_rec = nullptr;
  b->graph_moves = -1;
}

// LDS sizing follows the systems in use
void refresh_limits(fmc_batch* b) {
  invalidate_graph(b);
  b->max_nk = b->max_nlm = b->max_pos = 0;
  for (const auto& sh : b->systems) {
    if (!sh || sh->refs <= 0) continue;
    b->max_nk = std::max(b->max_nk, sh->dev.nk);
    b->max_nlm = std::max(b->max_nlm, sh->nlm);
    b->max_pos = std::max(b->max_pos, sh->npos);
  }
}

int sync_tables(fmc_batch* b) {
  if (b->dirty_sys) {
    invalidate_graph(b);
    const int n = static_cast<int>(b->systems.size());
    if (n > b->d_sys_cap) {
      if (b->d_sys) BCHECK(hipFree(b->d_sys));
      b->d_sys_cap = std::max(16, 2 * n);
      BCHECK(hipMalloc(&b->d_sys, b->d_sys_cap * sizeof(BSys)));
    }
    std::vector<BSys> h(n);
    for (int i = 0; i < n; ++i) {
      if (b->systems[i]) h[i] = b->systems[i]->dev;
      else std::memset(&h[i], 0, sizeof(BSys));
    }
    BCHECK(hipMemcpy(b->d_sys, h.data(), n * sizeof(BSys), hipMemcpyHostToDevice));
    b->dirty_sys = false;
  }
  b->active_ids.clear();
  for (int i = 0; i < b->max_chains; ++i)
    if (b->chains[i] && b->chains[i]->active) b->active_ids.push_back(i);
  if (!b->active_ids.empty())
    BCHECK(hipMemcpy(b->d_ids, b->active_ids.data(), b->active_ids.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  return FMC_OK;
}

// dynamic LDS of a form: (a_k, packed k index unless streaming) + X.Y
// tables; FMC_ERR_UNSUPPORTED when it and the kernel's static LDS exceed
// the device's per-work-group limit
int lds_need(fmc_batch* b, const KCfg* cfg, size_t* out) {
  const size_t elt = b->precision == FMC_PREC_MIXED ? 8 : 16;
  const int nkcap = b->max_nk;
  size_t lds = static_cast<size_t>(b->max_pos) * b->max_nlm * elt;
  if (cfg->kpt > 0)
    lds += static_cast<size_t>(nkcap) * sizeof(double) +
           static_cast<size_t>((nkcap + 3) & ~3) * sizeof(int);
  int dev = 0, lmax = 0;
  BCHECK(hipGetDevice(&dev));
  BCHECK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  hipFuncAttributes fa;
  BCHECK(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(cfg->fn)));
  if (lds + fa.sharedSizeBytes > static_cast<size_t>(lmax)) return FMC_ERR_UNSUPPORTED;
  *out = lds;
  return FMC_OK;
}

int launch(fmc_batch* b, long long n, int bt, int kpt, long long budget_ticks = 0) {
  const KCfg* cfg = find_cfg(bt, kpt, b->precision == FMC_PREC_MIXED);
  if (!cfg) return FMC_ERR_UNSUPPORTED;
  if (b->active_ids.empty()) return FMC_OK;
  if (kpt > 0 && bt * kpt < b->max_nk) return FMC_ERR_UNSUPPORTED;
  size_t lds = 0;
  int r;
  if ((r = lds_need(b, cfg, &lds))) return r;
  hipLaunchKernelGGL(cfg->fn, dim3(static_cast<unsigned>(b->active_ids.size())),
                     dim3(cfg->bt), lds, b->stream, b->d_sys, b->d_chains, b->d_ids, n,
                     b->max_nk, budget_ticks);
  BCHECK(hipGetLastError());
  return FMC_OK;
}

// the register form for block size bt if it covers the largest Nk and fits
// in LDS, else the streaming form (256/512 threads); 0 if neither
int pick_kpt(fmc_batch* b, int bt, bool allow_stream, int* kpt) {
  const bool mixed = b->precision == FMC_PREC_MIXED;
  const int k = kpt_for(bt, std::max(1, b->max_nk));
  size_t lds;
  if (k > 0 && lds_need(b, find_cfg(bt, k, mixed), &lds) == FMC_OK) {
    *kpt = k;
    return FMC_OK;
  }
  const KCfg* sc = find_cfg(bt, 0, mixed);
  if (allow_stream && sc && lds_need(b, sc, &lds) == FMC_OK) {
    *kpt = 0;
    return FMC_OK;
  }
  return FMC_ERR_UNSUPPORTED;
}

// snapshot / restore of every active chain's device state (autotuning runs
// real moves and must leave the chains untouched)
struct Snapshot {
  std::vector<BChain> chains;
  std::vector<std::vector<double>> data;
};

int snapshot(fmc_batch* b, Snapshot* s, bool restore) {
  // active chains only: a slot being admitted meanwhile is not touched
  const int nmax = b->max_chains;
  if (!restore) {
    s->chains.resize(nmax);
    for (int id : b->active_ids)
      BCHECK(hipMemcpy(&s->chains[id], b->d_chains + id, sizeof(BChain),
                       hipMemcpyDeviceToHost));
    s->data.clear();
  } else {
    for (int id : b->active_ids)
      BCHECK(hipMemcpy(b->d_chains + id, &s->chains[id], sizeof(BChain),
                       hipMemcpyHostToDevice));
  }
  size_t slot = 0;
  for (int id : b->active_ids) {
    const BChain& C = s->chains[id];
    const BSys& Y = b->systems[C.sys]->dev;
    // S(k) and the window snapshots that follow it, the guest coordinates
    std::vector<std::pair<double*, size_t>> bufs = {{C.sre, (size_t)Y.nk + kWinDoubles},
                                                    {C.sim, (size_t)Y.nk}};
    for (int g = 0; g < Y.q.ngt; ++g) {
      bufs.push_back({C.gx[g], (size_t)C.cap_atoms});
      bufs.push_back({C.gy[g], (size_t)C.cap_atoms});
      bufs.push_back({C.gz[g], (size_t)C.cap_atoms});
    }
    for (auto& pr : bufs) {
      if (!restore) {
        s->data.emplace_back(pr.second);
        BCHECK(hipMemcpy(s->data.back().data(), pr.first, pr.second * sizeof(double),
                         hipMemcpyDeviceToHost));
      } else {
        BCHECK(hipMemcpy(pr.first, s->data[slot].data(), pr.second * sizeof(double),
                         hipMemcpyHostToDevice));
      }
      ++slot;
    }
  }
  return FMC_OK;
}

// Project autotuner: time each block size (with the smallest register tile
// that covers the largest Nk) on real moves, keep the fastest, restore state.
int autotune(fmc_batch* b) {
  std::vector<int> bts = {256, 512};
  Snapshot snap;
  int r;
  if ((r = snapshot(b, &snap, false))) return r;
  double best = 1e300;
  int bbt = 0, bkpt = 0;
  for (int bt : bts) {
    int kpt = 0;
    if (pick_kpt(b, bt, true, &kpt)) continue;
    // warm-up + timed run
    r = launch(b, 4, bt, kpt);
    if (r == FMC_ERR_UNSUPPORTED) continue;
    if (r) return r;
    BCHECK(hipStreamSynchronize(b->stream));
    const auto t0 = std::chrono::steady_clock::now();
    if ((r = launch(b, 24, bt, kpt))) return r;
    BCHECK(hipStreamSynchronize(b->stream));
    const double dt =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (dt < best) {
      best = dt;
      bbt = bt;
      bkpt = kpt;
    }
    if ((r = snapshot(b, &snap, true))) return r;
  }
  if (!bbt) return FMC_ERR_UNSUPPORTED;
  b->bt = bbt;
  b->kpt = bkpt;
  b->tuned = true;
  return FMC_OK;
}

// before a launch: autotune once; a fixed tile follows the systems in use
// (register form while it covers the largest Nk and fits, else streaming)
int ensure_tile(fmc_batch* b) {
  if (!b->tuned) return autotune(b);
  if (b->force_stream) {
    size_t lds;
    const KCfg* sc = find_cfg(b->bt, 0, b->precision == FMC_PREC_MIXED);
    if (!sc) return FMC_ERR_UNSUPPORTED;
    b->kpt = 0;
    return lds_need(b, sc, &lds);
  }
  return pick_kpt(b, b->bt, true, &b->kpt);
}

}  // namespace

extern "C" int fmc_batch_create(fmc_batch** out, int max_chains, int precision) {
  if (!out || max_chains <= 0 || (precision != FMC_PREC_FP64 && precision != FMC_PREC_MIXED))
    return FMC_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FMC_ERR_NODEVICE;
  auto b = std::make_unique<fmc_batch>();
  b->max_chains = max_chains;
  b->precision = precision;
  b->chains.resize(max_chains);
  b->reserved.assign(max_chains, 0);
  b->h_chains.resize(max_chains);
  std::memset(b->h_chains.data(), 0, max_chains * sizeof(BChain));
  BCHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BCHECK(hipEventCreate(&b->ev0));
  BCHECK(hipEventCreate(&b->ev1));
  BCHECK(hipMalloc(&b->d_chains, max_chains * sizeof(BChain)));
  BCHECK(hipMemset(b->d_chains, 0, max_chains * sizeof(BChain)));
  BCHECK(hipMalloc(&b->d_ids, max_chains * sizeof(int)));
  BCHECK(hipMalloc(&b->d_dev, max_chains * sizeof(double)));
  BCHECK(hipMalloc(&b->d_dev1, sizeof(double)));
  BCHECK(hipStreamCreateWithFlags(&b->setup, hipStreamNonBlocking));
  *out = b.release();
  return FMC_OK;
}

extern "C" int fmc_batch_destroy(fmc_batch* b) {
  if (!b) return FMC_ERR_ARG;
  (void)hipStreamSynchronize(b->stream);
  if (b->setup) (void)hipStreamSynchronize(b->setup);
  b->chains.clear();
  b->systems.clear();
  if (b->d_sys) (void)hipFree(b->d_sys);
  if (b->d_chains) (void)hipFree(b->d_chains);
  if (b->d_ids) (void)hipFree(b->d_ids);
  if (b->d_dev) (void)hipFree(b->d_dev);
  if (b->d_dev1) (void)hipFree(b->d_dev1);
  invalidate_graph(b);
  if (b->setup) (void)hipStreamDestroy(b->setup);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  for (auto& e : b->cyc_ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return FMC_OK;
}

extern "C" int fmc_chain_footprint(const fmc_ctx* ctx, int precision, long long* bytes) {
  if (!ctx || !bytes || !ctx->sys.finalized) return FMC_ERR_ARG;
  (void)precision;
  const SystemDesc& s = ctx->sys;
  const long long nk = static_cast<long long>(s.kvecs.size());
  long long b = 4 * nk * 8;                                   // S(k), dS(k)
  b += static_cast<long long>(s.guests.size()) * 3 * s.max_guest_atoms * 8;  // guests
  b += sizeof(BChain);
  // system share: framework, tables, k-vectors, grid (counted per chain as
  // an upper bound when the framework is not shared)
  b += static_cast<long long>(s.fx.size()) * (4 * 8 + 4);
  b += static_cast<long long>(s.vvv.size() + s.erc.size()) * 12;
  b += nk * (4 + 8 + 16);
  // candidate records: ~(V / 1.5^3) cells x (atoms within rcut + 1.3 A),
  // at most 1 GiB
  {
    const double nfw = static_cast<double>(s.fx.size());
    const double reach = s.rcut + 1.3;
    const double cap = 1.2 * nfw / s.volm * 4.18879 * reach * reach * reach + 8.0;
    const double cells = s.volm / 3.375;
    b += static_cast<long long>(std::min(cells * cap * 36.0, 1073741824.0));
  }
  *bytes = b;
  return FMC_OK;
}

namespace {

// Admission, first step (under the batch lock): a free slot, and the
// context's system pinned if the batch already holds it (*sys = -1 if not).
int reserve_slot(fmc_batch* b, uint64_t key, int* slot, int* sys) {
  std::lock_guard<std::mutex> g(b->mu);
  *slot = -1;
  for (int i = 0; i < b->max_chains; ++i)
    if (!b->chains[i] && !b->reserved[i]) {
      *slot = i;
      break;
    }
  if (*slot < 0) return FMC_ERR_CAPACITY;
  b->reserved[*slot] = 1;
  *sys = find_system(b, key);
  if (*sys >= 0) b->systems[*sys]->refs++;
  return FMC_OK;
}

// Admission, second step (no lock held): build the system's tables and
// candidate records if the batch does not hold them yet.
int ensure_system(fmc_batch* b, const fmc_ctx* ctx, uint64_t key, int* sys,
                  BSys** d_self) {
  if (*sys < 0) {
    std::unique_ptr<SysHost> sh;
    int r = build_system(b, ctx, key, &sh);
    if (r) return r;
    std::lock_guard<std::mutex> g(b->mu);
    sh->refs = 1;
    *sys = insert_system(b, std::move(sh));
    refresh_limits(b);
  }
  std::lock_guard<std::mutex> g(b->mu);
  if (d_self) *d_self = b->systems[*sys]->d_self;
  return FMC_OK;
}

// Admission, last step: publish the chain in its slot, or give the slot
// and the system reference back when r reports a failure.
int publish_chain(fmc_batch* b, int slot, int sys, int r, std::unique_ptr<ChainHost> ch,
                  const BChain& C, int* chain_id) {
  std::lock_guard<std::mutex> g(b->mu);
  b->reserved[slot] = 0;
  if (r) {
    ch.reset();
    release_system(b, sys);
    refresh_limits(b);
    return r;
  }
  ch->active = true;
  b->chains[slot] = std::move(ch);
  b->h_chains[slot] = C;
  if (chain_id) *chain_id = slot;
  return FMC_OK;
}

}  // namespace

extern "C" int fmc_batch_add_chain(fmc_batch* b, const fmc_ctx* ctx,
                                   const fmc_chain_params* p, int* chain_id) {
  if (!b || !ctx || !p || !ctx->sys.finalized) return FMC_ERR_ARG;
  if (p->ntpguest != static_cast<int>(ctx->sys.guests.size()) || p->ntpguest < 1 ||
      p->ntpguest > kMaxGuestTypes)
    return FMC_ERR_ARG;
  for (int g = 0; g < p->ntpguest; ++g)
    if (p->freq[g][3] > 0.0 || p->freq[g][8] > 0.0) return FMC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> admit(b->admit_mu);
  const uint64_t key = system_key(ctx->sys);
  int slot = -1, sys = -1;
  int r = reserve_slot(b, key, &slot, &sys);
  if (r) return r;
  BSys* d_self = nullptr;
  std::unique_ptr<ChainHost> ch;
  BChain C;
  do {
    if ((r = ensure_system(b, ctx, key, &sys, &d_self))) break;
    ch = std::make_unique<ChainHost>();
    ch->sys = sys;
    std::memset(&C, 0, sizeof(C));
    C.sys = sys;
    C.ntpguest = p->ntpguest;
    C.cap_atoms = ctx->cap_atoms;
    const int nk = ctx->nk;
    auto& K = ch->keep;
    // S(k) is set by the admission recompute below; the window snapshots
    // follow it in the same buffer
    if ((r = dput(K, &C.sre, static_cast<double*>(nullptr), nk + kWinDoubles)) ||
        (r = dput(K, &C.sim, static_cast<double*>(nullptr), nk)) ||
        (r = dput(K, &C.dsre, static_cast<double*>(nullptr), nk)) ||
        (r = dput(K, &C.dsim, static_cast<double*>(nullptr), nk)))
      break;
    if (hipMemsetAsync(C.sre, 0, nk * sizeof(double), b->setup) != hipSuccess ||
        hipMemsetAsync(C.sim, 0, nk * sizeof(double), b->setup) != hipSuccess) {
      r = FMC_ERR_HIP;
      break;
    }
    for (int g = 0; g < p->ntpguest && !r; ++g) {
      if ((r = dput(K, &C.gx[g], ctx->hx[g].data(), C.cap_atoms)) ||
          (r = dput(K, &C.gy[g], ctx->hy[g].data(), C.cap_atoms)) ||
          (r = dput(K, &C.gz[g], ctx->hz[g].data(), C.cap_atoms)))
        break;
      C.nmol[g] = ctx->nmol[g];
    }
    if (r) break;
    C.temp = p->temp;
    C.beta = 1.0 / (kKboltz * p->temp);
    C.statvolm = ctx->sys.volm * 1e-30;
    C.engunit = p->engunit > 0 ? p->engunit : 418.4;
    C.eq_steps = p->eq_steps;
    for (int g = 0; g < p->ntpguest; ++g) {
      C.press[g] = p->press[g];
      double f[9], norm = 0.0;
      for (int i = 0; i < 9; ++i) {
        f[i] = p->freq[g][i];
        norm += f[i];
      }
      if (norm == 0.0) {
        for (int i = 0; i < 9; ++i) f[i] = 0.0;
        f[0] = f[1] = f[2] = 1.0 / 3.0;
        norm = 1.0;
      }
      C.cum[g][0] = f[0] / norm;
      for (int i = 1; i < 9; ++i) C.cum[g][i] = C.cum[g][i - 1] + (f[i] / norm);
      C.disp_ratio[g] = p->disp_ratio[g] > 0 ? p->disp_ratio[g] : 0.5;
      C.tran_ratio[g] = p->tran_ratio[g] > 0 ? p->tran_ratio[g] : 0.5;
      C.delrdisp[g] = p->delr[g];
      C.tran_delr[g] = p->delr[g];
    }
    C.rota_ratio = p->rota_ratio > 0 ? p->rota_ratio : 0.5;
    C.rotangle = kPi / 3.0;
    C.rota_rotangle = C.rotangle;
    C.elrc = 0.0;
    C.stop_at = p->max_moves > 0 ? p->max_moves : std::numeric_limits<long long>::max();
    {
      const long long prod = p->max_moves - std::max(0LL, p->eq_steps);
      C.win_len = (p->max_moves > 0 && prod >= kWin && prod <= kWinMax)
                      ? static_cast<int>(prod / kWin) : 0;
      C.win_snap = 0;
      if (hipMemsetAsync(C.sre + nk, 0, kWinDoubles * sizeof(double), b->setup) !=
          hipSuccess) {
        r = FMC_ERR_HIP;
        break;
      }
    }
    C.recomp_every = p->recompute_every > 0 ? p->recompute_every : 0;
    C.sel_nwind = p->window > 0 ? p->window : kSelWindow;
    duni_setup(p->seed[0], p->seed[1], p->seed[2], p->seed[3], &C);
    for (auto& kb : K) ch->bytes += kb->n;
    // the slot is outside any launch in flight: its struct and the
    // admission recompute (S(k), the full FP64 energy the accumulators
    // start from) go through the set-up stream
    if (hipMemcpyAsync(b->d_chains + slot, &C, sizeof(BChain), hipMemcpyHostToDevice,
                       b->setup) != hipSuccess) {
      r = FMC_ERR_HIP;
      break;
    }
    RecomputeParams P;
    P.sys = nullptr;
    P.chains = b->d_chains;
    P.ids = nullptr;
    P.maxdev = b->d_dev1;
    P.base = 1;
    P.one_id = slot;
    P.one = d_self;
    hipLaunchKernelGGL(recompute_kernel, dim3(1), dim3(kRecBT), 0, b->setup, P);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->setup) != hipSuccess)
      r = FMC_ERR_HIP;
  } while (false);
  return publish_chain(b, slot, sys, r, std::move(ch), C, chain_id);
}

extern "C" int fmc_batch_remove_chain(fmc_batch* b, int id) {
  if (!b || id < 0 || id >= b->max_chains) return FMC_ERR_ARG;
  BCHECK(hipStreamSynchronize(b->stream));
  std::lock_guard<std::mutex> g(b->mu);
  if (!b->chains[id] || !b->chains[id]->active) return FMC_ERR_ARG;
  const int sys = b->chains[id]->sys;
  b->chains[id].reset();
  release_system(b, sys);
  refresh_limits(b);
  return FMC_OK;
}

namespace {

// Chain checkpoint layout: this header, the BChain (device pointers
// cleared), the running S(k) (re, im), then per guest type the live sites'
// x, y, z.  Valid for the library build that wrote it (struct sizes and
// version are checked).
struct CkHeader {
  char magic[8];
  uint64_t key;          // system_key of the chain's system
  int version;
  int precision;         // the writing batch's (informational)
  int nk;
  int ntpguest;
  int cap_atoms;
  int ns[kMaxGuestTypes];
  int chain_bytes;       // sizeof(BChain)
};
constexpr char kCkMagic[8] = {'F', 'M', 'C', 'C', 'H', 'K', 0, 0};
constexpr int kCkVersion = 2;

// A duni state the generator can reach (gcmc.f:2760-2799): the table and c
// multiples of 2^-24 in [0, 1) and [0, cm), cd and cm its constants, and the
// two indices 64 apart (mod 97) as they start (97, 33) and step together.
// The kernel's lookahead draws rely on the last two.
bool rng_state_ok(const BChain& C) {
  const float q = 16777216.0f;
  auto grid = [q](float v) { return v >= 0.0f && v < 1.0f && v * q == std::floor(v * q); };
  if (C.cd != 7654321.0f / q || C.cm != 16777213.0f / q) return false;
  if (!grid(C.c) || !(C.c < C.cm)) return false;
  if (C.ir < 1 || C.ir > 97 || C.jr < 1 || C.jr > 97 || (C.ir - C.jr + 97) % 97 != 64)
    return false;
  for (int i = 0; i < 97; ++i)
    if (!grid(C.u[i])) return false;
  return true;
}

long long ck_size(const BChain& C, int nk, const int* ns) {
  long long n = sizeof(CkHeader) + sizeof(BChain) + 2LL * nk * sizeof(double) +
                kWinDoubles * sizeof(double);
  for (int g = 0; g < C.ntpguest; ++g)
    n += 3LL * C.nmol[g] * ns[g] * sizeof(double);
  return n;
}

void clear_pointers(BChain* C) {
  C->sre = C->sim = C->dsre = C->dsim = nullptr;
  for (int g = 0; g < kMaxGuestTypes; ++g) C->gx[g] = C->gy[g] = C->gz[g] = nullptr;
  C->log = nullptr;
  C->log_cap = C->log_n = 0;
}

}  // namespace

extern "C" int fmc_batch_checkpoint(fmc_batch* b, int id, void* buf, long long cap,
                                    long long* len) {
  if (!b || !len || id < 0 || id >= b->max_chains) return FMC_ERR_ARG;
  BCHECK(hipStreamSynchronize(b->stream));
  std::lock_guard<std::mutex> g(b->mu);
  if (!b->chains[id] || !b->chains[id]->active) return FMC_ERR_ARG;
  BChain C;
  BCHECK(hipMemcpy(&C, b->d_chains + id, sizeof(BChain), hipMemcpyDeviceToHost));
  const SysHost& sh = *b->systems[C.sys];
  const int nk = sh.dev.nk;
  const int* ns = sh.dev.q.ns;
  const long long need = ck_size(C, nk, ns);
  *len = need;
  if (!buf) return FMC_OK;
  if (cap < need) return FMC_ERR_ARG;
  char* out = static_cast<char*>(buf);
  CkHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kCkMagic, sizeof(h.magic));
  h.key = sh.key;
  h.version = kCkVersion;
  h.precision = b->precision;
  h.nk = nk;
  h.ntpguest = C.ntpguest;
  h.cap_atoms = C.cap_atoms;
  for (int t = 0; t < kMaxGuestTypes; ++t) h.ns[t] = t < C.ntpguest ? ns[t] : 0;
  h.chain_bytes = static_cast<int>(sizeof(BChain));
  std::memcpy(out, &h, sizeof(h));
  out += sizeof(h);
  BChain D = C;
  clear_pointers(&D);
  std::memcpy(out, &D, sizeof(D));
  out += sizeof(D);
  BCHECK(hipMemcpy(out, C.sre, nk * sizeof(double), hipMemcpyDeviceToHost));
  out += nk * sizeof(double);
  BCHECK(hipMemcpy(out, C.sim, nk * sizeof(double), hipMemcpyDeviceToHost));
  out += nk * sizeof(double);
  BCHECK(hipMemcpy(out, C.sre + nk, kWinDoubles * sizeof(double), hipMemcpyDeviceToHost));
  out += kWinDoubles * sizeof(double);
  for (int t = 0; t < C.ntpguest; ++t) {
    const size_t n = static_cast<size_t>(C.nmol[t]) * ns[t];
    for (double* src : {C.gx[t], C.gy[t], C.gz[t]}) {
      if (n) BCHECK(hipMemcpy(out, src, n * sizeof(double), hipMemcpyDeviceToHost));
      out += n * sizeof(double);
    }
  }
  return FMC_OK;
}

extern "C" int fmc_batch_restore_chain(fmc_batch* b, const fmc_ctx* ctx, const void* buf,
                                       long long len, int* chain_id) {
  if (!b || !ctx || !buf || !ctx->sys.finalized ||
      len < static_cast<long long>(sizeof(CkHeader) + sizeof(BChain)))
    return FMC_ERR_ARG;
  const char* in = static_cast<const char*>(buf);
  CkHeader h;
  std::memcpy(&h, in, sizeof(h));
  in += sizeof(h);
  BChain C;
  std::memcpy(&C, in, sizeof(C));
  in += sizeof(C);
  const uint64_t key = system_key(ctx->sys);
  if (std::memcmp(h.magic, kCkMagic, sizeof(h.magic)) != 0 || h.version != kCkVersion ||
      h.chain_bytes != static_cast<int>(sizeof(BChain)) || h.key != key ||
      h.nk != ctx->nk || h.ntpguest != static_cast<int>(ctx->sys.guests.size()) ||
      h.ntpguest != C.ntpguest || h.cap_atoms != C.cap_atoms || C.cap_atoms < 1 ||
      !rng_state_ok(C) || C.sel_nwind < 1 || C.sel_count < 0)
    return FMC_ERR_ARG;
  for (int t = 0; t < h.ntpguest; ++t)
    if (h.ns[t] != ctx->sys.guests[t].ns || C.nmol[t] < 0 ||
        static_cast<long long>(C.nmol[t]) * h.ns[t] > C.cap_atoms)
      return FMC_ERR_ARG;
  if (len != ck_size(C, h.nk, h.ns)) return FMC_ERR_ARG;
  std::lock_guard<std::mutex> admit(b->admit_mu);
  int slot = -1, sys = -1;
  int r = reserve_slot(b, key, &slot, &sys);
  if (r) return r;
  std::unique_ptr<ChainHost> ch;
  do {
    if ((r = ensure_system(b, ctx, key, &sys, nullptr))) break;
    ch = std::make_unique<ChainHost>();
    ch->sys = sys;
    C.sys = sys;
    clear_pointers(&C);
    auto& K = ch->keep;
    const double* sre = reinterpret_cast<const double*>(in);
    const double* sim = sre + h.nk;
    const double* win = sim + h.nk;
    std::vector<double> sre_win(sre, sre + h.nk);
    sre_win.insert(sre_win.end(), win, win + kWinDoubles);
    if ((r = dput(K, &C.sre, sre_win.data(), sre_win.size())) ||
        (r = dput(K, &C.sim, sim, h.nk)) ||
        (r = dput(K, &C.dsre, static_cast<double*>(nullptr), h.nk)) ||
        (r = dput(K, &C.dsim, static_cast<double*>(nullptr), h.nk)))
      break;
    const double* site = win + kWinDoubles;
    std::vector<double> stage(C.cap_atoms);
    for (int t = 0; t < h.ntpguest && !r; ++t) {
      const size_t n = static_cast<size_t>(C.nmol[t]) * h.ns[t];
      double** dst[3] = {&C.gx[t], &C.gy[t], &C.gz[t]};
      for (int c = 0; c < 3 && !r; ++c) {
        std::fill(stage.begin(), stage.end(), 0.0);
        std::memcpy(stage.data(), site, n * sizeof(double));
        site += n;
        r = dput(K, dst[c], stage.data(), C.cap_atoms);
      }
    }
    if (r) break;
    for (auto& kb : K) ch->bytes += kb->n;
    if (hipMemcpyAsync(b->d_chains + slot, &C, sizeof(BChain), hipMemcpyHostToDevice,
                       b->setup) != hipSuccess ||
        hipStreamSynchronize(b->setup) != hipSuccess)
      r = FMC_ERR_HIP;
  } while (false);
  return publish_chain(b, slot, sys, r, std::move(ch), C, chain_id);
}

extern "C" int fmc_batch_get_molecules(fmc_batch* b, int id, int gtype, double* x, double* y,
                                       double* z, int cap_sites, int* nmol) {
  if (!b || !nmol || id < 0 || id >= b->max_chains) return FMC_ERR_ARG;
  BCHECK(hipStreamSynchronize(b->stream));
  std::lock_guard<std::mutex> g(b->mu);
  if (!b->chains[id] || !b->chains[id]->active) return FMC_ERR_ARG;
  BChain C;
  BCHECK(hipMemcpy(&C, b->d_chains + id, sizeof(BChain), hipMemcpyDeviceToHost));
  if (gtype < 0 || gtype >= C.ntpguest) return FMC_ERR_ARG;
  const int ns = b->systems[C.sys]->dev.q.ns[gtype];
  *nmol = C.nmol[gtype];
  if (!x && !y && !z) return FMC_OK;
  const long long n = static_cast<long long>(C.nmol[gtype]) * ns;
  if (!x || !y || !z || cap_sites < n) return FMC_ERR_ARG;
  if (n) {
    BCHECK(hipMemcpy(x, C.gx[gtype], n * sizeof(double), hipMemcpyDeviceToHost));
    BCHECK(hipMemcpy(y, C.gy[gtype], n * sizeof(double), hipMemcpyDeviceToHost));
    BCHECK(hipMemcpy(z, C.gz[gtype], n * sizeof(double), hipMemcpyDeviceToHost));
  }
  return FMC_OK;
}

extern "C" int fmc_batch_set_tile(fmc_batch* b, int threads) {
  if (!b) return FMC_ERR_ARG;
  std::lock_guard<std::mutex> g(b->mu);
  if (threads == 0) {
    b->bt = b->kpt = 0;
    b->tuned = false;
    b->force_stream = false;
    return FMC_OK;
  }
  const bool stream = (threads & FMC_TILE_STREAM) != 0;
  const int bt = threads & ~FMC_TILE_STREAM;
  if (bt != 256 && bt != 512) return FMC_ERR_UNSUPPORTED;
  b->bt = bt;
  b->kpt = 0;
  b->tuned = true;
  b->force_stream = stream;
  return FMC_OK;
}

extern "C" int fmc_batch_get_tile(const fmc_batch* b, int* threads) {
  if (!b || !threads) return FMC_ERR_ARG;
  std::lock_guard<std::mutex> g(const_cast<fmc_batch*>(b)->mu);
  *threads = b->bt | ((b->tuned && b->kpt == 0 && b->bt) ? FMC_TILE_STREAM : 0);
  return FMC_OK;
}

extern "C" int fmc_batch_launch(fmc_batch* b, long long n) {
  if (!b || n < 0) return FMC_ERR_ARG;
  int r;
  if ((r = fmc_batch_sync(b))) return r;
  std::lock_guard<std::mutex> g(b->mu);
  if ((r = sync_tables(b))) return r;
  if (b->active_ids.empty()) return FMC_OK;
  if ((r = ensure_tile(b))) return r;
  // kernel time on the batch's own stream
  BCHECK(hipEventRecord(b->ev0, b->stream));
  if ((r = launch(b, n, b->bt, b->kpt))) return r;
  BCHECK(hipEventRecord(b->ev1, b->stream));
  b->timing_pending = true;
  b->timing_mode = 1;
  return FMC_OK;
}

extern "C" int fmc_batch_launch_budget(fmc_batch* b, long long max_moves,
                                       double budget_us) {
  if (!b || max_moves < 0 || budget_us <= 0.0) return FMC_ERR_ARG;
  int r;
  if ((r = fmc_batch_sync(b))) return r;
  std::lock_guard<std::mutex> g(b->mu);
  if ((r = sync_tables(b))) return r;
  if (b->active_ids.empty()) return FMC_OK;
  if ((r = ensure_tile(b))) return r;
  int dev = 0, khz = 0;
  BCHECK(hipGetDevice(&dev));
  BCHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) return FMC_ERR_UNSUPPORTED;
  const long long ticks = std::max(1LL, static_cast<long long>(budget_us * 1e-3 * khz));
  BCHECK(hipEventRecord(b->ev0, b->stream));
  if ((r = launch(b, max_moves, b->bt, b->kpt, ticks))) return r;
  BCHECK(hipEventRecord(b->ev1, b->stream));
  b->timing_pending = true;
  b->timing_mode = 1;
  return FMC_OK;
}

extern "C" int fmc_batch_sync(fmc_batch* b) {
  if (!b) return FMC_ERR_ARG;
  BCHECK(hipStreamSynchronize(b->stream));   // the wait holds no lock
  std::lock_guard<std::mutex> g(b->mu);
  if (b->timing_pending) {
    // mode 2: one batch-kernel graph replay per event pair
    float ms = 0.0f;
    if (b->timing_mode == 2) {
      for (long long c = 0; c < b->cyc_pending; ++c) {
        BCHECK(hipEventElapsedTime(&ms, b->cyc_ev[c].first, b->cyc_ev[c].second));
        b->kernel_ms += ms;
        b->launches++;
      }
    } else {
      BCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      b->kernel_ms += ms;
      b->launches++;
    }
    b->timing_pending = false;
  }
  return FMC_OK;
}

extern "C" int fmc_batch_query(fmc_batch* b, int* done) {
  if (!b || !done) return FMC_ERR_ARG;
  const hipError_t e = hipStreamQuery(b->stream);
  if (e == hipSuccess) {
    *done = 1;
    return FMC_OK;
  }
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();   // not-ready is not an error to keep
    *done = 0;
    return FMC_OK;
  }
  BCHECK(e);
  return FMC_OK;
}

extern "C" int fmc_batch_timing(fmc_batch* b, int reset, double* ms,
                                long long* launches) {
  if (!b) return FMC_ERR_ARG;
  std::lock_guard<std::mutex> g(b->mu);
  if (ms) *ms = b->kernel_ms;
  if (launches) *launches = b->launches;
  if (reset) {
    b->kernel_ms = 0.0;
    b->launches = 0;
  }
  return FMC_OK;
}

extern "C" int fmc_batch_run(fmc_batch* b, long long n) {
  int r = fmc_batch_launch(b, n);
  if (r) return r;
  return fmc_batch_sync(b);
}

extern "C" int fmc_batch_stats(fmc_batch* b, int id, fmc_chain_stats* st) {
  if (!b || !st || id < 0 || id >= b->max_chains) return FMC_ERR_ARG;
  // a launch in flight holds the chain struct in LDS and writes all of it
  // back at its end: read it only after the stream has drained
  BCHECK(hipStreamSynchronize(b->stream));
  std::lock_guard<std::mutex> g(b->mu);
  if (!b->chains[id]) return FMC_ERR_ARG;
  BChain C;
  BCHECK(hipMemcpy(&C, b->d_chains + id, sizeof(BChain), hipMemcpyDeviceToHost));
  std::memset(st, 0, sizeof(*st));
  st->steps = C.steps;
  st->prod_steps = C.prod_steps;
  for (int i = 0; i < 6; ++i) {
    st->accepted[i] = C.accepted[i];
    st->attempted[i] = C.attempted[i];
  }
  for (int g = 0; g < kMaxGuestTypes; ++g) {
    st->nmol[g] = C.nmol[g];
    st->energy[g] = C.energy[g];
    st->sum_n[g] = C.sum_n[g];
    st->sum_n2[g] = C.sum_n2[g];
    st->sum_e[g] = C.sum_e[g];
    st->sum_en[g] = C.sum_en[g];
    st->sum_e2[g] = C.sum_e2[g];
    st->delrdisp[g] = C.delrdisp[g];
    st->tran_delr[g] = C.tran_delr[g];
  }
  st->rotangle = C.rotangle;
  st->rota_rotangle = C.rota_rotangle;
  st->e_acc = C.e_acc;
  st->e_full = C.e_full;
  st->drift = C.drift;
  st->sk_dev = C.sk_dev;
  st->recomputes = C.n_recompute;
  st->sel_window = C.sel_nwind;
  st->sel_count = C.sel_count;
  for (int g = 0; g < kMaxGuestTypes; ++g) st->sel_base[g] = C.sel_base[g];
  for (int q = 0; q < kSelPairs; ++q) {
    st->sel_mean[q] = C.sel_mean[q];
    st->sel_second[q] = C.sel_sec[q];
    st->sel_var[q] = C.sel_var[q];
  }
  st->win_len = C.win_len;
  st->win_closed = 0;
  if (C.win_len > 0) {
    // running sums at each window bound: a snapshot, or the current sums
    // for a bound not reached (samples past the end read zero)
    double w[kWinDoubles];
    BCHECK(hipMemcpy(w, C.sre + b->systems[C.sys]->dev.nk, sizeof(w),
                     hipMemcpyDeviceToHost));
    const double* now[kWinQ] = {C.sum_n, C.sum_e, C.sum_en, C.sum_n2, C.sum_e2};
    auto at = [&](int k, int q, int g) {
      return k < C.win_snap ? w[(k * kWinQ + q) * kMaxGuestTypes + g] : now[q][g];
    };
    for (int j = 0; j < kWin; ++j) {
      const int k_end = j < kWin - 1 ? 2 * j + 1 : kWinSnap - 1;   // count jL+L+1
      const int k_beg = 2 * j - 2;                                  // count jL
      if (k_end < C.win_snap) st->win_closed = j + 1;
      for (int q = 0; q < kWinQ; ++q)
        for (int g = 0; g < kMaxGuestTypes; ++g)
          st->win[j][q][g] = at(k_end, q, g) - (j == 0 ? 0.0 : at(k_beg, q, g));
    }
  }
  return C.fault ? FMC_ERR_CAPACITY : FMC_OK;
}

extern "C" int fmc_batch_enable_log(fmc_batch* b, int id, long long cap) {
  if (!b || id < 0 || id >= b->max_chains || cap < 0) return FMC_ERR_ARG;
  // the struct is rewritten here; a launch in flight would overwrite it
  BCHECK(hipStreamSynchronize(b->stream));
  std::lock_guard<std::mutex> g(b->mu);
  if (!b->chains[id]) return FMC_ERR_ARG;
  BChain C;
  BCHECK(hipMemcpy(&C, b->d_chains + id, sizeof(BChain), hipMemcpyDeviceToHost));
  double* lg = nullptr;
  int r;
  if ((r = dput(b->chains[id]->keep, &lg, static_cast<double*>(nullptr),
                static_cast<size_t>(cap) * kLogFields)))
    return r;
  C.log = lg;
  C.log_cap = cap;
  C.log_n = 0;
  BCHECK(hipMemcpy(b->d_chains + id, &C, sizeof(BChain), hipMemcpyHostToDevice));
  return FMC_OK;
}

extern "C" int fmc_batch_get_log(fmc_batch* b, int id, long long n, double* rec) {
  if (!b || !rec || id < 0 || id >= b->max_chains) return FMC_ERR_ARG;
  BCHECK(hipStreamSynchronize(b->stream));
  std::lock_guard<std::mutex> g(b->mu);
  if (!b->chains[id]) return FMC_ERR_ARG;
  BChain C;
  BCHECK(hipMemcpy(&C, b->d_chains + id, sizeof(BChain), hipMemcpyDeviceToHost));
  if (!C.log || n > C.log_n) return FMC_ERR_ARG;
  BCHECK(hipMemcpy(rec, C.log, static_cast<size_t>(n) * kLogFields * sizeof(double),
                   hipMemcpyDeviceToHost));
  return FMC_OK;
}

namespace {

int enqueue_recompute(fmc_batch* b) {
  const int n = static_cast<int>(b->active_ids.size());
  if (!n) return FMC_OK;
  RecomputeParams P;
  P.sys = b->d_sys;
  P.chains = b->d_chains;
  P.ids = b->d_ids;
  P.maxdev = b->d_dev;
  P.base = 0;
  P.one_id = -1;
  P.one = nullptr;
  hipLaunchKernelGGL(recompute_kernel, dim3(n), dim3(kRecBT), 0, b->stream, P);
  BCHECK(hipGetLastError());
  return FMC_OK;
}

}  // namespace

extern "C" int fmc_batch_recompute(fmc_batch* b, double* max_sk_dev, double* max_drift) {
  if (!b) return FMC_ERR_ARG;
  int r;
  if ((r = fmc_batch_sync(b))) return r;
  std::lock_guard<std::mutex> g(b->mu);
  if ((r = sync_tables(b))) return r;
  if (max_sk_dev) *max_sk_dev = 0.0;
  if (max_drift) *max_drift = 0.0;
  const int n = static_cast<int>(b->active_ids.size());
  if (!n) return FMC_OK;
  if ((r = enqueue_recompute(b))) return r;
  std::vector<double> dev(n);
  BCHECK(hipMemcpyAsync(dev.data(), b->d_dev, n * sizeof(double), hipMemcpyDeviceToHost,
                        b->stream));
  BCHECK(hipMemcpyAsync(b->h_chains.data(), b->d_chains, b->max_chains * sizeof(BChain),
                        hipMemcpyDeviceToHost, b->stream));
  BCHECK(hipStreamSynchronize(b->stream));
  if (max_sk_dev) *max_sk_dev = *std::max_element(dev.begin(), dev.end());
  if (max_drift) {
    double m = 0.0;
    for (int id : b->active_ids) m = std::max(m, std::fabs(b->h_chains[id].drift));
    *max_drift = m;
  }
  return FMC_OK;
}

extern "C" int fmc_batch_resync(fmc_batch* b, double* max_dev) {
  return fmc_batch_recompute(b, max_dev, nullptr);
}

extern "C" int fmc_batch_launch_cycles(fmc_batch* b, long long n_cycles,
                                       long long moves_per_cycle) {
  if (!b || n_cycles < 0 || moves_per_cycle < 0) return FMC_ERR_ARG;
  int r;
  if ((r = fmc_batch_sync(b))) return r;
  std::lock_guard<std::mutex> g(b->mu);
  if ((r = sync_tables(b))) return r;
  if (b->active_ids.empty() || n_cycles == 0) return FMC_OK;
  if ((r = ensure_tile(b))) return r;
  if (!b->graph_exec || b->graph_ids != b->active_ids || b->graph_moves != moves_per_cycle ||
      b->graph_bt != b->bt || b->graph_kpt != b->kpt) {
    invalidate_graph(b);
    for (int part = 0; part < 2; ++part) {
      BCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int r1 = part == 0 ? launch(b, moves_per_cycle, b->bt, b->kpt)
                               : enqueue_recompute(b);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (r1) {
        if (g) (void)hipGraphDestroy(g);
        return r1;
      }
      BCHECK(e);
      hipGraphExec_t x = nullptr;
      const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (part == 0) {
        b->graph = g;
        b->graph_exec = x;
      } else {
        b->graph_rec = g;
        b->graph_rec_exec = x;
      }
      BCHECK(e2);
    }
    b->graph_ids = b->active_ids;
    b->graph_moves = moves_per_cycle;
    b->graph_bt = b->bt;
    b->graph_kpt = b->kpt;
  }
  while (static_cast<long long>(b->cyc_ev.size()) < n_cycles) {
    hipEvent_t e0, e1;
    BCHECK(hipEventCreate(&e0));
    BCHECK(hipEventCreate(&e1));
    b->cyc_ev.push_back({e0, e1});
  }
  // each cycle: [event] moves [event] FP64 recompute -- the events sit
  // between replays on the batch stream and time the batch kernel alone
  for (long long c = 0; c < n_cycles; ++c) {
    BCHECK(hipEventRecord(b->cyc_ev[c].first, b->stream));
    BCHECK(hipGraphLaunch(b->graph_exec, b->stream));
    BCHECK(hipEventRecord(b->cyc_ev[c].second, b->stream));
    BCHECK(hipGraphLaunch(b->graph_rec_exec, b->stream));
  }
  b->cyc_pending = n_cycles;
  b->timing_pending = true;
  b->timing_mode = 2;
  return FMC_OK;
}

extern "C" int fmc_batch_poll(fmc_batch* b, int* ids, int cap, int* n) {
  if (!b || !n || (cap > 0 && !ids)) return FMC_ERR_ARG;
  int r;
  if ((r = fmc_batch_sync(b))) return r;
  std::lock_guard<std::mutex> g(b->mu);
  *n = 0;
  BCHECK(hipMemcpy(b->h_chains.data(), b->d_chains, b->max_chains * sizeof(BChain),
                   hipMemcpyDeviceToHost));
  for (int i = 0; i < b->max_chains; ++i) {
    if (!b->chains[i] || !b->chains[i]->active) continue;
    const BChain& C = b->h_chains[i];
    if (C.steps >= C.stop_at || C.fault) {
      if (*n < cap) ids[*n] = i;
      ++*n;
    }
  }
  return FMC_OK;
}
