// Device side of the batched chains: the chain and system structures,
// the move loop's device routines and the persistent kernel template.
// fmc_batch.hip holds the host side; fmc_batch_tiles_*.hip instantiate
// the kernel's tile shapes.
//
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
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "fastmc_amd.h"
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
  double gce[kMaxGuestTypes][kMaxSites];   // q / epsq * r4pie0 per site
  // pairtab row of each guest site's type (ntpatm <= kLdsTypes)
  short sptab[kMaxGuestTypes][kMaxSites][kLdsTypes];
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
  long long pend;   // production samples of the current state not yet in the sums
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
  int tile_code;   // tile shape of the launches that moved the chain (pack_tile)
  int tile_pad;
};

// the tile shape a chain's launches ran with, as BChain::tile_code holds it
__host__ __device__ constexpr int pack_tile(int bt, int kpt, int kg, int ku) {
  return (bt / 64) | (kpt << 6) | (kg << 12) | (ku << 16);
}

// Dynamic LDS of the batch kernel, from a 16-byte aligned base: the erfc
// table, a_k and the packed k index (register form only), then the X.Y
// tables. Every region is padded to 16 bytes so the FP64 X.Y entries
// (complex doubles, read as 128-bit loads) stay aligned whatever the table
// sizes; an 8-byte aligned 128-bit LDS read is split by the hardware.
__host__ __device__ constexpr int dyn_erc_slots(int ercap) { return (ercap + 1) & ~1; }
__host__ __device__ constexpr int dyn_ak_slots(int nkcap) { return (nkcap + 1) & ~1; }
__host__ __device__ constexpr int dyn_kv_slots(int nkcap) { return (nkcap + 3) & ~3; }

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
static __device__ bool accept_test(double eng, double rande, double volm,
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

static __device__ double tail_of(const SysQ& Z, const int* n) {
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

// Stage clocks (a profiling build only, -DFMC_STAGE_CLOCKS): each chain's
// work-group sums s_memtime cycles per stage of its moves, and per wave the
// end of its parallel-stage work and of its decision-stage work, into
// g_stage_clocks[work-group][slot] (slots: kClk* below).  Each compiled
// tile has its own copy, read by fmc_stage_clocks_<threads>_<kpt>.
constexpr int kClkPropose = 0, kClkTables = 1, kClkParallel = 2, kClkDecision = 3,
              kClkCommit = 4, kClkMoves = 5, kClkFresh = 6, kClkDisplace = 7,
              kClkWaveParallel = 8, kClkWaveDecision = 16, kClkWaveRecip = 24,
              kClkBuild = 32, kClkPrep = 36, kClkSlots = 40;
constexpr int kClkGroups = 4096;
#ifdef FMC_STAGE_CLOCKS
static __device__ unsigned long long g_stage_clocks[kClkGroups][kClkSlots];
#define FMC_CLOCK(v) const unsigned long long v = __builtin_amdgcn_s_memtime()
#else
#define FMC_CLOCK(v)
#endif

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
// q X Y tabulated per (l, m) row, the positions of a row adjacent
// (xy[lm * pcap + p]), and Z per position and signed n (conjugates folded
// in, zt[p * kZRow + n]); four multiply-adds per position, accumulated in
// T over the positions in order.  Both kernel forms call these routines
// (contraction off, fused multiply-adds written out), so they round
// identically.
constexpr int kZRow = 2 * kBTab;   // Z table row: n = -16..16
constexpr int kMaxPos = 2 * kMaxSites;
template <typename T>
__device__ __forceinline__ void trial_ds(int np, const Cplx<T>* xy, int xo, const Cplx<T>* zt,
                                         int zo, T& dr, T& di) {
#pragma clang fp contract(off)
  T ar = 0, ai = 0;
  for (int p = 0; p < np; ++p) {
    const Cplx<T> a = xy[xo + p];
    const Cplx<T> c = zt[p * kZRow + zo];
    ar = fmat(a.x, c.x, ar);
    ar = fmat(-a.y, c.y, ar);
    ai = fmat(a.y, c.x, ai);
    ai = fmat(a.x, c.y, ai);
  }
  dr = ar;
  di = ai;
}

// trial_ds for the k-vectors a thread owns, kpt of its KPT slots (kpt the
// same on every thread), xo/zo their row offsets in the two tables.  The
// positions are unrolled (a uniform exit after the last), so every read
// takes its position's offset as an immediate; per position the LDS reads
// of a group of G k-vectors issue together and wait once.  Each k sums
// its positions in the same order as trial_ds.  A slot of a group beyond
// the thread's k-vectors reads row 0 and its sum is never used.
// An LDS table entry by byte address; the address is made opaque to the
// optimiser (an empty asm that only passes it through a register), so a
// constant added to it stays the read's immediate offset instead of being
// folded into a scalar base and re-added per read.
__device__ __forceinline__ unsigned lds_opaque(unsigned a) {
  asm volatile("" : "+v"(a));
  return a;
}
template <typename T>
__device__ __forceinline__ Cplx<T> lds_cplx(unsigned addr) {
  typedef T V2 __attribute__((ext_vector_type(2)));
  const V2 v = *reinterpret_cast<const __attribute__((address_space(3))) V2*>(
      static_cast<uintptr_t>(addr));
  return {v.x, v.y};
}

// Position P's terms for the k-vectors of one thread: per group of G
// k-vectors, every LDS read is issued before the first multiply-add waits
// on one (the scheduler would otherwise interleave them pairwise under
// register pressure).  A slot of a group beyond the thread's k-vectors reads
// row 0 and its sum is never used.
template <typename T, int KPT, int G, int P>
__device__ __forceinline__ void trial_ds_position(const unsigned (&xa)[KPT],
                                                  const unsigned (&za)[KPT], int kpt,
                                                  T (&dr)[KPT], T (&di)[KPT]) {
#pragma clang fp contract(off)
  constexpr int NG = (KPT + G - 1) / G;
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    if (g * G < kpt) {
      Cplx<T> a[G], c[G];
#pragma unroll
      for (int jj = 0; jj < G; ++jj) {
        const int j = g * G + jj;
        if (j >= KPT) break;
        a[jj] = lds_cplx<T>(xa[j] + P * sizeof(Cplx<T>));
        c[jj] = lds_cplx<T>(za[j] + P * kZRow * sizeof(Cplx<T>));
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int jj = 0; jj < G; ++jj) {
        const int j = g * G + jj;
        if (j >= KPT) break;
        dr[j] = fmat(a[jj].x, c[jj].x, dr[j]);
        dr[j] = fmat(-a[jj].y, c[jj].y, dr[j]);
        di[j] = fmat(a[jj].y, c[jj].x, di[j]);
        di[j] = fmat(a[jj].x, c[jj].y, di[j]);
      }
    }
  }
}

template <typename T, int KPT, int G, int... P>
__device__ __forceinline__ void trial_ds_positions(int np, const unsigned (&xa)[KPT],
                                                   const unsigned (&za)[KPT], int kpt,
                                                   T (&dr)[KPT], T (&di)[KPT],
                                                   std::integer_sequence<int, P...>) {
  // positions 0, 1, ... in order, stopping after the last (np is uniform)
  (void)((P < np ? (trial_ds_position<T, KPT, G, P>(xa, za, kpt, dr, di), true) : false) &&
         ...);
}

// trial_ds for the k-vectors a thread owns, kpt of its KPT slots (kpt the
// same on every thread), xo/zo their row offsets in the two tables.  The
// positions are unrolled, so every read takes its position's offset as an
// immediate from the k-vector's row address; each k sums its positions in
// the same order as trial_ds.
template <typename T, int KPT, int G>
__device__ __forceinline__ void trial_ds_tile(int np, const Cplx<T>* xy, const Cplx<T>* zt,
                                              int kpt, const int (&xo)[KPT],
                                              const int (&zo)[KPT], T (&dr)[KPT],
                                              T (&di)[KPT]) {
  unsigned xa[KPT], za[KPT];
  const unsigned xb = static_cast<unsigned>(reinterpret_cast<uintptr_t>(xy));
  const unsigned zb = static_cast<unsigned>(reinterpret_cast<uintptr_t>(zt));
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    dr[j] = 0;
    di[j] = 0;
    xa[j] = lds_opaque(xb + static_cast<unsigned>(xo[j]) * sizeof(Cplx<T>));
    za[j] = lds_opaque(zb + static_cast<unsigned>(zo[j]) * sizeof(Cplx<T>));
  }
  trial_ds_positions<T, KPT, G>(np, xa, za, kpt, dr, di,
                                std::make_integer_sequence<int, kMaxPos>{});
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
static __device__ void propose(const SysQ& Z, BChain& C, Rng& R, MoveDesc& D, int lane) {
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

// The production sums take every production move's sample (N_g, E_g)
// (gcmc.f:1814-1848).  The sample changes only when a move is accepted, so
// a move counts its sample in pend and the repeats are added in one step,
// r N, r N^2, r E, r E N, r E^2, just before an accepted move changes the
// state.  A reader of the sums in between (a statistics window, the host)
// takes the sums with the pending repeats added the same way without
// storing them, so the sums' bits follow from the trajectory alone, not
// from where launches split it or which windows are kept.  The N sums are
// integers and stay exact; an E sum differs from the sample-by-sample sum
// by rounding only.
struct SampleSums {
  double n, n2, e, en, e2;
};

__host__ __device__ inline SampleSums sums_with_pending(const BChain& C, int g) {
#pragma clang fp contract(off)
  const double r = static_cast<double>(C.pend);
  const double n = C.nmol[g];
  const double e = C.energy[g];
  if (C.pend == 0) return {C.sum_n[g], C.sum_n2[g], C.sum_e[g], C.sum_en[g], C.sum_e2[g]};
  return {C.sum_n[g] + r * n, C.sum_n2[g] + r * (n * n), C.sum_e[g] + r * e,
          C.sum_en[g] + r * (e * n), C.sum_e2[g] + r * (e * e)};
}

__host__ __device__ inline void fold_samples(BChain& C) {
  if (C.pend == 0) return;
  for (int g = 0; g < C.ntpguest; ++g) {
    const SampleSums t = sums_with_pending(C, g);
    C.sum_n[g] = t.n;
    C.sum_n2[g] = t.n2;
    C.sum_e[g] = t.e;
    C.sum_en[g] = t.en;
    C.sum_e2[g] = t.e2;
  }
  C.pend = 0;
}

static __device__ void window_snapshot(BChain& C, double* snaps) {
  while (C.win_snap < kWinSnap && C.prod_steps == win_target(C.win_snap, C.win_len)) {
    double* w = snaps + C.win_snap * kWinQ * kMaxGuestTypes;
    for (int g = 0; g < kMaxGuestTypes; ++g) {
      const SampleSums t = g < C.ntpguest ? sums_with_pending(C, g)
                                          : SampleSums{C.sum_n[g], C.sum_n2[g], C.sum_e[g],
                                                       C.sum_en[g], C.sum_e2[g]};
      gstore(w + 0 * kMaxGuestTypes + g, t.n);
      gstore(w + 1 * kMaxGuestTypes + g, t.e);
      gstore(w + 2 * kMaxGuestTypes + g, t.en);
      gstore(w + 3 * kMaxGuestTypes + g, t.n2);
      gstore(w + 4 * kMaxGuestTypes + g, t.e2);
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
  // the N sums with the pending samples (exact integers)
  double sn[kMaxGuestTypes];
  for (int g = 0; g < C.ntpguest; ++g) sn[g] = sums_with_pending(C, g).n;
  for (int i = 1; i < C.ntpguest; ++i) {
    const double aN = (sn[i] - C.sel_base[i]) / nw;
    if (aN == 0.0) continue;
    for (int j = 0; j < i; ++j) {
      const int p = i * (i - 1) / 2 + j;
      const double sij = ((sn[j] - C.sel_base[j]) / nw) / aN;
      const double delta = weight * (sij - C.sel_mean[p]);
      C.sel_mean[p] = C.sel_mean[p] + delta / av;
      C.sel_sec[p] = C.sel_sec[p] + weight * (sij * sij - C.sel_sec[p]) / (av - 1.0 + weight);
      const double w = av - 1.0 + weight;
      C.sel_var[p] = (w * C.sel_sec[p] - (w * C.sel_mean[p] * C.sel_mean[p])) / w;
    }
  }
  for (int g = 0; g < C.ntpguest; ++g) C.sel_base[g] = sn[g];
}

// Production statistics of one move (gcmc.f:1814-1848), after C.steps has
// counted it; true when it is a production sample.  The sample itself is
// counted in pend (fold_samples).
__device__ __forceinline__ bool production_sample(BChain& C) {
  if (C.steps <= C.eq_steps) return false;
  C.prod_steps++;
  C.pend++;
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
// insert or delete), so estep > E* = -ln(rande / pref) / beta rejects.  E* is
// estimated in FP32 (relative error ~1e-7 of |ln rande| + |ln pref|, over
// beta, in kcal/mol) and the bound adds a margin of 1e-4 (1 + |E*|), several
// hundred times that; a move at or below the bound is decided exactly.
// Returns false when no bound applies (rande 0, a zero or non-finite
// prefactor, the zero-pressure delete, a no-op).
static __device__ bool reject_bound(const BChain& C, const MoveDesc& D, double rande, double* thr) {
  const int ig = D.ig;
  float pref;
  if (D.kind == FMC_DISPLACE) {
    pref = 1.0f;
  } else if (D.kind == FMC_INSERT) {
    pref = static_cast<float>(C.statvolm) * static_cast<float>(C.press[ig]) /
           (static_cast<float>(C.nmol[ig] + 1) * static_cast<float>(kBoltz * C.temp));
  } else if (D.kind == FMC_DELETE) {
    if (fabs(C.press[ig]) < 1e-7) return false;
    pref = static_cast<float>(C.nmol[ig]) * static_cast<float>(kBoltz * C.temp) /
           (static_cast<float>(C.statvolm) * static_cast<float>(C.press[ig]));
  } else {
    return false;
  }
  if (!(rande > 0.0) || !(pref > 0.0f) || !(pref < 3.0e38f)) return false;
  const float e = -(logf(static_cast<float>(rande)) - logf(pref)) / static_cast<float>(C.beta);
  *thr = static_cast<double>(e) + 1e-4 * (1.0 + fabs(static_cast<double>(e)));
  return true;
}

__device__ __forceinline__ double tail_change(const SysQ& Z, const BChain& C, const MoveDesc& D) {
  int n2[kMaxGuestTypes];
  for (int g = 0; g < Z.ngt; ++g) n2[g] = C.nmol[g];
  n2[D.ig] += (D.kind == FMC_INSERT) ? 1 : -1;
  return tail_of(Z, n2) - C.elrc;
}

// The decision's inputs of move D on chain state C, from the duni state
// (table u, indices ir and jr, carry c0) its proposal left: the acceptance
// draw (read without taking it), the tail change of an insert or delete,
// and the estep above which the move is certainly rejected (fast: the
// bound applies).
struct DecisionIn {
  double rande, delrc, thr;
  int fast;
};

__device__ __forceinline__ DecisionIn decision_inputs(const SysQ& Z, const BChain& C,
                                                      const MoveDesc& D, const float* u,
                                                      int ir, int jr, float c0) {
  float uni = u[ir - 1] - u[jr - 1];
  if (uni < 0.0f) uni = uni + 1.0f;
  float c = c0 - C.cd;
  if (c < 0.0f) c = c + C.cm;
  uni = uni - c;
  if (uni < 0.0f) uni = uni + 1.0f;
  DecisionIn in;
  in.rande = uni;
  in.delrc = (D.kind == FMC_INSERT || D.kind == FMC_DELETE) ? tail_change(Z, C, D) : 0.0;
  in.thr = 0.0;
  in.fast = reject_bound(C, D, in.rande, &in.thr) ? 1 : 0;
  return in;
}

// Acceptance and bookkeeping on the serial lane (gcmc.f:938-1002, 1106-1172,
// 1326-1378, 1814-1870).
static __device__ void decide(const SysQ& Z, BChain& C, Rng& R, MoveDesc& D,
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
      fold_samples(C);   // the samples so far are of the state before the move
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
      fold_samples(C);
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

// The per-move tables of a trial's positions, built in two phases that a
// work-group barrier separates, each spread over the threads t < nt of the
// caller's range with every entry written by exactly one thread:
//   table_phases: the phase factors e^{i 2 pi l f} of every position, every
//     dimension and l (ewald.f:97-137), one sincospi per entry, into tab;
//   table_rows:   from them, each position's q X[l] Y[m] row and signed Z[n]
//     row for the reciprocal sum, its pair-table row, and its coordinates,
//     charge factors and candidate-list start for real space.
//
// A displacement in DELTA form (mixed precision) sums, per site, the change
// of q e^{ik.r} rather than the new term minus the old one:
//   q (P_n Z_n - P_o Z_o) = q P_o (Z_n - Z_o) + q (P_n - P_o) Z_n,
// with P = X[l] Y[m].  The old position's rows are (q P_o, Z_n - Z_o), the
// new position's (q (P_n - P_o), Z_n).  The differences are formed from
// FP64 phase factors in FP64 and only then rounded to T, so their rounding
// error is relative to the change of the term, not to the term: the FP32
// difference of two nearly equal phase factors would lose the digits that
// the framework's large |S(k)| amplifies in 2 a_k S.dS.  The partner
// position's phase factors each row needs go to tabp.
// Positions follow propose's layout: the nold old sites first (a delete or
// a displacement), then the new ones, so a position's set and site follow
// from its index.
__device__ __forceinline__ int pos_nold(const MoveDesc& D) {
  return D.kind != FMC_INSERT ? D.ns : 0;
}

template <bool DELTA>
__device__ __forceinline__ void table_phases(const SysQ& Z, const MoveDesc& D, int t, int nt,
                                             Cplx<double> (*tab)[3][kBTab],
                                             Cplx<double> (*tabp)[3][kBTab]) {
  const int np = D.np, ns = D.ns, nold = pos_nold(D);
  const bool delta = DELTA && D.kind == FMC_DISPLACE;
  const int k0 = Z.kmax[0] + 1, k1 = Z.kmax[1] + 1, k2 = Z.kmax[2] + 1;
  const int nown = k0 + k1 + k2;
  const float rnown = 1.0f / nown, rk2 = 1.0f / k2, rk01 = 1.0f / (k0 + k1);
  // items: every position's own entries (kmax+1 per dimension, packed),
  // then in delta form the partner entries the rows need: the new Z for
  // each old position, then the old X and Y for each new one
  const int nown_all = np * nown;
  const int total = nown_all + (delta ? ns * nown : 0);
  for (int i = t; i < total; i += nt) {
    int p, d, l;
    bool own = true;
    if (i < nown_all) {
      p = div_rc(i, nown, rnown);
      const int j = i - p * nown;
      d = j < k0 ? 0 : (j < k0 + k1 ? 1 : 2);
      l = j - (d == 0 ? 0 : (d == 1 ? k0 : k0 + k1));
    } else {
      own = false;
      int ii = i - nown_all;
      if (ii < ns * k2) {
        p = div_rc(ii, k2, rk2);   // old site p
        d = 2;
        l = ii - p * k2;
      } else {
        ii -= ns * k2;
        const int sn = div_rc(ii, k0 + k1, rk01);
        p = ns + sn;                // new site sn
        const int jj = ii - sn * (k0 + k1);
        d = jj < k0 ? 0 : 1;
        l = jj - (d == 1 ? k0 : 0);
      }
    }
    const bool set = p >= nold;
    const int s = set ? p - nold : p;
    const bool newpos = own ? set : !set;
    const double px = newpos ? D.nx[s] : D.ox[s];
    const double py = newpos ? D.ny[s] : D.oy[s];
    const double pz = newpos ? D.nz[s] : D.oz[s];
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
}

template <typename T, bool DELTA>
__device__ __forceinline__ void table_rows(const SysQ& Z, const BSys& Y, MoveDesc& D, int t,
                                           int nt, const Cplx<double> (*tab)[3][kBTab],
                                           const Cplx<double> (*tabp)[3][kBTab], Cplx<T>* xy_l,
                                           Cplx<T>* zt, int nlm, int pcap) {
  const int np = D.np, ig = D.ig, kind = D.kind, nold = pos_nold(D);
  const int ntp = Y.S.ntpatm;
  const int w2 = 2 * Z.kmax[1] + 1, k3 = Z.kmax[2], nz = 2 * k3 + 1;
  const float rw2 = 1.0f / w2, rnlm = 1.0f / nlm, rnz = 1.0f / nz;
  const bool delta = DELTA && kind == FMC_DISPLACE;
  // per position: coordinates, charge factors, site type, sign, cell, on
  // the range's last threads (the items below go round-robin, so those
  // have the fewest)
  for (int p = nt - 1 - t; p >= 0 && p < np; p += nt) {
    const bool set = p >= nold;
    const int s = set ? p - nold : p;
    const double x = set ? D.nx[s] : D.ox[s];
    const double y = set ? D.ny[s] : D.oy[s];
    const double z = set ? D.nz[s] : D.oz[s];
    const double q = Z.gq[ig][s];
    D.plo[p] = cell_of(Z, x, y, z) * Y.gcap;
    D.ppx[p] = x;
    D.ppy[p] = y;
    D.ppz[p] = z;
    D.pce[p] = Z.gce[ig][s];
    D.pq[p] = (set || delta) ? q : -q;
    D.pts[p] = Z.gtype[ig][s];
    D.psg[p] = (kind == FMC_DISPLACE && !set) ? -1 : 1;
  }
  // pair-table rows
  if (ntp <= kLdsTypes)
    for (int i = t; i < np * kLdsTypes; i += nt) {
      const int p = i / kLdsTypes, tj = i - p * kLdsTypes;
      if (tj < ntp) D.ptab[p][tj] = Z.sptab[ig][p >= nold ? p - nold : p][tj];
    }
  // X[l] * Y[m] for every (l, m) (m < 0 conjugates Y); in delta form the
  // new position's row is X_n Y_n - X_o Y_o.  Item i is row i mod nlm of
  // position i / nlm: neighbouring threads read the same or neighbouring
  // phase factors.
  for (int i = t; i < np * nlm; i += nt) {
    const int p = div_rc(i, nlm, rnlm), r = i - p * nlm;
    const bool set = p >= nold;
    const double q = Z.gq[ig][set ? p - nold : p];
    const double pq = (set || delta) ? q : -q;
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
    if (delta && set) {
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
    xy_l[r * pcap + p] = {static_cast<T>(pq * ox), static_cast<T>(pq * oy)};
  }
  // signed Z[n]; in delta form the old position's row is Z_n - Z_o
  for (int i = t; i < np * nz; i += nt) {
    const int p = div_rc(i, nz, rnz), k = i - p * nz;
    const int n = k - k3;
    const int an = n < 0 ? -n : n;
    Cplx<double> c = tab[p][2][an];
    if (delta && p < nold) {
      const Cplx<double> e = tabp[p][2][an];
      c = {e.x - c.x, e.y - c.y};
    }
    zt[p * kZRow + k] = {static_cast<T>(c.x), static_cast<T>(n < 0 ? -c.y : c.y)};
  }
}

// ---------------------------------------------------------------------------
// Real space of a trial: ewald2 + srfrce (ewald.f:632-694,
// vdw_terms.f:398-498) of every trial position against the framework atoms
// of its candidate cell and against every guest atom
// ---------------------------------------------------------------------------
// Minimum image (min_image's operations, utility_pack.f:1138-1238): a
// cubic or orthorhombic cell (imcon 1-2) from cell lengths and inverses
// held in registers, a parallelepiped (imcon 3) through fractional
// coordinates.  Rounding to even (one instruction) replaces Fortran's nint
// (half away from zero, several): the two differ only where the scaled
// separation is exactly +-1/2, where both images lie at the same distance
// to within one rounding.
struct OrthoCell {
  double a0, a1, a2, i0, i1, i2;
};

__device__ __forceinline__ OrthoCell ortho_cell(const DevSys& S) {
  const bool cubic = S.imcon != 2;   // imcon 1 takes cell(1) on every axis
  return {S.cell[0], cubic ? S.cell[0] : S.cell[4], cubic ? S.cell[0] : S.cell[8],
          S.inv[0], cubic ? S.inv[0] : S.inv[1], cubic ? S.inv[0] : S.inv[2]};
}

template <bool TRI>
__device__ __forceinline__ void image(const DevSys& S, const OrthoCell& o, double& x, double& y,
                                      double& z) {
#pragma clang fp contract(off)
  if (TRI) {
    const double sx = S.rcell[0] * x + S.rcell[3] * y + S.rcell[6] * z;
    const double sy = S.rcell[1] * x + S.rcell[4] * y + S.rcell[7] * z;
    const double sz = S.rcell[2] * x + S.rcell[5] * y + S.rcell[8] * z;
    const double ax = sx - rint(sx), ay = sy - rint(sy), az = sz - rint(sz);
    x = S.cell[0] * ax + S.cell[3] * ay + S.cell[6] * az;
    y = S.cell[1] * ax + S.cell[4] * ay + S.cell[7] * az;
    z = S.cell[2] * ax + S.cell[5] * ay + S.cell[8] * az;
  } else {
    x -= o.a0 * rint(o.i0 * x);
    y -= o.a1 * rint(o.i1 * y);
    z -= o.a2 * rint(o.i2 * z);
  }
}

// One pair's distance-dependent inputs: whether it counts (r^2 < rcut^2),
// and the interpolation rows and fractions of its erfc and vdW tables (a
// pair that does not count reads entry 0 and contributes nothing).
// two doubles from a 16-byte aligned global address, one 16-byte load
typedef double Dbl2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ Dbl2 gload2(const double* p) {
  return *(const __attribute__((address_space(1))) Dbl2*)(p);
}

struct PairSlot {
  double pe, pv;   // fractional grid offsets
  int le, lv;      // first table entries read
  bool in;
};

__device__ __forceinline__ PairSlot pair_slot(double rsq, bool ok, double rcut2, double rdrewd,
                                              double rdlrpot, int imax, bool same_grid) {
#pragma clang fp contract(off)
  PairSlot s;
  s.in = ok && rcut2 > rsq;
  const double r = sqrt(s.in ? rsq : 1.0);
  const double ue = r * rdrewd;
  int l = static_cast<int>(ue);
  l = l > imax ? imax : l;
  s.pe = ue - static_cast<double>(l);
  s.le = s.in ? l : 0;
  if (same_grid) {   // the usual case, rvdw = rcut: one grid for both tables
    s.pv = s.pe;
    s.lv = s.le;
    return s;
  }
  const double uv = r * rdlrpot;
  l = static_cast<int>(uv);
  l = l > imax ? imax : l;
  s.pv = uv - static_cast<double>(l);
  s.lv = s.in ? l : 0;
  return s;
}

// the vdW table of each item's (site type, atom type) pair, -1 for none:
// from the position's pairtab row in LDS, or from pairtab itself
template <int KU>
__device__ __forceinline__ void pair_rows(const DevSys& S, const MoveDesc& D, bool ldsrow,
                                          const int (&pp)[KU], const int (&ty)[KU],
                                          int (&vk)[KU]) {
  if (ldsrow) {
#pragma unroll
    for (int u = 0; u < KU; ++u) vk[u] = D.ptab[pp[u]][ty[u]];
  } else {
#pragma unroll
    for (int u = 0; u < KU; ++u) vk[u] = gload(S.pairtab + D.pts[pp[u]] * S.ntpatm + ty[u]);
  }
}

// interp3's formula on table values already loaded
__device__ __forceinline__ double interp3v(double v0, double v1, double v2, double p) {
#pragma clang fp contract(off)
  const double t1 = v0 + (v1 - v0) * p;
  const double t2 = v1 + (v2 - v1) * (p - 1.0);
  return t1 + (t2 - t1) * p * 0.5;
}

// Thread t of nt takes items t, t + nt, t + 2 nt, ... of the flattened list
// (the candidate records of every position, np * gcap of them, then every
// guest atom of every position) and adds their terms in that order, signed
// per position, into e2 (Coulomb) and ev (vdW).  KU items per pass: the
// pass issues every record load, then every table load, before it uses
// any of them, and no load sits under a branch (an item past the end or
// excluded reads a valid slot and is masked), so a pass waits for one
// record latency and one table latency.  The terms and their order are
// those of pair_fp64_vk per item, whatever KU is.
template <int KU, bool TRI>
__device__ __forceinline__ void real_space(const BSys& Y, const SysQ& Z, const BChain& C,
                                           const MoveDesc& D, const double* erc_l,
                                           bool erc_in_lds, int t,
                                           int nt, double& e2, double& ev) {
  // no contraction: an item's terms round the same way in every kernel
  // shape that inlines this (the oracle's build contracts nothing either)
#pragma clang fp contract(off)
  const DevSys& S = Y.S;
  const OrthoCell oc = ortho_cell(S);
  const double rcut2 = S.rcut2, rdrewd = S.rdrewd, rdlrpot = S.rdlrpot;
  const bool same_grid = rdrewd == rdlrpot;
  const int imax = S.mxgrid - 2;
  const size_t mx1 = static_cast<size_t>(S.mxgrid) + 1;
  // the erfc table: its LDS copy, or when that did not fit the table
  // itself (the same values)
  const double* __restrict__ erc = erc_l;
  const double* __restrict__ erc_g = S.erc;
  const double* __restrict__ vvv = S.vvv;
  const int ntp = S.ntpatm;
  const bool ldsrow = ntp <= kLdsTypes;
  const int gcap = Y.gcap;
  const int np = D.np;
  const int nfwp = np * gcap;
  const float rgcap = 1.0f / gcap;
  const Rec* __restrict__ grec = Y.grec;
  const int* __restrict__ grec_t = Y.grec_t;
  double a2 = 0.0, av = 0.0;
  // the terms of KU items whose separations are known
  auto finish = [&](const PairSlot (&sl)[KU], const double (&chg)[KU], const int (&vk)[KU],
                    const int (&pp)[KU]) {
    // every item reads its table entries, a masked one entry 0 of the
    // first table (valid, the same address on every masked lane), so no
    // value needs a default
    double e0[KU], e1[KU], e2v[KU], w0[KU], w1[KU], w2[KU];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const double* tv =
          vvv + static_cast<size_t>(vk[u] >= 0 && sl[u].in ? vk[u] : 0) * mx1 + sl[u].lv;
      if (erc_in_lds) {
        const double* te = erc + sl[u].le;
        e0[u] = te[0];
        e1[u] = te[1];
        e2v[u] = te[2];
      } else {
        const double* te = erc_g + sl[u].le;
        e0[u] = gload(te);
        e1[u] = gload(te + 1);
        e2v[u] = gload(te + 2);
      }
      w0[u] = gload(tv);
      w1[u] = gload(tv + 1);
      w2[u] = gload(tv + 2);
    }
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      // an item's terms (a masked item's are zero)
      double b2 = 0.0, bv = 0.0;
      if (sl[u].in && fabs(chg[u]) > 1e-10)
        b2 = interp3v(e0[u], e1[u], e2v[u], sl[u].pe) * chg[u];
      if (sl[u].in && vk[u] >= 0) bv = interp3v(w0[u], w1[u], w2[u], sl[u].pv);
      if (D.psg[pp[u]] < 0) {
        a2 -= b2;
        av -= bv;
      } else {
        a2 += b2;
        av += bv;
      }
    }
  };
  // framework candidates: item i is slot i mod gcap of position i / gcap
  for (int i = t; i < nfwp; i += KU * nt) {
    int pp[KU], slot[KU];
    bool ok[KU];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int idx = i + u * nt;
      ok[u] = idx < nfwp;
      const int ic = ok[u] ? idx : i;
      const int p = div_rc(ic, gcap, rgcap);
      pp[u] = p;
      slot[u] = D.plo[p] + (ic - p * gcap);
    }
    Dbl2 xy[KU], zq[KU];
    int ty[KU];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      // the wave's first active lane holds its smallest item: item 0 is
      // live there (the loop runs on live lanes only), and a later item
      // past the end there is past the end on every lane
      if (u == 0 || __builtin_amdgcn_readfirstlane(i + u * nt) < nfwp) {
        xy[u] = gload2(&grec[slot[u]].x);
        zq[u] = gload2(&grec[slot[u]].z);
        ty[u] = gload(grec_t + slot[u]);
      } else {
        xy[u] = Dbl2{0.0, 0.0};
        zq[u] = Dbl2{0.0, 0.0};
        ty[u] = 0;
      }
    }
    int vk[KU];
#pragma unroll
    for (int u = 0; u < KU; ++u) ty[u] = ty[u] < 0 ? 0 : (ty[u] >= ntp ? ntp - 1 : ty[u]);
    pair_rows(S, D, ldsrow, pp, ty, vk);
    PairSlot sl[KU];
    double chg[KU];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int p = pp[u];
      double dx = D.ppx[p] - xy[u].x, dy = D.ppy[p] - xy[u].y, dz = D.ppz[p] - zq[u].x;
      image<TRI>(S, oc, dx, dy, dz);
      const double rsq = dx * dx + dy * dy + dz * dz;
      sl[u] = pair_slot(rsq, ok[u], rcut2, rdrewd, rdlrpot, imax, same_grid);
      chg[u] = D.pce[p] * zq[u].y;
    }
    finish(sl, chg, vk, pp);
  }
  // guest atoms: item i is atom (i - nfwp) mod ng of position (i - nfwp) / ng,
  // the moving molecule's own atoms left out
  const int ng = D.nguest;
  const int total = nfwp + np * ng;
  // this thread's first item past the candidates (the last candidate pass
  // masked the ones it reached)
  int i = t < nfwp ? t + ((nfwp - t + nt - 1) / nt) * nt : t;
  if (i < total) {
    const int ig = D.ig, ns = D.ns;
    const int own0 = (D.kind != FMC_INSERT) ? D.gpre[ig] + D.imol * ns : -1;
    const float rng = 1.0f / max(ng, 1);
    for (; i < total; i += KU * nt) {
      int pp[KU], gg[KU], la[KU];
      bool ok[KU];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int idx = i + u * nt;
        const bool live = idx < total;
        const int r = (live ? idx : i) - nfwp;
        const int p = div_rc(r, ng, rng);
        const int a = r - p * ng;
        ok[u] = live && !(own0 >= 0 && a >= own0 && a < own0 + ns);
        int g = 0;
#pragma unroll
        for (int h = 1; h < kMaxGuestTypes; ++h) g += (h < Z.ngt && D.gpre[h] <= a) ? 1 : 0;
        pp[u] = p;
        gg[u] = g;
        la[u] = a - D.gpre[g];
      }
      double ax[KU], ay[KU], az[KU];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        ax[u] = gload(C.gx[gg[u]] + la[u]);
        ay[u] = gload(C.gy[gg[u]] + la[u]);
        az[u] = gload(C.gz[gg[u]] + la[u]);
      }
      PairSlot sl[KU];
      double chg[KU];
      int ty[KU], vk[KU];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int p = pp[u], g = gg[u];
        const int nsg = Z.ns[g];
        const int sa =
            la[u] - div_rc(la[u], nsg, __builtin_amdgcn_rcpf(static_cast<float>(nsg))) * nsg;
        double dx = D.ppx[p] - ax[u], dy = D.ppy[p] - ay[u], dz = D.ppz[p] - az[u];
        image<TRI>(S, oc, dx, dy, dz);
        const double rsq = dx * dx + dy * dy + dz * dz;
        sl[u] = pair_slot(rsq, ok[u], rcut2, rdrewd, rdlrpot, imax, same_grid);
        chg[u] = D.pce[p] * Z.gq[g][sa];
        ty[u] = Z.gtype[g][sa];
      }
      pair_rows(S, D, ldsrow, pp, ty, vk);
      finish(sl, chg, vk, pp);
    }
  }
  e2 = a2;
  ev = av;
}

// The candidate records of a prepared move, pulled into L2 while the move
// before it is decided, by threads t < nt of a wave that builds no tables
// for it: one dword per 128-byte line of each position's cell list.  The
// words feed a sum that is stored only under a condition no launch meets
// (nmoves < 0), which keeps the loads; the wave would otherwise wait at
// the decision's barrier.
__device__ __forceinline__ void prefetch_records(const BSys& Y, const SysQ& Z,
                                                 const MoveDesc& Dn, int t, int nt,
                                                 long long nmoves, int* sink) {
  const int np = Dn.np;
  const int nl = (Y.gcap * static_cast<int>(sizeof(Rec)) + 127) / 128;
  const float rnl = 1.0f / nl;
  const int total = np * nl;
  const char* base = reinterpret_cast<const char*>(Y.grec);
  constexpr int kPf = 8;
  int acc = 0;
  for (int j0 = t; j0 < total; j0 += kPf * nt) {
    int v[kPf];
#pragma unroll
    for (int u = 0; u < kPf; ++u) {
      const int j = j0 + u * nt;
      const int jc = j < total ? j : j0;
      const int p = div_rc(jc, nl, rnl);
      const int l = jc - p * nl;
      const int set = Dn.pset[p], s = Dn.psite[p];
      const double x = set ? Dn.nx[s] : Dn.ox[s];
      const double y = set ? Dn.ny[s] : Dn.oy[s];
      const double z = set ? Dn.nz[s] : Dn.oz[s];
      const size_t off =
          static_cast<size_t>(cell_of(Z, x, y, z)) * Y.gcap * sizeof(Rec) + l * 128;
      v[u] = gload(reinterpret_cast<const int*>(base + off));
    }
#pragma unroll
    for (int u = 0; u < kPf; ++u) acc ^= v[u];
  }
  if (acc == 0x5eed1e55 && nmoves < 0) *sink = acc;
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
//
// KG (k-vectors whose LDS reads issue together per position) and KU (real-
// space items per thread per pass) shape the latency hiding only: each k
// still sums its positions in order and each thread its items in order, so
// a chain's results depend on BT alone among the tile's parameters.
template <int BT, int KPT, bool MIXED, bool STREAM, int KG, int KU>
__global__ void __launch_bounds__(BT)
    batch_kernel(const BSys* __restrict__ systems, BChain* __restrict__ chains,
                 const int* __restrict__ ids, long long nmoves, int nkcap, int pcap, int ercap,
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
  // the decision's inputs of the move in descriptor i: its draw, tail
  // change and the estep above which it is rejected (pre_fast: that applies)
  __shared__ double pre_rande[2], pre_delrc[2], pre_thr[2];
  __shared__ int pre_fast[2];
  __shared__ double red[4][BT / kWave];
  __shared__ int stop_flag;
#ifdef FMC_STAGE_CLOCKS
  __shared__ unsigned long long clk[kClkSlots];
#endif
  __shared__ int shift_lo, shift_hi;
  __shared__ int pf_sink;   // written by no launch (keeps the record prefetch)
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
  if (tid == 0) C.tile_code = pack_tile(BT, STREAM ? 0 : KPT, KG, KU);
  // dynamic LDS (dyn_erc_slots): the erfc table, a_k [nkcap] and the
  // packed k index [nkcap] (register form), then the XY tables
  extern __shared__ __attribute__((aligned(16))) double dyn_lds[];
  // the system's erfc table (ercap entries reserved; 0 when it did not fit
  // and real space reads the table itself), read by real space
  double* erc_l = dyn_lds;
  double* ak_l = dyn_lds + dyn_erc_slots(ercap);
  int* kv_l = reinterpret_cast<int*>(ak_l + (STREAM ? 0 : dyn_ak_slots(nkcap)));
  // X[l]*Y[m] for every (l, m) of the half space, the trial positions of a
  // row adjacent: [l*(2*kmax2+1) + m + kmax2][p], pcap positions per row
  const int nlm = (Z.kmax[0] + 1) * (2 * Z.kmax[1] + 1);
  Cplx<T>* xy_l =
      reinterpret_cast<Cplx<T>*>(STREAM ? reinterpret_cast<int*>(ak_l)
                                        : kv_l + dyn_kv_slots(nkcap));
  // this thread owns k = tid + j*BT, j < kpt = ceil(nk/BT); the index
  // (l, m) is precomputed into kv's upper bits
  const int kpt = (nk + BT - 1) / BT;
  if (ercap > 0)
    for (int i = tid; i <= Y.S.mxgrid; i += BT) erc_l[i] = Y.S.erc[i];
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
#ifdef FMC_STAGE_CLOCKS
  if (tid < kClkSlots) clk[tid] = 0;
#endif
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
    FMC_CLOCK(clk_start);
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
        // the duni state after this proposal, published by the barrier
        // below (an adopted proposal's was published by the decision that
        // adopted it: no barrier follows here, and the waves that read it
        // may be ahead of wave 0)
        if (tid == 0) {
          cur_c = R.c;
          cur_ir = R.ir;
          cur_jr = R.jr;
        }
      }
    }
    if (fresh) __syncthreads();
    FMC_CLOCK(clk_proposed);
#ifdef FMC_STAGE_CLOCKS
    unsigned long long clk_tables = clk_proposed;
#endif
    // dsel is written only by thread 0's decision, after the reduction barrier
    const int dcur = dsel;
    // csel and rsel likewise (the decision may flip them while other waves
    // still read the prepared buffers)
    const int cnext = csel ^ 1, rnext = rsel ^ 1;
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
        table_phases<MIXED>(Z, D, tid, BT, tab, tabp);
        __syncthreads();
        table_rows<T, MIXED>(Z, Y, D, tid, BT, tab, tabp, xy_l, zt, nlm, pcap);
        __syncthreads();
#ifdef FMC_STAGE_CLOCKS
        clk_tables = __builtin_amdgcn_s_memtime();
#endif
      }
      // ---- reciprocal space (ewald1_guest, ewald.f:141-319) ----
      // dS(k) = sum_new q e^{ik.r} - sum_old q e^{ik.r}
      if (STREAM) {
        const int w2 = 2 * Z.kmax[1] + 1;
        for (int k = tid; k < nk; k += BT) {
          const int v = Y.kv[k];
          const int lm = ul(v) * w2 + um(v) + Z.kmax[1];
          T dr32, di32;
          trial_ds<T>(np, xy_l, lm * pcap, zt, un(v) + Z.kmax[2], dr32, di32);
          const double dr = dr32, di = di32;
          const double s0 = gload(C.sre + k), s1 = gload(C.sim + k);
          gstore(C.dsre + k, dr);
          gstore(C.dsim + k, di);
          e1 += de_k(Y.ak[k], s0, s1, dr, di);
        }
      } else {
        // fully unrolled (predicated, no early exit) so that sre/sim/dre/dim
        // keep static register indices
        // each k's row offsets in the q X Y table and the Z rows
        int xo[KPT], zo[KPT];
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
          const int k = tid + j * BT;
          const int v = (j < kpt && k < nk) ? kv_l[k] : 0;
          xo[j] = (v & 0xffff) * pcap;
          zo[j] = static_cast<signed char>((v >> 16) & 0xff) + Z.kmax[2];
        }
        trial_ds_tile<T, KPT, KG>(np, xy_l, zt, __builtin_amdgcn_readfirstlane(kpt), xo, zo, dre,
                                  dim);
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
          const int k = tid + j * BT;
          if (j < kpt && k < nk) {
            const double dr = dre[j], di = dim[j];
            e1 += de_k(ak_l[k], sre[j], sim[j], dr, di);
          }
        }
      }
#ifdef FMC_STAGE_CLOCKS
      if ((tid % kWave) == 0)
        clk[kClkWaveRecip + tid / kWave] += __builtin_amdgcn_s_memtime() - clk_tables;
#endif
      // ---- the next move's proposal, by the last wave, on copies of the
      //      chain state and duni table as this move leaves them if it is
      //      rejected (its acceptance draw taken, if it makes one) ----
      constexpr int RT = BT - kWave;   // threads on real space
      if (tid >= RT) {
        const int lane = tid - RT;
#ifdef FMC_STAGE_CLOCKS
        const unsigned long long t_p0 = __builtin_amdgcn_s_memtime();
#endif
        BChain& Cn = Cbuf[csel ^ 1];
        const long long* src = reinterpret_cast<const long long*>(&C);
        long long* dst = reinterpret_cast<long long*>(&Cn);
        static_assert(sizeof(BChain) % 8 == 0, "BChain copies in 8-byte words");
        {
          // every read of the copy issued before its writes
          constexpr int kW = static_cast<int>(sizeof(BChain) / 8);
          constexpr int kPer = (kW + kWave - 1) / kWave;
          long long w[kPer];
#pragma unroll
          for (int u = 0; u < kPer; ++u) {
            const int i = lane + u * kWave;
            w[u] = i < kW ? src[i] : 0;
          }
#pragma unroll
          for (int u = 0; u < kPer; ++u) {
            const int i = lane + u * kWave;
            if (i < kW) dst[i] = w[u];
          }
        }
        for (int i = lane; i < 97; i += kWave) rng_u[rsel ^ 1][i] = rng_u[rsel][i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
#ifdef FMC_STAGE_CLOCKS
        if (lane == 0) clk[kClkPrep] += __builtin_amdgcn_s_memtime() - t_p0;
#endif
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
#ifdef FMC_STAGE_CLOCKS
          if (lane == 0) clk[kClkPrep + 1] += __builtin_amdgcn_s_memtime() - t_p0;
#endif
          propose(Z, Cn, Rs, Dbuf[dsel ^ 1], lane);
#ifdef FMC_STAGE_CLOCKS
          if (lane == 0) clk[kClkPrep + 2] += __builtin_amdgcn_s_memtime() - t_p0;
#endif
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
      // the decision's inputs of a fresh proposal, by the first lane of the
      // wave before the last (a prepared proposal had them computed while
      // the move before it was decided)
      if (fresh && tid == RT - kWave) {
        const DecisionIn in = decision_inputs(Z, C, D, rng_u[rsel], cur_ir, cur_jr, cur_c);
        pre_rande[dcur] = in.rande;
        pre_delrc[dcur] = in.delrc;
        pre_thr[dcur] = in.thr;
        pre_fast[dcur] = in.fast;
      }
      if (tid < RT) {
        if (Y.S.imcon == 3)
          real_space<KU, true>(Y, Z, C, D, erc_l, ercap > 0, tid, RT, e2, ev);
        else
          real_space<KU, false>(Y, Z, C, D, erc_l, ercap > 0, tid, RT, e2, ev);
      }
    }
    // block reductions (fixed tree): DPP within each wave, then the waves'
    // totals in order
    FMC_CLOCK(clk_wave_done);
    e1 = wave_sum_dpp(e1);
    e2 = wave_sum_dpp(e2);
    ev = wave_sum_dpp(ev);
    if ((tid % kWave) == kWave - 1) {
      red[0][tid / kWave] = e1;
      red[1][tid / kWave] = e2;
      red[2][tid / kWave] = ev;
    }
    __syncthreads();
    FMC_CLOCK(clk_reduced);
    // the tables of the prepared next move (this move's are read no more),
    // in two phases: the phase factors on waves 1.., a barrier, then the
    // rows on waves 1 to the one before the last, while thread 0 decides
    // and the last wave computes the prepared move's decision inputs and
    // prefetches its candidate records.  The next move uses them if it
    // adopts the prepared proposal, and rebuilds them otherwise.
    MoveDesc& Dn = Dbuf[dcur ^ 1];
    const bool prep = kind != 0 && Dn.kind != 0;
    if (prep && tid >= kWave)
      table_phases<MIXED>(Z, Dn, tid - kWave, BT - kWave, tab, tabp);
    __syncthreads();
    FMC_CLOCK(clk_phases);
    if (prep && tid >= kWave) {
      if (tid < BT - kWave) {
        table_rows<T, MIXED>(Z, Y, Dn, tid - kWave, BT - 2 * kWave, tab, tabp, xy_l, zt, nlm,
                             pcap);
#ifdef FMC_STAGE_CLOCKS
        if (tid == kWave) clk[kClkBuild + 1] += __builtin_amdgcn_s_memtime() - clk_phases;
#endif
      } else {
        // the prepared move's decision inputs, on the first lane: the chain
        // state, duni table and duni state the preparing wave left (the
        // decision reads those and writes other fields)
        if (tid == BT - kWave) {
          const DecisionIn in =
              decision_inputs(Z, Cbuf[cnext], Dn, rng_u[rnext], spec_ir, spec_jr, spec_c);
          pre_rande[dcur ^ 1] = in.rande;
          pre_delrc[dcur ^ 1] = in.delrc;
          pre_thr[dcur ^ 1] = in.thr;
          pre_fast[dcur ^ 1] = in.fast;
        }
        prefetch_records(Y, Z, Dn, tid - (BT - kWave), kWave, nmoves, &pf_sink);
#ifdef FMC_STAGE_CLOCKS
        if (tid == BT - kWave) clk[kClkBuild + 2] += __builtin_amdgcn_s_memtime() - clk_phases;
#endif
      }
    }
#ifdef FMC_STAGE_CLOCKS
    if (tid == 0) clk[kClkBuild] += clk_phases - clk_reduced;
#endif
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
        const double delrc = pre_delrc[dcur], rande = pre_rande[dcur];
        if (pre_fast[dcur]) {
          // estep and the logged e1 as decide forms them; the bound first
          e1x = a;
          if (kind == FMC_INSERT) {
            e1x = e1x + Z.engsic[ig];
            estep = (e1x + b + Z.e3[ig] + c + delrc) / C.engunit;
          } else if (kind == FMC_DELETE) {
            e1x = e1x - Z.engsic[ig];
            estep = (e1x - b - Z.e3[ig] - c + delrc) / C.engunit;
          } else {
            estep = (e1x + b + c) / C.engunit;
          }
          rej = estep > pre_thr[dcur] ||
                !decision(Z, C, D, a, b, c, delrc, rande, estep, e1x);
        } else {
          rej = !decision(Z, C, D, a, b, c, delrc, rande, estep, e1x);
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
          cur_c = spec_c;
          cur_ir = spec_ir;
          cur_jr = spec_jr;
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
    FMC_CLOCK(clk_decided);
#ifdef FMC_STAGE_CLOCKS
    if (tid == 0) clk[kClkBuild + 3] += clk_decided - clk_phases;
#endif
    __syncthreads();
    FMC_CLOCK(clk_after_decision);
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
#ifdef FMC_STAGE_CLOCKS
    {
      const unsigned long long clk_end = __builtin_amdgcn_s_memtime();
      if (tid == 0) {
        clk[kClkPropose] += clk_proposed - clk_start;
        clk[kClkTables] += clk_tables - clk_proposed;
        clk[kClkParallel] += clk_reduced - clk_tables;
        clk[kClkDecision] += clk_after_decision - clk_reduced;
        clk[kClkCommit] += clk_end - clk_after_decision;
        clk[kClkMoves] += 1;
        clk[kClkFresh] += fresh ? 1 : 0;
        clk[kClkDisplace] += (kind == FMC_DISPLACE) ? 1 : 0;
      }
      if ((tid % kWave) == 0 && tid / kWave < 8) {
        clk[kClkWaveParallel + tid / kWave] += clk_wave_done - clk_tables;
        clk[kClkWaveDecision + tid / kWave] += clk_decided - clk_reduced;
      }
    }
#endif
  }
#ifdef FMC_STAGE_CLOCKS
  if (tid < kClkSlots && blockIdx.x < kClkGroups) g_stage_clocks[blockIdx.x][tid] += clk[tid];
#endif
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

// One compiled tile shape of the persistent kernel.
using BatchKernelFn = void (*)(const BSys*, BChain*, const int*, long long, int, int, int,
                               long long);
struct KCfg {
  int bt, kpt;   // kpt 0: the streaming form
  int kg, ku;
  bool mixed;
  BatchKernelFn fn;
};

}  // namespace fmc
