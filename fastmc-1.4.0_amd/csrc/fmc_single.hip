// Single-chain drop-in context: device-resident framework, guests, tables
// and running structure factor S(k); one fused kernel launch per trial move.
//
// Layout in HBM (one context):
//   framework  fx, fy, fz, fq (FP64 SoA), ftype (int, 0-based)
//   guests     per guest type a segment of capacity max_guest_atoms sites,
//              gx, gy, gz (FP64 SoA), molecule m site s at m*ns + s
//   tables     erc[mxgrid+1], vvv[ntable][mxgrid+1], pairtab[ntpatm^2]
//   k-vectors  kv[Nk] packed (l, m, n) int8 triples, ak[Nk] FP64
//   S(k)       two FP64 (re, im) buffers; a trial writes S_new into the
//              spare one and commit swaps them (ewald.f:260-293 ckcsnew and
//              gcmc.f:971-974 ckcsum <- ckcsnew)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fastmc_amd.h"
#include "fmc_ctx.h"
#include "fmc_device.h"
#include "fmc_host.h"

namespace fmc {

constexpr int kBS = 256;       // threads per block, 4 waves
constexpr int kTrialBS = 256;  // threads per trial work-group
constexpr int kTabMax = 33;    // phase table entries per dimension (kmax<=32)
constexpr int kMaxBlocks = 512;

struct GuestSeg {
  const double* x;
  const double* y;
  const double* z;
  int ns;
  int natoms;  // nmol * ns
  double q[kMaxSites];
  int t[kMaxSites];  // 0-based
};

struct TrialParams {
  DevSys S;
  int nk;
  const int* kv;
  const double* ak;
  const double* sre;
  const double* sim;
  double* nre;
  double* nim;
  int ngt;
  GuestSeg g[kMaxGuestTypes];
  int kind, ns, own_t, own_begin;
  int kmax[3];
  double nx[kMaxSites], ny[kMaxSites], nz[kMaxSites];
  double ox[kMaxSites], oy[kMaxSites], oz[kMaxSites];
  double q[kMaxSites];
  int t[kMaxSites];
  double rvolm2;   // 2*rvolm
  double engsic;   // self term of the moving molecule
  int nb_recip, nb_real;
  double* part;              // pinned host memory: [block][3] partial sums
  unsigned long long* word;  // pinned host memory: [block] sequence words
  unsigned long long seq;    // this trial's number, published after the sums
};

__device__ __forceinline__ int unpack_l(int v) { return static_cast<signed char>(v & 0xff); }
__device__ __forceinline__ int unpack_m(int v) { return static_cast<signed char>((v >> 8) & 0xff); }
__device__ __forceinline__ int unpack_n(int v) { return static_cast<signed char>((v >> 16) & 0xff); }

// Which atom is global index a: framework first, then guest segments.
__device__ __forceinline__ bool load_atom(const DevSys& S, const GuestSeg* g,
                                          int ngt, int a, double& x, double& y,
                                          double& z, double& q, int& t,
                                          int& gt, int& gi) {
  if (a < S.nfw) {
    x = S.fx[a]; y = S.fy[a]; z = S.fz[a]; q = S.fq[a]; t = S.ftype[a];
    gt = -1; gi = a;
    return true;
  }
  a -= S.nfw;
  for (int k = 0; k < ngt; ++k) {
    if (a < g[k].natoms) {
      const int s = a % g[k].ns;
      x = g[k].x[a]; y = g[k].y[a]; z = g[k].z[a];
      q = g[k].q[s]; t = g[k].t[s];
      gt = k; gi = a;
      return true;
    }
    a -= g[k].natoms;
  }
  return false;
}

// One atom against the trial's NS old and/or new sites, in the order of the
// site loop (per site: old, then new). Every table load is issued whatever
// the distance: a lane beyond the cutoff reads a clamped, valid entry (and a
// pair without a vdW table reads the erc table) and its terms are dropped,
// so the loads of all positions overlap instead of forming one dependent
// chain per position. Sums and their order are those of ewald.f:670-684 and
// vdw_terms.f:476-489 per pair.
template <int NS, bool OLD, bool NEW>
__device__ __forceinline__ void trial_atom(const TrialParams& P, int s0, double x, double y,
                                           double z, double q, int t, double& e2, double& ev) {
  const DevSys& S = P.S;
  const int imax = S.mxgrid - 2;
  const bool disp = P.kind == FMC_DISPLACE;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int s = s0 + j;
    const double ce = P.q[s] / S.epsq * kR4pie0;
    const double chgprd = ce * q;
    const bool qon = fabs(chgprd) > 1e-10;
    const int k = S.pairtab[P.t[s] * S.ntpatm + t];
    const double* vrow = k >= 0 ? S.vvv + static_cast<size_t>(k) * (S.mxgrid + 1) : S.erc;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if ((h == 0 && !OLD) || (h == 1 && !NEW)) continue;
      double dx = (h ? P.nx[s] : P.ox[s]) - x;
      double dy = (h ? P.ny[s] : P.oy[s]) - y;
      double dz = (h ? P.nz[s] : P.oz[s]) - z;
      min_image(S, dx, dy, dz);
      const double rsq = dx * dx + dy * dy + dz * dz;
      const bool in = S.rcut2 > rsq;
      const double r = sqrt(rsq);
      const double vc = interp3(S.erc, r, S.rdrewd, imax);
      const double vv = interp3(vrow, r, S.rdlrpot, imax);
      const double a2 = (in && qon) ? vc * chgprd : 0.0;
      const double av = (in && k >= 0) ? vv : 0.0;
      if (h == 0 && disp) {
        e2 -= a2;
        ev -= av;
      } else {
        e2 += a2;
        ev += av;
      }
    }
  }
}

// Real space over every atom (ewald2 + srfrce, ewald.f:632-694,
// vdw_terms.f:398-498) for a molecule of NS sites; pass 1 of a displacement
// with the minus sign.
// Real space for molecules of more than four sites: one position at a time
// (unrolling them all would not fit the registers).
__device__ __forceinline__ void trial_real_any(const TrialParams& P, int rb, bool use_old,
                                               bool use_new, double& e2, double& ev) {
  int ntot = P.S.nfw;
  for (int k = 0; k < P.ngt; ++k) ntot += P.g[k].natoms;
  for (int a = rb * kTrialBS + static_cast<int>(threadIdx.x); a < ntot;
       a += P.nb_real * kTrialBS) {
    double x, y, z, q;
    int t, gt, gi;
    load_atom(P.S, P.g, P.ngt, a, x, y, z, q, t, gt, gi);
    if (gt == P.own_t && gi >= P.own_begin && gi < P.own_begin + P.ns &&
        P.kind != FMC_INSERT)
      continue;
    for (int s = 0; s < P.ns; ++s) {
      const double ce = P.q[s] / P.S.epsq * kR4pie0;
      if (use_old) {
        double a2 = 0.0, av = 0.0;
        pair_fp64(P.S, ce, P.t[s], P.ox[s] - x, P.oy[s] - y, P.oz[s] - z, q, t, a2, av);
        if (P.kind == FMC_DISPLACE) { e2 -= a2; ev -= av; }
        else { e2 += a2; ev += av; }
      }
      if (use_new) {
        double a2 = 0.0, av = 0.0;
        pair_fp64(P.S, ce, P.t[s], P.nx[s] - x, P.ny[s] - y, P.nz[s] - z, q, t, a2, av);
        e2 += a2;
        ev += av;
      }
    }
  }
}

template <int NS, bool OLD, bool NEW>
__device__ __forceinline__ void trial_real(const TrialParams& P, int rb, double& e2,
                                           double& ev) {
  int ntot = P.S.nfw;
  for (int k = 0; k < P.ngt; ++k) ntot += P.g[k].natoms;
  for (int a = rb * kTrialBS + static_cast<int>(threadIdx.x); a < ntot;
       a += P.nb_real * kTrialBS) {
    double x, y, z, q;
    int t, gt, gi;
    load_atom(P.S, P.g, P.ngt, a, x, y, z, q, t, gt, gi);
    if (gt == P.own_t && gi >= P.own_begin && gi < P.own_begin + NS && P.kind != FMC_INSERT)
      continue;
    trial_atom<NS, OLD, NEW>(P, 0, x, y, z, q, t, e2, ev);
  }
}

// NS 1-4: the molecule's size, with the unrolled real-space form; NS 0:
// any size, one position at a time. One kernel per form keeps each one's
// registers to its own needs.
template <int NS>
__global__ void __launch_bounds__(kTrialBS) trial_kernel(TrialParams P) {
  __shared__ double2 tab[2][kMaxSites][3][kTabMax];
  __shared__ double2 red[kTrialBS / kWave];
  const int tid = threadIdx.x;
  const bool use_old = (P.kind != FMC_INSERT);
  const bool use_new = (P.kind != FMC_DELETE);
  double e1 = 0.0, e2 = 0.0, ev = 0.0;
  if (static_cast<int>(blockIdx.x) < P.nb_recip) {
    // ---- reciprocal part (ewald1_guest, ewald.f:97-331) ----
    // the first k of this thread is loaded before the phase tables are
    // built, so its loads overlap the tables' arithmetic
    const int k0 = blockIdx.x * kTrialBS + tid;
    int v0 = 0;
    double ak0 = 0.0, sre0 = 0.0, sim0 = 0.0;
    if (k0 < P.nk) {
      v0 = P.kv[k0];
      ak0 = P.ak[k0];
      sre0 = P.sre[k0];
      sim0 = P.sim[k0];
    }
    const int ntab = 2 * P.ns * 3;
    if (tid < ntab) {
      const int set = tid / (P.ns * 3);
      const int s = (tid / 3) % P.ns;
      const int d = tid % 3;
      const double x = set ? P.nx[s] : P.ox[s];
      const double y = set ? P.ny[s] : P.oy[s];
      const double z = set ? P.nz[s] : P.oz[s];
      const double f = P.S.rcell[d] * x + P.S.rcell[3 + d] * y + P.S.rcell[6 + d] * z;
      phase_table(f, P.kmax[d], &tab[set][s][d][0]);
    }
    __syncthreads();
    for (int k = k0; k < P.nk; k += P.nb_recip * kTrialBS) {
      const bool first = k == k0;
      const int v = first ? v0 : P.kv[k];
      const int l = unpack_l(v), m = unpack_m(v), n = unpack_n(v);
      const int am = m < 0 ? -m : m, an = n < 0 ? -n : n;
      double2 so = make_double2(0.0, 0.0), sn = make_double2(0.0, 0.0);
      if (use_old)
        for (int s = 0; s < P.ns; ++s) {
          const double2 t = structure_term(tab[0][s][0][l], tab[0][s][1][am], m,
                                           tab[0][s][2][an], n, P.q[s]);
          so.x = so.x + t.x;
          so.y = so.y + t.y;
        }
      if (use_new)
        for (int s = 0; s < P.ns; ++s) {
          const double2 t = structure_term(tab[1][s][0][l], tab[1][s][1][am], m,
                                           tab[1][s][2][an], n, P.q[s]);
          sn.x = sn.x + t.x;
          sn.y = sn.y + t.y;
        }
      const double cold = first ? sre0 : P.sre[k], sold = first ? sim0 : P.sim[k];
      double c, s;
      if (P.kind == FMC_INSERT) {
        c = sn.x + cold;
        s = sn.y + sold;
      } else if (P.kind == FMC_DELETE) {
        c = cold - so.x;
        s = sold - so.y;
      } else {
        c = (cold - so.x) + sn.x;
        s = (sold - so.y) + sn.y;
      }
      P.nre[k] = c;
      P.nim[k] = s;
      e1 += (first ? ak0 : P.ak[k]) * (c * c + s * s - cold * cold - sold * sold);
    }
  } else {
    // ---- real space: the move's halves (old sites for a delete, new for
    //      an insert, both for a displacement) select the unrolled form ----
    const int rb = blockIdx.x - P.nb_recip;
    if constexpr (NS == 0) {
      trial_real_any(P, rb, use_old, use_new, e2, ev);
    } else if (use_old && use_new) {
      trial_real<NS, true, true>(P, rb, e2, ev);
    } else if (use_old) {
      trial_real<NS, true, false>(P, rb, e2, ev);
    } else {
      trial_real<NS, false, true>(P, rb, e2, ev);
    }
  }
  // One reduction pass: a reciprocal block carries e1 only, a real-space
  // block e2 and ev (the other sums are zero). Waves are added in order.
  const bool rec = static_cast<int>(blockIdx.x) < P.nb_recip;
  const double ra = wave_sum(rec ? e1 : e2);
  const double rb = rec ? 0.0 : wave_sum(ev);
  if (tid % kWave == 0) red[tid / kWave] = make_double2(ra, rb);
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kTrialBS / kWave; ++w) {
      a += red[w].x;
      b += red[w].y;
    }
    e1 = rec ? a : 0.0;
    e2 = rec ? 0.0 : a;
    ev = rec ? 0.0 : b;
    // the block's sums reach host memory before its sequence word (release
    // at system scope); the host adds the blocks' sums in block order
    double* o = P.part + 3 * blockIdx.x;
    o[0] = e1;
    o[1] = e2;
    o[2] = ev;
    __hip_atomic_store(P.word + blockIdx.x, P.seq, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---- full structure factor (ewald1, ewald.f:368-630) ----
struct RecipFullParams {
  DevSys S;
  int ngt;
  GuestSeg g[kMaxGuestTypes];
  int natoms;   // atoms included (framework only or all)
  int nk;
  const int* kv;
  const double* ak;
  int chunk;
  double2* part;   // [nchunk][nk]
};

// A molecule's sites written into the resident guest arrays from kernel
// arguments (an accepted insert or displacement), so a commit enqueues no
// host-to-device copy; stream order puts it before the next trial.
struct SitesArgs {
  double* gx;
  double* gy;
  double* gz;
  int lo, ns;
  double x[kMaxSites], y[kMaxSites], z[kMaxSites];
};

__global__ void put_sites_kernel(SitesArgs A) {
  const int s = threadIdx.x;
  if (s < A.ns) {
    A.gx[A.lo + s] = A.x[s];
    A.gy[A.lo + s] = A.y[s];
    A.gz[A.lo + s] = A.z[s];
  }
}

// An accepted delete: the successors of the deleted molecule move down by
// ns sites (gcmc.f:1127-1131) in the resident arrays, [lo, hi) <- [lo + ns,
// hi + ns), by one work-group: each pass reads its span before a barrier
// and writes it after, so no read sees a slot already overwritten.
__global__ void __launch_bounds__(kBS) shift_sites_kernel(double* gx, double* gy, double* gz,
                                                          int lo, int hi, int ns) {
  for (int base = lo; base < hi; base += kBS) {
    const int i = base + threadIdx.x;
    double x = 0.0, y = 0.0, z = 0.0;
    if (i < hi) {
      x = gx[i + ns];
      y = gy[i + ns];
      z = gz[i + ns];
    }
    __syncthreads();
    if (i < hi) {
      gx[i] = x;
      gy[i] = y;
      gz[i] = z;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBS) recip_partial_kernel(RecipFullParams P) {
  __shared__ double4 at[kBS];
  const int c = blockIdx.y;
  const int a0 = c * P.chunk;
  const int na = min(P.chunk, P.natoms - a0);
  double2 acc = make_double2(0.0, 0.0);
  const int k = blockIdx.x * kBS + threadIdx.x;
  int l = 0, m = 0, n = 0;
  if (k < P.nk) {
    const int v = P.kv[k];
    l = unpack_l(v); m = unpack_m(v); n = unpack_n(v);
  }
  for (int base = 0; base < na; base += kBS) {
    __syncthreads();
    const int i = base + threadIdx.x;
    if (i < na) {
      double x, y, z, q;
      int t, gt, gi;
      load_atom(P.S, P.g, P.ngt, a0 + i, x, y, z, q, t, gt, gi);
      at[threadIdx.x] = make_double4(
          P.S.rcell[0] * x + P.S.rcell[3] * y + P.S.rcell[6] * z,
          P.S.rcell[1] * x + P.S.rcell[4] * y + P.S.rcell[7] * z,
          P.S.rcell[2] * x + P.S.rcell[5] * y + P.S.rcell[8] * z, q);
    }
    __syncthreads();
    const int cnt = min(kBS, na - base);
    if (k < P.nk)
      for (int j = 0; j < cnt; ++j) {
        const double4 a = at[j];
        double ph = l * a.x + m * a.y + n * a.z;
        ph -= rint(ph);
        double sv, cv;
        sincospi(2.0 * ph, &sv, &cv);
        acc.x += a.w * cv;
        acc.y += a.w * sv;
      }
  }
  if (k < P.nk) P.part[static_cast<size_t>(c) * P.nk + k] = acc;
}

__global__ void __launch_bounds__(kBS) recip_reduce_kernel(
    const double2* part, int nchunk, int nk, const double* ak, double* sre,
    double* sim, double* epart) {
  __shared__ double red[kBS / kWave];
  const int k = blockIdx.x * kBS + threadIdx.x;
  double e = 0.0;
  if (k < nk) {
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < nchunk; ++c) {
      const double2 p = part[static_cast<size_t>(c) * nk + k];
      s.x += p.x;
      s.y += p.y;
    }
    if (sre) {
      sre[k] = s.x;
      sim[k] = s.y;
    }
    e = ak[k] * (s.x * s.x + s.y * s.y);
  }
  e = block_sum(e, red);
  if (threadIdx.x == 0) epart[blockIdx.x] = e;
}

// Sum n partials in order (single block).
__global__ void __launch_bounds__(kBS) sum_kernel(const double* part, int n,
                                                  double* out) {
  __shared__ double red[kBS / kWave];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += kBS) a += part[i];
  a = block_sum(a, red);
  if (threadIdx.x == 0) *out = a;
}

// ---- single_point real space: every guest atom against the framework and
//      the later guest atoms (each pair once, gcmc.f:2644-2680) ----
struct PairAllParams {
  DevSys S;
  int ngt;
  GuestSeg g[kMaxGuestTypes];
  int nguest;   // guest atoms
  double* part; // [2][gridDim.x * gridDim.y]
};

__global__ void __launch_bounds__(kBS) pair_all_kernel(PairAllParams P) {
  __shared__ double red[kBS / kWave];
  const int i = blockIdx.y;  // guest atom (global guest index)
  double xi, yi, zi, qi;
  int ti, gti, gii;
  load_atom(P.S, P.g, P.ngt, P.S.nfw + i, xi, yi, zi, qi, ti, gti, gii);
  const int nsi = P.g[gti].ns;
  const int moli = gii / nsi;
  const double ce = qi / P.S.epsq * kR4pie0;
  const int ntot = P.S.nfw + P.nguest;
  double e2 = 0.0, ev = 0.0;
  for (int a = blockIdx.x * kBS + threadIdx.x; a < ntot; a += gridDim.x * kBS) {
    // framework atoms, or guest atoms after i in condensed order
    if (a >= P.S.nfw && a - P.S.nfw <= i) continue;
    double x, y, z, q;
    int t, gt, gi;
    load_atom(P.S, P.g, P.ngt, a, x, y, z, q, t, gt, gi);
    if (gt == gti && gi / P.g[gt].ns == moli) continue;
    pair_fp64(P.S, ce, ti, xi - x, yi - y, zi - z, q, t, e2, ev);
  }
  e2 = block_sum(e2, red);
  ev = block_sum(ev, red);
  if (threadIdx.x == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    const int nbt = gridDim.x * gridDim.y;
    P.part[b] = e2;
    P.part[nbt + b] = ev;
  }
}

}  // namespace fmc

using namespace fmc;

#define HIPCHECK(x)                                                     \
  do {                                                                  \
    hipError_t e_ = (x);                                                \
    if (e_ != hipSuccess) {                                             \
      std::fprintf(stderr, "fastmc_amd: HIP error %s at %s:%d\n",       \
                   hipGetErrorString(e_), __FILE__, __LINE__);          \
      return FMC_ERR_HIP;                                               \
    }                                                                   \
  } while (0)

namespace {

template <typename T>
int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  HIPCHECK(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
  return FMC_OK;
}

template <typename T>
int upload(T* d, const T* h, size_t n) {
  if (n) HIPCHECK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return FMC_OK;
}

void fill_devsys(fmc_ctx* c, DevSys* S) {
  const SystemDesc& s = c->sys;
  S->imcon = s.imcon;
  std::memcpy(S->cell, s.cell, sizeof(S->cell));
  std::memcpy(S->rcell, s.rcell, sizeof(S->rcell));
  S->inv[0] = 1.0 / s.cell[0];
  S->inv[1] = 1.0 / s.cell[4];
  S->inv[2] = 1.0 / s.cell[8];
  S->rcut2 = s.rcut * s.rcut;
  S->rdrewd = 1.0 / s.drewd;
  S->rdlrpot = 1.0 / s.dlrpot;
  S->mxgrid = s.mxgrid;
  S->ntpatm = s.ntpatm;
  S->epsq = s.epsq;
  S->erc = c->d_erc;
  S->vvv = c->d_vvv;
  S->pairtab = c->d_pairtab;
  S->nfw = static_cast<int>(s.fx.size());
  S->fx = c->d_fx;
  S->fy = c->d_fy;
  S->fz = c->d_fz;
  S->fq = c->d_fq;
  S->ftype = c->d_ftype;
}

void fill_segs(fmc_ctx* c, GuestSeg* g, int* ngt) {
  *ngt = static_cast<int>(c->sys.guests.size());
  for (int t = 0; t < *ngt; ++t) {
    const GuestTemplate& tp = c->sys.guests[t];
    g[t].x = c->d_gx[t];
    g[t].y = c->d_gy[t];
    g[t].z = c->d_gz[t];
    g[t].ns = tp.ns;
    g[t].natoms = c->nmol[t] * tp.ns;
    for (int s = 0; s < kMaxSites; ++s) {
      g[t].q[s] = s < tp.ns ? tp.q[s] : 0.0;
      g[t].t[s] = s < tp.ns ? tp.type[s] - 1 : 0;
    }
  }
}

// Device buffers one full-energy evaluation works in: the recip partial
// sums [nchunk][nk], the energy partials [nk/256], two output slots and the
// real-space partial sums [2][blocks]; the context's own buffers, or the
// library's static scratch array (fmc_energy_full_noalloc).
struct FullWork {
  double2* rpart;
  size_t rpart_cap;   // double2 entries
  double* epart;
  double* out;
  double* pair;       // nullptr: allocated per call (fmc_energy_full)
  size_t pair_cap;    // doubles
};

FullWork ctx_work(fmc_ctx* c) {
  FullWork w;
  w.rpart = c->d_rpart;
  w.rpart_cap = static_cast<size_t>(c->recip_nchunk_cap) * c->nk;
  w.epart = c->d_epart;
  w.out = c->d_out;
  w.pair = nullptr;
  w.pair_cap = 0;
  return w;
}

int recip_full_impl(fmc_ctx* c, const FullWork& w, bool all_atoms, bool set_running,
                    double* energy) {
  RecipFullParams P;
  fill_devsys(c, &P.S);
  fill_segs(c, P.g, &P.ngt);
  int nguest = 0;
  for (int t = 0; t < P.ngt; ++t) nguest += P.g[t].natoms;
  P.natoms = P.S.nfw + (all_atoms ? nguest : 0);
  P.nk = c->nk;
  P.kv = c->d_kv;
  P.ak = c->d_ak;
  P.chunk = c->recip_chunk;
  const int nchunk = std::max(1, (P.natoms + P.chunk - 1) / P.chunk);
  if (static_cast<size_t>(nchunk) * c->nk > w.rpart_cap) return FMC_ERR_CAPACITY;
  P.part = w.rpart;
  const int kb = (c->nk + kBS - 1) / kBS;
  hipLaunchKernelGGL(recip_partial_kernel, dim3(kb, nchunk), dim3(kBS), 0,
                     c->stream, P);
  HIPCHECK(hipGetLastError());
  double* sre = set_running ? c->d_sre[c->cur] : nullptr;
  double* sim = set_running ? c->d_sim[c->cur] : nullptr;
  hipLaunchKernelGGL(recip_reduce_kernel, dim3(kb), dim3(kBS), 0, c->stream,
                     w.rpart, nchunk, c->nk, c->d_ak, sre, sim, w.epart);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kBS), 0, c->stream, w.epart,
                     kb, w.out);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(c->h_out, w.out, sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  double sq = 0.0;
  for (double q : c->sys.fq) sq += q * q;
  if (all_atoms)
    for (size_t t = 0; t < c->sys.guests.size(); ++t)
      for (int s = 0; s < c->sys.guests[t].ns; ++s)
        sq += c->nmol[t] * c->sys.guests[t].q[s] * c->sys.guests[t].q[s];
  const SystemDesc& s = c->sys;
  const double engsic = -kR4pie0 / s.epsq * s.alpha * sq / kSqrpi;
  const double rvolm = 2.0 * kPi / s.volm;
  *energy = 2.0 * rvolm * kR4pie0 * c->h_out[0] / s.epsq + engsic;
  return FMC_OK;
}

// single_point (gcmc.f:2518-2710) on the work buffers w: framework and
// all-atom reciprocal sums (the latter resets the running S(k), as the
// ewald1 call at gcmc.f:522 does), every guest atom's real-space pairs,
// the ewald3 exclusions and the tail.
int energy_full_impl(fmc_ctx* c, const FullWork& w, double* out) {
  int r;
  double efw = 0.0, eall = 0.0;
  if ((r = recip_full_impl(c, w, false, false, &efw))) return r;
  if ((r = recip_full_impl(c, w, true, true, &eall))) return r;
  PairAllParams P;
  fill_devsys(c, &P.S);
  fill_segs(c, P.g, &P.ngt);
  P.nguest = 0;
  for (int t = 0; t < P.ngt; ++t) P.nguest += P.g[t].natoms;
  double e2 = 0.0, ev = 0.0, e3 = 0.0;
  if (P.nguest > 0) {
    const int ntot = P.S.nfw + P.nguest;
    const int gx = std::min(8, (ntot + kBS - 1) / kBS);
    const size_t nbt = static_cast<size_t>(gx) * P.nguest;
    double* part = w.pair;
    if (part) {
      if (2 * nbt > w.pair_cap) return FMC_ERR_CAPACITY;
    } else {
      HIPCHECK(hipMalloc(&part, 2 * nbt * sizeof(double)));
    }
    P.part = part;
    hipLaunchKernelGGL(pair_all_kernel, dim3(gx, P.nguest), dim3(kBS), 0,
                       c->stream, P);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kBS), 0, c->stream, part,
                       static_cast<int>(nbt), w.out);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kBS), 0, c->stream, part + nbt,
                       static_cast<int>(nbt), w.out + 1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(c->h_out, w.out, 2 * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (!w.pair) HIPCHECK(hipFree(part));
    e2 = c->h_out[0];
    ev = c->h_out[1];
    for (size_t t = 0; t < c->sys.guests.size(); ++t) {
      const int ns = c->sys.guests[t].ns;
      for (int m = 0; m < c->nmol[t]; ++m)
        e3 += ewald3_molecule(c->sys, c->sys.guests[t], &c->hx[t][m * ns],
                              &c->hy[t][m * ns], &c->hz[t][m * ns]);
    }
  }
  std::vector<int> nt, nf;
  c->sys.counts(c->nmol, &nt, &nf);
  out[0] = efw;
  out[1] = eall - efw;
  out[2] = e2;
  out[3] = ev;
  out[4] = e3;
  out[5] = c->sys.tail(nt, nf);
  return FMC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI: set-up
// ---------------------------------------------------------------------------

extern "C" int fmc_ctx_create(fmc_ctx** out) {
  if (!out) return FMC_ERR_ARG;
  *out = new fmc_ctx();
  return FMC_OK;
}

extern "C" int fmc_ctx_destroy(fmc_ctx* c) {
  if (!c) return FMC_ERR_ARG;
  c->release();
  delete c;
  return FMC_OK;
}

extern "C" int fmc_set_cell(fmc_ctx* c, int imcon, const double* cell) {
  if (!c || !cell) return FMC_ERR_ARG;
  if (imcon < 1 || imcon > 3) return FMC_ERR_UNSUPPORTED;
  c->sys.imcon = imcon;
  std::memcpy(c->sys.cell, cell, 9 * sizeof(double));
  return FMC_OK;
}

extern "C" int fmc_set_ewald(fmc_ctx* c, double alpha, int k1, int k2, int k3,
                             double epsq) {
  if (!c || alpha <= 0 || k1 < 0 || k2 < 0 || k3 < 0) return FMC_ERR_ARG;
  if (k1 >= kTabMax || k2 >= kTabMax || k3 >= kTabMax) return FMC_ERR_UNSUPPORTED;
  c->sys.alpha = alpha;
  c->sys.kmax[0] = k1;
  c->sys.kmax[1] = k2;
  c->sys.kmax[2] = k3;
  c->sys.epsq = epsq;
  c->sys.have_ewald = true;
  return FMC_OK;
}

extern "C" int fmc_set_ewald_precision(fmc_ctx* c, double eps, double epsq) {
  if (!c || c->sys.imcon == 0 || c->sys.rcut <= 0) return FMC_ERR_STATE;
  double alpha;
  int km[3];
  ewald_precision(eps, c->sys.rcut, c->sys.imcon, c->sys.cell, &alpha, km);
  return fmc_set_ewald(c, alpha, km[0], km[1], km[2], epsq);
}

extern "C" int fmc_set_cutoffs(fmc_ctx* c, double rcut, double delr,
                               double rvdw) {
  if (!c || rcut <= 0) return FMC_ERR_ARG;
  c->sys.rcut = rcut;
  c->sys.delr = delr;
  c->sys.rvdw = rvdw;
  return FMC_OK;
}

extern "C" int fmc_set_vdw(fmc_ctx* c, int ntpatm, int ntpvdw,
                           const int* a, const int* b, const int* keys,
                           const double* prm) {
  if (!c || ntpatm <= 0 || ntpvdw < 0) return FMC_ERR_ARG;
  SystemDesc& s = c->sys;
  s.ntpatm = ntpatm;
  s.ntpvdw = ntpvdw;
  s.vdw_a.assign(a, a + ntpvdw);
  s.vdw_b.assign(b, b + ntpvdw);
  s.vdw_key.assign(keys, keys + ntpvdw);
  s.vdw_prm.assign(prm, prm + 5 * ntpvdw);
  for (int t = 0; t < ntpvdw; ++t)
    if (keys[t] < 0 || keys[t] > 9) return FMC_ERR_UNSUPPORTED;
  return FMC_OK;
}

extern "C" int fmc_add_framework(fmc_ctx* c, int n, const double* x,
                                 const double* y, const double* z,
                                 const double* q, const int* type,
                                 const int* frozen) {
  if (!c || n < 0 || (n && (!x || !y || !z || !q || !type))) return FMC_ERR_ARG;
  if (c->sys.finalized) return FMC_ERR_STATE;
  SystemDesc& s = c->sys;
  for (int i = 0; i < n; ++i) {
    s.fx.push_back(x[i]);
    s.fy.push_back(y[i]);
    s.fz.push_back(z[i]);
    s.fq.push_back(q[i]);
    s.ftype.push_back(type[i]);
    s.ffrz.push_back(frozen ? frozen[i] : 1);
  }
  return FMC_OK;
}

extern "C" int fmc_add_guest_type(fmc_ctx* c, int ns, const double* x,
                                  const double* y, const double* z,
                                  const double* q, const int* type,
                                  const double* mass, int* gt) {
  if (!c || ns <= 0 || ns > kMaxSites || !x || !y || !z || !q || !type)
    return FMC_ERR_ARG;
  if (c->sys.finalized) return FMC_ERR_STATE;
  if (static_cast<int>(c->sys.guests.size()) >= kMaxGuestTypes)
    return FMC_ERR_CAPACITY;
  GuestTemplate g;
  std::memset(&g, 0, sizeof(g));
  g.ns = ns;
  for (int i = 0; i < ns; ++i) {
    g.x[i] = x[i];
    g.y[i] = y[i];
    g.z[i] = z[i];
    g.q[i] = q[i];
    g.type[i] = type[i];
    g.mass[i] = mass ? mass[i] : 1.0;
  }
  c->sys.guests.push_back(g);
  c->sys.gx0.emplace_back();
  c->sys.gy0.emplace_back();
  c->sys.gz0.emplace_back();
  if (gt) *gt = static_cast<int>(c->sys.guests.size()) - 1;
  return FMC_OK;
}

extern "C" int fmc_add_guest_molecule(fmc_ctx* c, int gt, const double* x,
                                      const double* y, const double* z) {
  if (!c || gt < 0 || gt >= static_cast<int>(c->sys.guests.size()))
    return FMC_ERR_ARG;
  if (c->sys.finalized) return FMC_ERR_STATE;
  for (int i = 0; i < c->sys.guests[gt].ns; ++i) {
    c->sys.gx0[gt].push_back(x[i]);
    c->sys.gy0[gt].push_back(y[i]);
    c->sys.gz0[gt].push_back(z[i]);
  }
  return FMC_OK;
}

extern "C" int fmc_set_capacity(fmc_ctx* c, int max_guest_atoms) {
  if (!c || max_guest_atoms <= 0) return FMC_ERR_ARG;
  if (c->sys.finalized) return FMC_ERR_STATE;
  c->sys.max_guest_atoms = max_guest_atoms;
  return FMC_OK;
}

// Host side of finalize: k-vector count, capacity and the host mirror of
// the guests read from CONFIG/REVIVE.  Nothing touches the device, so a
// context that only feeds batched chains never allocates device memory.
int fmc_ctx::host_state() {
  const SystemDesc& s = sys;
  const int ng = static_cast<int>(s.guests.size());
  nk = static_cast<int>(s.kvecs.size());
  cap_atoms = s.max_guest_atoms;
  nmol.assign(ng, 0);
  int guest_atoms = 0;
  for (int t = 0; t < ng; ++t) {
    const int ns = s.guests[t].ns;
    hx[t].assign(cap_atoms, 0.0);
    hy[t].assign(cap_atoms, 0.0);
    hz[t].assign(cap_atoms, 0.0);
    const int n0 = static_cast<int>(s.gx0[t].size());
    if (n0 > cap_atoms) return FMC_ERR_CAPACITY;
    for (int i = 0; i < n0; ++i) {
      hx[t][i] = s.gx0[t][i];
      hy[t][i] = s.gy0[t][i];
      hz[t][i] = s.gz0[t][i];
    }
    nmol[t] = n0 / ns;
    guest_atoms += n0;
  }
  if (guest_atoms > cap_atoms) return FMC_ERR_CAPACITY;
  return FMC_OK;
}

// Device side, on the first single-chain call: tables, framework, guests,
// S(k) buffers and the all-atom S(k) of gcmc.f:522.
int fmc_ctx::upload_state() {
  const SystemDesc& s = sys;
  const int ng = static_cast<int>(s.guests.size());
  const size_t nfw = s.fx.size();
  int r;
  if ((r = dalloc(&d_erc, s.erc.size())) || (r = upload(d_erc, s.erc.data(), s.erc.size())))
    return r;
  if ((r = dalloc(&d_vvv, s.vvv.size())) || (r = upload(d_vvv, s.vvv.data(), s.vvv.size())))
    return r;
  if ((r = dalloc(&d_pairtab, s.pairtab.size())) ||
      (r = upload(d_pairtab, s.pairtab.data(), s.pairtab.size())))
    return r;
  std::vector<int> ft0(nfw);
  for (size_t i = 0; i < nfw; ++i) ft0[i] = s.ftype[i] - 1;
  if ((r = dalloc(&d_fx, nfw)) || (r = upload(d_fx, s.fx.data(), nfw))) return r;
  if ((r = dalloc(&d_fy, nfw)) || (r = upload(d_fy, s.fy.data(), nfw))) return r;
  if ((r = dalloc(&d_fz, nfw)) || (r = upload(d_fz, s.fz.data(), nfw))) return r;
  if ((r = dalloc(&d_fq, nfw)) || (r = upload(d_fq, s.fq.data(), nfw))) return r;
  if ((r = dalloc(&d_ftype, nfw)) || (r = upload(d_ftype, ft0.data(), nfw))) return r;
  std::vector<int> kv(nk);
  std::vector<double> ak(nk);
  for (int k = 0; k < nk; ++k) {
    kv[k] = (s.kvecs[k].l & 0xff) | ((s.kvecs[k].m & 0xff) << 8) |
            ((s.kvecs[k].n & 0xff) << 16);
    ak[k] = s.kvecs[k].ak;
  }
  if ((r = dalloc(&d_kv, nk)) || (r = upload(d_kv, kv.data(), nk))) return r;
  if ((r = dalloc(&d_ak, nk)) || (r = upload(d_ak, ak.data(), nk))) return r;
  for (int b = 0; b < 2; ++b) {
    if ((r = dalloc(&d_sre[b], nk)) || (r = dalloc(&d_sim[b], nk))) return r;
  }
  for (int t = 0; t < ng; ++t) {
    const int n0 = nmol[t] * s.guests[t].ns;
    if ((r = dalloc(&d_gx[t], cap_atoms)) || (r = dalloc(&d_gy[t], cap_atoms)) ||
        (r = dalloc(&d_gz[t], cap_atoms)))
      return r;
    if ((r = upload(d_gx[t], hx[t].data(), n0)) || (r = upload(d_gy[t], hy[t].data(), n0)) ||
        (r = upload(d_gz[t], hz[t].data(), n0)))
      return r;
  }
  recip_chunk = 1024;
  recip_nchunk_cap = static_cast<int>((nfw + cap_atoms) / recip_chunk) + 2;
  if ((r = dalloc(&d_rpart, static_cast<size_t>(recip_nchunk_cap) * nk))) return r;
  if ((r = dalloc(&d_epart, static_cast<size_t>(nk / 64 + 64)))) return r;
  if ((r = dalloc(&d_out, 8))) return r;
  HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&h_out), 8 * sizeof(double),
                         hipHostMallocCoherent));
  HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&h_part), 3 * kMaxBlocks * sizeof(double),
                         hipHostMallocCoherent));
  HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&h_word),
                         kMaxBlocks * sizeof(unsigned long long), hipHostMallocCoherent));
  std::memset(h_word, 0, kMaxBlocks * sizeof(unsigned long long));
  HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  return FMC_OK;
}

void fmc_ctx::release() {
  // a trial returns before its kernel retires and a commit does not wait:
  // drain the stream before any buffer it uses is freed
  if (stream) (void)hipStreamSynchronize(stream);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(d_erc); f(d_vvv); f(d_pairtab); f(d_fx); f(d_fy); f(d_fz); f(d_fq);
  f(d_ftype); f(d_kv); f(d_ak);
  for (int b = 0; b < 2; ++b) { f(d_sre[b]); f(d_sim[b]); }
  for (int t = 0; t < kMaxGuestTypes; ++t) { f(d_gx[t]); f(d_gy[t]); f(d_gz[t]); }
  f(d_rpart); f(d_epart); f(d_out);
  if (h_out) (void)hipHostFree(h_out);
  if (h_part) (void)hipHostFree(h_part);
  if (h_word) (void)hipHostFree(h_word);
  if (stream) (void)hipStreamDestroy(stream);
  d_erc = d_vvv = d_fx = d_fy = d_fz = d_fq = d_ak = nullptr;
  d_pairtab = d_ftype = d_kv = nullptr;
  h_out = nullptr;
  h_part = nullptr;
  h_word = nullptr;
  stream = nullptr;
}

extern "C" int fmc_ctx_finalize(fmc_ctx* c) {
  if (!c) return FMC_ERR_ARG;
  if (c->sys.finalized) return FMC_ERR_STATE;
  std::string err;
  int r = c->sys.finalize(&err);
  if (r) {
    std::fprintf(stderr, "fastmc_amd: %s\n", err.c_str());
    return r;
  }
  return c->host_state();
}

// device state on the first call that needs it
static int ensure_device(fmc_ctx* c) {
  if (!c->sys.finalized) return FMC_ERR_STATE;
  if (c->dev_ready) return FMC_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FMC_ERR_NODEVICE;
  int r;
  if ((r = c->upload_state())) return r;
  double e;
  if ((r = recip_full_impl(c, ctx_work(c), true, true, &e))) return r;
  c->dev_ready = true;
  return FMC_OK;
}

extern "C" int fmc_get_info(const fmc_ctx* c, int* nk, double* volm,
                            int* mxgrid, double* drewd, double* dlrpot) {
  if (!c || !c->sys.finalized) return FMC_ERR_STATE;
  if (nk) *nk = c->nk;
  if (volm) *volm = c->sys.volm;
  if (mxgrid) *mxgrid = c->sys.mxgrid;
  if (drewd) *drewd = c->sys.drewd;
  if (dlrpot) *dlrpot = c->sys.dlrpot;
  return FMC_OK;
}

extern "C" int fmc_guest_ewald3(const fmc_ctx* c, int gt, double* e3) {
  if (!c || !c->sys.finalized || gt < 0 ||
      gt >= static_cast<int>(c->sys.guests.size()) || !e3)
    return FMC_ERR_ARG;
  *e3 = c->sys.guests[gt].e3;
  return FMC_OK;
}

extern "C" int fmc_guest_count(const fmc_ctx* c, int gt, int* n) {
  if (!c || !c->sys.finalized || gt < 0 ||
      gt >= static_cast<int>(c->sys.guests.size()) || !n)
    return FMC_ERR_ARG;
  *n = c->nmol[gt];
  return FMC_OK;
}

extern "C" int fmc_get_molecule(const fmc_ctx* c, int gt, int m, double* x,
                                double* y, double* z) {
  if (!c || !c->sys.finalized || gt < 0 ||
      gt >= static_cast<int>(c->sys.guests.size()) || m < 0 || m >= c->nmol[gt])
    return FMC_ERR_ARG;
  const int ns = c->sys.guests[gt].ns;
  for (int s = 0; s < ns; ++s) {
    x[s] = c->hx[gt][m * ns + s];
    y[s] = c->hy[gt][m * ns + s];
    z[s] = c->hz[gt][m * ns + s];
  }
  return FMC_OK;
}

extern "C" int fmc_get_table(const fmc_ctx* c, int which, int t, double* out) {
  if (!c || !c->sys.finalized || !out) return FMC_ERR_ARG;
  const int W = c->sys.mxgrid + 1;
  if (which == 0) {
    std::memcpy(out, c->sys.erc.data(), W * sizeof(double));
    return FMC_OK;
  }
  if (t < 0 || t >= c->sys.ntable) return FMC_ERR_ARG;
  std::memcpy(out, c->sys.vvv.data() + static_cast<size_t>(t) * W, W * sizeof(double));
  return FMC_OK;
}

extern "C" int fmc_get_sk(const fmc_ctx* c, double* re, double* im) {
  if (!c || !c->sys.finalized) return FMC_ERR_STATE;
  // building the device state on first use changes nothing observable
  int r;
  if ((r = ensure_device(const_cast<fmc_ctx*>(c)))) return r;
  // a trial returns before its kernel retires and a commit does not wait:
  // drain the stream before copying on another one
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(re, c->d_sre[c->cur], c->nk * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(im, c->d_sim[c->cur], c->nk * sizeof(double), hipMemcpyDeviceToHost));
  return FMC_OK;
}

// ---------------------------------------------------------------------------
// C ABI: hot path
// ---------------------------------------------------------------------------

extern "C" int fmc_recip_full(fmc_ctx* c, int all_atoms, double* energy) {
  if (!c || !c->sys.finalized || !energy) return FMC_ERR_STATE;
  int r;
  if ((r = ensure_device(c))) return r;
  c->pending = 0;
  return recip_full_impl(c, ctx_work(c), all_atoms != 0, all_atoms != 0, energy);
}

// Wait for a trial's sums: spin on the sequence word each work-group
// publishes after its partial sums, which land before the kernel retires,
// then add the partials in block order. The stream is queried every few
// thousand reads, so a kernel that ends (or fails) without publishing is
// reported instead of waited on.
static int wait_trial(fmc_ctx* c, unsigned long long seq, int nb, double v[3]) {
  for (int b = 0; b < nb; ++b) {
    unsigned long long* word = c->h_word + b;
    for (unsigned spin = 1; __atomic_load_n(word, __ATOMIC_ACQUIRE) != seq; ++spin) {
      __builtin_ia32_pause();
      if ((spin & 4095) == 0) {
        const hipError_t q = hipStreamQuery(c->stream);
        if (q == hipSuccess) {
          if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == seq) break;
          return FMC_ERR_HIP;
        }
        if (q != hipErrorNotReady) {
          std::fprintf(stderr, "fastmc_amd: HIP error %s in a trial\n", hipGetErrorString(q));
          return FMC_ERR_HIP;
        }
      }
    }
  }
  v[0] = v[1] = v[2] = 0.0;
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 3; ++k) v[k] += c->h_part[3 * b + k];
  return FMC_OK;
}

namespace {
double secs(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}
}  // namespace

extern "C" int fmc_trial(fmc_ctx* c, int kind, int gt, int imol,
                         const double* x, const double* y, const double* z,
                         double* out) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!c || !c->sys.finalized) return FMC_ERR_STATE;
  {
    const int r0 = ensure_device(c);
    if (r0) return r0;
  }
  if (gt < 0 || gt >= static_cast<int>(c->sys.guests.size()) || !out)
    return FMC_ERR_ARG;
  if (kind != FMC_INSERT && kind != FMC_DELETE && kind != FMC_DISPLACE)
    return FMC_ERR_ARG;
  const GuestTemplate& tp = c->sys.guests[gt];
  const int ns = tp.ns;
  if (kind != FMC_DELETE && (!x || !y || !z)) return FMC_ERR_ARG;
  if (kind != FMC_INSERT && (imol < 0 || imol >= c->nmol[gt])) return FMC_ERR_ARG;
  if (kind == FMC_INSERT) {
    int used = 0;
    for (size_t t = 0; t < c->sys.guests.size(); ++t)
      used += c->nmol[t] * c->sys.guests[t].ns;
    if (used + ns >= c->cap_atoms) return FMC_ERR_CAPACITY;  // error 75
  }
  TrialParams P;
  fill_devsys(c, &P.S);
  fill_segs(c, P.g, &P.ngt);
  P.nk = c->nk;
  P.kv = c->d_kv;
  P.ak = c->d_ak;
  P.sre = c->d_sre[c->cur];
  P.sim = c->d_sim[c->cur];
  P.nre = c->d_sre[1 - c->cur];
  P.nim = c->d_sim[1 - c->cur];
  P.kind = kind;
  P.ns = ns;
  P.own_t = gt;
  P.own_begin = imol * ns;
  for (int d = 0; d < 3; ++d) P.kmax[d] = c->sys.kmax[d];
  for (int s = 0; s < kMaxSites; ++s) {
    const bool in = s < ns;
    P.q[s] = in ? tp.q[s] : 0.0;
    P.t[s] = in ? tp.type[s] - 1 : 0;
    P.nx[s] = (in && kind != FMC_DELETE) ? x[s] : 0.0;
    P.ny[s] = (in && kind != FMC_DELETE) ? y[s] : 0.0;
    P.nz[s] = (in && kind != FMC_DELETE) ? z[s] : 0.0;
    P.ox[s] = (in && kind != FMC_INSERT) ? c->hx[gt][imol * ns + s] : 0.0;
    P.oy[s] = (in && kind != FMC_INSERT) ? c->hy[gt][imol * ns + s] : 0.0;
    P.oz[s] = (in && kind != FMC_INSERT) ? c->hz[gt][imol * ns + s] : 0.0;
  }
  P.rvolm2 = 2.0 * (2.0 * kPi / c->sys.volm);
  P.engsic = tp.engsic;
  P.nb_recip = std::max(1, std::min(c->nb_recip_max, (c->nk + kTrialBS - 1) / kTrialBS));
  int ntot = P.S.nfw;
  for (int t = 0; t < P.ngt; ++t) ntot += P.g[t].natoms;
  P.nb_real = std::max(1, std::min(c->nb_real_max, (ntot + kTrialBS - 1) / kTrialBS));
  if (P.nb_recip + P.nb_real > kMaxBlocks) return FMC_ERR_ARG;  // partial-sum slots
  // every work-group writes its sums straight into pinned host memory,
  // then the trial's sequence number
  P.part = c->h_part;
  P.word = c->h_word;
  P.seq = ++c->trial_seq;
  static void (*const kTrialForms[5])(TrialParams) = {trial_kernel<0>, trial_kernel<1>,
                                                      trial_kernel<2>, trial_kernel<3>,
                                                      trial_kernel<4>};
  const auto t1 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(kTrialForms[ns <= 4 ? ns : 0], dim3(P.nb_recip + P.nb_real), dim3(kTrialBS), 0,
                     c->stream, P);
  HIPCHECK(hipGetLastError());
  const auto t2 = std::chrono::steady_clock::now();
  double v[3];
  {
    const int rw = wait_trial(c, P.seq, P.nb_recip + P.nb_real, v);
    if (rw) return rw;
  }
  const auto t3 = std::chrono::steady_clock::now();
  double e1 = P.rvolm2 * kR4pie0 * v[0] / P.S.epsq;
  if (kind == FMC_INSERT) e1 = e1 + P.engsic;
  if (kind == FMC_DELETE) e1 = e1 - P.engsic;
  out[0] = e1;
  out[1] = v[1];
  out[2] = v[2];
  c->pending = kind;
  c->pend_gt = gt;
  c->pend_mol = imol;
  for (int s = 0; s < ns; ++s) {
    c->pend_x[s] = kind != FMC_DELETE ? x[s] : 0.0;
    c->pend_y[s] = kind != FMC_DELETE ? y[s] : 0.0;
    c->pend_z[s] = kind != FMC_DELETE ? z[s] : 0.0;
  }
  const auto t4 = std::chrono::steady_clock::now();
  c->n_trials++;
  c->t_prep += secs(t0, t1) + secs(t3, t4);
  c->t_launch += secs(t1, t2);
  c->t_wait += secs(t2, t3);
  return FMC_OK;
}

extern "C" int fmc_ctx_timing(fmc_ctx* c, int reset, double* out) {
  if (!c || !out) return FMC_ERR_ARG;
  out[0] = static_cast<double>(c->n_trials);
  out[1] = c->t_prep;
  out[2] = c->t_launch;
  out[3] = c->t_wait;
  out[4] = static_cast<double>(c->n_commits);
  out[5] = c->t_commit;
  if (reset) {
    c->n_trials = c->n_commits = 0;
    c->t_prep = c->t_launch = c->t_wait = c->t_commit = 0.0;
  }
  return FMC_OK;
}

extern "C" int fmc_commit(fmc_ctx* c) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!c || !c->pending) return FMC_ERR_STATE;
  const int gt = c->pend_gt;
  const int ns = c->sys.guests[gt].ns;
  int lo = 0, hi = 0;  // host-mirror range to upload
  if (c->pending == FMC_INSERT) {
    const int m = c->nmol[gt];
    for (int s = 0; s < ns; ++s) {
      c->hx[gt][m * ns + s] = c->pend_x[s];
      c->hy[gt][m * ns + s] = c->pend_y[s];
      c->hz[gt][m * ns + s] = c->pend_z[s];
    }
    c->nmol[gt] = m + 1;
    lo = m * ns;
    hi = (m + 1) * ns;
  } else if (c->pending == FMC_DELETE) {
    const int m = c->pend_mol;
    const int n = c->nmol[gt];
    for (int i = (m + 1) * ns; i < n * ns; ++i) {
      c->hx[gt][i - ns] = c->hx[gt][i];
      c->hy[gt][i - ns] = c->hy[gt][i];
      c->hz[gt][i - ns] = c->hz[gt][i];
    }
    c->nmol[gt] = n - 1;
    lo = m * ns;
    hi = (n - 1) * ns;
  } else {
    const int m = c->pend_mol;
    for (int s = 0; s < ns; ++s) {
      c->hx[gt][m * ns + s] = c->pend_x[s];
      c->hy[gt][m * ns + s] = c->pend_y[s];
      c->hz[gt][m * ns + s] = c->pend_z[s];
    }
    lo = m * ns;
    hi = (m + 1) * ns;
  }
  if (c->pending != FMC_DELETE) {
    // ns sites from kernel arguments
    SitesArgs A;
    A.gx = c->d_gx[gt];
    A.gy = c->d_gy[gt];
    A.gz = c->d_gz[gt];
    A.lo = lo;
    A.ns = ns;
    for (int s = 0; s < kMaxSites; ++s) {
      A.x[s] = s < ns ? c->pend_x[s] : 0.0;
      A.y[s] = s < ns ? c->pend_y[s] : 0.0;
      A.z[s] = s < ns ? c->pend_z[s] : 0.0;
    }
    hipLaunchKernelGGL(put_sites_kernel, dim3(1), dim3(kMaxSites), 0, c->stream, A);
    HIPCHECK(hipGetLastError());
  } else if (hi > lo) {
    // a delete shifts the successors down on the device
    hipLaunchKernelGGL(shift_sites_kernel, dim3(1), dim3(kBS), 0, c->stream, c->d_gx[gt],
                       c->d_gy[gt], c->d_gz[gt], lo, hi, ns);
    HIPCHECK(hipGetLastError());
  }
  c->cur = 1 - c->cur;  // S(k) <- S_new
  c->pending = 0;
  c->n_commits++;
  c->t_commit += secs(t0, std::chrono::steady_clock::now());
  return FMC_OK;
}

extern "C" int fmc_tail_delta(fmc_ctx* c, int gt, int sign, double elrc,
                              double* delrc) {
  if (!c || !c->sys.finalized || gt < 0 ||
      gt >= static_cast<int>(c->sys.guests.size()) || !delrc)
    return FMC_ERR_ARG;
  std::vector<int> nt, nf;
  c->sys.counts(c->nmol, &nt, &nf);
  for (int s = 0; s < c->sys.guests[gt].ns; ++s)
    nt[c->sys.guests[gt].type[s] - 1] += sign;
  *delrc = c->sys.tail(nt, nf) - elrc;
  return FMC_OK;
}

extern "C" int fmc_energy_full(fmc_ctx* c, double* out) {
  if (!c || !c->sys.finalized || !out) return FMC_ERR_STATE;
  int r;
  if ((r = ensure_device(c))) return r;
  c->pending = 0;
  return energy_full_impl(c, ctx_work(c), out);
}

// ---------------------------------------------------------------------------
// single_point without allocation.  Its scratch -- every partial sum and
// output slot of fmc_energy_full -- is one file-scope device array of
// kScratchBytes, held by each GPU the library runs on until the process
// ends.  Host code never passes the array's name to a kernel or a copy (on
// the host it names a stand-in, not device memory): it uses the device
// address hipGetSymbolAddress returns, looked up once per device with that
// device current.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kScratchBytes = size_t(4) << 20;
constexpr int kScratchDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {nullptr};

// the current device's address of scratch
int scratch_address(void** out) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return FMC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "fastmc_amd: hipGetSymbolAddress(scratch) failed: %s\n",
                   hipGetErrorString(e));
      return FMC_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FMC_OK;
}

}  // namespace

extern "C" int fmc_energy_full_noalloc(fmc_ctx* c, double* out) {
  if (!c || !c->sys.finalized || !out) return FMC_ERR_STATE;
  // building the device state allocates: it must exist already
  if (!c->dev_ready) return FMC_ERR_STATE;
  void* base = nullptr;
  int r;
  if ((r = scratch_address(&base))) return r;
  // layout: out[8] | epart[kb] | pair[2 * 8 * nguest] | rpart[nchunk * nk],
  // each from a 256-byte boundary
  auto up = [](size_t n) { return (n + 255) & ~size_t(255); };
  int nguest = 0;
  for (size_t t = 0; t < c->sys.guests.size(); ++t) nguest += c->nmol[t] * c->sys.guests[t].ns;
  const int natoms = static_cast<int>(c->sys.fx.size()) + nguest;
  const size_t nchunk = static_cast<size_t>(std::max(1, (natoms + c->recip_chunk - 1) /
                                                            c->recip_chunk));
  const size_t kb = static_cast<size_t>((c->nk + kBS - 1) / kBS);
  const size_t n_out = up(8 * sizeof(double));
  const size_t n_ep = up(kb * sizeof(double));
  const size_t n_pair = up(2 * 8 * static_cast<size_t>(nguest) * sizeof(double));
  const size_t n_rp = nchunk * c->nk * sizeof(double2);
  if (n_out + n_ep + n_pair + n_rp > kScratchBytes) {
    std::fprintf(stderr,
                 "fastmc_amd: fmc_energy_full_noalloc needs %zu bytes of scratch (%zu atoms, "
                 "%d k-vectors, %d guest atoms); the library holds %zu\n",
                 n_out + n_ep + n_pair + n_rp, static_cast<size_t>(natoms), c->nk, nguest,
                 kScratchBytes);
    return FMC_ERR_CAPACITY;
  }
  char* p = static_cast<char*>(base);
  FullWork w;
  w.out = reinterpret_cast<double*>(p);
  w.epart = reinterpret_cast<double*>(p + n_out);
  w.pair = reinterpret_cast<double*>(p + n_out + n_ep);
  w.pair_cap = n_pair / sizeof(double);
  w.rpart = reinterpret_cast<double2*>(p + n_out + n_ep + n_pair);
  w.rpart_cap = (kScratchBytes - n_out - n_ep - n_pair) / sizeof(double2);
  c->pending = 0;
  return energy_full_impl(c, w, out);
}

extern "C" int fmc_version(int* major, int* minor) {
  if (major) *major = 1;
  if (minor) *minor = 0;
  return FMC_OK;
}

extern "C" int fmc_device_ok(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 0;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 0;
  return std::strstr(p.gcnArchName, "gfx950") != nullptr ? 1 : 0;
}
