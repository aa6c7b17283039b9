// Batched device-resident chains, host side: admission, systems and their
// candidate records, launches (eager, time-budgeted, graph-captured
// cycles), the autotuner over the kernel's tile shapes, checkpoints,
// statistics and the FP64 recompute.  The device side -- the chain
// structures and the persistent kernel that runs the reference's move
// loop (gcmc.f:589-1960), one chain per work-group -- is fmc_batch_dev.h.
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
#include "fmc_batch_dev.h"
#include "fmc_ctx.h"
#include "fmc_device.h"
#include "fmc_host.h"

namespace fmc {

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
  // per chunk of 256 atoms, each wave's count of candidates (two buffers:
  // a chunk's counts are read before the barrier that lets the next chunk
  // but one write them again, so one barrier per chunk suffices)
  __shared__ int wcnt[2][4];
  const int c = blockIdx.x;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  double cx, cy, cz;
  grid_center(P, c, cx, cy, cz);
  Rec* rec = P.rec + static_cast<size_t>(c) * P.cap;
  int* rt = P.rtype + static_cast<size_t>(c) * P.cap;
  int base = 0;   // candidates before this chunk (the same on every thread)
  int buf = 0;
  for (int a0 = 0; a0 < P.S.nfw; a0 += 256, buf ^= 1) {
    const int a = a0 + threadIdx.x;
    bool f = false;
    if (a < P.S.nfw) f = within_reach(P, P.S.fx[a] - cx, P.S.fy[a] - cy, P.S.fz[a] - cz);
    // ascending atom order: the wave's candidates below this lane, then
    // the waves before this one
    const unsigned long long bal = __ballot(f);
    if (lane == 0) wcnt[buf][w] = __popcll(bal);
    __syncthreads();
    int off = base;
    for (int i = 0; i < w; ++i) off += wcnt[buf][i];
    const int slot = off + __popcll(bal & ((1ull << lane) - 1ull));
    if (f && slot < P.cap) {
      rec[slot] = {P.S.fx[a], P.S.fy[a], P.S.fz[a], P.S.fq[a]};
      rt[slot] = P.S.ftype[a];
    }
    base += wcnt[buf][0] + wcnt[buf][1] + wcnt[buf][2] + wcnt[buf][3];
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
// The framework's S(k): a block takes 16 k-vectors, its threads 16 slices
// of the atoms each (atom a in slice a mod 16), and the slices' partial
// sums are added in slice order, so the result is the same on every run.
constexpr int kSfacK = 16, kSfacSlices = 256 / kSfacK;
__global__ void __launch_bounds__(256) sfac_kernel(int nk, const int* kv,
                                                   const double* rcell, int n,
                                                   const double* x, const double* y,
                                                   const double* z, const double* q,
                                                   double* re, double* im) {
  __shared__ double part[2][kSfacSlices][kSfacK];
  const int kk = threadIdx.x % kSfacK, sl = threadIdx.x / kSfacK;
  const int k = blockIdx.x * kSfacK + kk;
  double sx = 0.0, sy = 0.0;
  if (k < nk) {
    const int v = kv[k];
    const int l = ul(v), m = um(v), nn = un(v);
    for (int a = sl; a < n; a += kSfacSlices) {
      double ph = l * (rcell[0] * x[a] + rcell[3] * y[a] + rcell[6] * z[a]) +
                  m * (rcell[1] * x[a] + rcell[4] * y[a] + rcell[7] * z[a]) +
                  nn * (rcell[2] * x[a] + rcell[5] * y[a] + rcell[8] * z[a]);
      ph -= rint(ph);
      double sv, cv;
      sincospi(2.0 * ph, &sv, &cv);
      sx += q[a] * cv;
      sy += q[a] * sv;
    }
  }
  part[0][sl][kk] = sx;
  part[1][sl][kk] = sy;
  __syncthreads();
  if (sl == 0 && k < nk) {
    double tx = 0.0, ty = 0.0;
    for (int s2 = 0; s2 < kSfacSlices; ++s2) {
      tx += part[0][s2][kk];
      ty += part[1][s2][kk];
    }
    re[k] = tx;
    im[k] = ty;
  }
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

// Device blocks that chains and systems leaving a batch released, kept for
// the next admissions.  hipFree waits for the device, so a serving loop
// that frees a chain's buffers between slices stalls on it; a block goes
// back here instead and is handed out again to a request of at least half
// its size.  Per device; at most kCacheBytes are held (blocks past that
// are freed), and the cache is emptied when the last batch is destroyed.
// A block is released only when no launch can reach it any more (its
// chain left the active set after its stream drained), and every buffer
// the batch allocates is written in full before it is read, so a reused
// block's old contents are never seen.
class DevCache {
 public:
  static constexpr size_t kCacheBytes = size_t(16) << 30;

  // a block of at least bytes on the current device (*cap: its size)
  hipError_t take(size_t bytes, void** p, size_t* cap, int* dev) {
    hipError_t e = hipGetDevice(dev);
    if (e != hipSuccess) return e;
    const size_t want = size_class(bytes);
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = free_.lower_bound({*dev, want});
      if (it != free_.end() && it->first.first == *dev && it->first.second <= 2 * want) {
        *p = it->second;
        *cap = it->first.second;
        held_ -= *cap;
        free_.erase(it);
        return hipSuccess;
      }
    }
    e = hipMalloc(p, want);
    if (e == hipErrorOutOfMemory) {   // cached blocks may be what is missing
      (void)hipGetLastError();
      drain();
      e = hipMalloc(p, want);
    }
    *cap = e == hipSuccess ? want : 0;
    if (e != hipSuccess) *p = nullptr;
    return e;
  }

  void give(int dev, void* p, size_t cap) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (held_ + cap <= kCacheBytes) {
        free_.emplace(std::make_pair(dev, cap), p);
        held_ += cap;
        return;
      }
    }
    (void)hipFree(p);
  }

  void drain() {
    std::multimap<std::pair<int, size_t>, void*> all;
    {
      std::lock_guard<std::mutex> g(mu_);
      all.swap(free_);
      held_ = 0;
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto& kv : all) {
      (void)hipSetDevice(kv.first.first);
      (void)hipFree(kv.second);
    }
    (void)hipSetDevice(cur);
  }

  // batches alive: the last one to go empties the cache
  void batch_created() {
    std::lock_guard<std::mutex> g(mu_);
    ++batches_;
  }
  void batch_destroyed() {
    bool last;
    {
      std::lock_guard<std::mutex> g(mu_);
      last = --batches_ == 0;
    }
    if (last) drain();
  }

 private:
  // powers of two up to 2 MiB, then whole 2 MiB steps
  static size_t size_class(size_t bytes) {
    constexpr size_t kStep = size_t(2) << 20;
    if (bytes >= kStep) return (bytes + kStep - 1) / kStep * kStep;
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  std::mutex mu_;
  std::multimap<std::pair<int, size_t>, void*> free_;
  size_t held_ = 0;
  int batches_ = 0;
};

// never destroyed: blocks may still be released while the process exits
DevCache& dev_cache() {
  static DevCache* c = new DevCache();
  return *c;
}

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;     // bytes asked for
  size_t cap = 0;   // bytes of the block
  int dev = 0;
  ~DevBuf() {
    if (p) dev_cache().give(dev, p, cap);
  }
};

template <typename T>
int dput(std::vector<std::unique_ptr<DevBuf>>& keep, T** out, const T* h,
         size_t n) {
  auto b = std::make_unique<DevBuf>();
  const size_t bytes = std::max<size_t>(1, n) * sizeof(T);
  BCHECK(dev_cache().take(bytes, &b->p, &b->cap, &b->dev));
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
  // tile shape: block size, register tile (0: streaming form), k-group
  // width and real-space items per pass
  int bt = 0, kpt = 0, kg = 4, ku = 1;
  bool tuned = false;    // false: autotune at the next launch
  bool force_stream = false;
  int tune_probes = 0;   // shapes the last autotune timed
  bool fixed_kpt = false; // the register tile was set, not derived
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
  hipEvent_t tev0 = nullptr, tev1 = nullptr;   // the autotuner's timings
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
  int graph_bt = 0, graph_kpt = -1, graph_kg = 0, graph_ku = 0;
  bool timing_pending = false;
  double kernel_ms = 0.0;
  long long launches = 0;
  int max_nk = 0;
  int max_nlm = 0;   // (kmax1+1)(2 kmax2+1) over the systems
  int max_pos = 0;   // trial positions: 2 x the largest guest
  int max_grid = 0;  // erfc table entries (mxgrid + 1) over the systems
  std::vector<int> active_ids;
};

// the tile shapes fmc_batch_tiles.hip instantiates, one object per
// (block size, register tile)
namespace fmc {
#define FMC_TILE_DECL(BT_, KPT_)                 \
  extern const KCfg kTiles_##BT_##_##KPT_[];    \
  extern const int kNTiles_##BT_##_##KPT_;
FMC_TILE_DECL(256, 8)
FMC_TILE_DECL(256, 16)
FMC_TILE_DECL(256, 28)
FMC_TILE_DECL(256, 0)
FMC_TILE_DECL(512, 4)
FMC_TILE_DECL(512, 8)
FMC_TILE_DECL(512, 12)
FMC_TILE_DECL(512, 14)
FMC_TILE_DECL(512, 0)
#undef FMC_TILE_DECL
}  // namespace fmc

namespace {

const std::vector<KCfg>& all_cfgs() {
  static const std::vector<KCfg> v = [] {
    std::vector<KCfg> out;
    auto add = [&out](const KCfg* t, int n) { out.insert(out.end(), t, t + n); };
    add(kTiles_256_8, kNTiles_256_8);
    add(kTiles_256_16, kNTiles_256_16);
    add(kTiles_256_28, kNTiles_256_28);
    add(kTiles_256_0, kNTiles_256_0);
    add(kTiles_512_4, kNTiles_512_4);
    add(kTiles_512_8, kNTiles_512_8);
    add(kTiles_512_12, kNTiles_512_12);
    add(kTiles_512_14, kNTiles_512_14);
    add(kTiles_512_0, kNTiles_512_0);
    return out;
  }();
  return v;
}

// kg is ignored by the streaming form (kpt 0)
const KCfg* find_cfg(int bt, int kpt, int kg, int ku, bool mixed) {
  for (const auto& c : all_cfgs())
    if (c.bt == bt && c.kpt == kpt && (kpt == 0 || c.kg == kg) && c.ku == ku &&
        c.mixed == mixed)
      return &c;
  return nullptr;
}

// smallest template KPT for block size bt that covers nk
int kpt_for(int bt, int nk) {
  int best = 0;
  for (const auto& c : all_cfgs())
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
  hipLaunchKernelGGL(sfac_kernel, dim3((nk + kSfacK - 1) / kSfacK), dim3(256), 0, b->setup,
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
      Y.q.gce[g][i] = t.q[i] / s.epsq * kR4pie0;
      Y.q.gtype[g][i] = t.type[i] - 1;
      Y.q.gmass[g][i] = t.mass[i];
      Y.q.tx[g][i] = t.x[i];
      Y.q.ty[g][i] = t.y[i];
      Y.q.tz[g][i] = t.z[i];
    }
    Y.q.e3[g] = t.e3;
    Y.q.engsic[g] = t.engsic;
    if (s.ntpatm <= kLdsTypes)
      for (int i = 0; i < t.ns; ++i)
        for (int tj = 0; tj < s.ntpatm; ++tj)
          Y.q.sptab[g][i][tj] =
              static_cast<short>(s.pairtab[static_cast<size_t>(t.type[i] - 1) * s.ntpatm + tj]);
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
  b->max_nk = b->max_nlm = b->max_pos = b->max_grid = 0;
  for (const auto& sh : b->systems) {
    if (!sh || sh->refs <= 0) continue;
    b->max_nk = std::max(b->max_nk, sh->dev.nk);
    b->max_nlm = std::max(b->max_nlm, sh->nlm);
    b->max_pos = std::max(b->max_pos, sh->npos);
    b->max_grid = std::max(b->max_grid, sh->dev.S.mxgrid + 1);
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

// dynamic LDS of a form: the erfc table, (a_k, packed k index unless
// streaming) + X.Y tables; FMC_ERR_UNSUPPORTED when it and the kernel's
// static LDS exceed the device's per-work-group limit
// The erfc table goes to LDS when it fits beside the rest (*ercap: its
// entries, 0 when real space reads it from global memory instead, which
// gives the same values): a system whose tables only just fit keeps the
// register form rather than the whole batch falling back to streaming.
int lds_need(fmc_batch* b, const KCfg* cfg, size_t* out, int* ercap = nullptr) {
  const size_t elt = b->precision == FMC_PREC_MIXED ? 8 : 16;
  const int nkcap = b->max_nk;
  size_t lds = static_cast<size_t>(b->max_pos) * b->max_nlm * elt;
  if (cfg->kpt > 0)
    lds += static_cast<size_t>(dyn_ak_slots(nkcap)) * sizeof(double) +
           static_cast<size_t>(dyn_kv_slots(nkcap)) * sizeof(int);
  const size_t erc = static_cast<size_t>(dyn_erc_slots(b->max_grid)) * sizeof(double);
  int dev = 0, lmax = 0;
  BCHECK(hipGetDevice(&dev));
  BCHECK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  hipFuncAttributes fa;
  BCHECK(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(cfg->fn)));
  const size_t cap = static_cast<size_t>(lmax);
  if (lds + fa.sharedSizeBytes > cap) return FMC_ERR_UNSUPPORTED;
  const bool with_erc = lds + erc + fa.sharedSizeBytes <= cap;
  *out = lds + (with_erc ? erc : 0);
  if (ercap) *ercap = with_erc ? b->max_grid : 0;
  return FMC_OK;
}

int launch(fmc_batch* b, long long n, int bt, int kpt, int kg, int ku,
           long long budget_ticks = 0) {
  const KCfg* cfg = find_cfg(bt, kpt, kg, ku, b->precision == FMC_PREC_MIXED);
  if (!cfg) return FMC_ERR_UNSUPPORTED;
  if (b->active_ids.empty()) return FMC_OK;
  if (kpt > 0 && bt * kpt < b->max_nk) return FMC_ERR_UNSUPPORTED;
  size_t lds = 0;
  int r, ercap = 0;
  if ((r = lds_need(b, cfg, &lds, &ercap))) return r;
  hipLaunchKernelGGL(cfg->fn, dim3(static_cast<unsigned>(b->active_ids.size())),
                     dim3(cfg->bt), lds, b->stream, b->d_sys, b->d_chains, b->d_ids, n,
                     b->max_nk, b->max_pos, ercap, budget_ticks);
  BCHECK(hipGetLastError());
  return FMC_OK;
}

int launch(fmc_batch* b, long long n, long long budget_ticks = 0) {
  return launch(b, n, b->bt, b->kpt, b->kg, b->ku, budget_ticks);
}

// the register form for block size bt if it covers the largest Nk and fits
// in LDS, else the streaming form (256/512 threads); 0 if neither
int pick_kpt(fmc_batch* b, int bt, bool allow_stream, int* kpt) {
  const bool mixed = b->precision == FMC_PREC_MIXED;
  const int k = kpt_for(bt, std::max(1, b->max_nk));
  size_t lds;
  const KCfg* rc = k > 0 ? find_cfg(bt, k, b->kg, b->ku, mixed) : nullptr;
  if (rc && lds_need(b, rc, &lds) == FMC_OK) {
    *kpt = k;
    return FMC_OK;
  }
  const KCfg* sc = find_cfg(bt, 0, b->kg, b->ku, mixed);
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

// Project autotuner.  Every compiled shape that can run the batch's systems
// is timed on real moves: block size 256 or 512, with the smallest register
// tile that covers the largest Nk or the streaming form, every k-group
// width (the k-vectors whose LDS reads issue together) and real-space items
// per pass.  Each shape first runs a short launch (loading its code), then
// a timed one of 64 moves, GPU time from HIP events on the batch stream;
// the three fastest are timed again over 256 moves, long enough that a
// launch's start-up (LDS tables, chain state) does not decide it, and the
// fastest of those is kept.  The chains are restored to their state before
// tuning, so the tuned run starts where it would have.  Only the block size changes a
// chain's arithmetic (the reductions over threads); the chains record the
// shape their launches used (fmc_chain_stats, checkpoints).
int autotune(fmc_batch* b) {
  const bool mixed = b->precision == FMC_PREC_MIXED;
  struct Shape {
    int bt, kpt, kg, ku;
    float ms;
  };
  std::vector<Shape> cand;
  for (int bt : {256, 512}) {
    const int k = kpt_for(bt, std::max(1, b->max_nk));
    for (const auto& c : all_cfgs()) {
      if (c.bt != bt || c.mixed != mixed || !(c.kpt == 0 || c.kpt == k)) continue;
      size_t lds;
      if (lds_need(b, &c, &lds) != FMC_OK) continue;
      cand.push_back({c.bt, c.kpt, c.kg, c.ku, 0.0f});
    }
  }
  if (cand.empty()) return FMC_ERR_UNSUPPORTED;
  Snapshot snap;
  int r;
  if ((r = snapshot(b, &snap, false))) return r;
  auto time_shape = [&](Shape& sh, long long moves) -> int {
    int rr = launch(b, 2, sh.bt, sh.kpt, sh.kg, sh.ku);
    if (rr) return rr;
    BCHECK(hipEventRecord(b->tev0, b->stream));
    if ((rr = launch(b, moves, sh.bt, sh.kpt, sh.kg, sh.ku))) return rr;
    BCHECK(hipEventRecord(b->tev1, b->stream));
    BCHECK(hipEventSynchronize(b->tev1));
    BCHECK(hipEventElapsedTime(&sh.ms, b->tev0, b->tev1));
    return FMC_OK;
  };
  for (auto& sh : cand)
    if ((r = time_shape(sh, 64))) return r;
  std::sort(cand.begin(), cand.end(), [](const Shape& x, const Shape& y) { return x.ms < y.ms; });
  const size_t top = std::min<size_t>(3, cand.size());
  for (size_t i = 0; i < top; ++i)
    if ((r = time_shape(cand[i], 256))) return r;
  std::sort(cand.begin(), cand.begin() + top,
            [](const Shape& x, const Shape& y) { return x.ms < y.ms; });
  if ((r = snapshot(b, &snap, true))) return r;
  b->bt = cand[0].bt;
  b->kpt = cand[0].kpt;
  b->kg = cand[0].kg;
  b->ku = cand[0].ku;
  b->force_stream = cand[0].kpt == 0;
  b->fixed_kpt = false;
  b->tune_probes = static_cast<int>(cand.size());
  b->tuned = true;
  return FMC_OK;
}

// before a launch: autotune once; a fixed tile follows the systems in use
// (register form while it covers the largest Nk and fits, else streaming)
int ensure_tile(fmc_batch* b) {
  if (!b->tuned) return autotune(b);
  if (b->force_stream) {
    size_t lds;
    const KCfg* sc = find_cfg(b->bt, 0, b->kg, b->ku, b->precision == FMC_PREC_MIXED);
    if (!sc) return FMC_ERR_UNSUPPORTED;
    b->kpt = 0;
    return lds_need(b, sc, &lds);
  }
  if (b->fixed_kpt) return b->bt * b->kpt >= b->max_nk ? FMC_OK : FMC_ERR_UNSUPPORTED;
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
  BCHECK(hipEventCreate(&b->tev0));
  BCHECK(hipEventCreate(&b->tev1));
  BCHECK(hipMalloc(&b->d_chains, max_chains * sizeof(BChain)));
  BCHECK(hipMemset(b->d_chains, 0, max_chains * sizeof(BChain)));
  BCHECK(hipMalloc(&b->d_ids, max_chains * sizeof(int)));
  BCHECK(hipMalloc(&b->d_dev, max_chains * sizeof(double)));
  BCHECK(hipMalloc(&b->d_dev1, sizeof(double)));
  BCHECK(hipStreamCreateWithFlags(&b->setup, hipStreamNonBlocking));
  dev_cache().batch_created();
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
  if (b->tev0) (void)hipEventDestroy(b->tev0);
  if (b->tev1) (void)hipEventDestroy(b->tev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  dev_cache().batch_destroyed();
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
  // a chain's arithmetic depends on the block size of the launches that
  // move it: a batch without a tile yet takes the chain's; a batch whose
  // tile has another block size refuses it
  if (C.tile_code) {
    const int cbt = (C.tile_code & 63) * 64;
    std::lock_guard<std::mutex> g(b->mu);
    if (b->tuned && b->bt != cbt) {
      std::fprintf(stderr,
                   "fastmc_amd batch: checkpoint of a chain moved by %d-thread launches, "
                   "batch tile %d threads\n", cbt, b->bt);
      return FMC_ERR_STATE;
    }
    if (!b->tuned) {
      const int ckpt = (C.tile_code >> 6) & 63;
      b->bt = cbt;
      b->kg = (C.tile_code >> 12) & 15;
      b->ku = (C.tile_code >> 16) & 15;
      b->kpt = ckpt;
      b->force_stream = ckpt == 0;
      b->fixed_kpt = false;
      b->tune_probes = 0;
      b->tuned = true;
    }
  }
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
  if (threads == 0) {
    std::lock_guard<std::mutex> g(b->mu);
    b->bt = b->kpt = 0;
    b->kg = 4;
    b->ku = 1;
    b->tuned = false;
    b->force_stream = false;
    b->fixed_kpt = false;
    return FMC_OK;
  }
  const bool stream = (threads & FMC_TILE_STREAM) != 0;
  // k-group width 4 is compiled for every register tile; one real-space
  // item per pass is the faster of the two
  return fmc_batch_set_tile_shape(b, threads & ~FMC_TILE_STREAM, stream ? 0 : -1, 4, 1);
}

extern "C" int fmc_batch_get_tile(const fmc_batch* b, int* threads) {
  if (!b || !threads) return FMC_ERR_ARG;
  std::lock_guard<std::mutex> g(const_cast<fmc_batch*>(b)->mu);
  *threads = b->bt | ((b->tuned && b->kpt == 0 && b->bt) ? FMC_TILE_STREAM : 0);
  return FMC_OK;
}

extern "C" int fmc_batch_set_tile_shape(fmc_batch* b, int threads, int kpt, int kgroup,
                                        int kunroll) {
  if (!b) return FMC_ERR_ARG;
  if (threads != 256 && threads != 512) return FMC_ERR_UNSUPPORTED;
  if (kgroup != 2 && kgroup != 4 && kgroup != 8) return FMC_ERR_UNSUPPORTED;
  if (kunroll != 2 && kunroll != 1) return FMC_ERR_UNSUPPORTED;
  if (kpt > 0) {
    bool have = false;
    for (const auto& c : all_cfgs())
      if (c.bt == threads && c.kpt == kpt && c.kg == kgroup) have = true;
    if (!have) return FMC_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> g(b->mu);
  b->bt = threads;
  b->kpt = kpt > 0 ? kpt : 0;
  b->kg = kgroup;
  b->ku = kunroll;
  b->tuned = true;
  b->force_stream = kpt == 0;
  b->fixed_kpt = kpt > 0;
  return FMC_OK;
}

extern "C" int fmc_batch_get_tile_shape(const fmc_batch* b, int* threads, int* kpt,
                                        int* kgroup, int* kunroll, int* tuned) {
  if (!b) return FMC_ERR_ARG;
  std::lock_guard<std::mutex> g(const_cast<fmc_batch*>(b)->mu);
  if (threads) *threads = b->bt;
  if (kpt) *kpt = b->kpt;
  if (kgroup) *kgroup = b->kg;
  if (kunroll) *kunroll = b->ku;
  if (tuned) *tuned = b->tuned ? (b->tune_probes > 0 ? 2 : 1) : 0;
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
  if ((r = launch(b, n))) return r;
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
  if ((r = launch(b, max_moves, ticks))) return r;
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
  fold_samples(C);   // the production samples still pending, as the kernel adds them
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
  st->tile_threads = (C.tile_code & 63) * 64;
  st->tile_kpt = (C.tile_code >> 6) & 63;
  st->tile_kgroup = (C.tile_code >> 12) & 15;
  st->tile_kunroll = (C.tile_code >> 16) & 15;
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
      b->graph_bt != b->bt || b->graph_kpt != b->kpt || b->graph_kg != b->kg ||
      b->graph_ku != b->ku) {
    invalidate_graph(b);
    for (int part = 0; part < 2; ++part) {
      BCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int r1 = part == 0 ? launch(b, moves_per_cycle)
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
    b->graph_kg = b->kg;
    b->graph_ku = b->ku;
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
