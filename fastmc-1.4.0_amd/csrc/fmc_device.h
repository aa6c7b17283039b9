// Device helpers shared by the single-chain and batched kernels (gfx950,
// wave64).
#pragma once

#include <hip/hip_runtime.h>

namespace fmc {

constexpr int kWave = 64;

// Static per-system data as the kernels see it.  Tables live in HBM and are
// read through L1/L2; the framework is SoA FP64.
struct DevSys {
  int imcon;
  double cell[9];
  double rcell[9];
  double inv[3];         // 1/cell(1), 1/cell(5), 1/cell(9) (images, imcon 1-2)
  double rcut2;
  double rdrewd, rdlrpot;
  int mxgrid;
  int ntpatm;
  double epsq;
  const double* erc;     // [mxgrid+1]
  const double* vvv;     // [table*(mxgrid+1) + i]
  const int* pairtab;    // [ti*ntpatm + tj] -> table or -1 (0-based types)
  int nfw;
  const double* fx;
  const double* fy;
  const double* fz;
  const double* fq;
  const int* ftype;      // 0-based
};

// Loads and stores through pointers known to be in global memory (the
// compiler cannot prove it for pointers read from structures, and flat
// accesses share the LDS wait counter).
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)(p);
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, const T& v) {
  *(__attribute__((address_space(1))) T*)(p) = v;
}

// Minimum image (utility_pack.f:1138-1238) for imcon 1, 2, 3.  Fortran nint
// rounds half away from zero, which is round().
__device__ __forceinline__ void min_image(const DevSys& S, double& x,
                                          double& y, double& z) {
  if (S.imcon == 3) {
    const double sx = S.rcell[0] * x + S.rcell[3] * y + S.rcell[6] * z;
    const double sy = S.rcell[1] * x + S.rcell[4] * y + S.rcell[7] * z;
    const double sz = S.rcell[2] * x + S.rcell[5] * y + S.rcell[8] * z;
    const double ax = sx - round(sx), ay = sy - round(sy), az = sz - round(sz);
    x = S.cell[0] * ax + S.cell[3] * ay + S.cell[6] * az;
    y = S.cell[1] * ax + S.cell[4] * ay + S.cell[7] * az;
    z = S.cell[2] * ax + S.cell[5] * ay + S.cell[8] * az;
  } else if (S.imcon == 2) {
    x -= S.cell[0] * round(S.inv[0] * x);
    y -= S.cell[4] * round(S.inv[1] * y);
    z -= S.cell[8] * round(S.inv[2] * z);
  } else {
    x -= S.cell[0] * round(S.inv[0] * x);
    y -= S.cell[0] * round(S.inv[0] * y);
    z -= S.cell[0] * round(S.inv[0] * z);
  }
}

// 3-point interpolation of ewald.f:670-684 / vdw_terms.f:476-489 on a table
// whose entry i sits at r = i*h (entry 0 duplicates entry 1).
__device__ __forceinline__ double interp3(const double* tab, double r,
                                          double rh, int imax) {
  const double u = r * rh;
  int l = static_cast<int>(u);
  l = l > imax ? imax : l;
  const double p = u - static_cast<double>(l);
  const double v0 = tab[l], v1 = tab[l + 1], v2 = tab[l + 2];
  const double t1 = v0 + (v1 - v0) * p;
  const double t2 = v1 + (v2 - v1) * (p - 1.0);
  return t1 + (t2 - t1) * p * 0.5;
}

// Real-space Coulomb (ewald2) + vdW (srfrce) of one site-atom pair, FP64.
// chgea = q_site / epsq * r4pie0, ts = site type (0-based).
__device__ __forceinline__ void pair_fp64(const DevSys& S, double chgea,
                                          int ts, double dx, double dy,
                                          double dz, double qj, int tj,
                                          double& e2, double& ev) {
  min_image(S, dx, dy, dz);
  const double rsq = dx * dx + dy * dy + dz * dz;
  if (!(S.rcut2 > rsq)) return;
  const double r = sqrt(rsq);
  const double chgprd = chgea * qj;
  const int imax = S.mxgrid - 2;
  if (fabs(chgprd) > 1e-10) e2 += interp3(S.erc, r, S.rdrewd, imax) * chgprd;
  const int k = S.pairtab[ts * S.ntpatm + tj];
  if (k >= 0)
    ev += interp3(S.vvv + static_cast<size_t>(k) * (S.mxgrid + 1), r,
                  S.rdlrpot, imax);
}

// The same pair with its vdW table index vk (pairtab, -1 for none) given.
__device__ __forceinline__ void pair_fp64_vk(const DevSys& S, double chgea, int vk,
                                             double dx, double dy, double dz, double qj,
                                             double& e2, double& ev) {
  min_image(S, dx, dy, dz);
  const double rsq = dx * dx + dy * dy + dz * dz;
  if (!(S.rcut2 > rsq)) return;
  const double r = sqrt(rsq);
  const double chgprd = chgea * qj;
  const int imax = S.mxgrid - 2;
  if (fabs(chgprd) > 1e-10) e2 += interp3(S.erc, r, S.rdrewd, imax) * chgprd;
  if (vk >= 0)
    ev += interp3(S.vvv + static_cast<size_t>(vk) * (S.mxgrid + 1), r, S.rdlrpot, imax);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Cross-lane move of a double by a DPP pattern (lanes the row mask leaves
// out read 0).
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_take(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROWS, 0xf, false);
  const int hi =
      __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Wave sum through DPP (no LDS traffic); the total is valid in lane 63.
// Fixed order: quad pairs, quads, half rows, rows, then rows 0+1 into row
// 1 and 2+3 into row 3, then row 1 into row 3.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_take<0xB1>(v);         // quad_perm [1,0,3,2]
  v += dpp_take<0x4E>(v);         // quad_perm [2,3,0,1]
  v += dpp_take<0x141>(v);        // row_half_mirror
  v += dpp_take<0x140>(v);        // row_mirror
  v += dpp_take<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
  v += dpp_take<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
  return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block sum with a fixed reduction tree (deterministic).  red must hold
// blockDim/64 doubles.  Result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int nw = blockDim.x / kWave;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

// Phase factors e^{i 2 pi l s} for l = 0..kmax by the recurrence of
// ewald.f:97-137 (base from cos/sin, then repeated complex products).
__device__ __forceinline__ void phase_table(double s, int kmax, double2* t) {
  const double twopi = 6.283185307179586;
  double c1, s1;
  sincos(twopi * s, &s1, &c1);
  t[0] = make_double2(1.0, 0.0);
  if (kmax >= 1) t[1] = make_double2(c1, s1);
  for (int l = 2; l <= kmax; ++l) {
    const double2 p = t[l - 1];
    t[l] = make_double2(p.x * c1 - p.y * s1, p.y * c1 + p.x * s1);
  }
}

// e^{i k.r} = X[l] Y[|m|]^(+/-) Z[|n|]^(+/-) with the operation order of
// ewald.f:174-252 (clm/slm first, then the n factor), times the charge.
__device__ __forceinline__ double2 structure_term(double2 X, double2 Y, int m,
                                                  double2 Z, int n, double q) {
  double clm, slm;
  if (m >= 0) {
    clm = X.x * Y.x - X.y * Y.y;
    slm = X.y * Y.x + Y.y * X.x;
  } else {
    clm = X.x * Y.x + X.y * Y.y;
    slm = X.y * Y.x - Y.y * X.x;
  }
  if (n >= 0)
    return make_double2(q * (clm * Z.x - slm * Z.y), q * (slm * Z.x + clm * Z.y));
  return make_double2(q * (clm * Z.x + slm * Z.y), q * (slm * Z.x - clm * Z.y));
}

}  // namespace fmc
