// The single-chain context behind the opaque fmc_ctx handle.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fmc_host.h"

struct fmc_ctx {
  fmc::SystemDesc sys;
  // host mirror of the resident guest coordinates, per guest type
  std::vector<double> hx[fmc::kMaxGuestTypes], hy[fmc::kMaxGuestTypes],
      hz[fmc::kMaxGuestTypes];
  std::vector<int> nmol;
  int cap_atoms = 0;
  int nk = 0;
  // device state
  double* d_erc = nullptr;
  double* d_vvv = nullptr;
  int* d_pairtab = nullptr;
  double *d_fx = nullptr, *d_fy = nullptr, *d_fz = nullptr, *d_fq = nullptr;
  int* d_ftype = nullptr;
  int* d_kv = nullptr;
  double* d_ak = nullptr;
  double* d_sre[2] = {nullptr, nullptr};
  double* d_sim[2] = {nullptr, nullptr};
  int cur = 0;
  double* d_gx[fmc::kMaxGuestTypes] = {nullptr};
  double* d_gy[fmc::kMaxGuestTypes] = {nullptr};
  double* d_gz[fmc::kMaxGuestTypes] = {nullptr};
  double2* d_rpart = nullptr;
  int recip_chunk = 1024, recip_nchunk_cap = 0;
  double* d_epart = nullptr;
  double* d_out = nullptr;
  double* h_out = nullptr;
  // pinned, coherent host memory a trial's work-groups write: three
  // partial sums each, then a sequence word once they are written
  double* h_part = nullptr;               // [block][3]
  unsigned long long* h_word = nullptr;   // [block]
  unsigned long long trial_seq = 0;
  hipStream_t stream = nullptr;
  int nb_recip_max = 64, nb_real_max = 64;
  // where fmc_trial / fmc_commit spend their wall time (fmc_ctx_timing):
  // argument set-up on the host, the launch call, the wait for the sums
  long long n_trials = 0, n_commits = 0;
  double t_prep = 0.0, t_launch = 0.0, t_wait = 0.0, t_commit = 0.0;
  // pending trial
  int pending = 0, pend_gt = 0, pend_mol = 0;
  double pend_x[fmc::kMaxSites], pend_y[fmc::kMaxSites], pend_z[fmc::kMaxSites];

  bool dev_ready = false;   // device state built (first single-chain call)

  int host_state();
  int upload_state();
  void release();
};
