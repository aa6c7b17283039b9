/*
 * The single-chain fmc_* C ABI (include/fastmc_amd.h) backed by the CPU
 * oracle.  TEST INFRASTRUCTURE ONLY: it lets the Fortran driver run the same
 * move loop against the oracle, for trajectory parity with the HIP build and
 * for the CPU baseline timing.  The product library never links it.
 */
#include <stdlib.h>
#include <string.h>

#include "fastmc_amd.h"
#include "fastmc_oracle.h"

struct fmc_ctx {
  orc_sys *s;
  int finalized;
  int pending;
};

int fmc_ctx_create(fmc_ctx **out) {
  if (!out) return FMC_ERR_ARG;
  fmc_ctx *c = calloc(1, sizeof(fmc_ctx));
  c->s = orc_sys_create();
  *out = c;
  return FMC_OK;
}

int fmc_ctx_destroy(fmc_ctx *c) {
  if (!c) return FMC_ERR_ARG;
  orc_sys_destroy(c->s);
  free(c);
  return FMC_OK;
}

int fmc_set_cell(fmc_ctx *c, int imcon, const double *cell) {
  return orc_sys_set_cell(c->s, imcon, cell);
}

int fmc_set_ewald(fmc_ctx *c, double alpha, int k1, int k2, int k3,
                  double epsq) {
  return orc_sys_set_ewald(c->s, alpha, k1, k2, k3, epsq);
}

int fmc_set_cutoffs(fmc_ctx *c, double rcut, double delr, double rvdw) {
  return orc_sys_set_cutoffs(c->s, rcut, delr, rvdw);
}

int fmc_set_ewald_precision(fmc_ctx *c, double eps, double epsq) {
  double cell[9], alpha, rcut;
  int kmax[3], imcon;
  orc_sys_cell_copy(c->s, &imcon, cell, &rcut);
  orc_ewald_precision(eps, rcut, imcon, cell, &alpha, kmax);
  return orc_sys_set_ewald(c->s, alpha, kmax[0], kmax[1], kmax[2], epsq);
}

int fmc_set_vdw(fmc_ctx *c, int ntpatm, int ntpvdw, const int *a, const int *b,
                const int *keys, const double *prm) {
  return orc_sys_set_vdw(c->s, ntpatm, ntpvdw, a, b, keys, prm);
}

int fmc_add_framework(fmc_ctx *c, int n, const double *x, const double *y,
                      const double *z, const double *q, const int *type,
                      const int *frozen) {
  return orc_sys_add_framework(c->s, n, x, y, z, q, type, frozen);
}

int fmc_add_guest_type(fmc_ctx *c, int ns, const double *x, const double *y,
                       const double *z, const double *q, const int *type,
                       const double *mass, int *gt) {
  int r = orc_sys_add_guest_type(c->s, ns, x, y, z, q, type, mass);
  if (r < 0) return FMC_ERR_ARG;
  if (gt) *gt = r;
  return FMC_OK;
}

int fmc_add_guest_molecule(fmc_ctx *c, int gt, const double *x,
                           const double *y, const double *z) {
  return orc_sys_add_guest_molecule(c->s, gt, x, y, z);
}

int fmc_set_capacity(fmc_ctx *c, int m) {
  (void)c;
  (void)m;
  return FMC_OK;
}

int fmc_ctx_finalize(fmc_ctx *c) {
  int r = orc_sys_finalize(c->s);
  c->finalized = 1;
  return r;
}

int fmc_get_info(const fmc_ctx *c, int *nk, double *volm, int *mxgrid,
                 double *drewd, double *dlrpot) {
  if (nk) *nk = orc_sys_nk(c->s);
  if (volm) *volm = orc_sys_volume(c->s);
  int mg;
  double d1, d2;
  orc_sys_grid_info(c->s, &mg, &d1, &d2);
  if (mxgrid) *mxgrid = mg;
  if (drewd) *drewd = d1;
  if (dlrpot) *dlrpot = d2;
  return FMC_OK;
}

int fmc_guest_ewald3(const fmc_ctx *c, int gt, double *e3) {
  *e3 = orc_sys_ewald3en(c->s, gt);
  return FMC_OK;
}

int fmc_guest_count(const fmc_ctx *c, int gt, int *n) {
  *n = orc_sys_nmol(c->s, gt);
  return FMC_OK;
}

int fmc_get_molecule(const fmc_ctx *c, int gt, int m, double *x, double *y,
                     double *z) {
  orc_sys_get_molecule(c->s, gt, m, x, y, z);
  return FMC_OK;
}

int fmc_get_table(const fmc_ctx *c, int which, int t, double *out) {
  return orc_sys_get_table(c->s, which, t, out) == 0 ? FMC_OK : FMC_ERR_ARG;
}

int fmc_get_sk(const fmc_ctx *c, double *re, double *im) {
  orc_sys_get_sk(c->s, re, im);
  return FMC_OK;
}

int fmc_recip_full(fmc_ctx *c, int all, double *e) {
  *e = orc_sys_recip_full(c->s, all);
  return FMC_OK;
}

int fmc_trial(fmc_ctx *c, int kind, int gt, int imol, const double *x,
              const double *y, const double *z, double *out) {
  int r = orc_sys_trial(c->s, kind, gt, imol, x, y, z, out);
  c->pending = (r == 0);
  return r == 0 ? FMC_OK : FMC_ERR_ARG;
}

int fmc_commit(fmc_ctx *c) {
  if (!c->pending) return FMC_ERR_STATE;
  c->pending = 0;
  return orc_sys_commit(c->s);
}

/* the CPU restatement keeps no per-call timers: zeros */
int fmc_ctx_timing(fmc_ctx *c, int reset, double *out) {
  (void)c;
  (void)reset;
  for (int i = 0; i < 6; i++) out[i] = 0.0;
  return FMC_OK;
}

int fmc_tail_delta(fmc_ctx *c, int gt, int sign, double elrc, double *d) {
  *d = orc_sys_tail_delta(c->s, gt, sign, elrc);
  return FMC_OK;
}

int fmc_energy_full(fmc_ctx *c, double *out) {
  orc_sys_energy_full(c->s, out);
  orc_sys_recip_full(c->s, 1);
  return FMC_OK;
}

int fmc_version(int *major, int *minor) {
  *major = 1;
  *minor = 0;
  return FMC_OK;
}

int fmc_device_ok(void) { return 0; }
