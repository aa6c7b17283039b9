/*
 * FastMC trial-move energy path: CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).
 *
 * This header and fastmc_oracle.c restate, in plain C, the algorithm of the
 * reference's Fortran routines on the trial-move path.  Only tests/, the
 * smoke check and bench.py's cpu_baseline leg may load or run this code, and
 * only as the checker; the product (libfastmc_amd.so) never links it.
 *
 * Index conventions: all arrays are 0-based C arrays.  The reference's
 * 1-based tables erc(1..mxgrid) and vvv(1..mxgrid,k) are stored with one
 * extra leading guard slot: tab[0] = tab[1] (the reference reads erc(0) /
 * vvv(0,k) out of bounds when r < one grid step; see ewald.f:670-678 and
 * vdw_terms.f:476-483).  Site types and vdW table indices are 1-based like
 * the reference (ltpsit/ltype hold 1..ntpatm, lstvdw holds 1..ntpvdw+1).
 */
#ifndef FASTMC_ORACLE_H
#define FASTMC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_R4PIE0 138935.4835
#define ORC_PI 3.141592653589793
#define ORC_SQRPI 1.7724538509055159
#define ORC_KBOLTZ 1.9872041e-3
#define ORC_BOLTZ 1.3806503e-23

/* ---- cell utilities (utility_pack.f:1351-1467, 1138-1349) ---- */
void orc_dcell(const double a[9], double b[10]);
void orc_invert(const double a[9], double b[9], double *det);
void orc_images(int imcon, int n, const double cell[9], double *x, double *y,
                double *z);
double orc_volume(int imcon, const double cell[9]);
/* ewald precision -> alpha, kmax (utility_pack.f:380-392) */
void orc_ewald_precision(double eps, double rcut, int imcon,
                         const double cell[9], double *alpha, int kmax[3]);

/* ---- tables (ewald.f:1-30, vdw_terms.f:189-396, utility_pack.f:725) ---- */
int orc_mxgrid(double rvdw);
/* erc[0..mxgrid]; returns drewd */
double orc_erfcgen(double alpha, double rcut, int mxgrid, double *erc);
/* prm[t*5+p], keys[t] (1..9), vvv[t*(mxgrid+1)+i] for t=0..ntpvdw (table
 * ntpvdw, i.e. reference index ntpvdw+1, is the zero table); returns dlrpot */
double orc_forgen(int ntpvdw, const int *keys, const double *prm, double rvdw,
                  int mxgrid, double *vvv);

/* ---- reciprocal space (ewald.f:32-630) ---- */
/* number of k-vectors kept by the spherical cut, in enumeration order */
int orc_kvec_count(const double cell[9], const int kmax[3]);
/* full structure factor over n atoms; writes S(k) (ckc, cks, length Nk),
 * returns E_recip + self (internal units); *engsic gets the self term */
double orc_ewald1(int imcon, const double cell[9], double alpha,
                  const int kmax[3], double epsq, int n, const double *x,
                  const double *y, const double *z, const double *q,
                  double *ckc, double *cks, int *newld, double *engsic);
/* per-move update for one rigid molecule; kind 1 insert, 2 delete,
 * 3 displace (pass 1 or 2).  Reads ckcsum/ckssum, writes ckcsnew/ckssnew. */
double orc_ewald1_guest(int imcon, const double cell[9], double alpha,
                        const int kmax[3], double epsq, int natms,
                        const double *x, const double *y, const double *z,
                        const double *q, int kind, int pass,
                        const double *ckcsum, const double *ckssum,
                        double *ckcsnew, double *ckssnew);

/* ---- real space (ewald.f:632-758, vdw_terms.f:398-498) ---- */
double orc_ewald2(double chg, int n, const int *ilist, const double *rsq,
                  const double *atmcharge, const double *erc, double drewd,
                  double rcut, double epsq);
double orc_srfrce(int sittyp, int n, const int *ilist, const double *rsq,
                  const int *ltype, const int *lstvdw, const int *ltpvdw,
                  const double *vvv, int mxgrid, double rcut, double dlrpot);
/* ewald3 for one charge over a list of separation vectors */
double orc_ewald3(double chg, int n, const double *chgj, const double *dx,
                  const double *dy, const double *dz, double alpha,
                  double epsq);
/* per guest-type exclusion constant (gcmc.f:529-552) */
double orc_ewald3_guest(int imcon, const double cell[9], int natms,
                        const double *gx, const double *gy, const double *gz,
                        const double *q, double alpha, double epsq);

/* ---- long-range vdW tail (vdw_terms.f:500-745) ---- */
/* steadd[loc8] (ntpatm*(ntpatm+1)/2 entries) filled; returns elrc */
double orc_lrcorrect(int imcon, int ntpatm, const int *numtyp,
                     const int *numfrz, const int *lstvdw, const int *ltpvdw,
                     const double *prm, double rcut, double volm,
                     double *steadd);
double orc_gstlrcorrect(int imcon, int ntpatm, const int *numtyp,
                        const int *numfrz, const double *steadd, double volm);

/* ---- neighbour list (utility_pack.f:3113-3177) ---- */
/* gstlist is natms x totatm (row i holds the list of site i) */
void orc_guestlistgen(int imcon, const double cell[9], double rcut,
                      double delr, int natms, const double *nx,
                      const double *ny, const double *nz, const int *ind,
                      int totatm, const double *xxx, const double *yyy,
                      const double *zzz, int *gstlentry, int *gstlist);

/* ---- RNG (gcmc.f:2712-2805) ---- */
typedef struct orc_duni {
  float u[97];
  float c, cd, cm;
  int ir, jr;
} orc_duni;
void orc_duni_init(orc_duni *s, int i, int j, int k, int l);
double orc_duni_next(orc_duni *s);
/* the reference's clock seeding (gcmc.f:2753-2758) from date_and_time
 * values(7) (seconds) and values(8) (milliseconds) */
void orc_duni_seed_from_clock(int seconds, int millis, int ijkl[4]);

/* ---- acceptance (utility_pack.f:3049-3095) ---- kind 1/2/3 */
int orc_energy_eval(double eng, double rande, double volm, double press,
                    int ngsts, double temp, double beta, int kind);

/* ---- proposals (utility_pack.f:2545-3047) ---- */
void orc_rotation(double *x, double *y, double *z, int n, double q1,
                  double q2, double q3, double q4);
void orc_random_ins(const double cell[9], int natms, const double *gx,
                    const double *gy, const double *gz, double ra, double rb,
                    double rc, double r1, double r2, double r3, double r4,
                    double *nx, double *ny, double *nz);
void orc_random_disp(const double cell[9], const double rcell[9], int natms,
                     const double *mass, double *nx, double *ny, double *nz,
                     double delr, double rotangle, double ra, double rb,
                     double rc, double r1, double r2, double r3, double r4);

/* ---- system state + trial/commit (gcmc.f:855-1380 energy assembly) ---- */
typedef struct orc_sys orc_sys;

orc_sys *orc_sys_create(void);
void orc_sys_destroy(orc_sys *s);
/* builder (mirrors the product's fmc_* set-up calls) */
int orc_sys_set_cell(orc_sys *s, int imcon, const double cell[9]);
int orc_sys_set_ewald(orc_sys *s, double alpha, int k1, int k2, int k3,
                      double epsq);
int orc_sys_set_cutoffs(orc_sys *s, double rcut, double delr, double rvdw);
/* vdw: ntpatm unique types, ntpvdw potentials given as (a,b,key,prm[5])
 * with prm energies already in internal units */
int orc_sys_set_vdw(orc_sys *s, int ntpatm, int ntpvdw, const int *ta,
                    const int *tb, const int *keys, const double *prm);
int orc_sys_add_framework(orc_sys *s, int n, const double *x, const double *y,
                          const double *z, const double *q, const int *type,
                          const int *frozen);
/* guest template (COM-centred) */
int orc_sys_add_guest_type(orc_sys *s, int nsites, const double *x,
                           const double *y, const double *z, const double *q,
                           const int *type, const double *mass);
int orc_sys_add_guest_molecule(orc_sys *s, int gtype, const double *x,
                               const double *y, const double *z);
int orc_sys_finalize(orc_sys *s);

int orc_sys_nk(const orc_sys *s);
void orc_sys_cell_copy(const orc_sys *s, int *imcon, double cell[9],
                       double *rcut);
void orc_sys_grid_info(const orc_sys *s, int *mxgrid, double *drewd,
                       double *dlrpot);
/* copies of the tables: which 0 erc(0..mxgrid), 1 vvv table t */
int orc_sys_get_table(const orc_sys *s, int which, int t, double *out);
int orc_sys_nmol(const orc_sys *s, int gtype);
int orc_sys_nsites(const orc_sys *s, int gtype);
double orc_sys_volume(const orc_sys *s);
double orc_sys_ewald3en(const orc_sys *s, int gtype);
void orc_sys_get_molecule(const orc_sys *s, int gtype, int imol, double *x,
                          double *y, double *z);
/* running structure factor (ckcsum/ckssum) */
void orc_sys_get_sk(const orc_sys *s, double *ckc, double *cks);

/* framework-only and all-atom ewald1; the all-atom call resets the running
 * S(k) like gcmc.f:522.  Returns energy. */
double orc_sys_recip_full(orc_sys *s, int all_atoms);
/* trial: kind 1 insert (x,y,z new coords), 2 delete (imol), 3 displace
 * (imol + new coords).  out = {ewld1eng, ewld2sum, vdwsum} (internal units,
 * signs as gcmc.f).  Returns 0 on success. */
int orc_sys_trial(orc_sys *s, int kind, int gtype, int imol, const double *x,
                  const double *y, const double *z, double out[3]);
int orc_sys_commit(orc_sys *s);
/* tail correction delta for inserting (+1) / deleting (-1) one molecule of
 * gtype relative to elrc (gstlrcorrect semantics; does not change state) */
double orc_sys_tail_delta(orc_sys *s, int gtype, int sign, double elrc);
/* single_point-style full energy (gcmc.f:2518-2710).  out = {ewald1sum
 * (framework recip+self), ewald1eng (all - fw), ewald2eng, vdweng, ewald3eng,
 * elrc}; internal units */
void orc_sys_energy_full(orc_sys *s, double out[6]);

/* ---- whole chain (gcmc.f move loop, restated for insert/delete/displace
 * incl. translate/rotate-only moves) ----
 *
 * Parity status of the chain: the per-move energies are pinned (they are
 * sums of the pinned pieces above; the bookkeeping invariant "accumulated
 * estep == full-energy difference" is tested), but the ORDER of the duni
 * draws, the adaptive-step schedule (gcmc.f:794-834), the forced insert at
 * N=0 and the production switch are "parity unpinned": the reference seeds
 * duni from the wall clock and its source does not compile unmodified
 * here, so no fixed-seed reference trajectory exists to compare with.  They
 * are pinned only by the statistical known answer (mean loading of the
 * reference's example run) and by restating gcmc.f line by line. */
typedef struct orc_chain_params {
  double temp;          /* K */
  int ntpguest;
  double press[8];      /* Pa (CONTROL bar * 1e5) */
  /* raw (unnormalised) move frequencies: ins, del, dis, jmp, flx, swp, tra,
   * rot, mjp */
  double freq[8][9];
  double delr[8];
  double disp_ratio[8], tran_ratio[8], rota_ratio;
  int seed[4];          /* duni (i,j,k,l) */
  double engunit;       /* 418.4 for kcal */
} orc_chain_params;

typedef struct orc_move_rec {
  int kind;             /* 1 ins 2 del 3 disp 4 tran 5 rota 0 none */
  int iguest;           /* 0-based */
  int accepted;
  int nmol_after;       /* molecules of iguest after the move */
  double estep;         /* kcal/mol (engunit units) */
  double e1, e2, vdw, e3, delrc; /* components, internal units */
} orc_move_rec;

typedef struct orc_chain orc_chain;
orc_chain *orc_chain_create(orc_sys *s, const orc_chain_params *p);
void orc_chain_destroy(orc_chain *c);
/* run nsteps moves; if rec != NULL fill nsteps records */
int orc_chain_run(orc_chain *c, long nsteps, orc_move_rec *rec);
double orc_chain_energy(const orc_chain *c, int gtype);
void orc_chain_adapt_state(const orc_chain *c, int gtype, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
