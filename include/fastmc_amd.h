/*
 * fastmc_amd.h -- C ABI of the MI355X (gfx950) trial-move energy path.
 *
 * This is the drop-in boundary between a GCMC driver (the reference's
 * Fortran move loop, or this repository's own Fortran driver) and the HIP
 * kernels.  Everything is plain C: opaque handles, plain pointers and sizes,
 * scalars by value, int status returns (0 = OK, negative = FMC_ERR_*).
 *
 * Units are the reference's internal DL_POLY units (energy 10 J/mol, length
 * Angstrom, charge e); conversion to kcal/mol (engunit 418.4) stays in the
 * driver, exactly as in the reference.  Atom-type indices are 1-based
 * (ltpsit/ltype); guest-type and molecule indices are 0-based.
 *
 * Citations below are to the reference checkout (fastmc/...).
 */
#ifndef FASTMC_AMD_H
#define FASTMC_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FMC_OK 0
#define FMC_ERR_ARG -1         /* bad argument (index, size, NULL) */
#define FMC_ERR_STATE -2       /* call out of order (e.g. commit w/o trial) */
#define FMC_ERR_UNSUPPORTED -3 /* imcon / potential not on this path */
#define FMC_ERR_CAPACITY -4    /* guest capacity exceeded (reference err 75) */
#define FMC_ERR_HIP -5         /* HIP runtime failure */
#define FMC_ERR_NODEVICE -6    /* no gfx950 device visible */
#define FMC_ERR_CELL -7        /* cutoff exceeds half cell width (err 95) */

/* trial kinds (gcmc.f insert / delete / displace blocks) */
#define FMC_INSERT 1
#define FMC_DELETE 2
#define FMC_DISPLACE 3

/* vdW potential keys (vdw_terms.f:48-83): 1 12-6, 2 lj, 3 nm, 4 buck,
 * 5 bhm, 6 hbnd, 7 snm/hcnm, 8 morse, 9 wca */

typedef struct fmc_ctx fmc_ctx;

/* ---------------------------------------------------------------------
 * Single-chain context: set-up.  Replaces the module-global state the
 * reference fills in initscan/readfield/readconfig/alloc_config_arrays
 * (utility_pack.f:8-56,176-457,706-905; readinputs.f:11-515) and the table
 * generators erfcgen (ewald.f:1-30) and forgen (vdw_terms.f:189-396).
 * ------------------------------------------------------------------- */
int fmc_ctx_create(fmc_ctx **out);
int fmc_ctx_destroy(fmc_ctx *ctx);
/* imcon 1 cubic, 2 orthorhombic, 3 parallelepiped (images,
 * utility_pack.f:1138-1349); cell = DL_POLY cell(1:9) */
int fmc_set_cell(fmc_ctx *ctx, int imcon, const double *cell);
/* "ewald sum alpha k1 k2 k3" (utility_pack.f:395-398) */
int fmc_set_ewald(fmc_ctx *ctx, double alpha, int kmax1, int kmax2,
                  int kmax3, double epsq);
/* "ewald precision eps" -> alpha, kmax (utility_pack.f:380-392); needs the
 * cell and cutoff set first */
int fmc_set_ewald_precision(fmc_ctx *ctx, double eps, double epsq);
/* cutoff, delr (neighbour skin, utility_pack.f:3130) and rvdw (table range,
 * utility_pack.f:725; pass 0 for rvdw = rcut) */
int fmc_set_cutoffs(fmc_ctx *ctx, double rcut, double delr, double rvdw);
/* VDW block of FIELD (vdw_terms.f:1-187): ntpatm unique atom types;
 * potential t acts between types atom_a[t], atom_b[t] (1-based) with key
 * keys[t] and parameters prm[5*t .. 5*t+4], energies already multiplied by
 * engunit (vdw_terms.f:119-139) */
int fmc_set_vdw(fmc_ctx *ctx, int ntpatm, int ntpvdw, const int *atom_a,
                const int *atom_b, const int *keys, const double *prm);
/* framework atoms (condensefram, utility_pack.f:908-947); frozen may be
 * NULL (all frozen, the FIELD convention for rigid frameworks) */
int fmc_add_framework(fmc_ctx *ctx, int n, const double *x, const double *y,
                      const double *z, const double *q, const int *type,
                      const int *frozen);
/* rigid guest template, COM-centred as readfield leaves it
 * (readinputs.f:226-253); *gtype_out gets the 0-based guest type */
int fmc_add_guest_type(fmc_ctx *ctx, int nsites, const double *x,
                       const double *y, const double *z, const double *q,
                       const int *type, const double *mass, int *gtype_out);
/* a guest molecule already present (CONFIG with NUMMOLS > 0 / REVIVE) */
int fmc_add_guest_molecule(fmc_ctx *ctx, int gtype, const double *x,
                           const double *y, const double *z);
/* maximum guest atoms ("max guest atoms", utility_pack.f:208) */
int fmc_set_capacity(fmc_ctx *ctx, int max_guest_atoms);
/* build tables, k-vectors and device state; computes S(k) of all atoms
 * like the ewald1 call at gcmc.f:522 */
int fmc_ctx_finalize(fmc_ctx *ctx);

/* info: nk = k-vectors kept (ewald.f:81-84), volm, mxgrid, ewald3en of
 * each guest type (gcmc.f:529-552) */
int fmc_get_info(const fmc_ctx *ctx, int *nk, double *volm, int *mxgrid,
                 double *drewd, double *dlrpot);
int fmc_guest_ewald3(const fmc_ctx *ctx, int gtype, double *e3);
int fmc_guest_count(const fmc_ctx *ctx, int gtype, int *nmol);
int fmc_get_molecule(const fmc_ctx *ctx, int gtype, int imol, double *x,
                     double *y, double *z);
/* copies of the tables as built: which 0 = erc(0..mxgrid), 1 = vvv table t
 * (0..mxgrid); entry 0 duplicates entry 1 (guard for r < one grid step) */
int fmc_get_table(const fmc_ctx *ctx, int which, int t, double *out);
/* running structure factor, Nk entries each (ckcsum/ckssum) */
int fmc_get_sk(const fmc_ctx *ctx, double *re, double *im);

/* ---------------------------------------------------------------------
 * Hot path.
 * ------------------------------------------------------------------- */
/* ewald1 (ewald.f:368-630): all_atoms 0 = framework only (no state change),
 * 1 = every atom; the all-atom call resets the running S(k) */
int fmc_recip_full(fmc_ctx *ctx, int all_atoms, double *energy);

/* One trial move: the fused replacement of guestlistgen
 * (utility_pack.f:3113-3177), ewald1_guest (ewald.f:32-365) for pass 1 and
 * pass 2, and the per-site ewald2 (ewald.f:632-694) + srfrce
 * (vdw_terms.f:398-498) loops of gcmc.f:906-936, 1076-1104, 1254-1324.
 *   kind FMC_INSERT:   x,y,z = new site coordinates; imol ignored
 *   kind FMC_DELETE:   molecule imol of gtype (resident); x,y,z ignored
 *   kind FMC_DISPLACE: molecule imol of gtype moves to x,y,z
 * out[0] = ewld1eng (reciprocal delta incl. the +/- self term of
 *          ewald.f:335-362), out[1] = ewld2sum, out[2] = vdwsum, with the
 * reference's signs, so the driver forms estep exactly as gcmc.f:946,1112,
 * 1328.  Blocking until out is filled; the move is pending until fmc_commit
 * or the next trial.  It returns once the sums have reached host memory,
 * which can be before its kernel has retired: later calls on the context
 * are ordered after it, and fmc_get_sk and fmc_ctx_destroy drain it.
 */
int fmc_trial(fmc_ctx *ctx, int kind, int gtype, int imol, const double *x,
              const double *y, const double *z, double *out);
/* accept the pending trial (gcmc.f:965-981, 1134-1151, 1355-1368 and
 * condense, utility_pack.f:3325-3423): S(k) <- S_new, coordinates updated;
 * a deleted molecule's successors shift down by one, as in gcmc.f:1127 */
int fmc_commit(fmc_ctx *ctx);
/* Wall time of this context's trials and commits (no reference
 * counterpart; the reference times only the whole run, gcmc.f:2433-2490):
 * out[0] trials, out[1] seconds of host work in fmc_trial (argument
 * set-up, reading the sums), out[2] seconds in the kernel launch call,
 * out[3] seconds waiting for the work-groups' sums, out[4] commits,
 * out[5] seconds in fmc_commit; reset != 0 clears them */
int fmc_ctx_timing(fmc_ctx *ctx, int reset, double *out);
/* gstlrcorrect (vdw_terms.f:684-745) for N(gtype)+sign molecules, minus
 * elrc (gcmc.f:891-901, 1056-1069) */
int fmc_tail_delta(fmc_ctx *ctx, int gtype, int sign, double elrc,
                   double *delrc);
/* single_point (gcmc.f:2518-2710): out = {framework recip+self,
 * recip(all)-recip(framework), real-space coulomb, vdW, ewald3 exclusions,
 * elrc}; internal units.  Also resets the running S(k) like gcmc.f:522. */
int fmc_energy_full(fmc_ctx *ctx, double *out);
/* fmc_energy_full without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Every
 * partial sum and output slot lives in one static device array of 4 MiB
 * that each GPU the library runs on holds until the process ends: a system
 * fits while 16 * ceil(atoms / 1024) * Nk + 128 * guest atoms + 8 KiB
 * stays within it, else FMC_ERR_CAPACITY (with a message on stderr).  The
 * context's device state must exist already (any earlier fmc_trial,
 * fmc_recip_full or fmc_energy_full builds it), else FMC_ERR_STATE.  Calls
 * share the array of their GPU and keep its last contents; calls that
 * overlap on one GPU race. */
int fmc_energy_full_noalloc(fmc_ctx *ctx, double *out);

/* ---------------------------------------------------------------------
 * Batched device-resident chains.  No reference counterpart call site: it
 * runs many independent copies of the gcmc.f:589-1960 move loop (insert,
 * delete, displace incl. translate/rotate-only moves, duni RNG, energy_eval
 * acceptance, adaptive steps) entirely on the GPU, one chain per work-group,
 * replacing the reference's one-chain-per-MPI-rank replica mode
 * (gcmc.f:366-368).
 * ------------------------------------------------------------------- */
#define FMC_PREC_FP64 0  /* everything FP64: same accept/reject stream */
#define FMC_PREC_MIXED 1 /* FP32 phase factors e^{ik.r} and dS(k), FP64
                          * S(k), real-space pair terms and totals */

typedef struct fmc_batch fmc_batch;

typedef struct fmc_chain_params {
  double temp;          /* K */
  int ntpguest;         /* guest types in the chain's context (<= 4) */
  int seed[4];          /* duni (i,j,k,l) (gcmc.f:2753-2758) */
  double press[4];      /* Pa per guest type (CONTROL bar * 1e5) */
  /* raw move frequencies per guest (readinputs.f:665-688): insert, delete,
   * displace, jump, flex, swap, translate, rotate, multi-jump; all zero
   * means 1/3 insert/delete/displace (gcmc.f:285-298) */
  double freq[4][9];
  double delr[4];       /* initial displacement length (CONTROL delr) */
  double disp_ratio[4]; /* target acceptance ratios (default 0.5) */
  double tran_ratio[4];
  double rota_ratio;
  double engunit;       /* 418.4 for kcal */
  long long eq_steps;   /* moves before production averaging starts */
  long long max_moves;  /* the chain stops after this many moves (0 = no
                         * limit); a stopped chain idles in later launches */
  long long recompute_every; /* > 0: FP64 recompute (as fmc_batch_recompute)
                         * inside the launches, every this many moves of
                         * the chain, so its cadence does not depend on how
                         * the moves are split into launches */
  long long window;     /* production samples per selectivity window
                         * (CONTROL "averaging window", readinputs.f:740-742;
                         * 0 = 100000 as readinputs.f:888-889) */
} fmc_chain_params;

typedef struct fmc_chain_stats {
  long long steps;            /* moves done */
  long long prod_steps;       /* production samples */
  long long accepted[6];      /* by kind: 0 none,1 ins,2 del,3 disp,4 tran,5 rota */
  long long attempted[6];
  int nmol[4];                /* current molecules per guest type */
  double energy[4];           /* accumulated energy (kcal/mol), gcmc.f:981 */
  double sum_n[4], sum_n2[4]; /* production sums of N, N^2 */
  double sum_e[4];            /* production sum of energy */
  double delrdisp[4], rotangle, tran_delr[4], rota_rotangle;
  /* FP64 full recompute (fmc_batch_recompute; also run at admission):
   * e_acc = running sum of accepted estep from the admission energy,
   * e_full = full guest energy at the last recompute (kcal/mol),
   * drift = e_acc - the recomputed energy with the same tail bookkeeping,
   * sk_dev = largest |S_running(k) - S_exact(k)| the recompute replaced */
  double e_acc, e_full, drift, sk_dev;
  long long recomputes;
  double sum_en[4], sum_e2[4]; /* production sums of E*N and E^2
                                * (gcmc.f:1838-1848; isosteric heat) */
  /* end-of-run windows (gcmc.f:1969-2031) for chains with a move limit of
   * at most 500000 production moves: win_len = int(production moves / 5);
   * win[j][q][g] sums quantity q (N, E, E*N, N^2, E^2) of guest g over the
   * production samples j*win_len ... j*win_len + win_len (the reference's
   * windows of win_len + 1 samples, to be divided by win_len); win_closed
   * windows are complete, the next one holds what has been sampled */
  long long win_len;
  int win_closed;
  double win[5][5][4];
  /* selectivity of guest j over guest i, j < i (gcmc.f:1887-1916): over
   * the closed windows of sel_window production samples, the running
   * window-weighted mean of <N_j>/<N_i> (sel_mean), of its square
   * (sel_second) and the variance the reference prints after +/-
   * (sel_var), pair p = i (i - 1) / 2 + j; sel_count windows are closed,
   * and sel_base holds the production sums of N when the last one closed
   * (the open window's samples are sum_n - sel_base).  The open window
   * enters at the end of the run with weight samples / sel_window. */
  long long sel_window;
  long long sel_count;
  double sel_base[4];
  double sel_mean[6], sel_second[6], sel_var[6];
  /* the tile shape of the launches that moved the chain (all 0 before its
   * first launch); see fmc_batch_set_tile_shape */
  int tile_threads, tile_kpt, tile_kgroup, tile_kunroll;
} fmc_chain_stats;

/* max_chains chains; precision FMC_PREC_*.  */
int fmc_batch_create(fmc_batch **out, int max_chains, int precision);
int fmc_batch_destroy(fmc_batch *b);
/* add a chain whose system is the finalized context ctx (copied; ctx may be
 * destroyed afterwards).  *chain_id gets the slot. */
int fmc_batch_add_chain(fmc_batch *b, const fmc_ctx *ctx,
                        const fmc_chain_params *p, int *chain_id);
/* remove a chain and free its slot (serving) */
int fmc_batch_remove_chain(fmc_batch *b, int chain_id);
/* Chain checkpoint, the batched form of the reference's restart state
 * (REVIVE, written by revive utility_pack.f:1800-1814 every 1000 moves,
 * gcmc.f:737-740, and read by revread utility_pack.f:1835-1862).  Beyond
 * REVIVE's molecule counts, coordinates, statistics and energies it holds
 * the duni table, the adaptive step sizes, the counters and the running
 * S(k), so a restored chain continues the identical trajectory (same
 * precision).  The bytes are opaque and valid for this library build.
 * buf == NULL: *len gets the size.  Waits for the launch in flight; call it
 * from the thread that enqueues launches (a launch enqueued meanwhile by
 * another thread races with the copy). */
int fmc_batch_checkpoint(fmc_batch *b, int chain_id, void *buf, long long cap,
                         long long *len);
/* add a chain from a checkpoint; ctx must describe the system the chain was
 * admitted with (framework, cell, Ewald and force-field set-up; its guest
 * molecules are ignored), and the duni state must be one the generator
 * reaches (table and carry on the 2^-24 grid, indices 64 apart), else
 * FMC_ERR_ARG */
int fmc_batch_restore_chain(fmc_batch *b, const fmc_ctx *ctx, const void *buf,
                            long long len, int *chain_id);
/* a chain's guest molecules of type gtype: *nmol molecules, sites
 * molecule-major as framwkxxx(mol, 1:nummols*numatoms) (utility_pack.f:
 * 1805-1810); x = y = z = NULL: count only.  Waits for the launch in
 * flight. */
int fmc_batch_get_molecules(fmc_batch *b, int chain_id, int gtype, double *x,
                            double *y, double *z, int cap_sites, int *nmol);
/* run n_moves moves on every active chain (one persistent launch) */
int fmc_batch_run(fmc_batch *b, long long n_moves);
/* enqueue without waiting / wait for completion */
int fmc_batch_launch(fmc_batch *b, long long n_moves);
int fmc_batch_sync(fmc_batch *b);
/* enqueue a launch in which each chain moves until budget_us of GPU time
 * has passed (checked after every move) or it has made max_moves moves;
 * the per-chain move counts differ, the chains' trajectories do not depend
 * on where the launches split them */
int fmc_batch_launch_budget(fmc_batch *b, long long max_moves, double budget_us);
/* *done = 1 when nothing enqueued on the batch is still running (does not
 * wait) */
int fmc_batch_query(fmc_batch *b, int *done);
/* a chain's counters and statistics; waits for the launch in flight */
int fmc_batch_stats(fmc_batch *b, int chain_id, fmc_chain_stats *st);
/* per-move log for one chain (parity checks): rec[8*i..8*i+7] = {kind,
 * iguest, accepted, nmol_after, estep (kcal/mol), ewld1eng incl. self term,
 * ewld2sum, vdwsum (internal units)}; enable before running, with capacity
 * moves.  Both wait for the launch in flight. */
int fmc_batch_enable_log(fmc_batch *b, int chain_id, long long capacity);
int fmc_batch_get_log(fmc_batch *b, int chain_id, long long n, double *rec);
/* FP64 full recompute of every active chain (periodic resync; no
 * reference counterpart, the reference only accumulates estep,
 * gcmc.f:981,1151,1368): S(k) <- S_framework + exact guest sum, full guest
 * energy as single_point (gcmc.f:2518-2710) computes it, drift against the
 * running sum.  Returns the largest |S_running - S_exact| and |drift|. */
int fmc_batch_recompute(fmc_batch *b, double *max_sk_dev, double *max_drift);
/* same, S(k) deviation only (kept for callers of the first version) */
int fmc_batch_resync(fmc_batch *b, double *max_dev);
/* enqueue n_cycles cycles of (moves_per_cycle moves on every active chain,
 * then the FP64 recompute); both launches are captured once as HIP graphs
 * and replayed each cycle, with a HIP event pair on the batch stream around
 * each move replay (fmc_batch_timing).  Complete with fmc_batch_sync.  The
 * graphs are re-captured when the active set, the tile or moves_per_cycle
 * change. */
int fmc_batch_launch_cycles(fmc_batch *b, long long n_cycles,
                            long long moves_per_cycle);
/* chains that reached max_moves: up to cap ids into ids, count in *n
 * (waits for the launch in flight) */
int fmc_batch_poll(fmc_batch *b, int *ids, int cap, int *n);
/* GPU time of the batch kernel launches since the last reset, measured
 * with HIP events on the batch's stream (one launch = one fmc_batch_launch
 * completed by fmc_batch_sync); reset != 0 clears the totals */
int fmc_batch_timing(fmc_batch *b, int reset, double *ms, long long *launches);
/* tile shape: threads per chain work-group (256 or 512); 0 =
 * autotuned.  S(k) is register-resident when the largest Nk fits the tile,
 * else the streaming form runs (S(k) in HBM/L2, any Nk); threads |
 * FMC_TILE_STREAM forces the streaming form (256 or 512), with k-group
 * width 4 and one real-space item per pass.  Results do not depend on the
 * form. */
#define FMC_TILE_STREAM 0x10000
int fmc_batch_set_tile(fmc_batch *b, int threads);
int fmc_batch_get_tile(const fmc_batch *b, int *threads);
/* the whole tile shape: threads per work-group (256, 512); kpt, the
 * register tile's k-vectors per thread (-1: the smallest that covers the
 * largest Nk while it fits, else the streaming form; 0: the streaming
 * form); kgroup, the k-vectors whose LDS reads issue together (2, 4, 8);
 * kunroll, the real-space items each thread loads per pass (1, 2).  Only
 * threads changes a chain's arithmetic.  fmc_batch_set_tile(b, 0) lets the
 * autotuner time every compiled shape on real moves before the next launch
 * (the k-vector enumeration it tiles: ewald.f:141-213); get reports the
 * shape in use and *tuned = 0 (none yet), 1 (set) or 2 (autotuned).  A
 * checkpoint restored into a batch without a tile sets the chain's; into a
 * batch whose tile has another thread count it is refused
 * (FMC_ERR_STATE). */
int fmc_batch_set_tile_shape(fmc_batch *b, int threads, int kpt, int kgroup,
                             int kunroll);
int fmc_batch_get_tile_shape(const fmc_batch *b, int *threads, int *kpt,
                             int *kgroup, int *kunroll, int *tuned);
/* bytes of device memory one chain of this context needs (admission) */
int fmc_chain_footprint(const fmc_ctx *ctx, int precision,
                        long long *bytes);

/* library version / device check */
int fmc_version(int *major, int *minor);
int fmc_device_ok(void);

#ifdef __cplusplus
}
#endif
#endif
