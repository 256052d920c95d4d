/*
 * cm3d_nb.h -- C ABI of the MI355X (gfx950) nonbonded force engine.
 *
 * Drop-in for cm3d's intermolecular force path.  Every array argument has
 * cm3d's own layout: positions and forces are 3N SoA doubles (x[N] y[N] z[N],
 * the px|py|pz and fxl|fyl|fzl blocks of utils/md_util.new.c:113-142),
 * tables are [ninter][ntable] doubles uniform in r^2, the type map is the
 * row-major ntypes x ntypes int matrix with -1 for null pairs.  Units are
 * cm3d's: Angstrom, Kelvin, charges in sqrt(K*A).  Forces are returned
 * *before* cm3d's rmass division (energy/control/md_fcntr.c:74-75).
 *
 * Entry points and the cm3d routine each replaces (reference tree paths):
 *   nb_create        search_ter_base + set_charges + set_exclude outputs
 *                    (input/md_searchbase.c:34, energy/control/elec_ngbr.c:36,
 *                     energy/control/md_setexclude.c:38) -> device context
 *   nb_build_lists   get_ngbr + get_ngbr_e   (energy/control/ngbr_list.c:297,
 *                                             energy/control/elec_ngbr.c:220)
 *   nb_need_rebuild  nrefsh's displacement criterion
 *                                            (energy/control/md_finter.c:43-73)
 *   nb_force         force(...,iflg) for iflg 0/1 minus rmass: force_ter +
 *                    force_ter_e (+ fk_ewald + ecorr when iflg==1, iperd==3)
 *                    (energy/control/md_fcntr.c:44-76)
 *   nb_energy fields the statics read by getvireal(...,1): getsvirter,
 *                    getsvirter_e, getvewald (energy/inter/md_fnopol.c:386,
 *                    energy/inter/finter_e.c:410, energy/inter/md_ewald.c:364)
 *   nb_table_pair / nb_table_coul / nb_kvectors / nb_ewald_consts
 *                    set-up helpers restating get_ter_parm
 *                    (input/md_getparm.c:41), set_charges
 *                    (energy/control/elec_ngbr.c:36) and ewald_setup
 *                    (energy/inter/md_ewald.c:48) for callers without
 *                    cm3d's input layer.
 *
 * Errors: every entry point returns NB_OK (0) or a negative NB_ERR_* code
 * and never exits; nb_last_error() gives the message.  cm3d's md_error maps
 * to a negative return, md_warning (a pair below the table minimum, which
 * cm3d clamps, md_fnopol.c:133-138) to nb_energy.nclamp > 0.
 *
 * Threading: one context is used by one host thread at a time (cm3d's own
 * path is non-reentrant through file statics; here all that state lives in
 * the context, so distinct contexts are independent).
 */
#ifndef CM3D_NB_H
#define CM3D_NB_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define NB_OK 0
#define NB_ERR_ARG (-1)
#define NB_ERR_HIP (-2)
#define NB_ERR_CAPACITY (-3)
#define NB_ERR_NODEVICE (-4)
#define NB_ERR_UNSUPPORTED (-5)

/* Arithmetic of the pair kernels.  NB_PREC_FP64 follows cm3d's FP64
 * arithmetic pair by pair; NB_PREC_MIXED evaluates pairs in FP32 from
 * 32-bit fixed-point fractional coordinates and accumulates forces in 64-bit
 * fixed point.  Ewald sums, energies and virials are FP64 in both. */
#define NB_PREC_FP64 0
#define NB_PREC_MIXED 1

typedef struct nb_ctx nb_ctx;

/* Everything cm3d's set-up leaves in INTER / COUL / SIMPARMS for the path
 * (include/md_typedef.h:59-69,120-131,255-268).  Host pointers, read during
 * nb_create only. */
typedef struct nb_params {
  int natoms, ntypes, ninter, ntable, iperd;
  const int *itype;                  /* [natoms]                         */
  const int *map;                    /* [ntypes*ntypes], -1 = null       */
  const double *vtab, *dvtab, *switchs;   /* [ninter*ntable]             */
  const double *rmins_cut2, *rmaxs_cut2, *rminl_cut2, *rmaxl_cut2;
  const double *dx2tab, *rmaxs_skin2, *rmaxl_skin2;  /* [ninter]         */
  double skin;                       /* inter->skin_ter                  */
  const int *nexclude, *exclude;     /* vdW exclusions: per atom i the   */
                                     /* sorted partners j>i, flattened   */
  const double *qch;                 /* [natoms], sqrt(K A)              */
  const double *coultab, *dcoultab, *switchse;       /* [ntable]         */
  double dx2coul, rmins_coul2, rmaxs_coul2, rminl_coul2, rmaxl_coul2;
  double rmaxs_skin2_e, rmaxl_skin2_e;
  const int *cnexclude, *cexclude;   /* coulomb exclusions, same layout  */
  int necorr;
  const int *iecorr, *jecorr;        /* [necorr] excluded charged pairs  */
  int ktot;                          /* half-space k vectors             */
  const int *ka, *kb, *kc;           /* [ktot], cm3d enumeration order   */
  double kappa, ewald_self, ewald_bkg;
  int precision;                     /* NB_PREC_*                        */
  int device;                        /* HIP device ordinal               */
  /* No cm3d counterpart: how many systems are expected to share the GPU
   * with this one (the members of a batch, a server's slots); 0 or 1 =
   * alone.  It sizes the work split of the reciprocal sum (fewer, larger
   * pieces when many systems run together, each piece capped so that the
   * pieces of small and large systems take about the same time) and, from
   * 4 up, the mixed pair kernel's atoms per workgroup; results change only
   * at rounding level and stay fixed for the context's lifetime, so a
   * system steps bit for bit the same alone and in a batch. */
  int sharing;
} nb_params;

/* By-products of force(...,1) that getvireal(...,1) consumes. */
typedef struct nb_energy {
  double spotl, spothl, sprsl, sprslt[9];  /* getsvirter               */
  double spote, spothe, sprse, sprset[9];  /* getsvirter_e             */
  double vk_ewald, vp_ewald, vp_ewald_tensor[9];
  double pot_ecorr, prs_ecorr, prs_ecorr_tensor[9];
  int nclamp;                              /* pairs below table min    */
  /* No cm3d counterpart: the real-space coulomb energies and virials of the
   * force path.  Equal to spote .. sprset except after nb_energy_only, whose
   * spote .. sprset are cm3d's vinter_e values (see there). */
  double spote_phys, spothe_phys, sprse_phys, sprset_phys[9];
} nb_energy;

int nb_create(const nb_params *p, nb_ctx **ctx);
void nb_destroy(nb_ctx *ctx);
const char *nb_last_error(void);

/* get_ngbr + get_ngbr_e on host positions; records the build positions. */
int nb_build_lists(nb_ctx *ctx, const double *pxyz, const double hmat[9]);

/* nrefsh criterion: *flag = 1 if any atom moved more than skin/2 since the
 * last build. */
int nb_need_rebuild(nb_ctx *ctx, const double *pxyz, int *flag);

/* force(...,iflg) minus rmass.  iflg 1: fl gets the full intermolecular
 * force (vdW + real-space coulomb + Ewald + ecorr), fs the RESPA-switched
 * part; iflg 0: fs gets the short-range force, fl is untouched.  fs/fl are
 * overwritten (cm3d zeroes them first, md_fcntr.c:55-63).  e may be NULL. */
int nb_force(nb_ctx *ctx, int iflg, const double *pxyz, const double hmat[9],
             double *fs, double *fl, nb_energy *e);

/* Energy-only evaluation: the energies and virials getvireal(...,0) sums
 * through getvirter / getvirter_e (vinter_npol, md_fnopol.c:257-383;
 * vinter_e, finter_e.c:260-388; over the current t / te lists), plus, for
 * iperd 3, the reciprocal-space and ecorr energies and virials fk_ewald /
 * ecorr would leave in COUL -- all of e's fields, no forces.  The vdW,
 * Ewald and ecorr values are those of nb_force(...,1), bit for bit.  The
 * coulomb fields spote .. sprset are cm3d's vinter_e values: a missing break
 * in its cutoff switch (finter_e.c:296-302) sets the table origin to
 * rmaxl_coul2, so every te pair inside the cutoff is clamped to the table
 * at k = 1, p = -1 (about coultab[0]); the force path's values are in
 * spote_phys .. sprset_phys. */
int nb_energy_only(nb_ctx *ctx, const double *pxyz, const double hmat[9], nb_energy *e);

/* Same on device-resident arrays (3N doubles in HBM) on the given HIP stream
 * (NULL = the context's stream).  fs may be NULL to skip the switched force. */
int nb_force_device(nb_ctx *ctx, int iflg, const double *d_pxyz,
                    const double hmat[9], double *d_fs, double *d_fl,
                    nb_energy *e, void *stream);
int nb_build_lists_device(nb_ctx *ctx, const double *d_pxyz,
                          const double hmat[9], void *stream);

/* nb_force_device that allocates nothing.  No counterpart in cm3d: it is for
 * callers that cannot allocate device memory where they evaluate forces
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The context (created beforehand) supplies the tables, lists
 * and records; the per-evaluation scratch comes from one static area per GPU
 * of 32 MiB, held until the process ends, shared by every call on that GPU.
 * A system whose scratch would not fit is refused with NB_ERR_CAPACITY
 * (about 17 MiB for 12,288 charges and 1535 k vectors; nb_force_device has
 * no such limit).  Calls that overlap on one GPU race on the shared area:
 * the caller serialises them (one stream, or waits between streams).
 * Results are bit for bit nb_force_device's. */
int nb_force_device_noalloc(nb_ctx *ctx, int iflg, const double *d_pxyz,
                            const double hmat[9], double *d_fs, double *d_fl,
                            nb_energy *e, void *stream);

/* Pair counts of the current lists as cm3d counts them (half pairs):
 * counts[0..3] = nprs, nprt, nprse, nprte. */
int nb_list_counts(nb_ctx *ctx, int counts[4]);
/* Device memory held by the context and any MD state attached to it, in
 * bytes (grows when a list build needs more neighbour capacity).  No
 * counterpart in cm3d: it is what a server admitting jobs by HBM footprint
 * accounts with. */
int nb_footprint(nb_ctx *ctx, size_t *bytes);

/* How the context evaluates the reciprocal sum (no counterpart in cm3d:
 * it reports which kernels a force call launches, for diagnostics and
 * measurement).  out[0]: bit 0 set when the structure factor runs on the
 * FP64 matrix cores, bit 1 when the reciprocal force pass does; out[1],
 * out[2]: matrix-core instructions (v_mfma_f64_16x16x4f64, 2048 FLOP each)
 * one launch of the structure factor / force pass issues (0 for the VALU
 * kernels); out[3]: 0. */
int nb_recip_info(nb_ctx *ctx, long long out[4]);

/* Neighbour slots per atom for the next list build (rounded up to 64).  No
 * counterpart in cm3d, whose get_ngbr sizes its lists by counting with a
 * x1.2 margin (energy/control/ngbr_list.c:331-335); here slots are sized
 * from the density and grown x1.25 when a build needs more.  A caller that
 * bounds list memory reserves them; the current lists are dropped (the next
 * device step rebuilds first, nb_force asks for nb_build_lists).  A host
 * build that needs more slots grows them; a device-side rebuild that needs
 * more is recovered by nb_batch_step and reported by nb_md_rebuilds. */
int nb_list_reserve(nb_ctx *ctx, int slots);

/* Copy one list as half pairs (i<j) in ascending (i,j) order.
 * which: 0 = s (vdW short), 1 = t (vdW), 2 = se, 3 = te.  Fails with
 * NB_ERR_CAPACITY (and sets *n) when cap is too small. */
int nb_get_pairs(nb_ctx *ctx, int which, int *ip, int *jp, int cap, int *n);

/* Per-phase timing of nb_force / nb_force_device with HIP events on the
 * stream the kernels run on (no counterpart in cm3d; diagnostics).
 * mode 1: start recording (up to 2048 calls), 0: stop, 2: read the averages
 * in ms: out = {calls, position pack, pair kernel(s), Ewald charge records,
 * structure factor, k reduction (and the matrix-core force pass's
 * coefficient image), reciprocal force, ecorr + force sum + energy
 * reductions, whole call}. */
#define NB_TIMING_N 9
int nb_timing(nb_ctx *ctx, int mode, double out[NB_TIMING_N]);

/* ---- device-resident MD step (cm3d's step(), integrate/md_step.c:62-159,
 * for nres_lrf == 1) ----
 * Velocity Verlet with the RESPA inner loop over harmonic bonds and bends
 * (energy/intra/md_fbond.c:76-128, md_fbend.c:126-260, harmonic cases),
 * nrefsh's rebuild test and the nonbonded force, all kept on the device.
 * Molecules are contiguous atom ranges of at most 8 atoms; bonded terms may
 * not cross molecules.  Units: dt in ps, masses in cm3d units (g/mol x
 * MCONV), bond k in K/A^2, bend k in K/rad^2, angles in radians. */
typedef struct nb_md nb_md;
typedef struct nb_md_params {
  const double *amass;                          /* [natoms]                */
  int nmol;
  const int *mol_start;                         /* [nmol+1]                */
  int nbond;
  const int *bond_i, *bond_j;
  const double *bond_k, *bond_r0;
  int nbend;
  const int *bend_i, *bend_j, *bend_k;          /* bend_j is the apex      */
  const double *bend_fk, *bend_th0;
  double dt;                                    /* outer step, ps          */
  int nres_intra;                               /* inner steps (ninnra)    */
  int nres_lrf;                                 /* must be 1               */
} nb_md_params;
typedef struct nb_md_energy {
  double vinter, vhinter, velec, vewald, vbond, vbend;
  double zke, poter, poth, potra;               /* geteng, md_fcntr.c:259  */
  double potn, zken;                            /* chain, getengn          */
  double ham;                                   /* poth+potra+zke+potn+zken */
  double winter, welec, wewald, wbond;          /* virials (getvireal)     */
  double prsi;                                  /* (p2m + W + WI) / (3 V)  */
  double prs_tensor[9];                         /* (p2m_t + W_t + WI_t) / V */
  double tiout;                                 /* p2m / ndof              */
  int overflow;                                 /* a list build overflowed */
} nb_md_energy;
int nb_md_create(nb_ctx *ctx, const nb_md_params *p, nb_md **md);
void nb_md_destroy(nb_md *md);
int nb_md_set_state(nb_md *md, const double *pos, const double *vel,
                    const double hmat[9]);
int nb_md_get_state(nb_md *md, double *pos, double *vel);
/* The forces of the current state as save_force reads them
 * (output/md_saveout.c:65-90): fl [3N] the nonbonded force before rmass
 * (cm3d's fxl times amass), aa [3N] the bonded accelerations (fxa); either
 * may be NULL. */
int nb_md_get_forces(nb_md *md, double *fl, double *aa);
/* step_init (md_step.c:32-58): list build, force(...,1), bonded forces. */
int nb_md_init(nb_md *md);
/* nsteps outer steps on the stream (NULL = the context's); use_graph
 * captures one step into a hipGraph once and replays it.  Asynchronous. */
int nb_md_step(nb_md *md, int nsteps, void *stream, int use_graph);
/* NVT: one Nose-Hoover chain of nchain links (cm3d's intnhc,
 * integrate/md_therm.new.c:278-332, applied at every inner RESPA substep as
 * md_step.c:113 does) over the atoms with ithm[i] == 0 (cm3d's "global"
 * option; -1 = not thermostatted).  gkt, mnh, veta0: [nchain] as set_therm /
 * samveta leave them (md_therm.new.c:158-169, 248-275); eta0 [nchain] or
 * NULL for zeros (a warm start reads both from the restart file,
 * md_readpos.c:243-262).  nchain 0 returns to NVE. */
int nb_md_set_thermostat(nb_md *md, int nchain, const double *gkt, const double *mnh,
                         const double *veta0, const double *eta0, const int *ithm);
/* The chain's current veta, eta ([nchain] each, either may be NULL), for
 * the restart file (writestart, output/md_writestart.c:67-73). */
int nb_md_get_thermostat(nb_md *md, double *veta, double *eta);
/* Energies, virials, pressure and temperature of the current state, as
 * getvireal + geteng leave them (md_fcntr.c:136-231, 259-306): vlrc, wlrc
 * are cm3d's simparms->vlrc / wlrc (applied when iperd == 3), ndof its
 * simparms->ndof (for tiout). */
int nb_md_energies(nb_md *md, double vlrc, double wlrc, int ndof, nb_md_energy *e);
/* Lists built so far by the device-side rebuild test (synchronises). */
int nb_md_rebuilds(nb_md *md, int *count);

/* ---- batches: many independent systems stepped together ----
 * No counterpart in cm3d (one system per process).  Each system is an
 * nb_ctx + nb_md created as above (same precision); the batch records one
 * step of each and launches every kernel once for all of them, so a step
 * costs the same number of launches for 1 or 1000 systems.  Systems keep
 * their own state and lists; nb_md_get_state / nb_md_energies still apply. */
typedef struct nb_batch nb_batch;
int nb_batch_create(nb_md **mds, int nsys, nb_batch **b);
void nb_batch_destroy(nb_batch *b);
/* nsteps steps of every system (one tick); returns when they are done.  The
 * tick snapshots each system's state first and checks after it whether any
 * of its device-side list builds wanted more neighbour slots than it has; a
 * system that did is restored, given the slots (x1.25), rebuilt and stepped
 * again alone, which is bitwise what the batch computes for it -- no system
 * continues on truncated lists.  A batch whose members were regrown or
 * re-thermostatted since it recorded its step records it again.
 * nsteps 0 only prepares the next tick: records the step again if a member
 * changed and, with use_graph, captures its graph, so that a caller can do
 * both on its own thread (a capture forbids legacy-stream work, such as
 * nb_create's copies, on other threads) and run ticks elsewhere. */
int nb_batch_step(nb_batch *b, int nsteps, void *stream, int use_graph);
/* Number of kernel launches one batched step issues. */
int nb_batch_launches(nb_batch *b, int *nkernels);
/* Systems redone after a list overflow since the batch was created. */
int nb_batch_recoveries(nb_batch *b, long long *count);

/* ---- set-up helpers (host only) ---- */
#define NB_POT_LJ 0
#define NB_POT_WILLIAMS 1
/* prm: LJ {eps, sig}; Williams {awill, bwill, c6, c8, c10}.  cut2 out:
 * rmins, rmaxs, rminl, rmaxl, rmaxs+skin, rmaxl+skin, rminl-skin squared. */
int nb_table_pair(int pot, const double *prm, double rmin, double rmax,
                  double rres, double rheal, double skin, double scaleeps,
                  int ntable, double *vtab, double *dvtab, double *switchs,
                  double cut2[7], double *dx2tab, double *vlrc, double *wlrc);
/* cut2 out: rmins, rmaxs, rminl, rmaxl, rmaxs+skin, rmaxl+skin squared. */
int nb_table_coul(int ntable, int iperd, double alpha, double rcute_min,
                  double rcute_resp, double rcute_max, double rheal,
                  double skin, int ishift, double *coultab, double *dcoultab,
                  double *switchse, double cut2[6], double *dx2coul);
/* Returns ktot; arrays may be NULL to count. */
int nb_kvectors(int kmax, int *ka, int *kb, int *kc);
int nb_ewald_consts(int natoms, const double *qch, double alpha,
                    double *self, double *bkg);

#ifdef __cplusplus
}
#endif
#endif
