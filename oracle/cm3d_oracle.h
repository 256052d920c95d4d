/*
 * cm3d_oracle.h -- clean-room CPU restatement of cm3d's nonbonded force path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker (or as the
 * timed CPU baseline); the shipped path never calls it.
 *
 * Every function restates the algorithm of the cm3d routine it cites
 * (paths relative to the reference tree).  It follows the reference's
 * operation order (the -DLINUX truncating anint, no FMA, the same chunking
 * of partial energy sums) so that, built with gcc -O2 like cm3d itself, its
 * lists are identical and its forces agree bit for bit with the compiled
 * reference on the same inputs; tests/test_oracle_*.py pin that against the
 * fixtures written by oracle/ref_dump.c.
 */
#ifndef CM3D_ORACLE_H
#define CM3D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Everything the nonbonded path reads, flat (cm3d's INTER/COUL/COORDS
 * fields, md_typedef.h:59-69,201-268). */
typedef struct orc_sys {
  int natoms, ntypes, ninter, ntable, iperd, ncell, nlen;
  const int *itype;             /* [natoms]                               */
  const int *map;               /* [ntypes*ntypes], -1 = null pair        */
  const double *vtab, *dvtab, *switchs;           /* [ninter*ntable]     */
  const double *rmins_cut2, *rmaxs_cut2, *rminl_cut2, *rmaxl_cut2;
  const double *dx2tab, *rmaxs_skin2, *rmaxl_skin2;  /* [ninter]         */
  double skin;
  const int *nexclude, *exclude;   /* vdW exclusions, flattened, j>i     */
  const double *qch;               /* [natoms] in sqrt(K A)              */
  const double *coultab, *dcoultab, *switchse;       /* [ntable]         */
  double dx2coul, rmins_coul2, rmaxs_coul2, rminl_coul2, rmaxl_coul2;
  double rmaxs_skin2_e, rmaxl_skin2_e;
  const int *cnexclude, *cexclude; /* coulomb exclusions, flattened      */
  int necorr;
  const int *iecorr, *jecorr;
  int ktot;
  const int *ka, *kb, *kc, *kup;
  double kappa, ewald_self, ewald_bkg;
} orc_sys;

typedef struct orc_lists {
  int nprt, nprs, nprte, nprse;
  int *inpt, *jnpt, *inps, *jnps, *inpte, *jnpte, *inpse, *jnpse;
} orc_lists;

typedef struct orc_energy {
  double spotl, spothl, sprsl, sprslt[9];
  double spote, spothe, sprse, sprset[9];
  double vk_ewald, vp_ewald, vp_ewald_tensor[9];
  double pot_ecorr, prs_ecorr, prs_ecorr_tensor[9];
  int nclamp;
} orc_energy;

/* ---- set-up (input/md_getparm.c, input/md_searchbase.c,
 *      energy/control/elec_ngbr.c, energy/inter/md_ewald.c) ---- */
enum { ORC_POT_LJ = 0, ORC_POT_WILLIAMS = 1 };
void orc_table_pair(int pot, const double *prm, double rmin, double rmax,
                    double rres, double rheal, double skin, double scaleeps,
                    int ntable, double *vtab, double *dvtab, double *switchs,
                    double cut2[7], double *dx2tab, double *vlrc,
                    double *wlrc);
void orc_table_shift(int ntable, double *vtab);
void orc_table_coul(int ntable, int iperd, double alpha, double rcute_min,
                    double rcute_resp, double rcute_max, double rheal,
                    double skin, int ishift, double *coultab,
                    double *dcoultab, double *switchse, double cut2[6],
                    double *dx2coul);
int orc_kvectors(int kmax, int *ka, int *kb, int *kc, int *kup);
void orc_ewald_consts(int natoms, const double *qch, double alpha,
                      double *self, double *bkg);
void orc_lrc_total(int natoms, int ntypes, const int *itype, const int *map,
                   const double *vlrc, const double *wlrc, double *vtot,
                   double *wtot);

/* ---- lists (energy/control/ngbr_list.c, elec_ngbr.c, set_link.c) ---- */
int orc_build_lists(const orc_sys *s, const double *pos, const double *hmat,
                    int use_cells, orc_lists *L);
void orc_free_lists(orc_lists *L);
int orc_need_rebuild(int natoms, double skin, const double *pos,
                     const double *pos_build);

/* ---- forces (energy/control/md_fcntr.c force(...,iflg) minus rmass) ---- */
void orc_force(const orc_sys *s, const orc_lists *L, int iflg,
               const double *pos, const double *hmat, double *fs,
               double *fl, orc_energy *e);

/* ---- independent checks that do not share the table path ---- */
double orc_ewald_recip_direct(const orc_sys *s, const double *pos,
                              const double *hmat);
double orc_coul_real_direct(const orc_sys *s, const double *pos,
                            const double *hmat, double rcut);
double orc_lj_direct(const orc_sys *s, const double *pos, const double *hmat,
                     const double *eps, const double *sig, double *fx);

/* ---- bonded water terms and cm3d's RESPA step (integrate/md_step.c) ---- */
typedef struct orc_bonded {
  int nbond, nbend;
  const int *ib, *jb;            /* bonds                                 */
  const double *bk, *beq;        /* V = 1/2 k (r-r0)^2                    */
  const int *ia, *ja, *ka;       /* bends, ja is the centre               */
  const double *ak, *aeq;        /* V = 1/2 k (theta-theta0)^2, radians   */
} orc_bonded;
/* parts (may be NULL) gets {vbond, vbend}, summed separately as getvbond /
 * getvbend do; the return value is their sum, cm3d's potra for this model. */
double orc_bonded_force(int natoms, const orc_bonded *b, const double *pos,
                        double *f, double parts[2]);
#ifdef __cplusplus
}
#endif
#endif
