/*
 * ref_dump.c -- dump harness around the compiled cm3d reference.
 *
 * Test infrastructure only.  Linked against cm3d's own objects (built by
 * `make ref` from the sources under the reference tree, never copied), it
 * reads a cm3d deck exactly as the cm3d executable does and writes, in a
 * small tagged binary format, everything the nonbonded path produces:
 *
 *   - the Verlet lists after get_ngbr (t, s, te, se), as cm3d orders them;
 *   - fxl / fxs after force(...,1) *before* the rmass division (masses are set
 *     to 1.0 around the call so rmass is the identity; freeze_atom is held
 *     off around it);
 *   - the energy / virial by-products read by getvireal(...,1);
 *   - the tables, cutoffs, k-vectors and exclusion lists built at set-up;
 *   - optionally an nstep trajectory of cm3d's own step() (RESPA + bonded).
 *
 * With npre > 0 the snapshot is not taken on the deck's own configuration
 * but after npre steps of cm3d's step() from it (cold-start velocities
 * included), so that a fixture can hold a thermalised state; the deck's
 * positions are then dumped as pos_deck and the trajectory's step count
 * continues from npre (nrefsh keeps the step of the last rebuild in a
 * static, md_finter.c:48).
 *
 * Call sites it drives (reference file:line):
 *   md_startup   main/md_startup.new.c:31
 *   readpos      input/md_readpos.c:32
 *   get_ngbr     energy/control/ngbr_list.c:297
 *   force        energy/control/md_fcntr.c:44
 *   getvireal    energy/control/md_fcntr.c:136 (iflg 1, and iflg 0 for the
 *                energy-only sums)
 *   step_init    integrate/md_step.c:32,  step integrate/md_step.c:62
 *   geteng       energy/control/md_fcntr.c:259, getengn integrate/md_therm.c
 *
 * usage: cm3d_ref_dump deck out.bin [nsteps [npre]]
 */
#include "md.h"
#include <stdint.h>

static FILE *fout;

/* main/md_main.c (which also holds cm3d's main) is not linked; its one other
   symbol, used by main_new in the same object as args_filename_new, is this
   forwarder. */
void args_filename(FILENAMES *filenames, int argc, char *argv[])
{
  args_filename_new(filenames, argc, argv);
}

static void put(const char *name, char t, long n, const void *p)
{
  int32_t len = (int32_t)strlen(name);
  int64_t cnt = n;
  fwrite(&len, 4, 1, fout);
  fwrite(name, 1, len, fout);
  fwrite(&t, 1, 1, fout);
  fwrite(&cnt, 8, 1, fout);
  if (n > 0) fwrite(p, t == 'd' ? 8 : 4, n, fout);
}
static void putd(const char *name, double v) { put(name, 'd', 1, &v); }
static void puti(const char *name, int v) { put(name, 'i', 1, &v); }

static void dump_energies(const char *pfx, ENERGIES *e)
{
  char nm[64];
#define D1(f) sprintf(nm, "%s" #f, pfx); putd(nm, e->f)
#define D9(f) sprintf(nm, "%s" #f, pfx); put(nm, 'd', 9, e->f)
  D1(vinter); D1(vhinter); D1(winter); D9(wintertensor);
  D1(velec); D1(welec); D9(welectensor);
  D1(vewald); D1(wewald); D9(wewaldtensor);
  D1(vbond); D1(vbend); D1(wbond); D1(wbend);
  D1(US); D1(U); D1(W); D9(Wtensor);
#undef D1
#undef D9
}

int main(int argc, char *argv[])
{
  static FILENAMES filenames;
  static SIMPARMS simparms;
  static SUBSPACE subspace;
  static WRITE_STEP write_step;
  static NGBR ngbr;
  static INTER inter;
  static COORDS coords;
  static ENERGIES energies;
  char *fargv[2];
  int i, n3, nsteps = 0, npre = 0, ninter;
  double *msave;

  if (argc < 3) {
    fprintf(stderr, "usage: %s deck out.bin [nsteps [npre]]\n", argv[0]);
    return 2;
  }
  if (argc > 3) nsteps = atoi(argv[3]);
  if (argc > 4) npre = atoi(argv[4]);
  fout = fopen(argv[2], "wb");
  if (!fout) { perror("out"); return 2; }

  simparms.mem_bytes = 0;
  simparms.rank = 0;
  simparms.size = 1;
  fargv[0] = argv[0];
  fargv[1] = argv[1];
  args_filename_new(&filenames, 2, fargv);
  md_startup(&filenames, &simparms, &subspace, &write_step, &inter, &ngbr,
             &coords);
  readpos(argv[0], filenames.initfile, &simparms, &energies.estep, &energies,
          &coords);
  if (simparms.istart == 1) {
    samvel(&simparms, &coords, 1);
    samveta(&simparms, &coords);
  }
  gethinv9(coords.hmat, coords.hmati);

  n3 = 3 * simparms.natoms;
  ninter = inter.ntype;
  if (npre > 0) {
    int s;
    put("pos_deck", 'd', n3, coords.px);
    step_init(&simparms, &coords, &energies, &inter, &ngbr);
    for (s = 1; s <= npre; s++) {
      simparms.istep = s;
      step(&simparms, &coords, &energies, &inter, &ngbr);
    }
    puti("npre", npre);
  }
  puti("natoms", simparms.natoms);
  puti("ntypes", simparms.ntypes);
  puti("ninter", ninter);
  puti("ntable", inter.ntable);
  puti("ncell", ngbr.ncells);
  puti("nlen", simparms.nlen);
  puti("iperd", simparms.iperd);
  put("hmat", 'd', 9, coords.hmat);
  put("pos", 'd', n3, coords.px);
  put("vel", 'd', n3, coords.vx);
  puti("ndof", simparms.ndof);
  puti("ntherm", simparms.ntherm);
  if (simparms.ntherm > 0) {
    /* the chain after set_therm + samveta (md_therm.new.c:158-169, 248) */
    puti("nchain", simparms.nchain);
    put("ithm", 'i', simparms.natoms, coords.ithm);
    put("gkt", 'd', simparms.nchain, coords.gkt[0]);
    put("mnh", 'd', simparms.nchain, coords.mnh[0]);
    put("veta0", 'd', simparms.nchain, coords.veta[0]);
    put("eta0", 'd', simparms.nchain, coords.eta[0]);
  }
  put("qch", 'd', simparms.natoms, coords.qch);
  put("amass", 'd', simparms.natoms, coords.amass);
  put("itype", 'i', simparms.natoms, inter.itype);
  put("map", 'i', simparms.ntypes * simparms.ntypes, inter.map[0]);
  putd("skin", inter.skin_ter);
  putd("rheal", inter.rheal);
  putd("vlrc", simparms.vlrc);
  putd("wlrc", simparms.wlrc);
  put("vtab", 'd', (long)ninter * inter.ntable, inter.vtab[0]);
  put("dvtab", 'd', (long)ninter * inter.ntable, inter.dvtab[0]);
  put("switchs", 'd', (long)ninter * inter.ntable, inter.switchs[0]);
  put("rmins_cut2", 'd', ninter, inter.rmins_cut2);
  put("rmaxs_cut2", 'd', ninter, inter.rmaxs_cut2);
  put("rminl_cut2", 'd', ninter, inter.rminl_cut2);
  put("rmaxl_cut2", 'd', ninter, inter.rmaxl_cut2);
  put("rmaxs_skin2", 'd', ninter, inter.rmaxs_skin2);
  put("rmaxl_skin2", 'd', ninter, inter.rmaxl_skin2);
  put("dx2tab", 'd', ninter, inter.dx2tab);
  if (coords.coul.ncharge > 0 || simparms.iperd == 3) {
    put("coultab", 'd', inter.ntable, inter.coultab);
    put("dcoultab", 'd', inter.ntable, inter.dcoultab);
    put("switchse", 'd', inter.ntable, inter.switchse);
  }
  putd("dx2coul", inter.dx2coul);
  putd("rmins_coul2", inter.rmins_coul2);
  putd("rmaxs_coul2", inter.rmaxs_coul2);
  putd("rminl_coul2", inter.rminl_coul2);
  putd("rmaxl_coul2", inter.rmaxl_coul2);
  putd("rmaxs_skin2_e", inter.rmaxs_skin2_e);
  putd("rmaxl_skin2_e", inter.rmaxl_skin2_e);
  puti("ncharge", coords.coul.ncharge);
  if (coords.coul.ncharge > 0)
    put("icharge", 'i', coords.coul.ncharge, coords.coul.icharge);
  puti("ktot", coords.coul.ktot);
  if (coords.coul.ktot > 0) {
    put("ka", 'i', coords.coul.ktot, coords.coul.ka);
    put("kb", 'i', coords.coul.ktot, coords.coul.kb);
    put("kc", 'i', coords.coul.ktot, coords.coul.kc);
    put("kup", 'i', coords.coul.ktot + 1, coords.coul.kup);
  }
  putd("kappa", coords.coul.kappa);
  putd("ewald_self", coords.coul.ewald_self);
  putd("ewald_bkg", coords.coul.ewald_bkg);
  puti("necorr", coords.coul.necorr);
  if (coords.coul.necorr > 0) {
    put("iecorr", 'i', coords.coul.necorr, coords.coul.iecorr);
    put("jecorr", 'i', coords.coul.necorr, coords.coul.jecorr);
  }
  /* vdW exclusion lists, flattened */
  {
    long tot = 0, k = 0;
    int *flat;
    for (i = 0; i < simparms.natoms; i++) tot += inter.nexclude[i];
    flat = (int *)malloc((tot + 1) * sizeof(int));
    for (i = 0; i < simparms.natoms; i++) {
      int j;
      for (j = 0; j < inter.nexclude[i]; j++) flat[k++] = inter.exclude[i][j];
    }
    put("nexclude", 'i', simparms.natoms, inter.nexclude);
    put("exclude", 'i', tot, flat);
    free(flat);
  }
  /* coulomb exclusion lists (charged atoms only), flattened */
  {
    long tot = 0, k = 0;
    int *flat;
    for (i = 0; i < simparms.natoms; i++) tot += coords.coul.nexclude[i];
    flat = (int *)malloc((tot + 1) * sizeof(int));
    for (i = 0; i < simparms.natoms; i++) {
      int j;
      for (j = 0; j < coords.coul.nexclude[i]; j++)
        flat[k++] = coords.coul.exclude[i][j];
    }
    put("cnexclude", 'i', simparms.natoms, coords.coul.nexclude);
    put("cexclude", 'i', tot, flat);
    free(flat);
  }

  /* ---- snapshot: lists, forces, energies ---- */
  ngbr.update = -1;
  get_ngbr(&simparms, &coords, &inter, &ngbr);
  put("inpt", 'i', ngbr.nprt, ngbr.inpt);
  put("jnpt", 'i', ngbr.nprt, ngbr.jnpt);
  put("inps", 'i', ngbr.nprs, ngbr.inps);
  put("jnps", 'i', ngbr.nprs, ngbr.jnps);
  put("inpte", 'i', ngbr.nprte, ngbr.inpte);
  put("jnpte", 'i', ngbr.nprte, ngbr.jnpte);
  put("inpse", 'i', ngbr.nprse, ngbr.inpse);
  put("jnpse", 'i', ngbr.nprse, ngbr.jnpse);

  msave = (double *)malloc(simparms.natoms * sizeof(double));
  for (i = 0; i < simparms.natoms; i++) {
    msave[i] = coords.amass[i];
    coords.amass[i] = 1.0;
  }
  {
    /* freeze_atom (md_fcntr.c:310) post-processes force() for frozen atoms;
       it is the caller's step, not the nonbonded path, so it is held off. */
    int nfz = simparms.nfreeze;
    simparms.nfreeze = 0;
    force(&simparms, &coords, &inter, &ngbr, 1);
    simparms.nfreeze = nfz;
  }
  for (i = 0; i < simparms.natoms; i++) coords.amass[i] = msave[i];
  put("fxl", 'd', n3, coords.fxl);
  put("fxs", 'd', n3, coords.fxs);
  getvireal(&simparms, &coords, &energies, &inter, &ngbr, 1);
  dump_energies("e_", &energies);
  putd("vk_ewald", coords.coul.vk_ewald);
  putd("vp_ewald", coords.coul.vp_ewald);
  put("vp_ewald_tensor", 'd', 9, coords.coul.vp_ewald_tensor);
  putd("pot_ecorr", coords.coul.pot_ecorr);
  putd("prs_ecorr", coords.coul.prs_ecorr);
  put("prs_ecorr_tensor", 'd', 9, coords.coul.prs_ecorr_tensor);
  /* the energy-only sums (getvirter / getvirter_e, vinter_npol / vinter_e)
     of getvireal(...,0) on the same lists and positions */
  puti("ilist", ngbr.ilist);
  getvireal(&simparms, &coords, &energies, &inter, &ngbr, 0);
  dump_energies("e0_", &energies);

  /* ---- optional trajectory through cm3d's own integrator ---- */
  if (nsteps > 0) {
    int s;
    char nm[64];
    put("traj_pos0", 'd', n3, coords.px);
    put("traj_vel0", 'd', n3, coords.vx);
    puti("ninter_respa", simparms.ninter);
    puti("ninnra", simparms.ninnra);
    puti("ninrra", simparms.ninrra);
    putd("dt", simparms.dt);
    step_init(&simparms, &coords, &energies, &inter, &ngbr);
    geteng(&simparms, &coords, &energies);
    putd("traj_poter0", energies.poter);
    putd("traj_zke0", energies.zke);
    putd("traj_poth0", energies.poth);
    putd("traj_potra0", energies.potra);
    putd("traj_prsi0", energies.prsi);
    putd("traj_tiout0", energies.tiout);
    for (s = 1; s <= nsteps; s++) {
      simparms.istep = npre + s;
      step(&simparms, &coords, &energies, &inter, &ngbr);
      geteng(&simparms, &coords, &energies);
      sprintf(nm, "traj_pos%d", s); put(nm, 'd', n3, coords.px);
      sprintf(nm, "traj_vel%d", s); put(nm, 'd', n3, coords.vx);
      sprintf(nm, "traj_poter%d", s); putd(nm, energies.poter);
      sprintf(nm, "traj_potra%d", s); putd(nm, energies.potra);
      sprintf(nm, "traj_zke%d", s); putd(nm, energies.zke);
      sprintf(nm, "traj_nupdate%d", s); puti(nm, ngbr.update);
      sprintf(nm, "traj_poth%d", s); putd(nm, energies.poth);
      sprintf(nm, "traj_prsi%d", s); putd(nm, energies.prsi);
      sprintf(nm, "traj_tiout%d", s); putd(nm, energies.tiout);
      if (simparms.ntherm > 0) {
        double zken, potn;
        getengn(&simparms, &coords, &zken, &potn);
        sprintf(nm, "traj_veta%d", s); put(nm, 'd', simparms.nchain, coords.veta[0]);
        sprintf(nm, "traj_eta%d", s); put(nm, 'd', simparms.nchain, coords.eta[0]);
        sprintf(nm, "traj_zken%d", s); putd(nm, zken);
        sprintf(nm, "traj_potn%d", s); putd(nm, potn);
      }
    }
    puti("traj_nsteps", nsteps);
  }
  fclose(fout);
  free(msave);
  return 0;
}
