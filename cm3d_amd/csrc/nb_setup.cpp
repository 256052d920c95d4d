// nb_setup.cpp -- host-side set-up helpers of the C ABI: the potential and
// coulomb tables, the half-space k-vector list and the Ewald constants that
// cm3d's input layer builds before the force path runs.  They exist so a
// caller without cm3d's deck reader can drive the engine; cm3d itself passes
// its own tables to nb_create.
//
// Restated algorithms (reference tree paths):
//   nb_table_pair   get_ter_parm, input/md_getparm.c:222-366 (RESPA split,
//                   grid dx2 = (ntable-3)/(rmax^2-rmin^2), tail corrections,
//                   table fill and smoothstep switch); func_lj / func_will,
//                   energy/inter/md_funcvdw.c:40-46,104-114; cut^2 values as
//                   search_ter_base stores them (input/md_searchbase.c:109-121)
//   nb_table_coul   set_charges, energy/control/elec_ngbr.c:47-103; func_kerf /
//                   func_coul, md_funcvdw.c:219-249
//   nb_kvectors     ewald_setup's integer sphere, energy/inter/md_ewald.c:93-114
//   nb_ewald_consts self and background terms, md_ewald.c:59-79
// The tables are evaluated with the C library's pow/exp/erfc exactly as cm3d
// does on the host, so a table built here equals cm3d's own.
#include <math.h>

#include "../../include/cm3d_nb.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static void pot_eval(int pot, const double *prm, double x, double *f, double *dfr) {
  if (pot == NB_POT_LJ) {
    const double eps = prm[0], sig = prm[1];
    *f = 4. * eps * (pow(sig / x, 12.) - pow(sig / x, 6.));
    *dfr = 4. * eps * (6. * pow(sig, 6.) * pow(x, -7.) - 12. * pow(sig, 12.) * pow(x, -13.));
  } else {
    const double a = prm[0], b = prm[1], c6 = prm[2], c8 = prm[3], c10 = prm[4];
    *f = a * exp(-b * x) - c10 * pow(x, -10.) - c8 * pow(x, -8.) - c6 * pow(x, -6.);
    *dfr = (-a * b * exp(-b * x) + 10. * c10 * pow(x, -11.) + 8. * c8 * pow(x, -9.) +
            6. * c6 * pow(x, -7.));
  }
  *dfr /= x;
}

extern "C" int nb_table_pair(int pot, const double *prm, double rmin, double rmax,
                             double rres, double rheal, double skin, double scaleeps,
                             int ntable, double *vtab, double *dvtab, double *switchs,
                             double cut2[7], double *dx2tab, double *vlrc, double *wlrc) {
  if (!prm || ntable < 4 || !(rmax > rmin) || (pot != NB_POT_LJ && pot != NB_POT_WILLIAMS))
    return NB_ERR_ARG;
  double rmins = rmin, rmaxs, rminl, rmaxl = rmax;
  if (rres < rmax) {
    rmaxs = rres;
    rminl = rres - rheal;
  } else {  // no RESPA split for this pair
    rmaxs = rmax;
    rminl = rmax + 2. * skin;
    rres = rmax + 2. * rheal;
  }
  if (rminl - skin < rmin) return NB_ERR_ARG;  // md_getparm.c:240-248
  const double dx2 = ((double)ntable - 3.) / (rmax * rmax - rmin * rmin);
  *dx2tab = dx2;
  const double rch = rres - rheal, rhl = (rheal == 0.) ? 1. : 1. / rheal;
  double vl, wl;
  if (pot == NB_POT_LJ) {
    const double eps = prm[0], sig = prm[1];
    vl = 4. * eps * pow(sig, 6.) *
         ((1. / 9.) * pow(sig, 6.) * pow(1. / rmax, 9.) - (1. / 3.) * pow(1. / rmax, 3.));
    wl = 8. * eps * pow(sig, 6.) * (pow(1. / rmax, 3.) - (2. / 3.) * pow(sig, 6.) * pow(1. / rmax, 9.));
  } else {
    const double a = prm[0], b = prm[1], c6 = prm[2], c8 = prm[3], c10 = prm[4], c9 = 0.;
    const double sm = b * rmax;
    vl = ((a * (sm * sm + 2 * sm + 2.) / pow(b, 3.)) * exp(-b * rmax) -
          c10 * pow(rmax, -7.) / 7. - c9 * pow(rmax, -6.) / 6. - c8 * pow(rmax, -5.) / 5. -
          c6 * pow(rmax, -3.) / 3.);
    wl = (-(a * exp(-b * rmax) * (sm * sm * sm + 3. * sm * sm + 6. * sm + 6.) / pow(b, 3.)) +
          c10 * pow(rmax, -7.) * 10. / 7. + c9 * pow(rmax, -6.) * 9. / 6. +
          c8 * pow(rmax, -5.) * 8. / 5. + c6 * pow(rmax, -3.) * 6. / 3.);
  }
  if (vlrc) *vlrc = vl * scaleeps;
  if (wlrc) *wlrc = wl * scaleeps;
  for (int i = 0; i < ntable; i++) {
    const double xt = i / dx2 + rmins * rmins;
    double f, df;
    pot_eval(pot, prm, sqrt(xt), &f, &df);
    vtab[i] = f * scaleeps;
    dvtab[i] = df * scaleeps;
    if (xt <= rch * rch) switchs[i] = 1.0;
    if (xt > rch * rch && xt < rres * rres) {
      const double br = (sqrt(xt) - rch) * rhl;
      switchs[i] = (1. + br * br * (2. * br - 3.));
    }
    if (xt >= rres * rres) switchs[i] = 0.0;
  }
  if (cut2) {
    cut2[0] = rmins * rmins;
    cut2[1] = rmaxs * rmaxs;
    cut2[2] = rminl * rminl;
    cut2[3] = rmaxl * rmaxl;
    cut2[4] = (rmaxs + skin) * (rmaxs + skin);
    cut2[5] = (rmaxl + skin) * (rmaxl + skin);
    cut2[6] = (rminl - skin) * (rminl - skin);
  }
  return NB_OK;
}

extern "C" int nb_table_coul(int ntable, int iperd, double alpha, double rcute_min,
                             double rcute_resp, double rcute_max, double rheal, double skin,
                             int ishift, double *coultab, double *dcoultab, double *switchse,
                             double cut2[6], double *dx2coul) {
  if (ntable < 4 || !(rcute_max > rcute_min)) return NB_ERR_ARG;
  const double rmins2 = rcute_min * rcute_min, rmaxs2 = rcute_resp * rcute_resp;
  const double rminl2 = (rcute_resp - rheal) * (rcute_resp - rheal);
  const double rmaxl2 = rcute_max * rcute_max;
  const double dx2 = (((double)ntable - 3.) / (rmaxl2 - rmins2));
  *dx2coul = dx2;
  if (cut2) {
    cut2[0] = rmins2;
    cut2[1] = rmaxs2;
    cut2[2] = rminl2;
    cut2[3] = rmaxl2;
    cut2[4] = (rcute_resp + skin) * (rcute_resp + skin);
    cut2[5] = (rcute_max + skin) * (rcute_max + skin);
  }
  for (int i = 0; i < ntable; i++) {
    const double xt = i / dx2 + rmins2;
    if (xt < 0) return NB_ERR_ARG;
    const double x = sqrt(xt);
    double f, df;
    if (iperd == 3) {
      f = erfc(alpha * x) / x;
      df = -erfc(alpha * x) / (x * x) - 2. * alpha * exp(-alpha * alpha * x * x) / (sqrt(M_PI) * x);
    } else {
      f = 1. / x;
      df = -1. / (x * x);
    }
    df /= x;
    coultab[i] = f;
    dcoultab[i] = df;
    if (xt <= rminl2) switchse[i] = 1.;
    if (xt > rminl2 && xt < rmaxs2) {
      const double br = (sqrt(xt) - sqrt(rminl2)) / rheal;
      switchse[i] = (1. + br * br * (2. * br - 3.));
    }
    if (xt >= rmaxs2) switchse[i] = 0.;
  }
  if (ishift) {
    const double v = coultab[ntable - 3];
    for (int i = 0; i < ntable; i++) coultab[i] -= v;
  }
  return NB_OK;
}

extern "C" int nb_kvectors(int kmax, int *ka, int *kb, int *kc) {
  if (kmax < 0) return NB_ERR_ARG;
  int n = 0;
  const int k2max = kmax * kmax;
  for (int i = 0; i <= kmax; i++)
    for (int j = (i == 0) ? 0 : -kmax; j <= kmax; j++)
      for (int k = (i == 0 && j == 0) ? 1 : -kmax; k <= kmax; k++)
        if (i * i + j * j + k * k <= k2max) {
          if (ka) { ka[n] = i; kb[n] = j; kc[n] = k; }
          n++;
        }
  return n;
}

extern "C" int nb_ewald_consts(int natoms, const double *qch, double alpha, double *self,
                               double *bkg) {
  if (natoms < 0 || (natoms > 0 && !qch) || !(alpha > 0)) return NB_ERR_ARG;
  double s = 0., b = 0.;
  int nq = 0;
  for (int i = 0; i < natoms; i++)
    if (qch[i] != 0.) { s += qch[i] * qch[i]; b += qch[i]; nq++; }
  if (nq == 0) { *self = 0.; *bkg = 0.; return NB_OK; }
  *self = s * (-alpha / sqrt(M_PI));
  *bkg = -.5 * b * b * M_PI / (alpha * alpha);
  return NB_OK;
}
