/*
 * cm3d_oracle.c -- clean-room CPU restatement of cm3d's nonbonded force path.
 *
 * TEST INFRASTRUCTURE (the parity oracle and the CPU baseline).  Nothing in
 * the shipped library calls this file.
 *
 * Each function names the cm3d routine whose algorithm it restates.  The
 * restatement keeps the reference's arithmetic order wherever rounding is
 * observable (minimum image, r^2, table interpolation, per-chunk partial
 * energy sums, the k-space recursion), because the tests compare it to the
 * compiled reference bit for bit.  Build: gcc -O2 -ffp-contract=off.
 */
#include "cm3d_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* cm3d's -DLINUX anint: round half away from zero by truncation
 * (include/md_macros.h:31-52). */
static inline int orc_anint(double a) { return (int)(a + (a >= 0 ? .5 : -.5)); }

/* Inverse of the 3x3 box matrix by cofactors; returns the determinant
 * (utils/md_mathlib.c:30-50, gethinv9). */
static double orc_hinv(const double h[9], double hi[9])
{
  double c0 = h[4] * h[8] - h[5] * h[7];
  double c1 = h[5] * h[6] - h[3] * h[8];
  double c2 = h[3] * h[7] - h[4] * h[6];
  double rvol = 1. / (h[0] * c0 + h[1] * c1 + h[2] * c2);
  hi[0] = c0 * rvol;
  hi[1] = (h[2] * h[7] - h[1] * h[8]) * rvol;
  hi[2] = (h[1] * h[5] - h[2] * h[4]) * rvol;
  hi[3] = c1 * rvol;
  hi[4] = (h[0] * h[8] - h[2] * h[6]) * rvol;
  hi[5] = (h[2] * h[3] - h[0] * h[5]) * rvol;
  hi[6] = c2 * rvol;
  hi[7] = (h[1] * h[6] - h[0] * h[7]) * rvol;
  hi[8] = (h[0] * h[4] - h[1] * h[3]) * rvol;
  return 1. / rvol;
}

static double orc_det(const double h[9])
{
  return h[0] * (h[4] * h[8] - h[5] * h[7]) + h[1] * (h[5] * h[6] - h[3] * h[8]) +
         h[2] * (h[3] * h[7] - h[4] * h[6]);
}

/* Minimum image of one displacement, iperd==3 branch of period_new
 * (integrate/md_period.new.c:39-51): s = hi.d, s -= anint(s), d = h.s. */
static inline void orc_minimg(double *dx, double *dy, double *dz,
                              const double h[9], const double hi[9])
{
  double sx = *dx * hi[0] + *dy * hi[1] + *dz * hi[2];
  double sy = *dx * hi[3] + *dy * hi[4] + *dz * hi[5];
  double sz = *dx * hi[6] + *dy * hi[7] + *dz * hi[8];
  sx -= orc_anint(sx);
  sy -= orc_anint(sy);
  sz -= orc_anint(sz);
  *dx = sx * h[0] + sy * h[1] + sz * h[2];
  *dy = sx * h[3] + sy * h[4] + sz * h[5];
  *dz = sx * h[6] + sy * h[7] + sz * h[8];
}

/* 3-point quadratic interpolation in r^2 (energy/inter/md_fnopol.c:146). */
static inline double orc_interp(const double *t, int k, double p)
{
  double fm1 = t[k - 1], f0 = t[k], f1 = t[k + 1];
  return f0 + .5 * p * (f1 - fm1 + p * (fm1 - 2. * f0 + f1));
}

/* ======================================================================
 * Set-up: potential tables, coulomb table, k-vectors, LRC
 * ==================================================================== */

/* V, V'/r of the pair potentials (energy/inter/md_funcvdw.c:40-46 func_lj,
 * :104-114 func_will). */
static void orc_pot(int pot, const double *prm, double x, double *f,
                    double *dfr)
{
  if (pot == ORC_POT_LJ) {
    double eps = prm[0], sig = prm[1];
    *f = 4. * eps * (pow(sig / x, 12.) - pow(sig / x, 6.));
    *dfr = 4. * eps * (6. * pow(sig, 6.) * pow(x, -7.) -
                       12. * pow(sig, 12.) * pow(x, -13.));
  } else {
    double a = prm[0], b = prm[1], c6 = prm[2], c8 = prm[3], c10 = prm[4];
    *f = a * exp(-b * x) - c10 * pow(x, -10.) - c8 * pow(x, -8.) -
         c6 * pow(x, -6.);
    *dfr = (-a * b * exp(-b * x) + 10. * c10 * pow(x, -11.) +
            8. * c8 * pow(x, -9.) + 6. * c6 * pow(x, -7.));
  }
  *dfr /= x;
}

/* One interaction's table, cutoffs and tail correction
 * (input/md_getparm.c:222-235 RESPA split, :236-237 grid, :259-266 LJ tail,
 *  :269-281 Williams tail, :321-366 fill; input/md_searchbase.c:109-121 cut^2).
 * prm: LJ {eps, sig}; Williams {awill, bwill, c6, c8, c10} (c9 unused in V).
 * cut2 out: rmins, rmaxs, rminl, rmaxl, rmaxs_skin, rmaxl_skin, rminl_skin. */
void orc_table_pair(int pot, const double *prm, double rmin, double rmax,
                    double rres, double rheal, double skin, double scaleeps,
                    int ntable, double *vtab, double *dvtab, double *switchs,
                    double cut2[7], double *dx2tab, double *vlrc,
                    double *wlrc)
{
  double rmins, rmaxs, rminl, rmaxl, rch, rhl, dx2;
  int i;
  if (rres < rmax) {
    rmins = rmin; rmaxs = rres; rminl = rres - rheal; rmaxl = rmax;
  } else {
    rmins = rmin; rmaxs = rmax; rmaxl = rmax; rminl = rmax + 2. * skin;
    rres = rmax + 2. * rheal;
  }
  dx2 = ((double)ntable - 3.) / (rmax * rmax - rmin * rmin);
  *dx2tab = dx2;
  rch = rres - rheal;
  rhl = (rheal == 0.) ? 1. : 1. / rheal;
  *vlrc = *wlrc = 0.;
  if (pot == ORC_POT_LJ) {
    double eps = prm[0], sig = prm[1];
    *vlrc = 4. * eps * pow(sig, 6.) *
            ((1. / 9.) * pow(sig, 6.) * pow(1. / rmax, 9.) -
             (1. / 3.) * pow(1. / rmax, 3.));
    *wlrc = 8. * eps * pow(sig, 6.) *
            (pow(1. / rmax, 3.) - (2. / 3.) * pow(sig, 6.) * pow(1. / rmax, 9.));
  } else {
    double a = prm[0], b = prm[1], c6 = prm[2], c8 = prm[3], c10 = prm[4];
    double c9 = 0., sm = b * rmax;
    *vlrc = ((a * (sm * sm + 2 * sm + 2.) / pow(b, 3.)) * exp(-b * rmax) -
             c10 * pow(rmax, -7.) / 7. - c9 * pow(rmax, -6.) / 6. -
             c8 * pow(rmax, -5.) / 5. - c6 * pow(rmax, -3.) / 3.);
    *wlrc = (-(a * exp(-b * rmax) * (sm * sm * sm + 3. * sm * sm + 6. * sm + 6.) /
               pow(b, 3.)) +
             c10 * pow(rmax, -7.) * 10. / 7. + c9 * pow(rmax, -6.) * 9. / 6. +
             c8 * pow(rmax, -5.) * 8. / 5. + c6 * pow(rmax, -3.) * 6. / 3.);
  }
  *vlrc *= scaleeps;
  *wlrc *= scaleeps;
  for (i = 0; i < ntable; i++) {
    double xt = i / dx2 + rmins * rmins, f, df, br, sm;
    orc_pot(pot, prm, sqrt(xt), &f, &df);
    vtab[i] = f * scaleeps;
    dvtab[i] = df * scaleeps;
    if (xt <= rch * rch) switchs[i] = 1.0;
    if (xt > rch * rch && xt < rres * rres) {
      br = (sqrt(xt) - rch) * rhl;
      sm = (1. + br * br * (2. * br - 3.));
      switchs[i] = sm;
    }
    if (xt >= rres * rres) switchs[i] = 0.0;
  }
  cut2[0] = rmins * rmins;
  cut2[1] = rmaxs * rmaxs;
  cut2[2] = rminl * rminl;
  cut2[3] = rmaxl * rmaxl;
  cut2[4] = (rmaxs + skin) * (rmaxs + skin);
  cut2[5] = (rmaxl + skin) * (rmaxl + skin);
  cut2[6] = (rminl - skin) * (rminl - skin);
}

/* \shift{on}: subtract the table value at ntable-3 (md_searchbase.c:181-188). */
void orc_table_shift(int ntable, double *vtab)
{
  double v = vtab[ntable - 3];
  int i;
  for (i = 0; i < ntable; i++) vtab[i] -= v;
}

/* erfc(a r)/r table for Ewald real space, or 1/r without periodicity
 * (energy/control/elec_ngbr.c:47-103; func_kerf md_funcvdw.c:219-227,
 *  func_coul :243-249).  cut2 out: rmins, rmaxs, rminl, rmaxl, rmaxs_skin,
 *  rmaxl_skin. */
void orc_table_coul(int ntable, int iperd, double alpha, double rcute_min,
                    double rcute_resp, double rcute_max, double rheal,
                    double skin, int ishift, double *coultab,
                    double *dcoultab, double *switchse, double cut2[6],
                    double *dx2coul)
{
  double rmins2 = rcute_min * rcute_min, rmaxs2 = rcute_resp * rcute_resp;
  double rminl2 = (rcute_resp - rheal) * (rcute_resp - rheal);
  double rmaxl2 = rcute_max * rcute_max, dx2;
  int i;
  cut2[0] = rmins2; cut2[1] = rmaxs2; cut2[2] = rminl2; cut2[3] = rmaxl2;
  cut2[4] = (rcute_resp + skin) * (rcute_resp + skin);
  cut2[5] = (rcute_max + skin) * (rcute_max + skin);
  dx2 = (((double)ntable - 3.) / (rmaxl2 - rmins2));
  *dx2coul = dx2;
  for (i = 0; i < ntable; i++) {
    double xt = i / dx2 + rmins2, x = sqrt(xt), f, df, br, sm;
    if (iperd == 3) {
      f = erfc(alpha * x) / x;
      df = -erfc(alpha * x) / (x * x) -
           2. * alpha * exp(-alpha * alpha * x * x) / (sqrt(M_PI) * x);
    } else {
      f = 1. / x;
      df = -1. / (x * x);
    }
    df /= x;
    coultab[i] = f;
    dcoultab[i] = df;
    if (xt <= rminl2) switchse[i] = 1.;
    if (xt > rminl2 && xt < rmaxs2) {
      br = (sqrt(xt) - sqrt(rminl2)) / rheal;
      sm = (1. + br * br * (2. * br - 3.));
      switchse[i] = sm;
    }
    if (xt >= rmaxs2) switchse[i] = 0.;
  }
  if (ishift) {
    double v = coultab[ntable - 3];
    for (i = 0; i < ntable; i++) coultab[i] -= v;
  }
}

/* Half-space integer k-sphere, i>=0, kup[] marks the first k of every (i,j)
 * row so the c-axis recursion restarts there (energy/inter/md_ewald.c:93-114).
 * Arrays may be NULL to count only. */
int orc_kvectors(int kmax, int *ka, int *kb, int *kc, int *kup)
{
  int i, j, k, n = 0, k2max = kmax * kmax;
  for (i = 0; i <= kmax; i++) {
    for (j = (i == 0) ? 0 : -kmax; j <= kmax; j++) {
      if (kup) kup[n] = 1;
      for (k = (i == 0 && j == 0) ? 1 : -kmax; k <= kmax; k++) {
        if (i * i + j * j + k * k <= k2max) {
          if (ka) { ka[n] = i; kb[n] = j; kc[n] = k; }
          if (kup) kup[n + 1] = 0;
          n++;
        }
      }
    }
  }
  return n;
}

/* Self and neutralising-background constants (md_ewald.c:59-79). */
void orc_ewald_consts(int natoms, const double *qch, double alpha,
                      double *self, double *bkg)
{
  double s = 0., b = 0.;
  int i, nq = 0;
  for (i = 0; i < natoms; i++) {
    if (qch[i] != 0.) { s += qch[i] * qch[i]; b += qch[i]; nq++; }
  }
  if (nq == 0) { *self = 0.; *bkg = 0.; return; }
  *self = s * (-alpha / sqrt(M_PI));
  *bkg = -.5 * b * b * M_PI / (alpha * alpha);
}

/* Sum of per-pair tail corrections over type pairs (md_searchbase.c:212-224). */
void orc_lrc_total(int natoms, int ntypes, const int *itype, const int *map,
                   const double *vlrc, const double *wlrc, double *vtot,
                   double *wtot)
{
  int i, j, k;
  double v = 0., w = 0.;
  for (i = 0; i < ntypes; i++) {
    int ni = 0;
    for (k = 0; k < natoms; k++) if (itype[k] == i) ni++;
    for (j = 0; j < ntypes; j++) {
      int nj = 0, l = map[i * ntypes + j];
      for (k = 0; k < natoms; k++) if (itype[k] == j) nj++;
      if (l != -1) v += ni * nj * M_PI * 2. * vlrc[l];
      if (l != -1) w += ni * nj * M_PI * 2. * wlrc[l] / 3.;
    }
  }
  *vtot = v;
  *wtot = w;
}

/* ======================================================================
 * Neighbor lists
 * ==================================================================== */

typedef struct { int n, cap; int *i, *j; } pairvec;

static void pv_push(pairvec *v, int a, int b)
{
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 1024;
    v->i = (int *)realloc(v->i, v->cap * sizeof(int));
    v->j = (int *)realloc(v->j, v->cap * sizeof(int));
  }
  v->i[v->n] = a;
  v->j[v->n] = b;
  v->n++;
}

/* Is b in a's sorted exclusion row (min index row, utils/utils.c:163-172)? */
static int orc_excluded(const int *nex, const int *off, const int *ex, int a,
                        int b)
{
  int lo, hi, k;
  if (a > b) { k = a; a = b; b = k; }
  lo = off[a];
  hi = off[a] + nex[a];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ex[mid] == b) return 1;
    if (ex[mid] < b) lo = mid + 1; else hi = mid;
  }
  return 0;
}

/* Interacting cell pairs for a half-space stencil on an odd ncell^3 grid, in
 * the reference's enumeration order (energy/control/set_link.c:84-178). */
static void orc_cell_pairs(int ncell, const double h[9], double rmax,
                           pairvec *out)
{
  int nmax = ncell / 2, n2 = ncell * ncell, n3 = n2 * ncell, c;
  double rc = 1. / (double)ncell, sa, d;
  d = (h[0] * h[3] + h[1] * h[4] + h[2] * h[5]) /
      (sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]) *
       sqrt(h[3] * h[3] + h[4] * h[4] + h[5] * h[5]));
  sa = sin(acos(d));
  d = (h[0] * h[6] + h[1] * h[7] + h[2] * h[8]) /
      (sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]) *
       sqrt(h[6] * h[6] + h[7] * h[7] + h[8] * h[8]));
  d = sin(acos(d));
  if (d < sa) sa = d;
  d = (h[3] * h[6] + h[4] * h[7] + h[5] * h[8]) /
      (sqrt(h[3] * h[3] + h[4] * h[4] + h[5] * h[5]) *
       sqrt(h[6] * h[6] + h[7] * h[7] + h[8] * h[8]));
  d = sin(acos(d));
  if (d < sa) sa = d;
  for (c = 0; c < n3; c++) {
    int ic = c / n2, jc = (c - ic * n2) / ncell, kc = c - ic * n2 - jc * ncell;
    int i, j, k;
    for (i = 0; i <= nmax; i++)
      for (j = (i == 0) ? 0 : -nmax; j <= nmax; j++)
        for (k = (i == 0 && j == 0) ? 0 : -nmax; k <= nmax; k++) {
          int di = abs(ic - (ic + i) % ncell), dj = abs(jc - (jc + j) % ncell);
          int dk = abs(kc - (kc + k) % ncell);
          double dx, dy, dz, r;
          if (di > ncell / 2) di = ncell - di;
          if (dj > ncell / 2) dj = ncell - dj;
          if (dk > ncell / 2) dk = ncell - dk;
          if (di > 0) di--;
          if (dj > 0) dj--;
          if (dk > 0) dk--;
          dx = h[0] * di * rc + h[1] * dj * rc + h[2] * dk * rc;
          dy = h[3] * di * rc + h[4] * dj * rc + h[5] * dk * rc;
          dz = h[6] * di * rc + h[7] * dj * rc + h[8] * dk * rc;
          r = sqrt(dx * dx + dy * dy + dz * dz) * sa;
          if (r < rmax) {
            int a = ic + i, b = jc + j, e = kc + k;
            if (a < 0) a += ncell;
            if (b < 0) b += ncell;
            if (e < 0) e += ncell;
            if (a >= ncell) a -= ncell;
            if (b >= ncell) b -= ncell;
            if (e >= ncell) e -= ncell;
            pv_push(out, c, a * n2 + b * ncell + e);
          }
        }
  }
}

/* Head/link binning in wrapped fractional coordinates; atoms inserted from
 * the top so each cell chain ascends (set_link.c:203-300). */
static void orc_bin(int natoms, int ncell, const double *pos, const double h[9],
                    const double *q, int *head, int *link)
{
  double hi[9];
  int n2 = ncell * ncell, n3 = n2 * ncell, i;
  const double *x = pos, *y = pos + natoms, *z = pos + 2 * natoms;
  orc_hinv(h, hi);
  for (i = 0; i < n3; i++) head[i] = -1;
  for (i = natoms - 1; i >= 0; i--) {
    double sx, sy, sz;
    int a, b, c, cell;
    if (q && q[i] == 0.0) continue;
    sx = x[i] * hi[0] + y[i] * hi[1] + z[i] * hi[2];
    sy = x[i] * hi[3] + y[i] * hi[4] + z[i] * hi[5];
    sz = x[i] * hi[6] + y[i] * hi[7] + z[i] * hi[8];
    sx -= orc_anint(sx);
    sy -= orc_anint(sy);
    sz -= orc_anint(sz);
    a = (int)((sx + .5) * ncell);
    b = (int)((sy + .5) * ncell);
    c = (int)((sz + .5) * ncell);
    if (a < 0) a = 0;
    if (b < 0) b = 0;
    if (c < 0) c = 0;
    if (a > ncell - 1) a = ncell - 1;
    if (b > ncell - 1) b = ncell - 1;
    if (c > ncell - 1) c = ncell - 1;
    cell = a * n2 + b * ncell + c;
    link[i] = head[cell];
    head[cell] = i;
  }
}

/* Candidate filter: minimum image, r^2 = x^2+y^2+z^2, strict compare against
 * the (rc+skin)^2 of the pair's interaction (ngbr_list.c:35-101,103-143). */
static void orc_keep_vdw(const orc_sys *s, const double *pos, const double *h,
                         const double *hi, int a, int b, pairvec *t,
                         pairvec *sl)
{
  int n = s->natoms, in = s->map[s->itype[a] * s->ntypes + s->itype[b]];
  double dx = pos[a] - pos[b], dy = pos[n + a] - pos[n + b];
  double dz = pos[2 * n + a] - pos[2 * n + b], r2;
  if (s->iperd == 3) orc_minimg(&dx, &dy, &dz, h, hi);
  r2 = dx * dx + dy * dy + dz * dz;
  if (r2 < s->rmaxl_skin2[in]) {
    pv_push(t, a, b);
    if (r2 < s->rmaxs_skin2[in]) pv_push(sl, a, b);
  }
}

static void orc_keep_elec(const orc_sys *s, const double *pos, const double *h,
                          const double *hi, int a, int b, pairvec *t,
                          pairvec *sl)
{
  int n = s->natoms;
  double dx = pos[a] - pos[b], dy = pos[n + a] - pos[n + b];
  double dz = pos[2 * n + a] - pos[2 * n + b], r2;
  if (s->iperd == 3) orc_minimg(&dx, &dy, &dz, h, hi);
  r2 = dx * dx + dy * dy + dz * dz;
  if (r2 < s->rmaxl_skin2_e) {
    pv_push(t, a, b);
    if (r2 < s->rmaxs_skin2_e) pv_push(sl, a, b);
  }
}

/* Verlet lists as get_ngbr + get_ngbr_e build them (ngbr_list.c:297-581,
 * elec_ngbr.c:220-358).  use_cells=1 walks the reference's link-cell pairs in
 * its order (nocell_all=1, the default); 0 is the O(N^2) i<j walk. */
int orc_build_lists(const orc_sys *s, const double *pos, const double *h,
                    int use_cells, orc_lists *L)
{
  int n = s->natoms, i, k, *xoff, *cxoff;
  double hi[9];
  pairvec t = {0}, sl = {0}, te = {0}, se = {0};
  orc_hinv(h, hi);
  xoff = (int *)malloc((n + 1) * sizeof(int));
  cxoff = (int *)malloc((n + 1) * sizeof(int));
  xoff[0] = cxoff[0] = 0;
  for (i = 0; i < n; i++) {
    xoff[i + 1] = xoff[i] + s->nexclude[i];
    cxoff[i + 1] = cxoff[i] + (s->cnexclude ? s->cnexclude[i] : 0);
  }
  if (!use_cells) {
    for (i = 0; i < n; i++) {
      int j;
      for (j = i + 1; j < n; j++) {
        if (s->map[s->itype[i] * s->ntypes + s->itype[j]] == -1) continue;
        if (orc_excluded(s->nexclude, xoff, s->exclude, i, j)) continue;
        orc_keep_vdw(s, pos, h, hi, i, j, &t, &sl);
      }
    }
    for (i = 0; i < n; i++) {
      int j;
      if (s->qch[i] == 0.) continue;
      for (j = i + 1; j < n; j++) {
        if (s->qch[j] == 0.) continue;
        if (s->cnexclude && orc_excluded(s->cnexclude, cxoff, s->cexclude, i, j))
          continue;
        orc_keep_elec(s, pos, h, hi, i, j, &te, &se);
      }
    }
  } else {
    int nc = s->ncell, n3 = nc * nc * nc;
    int *head = (int *)malloc(n3 * sizeof(int)), *link = (int *)malloc(n * sizeof(int));
    pairvec cp = {0}, cpe = {0};
    double rmx = 0., rv, re;
    for (k = 0; k < s->ninter; k++)
      if (s->rmaxl_skin2[k] > rmx) rmx = s->rmaxl_skin2[k];
    rv = sqrt(rmx) + s->skin;
    re = sqrt(s->rmaxl_skin2_e) + s->skin;
    orc_cell_pairs(nc, h, rv, &cp);
    orc_cell_pairs(nc, h, re, &cpe);
    orc_bin(n, nc, pos, h, NULL, head, link);
    for (k = 0; k < cp.n; k++) {
      int ic = cp.i[k], jc = cp.j[k], a = head[ic];
      while (a != -1) {
        int b = (ic == jc) ? link[a] : head[jc];
        const int *mrow = s->map + s->itype[a] * s->ntypes;
        while (b != -1) {
          if (mrow[s->itype[b]] != -1 &&
              !orc_excluded(s->nexclude, xoff, s->exclude, a, b))
            orc_keep_vdw(s, pos, h, hi, a, b, &t, &sl);
          b = link[b];
        }
        a = link[a];
      }
    }
    orc_bin(n, nc, pos, h, s->qch, head, link);
    for (k = 0; k < cpe.n; k++) {
      int ic = cpe.i[k], jc = cpe.j[k], a = head[ic];
      while (a != -1) {
        int b = (ic == jc) ? link[a] : head[jc];
        while (b != -1) {
          if (!s->cnexclude ||
              !orc_excluded(s->cnexclude, cxoff, s->cexclude, a, b))
            orc_keep_elec(s, pos, h, hi, a, b, &te, &se);
          b = link[b];
        }
        a = link[a];
      }
    }
    free(head); free(link);
    free(cp.i); free(cp.j); free(cpe.i); free(cpe.j);
  }
  free(xoff); free(cxoff);
  L->nprt = t.n; L->inpt = t.i; L->jnpt = t.j;
  L->nprs = sl.n; L->inps = sl.i; L->jnps = sl.j;
  L->nprte = te.n; L->inpte = te.i; L->jnpte = te.j;
  L->nprse = se.n; L->inpse = se.i; L->jnpse = se.j;
  return 0;
}

void orc_free_lists(orc_lists *L)
{
  free(L->inpt); free(L->jnpt); free(L->inps); free(L->jnps);
  free(L->inpte); free(L->jnpte); free(L->inpse); free(L->jnpse);
  memset(L, 0, sizeof(*L));
}

/* Rebuild trigger: any atom farther than skin/2 from its build position
 * (energy/control/md_finter.c:61-73).  RESETPOS is not restated. */
int orc_need_rebuild(int natoms, double skin, const double *pos,
                     const double *pb)
{
  double s2 = .25 * skin * skin;
  int i;
  for (i = 0; i < natoms; i++) {
    double x = pb[i] - pos[i], y = pb[natoms + i] - pos[natoms + i];
    double z = pb[2 * natoms + i] - pos[2 * natoms + i];
    if (x * x + y * y + z * z > s2) return 1;
  }
  return 0;
}

/* ======================================================================
 * Forces
 * ==================================================================== */

/* vdW table kernel over one chunk of at most nlen pairs
 * (energy/inter/md_fnopol.c:46-255).  scr gets the full force for iflg 1 and
 * the switched force for iflg 0; rec gets the switched force (iflg 1). */
static void orc_vdw_chunk(const orc_sys *s, int n0, const int *ip,
                          const int *jp, int iflg, const double *pos,
                          const double *h, const double *hi, double *scr,
                          double *rec, orc_energy *e)
{
  int n = s->natoms, j;
  int ntm4 = s->ntable - 4;
  const double *rmax2 = iflg ? s->rmaxl_cut2 : s->rmaxs_cut2;
  double spotlt = 0., spothlt = 0.;
  for (j = 0; j < n0; j++) {
    int a = ip[j], b = jp[j], in = s->map[s->itype[a] * s->ntypes + s->itype[b]];
    double dx = pos[a] - pos[b], dy = pos[n + a] - pos[n + b];
    double dz = pos[2 * n + a] - pos[2 * n + b], d2, p, f, swv, fx, fy, fz;
    double sx = 0., sy = 0., sz = 0.;
    int k;
    if (s->iperd == 3) orc_minimg(&dx, &dy, &dz, h, hi);
    d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 < rmax2[in])) continue;
    p = (d2 - s->rmins_cut2[in]) * s->dx2tab[in];
    k = (int)p;
    if (k < 1) { e->nclamp++; k = 1; p = 1.; }
    p = p - (double)k;
    f = orc_interp(s->dvtab + (long)in * s->ntable, k, p);
    if (d2 > s->rminl_cut2[in])
      swv = orc_interp(s->switchs + (long)in * s->ntable, k, p);
    else
      swv = 1.;
    if (iflg == 1) {
      double v;
      e->sprsl -= d2 * f;
      e->sprslt[0] -= dx * dx * f;
      e->sprslt[1] -= dy * dx * f;
      e->sprslt[2] -= dz * dx * f;
      e->sprslt[4] -= dy * dy * f;
      e->sprslt[5] -= dz * dy * f;
      e->sprslt[8] -= dz * dz * f;
      fx = dx * f; fy = dy * f; fz = dz * f;
      sx = fx * swv; sy = fy * swv; sz = fz * swv;
      v = orc_interp(s->vtab + (long)in * s->ntable, k, p);
      spotlt += v;
      spothlt += v - s->vtab[(long)in * s->ntable + ntm4];
    } else {
      f *= swv;
      fx = dx * f; fy = dy * f; fz = dz * f;
    }
    scr[a] -= fx; scr[n + a] -= fy; scr[2 * n + a] -= fz;
    scr[b] += fx; scr[n + b] += fy; scr[2 * n + b] += fz;
    if (iflg == 1) {
      rec[a] -= sx; rec[n + a] -= sy; rec[2 * n + a] -= sz;
      rec[b] += sx; rec[n + b] += sy; rec[2 * n + b] += sz;
    }
  }
  if (iflg == 1) {
    e->sprslt[3] = e->sprslt[1];
    e->sprslt[6] = e->sprslt[2];
    e->sprslt[7] = e->sprslt[5];
  }
  e->spotl += spotlt;
  e->spothl += spothlt;
}

/* Real-space Coulomb table kernel over one chunk (energy/inter/finter_e.c:
 * 42-233).  The switch is inert: both RESPA levels use the long cutoff
 * (the reference's case 0 falls through into case 1). */
static void orc_elec_chunk(const orc_sys *s, int n0, const int *ip,
                           const int *jp, int iflg, const double *pos,
                           const double *h, const double *hi, double *scr,
                           double *rec, orc_energy *e)
{
  int n = s->natoms, j;
  double rmax2 = s->rmaxl_coul2, rmin2 = s->rmaxl_coul2;
  double tablast = s->coultab[s->ntable - 4], sphet = 0., spet = 0.;
  for (j = 0; j < n0; j++) {
    int a = ip[j], b = jp[j], k;
    double dx = pos[a] - pos[b], dy = pos[n + a] - pos[n + b];
    double dz = pos[2 * n + a] - pos[2 * n + b], d2, p, f, swv, qij, fcc;
    double fx, fy, fz, sx = 0., sy = 0., sz = 0.;
    qij = s->qch[a] * s->qch[b];
    if (s->iperd == 3) orc_minimg(&dx, &dy, &dz, h, hi);
    d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 < rmax2)) continue;
    p = (d2 - s->rmins_coul2) * s->dx2coul;
    k = (int)p;
    if (k < 1) { e->nclamp++; k = 1; p = 1; }
    p = p - (double)k;
    f = orc_interp(s->dcoultab, k, p);
    swv = (d2 > rmin2) ? orc_interp(s->switchse, k, p) : 1.;
    fcc = qij * f;
    if (iflg == 1) {
      double v;
      e->sprse -= d2 * fcc;
      e->sprset[0] -= dx * dx * fcc;
      e->sprset[1] -= dy * dx * fcc;
      e->sprset[2] -= dz * dx * fcc;
      e->sprset[4] -= dy * dy * fcc;
      e->sprset[5] -= dz * dy * fcc;
      e->sprset[8] -= dz * dz * fcc;
      fx = dx * fcc; fy = dy * fcc; fz = dz * fcc;
      sx = fx * swv; sy = fy * swv; sz = fz * swv;
      v = orc_interp(s->coultab, k, p);
      sphet += qij * (v - tablast);
      spet += qij * v;
    } else {
      fcc *= swv;
      fx = dx * fcc; fy = dy * fcc; fz = dz * fcc;
    }
    scr[a] -= fx; scr[n + a] -= fy; scr[2 * n + a] -= fz;
    scr[b] += fx; scr[n + b] += fy; scr[2 * n + b] += fz;
    if (iflg == 1) {
      rec[a] -= sx; rec[n + a] -= sy; rec[2 * n + a] -= sz;
      rec[b] += sx; rec[n + b] += sy; rec[2 * n + b] += sz;
    }
  }
  e->spothe += sphet;
  e->spote += spet;
  if (iflg == 1) {
    e->sprset[3] = e->sprset[1];
    e->sprset[6] = e->sprset[2];
    e->sprset[7] = e->sprset[5];
  }
}

/* Reciprocal-space Ewald sum over the half-space k list, with the
 * cos/sin recursion along the third reciprocal axis inside each (i,j) row
 * (energy/inter/md_ewald.c:185-361, fk_ewald).  Adds into fl. */
static void orc_fk_ewald(const orc_sys *s, const double *pos, const double *h,
                         double *fl, orc_energy *e)
{
  int n = s->natoms, nq = 0, i, j;
  double hi[9], rvol, fact = sqrt(4.0 * M_PI), tpi = 2. * M_PI;
  double falp2 = 1. / (4. * s->kappa * s->kappa);
  double vk = 0., vp = 0., vt[9] = {0}, bk;
  int *iq;
  double *tx, *ty, *tz, *tq, *cs, *sn, *hr, *hm;
  rvol = 1. / orc_hinv(h, hi);
  iq = (int *)malloc(n * sizeof(int));
  for (i = 0; i < n; i++) if (s->qch[i] != 0.) iq[nq++] = i;
  tx = (double *)malloc(8 * (nq + 1) * sizeof(double));
  ty = tx + nq; tz = tx + 2 * nq; tq = tx + 3 * nq;
  cs = tx + 4 * nq; sn = tx + 5 * nq; hr = tx + 6 * nq; hm = tx + 7 * nq;
  for (j = 0; j < nq; j++) {
    int a = iq[j];
    tx[j] = pos[a]; ty[j] = pos[n + a]; tz[j] = pos[2 * n + a];
    tq[j] = s->qch[a] * fact;
  }
  for (j = 0; j < nq; j++) {
    double akc = (tx[j] * hi[6] + ty[j] * hi[7] + tz[j] * hi[8]) * tpi;
    cs[j] = cos(akc);
    sn[j] = sin(akc);
  }
  for (i = 0; i < s->ktot; i++) {
    double six = tpi * s->ka[i], siy = tpi * s->kb[i], siz = tpi * s->kc[i];
    double aka = (six * hi[0] + siy * hi[3] + siz * hi[6]);
    double akb = (six * hi[1] + siy * hi[4] + siz * hi[7]);
    double akc = (six * hi[2] + siy * hi[5] + siz * hi[8]);
    double g2 = aka * aka + akb * akb + akc * akc;
    double preg = exp(-g2 * falp2) * rvol / g2;
    double prep = -2. * preg * (g2 * falp2 + 1.) / g2;
    double sr = 0., si = 0., sm;
    if (s->kup[i] || i == 0) {
      for (j = 0; j < nq; j++) {
        double rk = (tx[j] * aka + ty[j] * akb + tz[j] * akc);
        hr[j] = cos(rk);
        hm[j] = sin(rk);
      }
    } else {
      for (j = 0; j < nq; j++) {
        double r0 = hr[j], r1 = hm[j];
        hr[j] = r0 * cs[j] - r1 * sn[j];
        hm[j] = r1 * cs[j] + r0 * sn[j];
      }
    }
    for (j = 0; j < nq; j++) {
      sr += tq[j] * hr[j];
      si += tq[j] * hm[j];
    }
    sm = sr * sr + si * si;
    vk += preg * sm;
    vp += prep * sm * g2;
    vt[0] += prep * sm * aka * aka;
    vt[1] += prep * sm * akb * aka;
    vt[2] += prep * sm * akc * aka;
    vt[3] += prep * sm * aka * akb;
    vt[4] += prep * sm * akb * akb;
    vt[5] += prep * sm * akc * akb;
    vt[6] += prep * sm * aka * akc;
    vt[7] += prep * sm * akb * akc;
    vt[8] += prep * sm * akc * akc;
    sr *= 2. * preg;
    si *= 2. * preg;
    for (j = 0; j < nq; j++) {
      int a = iq[j];
      double q = tq[j], rr = hr[j], rm = hm[j];
      double srx = aka * sr * q, sry = akb * sr * q, srz = akc * sr * q;
      double sxx = aka * si * q, syy = akb * si * q, szz = akc * si * q;
      fl[a] += srx * rm - sxx * rr;
      fl[n + a] += sry * rm - syy * rr;
      fl[2 * n + a] += srz * rm - szz * rr;
    }
  }
  bk = s->ewald_bkg / orc_det(h);
  e->vk_ewald = vk + bk + s->ewald_self;
  e->vp_ewald = vp + 3. * (vk + bk);
  for (i = 0; i < 9; i++) e->vp_ewald_tensor[i] = vt[i];
  e->vp_ewald_tensor[0] += vk + bk;
  e->vp_ewald_tensor[4] += vk + bk;
  e->vp_ewald_tensor[8] += vk + bk;
  free(iq);
  free(tx);
}

/* Remove erf(a r)/r between excluded charged pairs, no minimum image
 * (energy/inter/md_ewald.c:385-453, ecorr). */
static void orc_ecorr(const orc_sys *s, const double *pos, double *fl,
                      orc_energy *e)
{
  int n = s->natoms, i;
  double spi = 1. / sqrt(M_PI), ka = s->kappa, ka2 = ka * ka;
  double spot = 0., sprs = 0., t[9] = {0};
  for (i = 0; i < s->necorr; i++) {
    int a = s->iecorr[i], b = s->jecorr[i];
    double dx = pos[a] - pos[b], dy = pos[n + a] - pos[n + b];
    double dz = pos[2 * n + a] - pos[2 * n + b];
    double d2 = dx * dx + dy * dy + dz * dz, d = sqrt(d2);
    double qr = s->qch[a] * s->qch[b] / d, er = erf(d * ka), dv;
    spot += -qr * er;
    dv = -qr * (-2. * ka * exp(-ka2 * d2) * spi + er / d);
    sprs += d * dv;
    dv /= d;
    t[0] += dv * dx * dx; t[1] += dv * dx * dy; t[2] += dv * dx * dz;
    t[3] += dv * dy * dx; t[4] += dv * dy * dy; t[5] += dv * dy * dz;
    t[6] += dv * dz * dx; t[7] += dv * dz * dy; t[8] += dv * dz * dz;
    dx *= dv; dy *= dv; dz *= dv;
    fl[a] += dx; fl[n + a] += dy; fl[2 * n + a] += dz;
    fl[b] -= dx; fl[n + b] -= dy; fl[2 * n + b] -= dz;
  }
  e->pot_ecorr = spot;
  e->prs_ecorr = sprs;
  for (i = 0; i < 9; i++) e->prs_ecorr_tensor[i] = t[i];
}

/* force(...,iflg) for the intermolecular levels, before rmass
 * (energy/control/md_fcntr.c:53-76; drivers md_finter.c:110-310 and
 *  md_finter_e.c:78-233 in verlist mode, chunked by nlen). */
void orc_force(const orc_sys *s, const orc_lists *L, int iflg,
               const double *pos, const double *h, double *fs, double *fl,
               orc_energy *e)
{
  int n3 = 3 * s->natoms, i, np;
  int nlen = s->nlen > 0 ? s->nlen : 500;
  double hi[9], *scr, *rec, *fx;
  const int *ip, *jp;
  memset(e, 0, sizeof(*e));
  orc_hinv(h, hi);
  scr = (double *)calloc(2 * n3, sizeof(double));
  rec = scr + n3;
  for (i = 0; i < n3; i++) fs[i] = 0.;
  if (iflg == 1) for (i = 0; i < n3; i++) fl[i] = 0.;
  fx = iflg ? fl : fs;
  /* vdW */
  np = iflg ? L->nprt : L->nprs;
  ip = iflg ? L->inpt : L->inps;
  jp = iflg ? L->jnpt : L->jnps;
  for (i = 0; i < np; i += nlen) {
    int k = np - i < nlen ? np - i : nlen;
    orc_vdw_chunk(s, k, ip + i, jp + i, iflg, pos, h, hi, scr, rec, e);
  }
  for (i = 0; i < n3; i++) fx[i] += scr[i];
  if (iflg == 1) for (i = 0; i < n3; i++) fs[i] += rec[i];
  /* real-space coulomb */
  memset(scr, 0, 2 * n3 * sizeof(double));
  np = iflg ? L->nprte : L->nprse;
  ip = iflg ? L->inpte : L->inpse;
  jp = iflg ? L->jnpte : L->jnpse;
  for (i = 0; i < np; i += nlen) {
    int k = np - i < nlen ? np - i : nlen;
    orc_elec_chunk(s, k, ip + i, jp + i, iflg, pos, h, hi, scr, rec, e);
  }
  for (i = 0; i < n3; i++) fx[i] += scr[i];
  if (iflg == 1) for (i = 0; i < n3; i++) fs[i] += rec[i];
  /* md_fcntr.c:65-68: with iperd 3 fk_ewald (background and self even with
     no k-vectors) and ecorr always run */
  if (iflg == 1 && s->iperd == 3) {
    orc_fk_ewald(s, pos, h, fl, e);
    orc_ecorr(s, pos, fl, e);
  }
  free(scr);
}

/* ======================================================================
 * Independent checks
 * ==================================================================== */

/* Reciprocal energy with direct cos/sin for every k (md_ewald.c:144-183,
 * k_ewald) plus the self and background constants. */
double orc_ewald_recip_direct(const orc_sys *s, const double *pos,
                              const double *h)
{
  int n = s->natoms, i, j;
  double hi[9], vol, vk = 0., tpi = 2. * M_PI;
  double falp2 = 1. / (4. * s->kappa * s->kappa);
  vol = orc_hinv(h, hi);
  for (i = 0; i < s->ktot; i++) {
    double aka = (s->ka[i] * hi[0] + s->kb[i] * hi[3] + s->kc[i] * hi[6]) * tpi;
    double akb = (s->ka[i] * hi[1] + s->kb[i] * hi[4] + s->kc[i] * hi[7]) * tpi;
    double akc = (s->ka[i] * hi[2] + s->kb[i] * hi[5] + s->kc[i] * hi[8]) * tpi;
    double g2 = aka * aka + akb * akb + akc * akc;
    /* half-space k list: E = (4 pi / V) sum_k e^{-k^2/4a^2} / k^2 |S(k)|^2 */
    double preg = 4. * M_PI * exp(-g2 * falp2) / (g2 * vol), sr = 0., si = 0.;
    for (j = 0; j < n; j++) {
      double q = s->qch[j], rk;
      if (q == 0.) continue;
      rk = pos[j] * aka + pos[n + j] * akb + pos[2 * n + j] * akc;
      sr += q * cos(rk);
      si += q * sin(rk);
    }
    vk += preg * (sr * sr + si * si);
  }
  return vk + s->ewald_self + s->ewald_bkg / vol;
}

/* Real-space Ewald energy with erfc evaluated directly, all non-excluded
 * charged pairs within rcut, minimum image (md_ewald.c:455-483, r_ewald). */
double orc_coul_real_direct(const orc_sys *s, const double *pos,
                            const double *h, double rcut)
{
  int n = s->natoms, i, j, *xoff;
  double hi[9], v = 0.;
  orc_hinv(h, hi);
  xoff = (int *)malloc((n + 1) * sizeof(int));
  xoff[0] = 0;
  for (i = 0; i < n; i++) xoff[i + 1] = xoff[i] + (s->cnexclude ? s->cnexclude[i] : 0);
  for (i = 0; i < n; i++) {
    if (s->qch[i] == 0.) continue;
    for (j = i + 1; j < n; j++) {
      double dx, dy, dz, r;
      if (s->qch[j] == 0.) continue;
      if (s->cnexclude && orc_excluded(s->cnexclude, xoff, s->cexclude, i, j)) continue;
      dx = pos[i] - pos[j]; dy = pos[n + i] - pos[n + j]; dz = pos[2 * n + i] - pos[2 * n + j];
      if (s->iperd == 3) orc_minimg(&dx, &dy, &dz, h, hi);
      r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r < rcut) v += s->qch[i] * s->qch[j] * erfc(s->kappa * r) / r;
    }
  }
  free(xoff);
  return v;
}

/* Analytic Lennard-Jones energy and forces over the vdW pairs within each
 * interaction's long cutoff, minimum image, no table; eps/sig per
 * interaction index. */
double orc_lj_direct(const orc_sys *s, const double *pos, const double *h,
                     const double *eps, const double *sig, double *fx)
{
  int n = s->natoms, i, j, *xoff;
  double hi[9], v = 0.;
  orc_hinv(h, hi);
  xoff = (int *)malloc((n + 1) * sizeof(int));
  xoff[0] = 0;
  for (i = 0; i < n; i++) xoff[i + 1] = xoff[i] + s->nexclude[i];
  for (i = 0; i < 3 * n; i++) fx[i] = 0.;
  for (i = 0; i < n; i++) {
    for (j = i + 1; j < n; j++) {
      int in = s->map[s->itype[i] * s->ntypes + s->itype[j]];
      double dx, dy, dz, r2, sr6, f;
      if (in == -1) continue;
      if (orc_excluded(s->nexclude, xoff, s->exclude, i, j)) continue;
      dx = pos[i] - pos[j]; dy = pos[n + i] - pos[n + j]; dz = pos[2 * n + i] - pos[2 * n + j];
      if (s->iperd == 3) orc_minimg(&dx, &dy, &dz, h, hi);
      r2 = dx * dx + dy * dy + dz * dz;
      if (!(r2 < s->rmaxl_cut2[in])) continue;
      sr6 = pow(sig[in] * sig[in] / r2, 3.);
      v += 4. * eps[in] * (sr6 * sr6 - sr6);
      f = 4. * eps[in] * (12. * sr6 * sr6 - 6. * sr6) / r2; /* -dV/dr / r */
      fx[i] += f * dx; fx[n + i] += f * dy; fx[2 * n + i] += f * dz;
      fx[j] -= f * dx; fx[n + j] -= f * dy; fx[2 * n + j] -= f * dz;
    }
  }
  free(xoff);
  return v;
}

/* Harmonic bonds and bends (energy/intra/md_fbond.c:76-128 case 0,
 * energy/intra/md_fbend.c:126-260 case 0); adds forces, returns energy.
 * The energies follow getvbond (md_fbond.c:131-200, case 0: .5*fk*dr*dr)
 * and getvbend (md_fbend.c:262-340, case 0, with th = 1 when |cos| >= 1),
 * each in its own running sum, added as geteng adds them
 * (energy/control/md_fcntr.c:206-207). */
double orc_bonded_force(int natoms, const orc_bonded *b, const double *pos,
                        double *f, double parts[2])
{
  const double *x = pos, *y = pos + natoms, *z = pos + 2 * natoms;
  double *fx = f, *fy = f + natoms, *fz = f + 2 * natoms, v = 0., va = 0.;
  int nb;
  for (nb = 0; nb < b->nbond; nb++) {
    int i = b->ib[nb], j = b->jb[nb];
    double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
    double r = sqrt(dx * dx + dy * dy + dz * dz), dr = r - b->beq[nb];
    double fr = b->bk[nb] * dr / r;
    v += .5 * b->bk[nb] * dr * dr;
    dx *= fr; dy *= fr; dz *= fr;
    fx[i] -= dx; fy[i] -= dy; fz[i] -= dz;
    fx[j] += dx; fy[j] += dy; fz[j] += dz;
  }
  for (nb = 0; nb < b->nbend; nb++) {
    int i = b->ia[nb], j = b->ja[nb], k = b->ka[nb];
    double ax = x[i] - x[j], ay = y[i] - y[j], az = z[i] - z[j];
    double bx = x[k] - x[j], by = y[k] - y[j], bz = z[k] - z[j];
    double al = ax * ax + ay * ay + az * az, bl = bx * bx + by * by + bz * bz;
    double ab = ax * bx + ay * by + az * bz, r11 = 1. / sqrt(al * bl);
    double c = ab * r11, dudc, d1x, d1y, d1z, d3x, d3y, d3z;
    if (fabs(c) >= 1.) {
      dudc = 0.;
      va += .5 * b->ak[nb] * (1. - b->aeq[nb]) * (1. - b->aeq[nb]);
    } else {
      double th = acos(c), dth = th - b->aeq[nb];
      dudc = (b->ak[nb] * dth) * (-1. / sqrt(1. - c * c));
      va += .5 * b->ak[nb] * dth * dth;
    }
    d1x = c * ax / al - bx * r11; d1y = c * ay / al - by * r11; d1z = c * az / al - bz * r11;
    d3x = c * bx / bl - ax * r11; d3y = c * by / bl - ay * r11; d3z = c * bz / bl - az * r11;
    fx[i] += dudc * d1x; fy[i] += dudc * d1y; fz[i] += dudc * d1z;
    fx[j] += dudc * (-d1x - d3x); fy[j] += dudc * (-d1y - d3y); fz[j] += dudc * (-d1z - d3z);
    fx[k] += dudc * d3x; fy[k] += dudc * d3y; fz[k] += dudc * d3z;
  }
  if (parts) { parts[0] = v; parts[1] = va; }
  return v + va;
}
