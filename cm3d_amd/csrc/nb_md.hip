// nb_md.hip -- device-resident velocity-Verlet / RESPA step around the
// nonbonded engine: cm3d's step() (integrate/md_step.c:62-159) for the
// nres_lrf = 1 case every configuration uses, with harmonic bonds and bends
// (energy/intra/md_fbond.c:76-128, md_fbend.c:126-260) on the inner levels.
//
// Per outer step, all on one stream with no host round trip:
//   k_md_inner    one thread per molecule: v += dter*al; then nres_intra times
//                 { v += dta2*aa; r += dta2*v; r += dta2*v; aa = bonded(r)/m;
//                   v += dta2*aa } -- molecules never share bonded terms, so
//                 the whole inner RESPA loop runs in one launch with the
//                 molecule's atoms in LDS; its epilogue applies nrefsh's
//                 test |r - r_build| > skin/2 (flag in HBM) and writes the
//                 sorted pair records and the per-charge Ewald data
//   list build    the cell sort + list kernels, each returning at once when
//                 the flag is clear (so the step is graph-capturable)
//   force(...,1)  the pair and Ewald (+ ecorr) kernels
//   k_md_kick     v += dter * fl / m, plus the energy-column reductions and
//                 the flag reset for the next step
// Positions and velocities stay FP64 (cm3d's integrator arithmetic, no FMA).
// A step can be captured once into a hipGraph and replayed.
//
// NVT (a Nose-Hoover chain, cm3d's intnhc at every inner substep) splits the
// inner loop at each chain update, which needs the kinetic sum of the whole
// system: k_md_nhc launches 0..nres_intra replace k_md_inner (see there).
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "nb_device.h"

#define MD_MAXA 8          // atoms per molecule held in LDS by k_md_inner
#define MD_TPB 64          // molecules per workgroup
#define MD_MAXCHAIN 16     // Nose-Hoover chain links

struct nb_md {
  nb_ctx *c;
  int n, nmol, nbond, nbend, ninnra;
  double dt, dter, dta2;
  double *d_vel, *d_rm, *d_aa, *d_amass;
  int *d_moff, *d_boff, *d_aoff;
  int *d_bi, *d_bj, *d_ai, *d_aj, *d_ak;
  double *d_bk, *d_br0, *d_afk, *d_ath0;
  double *d_ep, *d_eo;
  int ep_blocks;
  Box box;
  double hmat[9];
  hipGraph_t graph;
  hipGraphExec_t gexec;
  int have_graph;
  int nrebuild_host;
  int maxa;      // atoms in the largest molecule
  size_t bytes;  // device memory of this MD state (counted in its context's footprint)
  // Nose-Hoover chain (nchain 0: NVE).  d_nhc: veta[nc], eta[nc], gkt[nc],
  // mnh[nc]; d_ppart: per-workgroup kinetic sums; d_ticket: last-workgroup
  // counter (zero between launches)
  int fork;      // run the reciprocal sum on the context's second stream
  // what the captured graph was recorded against: the context's generation
  // and this state's (bumped when the thermostat changes the kernel
  // sequence or reallocates its buffers); a mismatch re-captures
  unsigned tgen, graph_cgen, graph_tgen;
  // snapshot of the dynamic state (pos, vel, aa, fl: 12N; then veta, eta)
  // taken at the start of a batch tick, restored when a list build in the
  // tick overflowed
  double *d_save;
  int used;      // a step has been recorded or launched (it may be in flight)
  int nchain;
  double *d_nhc, *d_ppart;
  int *d_ithm, *d_ticket;
  size_t nhc_bytes;
};

// Harmonic bonds then bends of one molecule, in cm3d's term order, on the
// molecule's atoms in LDS (p: [MD_MAXA][3] of this thread); f is zeroed here.
static __device__ void mol_bonded(const double (*p)[3], double (*f)[3], int na, int a0,
                                  int b0, int b1, int e0, int e1, const int *bi, const int *bj,
                                  const double *bk, const double *br0, const int *ai,
                                  const int *aj, const int *ak, const double *afk,
                                  const double *ath0) {
  for (int a = 0; a < na; a++) f[a][0] = f[a][1] = f[a][2] = 0.;
  for (int b = b0; b < b1; b++) {
    int i = bi[b] - a0, j = bj[b] - a0;
    double dx = p[i][0] - p[j][0], dy = p[i][1] - p[j][1], dz = p[i][2] - p[j][2];
    double r = sqrt(dx * dx + dy * dy + dz * dz), dr = r - br0[b];
    double fr = bk[b] * dr / r;
    dx *= fr; dy *= fr; dz *= fr;
    f[i][0] -= dx; f[i][1] -= dy; f[i][2] -= dz;
    f[j][0] += dx; f[j][1] += dy; f[j][2] += dz;
  }
  for (int e = e0; e < e1; e++) {
    int i = ai[e] - a0, j = aj[e] - a0, k = ak[e] - a0;
    double ax = p[i][0] - p[j][0], ay = p[i][1] - p[j][1], az = p[i][2] - p[j][2];
    double bx = p[k][0] - p[j][0], by = p[k][1] - p[j][1], bz = p[k][2] - p[j][2];
    double al = ax * ax + ay * ay + az * az, bl = bx * bx + by * by + bz * bz;
    double ab = ax * bx + ay * by + az * bz, r11 = 1. / sqrt(al * bl);
    double c = ab * r11, dudc;
    if (fabs(c) >= 1.) {
      dudc = 0. * -1.;
    } else {
      double th = acos(c), dth = th - ath0[e];
      double dudth = afk[e] * dth, dthdcos = -1. / sqrt(1. - c * c);
      dudc = dudth * dthdcos;
    }
    double d1x = c * ax / al - bx * r11, d1y = c * ay / al - by * r11, d1z = c * az / al - bz * r11;
    double d3x = c * bx / bl - ax * r11, d3y = c * by / bl - ay * r11, d3z = c * bz / bl - az * r11;
    double d2x = -d1x - d3x, d2y = -d1y - d3y, d2z = -d1z - d3z;
    f[i][0] += dudc * d1x; f[i][1] += dudc * d1y; f[i][2] += dudc * d1z;
    f[j][0] += dudc * d2x; f[j][1] += dudc * d2y; f[j][2] += dudc * d2z;
    f[k][0] += dudc * d3x; f[k][1] += dudc * d3y; f[k][2] += dudc * d3z;
  }
}

// The same harmonic bond / bend arithmetic on a molecule held in registers
// (at most NA atoms): atom slots are picked by unrolled compares, so every
// index is static; operations and their order are those of mol_bonded.
template <int NA>
static __device__ __forceinline__ void sel3(const double (&p)[NA][3], int i, double &x,
                                            double &y, double &z) {
  x = p[0][0]; y = p[0][1]; z = p[0][2];
#pragma unroll
  for (int a = 1; a < NA; a++) {
    const bool h = i == a;
    x = h ? p[a][0] : x;
    y = h ? p[a][1] : y;
    z = h ? p[a][2] : z;
  }
}
template <int NA>
static __device__ __forceinline__ void add3(double (&f)[NA][3], int i, double x, double y,
                                            double z) {
#pragma unroll
  for (int a = 0; a < NA; a++) {  // +0. leaves the other slots' values unchanged
    const bool h = i == a;
    f[a][0] += h ? x : 0.;
    f[a][1] += h ? y : 0.;
    f[a][2] += h ? z : 0.;
  }
}
template <int NA>
static __device__ __forceinline__ void mol_bonded_r(const double (&p)[NA][3], double (&f)[NA][3],
                                                    int a0, int b0, int b1, int e0, int e1,
                                                    const int *bi, const int *bj,
                                                    const double *bk, const double *br0,
                                                    const int *ai, const int *aj,
                                                    const int *ak, const double *afk,
                                                    const double *ath0) {
#pragma unroll
  for (int a = 0; a < NA; a++) f[a][0] = f[a][1] = f[a][2] = 0.;
  for (int b = b0; b < b1; b++) {
    const int i = bi[b] - a0, j = bj[b] - a0;
    double xi, yi, zi, xj, yj, zj;
    sel3(p, i, xi, yi, zi);
    sel3(p, j, xj, yj, zj);
    double dx = xi - xj, dy = yi - yj, dz = zi - zj;
    double r = sqrt(dx * dx + dy * dy + dz * dz), dr = r - br0[b];
    double fr = bk[b] * dr / r;
    dx *= fr; dy *= fr; dz *= fr;
    add3(f, i, -dx, -dy, -dz);
    add3(f, j, dx, dy, dz);
  }
  for (int e = e0; e < e1; e++) {
    const int i = ai[e] - a0, j = aj[e] - a0, k = ak[e] - a0;
    double xi, yi, zi, xj, yj, zj, xk, yk, zk;
    sel3(p, i, xi, yi, zi);
    sel3(p, j, xj, yj, zj);
    sel3(p, k, xk, yk, zk);
    double ax = xi - xj, ay = yi - yj, az = zi - zj;
    double bx = xk - xj, by = yk - yj, bz = zk - zj;
    double al = ax * ax + ay * ay + az * az, bl = bx * bx + by * by + bz * bz;
    double ab = ax * bx + ay * by + az * bz, r11 = 1. / sqrt(al * bl);
    double c = ab * r11, dudc;
    if (fabs(c) >= 1.) {
      dudc = 0. * -1.;
    } else {
      double th = acos(c), dth = th - ath0[e];
      double dudth = afk[e] * dth, dthdcos = -1. / sqrt(1. - c * c);
      dudc = dudth * dthdcos;
    }
    double d1x = c * ax / al - bx * r11, d1y = c * ay / al - by * r11, d1z = c * az / al - bz * r11;
    double d3x = c * bx / bl - ax * r11, d3y = c * by / bl - ay * r11, d3z = c * bz / bl - az * r11;
    double d2x = -d1x - d3x, d2y = -d1y - d3y, d2z = -d1z - d3z;
    add3(f, i, dudc * d1x, dudc * d1y, dudc * d1z);
    add3(f, j, dudc * d2x, dudc * d2y, dudc * d2z);
    add3(f, k, dudc * d3x, dudc * d3y, dudc * d3z);
  }
}

struct InnerArgs {
  int n, nmol, ninnra, kick_first;
  double dter, dta2;
  double *pos, *vel, *aa;
  const double *fl, *rm;
  const int *moff, *boff, *aoff, *bi, *bj, *ai, *aj, *ak;
  const double *bk, *br0, *afk, *ath0;
  // epilogue, for the force evaluation that follows: the nrefsh test, the
  // sorted record of each atom, the Ewald data of each charge
  int epilogue;
  const double *posb; double skin2q; int *flag;
  const int *iperm, *qpos; const double *q;
  double4 *rec64; uint4 *recf;
  double4 *chg, *stp;
  Box b;
};

// End of the inner loop for one atom: write it back and, for the force
// evaluation that follows, apply nrefsh's test (md_finter.c:61-73:
// |r - r_build|^2 > skin^2/4) and write its sorted record and Ewald data.
static __device__ __forceinline__ bool md_atom_out(const InnerArgs &A, int g, double x, double y,
                                                   double z, double vx, double vy, double vz,
                                                   double ax, double ay, double az) {
  const int n = A.n;
  A.pos[g] = x; A.pos[n + g] = y; A.pos[2 * n + g] = z;
  A.vel[g] = vx; A.vel[n + g] = vy; A.vel[2 * n + g] = vz;
  A.aa[g] = ax; A.aa[n + g] = ay; A.aa[2 * n + g] = az;
  if (!A.epilogue) return false;
  const double ex = A.posb[g] - x, ey = A.posb[n + g] - y, ez = A.posb[2 * n + g] - z;
  const bool moved = ex * ex + ey * ey + ez * ez > A.skin2q;
  const int sl = A.iperm[g];
  const double qg = A.q[g];
  if (A.recf) A.recf[sl] = nb_mixed_record(x, y, z, qg, A.b.hi);
  else A.rec64[sl] = make_double4(x, y, z, qg);
  if (A.chg) {
    const int c = A.qpos[g];
    if (c >= 0) nb_charge_record(x, y, z, qg, A.b.hi, A.chg + c, A.stp + c);
  }
  return moved;
}

// Molecules of at most NA atoms, all in registers.
template <int NA>
static __device__ __forceinline__ bool mdinner_regs(const InnerArgs &A, int m, int a0, int na) {
  const int n = A.n;
  double p[NA][3], v[NA][3], aa[NA][3];
#pragma unroll
  for (int a = 0; a < NA; a++) {
    if (a >= na) break;
    const int g = a0 + a;
    const double rm = A.rm[g];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      p[a][d] = A.pos[d * n + g];
      v[a][d] = A.vel[d * n + g];
      aa[a][d] = A.aa[d * n + g];
      if (A.kick_first) v[a][d] += A.dter * (A.fl[d * n + g] * rm);
    }
  }
  const int b0 = A.boff[m], b1 = A.boff[m + 1], e0 = A.aoff[m], e1 = A.aoff[m + 1];
  const bool bonded = (b1 > b0) || (e1 > e0);
  const double dta2 = A.dta2;
  for (int it = 0; it < A.ninnra; it++) {
#pragma unroll
    for (int a = 0; a < NA; a++) {
      if (a >= na) break;
  #pragma unroll
    for (int d = 0; d < 3; d++) {
        v[a][d] += dta2 * aa[a][d];
        p[a][d] += dta2 * v[a][d];
        p[a][d] += dta2 * v[a][d];
      }
    }
    if (bonded) {
      mol_bonded_r<NA>(p, aa, a0, b0, b1, e0, e1, A.bi, A.bj, A.bk, A.br0, A.ai, A.aj, A.ak,
                       A.afk, A.ath0);
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (a >= na) break;
        const double rm = A.rm[a0 + a];
#pragma unroll
        for (int d = 0; d < 3; d++) aa[a][d] *= rm;
      }
    }
#pragma unroll
    for (int a = 0; a < NA; a++) {
      if (a >= na) break;
#pragma unroll
      for (int d = 0; d < 3; d++) v[a][d] += dta2 * aa[a][d];
    }
  }
  bool moved = false;
#pragma unroll
  for (int a = 0; a < NA; a++) {
    if (a >= na) break;
    moved |= md_atom_out(A, a0 + a, p[a][0], p[a][1], p[a][2], v[a][0], v[a][1], v[a][2],
                         aa[a][0], aa[a][1], aa[a][2]);
  }
  return moved;
}

// Larger molecules (up to MD_MAXA atoms) in this thread's slice of the
// workgroup's dynamic LDS.
static __device__ __forceinline__ bool mdinner_lds(const InnerArgs &A, int m, int a0, int na) {
  extern __shared__ double mdl[];
  double (*p)[3] = (double (*)[3])(mdl + (size_t)threadIdx.x * 9 * MD_MAXA);
  double (*v)[3] = p + MD_MAXA, (*aa)[3] = p + 2 * MD_MAXA;
  const int n = A.n;
  for (int a = 0; a < na; a++) {
    int g = a0 + a;
    double rm = A.rm[g];
    for (int d = 0; d < 3; d++) {
      p[a][d] = A.pos[d * n + g];
      v[a][d] = A.vel[d * n + g];
      aa[a][d] = A.aa[d * n + g];
      if (A.kick_first) v[a][d] += A.dter * (A.fl[d * n + g] * rm);
    }
  }
  const int b0 = A.boff[m], b1 = A.boff[m + 1], e0 = A.aoff[m], e1 = A.aoff[m + 1];
  const bool bonded = (b1 > b0) || (e1 > e0);
  const double dta2 = A.dta2;
  for (int it = 0; it < A.ninnra; it++) {
    for (int a = 0; a < na; a++)
      for (int d = 0; d < 3; d++) {
        v[a][d] += dta2 * aa[a][d];
        p[a][d] += dta2 * v[a][d];
        p[a][d] += dta2 * v[a][d];
      }
    if (bonded) {
      mol_bonded(p, aa, na, a0, b0, b1, e0, e1, A.bi, A.bj, A.bk, A.br0, A.ai, A.aj, A.ak,
                 A.afk, A.ath0);
      for (int a = 0; a < na; a++) {
        double rm = A.rm[a0 + a];
        for (int d = 0; d < 3; d++) aa[a][d] *= rm;
      }
    }
    for (int a = 0; a < na; a++)
      for (int d = 0; d < 3; d++) v[a][d] += dta2 * aa[a][d];
  }
  bool moved = false;
  for (int a = 0; a < na; a++)
    moved |= md_atom_out(A, a0 + a, p[a][0], p[a][1], p[a][2], v[a][0], v[a][1], v[a][2],
                         aa[a][0], aa[a][1], aa[a][2]);
  return moved;
}

static __device__ __forceinline__ void k_mdinner_body(const InnerArgs &A, int bid) {
  const int m = bid * MD_TPB + threadIdx.x;
  bool moved = false;
  if (m < A.nmol) {
    const int a0 = A.moff[m], na = A.moff[m + 1] - a0;
    if (na == 1) moved = mdinner_regs<1>(A, m, a0, na);
    else if (na <= 3) moved = mdinner_regs<3>(A, m, a0, na);
    else moved = mdinner_lds(A, m, a0, na);
  }
  if (__ballot(moved) != 0 && lane_id() == 0) atomicOr(A.flag, 1);
}
NB_KERNEL_PAIR(k_mdinner, InnerArgs, MD_TPB)

// ---------------------------------------------------------------------------
// NVT inner loop.  cm3d's substep (md_step.c:108-128) is
//   v += dta2 aa; r += dta2 v; intnhc; r += dta2 v; aa = bonded(r)/m;
//   v += dta2 aa
// and intnhc (intnhc_new, md_therm.new.c:278-332) scales the thermostatted
// velocities by exp(-dta2 veta0), updates the chain from their kinetic sum
// and scales them again by exp(-dta2 veta0) with the new veta0.  The chain
// is not touched between two intnhc calls, so the second scale of one
// substep and the first of the next use the same factor.  Launch `phase`:
//   0          v += dter fl/m (the outer kick), then substep 0 up to its
//              chain update
//   1..nri-1   finish substep phase-1, then substep `phase` up to its update
//   nri        finish the last substep, then the epilogue of k_md_inner
// Launches 0..nri-1 end with the chain update, done by the last workgroup to
// finish from the per-workgroup kinetic sums, summed in workgroup order, so
// the result is independent of scheduling (and of batching).
struct NhcArgs {
  InnerArgs I;
  int phase, nchain, nblocks;
  const int *ithm;
  const double *amass;
  double *nhc, *ppart;
  int *ticket;
};

// The chain update of intnhc_new (md_therm.new.c:294-326) with forcen_new's
// forces (md_therm.new.c:174-197); nhc = veta[nc], eta[nc], gkt[nc], mnh[nc].
static __device__ void nhc_update(double *nhc, int nc, double p2m, double dta2) {
  double *veta = nhc, *eta = nhc + nc;
  const double *gkt = nhc + 2 * nc, *mnh = nhc + 3 * nc;
  const double dta = 2. * dta2;
  double feta[MD_MAXCHAIN];
  for (int i = 0; i < nc - 1; i++) veta[i] *= exp(-dta2 * veta[i + 1]);
  feta[0] = (p2m - gkt[0]) / mnh[0];
  for (int i = 1; i < nc; i++) feta[i] = (mnh[i - 1] * veta[i - 1] * veta[i - 1] - gkt[i]) / mnh[i];
  for (int i = 0; i < nc; i++) veta[i] += dta2 * feta[i];
  for (int i = 0; i < nc; i++) eta[i] += dta * veta[i];
  veta[0] += dta2 * feta[0];
  for (int i = 1; i < nc; i++)
    veta[i] += dta2 * (mnh[i - 1] * veta[i - 1] * veta[i - 1] - gkt[i]) / mnh[i];
  for (int i = nc - 2; i >= 0; i--) veta[i] *= exp(-dta2 * veta[i + 1]);
}

// One molecule through launch N.phase; p, v, aa: NA x 3 storage (registers
// for NA <= 3, this thread's LDS slice otherwise).  s = exp(-dta2 veta0).
template <int NA>
static __device__ __forceinline__ bool mdnhc_mol(const NhcArgs &N, int m, int a0, int na,
                                                 double s, double &p2m, double (*p)[3],
                                                 double (*v)[3], double (*aa)[3]) {
  const InnerArgs &A = N.I;
  const int n = A.n;
  const double dta2 = A.dta2;
  double sc[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) {
    if (a >= na) break;
    const int g = a0 + a;
    const double rm = A.rm[g];
    sc[a] = N.ithm[g] == 0 ? s : 1.;  // exp(-dta2 * 0) for the others
#pragma unroll
    for (int d = 0; d < 3; d++) {
      p[a][d] = A.pos[d * n + g];
      v[a][d] = A.vel[d * n + g];
      aa[a][d] = A.aa[d * n + g];
      if (N.phase == 0) v[a][d] += A.dter * (A.fl[d * n + g] * rm);
    }
  }
  if (N.phase > 0) {
#pragma unroll
    for (int a = 0; a < NA; a++) {
      if (a >= na) break;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        v[a][d] *= sc[a];
        p[a][d] += dta2 * v[a][d];
      }
    }
    const int b0 = A.boff[m], b1 = A.boff[m + 1], e0 = A.aoff[m], e1 = A.aoff[m + 1];
    if ((b1 > b0) || (e1 > e0)) {
      if constexpr (NA <= 3)
        mol_bonded_r<NA>(*reinterpret_cast<double (*)[NA][3]>(p),
                         *reinterpret_cast<double (*)[NA][3]>(aa), a0, b0, b1, e0, e1, A.bi,
                         A.bj, A.bk, A.br0, A.ai, A.aj, A.ak, A.afk, A.ath0);
      else
        mol_bonded(p, aa, na, a0, b0, b1, e0, e1, A.bi, A.bj, A.bk, A.br0, A.ai, A.aj, A.ak,
                   A.afk, A.ath0);
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (a >= na) break;
        const double rm = A.rm[a0 + a];
#pragma unroll
        for (int d = 0; d < 3; d++) aa[a][d] *= rm;
      }
    }
#pragma unroll
    for (int a = 0; a < NA; a++) {
      if (a >= na) break;
#pragma unroll
      for (int d = 0; d < 3; d++) v[a][d] += dta2 * aa[a][d];
    }
  }
  if (N.phase == A.ninnra) {
    bool moved = false;
#pragma unroll
    for (int a = 0; a < NA; a++) {
      if (a >= na) break;
      moved |= md_atom_out(A, a0 + a, p[a][0], p[a][1], p[a][2], v[a][0], v[a][1], v[a][2],
                           aa[a][0], aa[a][1], aa[a][2]);
    }
    return moved;
  }
#pragma unroll
  for (int a = 0; a < NA; a++) {
    if (a >= na) break;
    const int g = a0 + a;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      v[a][d] += dta2 * aa[a][d];
      p[a][d] += dta2 * v[a][d];
      v[a][d] *= sc[a];
    }
    if (N.ithm[g] == 0)  // forcen_new's p2m (md_therm.new.c:186-188)
      p2m += N.amass[g] * (v[a][0] * v[a][0] + v[a][1] * v[a][1] + v[a][2] * v[a][2]);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      A.pos[d * n + g] = p[a][d];
      A.vel[d * n + g] = v[a][d];
      A.aa[d * n + g] = aa[a][d];
    }
  }
  return false;
}

static __device__ __forceinline__ void k_mdnhc_body(const NhcArgs &N, int bid) {
  const InnerArgs &A = N.I;
  const int m = bid * MD_TPB + threadIdx.x;
  const double s = exp(-A.dta2 * N.nhc[0]);
  bool moved = false;
  double p2m = 0.;
  if (m < A.nmol) {
    const int a0 = A.moff[m], na = A.moff[m + 1] - a0;
    if (na == 1) {
      double p[1][3], v[1][3], aa[1][3];
      moved = mdnhc_mol<1>(N, m, a0, na, s, p2m, p, v, aa);
    } else if (na <= 3) {
      double p[3][3], v[3][3], aa[3][3];
      moved = mdnhc_mol<3>(N, m, a0, na, s, p2m, p, v, aa);
    } else {
      extern __shared__ double mdl[];
      double (*p)[3] = (double (*)[3])(mdl + (size_t)threadIdx.x * 9 * MD_MAXA);
      moved = mdnhc_mol<MD_MAXA>(N, m, a0, na, s, p2m, p, p + MD_MAXA, p + 2 * MD_MAXA);
    }
  }
  if (N.phase == A.ninnra) {
    if (__ballot(moved) != 0 && lane_id() == 0) atomicOr(A.flag, 1);
    return;
  }
  // workgroup = one wave: its kinetic sum, then the chain update in the last
  // workgroup to arrive
  static_assert(MD_TPB == NB_WAVE, "one wave per workgroup");
  __shared__ int last;
  const double tot = nb_wave_sum(p2m);
  if (threadIdx.x == 0) N.ppart[bid] = tot;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(N.ticket, 1) == N.nblocks - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  double q = 0.;
  for (int b = threadIdx.x; b < N.nblocks; b += MD_TPB)
    q += __hip_atomic_load(&N.ppart[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  q = nb_wave_sum(q);
  if (threadIdx.x == 0) {
    nhc_update(N.nhc, N.nchain, q, A.dta2);
    *N.ticket = 0;
  }
}
NB_KERNEL_PAIR(k_mdnhc, NhcArgs, MD_TPB)

// cm3d's force(...,2) at step_init: bonded accelerations at the current
// positions (md_step.c:48, md_fcntr.c:79-105 with rmass).
__global__ __launch_bounds__(MD_TPB) void k_md_bonded(InnerArgs A) {
  __shared__ double sp[MD_TPB][MD_MAXA][3], sa[MD_TPB][MD_MAXA][3];
  int m = blockIdx.x * MD_TPB + threadIdx.x;
  if (m >= A.nmol) return;
  const int n = A.n, a0 = A.moff[m], na = A.moff[m + 1] - a0;
  double (*p)[3] = sp[threadIdx.x], (*aa)[3] = sa[threadIdx.x];
  for (int a = 0; a < na; a++)
    for (int d = 0; d < 3; d++) p[a][d] = A.pos[d * n + a0 + a];
  mol_bonded(p, aa, na, a0, A.boff[m], A.boff[m + 1], A.aoff[m], A.aoff[m + 1], A.bi, A.bj,
             A.bk, A.br0, A.ai, A.aj, A.ak, A.afk, A.ath0);
  for (int a = 0; a < na; a++) {
    double rm = A.rm[a0 + a];
    for (int d = 0; d < 3; d++) A.aa[d * n + a0 + a] = aa[a][d] * rm;
  }
}

// v += dter * fl / m; the workgroups past the kick reduce the energy
// columns of this step's force(...,1) and the first clears the rebuild flag
// for the next step.
struct KickArgs { double *vel; const double *fl, *rm; int n, nkick; double dter; int *flag; RedSet R; };
static __device__ __forceinline__ void k_mdkick_body(const KickArgs &A, int bid) {
  if (bid >= A.nkick) {
    reduce_col(A.R, bid - A.nkick);
    return;
  }
  if (bid == 0 && threadIdx.x == 0) *A.flag = 0;
  int i = bid * blockDim.x + threadIdx.x;
  if (i >= 3 * A.n) return;
  A.vel[i] += A.dter * (A.fl[i] * A.rm[i % A.n]);
}
NB_KERNEL_PAIR(k_mdkick, KickArgs, 256)

// Kinetic tensor, bond and bend energies and the bond virial (get_p2mt,
// md_fcntr.c:332-354; getvbond / getvbend harmonic cases, md_fbond.c:165-207,
// md_fbend.c:318-338 -- the bend virial is compiled only with -DPRESSURE,
// which cm3d's Linux builds leave off).  Columns: kinetic xx yy zz xy xz yz,
// vbond, vbend, wbond, wbond tensor [9].
#define MD_NE 18
__global__ __launch_bounds__(256) void k_md_energy(InnerArgs A, const double *__restrict__ amass,
                                                   double *__restrict__ epart) {
  double acc[MD_NE];
#pragma unroll
  for (int k = 0; k < MD_NE; k++) acc[k] = 0.;
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < A.nmol) {
    const int n = A.n, a0 = A.moff[m], a1 = A.moff[m + 1];
    for (int g = a0; g < a1; g++) {
      const double am = amass[g];
      const double vx = A.vel[g], vy = A.vel[n + g], vz = A.vel[2 * n + g];
      acc[0] += am * (vx * vx);
      acc[1] += am * (vy * vy);
      acc[2] += am * (vz * vz);
      acc[3] += am * (vy * vx);
      acc[4] += am * (vz * vx);
      acc[5] += am * (vz * vy);
    }
    for (int b = A.boff[m]; b < A.boff[m + 1]; b++) {
      int i = A.bi[b], j = A.bj[b];
      double dx = A.pos[i] - A.pos[j], dy = A.pos[n + i] - A.pos[n + j];
      double dz = A.pos[2 * n + i] - A.pos[2 * n + j];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      double dr = r - A.br0[b];
      acc[6] += .5 * A.bk[b] * dr * dr;
      const double fk = A.bk[b] * dr / r;
      acc[8] -= fk * r * r;
      const double d[3] = {dx, dy, dz};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) acc[9 + 3 * a + c] -= fk * d[c] * d[a];
    }
    for (int e = A.aoff[m]; e < A.aoff[m + 1]; e++) {
      int i = A.ai[e], j = A.aj[e], k = A.ak[e];
      double ax = A.pos[i] - A.pos[j], ay = A.pos[n + i] - A.pos[n + j], az = A.pos[2 * n + i] - A.pos[2 * n + j];
      double bx = A.pos[k] - A.pos[j], by = A.pos[n + k] - A.pos[n + j], bz = A.pos[2 * n + k] - A.pos[2 * n + j];
      double c = (ax * bx + ay * by + az * bz) / sqrt((ax * ax + ay * ay + az * az) * (bx * bx + by * by + bz * bz));
      c = c < 1. ? c : 1.;
      c = c > -1. ? c : -1.;
      double th = fabs(c) >= 1. ? 1. : acos(c), dth = th - A.ath0[e];
      acc[7] += .5 * A.afk[e] * dth * dth;
    }
  }
  __shared__ double red[4][MD_NE];
  int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = 0; k < MD_NE; k++) {
    double v = nb_wave_sum(acc[k]);
    if (lane == 0) red[w][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < MD_NE)
    epart[blockIdx.x * MD_NE + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// ---------------------------------------------------------------------------
static thread_local char g_dummy;  // keeps this TU's statics distinct

#define MCHK(x)                                                        \
  do {                                                                 \
    hipError_t _e = (x);                                               \
    if (_e != hipSuccess) {                                            \
      nb_set_error("%s: %s", #x, hipGetErrorString(_e));               \
      return NB_ERR_HIP;                                               \
    }                                                                  \
  } while (0)

template <typename T>
static int mup(nb_ctx *c, T **p, const T *src, size_t cnt) {
  if (cnt == 0) cnt = 1;
  MCHK(hipMalloc((void **)p, cnt * sizeof(T)));
  c->dev_bytes += cnt * sizeof(T);
  if (src) MCHK(hipMemcpy(*p, src, cnt * sizeof(T), hipMemcpyHostToDevice));
  return NB_OK;
}

static InnerArgs inner_args(nb_md *md) {
  InnerArgs A;
  A.n = md->n; A.nmol = md->nmol; A.ninnra = md->ninnra; A.kick_first = 1;
  A.dter = md->dter; A.dta2 = md->dta2;
  A.pos = md->c->d_pos; A.vel = md->d_vel; A.aa = md->d_aa;
  A.fl = md->c->d_fl; A.rm = md->d_rm;
  A.moff = md->d_moff; A.boff = md->d_boff; A.aoff = md->d_aoff;
  A.bi = md->d_bi; A.bj = md->d_bj; A.ai = md->d_ai; A.aj = md->d_aj; A.ak = md->d_ak;
  A.bk = md->d_bk; A.br0 = md->d_br0; A.afk = md->d_afk; A.ath0 = md->d_ath0;
  nb_ctx *c = md->c;
  A.epilogue = 1;
  A.posb = c->d_posb; A.skin2q = .25 * c->skin * c->skin; A.flag = c->d_flags + 1;
  A.iperm = c->d_iperm; A.qpos = c->d_qpos; A.q = c->d_qch;
  A.rec64 = (double4 *)c->d_rec64;
  A.recf = c->prec == NB_PREC_MIXED ? (uint4 *)c->d_recf : nullptr;
  const bool ew = c->iperd == 3 && c->ktot > 0 && c->ncharge > 0;
  A.chg = ew ? (double4 *)c->d_chg : nullptr;
  A.stp = ew ? (double4 *)c->d_stp : nullptr;
  A.b = md->box;
  return A;
}

static void md_drop_graph(nb_md *md) {
  if (md->have_graph) {
    hipGraphExecDestroy(md->gexec);
    hipGraphDestroy(md->graph);
    md->have_graph = 0;
  }
}

static void md_free_nhc(nb_md *md) {
  void *ptrs[] = {md->d_nhc, md->d_ppart, md->d_ithm, md->d_ticket};
  for (void *q : ptrs)
    if (q) hipFree(q);
  md->d_nhc = md->d_ppart = nullptr;
  md->d_ithm = md->d_ticket = nullptr;
  md->c->dev_bytes -= md->nhc_bytes;
  md->nhc_bytes = 0;
  md->nchain = 0;
}

extern "C" void nb_md_destroy(nb_md *md) {
  if (!md) return;
  hipSetDevice(md->c->dev);
  hipDeviceSynchronize();  // no step of this state may still be running
  if (md->have_graph) { hipGraphExecDestroy(md->gexec); hipGraphDestroy(md->graph); }
  md->c->dev_bytes -= md->bytes;
  md->c->dev_bytes -= md->nhc_bytes;
  void *ptrs[] = {md->d_vel, md->d_rm, md->d_aa, md->d_amass, md->d_moff, md->d_boff,
                  md->d_aoff, md->d_bi, md->d_bj, md->d_ai, md->d_aj, md->d_ak, md->d_bk,
                  md->d_br0, md->d_afk, md->d_ath0, md->d_ep, md->d_eo, md->d_nhc,
                  md->d_ppart, md->d_ithm, md->d_ticket, md->d_save};
  for (void *p : ptrs)
    if (p) hipFree(p);
  delete md;
}

extern "C" int nb_md_create(nb_ctx *c, const nb_md_params *p, nb_md **out) {
  if (!c || !p || !out || !p->amass || p->nmol <= 0 || !p->mol_start) {
    nb_set_error("nb_md_create: bad argument");
    return NB_ERR_ARG;
  }
  if (p->nres_lrf != 1) {
    nb_set_error("nb_md_create: only nres_lrf == 1 is supported");
    return NB_ERR_UNSUPPORTED;
  }
  const int n = c->natoms, nm = p->nmol;
  if (p->mol_start[0] != 0 || p->mol_start[nm] != n) {
    nb_set_error("nb_md_create: molecules must cover atoms 0..natoms-1 contiguously");
    return NB_ERR_ARG;
  }
  std::vector<int> molof(n);
  int maxa = 1;
  for (int m = 0; m < nm; m++) {
    int na = p->mol_start[m + 1] - p->mol_start[m];
    if (na <= 0 || na > MD_MAXA) {
      nb_set_error("nb_md_create: molecule %d has %d atoms (1..%d supported)", m, na, MD_MAXA);
      return NB_ERR_UNSUPPORTED;
    }
    for (int a = p->mol_start[m]; a < p->mol_start[m + 1]; a++) molof[a] = m;
    if (na > maxa) maxa = na;
  }
  // group bonded terms by molecule, keeping their order within a molecule
  std::vector<int> boff(nm + 1, 0), aoff(nm + 1, 0), border, aorder;
  for (int b = 0; b < p->nbond; b++) {
    int i = p->bond_i[b], j = p->bond_j[b];
    if (i < 0 || i >= n || j < 0 || j >= n || molof[i] != molof[j]) {
      nb_set_error("nb_md_create: bond %d crosses molecules or is out of range", b);
      return NB_ERR_ARG;
    }
    boff[molof[i] + 1]++;
  }
  for (int e = 0; e < p->nbend; e++) {
    int i = p->bend_i[e], j = p->bend_j[e], k = p->bend_k[e];
    if (i < 0 || i >= n || j < 0 || j >= n || k < 0 || k >= n || molof[i] != molof[j] ||
        molof[k] != molof[j]) {
      nb_set_error("nb_md_create: bend %d crosses molecules or is out of range", e);
      return NB_ERR_ARG;
    }
    aoff[molof[j] + 1]++;
  }
  for (int m = 0; m < nm; m++) { boff[m + 1] += boff[m]; aoff[m + 1] += aoff[m]; }
  std::vector<int> bi(p->nbond), bj(p->nbond), ai(p->nbend), aj(p->nbend), ak(p->nbend);
  std::vector<double> bk(p->nbond), br0(p->nbond), afk(p->nbend), ath0(p->nbend);
  {
    std::vector<int> fill(boff.begin(), boff.end() - 1);
    for (int b = 0; b < p->nbond; b++) {
      int s = fill[molof[p->bond_i[b]]]++;
      bi[s] = p->bond_i[b]; bj[s] = p->bond_j[b]; bk[s] = p->bond_k[b]; br0[s] = p->bond_r0[b];
    }
    std::vector<int> fa(aoff.begin(), aoff.end() - 1);
    for (int e = 0; e < p->nbend; e++) {
      int s = fa[molof[p->bend_j[e]]]++;
      ai[s] = p->bend_i[e]; aj[s] = p->bend_j[e]; ak[s] = p->bend_k[e];
      afk[s] = p->bend_fk[e]; ath0[s] = p->bend_th0[e];
    }
  }
  nb_md *md = new nb_md();
  memset(md, 0, sizeof(*md));
  md->c = c; md->n = n; md->nmol = nm; md->nbond = p->nbond; md->nbend = p->nbend;
  md->maxa = maxa;
  // the structure factor on the matrix cores keeps every SIMD's registers
  // (two 242-register waves) for most of the step: a list build and pair
  // kernel beside it wait for slots and slow it down, so then the two
  // branches run one after the other (98,304-atom water: 2.49 ms a step
  // against 2.57 forked; 12,288 atoms with the VALU passes: 0.112 forked
  // against 0.121)
  md->fork = c->sfm ? 0 : 1;
  md->ninnra = p->nres_intra > 0 ? p->nres_intra : 1;
  // md_step.c:73-78 with ninter = ninrra = 1
  md->dt = p->dt;
  double dtl2 = .5 * p->dt, dts2 = dtl2 / 1., dtr2 = dts2 / 1.;
  md->dta2 = dtr2 / (double)md->ninnra;
  md->dter = 1. * dts2;
  std::vector<double> rm(n);
  for (int i = 0; i < n; i++) rm[i] = 1. / p->amass[i];
  int rc;
  hipSetDevice(c->dev);
  const size_t bytes0 = c->dev_bytes;
#define MUP(ptr, src, cnt) \
  if ((rc = mup(c, ptr, src, cnt))) { md->bytes = c->dev_bytes - bytes0; nb_md_destroy(md); return rc; }
  MUP(&md->d_vel, (const double *)nullptr, 3 * (size_t)n);
  MUP(&md->d_aa, (const double *)nullptr, 3 * (size_t)n);
  MUP(&md->d_rm, rm.data(), n);
  MUP(&md->d_amass, p->amass, n);
  MUP(&md->d_moff, p->mol_start, nm + 1);
  MUP(&md->d_boff, boff.data(), nm + 1);
  MUP(&md->d_aoff, aoff.data(), nm + 1);
  MUP(&md->d_bi, bi.data(), bi.size()); MUP(&md->d_bj, bj.data(), bj.size());
  MUP(&md->d_bk, bk.data(), bk.size()); MUP(&md->d_br0, br0.data(), br0.size());
  MUP(&md->d_ai, ai.data(), ai.size()); MUP(&md->d_aj, aj.data(), aj.size());
  MUP(&md->d_ak, ak.data(), ak.size()); MUP(&md->d_afk, afk.data(), afk.size());
  MUP(&md->d_ath0, ath0.data(), ath0.size());
  md->ep_blocks = (nm + 255) / 256;
  MUP(&md->d_ep, (const double *)nullptr, MD_NE * (size_t)md->ep_blocks);
  MUP(&md->d_eo, (const double *)nullptr, 8);
  // a batch tick's snapshot: pos, vel, aa, fl (12N) and the chain's veta, eta
  MUP(&md->d_save, (const double *)nullptr, 12 * (size_t)n + 2 * MD_MAXCHAIN);
#undef MUP
  md->bytes = c->dev_bytes - bytes0;
  (void)g_dummy;
  *out = md;
  return NB_OK;
}

extern "C" int nb_md_set_state(nb_md *md, const double *pos, const double *vel,
                               const double hmat[9]) {
  if (!md || !pos || !vel || !hmat) { nb_set_error("nb_md_set_state: null argument"); return NB_ERR_ARG; }
  nb_ctx *c = md->c;
  MCHK(hipSetDevice(c->dev));
  // steps may still run on any stream (a batch's, a caller's): wait for the
  // device before overwriting the state and dropping the captured graph (a
  // state that never stepped has nothing in flight, and a server admitting
  // it must not wait for the running batch)
  if (md->used) MCHK(hipDeviceSynchronize());
  MCHK(hipMemcpyAsync(c->d_pos, pos, sizeof(double) * 3 * md->n, hipMemcpyHostToDevice,
                      c->stream));
  MCHK(hipMemcpyAsync(md->d_vel, vel, sizeof(double) * 3 * md->n, hipMemcpyHostToDevice,
                      c->stream));
  MCHK(hipStreamSynchronize(c->stream));
  memcpy(md->hmat, hmat, sizeof(md->hmat));
  md_drop_graph(md);
  c->recip_owner = nullptr;
  return NB_OK;
}

extern "C" int nb_md_get_state(nb_md *md, double *pos, double *vel) {
  if (!md) { nb_set_error("nb_md_get_state: null argument"); return NB_ERR_ARG; }
  nb_ctx *c = md->c;
  MCHK(hipSetDevice(c->dev));
  MCHK(hipDeviceSynchronize());  // the step may have run on another stream
  if (pos) MCHK(hipMemcpy(pos, c->d_pos, sizeof(double) * 3 * md->n, hipMemcpyDeviceToHost));
  if (vel) MCHK(hipMemcpy(vel, md->d_vel, sizeof(double) * 3 * md->n, hipMemcpyDeviceToHost));
  return NB_OK;
}

// The forces save_force prints (output/md_saveout.c:65-90): fl, the long
// (nonbonded) force of the current state before rmass, and aa, the
// intramolecular accelerations (cm3d's fxl * amass and fxa).
extern "C" int nb_md_get_forces(nb_md *md, double *fl, double *aa) {
  if (!md) { nb_set_error("nb_md_get_forces: null argument"); return NB_ERR_ARG; }
  nb_ctx *c = md->c;
  MCHK(hipSetDevice(c->dev));
  MCHK(hipDeviceSynchronize());  // the step may have run on another stream
  if (fl) MCHK(hipMemcpy(fl, c->d_fl, sizeof(double) * 3 * md->n, hipMemcpyDeviceToHost));
  if (aa) MCHK(hipMemcpy(aa, md->d_aa, sizeof(double) * 3 * md->n, hipMemcpyDeviceToHost));
  return NB_OK;
}

// set_therm + samveta's result (md_therm.new.c:158-169, 248-275) for one
// "global" thermostat: eta starts at 0 and veta at veta0.
extern "C" int nb_md_set_thermostat(nb_md *md, int nchain, const double *gkt, const double *mnh,
                                    const double *veta0, const double *eta0, const int *ithm) {
  if (!md || nchain < 0 || nchain > MD_MAXCHAIN) {
    nb_set_error("nb_md_set_thermostat: bad argument (nchain 0..%d)", MD_MAXCHAIN);
    return NB_ERR_ARG;
  }
  nb_ctx *c = md->c;
  MCHK(hipSetDevice(c->dev));
  // a step of this state may be running on any stream (a batch's, a
  // caller's): wait for the device before freeing the chain's buffers
  MCHK(hipDeviceSynchronize());
  md_drop_graph(md);
  md_free_nhc(md);
  md->tgen++;  // the kernel sequence and the chain's buffers change
  if (nchain == 0) return NB_OK;
  if (!gkt || !mnh || !veta0 || !ithm) {
    nb_set_error("nb_md_set_thermostat: null argument");
    return NB_ERR_ARG;
  }
  for (int i = 0; i < nchain; i++)
    if (!(mnh[i] > 0.)) {
      nb_set_error("nb_md_set_thermostat: mnh[%d] = %g must be positive", i, mnh[i]);
      return NB_ERR_ARG;
    }
  for (int i = 0; i < md->n; i++)
    if (ithm[i] != 0 && ithm[i] != -1) {
      nb_set_error("nb_md_set_thermostat: ithm[%d] = %d (0: global thermostat, -1: none)", i,
                   ithm[i]);
      return NB_ERR_UNSUPPORTED;
    }
  std::vector<double> h(4 * nchain, 0.);
  for (int i = 0; i < nchain; i++) {
    h[i] = veta0[i];
    h[nchain + i] = eta0 ? eta0[i] : 0.;
    h[2 * nchain + i] = gkt[i];
    h[3 * nchain + i] = mnh[i];
  }
  const int nblk = (md->nmol + MD_TPB - 1) / MD_TPB;
  const size_t b0 = c->dev_bytes;
  int rc = 0;
  const int zero = 0;
  if ((rc = mup(c, &md->d_nhc, h.data(), h.size())) || (rc = mup(c, &md->d_ithm, ithm, md->n)) ||
      (rc = mup(c, &md->d_ppart, (const double *)nullptr, nblk)) ||
      (rc = mup(c, &md->d_ticket, &zero, 1))) {
    md->nhc_bytes = c->dev_bytes - b0;
    md_free_nhc(md);
    return rc;
  }
  md->nhc_bytes = c->dev_bytes - b0;
  md->nchain = nchain;
  return NB_OK;
}

extern "C" int nb_md_get_thermostat(nb_md *md, double *veta, double *eta) {
  if (!md) { nb_set_error("nb_md_get_thermostat: null argument"); return NB_ERR_ARG; }
  const int nc = md->nchain;
  if (nc == 0) return NB_OK;
  nb_ctx *c = md->c;
  MCHK(hipSetDevice(c->dev));
  MCHK(hipStreamSynchronize(c->stream));
  std::vector<double> h(2 * nc);
  MCHK(hipMemcpy(h.data(), md->d_nhc, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < nc; i++) {
    if (veta) veta[i] = h[i];
    if (eta) eta[i] = h[nc + i];
  }
  return NB_OK;
}


// step_init (md_step.c:32-58): lists, force(...,1), force(...,2).
extern "C" int nb_md_init(nb_md *md) {
  if (!md) { nb_set_error("nb_md_init: null argument"); return NB_ERR_ARG; }
  nb_ctx *c = md->c;
  MCHK(hipSetDevice(c->dev));
  if (md->used) {  // a step of this state may be running on any stream
    MCHK(hipDeviceSynchronize());
    md_drop_graph(md);
  }
  int rc = nb_build_now(c, md->hmat, c->stream);
  if (rc) return rc;
  md->box = c->box_built;
  rc = nb_launch_force(c, 1, md->box, nullptr, c->d_fl, c->stream, 0);
  if (rc) return rc;
  c->recip_owner = md;
  InnerArgs A = inner_args(md);
  k_md_bonded<<<(md->nmol + MD_TPB - 1) / MD_TPB, MD_TPB, 0, c->stream>>>(A);
  MCHK(hipGetLastError());
  // the step's rebuild flag starts clear (its inner kernel sets it)
  MCHK(hipMemsetAsync(c->d_flags + 1, 0, sizeof(int), c->stream));
  MCHK(hipStreamSynchronize(c->stream));
  md->nrebuild_host = 0;
  return NB_OK;
}

int md_enqueue_step(nb_md *md, hipStream_t s) {
  nb_ctx *c = md->c;
  md->used = 1;
  // inner RESPA loop + nrefsh test + records for the force evaluation
  InnerArgs A = inner_args(md);
  const int lds = md->maxa > 3 ? MD_TPB * 9 * MD_MAXA * (int)sizeof(double) : 0;
  const int nblk = (md->nmol + MD_TPB - 1) / MD_TPB;
  if (md->nchain == 0) {
    NB_GO(c, K_MDINNER, k_mdinner, nblk, MD_TPB, lds, A, s);
  } else {
    NhcArgs N;
    N.I = A;
    N.nchain = md->nchain; N.nblocks = nblk;
    N.ithm = md->d_ithm; N.amass = md->d_amass;
    N.nhc = md->d_nhc; N.ppart = md->d_ppart; N.ticket = md->d_ticket;
    for (int ph = 0; ph <= md->ninnra; ph++) {
      N.phase = ph;
      NB_GO(c, K_MDNHC, k_mdnhc, nblk, MD_TPB, lds, N, s);
    }
  }
  // the reciprocal sum needs only the charge records the inner kernel just
  // wrote, not the lists: it runs on a second stream beside the conditional
  // list build and the pair kernel (a fork / join in the captured graph),
  // and k_ewfin joins the two.  A batch records the launches in one
  // sequence (its reciprocal sum runs after the pair kernel).
  // forces only: the step's energies are evaluated when asked for
  // (nb_md_energies), as cm3d's output steps read them
  const int fp = NB_FORCE_PACKED | NB_FORCE_NORED | NB_FORCE_NOENER;
  int rc;
  if (md->fork && !c->rec) {
    MCHK(hipEventRecord(c->fork_ev[0], s));
    MCHK(hipStreamWaitEvent(c->stream2, c->fork_ev[0], 0));
    rc = nb_launch_force(c, 1, md->box, nullptr, c->d_fl, c->stream2,
                         fp | NB_FORCE_NO_PAIR | NB_FORCE_NO_FIN);
    if (rc) return rc;
    MCHK(hipEventRecord(c->fork_ev[1], c->stream2));
    rc = nb_launch_build_cond(c, md->box, s);
    if (rc) return rc;
    rc = nb_launch_force(c, 1, md->box, nullptr, c->d_fl, s,
                         fp | NB_FORCE_NO_RECIP | NB_FORCE_NO_FIN);
    if (rc) return rc;
    MCHK(hipStreamWaitEvent(s, c->fork_ev[1], 0));
    rc = nb_launch_force(c, 1, md->box, nullptr, c->d_fl, s, fp | NB_FORCE_NO_PAIR | NB_FORCE_NO_RECIP);
    if (rc) return rc;
  } else {
    rc = nb_launch_build_cond(c, md->box, s);
    if (rc) return rc;
    rc = nb_launch_force(c, 1, md->box, nullptr, c->d_fl, s, fp);
    if (rc) return rc;
  }
  // kick + energy-column reductions + flag reset
  const int nkick = (3 * md->n + 255) / 256;
  KickArgs K = {md->d_vel, c->d_fl, md->d_rm, md->n, nkick, md->dter, c->d_flags + 1,
                nb_red_set(c, 1, false)};
  NB_GO(c, K_MDKICK, k_mdkick, nkick + K.R.ncols, 256, 0, K, s);
  if (!c->rec) MCHK(hipGetLastError());
  return NB_OK;
}

// Snapshot / restore of the dynamic state around a batch tick: pos, vel,
// aa (inner accelerations), fl (the long force the next step's first kick
// uses) and the chain's veta, eta.  The snapshot also clears the sticky
// overflow flag, so after the tick it says whether any list build of the
// tick overflowed.
struct SaveArgs {
  double *pos, *vel, *aa, *fl, *nhc, *save;
  int n, nnhc, restore, nblocks;
  int *flags;
};
static __device__ __forceinline__ void k_mdsave_body(const SaveArgs &A, int bid) {
  if (bid == 0 && threadIdx.x == 0 && !A.restore) A.flags[2] = 0;
  const int i = bid * blockDim.x + threadIdx.x, n3 = 3 * A.n;
  double *st;
  int k;
  if (i < n3) { st = A.pos; k = i; }
  else if (i < 2 * n3) { st = A.vel; k = i - n3; }
  else if (i < 3 * n3) { st = A.aa; k = i - 2 * n3; }
  else if (i < 4 * n3) { st = A.fl; k = i - 3 * n3; }
  else if (i < 4 * n3 + A.nnhc) { st = A.nhc; k = i - 4 * n3; }
  else return;
  if (A.restore) st[k] = A.save[i];
  else A.save[i] = st[k];
}
NB_KERNEL_PAIR(k_mdsave, SaveArgs, 256)

// Per system of a batch: out[0] = the largest row a list build of the tick
// wanted, if it exceeded the slots (0 otherwise).
struct OvfArgs { const int *flags; int cap; int *out; };
static __device__ __forceinline__ void k_ovfchk_body(const OvfArgs &A, int bid) {
  if (bid == 0 && threadIdx.x == 0) {
    const int m = A.flags[2];
    *A.out = m > A.cap ? m : 0;
  }
}
NB_KERNEL_PAIR(k_ovfchk, OvfArgs, 64)

// The snapshot buffer is allocated with the state (nb_md_create), sized
// for the longest chain, so a server's footprint accounting sees it.
int md_ensure_save(nb_md *md) {
  if (!md->d_save) {
    nb_set_error("nb_batch: MD state has no snapshot buffer");
    return NB_ERR_ARG;
  }
  return NB_OK;
}

// Snapshot (restore 0) or restore (1) on stream s; recordable by a batch.
int md_enqueue_save(nb_md *md, hipStream_t s, int restore) {
  nb_ctx *c = md->c;
  SaveArgs A;
  A.pos = c->d_pos; A.vel = md->d_vel; A.aa = md->d_aa; A.fl = c->d_fl; A.nhc = md->d_nhc;
  A.save = md->d_save; A.n = md->n; A.nnhc = 2 * md->nchain; A.restore = restore;
  A.flags = c->d_flags;
  A.nblocks = (int)((12 * (size_t)md->n + A.nnhc + 255) / 256);
  NB_GO(c, K_MDSAVE, k_mdsave, A.nblocks, 256, 0, A, s);
  if (!c->rec) MCHK(hipGetLastError());
  return NB_OK;
}

int md_enqueue_ovfchk(nb_md *md, hipStream_t s, int *d_out) {
  nb_ctx *c = md->c;
  OvfArgs A = {c->d_flags, c->cap, d_out};
  NB_GO(c, K_OVFCHK, k_ovfchk, 1, 64, 0, A, s);
  if (!c->rec) MCHK(hipGetLastError());
  return NB_OK;
}

void nb_md_gens(nb_md *md, unsigned *cgen, unsigned *tgen) {
  *cgen = md->c->gen;
  *tgen = md->tgen;
}

int nb_grow_capacity(nb_ctx *c, int want);

// After an overflowed tick: back to the snapshot, neighbour slots for the
// largest row seen (x1.25 + 64, as a host build grows them), lists rebuilt
// at the restored positions, and the nsteps redone for this system alone
// (bitwise what the batch computes for it).  Synchronous.
int md_recover(nb_md *md, int nsteps, int want, hipStream_t s) {
  nb_ctx *c = md->c;
  for (int attempt = 0; attempt < 4; attempt++) {
    int rc = md_enqueue_save(md, s, 1);
    if (rc) return rc;
    MCHK(hipStreamSynchronize(s));
    rc = nb_grow_capacity(c, (int)(want * 1.25) + 64);
    if (rc) return rc;
    rc = nb_build_now(c, md->hmat, s);
    if (rc) return rc;
    MCHK(hipMemsetAsync(c->d_flags + 1, 0, 2 * sizeof(int), s));  // rebuild + sticky flags
    md_drop_graph(md);
    rc = nb_md_step(md, nsteps, s, 0);
    if (rc) return rc;
    int f2 = 0;
    MCHK(hipMemcpyAsync(&f2, c->d_flags + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    MCHK(hipStreamSynchronize(s));
    if (f2 <= c->cap) return NB_OK;
    want = f2;
  }
  nb_set_error("neighbour list overflow persists after regrowing the slots (%d > %d)", want,
               c->cap);
  return NB_ERR_CAPACITY;
}

nb_ctx *nb_md_ctx(nb_md *md) { return md->c; }

static struct RegisterMD {
  RegisterMD() {
    g_batch_fn[K_MDINNER] = k_mdinner_bl;
    g_batch_fn[K_MDKICK] = k_mdkick_bl;
    g_batch_fn[K_MDNHC] = k_mdnhc_bl;
    g_batch_fn[K_MDSAVE] = k_mdsave_bl;
    g_batch_fn[K_OVFCHK] = k_ovfchk_bl;
  }
} g_register_md;

extern "C" int nb_md_step(nb_md *md, int nsteps, void *stream, int use_graph) {
  if (!md || nsteps < 0) { nb_set_error("nb_md_step: bad argument"); return NB_ERR_ARG; }
  nb_ctx *c = md->c;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  MCHK(hipSetDevice(c->dev));
  if (md->have_graph && (md->graph_cgen != c->gen || md->graph_tgen != md->tgen)) {
    // recorded against buffers or a launch shape that have since changed
    MCHK(hipDeviceSynchronize());
    md_drop_graph(md);
  }
  if (use_graph && !md->have_graph) {
    MCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = md_enqueue_step(md, s);
    hipGraph_t g;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    MCHK(e);
    MCHK(hipGraphInstantiate(&md->gexec, g, nullptr, nullptr, 0));
    md->graph = g;
    md->have_graph = 1;
    md->graph_cgen = c->gen;
    md->graph_tgen = md->tgen;
  }
  for (int k = 0; k < nsteps; k++) {
    if (use_graph) {
      MCHK(hipGraphLaunch(md->gexec, s));
    } else {
      int rc = md_enqueue_step(md, s);
      if (rc) return rc;
    }
  }
  // the step's last force(...,1) is at the positions it leaves
  if (nsteps > 0) c->recip_owner = md;
  return NB_OK;
}

// Energies of the current state (getvireal + geteng, md_fcntr.c:136-231 and
// 259-306): poter = U (+ vlrc/V), poth = US (+ vlrc/V, US with vhinter),
// potra = vbond + vbend, zke, the chain's potn / zken, cm3d's conserved
// ham = poth + potra + zke + potn + zken (md_averages.new.c:100), and
// prsi = (p2m + W + WI) / (3V) with W less 3 wlrc/V.
extern "C" int nb_md_energies(nb_md *md, double vlrc, double wlrc, int ndof, nb_md_energy *e) {
  if (!md || !e) { nb_set_error("nb_md_energies: null argument"); return NB_ERR_ARG; }
  nb_ctx *c = md->c;
  hipStream_t s = c->stream;
  MCHK(hipSetDevice(c->dev));
  // the last step may have run on another stream (a batch's, a caller's);
  // the force rerun below rewrites the records and energy columns it uses
  MCHK(hipDeviceSynchronize());
  // nonbonded energies and virials of the current positions: force(...,1)'s
  // energy pass (FP64 pair sums, the reciprocal sum without its force pass,
  // ecorr) into scratch (fs), so the integrator's fl is untouched.  When the
  // last force(...,1) on this context was this state's, at these positions
  // (a step's, or init's), its reciprocal-sum and ecorr energy columns are
  // reduced as they are: the same kernels on the same inputs would
  // recompute them bit for bit
  const int reuse = c->recip_owner == md ? NB_FORCE_REUSE_RECIP : 0;
  int rc = nb_launch_force(c, 1, md->box, nullptr, c->d_fs, s, NB_FORCE_EONLY | reuse);
  if (rc) return rc;
  c->recip_owner = md;
  InnerArgs A = inner_args(md);
  k_md_energy<<<md->ep_blocks, 256, 0, s>>>(A, md->d_amass, md->d_ep);
  MCHK(hipGetLastError());
  std::vector<double> ep(MD_NE * md->ep_blocks);
  double E[64];
  MCHK(hipMemcpyAsync(ep.data(), md->d_ep, ep.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  MCHK(hipMemcpyAsync(E, c->d_eout, sizeof(E), hipMemcpyDeviceToHost, s));
  MCHK(hipMemcpyAsync(c->h_flags, c->d_flags, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
  MCHK(hipStreamSynchronize(s));
  double col[MD_NE] = {0.};
  for (int b = 0; b < md->ep_blocks; b++)
    for (int k = 0; k < MD_NE; k++) col[k] += ep[MD_NE * (size_t)b + k];
  // p2m_tensor (get_p2mt) and its trace, as geteng forms it
  const double p2mt[9] = {col[0], col[3], col[4], col[3], col[1], col[5], col[4], col[5], col[2]};
  const double p2m = p2mt[0] + p2mt[4] + p2mt[8];
  const double vb = col[6], va = col[7], wb = col[8];
  const int nc = md->nchain;
  std::vector<double> nhc(4 * nc);
  if (nc) MCHK(hipMemcpy(nhc.data(), md->d_nhc, nhc.size() * sizeof(double), hipMemcpyDeviceToHost));
  nb_energy ne;
  nb_fill_energy(c, 1, E, md->hmat, &ne);
  const double vol = nb_det9(md->hmat);
  memset(e, 0, sizeof(*e));
  e->vinter = ne.spotl;
  e->vhinter = ne.spothl;
  e->velec = ne.spote;
  e->winter = ne.sprsl;
  e->welec = ne.sprse;
  if (c->iperd == 3) {  // getvewald (md_ewald.c:364-373)
    e->vewald = ne.vk_ewald + ne.pot_ecorr;
    e->wewald = ne.vp_ewald + ne.prs_ecorr;
  }
  e->vbond = vb;
  e->vbend = va;
  e->wbond = wb;
  e->zke = .5 * p2m;
  e->poter = e->vinter + e->velec + e->vewald;
  e->poth = e->vhinter + e->velec + e->vewald;
  double W = e->winter + e->welec + e->wewald;
  double Wt[9];
  for (int k = 0; k < 9; k++) {
    Wt[k] = ne.sprslt[k] + ne.sprset[k];
    if (c->iperd == 3) Wt[k] += ne.vp_ewald_tensor[k] + ne.prs_ecorr_tensor[k];
  }
  if (c->iperd == 3) {
    e->poter += vlrc / vol;
    e->poth += vlrc / vol;
    const double rvol = 1. / vol;
    W -= 3. * wlrc * rvol;
    Wt[0] -= wlrc * rvol;
    Wt[4] -= wlrc * rvol;
    Wt[8] -= wlrc * rvol;
  }
  for (int k = 0; k < 9; k++) e->prs_tensor[k] = (p2mt[k] + Wt[k] + col[9 + k]) / vol;
  e->potra = vb + va;
  // getengn_new (md_therm.new.c:200-222)
  for (int i = 0; i < nc; i++) e->potn += nhc[2 * nc + i] * nhc[nc + i];
  double sv = 0.;
  for (int i = 0; i < nc; i++) sv += nhc[3 * nc + i] * nhc[i] * nhc[i];
  e->zken = .5 * sv;
  e->ham = e->poth + e->potra + e->zke + e->potn + e->zken;
  e->prsi = (p2m + W + wb) / (3. * vol);
  e->tiout = ndof > 0 ? p2m / (double)ndof : 0.;
  // the last build's row maximum, or any since the last batch tick began
  e->overflow = c->h_flags[0] > c->cap || c->h_flags[2] > c->cap;
  return NB_OK;
}

extern "C" int nb_md_rebuilds(nb_md *md, int *count) {
  if (!md || !count) { nb_set_error("nb_md_rebuilds: null argument"); return NB_ERR_ARG; }
  nb_ctx *c = md->c;
  MCHK(hipSetDevice(c->dev));
  MCHK(hipStreamSynchronize(c->stream));
  MCHK(hipDeviceSynchronize());
  int f[8];
  MCHK(hipMemcpy(f, c->d_flags, sizeof(f), hipMemcpyDeviceToHost));
  *count = f[6];
  if (f[0] > c->cap || f[2] > c->cap) {
    nb_set_error("neighbour list overflow in a device-side rebuild (%d > %d slots)",
                 f[0] > f[2] ? f[0] : f[2], c->cap);
    return NB_ERR_CAPACITY;
  }
  return NB_OK;
}
