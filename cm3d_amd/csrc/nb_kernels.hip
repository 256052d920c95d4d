// nb_kernels.hip -- gfx950 kernels of the nonbonded path (FP64 form) and
// their launchers.
//
//   list build   k_cell -> k_scan -> k_scat -> k_order -> k_pack64 -> k_list
//                (cm3d get_ngbr / get_ngbr_e, energy/control/ngbr_list.c:297,
//                 elec_ngbr.c:220, set_link.c:203)
//   pair forces  k_pair64_* (force_npol md_fnopol.c:46, force_e finter_e.c:42),
//                k_ecorr (ecorr md_ewald.c:385); the mixed form is nb_mixed.hip
//   reciprocal   k_ewchg -> k_ewsfac -> k_ewkred -> k_ewforce
//                (fk_ewald md_ewald.c:185)
//   reductions   k_red_* (the getsvirter / getsvirter_e / getvewald sums)
//
// Every kernel is a __device__ body plus a single-system and a batched entry
// point (nb_launch.h); none allocates.
#include <stdarg.h>
#include <stdio.h>

#include "nb_launch.h"

BatchFn g_batch_fn[K_COUNT];

static __device__ __forceinline__ int lane_id() { return threadIdx.x & (NB_WAVE - 1); }

static __device__ __forceinline__ int mbcnt(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Sum NACC per-thread accumulators over a block of NW wavefronts (wave
// shuffles, then the wave totals in LDS in a fixed order) and store them to
// out[k * stride], k < NACC -- column-major partial tables, one column per
// accumulator, so the final reduction reads each column contiguously.
template <int NACC, int NW = 4>
static __device__ __forceinline__ void block_reduce_store(const double *acc, double *out,
                                                          int stride) {
  __shared__ double red[NW][NACC];
  const int w = threadIdx.x >> 6, lane = lane_id();
#pragma unroll
  for (int k = 0; k < NACC; k++) {
    double v = nb_wave_sum(acc[k]);
    if (NW == 1) {
      if (lane == 0) out[(size_t)k * stride] = v;
    } else if (lane == 0) {
      red[w][k] = v;
    }
  }
  if (NW == 1) return;
  __syncthreads();
  if (threadIdx.x < NACC) {
    double v = red[0][threadIdx.x];
#pragma unroll
    for (int q = 1; q < NW; q++) v += red[q][threadIdx.x];
    out[(size_t)threadIdx.x * stride] = v;
  }
}

// Every list-build kernel takes `cond`: when non-null and *cond == 0 it
// returns at once, so a device-side rebuild test can gate a captured build.
#define NB_SKIP(cond) \
  if ((cond) && *(cond) == 0) return

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------
struct ZeroArgs { int *p; int n; const int *cond; };
static __device__ __forceinline__ void k_zero_body(const ZeroArgs &A, int bid) {
  NB_SKIP(A.cond);
  int i = bid * blockDim.x + threadIdx.x;
  if (i < A.n) A.p[i] = 0;
}
NB_KERNEL_PAIR(k_zero, ZeroArgs, 256)

struct CopyArgs { double *dst; const double *src; int n; const int *cond; int *counter; };
static __device__ __forceinline__ void k_copy_body(const CopyArgs &A, int bid) {
  NB_SKIP(A.cond);
  int i = bid * blockDim.x + threadIdx.x;
  if (i < A.n) A.dst[i] = A.src[i];
  if (i == 0 && A.counter) atomicAdd(A.counter, 1);
}
NB_KERNEL_PAIR(k_copy, CopyArgs, 256)

// ---------------------------------------------------------------------------
// Cell sort.  Cells tile the box in fractional coordinates with widths no
// smaller than the longest list cutoff, so every pair within it under cm3d's
// minimum image lies in the 27 (or fewer, for thin boxes) neighbouring cells.
// Binning follows get_lnklist (set_link.c:218-236): wrap s -= anint(s), then
// (int)((s+.5)*n) clamped.
// ---------------------------------------------------------------------------
struct CellArgs {
  const double *pos; int n; Box b; int ncx, ncy, ncz; int *cell; int *count; const int *cond;
};
static __device__ __forceinline__ void k_cell_body(const CellArgs &A, int bid) {
  NB_SKIP(A.cond);
  int i = bid * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const int n = A.n;
  double x = A.pos[i], y = A.pos[n + i], z = A.pos[2 * n + i];
  const double *hi = A.b.hi;
  double sx = x * hi[0] + y * hi[1] + z * hi[2];
  double sy = x * hi[3] + y * hi[4] + z * hi[5];
  double sz = x * hi[6] + y * hi[7] + z * hi[8];
  sx -= nb_anint(sx);
  sy -= nb_anint(sy);
  sz -= nb_anint(sz);
  int a = (int)((sx + .5) * A.ncx), c = (int)((sy + .5) * A.ncy), e = (int)((sz + .5) * A.ncz);
  a = a < 0 ? 0 : (a > A.ncx - 1 ? A.ncx - 1 : a);
  c = c < 0 ? 0 : (c > A.ncy - 1 ? A.ncy - 1 : c);
  e = e < 0 ? 0 : (e > A.ncz - 1 ? A.ncz - 1 : e);
  int id = (a * A.ncy + c) * A.ncz + e;
  A.cell[i] = id;
  atomicAdd(&A.count[id], 1);
}
NB_KERNEL_PAIR(k_cell, CellArgs, 256)

// Exclusive scan of the cell counts in one workgroup of 1024 threads.
struct ScanArgs { const int *count; int ncells; int *start; int *fill; const int *cond; };
static __device__ __forceinline__ void k_scan_body(const ScanArgs &A, int) {
  NB_SKIP(A.cond);
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < A.ncells; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < A.ncells ? A.count[i] : 0;
    int x = v;
#pragma unroll
    for (int o = 1; o < NB_WAVE; o <<= 1) {
      int y = __shfl_up(x, o, NB_WAVE);
      if (lane_id() >= o) x += y;
    }
    int w = threadIdx.x >> 6;
    if (lane_id() == NB_WAVE - 1) wsum[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int k = 0; k < 16; k++) { int t = wsum[k]; wsum[k] = acc; acc += t; }
    }
    __syncthreads();
    int excl = carry + wsum[w] + x - v;
    if (i < A.ncells) { A.start[i] = excl; A.fill[i] = excl; }
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
}
NB_KERNEL_PAIR(k_scan, ScanArgs, 1024)

struct ScatArgs { const int *cell; int n; int *fill; int *tmp; const int *cond; };
static __device__ __forceinline__ void k_scat_body(const ScatArgs &A, int bid) {
  NB_SKIP(A.cond);
  int i = bid * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  int slot = atomicAdd(&A.fill[A.cell[i]], 1);
  A.tmp[slot] = i;
}
NB_KERNEL_PAIR(k_scat, ScatArgs, 256)

// Order each cell's atoms by atom index (the scatter's atomics leave them in
// arbitrary order): every member's rank is the number of smaller indices in
// the cell.  Makes the sorted order, and so every list, deterministic.
#define NB_CELL_LDS 4096
struct OrderArgs {
  const int *count, *start, *tmp; int *perm, *iperm, *scell; const int *cond;
};
static __device__ __forceinline__ void k_order_body(const OrderArgs &A, int bid) {
  NB_SKIP(A.cond);
  __shared__ int seg[NB_CELL_LDS];
  int c = bid, n = A.count[c], st = A.start[c];
  const int *src = A.tmp + st;
  if (n <= NB_CELL_LDS) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) seg[t] = src[t];
    __syncthreads();
    src = seg;
  }
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    int a = src[t], r = 0;
    for (int u = 0; u < n; u++) r += src[u] < a;
    A.perm[st + r] = a;
    A.iperm[a] = st + r;
    A.scell[st + r] = c;
  }
}
NB_KERNEL_PAIR(k_order, OrderArgs, 256)

// Sorted-order records refreshed at every evaluation: {x,y,z,q} in FP64.
struct Pack64Args {
  const double *pos, *q; const int *itype, *perm; int n; double4 *rec; int *stype;
  const int *cond;
};
static __device__ __forceinline__ void k_pack64_body(const Pack64Args &A, int bid) {
  NB_SKIP(A.cond);
  int s = bid * blockDim.x + threadIdx.x;
  if (s >= A.n) return;
  const int n = A.n, a = A.perm[s];
  A.rec[s] = make_double4(A.pos[a], A.pos[n + a], A.pos[2 * n + a], A.q[a]);
  if (A.stype) A.stype[s] = A.itype[a];
}
NB_KERNEL_PAIR(k_pack64, Pack64Args, 256)

// ---------------------------------------------------------------------------
// Verlet list build, one wavefront per (sorted) atom.  Each wave walks the
// neighbouring cells 64 candidates at a time, applies cm3d's filters in FP64
// exactly as ngbr_list / ngbr_liste do (null map, exclusion, minimum image,
// strict r^2 < (rc+skin)^2, s nested in t), and compacts the survivors with a
// ballot + mbcnt prefix into the atom's slots, so the order is the candidate
// order.  Both directions of every pair are stored.
// ---------------------------------------------------------------------------
struct ListArgs {
  int n, ntypes, cap, ncx, ncy, ncz;
  const double4 *rec;
  const int *stype, *perm, *scell, *cstart, *ccount;
  const int *xoff, *xlist, *map;
  const InterPar *ip;
  double rmaxs_skin2_e, rmaxl_skin2_e;
  Box b;
  uint32_t *nbr;
  int *nnb, *flags;
  const int *cond;
};

static __device__ __forceinline__ int wrapc(int v, int n) {
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

static __device__ __forceinline__ void k_list_body(const ListArgs &A, int bid) {
  NB_SKIP(A.cond);
  const int lane = lane_id();
  const int s = bid * 4 + (threadIdx.x >> 6);
  if (s >= A.n) return;
  const double4 ri = A.rec[s];
  const int ti = A.stype[s], ai = A.perm[s];
  const int c = A.scell[s];
  const int cz = c % A.ncz, cy = (c / A.ncz) % A.ncy, cx = c / (A.ncy * A.ncz);
  const int x0 = A.xoff[ai], x1 = A.xoff[ai + 1];
  const bool qi = ri.w != 0.0;
  const int rx = A.ncx >= 3 ? 1 : 0, ry = A.ncy >= 3 ? 1 : 0, rz = A.ncz >= 3 ? 1 : 0;
  int count = 0, hs = 0, ht = 0, hse = 0, hte = 0;
  uint32_t *out = A.nbr + (size_t)s * A.cap;
  for (int ox = -rx; ox <= rx; ox++)
    for (int oy = -ry; oy <= ry; oy++)
      for (int oz = -rz; oz <= rz; oz++) {
        int cc = (wrapc(cx + ox, A.ncx) * A.ncy + wrapc(cy + oy, A.ncy)) * A.ncz +
                 wrapc(cz + oz, A.ncz);
        int st = A.cstart[cc], nn = A.ccount[cc];
        for (int b0 = 0; b0 < nn; b0 += NB_WAVE) {
          int t = b0 + lane, s2 = st + t, aj = -1;
          unsigned fl = 0;
          if (t < nn && s2 != s) {
            double4 rj = A.rec[s2];
            int tj = A.stype[s2];
            aj = A.perm[s2];
            double dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
            nb_minimg_box(dx, dy, dz, A.b);
            double d2 = dx * dx + dy * dy + dz * dz;
            unsigned ex = 0;
            for (int e = x0; e < x1; e++) {
              unsigned v = (unsigned)A.xlist[e];
              if ((int)(v & 0x3FFFFFFFu) == aj) ex |= v >> 30;
            }
            int in = A.map[ti * A.ntypes + tj];
            if (in >= 0 && !(ex & 1u)) {
              const InterPar &P = A.ip[in];
              if (d2 < P.rmaxl_skin2) {
                fl |= NB_FLAG_VT;
                if (d2 < P.rmaxs_skin2) fl |= NB_FLAG_VS;
              }
            }
            if (qi && rj.w != 0.0 && !(ex & 2u)) {
              if (d2 < A.rmaxl_skin2_e) {
                fl |= NB_FLAG_ET;
                if (d2 < A.rmaxs_skin2_e) fl |= NB_FLAG_ES;
              }
            }
          }
          unsigned long long m = __ballot(fl != 0);
          if (fl) {
            int slot = count + mbcnt(m);
            if (slot < A.cap) out[slot] = (uint32_t)s2 | (fl << 28);
            if (ai < aj) {
              hs += (fl & NB_FLAG_VS) != 0;
              ht += (fl & NB_FLAG_VT) != 0;
              hse += (fl & NB_FLAG_ES) != 0;
              hte += (fl & NB_FLAG_ET) != 0;
            }
          }
          count += __popcll(m);
        }
      }
  hs = nb_wave_sum(hs);
  ht = nb_wave_sum(ht);
  hse = nb_wave_sum(hse);
  hte = nb_wave_sum(hte);
  if (lane == 0) {
    A.nnb[s] = count < A.cap ? count : A.cap;
    if (count > A.cap) atomicMax(&A.flags[0], count);
    if (hs) atomicAdd(&A.flags[2], hs);
    if (ht) atomicAdd(&A.flags[3], ht);
    if (hse) atomicAdd(&A.flags[4], hse);
    if (hte) atomicAdd(&A.flags[5], hte);
  }
}
NB_KERNEL_PAIR(k_list, ListArgs, 256)

// nrefsh criterion (md_finter.c:61-73): any |r - r_build|^2 > skin^2/4.
struct CheckArgs { const double *pos, *posb; int n; double skin2q; int *flag; };
static __device__ __forceinline__ void k_check_body(const CheckArgs &A, int bid) {
  int i = bid * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const int n = A.n;
  double x = A.posb[i] - A.pos[i], y = A.posb[n + i] - A.pos[n + i];
  double z = A.posb[2 * n + i] - A.pos[2 * n + i];
  if (x * x + y * y + z * z > A.skin2q) atomicOr(A.flag, 1);
}
NB_KERNEL_PAIR(k_check, CheckArgs, 256)

// ---------------------------------------------------------------------------
// Pair forces, FP64.  One wavefront per sorted atom i walks i's full list,
// 64 neighbours at a time; each lane evaluates cm3d's table kernels for its
// pair with cm3d's arithmetic (so each pair force is bitwise the reference's)
// and accumulates F_i; the wave then reduces F_i with shuffles.  Energies and
// virials count every pair twice and are halved at the end.  The tables
// (interleaved {dV/r, switch, V} per r^2 node) are staged in LDS when they
// fit.
// IFLG 1: fl = full force, fs = switched force (when WANT_FS), energies and
// virials; IFLG 0: fs = short-range force over the s lists, no energies
// (cm3d computes them only at the long level).
// ---------------------------------------------------------------------------
// accumulator slots (NB_NE_PAIR = 19 with the clamp count)
enum {
  E_SPOTL = 0, E_SPOTHL, E_SPRSL, E_SL0, E_SL1, E_SL2, E_SL4, E_SL5, E_SL8,
  E_SPOTE, E_SPOTHE, E_SPRSE, E_SE0, E_SE1, E_SE2, E_SE4, E_SE5, E_SE8, E_CLAMP
};

template <int IFLG, bool WANT_FS>
static __device__ __forceinline__ void pair64_body(const PairArgs &A, int bid) {
  extern __shared__ double tabs[];
  const double *vt = A.vtab3, *ct = A.ctab3;
  if (A.tab_lds) {
    int nv = A.ninter * A.ntable * 3, nc = A.ntable * 3;
    for (int t = threadIdx.x; t < nv; t += blockDim.x) tabs[t] = A.vtab3[t];
    for (int t = threadIdx.x; t < nc; t += blockDim.x) tabs[A.ctab_off + t] = A.ctab3[t];
    __syncthreads();
    vt = tabs;
    ct = tabs + A.ctab_off;
  }
  double acc[NB_NE_PAIR];
#pragma unroll
  for (int k = 0; k < NB_NE_PAIR; k++) acc[k] = 0.;
  const int lane = lane_id();
  const int nw = A.nblocks * 4;
  constexpr unsigned fv = IFLG ? NB_FLAG_VT : NB_FLAG_VS;
  constexpr unsigned fe = IFLG ? NB_FLAG_ET : NB_FLAG_ES;
  const int ntm4 = A.ntable - 4;
  for (int s = bid * 4 + (threadIdx.x >> 6); s < A.n; s += nw) {
    const double4 ri = A.rec[s];
    const int ti = A.stype[s];
    const int nb = A.nnb[s];
    const uint32_t *lst = A.nbr + (size_t)s * A.cap;
    double fx = 0., fy = 0., fz = 0., sx = 0., sy = 0., sz = 0.;
    for (int k = lane; k < nb; k += NB_WAVE) {
      uint32_t e = lst[k];
      unsigned fl = e >> 28;
      if (!(fl & (fv | fe))) continue;
      int j = (int)(e & NB_IDX_MASK);
      double4 rj = A.rec[j];
      double dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
      nb_minimg_box(dx, dy, dz, A.b);
      double d2 = dx * dx + dy * dy + dz * dz;
      if (fl & fv) {
        int in = A.map[ti * A.ntypes + A.stype[j]];
        const InterPar &P = A.ip[in];
        double rmax2 = IFLG ? P.rmaxl2 : P.rmaxs2;
        if (d2 < rmax2) {
          double p = (d2 - P.rmins2) * P.dx2;
          int kk = (int)p;
          if (kk < 1) { acc[E_CLAMP] += 1.; kk = 1; p = 1.; }
          p = p - (double)kk;
          const double *T = vt + (size_t)in * A.ntable * 3;
          double f = nb_interp3(T, kk, 0, p);
          double swv = 1.;
          if ((WANT_FS || !IFLG) && d2 > P.rminl2) swv = nb_interp3(T, kk, 1, p);
          if (IFLG) {
            acc[E_SPRSL] -= d2 * f;
            acc[E_SL0] -= dx * dx * f;
            acc[E_SL1] -= dy * dx * f;
            acc[E_SL2] -= dz * dx * f;
            acc[E_SL4] -= dy * dy * f;
            acc[E_SL5] -= dz * dy * f;
            acc[E_SL8] -= dz * dz * f;
            double gx = dx * f, gy = dy * f, gz = dz * f;
            fx -= gx; fy -= gy; fz -= gz;
            if (WANT_FS) { sx -= gx * swv; sy -= gy * swv; sz -= gz * swv; }
            double v = nb_interp3(T, kk, 2, p);
            acc[E_SPOTL] += v;
            acc[E_SPOTHL] += v - T[3 * ntm4 + 2];
          } else {
            f *= swv;
            sx -= dx * f; sy -= dy * f; sz -= dz * f;
          }
        }
      }
      if ((fl & fe) && d2 < A.rmaxl_coul2) {
        double p = (d2 - A.rmins_coul2) * A.dx2coul;
        int kk = (int)p;
        if (kk < 1) { acc[E_CLAMP] += 1.; kk = 1; p = 1; }
        p = p - (double)kk;
        double qij = ri.w * rj.w;
        double fcc = qij * nb_interp3(ct, kk, 0, p);
        double gx = dx * fcc, gy = dy * fcc, gz = dz * fcc;
        if (IFLG) {
          acc[E_SPRSE] -= d2 * fcc;
          acc[E_SE0] -= dx * dx * fcc;
          acc[E_SE1] -= dy * dx * fcc;
          acc[E_SE2] -= dz * dx * fcc;
          acc[E_SE4] -= dy * dy * fcc;
          acc[E_SE5] -= dz * dy * fcc;
          acc[E_SE8] -= dz * dz * fcc;
          fx -= gx; fy -= gy; fz -= gz;
          double v = nb_interp3(ct, kk, 2, p);
          acc[E_SPOTHE] += qij * (v - A.clast);
          acc[E_SPOTE] += qij * v;
        }
        if (WANT_FS || !IFLG) { sx -= gx; sy -= gy; sz -= gz; }
      }
    }
    const int ai = A.perm[s];
    if (IFLG) { fx = nb_wave_sum(fx); fy = nb_wave_sum(fy); fz = nb_wave_sum(fz); }
    if (WANT_FS || !IFLG) { sx = nb_wave_sum(sx); sy = nb_wave_sum(sy); sz = nb_wave_sum(sz); }
    if (lane == 0) {
      const int n = A.n;
      if (IFLG) { A.fl[ai] = fx; A.fl[n + ai] = fy; A.fl[2 * n + ai] = fz; }
      if (WANT_FS || !IFLG) { A.fs[ai] = sx; A.fs[n + ai] = sy; A.fs[2 * n + ai] = sz; }
    }
  }
  block_reduce_store<NB_NE_PAIR>(acc, A.epart + bid, A.nblocks);
}
#define k_pair64_0t_body pair64_body<0, true>
#define k_pair64_1t_body pair64_body<1, true>
#define k_pair64_1f_body pair64_body<1, false>
NB_KERNEL_PAIR(k_pair64_0t, PairArgs, 256)
NB_KERNEL_PAIR(k_pair64_1t, PairArgs, 256)
NB_KERNEL_PAIR(k_pair64_1f, PairArgs, 256)

// ecorr (md_ewald.c:385-453): one thread per atom over its excluded charged
// partners (both directions, so energies and virials are halved later), on
// the caller's unwrapped positions with no minimum image.  Adds to fl.
struct EcorrArgs {
  const int *eoff, *elist; const double *pos, *q; int n, nblocks; double kappa; double *fl;
  double *epart;
};
static __device__ __forceinline__ void k_ecorr_body(const EcorrArgs &A, int bid) {
  double acc[NB_NE_ECORR];
#pragma unroll
  for (int k = 0; k < NB_NE_ECORR; k++) acc[k] = 0.;
  int a = bid * blockDim.x + threadIdx.x;
  if (a < A.n) {
    const int n = A.n;
    const double spi = 1. / sqrt(M_PI), kappa = A.kappa, ka2 = kappa * kappa;
    const double *pos = A.pos, *q = A.q;
    double fx = 0., fy = 0., fz = 0.;
    for (int e = A.eoff[a]; e < A.eoff[a + 1]; e++) {
      int b = A.elist[e];
      double dx = pos[a] - pos[b], dy = pos[n + a] - pos[n + b], dz = pos[2 * n + a] - pos[2 * n + b];
      double d2 = dx * dx + dy * dy + dz * dz, d = sqrt(d2);
      double qr = q[a] * q[b] / d, er = erf(d * kappa);
      acc[0] += -qr * er;
      double dv = -qr * (-2. * kappa * exp(-ka2 * d2) * spi + er / d);
      acc[1] += d * dv;
      dv /= d;
      acc[2] += dv * dx * dx; acc[3] += dv * dx * dy; acc[4] += dv * dx * dz;
      acc[5] += dv * dy * dx; acc[6] += dv * dy * dy; acc[7] += dv * dy * dz;
      acc[8] += dv * dz * dx; acc[9] += dv * dz * dy; acc[10] += dv * dz * dz;
      fx += dx * dv; fy += dy * dv; fz += dz * dv;
    }
    if (A.eoff[a + 1] > A.eoff[a]) { A.fl[a] += fx; A.fl[n + a] += fy; A.fl[2 * n + a] += fz; }
  }
  block_reduce_store<NB_NE_ECORR, 1>(acc, A.epart + bid, A.nblocks);
}
NB_KERNEL_PAIR(k_ecorr, EcorrArgs, 64)

// ---------------------------------------------------------------------------
// Reciprocal-space Ewald, FP64 throughout.
// k_ewchg: per charge, fractional coordinates s = h^-1 r wrapped to
//   [-1/2,1/2) (exact for e^{2 pi i k.s} with integer k), stored as 2 pi s,
//   the charge times sqrt(4 pi) (md_ewald.c:220-224), and e^{2 pi i s_c}, the
//   c-axis step of cm3d's recursion (md_ewald.c:227-232, 273-278).
// k_ewsfac: one workgroup per (segment, charge chunk); each thread starts a
//   segment with one sincos and walks its <= 16 k vectors by the recursion,
//   accumulating q e^{i k.r} over its charges; a block reduction writes the
//   chunk's partial S(k).
// k_ewkred: per k, sum the chunk partials in fixed order, form preg, prep,
//   |S|^2 and the energy / virial terms (md_ewald.c:247-302).
// k_ewforce: per charge, walk all segments (staged in LDS) and add
//   q k (Re S' sin - Im S' cos) to fl (md_ewald.c:303-323).
// ---------------------------------------------------------------------------
struct EwChgArgs {
  const double *pos, *q; const int *qidx; int nq, n; Box b; double4 *chg; double2 *stp;
};
static __device__ __forceinline__ void k_ewchg_body(const EwChgArgs &A, int bid) {
  int c = bid * blockDim.x + threadIdx.x;
  if (c >= A.nq) return;
  const int n = A.n, a = A.qidx[c];
  double x = A.pos[a], y = A.pos[n + a], z = A.pos[2 * n + a];
  const double *hi = A.b.hi;
  double sa = x * hi[0] + y * hi[1] + z * hi[2];
  double sb = x * hi[3] + y * hi[4] + z * hi[5];
  double sc = x * hi[6] + y * hi[7] + z * hi[8];
  sa -= rint(sa);
  sb -= rint(sb);
  sc -= rint(sc);
  const double tpi = 2. * M_PI;
  A.chg[c] = make_double4(tpi * sa, tpi * sb, tpi * sc, A.q[a] * sqrt(4.0 * M_PI));
  double sn, cs;
  sincos(tpi * sc, &sn, &cs);
  A.stp[c] = make_double2(cs, sn);
}
NB_KERNEL_PAIR(k_ewchg, EwChgArgs, 256)

struct EwSfacArgs {
  const EwSeg *seg; const double4 *chg; const double2 *stp; int nq, chunk, ktot, nseg;
  double2 *part;
};
static __device__ __forceinline__ void k_ewsfac_body(const EwSfacArgs &A, int bid) {
  const int gseg = bid % A.nseg, gchunk = bid / A.nseg;
  const EwSeg sg = A.seg[gseg];
  const int c0 = gchunk * A.chunk;
  const int c1 = min(A.nq, c0 + A.chunk);
  double re[NB_EW_SEG], im[NB_EW_SEG];
#pragma unroll
  for (int m = 0; m < NB_EW_SEG; m++) { re[m] = 0.; im[m] = 0.; }
  for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
    double4 r = A.chg[c];
    double2 st = A.stp[c];
    double th = sg.ka * r.x + sg.kb * r.y + sg.kc0 * r.z;
    double sn, cs;
    sincos(th, &sn, &cs);
#pragma unroll
    for (int m = 0; m < NB_EW_SEG; m++) {
      if (m < sg.cnt) {
        re[m] += r.w * cs;
        im[m] += r.w * sn;
        double c2 = cs * st.x - sn * st.y;
        sn = sn * st.x + cs * st.y;
        cs = c2;
      }
    }
  }
  __shared__ double red[4][2 * NB_EW_SEG];
  const int w = threadIdx.x >> 6, lane = lane_id();
#pragma unroll
  for (int m = 0; m < NB_EW_SEG; m++) {
    double a = nb_wave_sum(re[m]), bb = nb_wave_sum(im[m]);
    if (lane == 0) { red[w][2 * m] = a; red[w][2 * m + 1] = bb; }
  }
  __syncthreads();
  if (threadIdx.x < sg.cnt) {
    int m = threadIdx.x;
    double a = red[0][2 * m] + red[1][2 * m], bb = red[0][2 * m + 1] + red[1][2 * m + 1];
    a += red[2][2 * m] + red[3][2 * m];
    bb += red[2][2 * m + 1] + red[3][2 * m + 1];
    A.part[(size_t)gchunk * A.ktot + sg.k0 + m] = make_double2(a, bb);
  }
}
NB_KERNEL_PAIR(k_ewsfac, EwSfacArgs, 256)

struct KData { double kx, ky, kz, cr, ci; };

struct EwKredArgs {
  const int *kabc; int ktot; const double2 *part; int nchunk, nblocks; Box b; double rvol, falp2;
  KData *kd; double *epart;
};
static __device__ __forceinline__ void k_ewkred_body(const EwKredArgs &A, int bid) {
  double acc[NB_NE_EW];
#pragma unroll
  for (int k = 0; k < NB_NE_EW; k++) acc[k] = 0.;
  int i = bid * blockDim.x + threadIdx.x;
  if (i < A.ktot) {
    double sr = 0., si = 0.;
    for (int c = 0; c < A.nchunk; c++) {
      double2 p = A.part[(size_t)c * A.ktot + i];
      sr += p.x;
      si += p.y;
    }
    const double tpi = 2. * M_PI, *hi = A.b.hi;
    double six = tpi * A.kabc[3 * i], siy = tpi * A.kabc[3 * i + 1], siz = tpi * A.kabc[3 * i + 2];
    double aka = six * hi[0] + siy * hi[3] + siz * hi[6];
    double akb = six * hi[1] + siy * hi[4] + siz * hi[7];
    double akc = six * hi[2] + siy * hi[5] + siz * hi[8];
    double g2 = aka * aka + akb * akb + akc * akc;
    double preg = exp(-g2 * A.falp2) * A.rvol / g2;
    double prep = -2. * preg * (g2 * A.falp2 + 1.) / g2;
    double sm = sr * sr + si * si;
    acc[0] = preg * sm;
    acc[1] = prep * sm * g2;
    acc[2] = prep * sm * aka * aka; acc[3] = prep * sm * akb * aka; acc[4] = prep * sm * akc * aka;
    acc[5] = prep * sm * aka * akb; acc[6] = prep * sm * akb * akb; acc[7] = prep * sm * akc * akb;
    acc[8] = prep * sm * aka * akc; acc[9] = prep * sm * akb * akc; acc[10] = prep * sm * akc * akc;
    KData d;
    d.kx = aka; d.ky = akb; d.kz = akc;
    d.cr = sr * 2. * preg;
    d.ci = si * 2. * preg;
    A.kd[i] = d;
  }
  block_reduce_store<NB_NE_EW>(acc, A.epart + bid, A.nblocks);
}
NB_KERNEL_PAIR(k_ewkred, EwKredArgs, 256)

// One wavefront per charge; its lanes take the k segments round-robin, each
// starting its segment with one sincos and walking it by the recursion, and
// the wave sums the three force components at the end.  The k data (40 B
// per k) are read by every wave and stay cache-resident.
struct EwForceArgs {
  const EwSeg *seg; int nseg; const KData *kd; const double4 *chg; const double2 *stp;
  const int *qidx; int nq, n; double *fl;
};
static __device__ __forceinline__ void k_ewforce_body(const EwForceArgs &A, int bid) {
  const int c = bid * 4 + (threadIdx.x >> 6), lane = lane_id();
  if (c >= A.nq) return;
  const double4 r = A.chg[c];
  const double2 st = A.stp[c];
  double fx = 0., fy = 0., fz = 0.;
  for (int g = lane; g < A.nseg; g += NB_WAVE) {
    const EwSeg sg = A.seg[g];
    double th = sg.ka * r.x + sg.kb * r.y + sg.kc0 * r.z;
    double sn, cs;
    sincos(th, &sn, &cs);
    const KData *K = A.kd + sg.k0;
    for (int m = 0; m < sg.cnt; m++) {
      const KData kk = K[m];
      double tt = kk.cr * sn - kk.ci * cs;
      fx += kk.kx * tt;
      fy += kk.ky * tt;
      fz += kk.kz * tt;
      double c2 = cs * st.x - sn * st.y;
      sn = sn * st.x + cs * st.y;
      cs = c2;
    }
  }
  fx = nb_wave_sum(fx);
  fy = nb_wave_sum(fy);
  fz = nb_wave_sum(fz);
  if (lane == 0) {
    const int a = A.qidx[c], n = A.n;
    A.fl[a] += r.w * fx;
    A.fl[n + a] += r.w * fy;
    A.fl[2 * n + a] += r.w * fz;
  }
}
NB_KERNEL_PAIR(k_ewforce, EwForceArgs, 256)

// Column sums of a column-major [COLS][rows] partial table: one block per
// column, a fixed assignment of rows to threads and a fixed tree
// (deterministic).
static __device__ __forceinline__ void reduce_body(const RedArgs &A, int bid) {
  const double *col = A.part + (size_t)bid * A.rows;
  double acc = 0.;
  for (int r = threadIdx.x; r < A.rows; r += blockDim.x) acc += col[r];
  block_reduce_store<1>(&acc, A.out + bid, 1);
}
#define k_red_pair_body reduce_body
#define k_red_ecorr_body reduce_body
#define k_red_ew_body reduce_body
NB_KERNEL_PAIR(k_red_pair, RedArgs, 256)
NB_KERNEL_PAIR(k_red_ecorr, RedArgs, 256)
NB_KERNEL_PAIR(k_red_ew, RedArgs, 256)

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define LCHK(x)                                                        \
  do {                                                                 \
    hipError_t _e = (x);                                               \
    if (_e != hipSuccess) {                                            \
      nb_set_error("%s: %s", #x, hipGetErrorString(_e));               \
      return NB_ERR_HIP;                                               \
    }                                                                  \
  } while (0)

static int blocks(int n, int t) { return (n + t - 1) / t; }

// The list build.  cond == nullptr: unconditional; otherwise every kernel
// checks *cond (the device rebuild flag) and does nothing when it is clear.
// No memset nodes: zeroing is done by kernels so the sequence can be captured.
static int launch_build_impl(nb_ctx *c, const Box &b, hipStream_t s, const int *cond) {
  const int n = c->natoms;
  ZeroArgs z1 = {c->d_ccount, c->ncells, cond};
  NB_GO(c, K_ZERO, k_zero, blocks(c->ncells, 256), 256, 0, z1, s);
  // overflow max (flags[0]) and half-pair counts (flags[2..5]); flags[1] is
  // the rebuild flag `cond` itself and must survive the whole build
  ZeroArgs z2 = {c->d_flags + 2, 4, cond};
  NB_GO(c, K_ZERO, k_zero, 1, 256, 0, z2, s);
  ZeroArgs z3 = {c->d_flags, 1, cond};
  NB_GO(c, K_ZERO, k_zero, 1, 256, 0, z3, s);
  CellArgs ca = {c->d_pos, n, b, c->ncx, c->ncy, c->ncz, c->d_cell, c->d_ccount, cond};
  NB_GO(c, K_CELL, k_cell, blocks(n, 256), 256, 0, ca, s);
  ScanArgs sa = {c->d_ccount, c->ncells, c->d_cstart, c->d_cfill, cond};
  NB_GO(c, K_SCAN, k_scan, 1, 1024, 0, sa, s);
  ScatArgs sc = {c->d_cell, n, c->d_cfill, c->d_tmp, cond};
  NB_GO(c, K_SCAT, k_scat, blocks(n, 256), 256, 0, sc, s);
  // the sorted cell id of each sorted slot overwrites d_cell
  OrderArgs oa = {c->d_ccount, c->d_cstart, c->d_tmp, c->d_perm, c->d_iperm, c->d_cell, cond};
  NB_GO(c, K_ORDER, k_order, c->ncells, 256, 0, oa, s);
  Pack64Args pa = {c->d_pos, c->d_qch, c->d_itype, c->d_perm, n, (double4 *)c->d_rec64,
                   c->d_stype, cond};
  NB_GO(c, K_PACK64, k_pack64, blocks(n, 256), 256, 0, pa, s);
  ListArgs A;
  A.n = n; A.ntypes = c->ntypes; A.cap = c->cap;
  A.ncx = c->ncx; A.ncy = c->ncy; A.ncz = c->ncz;
  A.rec = (const double4 *)c->d_rec64; A.stype = c->d_stype; A.perm = c->d_perm;
  A.scell = c->d_cell; A.cstart = c->d_cstart; A.ccount = c->d_ccount;
  A.xoff = c->d_xoff; A.xlist = c->d_xlist; A.map = c->d_map; A.ip = c->d_ip;
  A.rmaxs_skin2_e = c->rmaxs_skin2_e; A.rmaxl_skin2_e = c->rmaxl_skin2_e;
  A.b = b; A.nbr = c->d_nbr; A.nnb = c->d_nnb; A.flags = c->d_flags; A.cond = cond;
  NB_GO(c, K_LIST, k_list, blocks(n, 4), 256, 0, A, s);
  CopyArgs cp = {c->d_posb, c->d_pos, 3 * n, cond, c->d_flags + 6};
  NB_GO(c, K_COPY, k_copy, blocks(3 * n, 256), 256, 0, cp, s);
  if (!c->rec) LCHK(hipGetLastError());
  return NB_OK;
}

int nb_launch_build(nb_ctx *c, const Box &b, hipStream_t s) {
  return launch_build_impl(c, b, s, nullptr);
}

int nb_launch_build_cond(nb_ctx *c, const Box &b, hipStream_t s) {
  return launch_build_impl(c, b, s, c->d_flags + 1);
}

int nb_launch_rebuild_check(nb_ctx *c, const double *d_pos, hipStream_t s) {
  ZeroArgs z = {c->d_flags + 1, 1, nullptr};
  NB_GO(c, K_ZERO, k_zero, 1, 256, 0, z, s);
  CheckArgs ch = {d_pos, c->d_posb, c->natoms, .25 * c->skin * c->skin, c->d_flags + 1};
  NB_GO(c, K_CHECK, k_check, blocks(c->natoms, 256), 256, 0, ch, s);
  if (!c->rec) LCHK(hipGetLastError());
  return NB_OK;
}

int nb_launch_force(nb_ctx *c, int iflg, const Box &b, double *d_fs, double *d_fl,
                    hipStream_t s) {
  const int n = c->natoms;
  nb_time_mark(c, 0, s);
  if (c->prec == NB_PREC_MIXED) {
    nb_time_mark(c, 1, s);
    int rc = nb_launch_pair_mixed(c, iflg, b, d_fs, d_fl, s);
    if (rc) return rc;
  } else {
    Pack64Args pa = {c->d_pos, c->d_qch, c->d_itype, c->d_perm, n, (double4 *)c->d_rec64,
                     nullptr, nullptr};
    NB_GO(c, K_PACK64, k_pack64, blocks(n, 256), 256, 0, pa, s);
    nb_time_mark(c, 1, s);
    PairArgs P;
    P.n = n; P.ntypes = c->ntypes; P.ninter = c->ninter; P.ntable = c->ntable;
    P.cap = c->cap; P.tab_lds = c->tab_lds_bytes > 0; P.ctab_off = c->ctab_off;
    P.nblocks = c->npair_blocks;
    P.nbr = c->d_nbr; P.nnb = c->d_nnb; P.stype = c->d_stype; P.perm = c->d_perm;
    P.map = c->d_map; P.rec = (const double4 *)c->d_rec64; P.ip = c->d_ip;
    P.vtab3 = c->d_vtab3; P.ctab3 = c->d_ctab3;
    P.rmins_coul2 = c->rmins_coul2; P.dx2coul = c->dx2coul; P.rmaxl_coul2 = c->rmaxl_coul2;
    P.clast = c->clast;
    P.b = b; P.fl = d_fl; P.fs = d_fs; P.epart = c->d_epair;
    const int g = c->npair_blocks, lds = c->tab_lds_bytes;
    if (iflg == 0)
      NB_GO(c, K_PAIR64_0T, k_pair64_0t, g, 256, lds, P, s);
    else if (d_fs)
      NB_GO(c, K_PAIR64_1T, k_pair64_1t, g, 256, lds, P, s);
    else
      NB_GO(c, K_PAIR64_1F, k_pair64_1f, g, 256, lds, P, s);
    if (!c->rec) LCHK(hipGetLastError());
  }
  nb_time_mark(c, 2, s);
  RedArgs rp = {c->d_epair, c->npair_blocks, c->d_eout};
  NB_GO(c, K_RED_PAIR, k_red_pair, NB_NE_PAIR, 256, 0, rp, s);
  const bool ec = iflg == 1 && c->iperd == 3 && c->necorr > 0;
  EcorrArgs ea = {c->d_eoff, c->d_elist, c->d_pos, c->d_qch, n, c->necorr_blocks, c->kappa,
                  d_fl, c->d_eec};
  NB_GO(c, K_ECORR, k_ecorr, ec ? c->necorr_blocks : 0, 64, 0, ea, s);
  RedArgs re = {c->d_eec, c->necorr_blocks, c->d_eout + 20};
  NB_GO(c, K_RED_ECORR, k_red_ecorr, ec ? NB_NE_ECORR : 0, 256, 0, re, s);
  nb_time_mark(c, 3, s);
  const bool ew = iflg == 1 && c->iperd == 3 && c->ktot > 0 && c->ncharge > 0;
  const int nq = c->ncharge;
  EwChgArgs ch = {c->d_pos, c->d_qch, c->d_qidx, nq, n, b, (double4 *)c->d_chg,
                  (double2 *)c->d_stp};
  NB_GO(c, K_EWCHG, k_ewchg, ew ? blocks(nq, 256) : 0, 256, 0, ch, s);
  EwSfacArgs sf = {c->d_seg, (const double4 *)c->d_chg, (const double2 *)c->d_stp, nq,
                   (nq + c->nchunk - 1) / c->nchunk, c->ktot, c->nseg, (double2 *)c->d_spart};
  NB_GO(c, K_EWSFAC, k_ewsfac, ew ? c->nseg * c->nchunk : 0, 256, 0, sf, s);
  double det = b.h[0] * (b.h[4] * b.h[8] - b.h[5] * b.h[7]) +
               b.h[1] * (b.h[5] * b.h[6] - b.h[3] * b.h[8]) +
               b.h[2] * (b.h[3] * b.h[7] - b.h[4] * b.h[6]);
  double falp2 = c->kappa > 0 ? 1. / (4. * c->kappa * c->kappa) : 0.;
  EwKredArgs kr = {c->d_kabc, c->ktot, (const double2 *)c->d_spart, c->nchunk, c->new_blocks,
                   b, 1. / det, falp2, (KData *)c->d_kd, c->d_eew};
  NB_GO(c, K_EWKRED, k_ewkred, ew ? c->new_blocks : 0, 256, 0, kr, s);
  RedArgs rw = {c->d_eew, c->new_blocks, c->d_eout + 32};
  NB_GO(c, K_RED_EW, k_red_ew, ew ? NB_NE_EW : 0, 256, 0, rw, s);
  EwForceArgs ef = {c->d_seg, c->nseg, (const KData *)c->d_kd, (const double4 *)c->d_chg,
                    (const double2 *)c->d_stp, c->d_qidx, nq, n, d_fl};
  NB_GO(c, K_EWFORCE, k_ewforce, ew ? blocks(nq, 4) : 0, 256, 0, ef, s);
  if (!c->rec) LCHK(hipGetLastError());
  nb_time_mark(c, 4, s);
  nb_time_mark(c, 5, s);
  if (c->timing && !c->rec) c->trec++;
  return NB_OK;
}

size_t nb_kdata_bytes() { return sizeof(KData); }

static struct RegisterKernels {
  RegisterKernels() {
    g_batch_fn[K_ZERO] = k_zero_bl;
    g_batch_fn[K_COPY] = k_copy_bl;
    g_batch_fn[K_CELL] = k_cell_bl;
    g_batch_fn[K_SCAN] = k_scan_bl;
    g_batch_fn[K_SCAT] = k_scat_bl;
    g_batch_fn[K_ORDER] = k_order_bl;
    g_batch_fn[K_PACK64] = k_pack64_bl;
    g_batch_fn[K_LIST] = k_list_bl;
    g_batch_fn[K_CHECK] = k_check_bl;
    g_batch_fn[K_PAIR64_0T] = k_pair64_0t_bl;
    g_batch_fn[K_PAIR64_1T] = k_pair64_1t_bl;
    g_batch_fn[K_PAIR64_1F] = k_pair64_1f_bl;
    g_batch_fn[K_RED_PAIR] = k_red_pair_bl;
    g_batch_fn[K_RED_ECORR] = k_red_ecorr_bl;
    g_batch_fn[K_RED_EW] = k_red_ew_bl;
    g_batch_fn[K_ECORR] = k_ecorr_bl;
    g_batch_fn[K_EWCHG] = k_ewchg_bl;
    g_batch_fn[K_EWSFAC] = k_ewsfac_bl;
    g_batch_fn[K_EWKRED] = k_ewkred_bl;
    g_batch_fn[K_EWFORCE] = k_ewforce_bl;
  }
} g_register_kernels;
