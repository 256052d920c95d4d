// nb_kernels.hip -- gfx950 kernels of the nonbonded path (FP64 form) and
// their launchers.
//
//   list build   k_cell -> k_scat -> k_order -> k_list
//                (cm3d get_ngbr / get_ngbr_e, energy/control/ngbr_list.c:297,
//                 elec_ngbr.c:220, set_link.c:203)
//   pair forces  k_pair64_* (force_npol md_fnopol.c:46, force_e finter_e.c:42);
//                the mixed form is nb_mixed.hip
//   reciprocal   k_ewchg -> k_ewsfac -> k_ewkred -> k_ewforce -> k_ewfin
//                (fk_ewald md_ewald.c:185; k_ewfin also does ecorr,
//                md_ewald.c:385)
//   reductions   k_red (the getsvirter / getsvirter_e / getvewald sums)
//
// Every kernel is a __device__ body plus a single-system and a batched entry
// point (nb_launch.h); none allocates.
#include <stdarg.h>
#include <string.h>
#include <stdio.h>

#include "nb_device.h"

BatchFn g_batch_fn[K_COUNT];

static __device__ __forceinline__ int mbcnt(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// record of sorted atom j (j < NB_MAX_ATOMS: a 32-bit byte offset)
static __device__ __forceinline__ double4 rec64_at(const double4 *rec, uint32_t j) {
  return *(const double4 *)((const char *)rec + (j << 5));
}
static __device__ __forceinline__ int type_at(const int *t, uint32_t j) {
  return *(const int *)((const char *)t + (j << 2));
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


// ---------------------------------------------------------------------------
// Cell sort.  Cells tile the box in fractional coordinates; along each box
// vector they are either no narrower than a third of the longest list
// cutoff (reach 3: the 7 cells around a cell's own), half of it (reach 2),
// the cutoff (reach 1) or one cell for the whole box (reach 0), so every pair within
// the cutoff under cm3d's minimum image lies in the (2rx+1)(2ry+1)(2rz+1)
// cells around.  Binning follows get_lnklist (set_link.c:218-236): wrap
// s -= anint(s), then (int)((s+.5)*n) clamped.
//   k_cell   cell of each atom, counts by atomics; the last workgroup to
//            finish scans the counts into cell starts (and clears the
//            atomics' array and its own ticket for the next build)
//   k_scat   scatter atom indices into their cell's slots (arbitrary order)
//   k_order  per cell: order its atoms by index (deterministic sorted
//            order), then write the sorted records {x,y,z,q} (and the mixed
//            kernel's), types, the permutation both ways, and the build
//            positions (r_build)
//   k_list   the Verlet lists
// ---------------------------------------------------------------------------
struct CellArgs {
  const double *pos; int n, nblocks, ncells; Box b; int ncx, ncy, ncz;
  int *cell, *work, *count, *start, *fill, *flags; const int *cond;
};
static __device__ __forceinline__ void k_cell_body(const CellArgs &A, int bid) {
  NB_SKIP(A.cond);
  __shared__ int wsum[4];
  __shared__ int last, carry;
  const int i = bid * blockDim.x + threadIdx.x;
  if (bid == 0 && threadIdx.x == 0) A.flags[0] = 0;  // overflow max of this build
  if (i < A.n) {
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
    const int id = (a * A.ncy + c) * A.ncz + e;
    A.cell[i] = id;
    atomicAdd(&A.work[id], 1);
  }
  // last workgroup: exclusive scan of the counts (flags[7] is the ticket)
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(&A.flags[7], 1) == A.nblocks - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x >> 6;
  for (int base = 0; base < A.ncells; base += blockDim.x) {
    const int k = base + threadIdx.x;
    const int v = k < A.ncells ? __hip_atomic_load(&A.work[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    int x = v;
#pragma unroll
    for (int o = 1; o < NB_WAVE; o <<= 1) {
      int y = __shfl_up(x, o, NB_WAVE);
      if (lane >= o) x += y;
    }
    if (lane == NB_WAVE - 1) wsum[w] = x;
    __syncthreads();
    int pre = carry;
    for (int q = 0; q < w; q++) pre += wsum[q];
    const int excl = pre + x - v;
    if (k < A.ncells) { A.start[k] = excl; A.fill[k] = excl; A.count[k] = v; A.work[k] = 0; }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) A.flags[7] = 0;
}
NB_KERNEL_PAIR(k_cell, CellArgs, 256)

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
  const int *count, *start, *tmp; int *perm, *iperm, *scell;
  const double *pos, *q; const int *itype; double4 *rec; int *stype; double *posb; int n;
  int *builds; uint4 *recf; Box b; const int *cond;
  int ncells, packed;  // packed: 16 cells a workgroup, 16 lanes a cell
};
// atom a of cell c goes to sorted slot st + (its rank by index in the cell)
static __device__ __forceinline__ void order_atom(const OrderArgs &A, int c, int st, int cnt,
                                                  const int *src, int a) {
  const int n = A.n;
  int r = 0;
  for (int u = 0; u < cnt; u++) r += src[u] < a;
  const int sl = st + r;
  A.perm[sl] = a;
  A.iperm[a] = sl;
  A.scell[sl] = c;
  const double x = A.pos[a], y = A.pos[n + a], z = A.pos[2 * n + a];
  A.rec[sl] = make_double4(x, y, z, A.q[a]);
  if (A.recf) A.recf[sl] = nb_mixed_record(x, y, z, A.q[a], A.b.hi);
  A.stype[sl] = A.itype[a];
  A.posb[a] = x; A.posb[n + a] = y; A.posb[2 * n + a] = z;
}
static __device__ __forceinline__ void k_order_body(const OrderArgs &A, int bid) {
  NB_SKIP(A.cond);
  __shared__ int seg[NB_CELL_LDS];
  if (bid == 0 && threadIdx.x == 0) atomicAdd(A.builds, 1);
  if (A.packed) {
    // many small cells: a workgroup per 16 cells keeps the launch (and the
    // skipped launch of a step without a rebuild) small
    const int c = bid * 16 + (threadIdx.x >> 4), l = threadIdx.x & 15;
    if (c >= A.ncells) return;
    const int cnt = A.count[c], st = A.start[c];
    const int *src = A.tmp + st;
    for (int t = l; t < cnt; t += 16) order_atom(A, c, st, cnt, src, src[t]);
    return;
  }
  const int c = bid, cnt = A.count[c], st = A.start[c];
  const int *src = A.tmp + st;
  if (cnt <= NB_CELL_LDS) {
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) seg[t] = src[t];
    __syncthreads();
    src = seg;
  }
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) order_atom(A, c, st, cnt, src, src[t]);
}
NB_KERNEL_PAIR(k_order, OrderArgs, 256)

// Sorted-order records refreshed at every evaluation: {x,y,z,q} in FP64.
struct Pack64Args {
  const double *pos, *q; const int *perm; int n; double4 *rec;
};
static __device__ __forceinline__ void k_pack64_body(const Pack64Args &A, int bid) {
  int s = bid * blockDim.x + threadIdx.x;
  if (s >= A.n) return;
  const int n = A.n, a = A.perm[s];
  A.rec[s] = make_double4(A.pos[a], A.pos[n + a], A.pos[2 * n + a], A.q[a]);
}
NB_KERNEL_PAIR(k_pack64, Pack64Args, 256)

// ---------------------------------------------------------------------------
// Verlet list build, one wavefront per (sorted) atom.  The candidate cells
// around the atom's own are runs of consecutive cells along c (at most 2
// runs per (a,b) column, consecutive in sorted order); the runs of all
// columns, concatenated, are walked 64 candidates at a time; each lane
// applies cm3d's filters in FP64 exactly as ngbr_list /
// ngbr_liste do (null map, exclusion, minimum image, strict
// r^2 < (rc+skin)^2, s nested in t), and the survivors are compacted with a
// ballot + mbcnt prefix into the atom's slots.  Both directions of every pair
// are stored; the per-atom half-pair counts (each pair once) go to hcnt.
// ---------------------------------------------------------------------------
#define NB_PCUT_LDS 1024  // type pairs whose list cutoffs are staged in LDS
struct ListArgs {
  int n, ntypes, cap, ncx, ncy, ncz, rx, ry, rz, pcut_lds, prune;
  double rprune2, wcx, wcy, wcz;  // pruning bound^2, cell widths (A) of an orthorhombic box
  const double4 *rec;
  const int *stype, *perm, *iperm, *scell, *cstart, *ccount;
  const int *xoff, *xlist;
  const double2 *pcut;  // {(rmaxl+skin)^2, (rmaxs+skin)^2} per (ti,tj); -1 for null
  double rmaxs_skin2_e, rmaxl_skin2_e;
  Box b;
  uint32_t *nbr;
  int *nnb, *flags;
  int4 *hcnt;
  const int *cond;
  int ibits, etype;  // entry layout: the partner's type above ibits index bits (when etype)
};

static __device__ __forceinline__ int wrapc(int v, int n) {
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

static __device__ __forceinline__ void k_list_body(const ListArgs &A, int bid) {
  NB_SKIP(A.cond);
  __shared__ double2 pc[NB_PCUT_LDS];
  __shared__ int colrun[4][4 * NB_WAVE];  // per wave: column ends, rs0, rl0, rs1
  if (A.pcut_lds) {
    for (int t = threadIdx.x; t < A.ntypes * A.ntypes; t += blockDim.x) pc[t] = A.pcut[t];
    __syncthreads();
  }
  const int lane = lane_id();
  const int s = bid * 4 + (threadIdx.x >> 6);
  if (s >= A.n) return;
  const double4 ri = A.rec[s];
  const int ti = A.stype[s], ai = A.perm[s];
  const int c = A.scell[s];
  const int cz = c % A.ncz, cy = (c / A.ncz) % A.ncy, cx = c / (A.ncy * A.ncz);
  // this atom's exclusions (kind bits 30-31 over an atom index), the first
  // four as wave-uniform values with the partner's sorted index, so that a
  // candidate is tested by its sorted index and no atom index is gathered
  const int x0 = A.xoff[ai], x1 = A.xoff[ai + 1], nx = x1 - x0;
  const unsigned none = 0xFFFFFFFFu;  // index bits 0x3FFFFFFF match no atom
  auto xsorted = [&](int e) {
    const unsigned v = (unsigned)A.xlist[e];
    return (v & 0xC0000000u) | (unsigned)A.iperm[v & 0x3FFFFFFFu];
  };
  const unsigned ex0 = nx > 0 ? xsorted(x0) : none;
  const unsigned ex1 = nx > 1 ? xsorted(x0 + 1) : none;
  const unsigned ex2 = nx > 2 ? xsorted(x0 + 2) : none;
  const unsigned ex3 = nx > 3 ? xsorted(x0 + 3) : none;
  const bool qi = ri.w != 0.0;
  // candidate runs: lane l < ncol takes column l (offsets ox, oy) and
  // finds its <= 2 slot runs
  const int wy = 2 * A.ry + 1, ncol = (2 * A.rx + 1) * wy;
  int rs0 = 0, rl0 = 0, rs1 = 0, rl1 = 0;
  if (lane < ncol) {
    const int ox = lane / wy - A.rx, oy = lane % wy - A.ry;
    const int xx = wrapc(cx + ox, A.ncx), yy = wrapc(cy + oy, A.ncy);
    const int base = (xx * A.ncy + yy) * A.ncz;
    int zlo = cz - A.rz, zhi = cz + A.rz;
    if (A.prune) {
      // orthorhombic box: the column's nearest point lies gx, gy away in x,
      // y; columns beyond every list cutoff are dropped and the rest keep
      // only the cells within reach in z.  The bound is the list cutoff
      // widened by 1e-6, so no pair the exact test below would keep is lost.
      const double *hi = A.b.hi;
      double sx = ri.x * hi[0], sy = ri.y * hi[4], sz = ri.z * hi[8];
      sx -= nb_anint(sx);
      sy -= nb_anint(sy);
      sz -= nb_anint(sz);
      const double ux = (sx + .5) * A.ncx, uy = (sy + .5) * A.ncy, uz = (sz + .5) * A.ncz;
      double gx = ox > 0 ? (double)(cx + ox) - ux : (ox < 0 ? ux - (double)(cx + ox + 1) : 0.);
      double gy = oy > 0 ? (double)(cy + oy) - uy : (oy < 0 ? uy - (double)(cy + oy + 1) : 0.);
      gx = gx > 0. ? gx * A.wcx : 0.;
      gy = gy > 0. ? gy * A.wcy : 0.;
      const double rest = A.rprune2 - (gx * gx + gy * gy);
      if (rest < 0.) {
        zhi = zlo - 1;
      } else {
        const double dz = sqrt(rest) / A.wcz;
        const int z0 = (int)floor(uz - dz), z1 = (int)floor(uz + dz);
        zlo = z0 > zlo ? z0 : zlo;
        zhi = z1 < zhi ? z1 : zhi;
      }
    }
    if (zlo <= zhi) {
      const int a0 = zlo < 0 ? 0 : zlo, a1 = zhi > A.ncz - 1 ? A.ncz - 1 : zhi;
      if (a0 <= a1) {
        rs0 = A.cstart[base + a0];
        rl0 = A.cstart[base + a1] + A.ccount[base + a1] - rs0;
      }
      int b0 = -1, b1 = -1;
      if (zlo < 0) { b0 = A.ncz + zlo; b1 = A.ncz - 1; }
      else if (zhi > A.ncz - 1) { b0 = 0; b1 = zhi - A.ncz; }
      if (b0 >= 0) {
        rs1 = A.cstart[base + b0];
        rl1 = A.cstart[base + b1] + A.ccount[base + b1] - rs1;
      }
    }
  }
  const double2 *pcut = A.pcut + ti * A.ntypes;
  // does this atom's type have any vdW partner type (null pairs carry -1)
  bool rowv = false;
  for (int tj = 0; tj < A.ntypes; tj++)
    rowv |= (A.pcut_lds ? pc[ti * A.ntypes + tj].x : pcut[tj].x) > 0.;
  int nv = 0, nc = 0;
  uint32_t *out = A.nbr + (size_t)s * A.cap;
  // The candidates are the concatenation, column by column, of each
  // column's runs; 64 consecutive candidates make a chunk whatever runs
  // they come from, so short runs (small cells) leave no lanes idle.  The
  // order is the concatenation order either way.  Column l's runs and the
  // end of its candidates in the concatenation go to this wave's LDS rows;
  // each lane keeps a cursor on the column its candidate falls in.
  int cend = rl0 + rl1;  // inclusive prefix over the columns (lanes)
#pragma unroll
  for (int o = 1; o < NB_WAVE; o <<= 1) {
    const int y = __shfl_up(cend, o, NB_WAVE);
    if (lane >= o) cend += y;
  }
  const int total = __builtin_amdgcn_readlane(cend, NB_WAVE - 1);
  int *lc = colrun[threadIdx.x >> 6];
  lc[lane] = cend;
  lc[NB_WAVE + lane] = rs0;
  lc[2 * NB_WAVE + lane] = rl0;
  lc[3 * NB_WAVE + lane] = rs1;
  // the wave's LDS rows are read by other lanes of the same wave
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  __builtin_amdgcn_wave_barrier();
  // a lane's cursor: col, the column of its next candidate, and colend,
  // that column's end; it only moves forward.  A lane past the total keeps
  // its cursor on the last candidate (colrun[63] is the total, so the
  // cursor stops by lane 63's entry) and gathers that candidate's record,
  // unused: every lane gathers, no exec-mask change guards the loads.
  int col = 0, colend = lc[0];
  auto cand = [&](int c) {
    const int cc = min(c, total - 1);
    while (colend <= cc) colend = lc[++col];
    const int cstart = col > 0 ? lc[col - 1] : 0;
    const int loc = cc - cstart, r0 = lc[2 * NB_WAVE + col];
    return loc < r0 ? lc[NB_WAVE + col] + loc : lc[3 * NB_WAVE + col] + (loc - r0);
  };
  int s2c = cand(lane);
  bool okc = lane < total && s2c != s;
  double4 rjc = rec64_at(A.rec, (uint32_t)s2c);
  int tjc = type_at(A.stype, (uint32_t)s2c);
  // half pairs (each once, by sorted index) counted from ballots: wave-wide
  // totals, no per-lane sums
  int hs = 0, ht = 0, hse = 0, hte = 0;
#pragma unroll 2
  for (int cb = 0; cb < total; cb += NB_WAVE) {
    const int cn = cb + NB_WAVE + lane;
    const int s2n = cand(cn);
    const bool okn = cn < total && s2n != s;
    const double4 rjn = rec64_at(A.rec, (uint32_t)s2n);
    const int tjn = type_at(A.stype, (uint32_t)s2n);
    unsigned fl = 0;
    if (okc) {
      double dx = ri.x - rjc.x, dy = ri.y - rjc.y, dz = ri.z - rjc.z;
      nb_minimg_box(dx, dy, dz, A.b);
      const double d2 = dx * dx + dy * dy + dz * dz;
      const unsigned uj = (unsigned)s2c;
      unsigned ex = 0;
      // only as many tests as the atom has exclusions (wave-uniform)
      if (nx > 0) {
        if ((ex0 & 0x3FFFFFFFu) == uj) ex |= ex0 >> 30;
        if (nx > 1) {
          if ((ex1 & 0x3FFFFFFFu) == uj) ex |= ex1 >> 30;
          if (nx > 2) {
            if ((ex2 & 0x3FFFFFFFu) == uj) ex |= ex2 >> 30;
            if ((ex3 & 0x3FFFFFFFu) == uj) ex |= ex3 >> 30;
            for (int e = x0 + 4; e < x1; e++) {
              const unsigned v = xsorted(e);
              if ((v & 0x3FFFFFFFu) == uj) ex |= v >> 30;
            }
          }
        }
      }
      const double2 cut = A.pcut_lds ? pc[ti * A.ntypes + tjc] : pcut[tjc];
      if (!(ex & 1u) && d2 < cut.x) {  // null pairs carry -1: never true
        fl |= NB_FLAG_VT;
        if (d2 < cut.y) fl |= NB_FLAG_VS;
      }
      if (qi && rjc.w != 0.0 && !(ex & 2u)) {
        if (d2 < A.rmaxl_skin2_e) {
          fl |= NB_FLAG_ET;
          if (d2 < A.rmaxs_skin2_e) fl |= NB_FLAG_ES;
        }
      }
    }
    // entries with a vdW flag fill the row from the front, Coulomb-only
    // entries from the back, each in candidate order (all from the front in
    // a row whose atom type has no vdW partner type)
    const bool isv = (fl & NB_FLAG_VT) != 0 || (fl != 0 && !rowv);
    const unsigned long long mv = __ballot(isv), mc = __ballot(fl != 0 && !isv);
    if (fl) {
      const uint32_t ent = (uint32_t)s2c | (A.etype ? (uint32_t)tjc << A.ibits : 0u) | (fl << 28);
      if (isv) {
        const int slot = nv + mbcnt(mv);
        if (slot < A.cap) out[slot] = ent;
      } else {
        const int slot = A.cap - 1 - (nc + mbcnt(mc));
        if (slot >= 0) out[slot] = ent;
      }
    }
    // each unordered pair once (the filters are symmetric in i and j, so
    // counting by sorted index gives cm3d's i < j totals)
    const unsigned hf = s < s2c ? fl : 0u;
    hs += __popcll(__ballot((hf & NB_FLAG_VS) != 0));
    ht += __popcll(__ballot((hf & NB_FLAG_VT) != 0));
    hse += __popcll(__ballot((hf & NB_FLAG_ES) != 0));
    hte += __popcll(__ballot((hf & NB_FLAG_ET) != 0));
    nv += __popcll(mv);
    nc += __popcll(mc);
    rjc = rjn; tjc = tjn; s2c = s2n; okc = okn;
  }
  const int count = nv + nc;
  // close the row: the Coulomb-only entries move down behind the vdW ones,
  // so the row is [vdW entries | Coulomb-only entries] and the pair kernels'
  // chunks past the vdW entries skip the vdW branch as a whole wave.  The
  // destination lies below the source, and each chunk is read before it is
  // written, so the move is safe in place.
  if (count <= A.cap) {
    // the back entries were stored by other lanes of this wave
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    for (int i0 = 0; i0 < nc; i0 += NB_WAVE) {
      const int i = i0 + lane;
      const uint32_t v = i < nc ? out[A.cap - nc + i] : 0u;
      if (i < nc) out[nv + i] = v;
    }
  }
  if (lane == 0) {
    A.nnb[s] = count < A.cap ? count : A.cap;
    if (count > A.cap) {
      atomicMax(&A.flags[0], count);
      atomicMax(&A.flags[2], count);  // sticky until the host clears it
    }
    A.hcnt[s] = make_int4(hs, ht, hse, hte);
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
// accumulator slots (NB_NE_PAIR = 19 with the clamp count; NB_NE_PAIRQ = 27
// with the sums of cm3d's energy-only coulomb routine: sum q_i q_j, and
// sum q_i q_j r^2 and its tensor, over the te pairs inside the cutoff)
enum {
  E_SPOTL = 0, E_SPOTHL, E_SPRSL, E_SL0, E_SL1, E_SL2, E_SL4, E_SL5, E_SL8,
  E_SPOTE, E_SPOTHE, E_SPRSE, E_SE0, E_SE1, E_SE2, E_SE4, E_SE5, E_SE8, E_CLAMP,
  E_QQ, E_QR2, E_Q0, E_Q1, E_Q2, E_Q4, E_Q5, E_Q8
};

#define NB_IP_LDS 16    // interactions staged in LDS
#define NB_MAP_LDS 256  // type-map entries staged in LDS
// Table interpolation and the force / energy / virial sums may contract to
// FMA; the minimum image, r^2, the cutoff tests and the table index stay in
// cm3d's operation order (no contraction), so that which pairs enter and at
// which table node is decided exactly as cm3d decides it.
static __device__ __forceinline__ double interp3c(const double *t, int k, int c, double p) {
#pragma clang fp contract(fast)
  const double fm1 = t[3 * (k - 1) + c], f0 = t[3 * k + c], f1 = t[3 * (k + 1) + c];
  return f0 + .5 * p * (f1 - fm1 + p * (fm1 - 2. * f0 + f1));
}

// The atom loop.  TLDS: the tables (vt / ct) and the interaction parameters
// (sip / smap) are all in LDS, so the reads compile to LDS reads rather than
// generic (flat) ones; otherwise any of them may be in HBM.  ENER: energies,
// virials and the clamp count are summed (else forces only: the MD step,
// whose energies are evaluated when they are asked for); WF: forces are
// written (else energies only).
// QK: also the energy-only coulomb sums of cm3d's vinter_e.
template <int IFLG, bool WANT_FS, bool TLDS, bool ENER, bool WF, bool QK = false>
static __device__ __forceinline__ void pair64_run(const PairArgs &A, int bid, const double *vt,
                                                  const double *ct, const InterPar *sip,
                                                  const int *smap, bool ip_lds) {
  constexpr int NACC = QK ? NB_NE_PAIRQ : NB_NE_PAIR;
  double acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; k++) acc[k] = 0.;
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
    // chunks of 64 list entries; list entries are loaded two chunks ahead
    // of the evaluation and partner records one chunk ahead, so a record
    // gather never waits on the list load it depends on
    uint32_t ec = lane < nb ? lst[lane] : 0u;
    uint32_t en = NB_WAVE + lane < nb ? lst[NB_WAVE + lane] : 0u;
    double4 rjc = make_double4(0., 0., 0., 0.);
    int tjc = 0;
    if ((ec >> 28) & (fv | fe)) {
      const int j = (int)(ec & A.imask);
      rjc = A.rec[j];
      tjc = A.stype[j];
    }
#pragma unroll 2
    for (int k0 = 0; k0 < nb; k0 += NB_WAVE) {
      const int k2 = k0 + 2 * NB_WAVE + lane;
      const uint32_t e2 = k2 < nb ? lst[k2] : 0u;
      double4 rjn = make_double4(0., 0., 0., 0.);
      int tjn = 0;
      if ((en >> 28) & (fv | fe)) {
        const int j = (int)(en & A.imask);
        rjn = A.rec[j];
        tjn = A.stype[j];
      }
      const unsigned fl = ec >> 28;
      if (fl & (fv | fe)) {
        const double4 rj = rjc;
        double dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
        nb_minimg_box(dx, dy, dz, A.b);
        double d2 = dx * dx + dy * dy + dz * dz;
        if (fl & fv) {
          const int in = TLDS ? smap[ti * A.ntypes + tjc]
                              : (ip_lds ? smap[ti * A.ntypes + tjc] : A.map[ti * A.ntypes + tjc]);
          const InterPar &P = TLDS ? sip[in] : (ip_lds ? sip[in] : A.ip[in]);
          double rmax2 = IFLG ? P.rmaxl2 : P.rmaxs2;
          if (d2 < rmax2) {
            double p = (d2 - P.rmins2) * P.dx2;
            int kk = (int)p;
            if (kk < 1) { if (ENER) acc[E_CLAMP] += 1.; kk = 1; p = 1.; }
            p = p - (double)kk;
            const double *T = vt + (size_t)in * A.ntable * 3;
            double f = interp3c(T, kk, 0, p);
            double swv = 1.;
            if ((WANT_FS || !IFLG) && d2 > P.rminl2) swv = interp3c(T, kk, 1, p);
            {
#pragma clang fp contract(fast)
              if (IFLG) {
                if (ENER) {
                  acc[E_SPRSL] -= d2 * f;
                  acc[E_SL0] -= dx * dx * f;
                  acc[E_SL1] -= dy * dx * f;
                  acc[E_SL2] -= dz * dx * f;
                  acc[E_SL4] -= dy * dy * f;
                  acc[E_SL5] -= dz * dy * f;
                  acc[E_SL8] -= dz * dz * f;
                  double v = interp3c(T, kk, 2, p);
                  acc[E_SPOTL] += v;
                  acc[E_SPOTHL] += v - T[3 * ntm4 + 2];
                }
                if (WF) { fx -= dx * f; fy -= dy * f; fz -= dz * f; }
                if (WANT_FS) { sx -= dx * f * swv; sy -= dy * f * swv; sz -= dz * f * swv; }
              } else {
                f *= swv;
                sx -= dx * f; sy -= dy * f; sz -= dz * f;
              }
            }
          }
        }
        if ((fl & fe) && d2 < A.rmaxl_coul2) {
          double p = (d2 - A.rmins_coul2) * A.dx2coul;
          int kk = (int)p;
          if (kk < 1) { if (ENER) acc[E_CLAMP] += 1.; kk = 1; p = 1; }
          p = p - (double)kk;
          {
#pragma clang fp contract(fast)
            double qij = ri.w * rj.w;
            double fcc = qij * interp3c(ct, kk, 0, p);
            if (IFLG) {
              if (ENER) {
                acc[E_SPRSE] -= d2 * fcc;
                acc[E_SE0] -= dx * dx * fcc;
                acc[E_SE1] -= dy * dx * fcc;
                acc[E_SE2] -= dz * dx * fcc;
                acc[E_SE4] -= dy * dy * fcc;
                acc[E_SE5] -= dz * dy * fcc;
                acc[E_SE8] -= dz * dz * fcc;
                double v = interp3c(ct, kk, 2, p);
                acc[E_SPOTHE] += qij * (v - A.clast);
                acc[E_SPOTE] += qij * v;
                if (QK) {
                  acc[E_QQ] += qij;
                  acc[E_QR2] += qij * d2;
                  acc[E_Q0] += qij * dx * dx;
                  acc[E_Q1] += qij * dy * dx;
                  acc[E_Q2] += qij * dz * dx;
                  acc[E_Q4] += qij * dy * dy;
                  acc[E_Q5] += qij * dz * dy;
                  acc[E_Q8] += qij * dz * dz;
                }
              }
              if (WF) { fx -= dx * fcc; fy -= dy * fcc; fz -= dz * fcc; }
            }
            if (WANT_FS || !IFLG) { sx -= dx * fcc; sy -= dy * fcc; sz -= dz * fcc; }
          }
        }
      }
      ec = en;
      en = e2;
      rjc = rjn;
      tjc = tjn;
    }
    if (WF) {
      const int ai = A.perm[s];
      if (IFLG) { fx = nb_wave_sum(fx); fy = nb_wave_sum(fy); fz = nb_wave_sum(fz); }
      if (WANT_FS || !IFLG) { sx = nb_wave_sum(sx); sy = nb_wave_sum(sy); sz = nb_wave_sum(sz); }
      if (lane == 0) {
        const int n = A.n;
        if (IFLG) { A.fl[ai] = fx; A.fl[n + ai] = fy; A.fl[2 * n + ai] = fz; }
        if (WANT_FS || !IFLG) { A.fs[ai] = sx; A.fs[n + ai] = sy; A.fs[2 * n + ai] = sz; }
      }
    }
  }
  if (ENER) block_reduce_store<NACC>(acc, A.epart + bid, A.nblocks);
}

template <int IFLG, bool WANT_FS, bool ENER, bool WF, bool QK = false>
static __device__ __forceinline__ void pair64_body(const PairArgs &A, int bid) {
  extern __shared__ double tabs[];
  __shared__ InterPar sip[NB_IP_LDS];
  __shared__ int smap[NB_MAP_LDS];
  if (A.tab_lds) {
    int nv = A.ninter * A.ntable * 3, nc = A.ntable * 3;
    for (int t = threadIdx.x; t < nv; t += blockDim.x) tabs[t] = A.vtab3[t];
    for (int t = threadIdx.x; t < nc; t += blockDim.x) tabs[A.ctab_off + t] = A.ctab3[t];
  }
  const bool ip_lds = A.ninter <= NB_IP_LDS && A.ntypes * A.ntypes <= NB_MAP_LDS;
  if (ip_lds) {
    for (int t = threadIdx.x; t < A.ninter; t += blockDim.x) sip[t] = A.ip[t];
    for (int t = threadIdx.x; t < A.ntypes * A.ntypes; t += blockDim.x) smap[t] = A.map[t];
  }
  __syncthreads();
  if (A.tab_lds && ip_lds)
    pair64_run<IFLG, WANT_FS, true, ENER, WF, QK>(A, bid, tabs, tabs + A.ctab_off, sip, smap,
                                                  ip_lds);
  else if (A.tab_lds)
    pair64_run<IFLG, WANT_FS, false, ENER, WF, QK>(A, bid, tabs, tabs + A.ctab_off, sip, smap,
                                                   ip_lds);
  else
    pair64_run<IFLG, WANT_FS, false, ENER, WF, QK>(A, bid, A.vtab3, A.ctab3, sip, smap, ip_lds);
}
// ---------------------------------------------------------------------------
// Forces-only pair kernel of the MD step (k_pair64_1n): fl over the t / te
// lists, no energies.  The same pairs, minimum image, r^2, cut-off tests and
// table node as pair64_run (cm3d's operation order, no contraction), in
// fewer instructions per pair:
//   - a chunk of entries is loaded when it starts inside the row (a
//     wave-uniform test); its entries past the row's end read the slots
//     that follow (the neighbour array ends in NB_LIST_PAD spare slots) and
//     a select zeroes them; every lane gathers its partner's record (entry
//     0 reads atom 0's), so no exec-mask change guards a load; partner
//     types are gathered
//     only for chunks that hold a vdW entry.  The gathers bound the kernel:
//     a lane's record is a cache line of its own, and the L1 serves about
//     one line a cycle;
//   - the coulomb term is evaluated on every lane and kept by a select: a
//     lane whose entry is not in te, or whose pair lies beyond the cut-off,
//     adds zero; its table node is clamped into the table so that the read
//     stays in bounds.  The vdW term runs only where an entry carries the
//     vdW flag (they come first in a row, so later chunks skip it);
//   - the force column is interpolated from per-node coefficients
//     (c0 + p (c1 + p c2): cm3d's quadratic interpolation to rounding);
//   - an orthorhombic box takes the diagonal minimum image.
// ---------------------------------------------------------------------------
template <bool TLDS, bool ORTHO>
static __device__ __forceinline__ void pair64_fonly_run(const PairArgs &A, int bid,
                                                        const double *fc, const InterPar *sip,
                                                        const int *smap) {
  const int lane = lane_id();
  const int nw = A.nblocks * 4;
  const double *cc = fc + A.ctab_off;
  const int kmax = A.ntable - 2;
  // the wave's atom (and so its row length) is wave-uniform: scalar loads
  // and branches
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int s = bid * 4 + wv; s < A.n; s += nw) {
    const double4 ri = A.rec[s];
    const int ti = A.stype[s];
    const int nb = A.nnb[s];
    const uint32_t *lst = A.nbr + (size_t)s * A.cap + lane;
    double fx = 0., fy = 0., fz = 0.;
    // entries two chunks ahead of the evaluation, partner records and
    // types one chunk ahead
    // a chunk that starts past the row's end is not loaded (a wave-uniform
    // branch); entries of a loaded chunk past the row's end are zeroed
    uint32_t ec = lst[0], en = 0u;
    ec = lane < nb ? ec : 0u;
    if (NB_WAVE < nb) {
      en = lst[NB_WAVE];
      en = NB_WAVE + lane < nb ? en : 0u;
    }
    const uint32_t im = A.imask;
    double4 rjc = rec64_at(A.rec, ec & im);
    int tjc = 0;
    if (!A.etype && __any((ec >> 28) & NB_FLAG_VT)) tjc = type_at(A.stype, ec & im);
#pragma unroll 2
    for (int k0 = 0; k0 < nb; k0 += NB_WAVE) {
      // loaded raw; the select that zeroes entries past the row's end is
      // applied a chunk later, when the load has long completed
      uint32_t e2 = 0u;
      if (k0 + 2 * NB_WAVE < nb) e2 = lst[k0 + 2 * NB_WAVE];
      const double4 rjn = rec64_at(A.rec, en & im);
      // partner types from the entry, else gathered only for chunks holding
      // vdW entries (the gathers, one cache line a lane, bound this kernel)
      int tjn = 0;
      if (!A.etype && __any((en >> 28) & NB_FLAG_VT)) tjn = type_at(A.stype, en & im);
      const unsigned fl = ec >> 28;
      double dx = ri.x - rjc.x, dy = ri.y - rjc.y, dz = ri.z - rjc.z;
      if (ORTHO) {
        double sx = dx * A.b.hi[0], sy = dy * A.b.hi[4], sz = dz * A.b.hi[8];
        sx -= nb_anint(sx);
        sy -= nb_anint(sy);
        sz -= nb_anint(sz);
        dx = sx * A.b.h[0];
        dy = sy * A.b.h[4];
        dz = sz * A.b.h[8];
      } else {
        nb_minimg(dx, dy, dz, A.b.h, A.b.hi);
      }
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (fl & NB_FLAG_VT) {
        const int tj = A.etype ? (int)((ec & NB_IDX_MASK) >> A.ibits) : tjc;
        const int in = TLDS ? smap[ti * A.ntypes + tj] : A.map[ti * A.ntypes + tj];
        const InterPar &P = TLDS ? sip[in] : A.ip[in];
        if (d2 < P.rmaxl2) {
          double p = (d2 - P.rmins2) * P.dx2;
          int kk = (int)p;
          if (kk < 1) { kk = 1; p = 1.; }
          kk = min(kk, kmax);
          p = p - (double)kk;
          const double *T = fc + ((size_t)in * A.ntable + kk) * 3;
          {
#pragma clang fp contract(fast)
            const double f = T[0] + p * (T[1] + p * T[2]);
            fx -= dx * f;
            fy -= dy * f;
            fz -= dz * f;
          }
        }
      }
      {
        double p = (d2 - A.rmins_coul2) * A.dx2coul;
        int kk = (int)p;
        const bool lo = kk < 1;
        p = lo ? 1. : p;
        kk = lo ? 1 : min(kk, kmax);
        p = p - (double)kk;
        const double *T = cc + kk * 3;
        const bool on = (fl & NB_FLAG_ET) && d2 < A.rmaxl_coul2;
        {
#pragma clang fp contract(fast)
          const double f = T[0] + p * (T[1] + p * T[2]);
          const double fcc = on ? ri.w * rjc.w * f : 0.;
          fx -= dx * fcc;
          fy -= dy * fcc;
          fz -= dz * fcc;
        }
      }
      ec = en;
      en = k0 + 2 * NB_WAVE + lane < nb ? e2 : 0u;
      rjc = rjn;
      tjc = tjn;
    }
    fx = nb_wave_sum(fx);
    fy = nb_wave_sum(fy);
    fz = nb_wave_sum(fz);
    if (lane == 0) {
      const int ai = A.perm[s], n = A.n;
      A.fl[ai] = fx;
      A.fl[n + ai] = fy;
      A.fl[2 * n + ai] = fz;
    }
  }
}

static __device__ __forceinline__ void pair64_fonly_body(const PairArgs &A, int bid) {
  extern __shared__ double tabs[];
  __shared__ InterPar sip[NB_IP_LDS];
  __shared__ int smap[NB_MAP_LDS];
  const bool lds = A.fc_lds && A.ninter <= NB_IP_LDS && A.ntypes * A.ntypes <= NB_MAP_LDS;
  if (lds) {
    const int nd = A.fc_lds / (int)sizeof(double);
    for (int t = threadIdx.x; t < nd; t += blockDim.x) tabs[t] = A.fcoef[t];
    for (int t = threadIdx.x; t < A.ninter; t += blockDim.x) sip[t] = A.ip[t];
    for (int t = threadIdx.x; t < A.ntypes * A.ntypes; t += blockDim.x) smap[t] = A.map[t];
  }
  __syncthreads();
  if (lds) {
    if (A.b.ortho) pair64_fonly_run<true, true>(A, bid, tabs, sip, smap);
    else pair64_fonly_run<true, false>(A, bid, tabs, sip, smap);
  } else {
    if (A.b.ortho) pair64_fonly_run<false, true>(A, bid, A.fcoef, sip, smap);
    else pair64_fonly_run<false, false>(A, bid, A.fcoef, sip, smap);
  }
}

// 0t: force(...,0) (fs, clamp count); 1t: force(...,1) with fs and energies;
// 1f: fl and energies; 1n: fl only (the MD step); 1e: energies only
#define k_pair64_0t_body pair64_body<0, true, true, true>
#define k_pair64_1t_body pair64_body<1, true, true, true>
#define k_pair64_1f_body pair64_body<1, false, true, true>
#define k_pair64_1n_body pair64_fonly_body
#define k_pair64_1e_body pair64_body<1, false, true, false>
// 1q: 1e + the sums of cm3d's energy-only coulomb routine (nb_energy_only)
#define k_pair64_1q_body pair64_body<1, false, true, false, true>
NB_KERNEL_PAIR_W(k_pair64_0t, PairArgs, 256, 4)
NB_KERNEL_PAIR_W(k_pair64_1t, PairArgs, 256, 3)
NB_KERNEL_PAIR_W(k_pair64_1f, PairArgs, 256, 4)
NB_KERNEL_PAIR_W(k_pair64_1n, PairArgs, 256, 4)
NB_KERNEL_PAIR_W(k_pair64_1e, PairArgs, 256, 4)
NB_KERNEL_PAIR_W(k_pair64_1q, PairArgs, 256, 4)

// ---------------------------------------------------------------------------
// Reciprocal-space Ewald, FP64 throughout (fk_ewald, md_ewald.c:185-361).
// The k vectors are taken as rows: runs of consecutive kc at one (ka, kb),
// along which e^{i k.r} advances by the c-axis step e^{2 pi i s_c} (cm3d's
// recursion, md_ewald.c:261-279).  Both passes walk a row by m = |kc| and
// take kc = +m and -m together: e^{i k.r} = w e^{+-i m theta_c} with w the
// row's (ka, kb) phase, so one step of e^{i m theta_c} serves two k
// vectors.  Sines and cosines come from nb_sincos.h.  What a whole wavefront
// reads at once (the charges in k_ewsfac, the k data in k_ewforce) is staged
// in LDS and read as broadcasts.  FMA contraction is allowed in these
// kernels: the sums only have to agree with cm3d's to rounding.
// k_ewchg: per charge, fractional coordinates s = h^-1 r wrapped to
//   [-1/2,1/2) (exact for e^{2 pi i k.s} with integer k), stored as 2 pi s,
//   the charge times sqrt(4 pi) (md_ewald.c:220-224), e^{2 pi i s_c} and
//   e^{2 pi i NB_EW_SYM s_c}.
// k_ewsfac: lanes = runs of <= NB_EW_SYM values of m in one row, or, in a
//   system whose rows are all short, 2 / 4 rows of consecutive kb with <= 8
//   / 4 values of m each (sorted longest first, so a wave's trip count is
//   its first lane's); a workgroup walks one chunk of charges and each lane sums
//   q w e^{i m theta_c} split into its cosine and sine parts in registers --
//   no cross-lane reduction -- then the 4 waves combine and write the
//   chunk's partial S(k) for +m and -m.
// k_ewkred: per k, the chunk partials summed in fixed order, then preg, prep,
//   |S|^2 and the energy / virial terms (md_ewald.c:247-302).
// k_ewforce: lanes = charges; a workgroup walks one range of rows in k
//   order, tile by tile (the next kb of a ka plane steps w by e^{i theta_b}),
//   summing k (Re S' sin - Im S' cos) (md_ewald.c:303-323) into a per-range
//   partial.
// k_ewfin: per charge, the range partials in fixed order, times q, added
//   to fl, with the excluded-pair correction (ecorr).
// ---------------------------------------------------------------------------
template <typename T>
using cptr4 = __attribute__((address_space(4))) const T *;
// The constant address space: uniform loads through it become scalar loads.
template <typename T>
static __device__ __forceinline__ cptr4<T> as_const4(const T *p) {
  return (cptr4<T>)(unsigned long long)p;
}

struct EwChgArgs {
  const double *pos, *q; const int *qidx; int nq, n; Box b; double4 *chg, *stp;
};
static __device__ __forceinline__ void k_ewchg_body(const EwChgArgs &A, int bid) {
  int c = bid * blockDim.x + threadIdx.x;
  if (c >= A.nq) return;
  const int n = A.n, a = A.qidx[c];
  nb_charge_record(A.pos[a], A.pos[n + a], A.pos[2 * n + a], A.q[a], A.b.hi, A.chg + c, A.stp + c);
}
NB_KERNEL_PAIR(k_ewchg, EwChgArgs, 256)

struct EwSfacArgs {
  const EwSym *sym; const double4 *chg, *stp; int nq, qc, nsym, nsw, ktot;
  double2 *part;
};
// LDS of k_ewsfac: the charge tile (8 doubles a charge) during the sweep,
// then one wave's sums [4 NB_EW_SYM][64] at a time in the cross-wave
// reduction
#define EW_SFAC_LDS 4096
#define EW_QT (EW_SFAC_LDS / 8)  // charges per tile
// One charge against a lane's run of M values of m from m0 in one row, with
// w = q e^{i (ka theta_a + kb theta_b)} given: the phase z^m = e^{i m
// theta_c} from z^{m0} (1 or z times powers of z^NB_EW_SYM, or a sine /
// cosine when m0 is neither) by the three-term recurrence of cos(m theta)
// and sin(m theta) (an error introduced at step j reaches step m at most
// m - j times larger: |z^m| error <= m^2 eps / 2, 1e-13 at m = 33), and
// A_m += w Re z^m, B_m += w Im z^m: then S(+m) = A_m + i B_m and
// S(-m) = A_m - i B_m.  Runs past the lane's own count accumulate into
// slots that are never stored.
template <int M>
static __device__ __forceinline__ void sfac_charge(const double *cq, double wr, double wi,
                                                   int m0, int m16, int mb,
                                                   double (&ar)[NB_EW_SYM], double (&ai)[NB_EW_SYM],
                                                   double (&br)[NB_EW_SYM], double (&bi)[NB_EW_SYM]) {
  const double zr = cq[4], zi = cq[5];
  double c = 1., s = 0.;
  if (mb == 1) { c = zr; s = zi; }
  if (mb > 1) {
    nb_sincos(m0 * cq[2], &s, &c);
  } else {
    for (int j = 0; j < m16; j++) {
      const double c2 = c * cq[6] - s * cq[7];
      s = s * cq[6] + c * cq[7];
      c = c2;
    }
  }
  // z^{m0 - 1}, and the recurrence z^{m+1} = 2 cos(theta_c) z^m - z^{m-1},
  // component by component
  double cp = c * zr + s * zi, sp = s * zr - c * zi;
  const double t2 = 2. * zr;
#pragma unroll
  for (int m = 0; m < M; m++) {
    ar[m] = __builtin_fma(wr, c, ar[m]);
    ai[m] = __builtin_fma(wi, c, ai[m]);
    br[m] = __builtin_fma(wr, s, br[m]);
    bi[m] = __builtin_fma(wi, s, bi[m]);
    if (m + 1 < M) {
      const double cn = __builtin_fma(t2, c, -cp), sn = __builtin_fma(t2, s, -sp);
      cp = c; sp = s;
      c = cn; s = sn;
    }
  }
}
// One charge against a lane of R rows (ka, kb + r) x M values of m from 0:
// w_r = w e^{i r theta_b} by steps of e^{i theta_b} (eb), z^m from 1 by the
// three-term recurrence, shared by the R rows.  Accumulator r * M + m.
template <int R, int M>
static __device__ __forceinline__ void sfac_charge_rows(const double *cq, const double *eb,
                                                        double wr, double wi,
                                                        double (&ar)[NB_EW_SYM], double (&ai)[NB_EW_SYM],
                                                        double (&br)[NB_EW_SYM], double (&bi)[NB_EW_SYM]) {
  const double zr = cq[4], zi = cq[5];
  double wrr[R], wir[R];
  wrr[0] = wr;
  wir[0] = wi;
#pragma unroll
  for (int r = 1; r < R; r++) {
    wrr[r] = wrr[r - 1] * eb[0] - wir[r - 1] * eb[1];
    wir[r] = wir[r - 1] * eb[0] + wrr[r - 1] * eb[1];
  }
  // z^0 = 1 and z^-1 = conj z
  double c = 1., s = 0., cp = zr, sp = -zi;
  const double t2 = 2. * zr;
#pragma unroll
  for (int m = 0; m < M; m++) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      ar[r * M + m] = __builtin_fma(wrr[r], c, ar[r * M + m]);
      ai[r * M + m] = __builtin_fma(wir[r], c, ai[r * M + m]);
      br[r * M + m] = __builtin_fma(wrr[r], s, br[r * M + m]);
      bi[r * M + m] = __builtin_fma(wir[r], s, bi[r * M + m]);
    }
    if (m + 1 < M) {
      const double cn = __builtin_fma(t2, c, -cp), sn = __builtin_fma(t2, s, -sp);
      cp = c; sp = s;
      c = cn; s = sn;
    }
  }
}
// w = q e^{i (ka theta_a + kb theta_b)} for charge t of the tile, from one
// sine / cosine
static __device__ __forceinline__ void sfac_w(const double *lds, int t, double ka, double kb,
                                              double *wr, double *wi) {
  const double *cq = lds + 8 * t;
  double sn, cs;
  nb_sincos(ka * cq[0] + kb * cq[1], &sn, &cs);
  *wr = cq[3] * cs;
  *wi = cq[3] * sn;
}
template <int R, int M>
static __device__ __forceinline__ void sfac_run(const EwSfacArgs &A, double *lds, double *ldsb,
                                                int chunk, const EwSym &y, bool live) {
  static_assert(R * M <= NB_EW_SYM, "a lane holds NB_EW_SYM (row, m) accumulators");
  constexpr int RM = R * M;
  const int lane = lane_id(), w = threadIdx.x >> 6;
  double ar[NB_EW_SYM], ai[NB_EW_SYM], br[NB_EW_SYM], bi[NB_EW_SYM];
#pragma unroll
  for (int m = 0; m < RM; m++) { ar[m] = 0.; ai[m] = 0.; br[m] = 0.; bi[m] = 0.; }
  const double ka = y.ka, kb = y.kb;
  const int mb = y.m0 % NB_EW_SYM, m16 = y.m0 / NB_EW_SYM;
  const int c0 = chunk * A.qc, c1 = min(A.nq, c0 + A.qc);
  for (int t0 = c0; t0 < c1; t0 += EW_QT) {
    const int nt = min(EW_QT, c1 - t0);
    __syncthreads();
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
      const double4 r = A.chg[t0 + t];
      const double4 st = A.stp[t0 + t];
      double *d = lds + 8 * t;
      d[0] = r.x; d[1] = r.y; d[2] = r.z; d[3] = r.w;
      d[4] = st.x; d[5] = st.y; d[6] = st.z; d[7] = st.w;
      if constexpr (R > 1) nb_sincos(r.y, &ldsb[2 * t + 1], &ldsb[2 * t]);  // e^{i theta_b}
    }
    __syncthreads();
    int t = w;
    double wr, wi, wr2, wi2;
#if NB_SFAC_PAIR
    for (; t + 4 < nt; t += 8) {  // two charges per trip for ILP
      sfac_w(lds, t, ka, kb, &wr, &wi);
      sfac_w(lds, t + 4, ka, kb, &wr2, &wi2);
      if constexpr (R == 1) {
        sfac_charge<M>(lds + 8 * t, wr, wi, y.m0, m16, mb, ar, ai, br, bi);
        sfac_charge<M>(lds + 8 * (t + 4), wr2, wi2, y.m0, m16, mb, ar, ai, br, bi);
      } else {
        sfac_charge_rows<R, M>(lds + 8 * t, ldsb + 2 * t, wr, wi, ar, ai, br, bi);
        sfac_charge_rows<R, M>(lds + 8 * (t + 4), ldsb + 2 * (t + 4), wr2, wi2, ar, ai, br, bi);
      }
    }
#endif
    for (; t < nt; t += 4) {
      sfac_w(lds, t, ka, kb, &wr, &wi);
      if constexpr (R == 1) sfac_charge<M>(lds + 8 * t, wr, wi, y.m0, m16, mb, ar, ai, br, bi);
      else sfac_charge_rows<R, M>(lds + 8 * t, ldsb + 2 * t, wr, wi, ar, ai, br, bi);
    }
  }
  // ((w3 + w2) + w1) + w0, one wave's sums at a time through the same LDS
  double (*red)[NB_WAVE] = (double (*)[NB_WAVE])lds;
#pragma unroll
  for (int src = 3; src >= 1; src--) {
    __syncthreads();
    if (w == src) {
#pragma unroll
      for (int m = 0; m < RM; m++) {
        if (src < 3) {
          ar[m] += red[4 * m][lane]; ai[m] += red[4 * m + 1][lane];
          br[m] += red[4 * m + 2][lane]; bi[m] += red[4 * m + 3][lane];
        }
        red[4 * m][lane] = ar[m]; red[4 * m + 1][lane] = ai[m];
        red[4 * m + 2][lane] = br[m]; red[4 * m + 3][lane] = bi[m];
      }
    }
  }
  __syncthreads();
  if (w == 0 && live) {
    double2 *out = A.part + (size_t)chunk * A.ktot;
#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int m = 0; m < M; m++) {
        const int q = r * M + m;
        if (m < y.cnt) {
          const double sar = ar[q] + red[4 * q][lane], sai = ai[q] + red[4 * q + 1][lane];
          const double sbr = br[q] + red[4 * q + 2][lane], sbi = bi[q] + red[4 * q + 3][lane];
          const int mm = y.m0 + m;
          if (mm >= y.kc_lo[r] && mm <= y.kc_hi[r])
            out[y.k_lo[r] + mm - y.kc_lo[r]] = make_double2(sar - sbi, sai + sbr);
          if (mm > 0 && -mm >= y.kc_lo[r] && -mm <= y.kc_hi[r])
            out[y.k_lo[r] - mm - y.kc_lo[r]] = make_double2(sar + sbi, sai - sbr);
        }
      }
    }
  }
}
// block = 4 waves on the same 64 lanes; the block's chunk of charges is
// staged through LDS in tiles and each wave takes every 4th charge; the 4
// waves' sums are then combined in LDS in a fixed order.  Lane 0 (always a
// real lane) has the group's longest run; the 4 waves of the block share the
// lanes, so the trip counts are uniform over the block.  Lanes of one row
// (k_ewsfac) and of several rows (k_ewsfacr) are separate kernels, so that
// the one-row kernel keeps its smaller register and LDS footprint (its
// waves then leave room on a SIMD for the list build and the pair kernel on
// the MD step's other branch).
static __device__ __forceinline__ EwSym sfac_lane(const EwSfacArgs &A, int bid, int *chunk,
                                                  bool *live) {
  const int lane = lane_id();
  const int sw = bid % A.nsw;
  *chunk = bid / A.nsw;
  const int g = sw * NB_WAVE + lane;
  EwSym y;
  y.ka = y.kb = y.m0 = y.cnt = 0;
  y.nr = 1;
  for (int q = 0; q < 4; q++) { y.k_lo[q] = 0; y.kc_lo[q] = 1; y.kc_hi[q] = 0; }  // no k: nothing stored
  *live = g < A.nsym;
  if (*live) y = A.sym[g];
  return y;
}
static __device__ __forceinline__ void k_ewsfac_body(const EwSfacArgs &A, int bid) {
  __shared__ double lds[EW_SFAC_LDS];
  static_assert(8 * EW_QT <= EW_SFAC_LDS && 4 * NB_EW_SYM * NB_WAVE <= EW_SFAC_LDS,
                "charge tile or reduction exceeds the LDS buffer");
  static_assert(NB_EW_SYM == 16, "the dispatch below covers runs of 1..16 values of m");
  int chunk;
  bool live;
  const EwSym y = sfac_lane(A, bid, &chunk, &live);
  switch (__builtin_amdgcn_readfirstlane(y.cnt)) {
#define EW_SFAC_CASE(M) \
  case M: sfac_run<1, M>(A, lds, nullptr, chunk, y, live); break;
    EW_SFAC_CASE(1) EW_SFAC_CASE(2) EW_SFAC_CASE(3) EW_SFAC_CASE(4)
    EW_SFAC_CASE(5) EW_SFAC_CASE(6) EW_SFAC_CASE(7) EW_SFAC_CASE(8)
    EW_SFAC_CASE(9) EW_SFAC_CASE(10) EW_SFAC_CASE(11) EW_SFAC_CASE(12)
    EW_SFAC_CASE(13) EW_SFAC_CASE(14) EW_SFAC_CASE(15) EW_SFAC_CASE(16)
#undef EW_SFAC_CASE
    default: break;
  }
}
static __device__ __forceinline__ void k_ewsfacr_body(const EwSfacArgs &A, int bid) {
  __shared__ double lds[EW_SFAC_LDS];
  __shared__ double ldsb[2 * EW_QT];  // e^{i theta_b} of the tile's charges
  int chunk;
  bool live;
  const EwSym y = sfac_lane(A, bid, &chunk, &live);
  switch (__builtin_amdgcn_readfirstlane(y.nr * 32 + y.cnt)) {
#define EW_SFAC_CASE(R, M) \
  case R * 32 + M: sfac_run<R, M>(A, lds, ldsb, chunk, y, live); break;
    EW_SFAC_CASE(2, 1) EW_SFAC_CASE(2, 2) EW_SFAC_CASE(2, 3) EW_SFAC_CASE(2, 4)
    EW_SFAC_CASE(2, 5) EW_SFAC_CASE(2, 6) EW_SFAC_CASE(2, 7) EW_SFAC_CASE(2, 8)
    EW_SFAC_CASE(4, 1) EW_SFAC_CASE(4, 2) EW_SFAC_CASE(4, 3) EW_SFAC_CASE(4, 4)
#undef EW_SFAC_CASE
    default: break;
  }
}
#ifndef NB_SFAC_WPE
#define NB_SFAC_WPE 1
#endif
NB_KERNEL_PAIR_W(k_ewsfac, EwSfacArgs, 256, NB_SFAC_WPE)
NB_KERNEL_PAIR_W(k_ewsfacr, EwSfacArgs, 256, NB_SFAC_WPE)

// k_ewsfacm: the structure factor on the FP64 matrix cores.  The MFMA
// v_mfma_f64_16x16x4f64 adds A[16][4] B[4][16] to D[16][16]; lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16], and D[l / 16 + 4 r][l % 16] as the
// r-th double of the result.  Here the K index is a charge (4 a step), A's
// row i = c 8 + g is Re (c = 0) or Im (c = 1) of w = q e^{i (ka theta_a +
// kb theta_b)} for row g of a tile of 8 k rows, and B's column j = s 8 + u
// is cos (s = 0) or sin (s = 1) of m theta_c for m = 8 b + u of m tile b,
// so D[c 8 + g][s 8 + u] is Re / Im of A_m (s = 0) or B_m (s = 1) of
// k_ewsfac for 8 rows x 8 values of m.  The phases come from per-charge
// tables (k_ewsfmtab): q e^{i ka theta_a} stored {im, re, -im}, so that
// lane c reads at 1 - c the pair (p, q) with w's Re / Im = p Re t + q Im t
// for t = e^{i kb theta_b} (stored for |kb|, conjugated by the row's sign),
// and e^{i m theta_c}.  A tile of NB_SFM_NQ charges' tables is copied HBM
// -> LDS by direct loads (no registers) into one of two buffers while the
// products of the previous tile run, and within a tile the operands of the
// next K step are loaded while the current one's products issue.  A
// workgroup takes 4 waves' row tiles (runs of consecutive kb in one ka
// plane, sorted by m range, longest first: a tile's rows read one a-table
// entry and 8 consecutive b-table entries) and one chunk of charges; each
// wave owns its tiles and every charge of the chunk, so the chunk's
// partial S(k) leaves the registers with no cross-wave sum.
struct EwSfmArgs {
  const EwRowM *rowm; const EwSfmWave *wave; const double *tab;
  int nq, qc, nrwg, nchunk, la, lb, lz, ktot;
  double2 *part;
};
// k_ewsfmtab: one table entry per thread, from one sine / cosine of the
// entry's multiple of the charge's phase (padding charges past nq: zero
// charge, unit phases); SFM_NE_MAX threads a charge, the ones past its
// entries idle
#define SFM_NE_MAX (2 * NB_SFM_TAB + 8 * NB_SFM_NT)
struct EwSfmTabArgs {
  const double4 *chg; int nq, nqp, la, lb, lz; double *tab;
};
static __device__ __forceinline__ void k_ewsfmtab_body(const EwSfmTabArgs &A, int bid) {
  const int sc = nb_sfm_stride(A.la, A.lb, A.lz);
  const unsigned idx = (unsigned)bid * 256u + threadIdx.x;
  const int t = (int)(idx / (unsigned)SFM_NE_MAX), e = (int)(idx % (unsigned)SFM_NE_MAX);
  if (t >= A.nqp || e >= A.la + A.lb + A.lz) return;
  const double4 r = t < A.nq ? A.chg[t] : make_double4(0., 0., 0., 0.);
  double *d = A.tab + (size_t)t * sc;
  // entry e: the a table (ka = e), the b table (kb = e - la), the c table
  // (m = e - la - lb); one sine / cosine for all three
  const int ax = e < A.la ? 0 : (e < A.la + A.lb ? 1 : 2);
  const int n = ax == 0 ? e : (ax == 1 ? e - A.la : e - A.la - A.lb);
  const double th = ax == 0 ? r.x : (ax == 1 ? r.y : r.z);
  double sn, cs;
  nb_sincos(n * th, &sn, &cs);
  if (ax == 0) {
    const double q = t < A.nq ? r.w : 0.;
    d[3 * n] = q * sn; d[3 * n + 1] = q * cs; d[3 * n + 2] = -(q * sn);
  } else {
    *(double2 *)(d + (ax == 1 ? nb_sfm_tb(A.la) : nb_sfm_tz(A.la, A.lb)) + 2 * n) =
        make_double2(cs, sn);
  }
}
NB_KERNEL_PAIR(k_ewsfmtab, EwSfmTabArgs, 256)

typedef double sfm_d4 __attribute__((ext_vector_type(4)));
#define SFM_SC_MAX (3 * NB_SFM_TAB + 2 * NB_SFM_TAB + 2 * 8 * NB_SFM_NT + 4)
typedef __attribute__((address_space(3))) void *sfm_lptr;
// tile tl of the chunk (tables of charges c0 + NB_SFM_NQ tl ..) into an LDS
// buffer: 1 KB per wave-load, 16 B a lane, waited for by the caller
static __device__ __forceinline__ void sfm_load(const EwSfmArgs &A, double *dst, int c0, int tl,
                                                int sc) {
  const int bytes = NB_SFM_NQ * sc * 8, lane = lane_id();
  const char *src = (const char *)(A.tab + (size_t)(c0 + tl * NB_SFM_NQ) * sc);
  for (int q = threadIdx.x >> 6; q * 1024 < bytes; q += 4) {
    const int off = q * 1024 + lane * 16;
    if (off < bytes)
      __builtin_amdgcn_global_load_lds(src + off, (sfm_lptr)((char *)dst + q * 1024), 16, 0, 0);
  }
}
// the operands of one K step: per row tile the a-table pair and the b-table
// entry, per m tile the cos / sin
template <int NT, int R>
struct SfmOps { double y[NT]; double p[R], q[R]; double2 t[R]; };
template <int NT, int R>
static __device__ __forceinline__ void sfm_ops(SfmOps<NT, R> &o, const double *cb,
                                               const int *aoff, const int *boff, int zoff) {
#pragma unroll
  for (int b = 0; b < NT; b++) o.y[b] = cb[zoff + 16 * b];
#pragma unroll
  for (int r = 0; r < R; r++) {
    o.p[r] = cb[aoff[r]];
    o.q[r] = cb[aoff[r] + 1];
    o.t[r] = *(const double2 *)(cb + boff[r]);
  }
}
// the products of one K step from its operands
template <int NT, int R>
static __device__ __forceinline__ void sfm_step(const SfmOps<NT, R> &o, const double *sg,
                                                sfm_d4 (&acc)[R][NT]) {
#pragma clang fp contract(fast)
  double x[R];
#pragma unroll
  for (int r = 0; r < R; r++) x[r] = o.p[r] * o.t[r].x + (o.q[r] * sg[r]) * o.t[r].y;
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int b = 0; b < NT; b++)
      acc[r][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[r], o.y[b], acc[r][b], 0, 0, 0);
}
// the products of one staged tile (NB_SFM_NQ / 4 matrix-core K steps), two
// operand sets in turn: each step's loads are issued before the previous
// step's products
template <int NT, int R>
static __device__ __forceinline__ void sfm_tile(const double *tab, int sc, const int *aoff,
                                                const int *boff, const double *sg, int zoff,
                                                sfm_d4 (&acc)[R][NT]) {
  constexpr int KS = NB_SFM_NQ / 4;
  static_assert(KS % 2 == 0, "k_ewsfacm: K steps in pairs");
  const int kq = lane_id() >> 4;
  SfmOps<NT, R> o0, o1;
  sfm_ops<NT, R>(o0, tab + kq * sc, aoff, boff, zoff);
#pragma unroll 1
  for (int ks = 0; ks < KS; ks += 2) {
    sfm_ops<NT, R>(o1, tab + ((ks + 1) * 4 + kq) * sc, aoff, boff, zoff);
    sfm_step<NT, R>(o0, sg, acc);
    if (ks + 2 < KS) sfm_ops<NT, R>(o0, tab + ((ks + 2) * 4 + kq) * sc, aoff, boff, zoff);
    sfm_step<NT, R>(o1, sg, acc);
  }
}
template <int NT, int R>
static __device__ __forceinline__ void sfm_run(const EwSfmArgs &A, double *const (&buf)[2],
                                               int chunk, const EwSfmWave &wv) {
  const int sc = nb_sfm_stride(A.la, A.lb, A.lz);
  const int lane = lane_id(), kq = lane >> 4, i = lane & 15;
  const int tb = nb_sfm_tb(A.la);
  int aoff[R], boff[R];
  double sg[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const EwRowM e = A.rowm[wv.t[r] * 8 + (i & 7)];
    aoff[r] = 3 * e.ka + 1 - (i >> 3);
    boff[r] = tb + 2 * abs(e.kb);
    sg[r] = e.kb < 0 ? -1. : 1.;
  }
  const int zoff = nb_sfm_tz(A.la, A.lb) + 2 * (i & 7) + (i >> 3);
  sfm_d4 acc[R][NT];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int b = 0; b < NT; b++) acc[r][b] = sfm_d4{0., 0., 0., 0.};
  const int c0 = chunk * A.qc, nc = min(A.nq, c0 + A.qc) - c0;
  const int ntl = (nc + NB_SFM_NQ - 1) / NB_SFM_NQ;  // tiles (the tables are padded to whole ones)
  sfm_load(A, buf[0], c0, 0, sc);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  // two tiles a trip so that each buffer is one LDS object at every access
  for (int tl = 0; tl < ntl; tl += 2) {
    if (tl + 1 < ntl) sfm_load(A, buf[1], c0, tl + 1, sc);
    if constexpr (NT > 0) sfm_tile<NT, R>(buf[0], sc, aoff, boff, sg, zoff, acc);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tl + 1 >= ntl) break;
    if (tl + 2 < ntl) sfm_load(A, buf[0], c0, tl + 2, sc);
    if constexpr (NT > 0) sfm_tile<NT, R>(buf[1], sc, aoff, boff, sg, zoff, acc);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  if constexpr (NT > 0) {
    // lane: m = 8 b + (i & 7); s = i >> 3 (0: A_m, 1: B_m, the other in
    // lane ^ 8); rows kq and kq + 4 of each tile (doubles h and 2 + h:
    // Re and Im)
    double2 *out = A.part + (size_t)chunk * A.ktot;
    const int s = i >> 3;
#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const EwRowM e = A.rowm[wv.t[r] * 8 + kq + 4 * h];
#pragma unroll
        for (int b = 0; b < NT; b++) {
          const double xr = acc[r][b][h], xi = acc[r][b][2 + h];
          const double yr = __shfl_xor(xr, 8), yi = __shfl_xor(xi, 8);
          const int m = 8 * b + (i & 7);
          if (s == 0) {
            if (m >= e.kc_lo && m <= e.kc_hi)
              out[e.k_lo + m - e.kc_lo] = make_double2(xr - yi, xi + yr);
          } else if (m > 0 && -m >= e.kc_lo && -m <= e.kc_hi) {
            out[e.k_lo - m - e.kc_lo] = make_double2(yr + xi, yi - xr);
          }
        }
      }
    }
  }
}
static __device__ __forceinline__ void k_ewsfacm_body(const EwSfmArgs &A, int bid) {
  __shared__ double sfm_b0[NB_SFM_NQ * SFM_SC_MAX], sfm_b1[NB_SFM_NQ * SFM_SC_MAX];
  double *const buf[2] = {sfm_b0, sfm_b1};
  // blocks are dealt to the 8 XCDs in turn: the nrwg workgroups of a chunk
  // go to one XCD, so its tables are read from that XCD's L2
  const int xcd = bid & 7, j = bid >> 3;
  const int chunk = (j / A.nrwg) * 8 + xcd, rwg = j % A.nrwg;
  if (chunk >= A.nchunk) return;
  const EwSfmWave wv = A.wave[rwg * 4 + (threadIdx.x >> 6)];
  switch (__builtin_amdgcn_readfirstlane(wv.nt)) {
    case 1: sfm_run<1, NB_SFM_R>(A, buf, chunk, wv); break;
    case 2: sfm_run<2, NB_SFM_R>(A, buf, chunk, wv); break;
    case 3: sfm_run<3, NB_SFM_R>(A, buf, chunk, wv); break;
    case 4: sfm_run<4, nb_sfm_wave_tiles(4)>(A, buf, chunk, wv); break;
    case 5: sfm_run<5, nb_sfm_wave_tiles(5)>(A, buf, chunk, wv); break;
    default: sfm_run<0, 1>(A, buf, chunk, wv); break;  // loads with the others
  }
}
static_assert(NB_SFM_NT == 5, "k_ewsfacm dispatches m-tile counts 1..5");
static_assert(NB_SFM_NQ % 4 == 0, "k_ewsfacm: whole K steps of 4 charges a tile");
NB_KERNEL_PAIR_W(k_ewsfacm, EwSfmArgs, 256, 2)

// per k: 2 preg S(k), the amplitude the force pass needs
struct KData { double cr, ci; };

// per k: the chunk partials in fixed order (a workgroup = 64 k x 4 chunk
// quarters, the quarters combined in order), then the k-space terms
struct EwKredArgs {
  const int *kabc; int ktot; const double2 *part; int nchunk, nblocks; Box b; double rvol, falp2;
  KData *kd; double *epart;
};
static __device__ __forceinline__ void k_ewkred_body(const EwKredArgs &A, int bid) {
  __shared__ double2 qs[4][NB_WAVE];
  double acc[NB_NE_EW];
#pragma unroll
  for (int k = 0; k < NB_NE_EW; k++) acc[k] = 0.;
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const int i = bid * NB_WAVE + lane;
  {
    double sr = 0., si = 0.;
    if (i < A.ktot) {
      const int q0 = (A.nchunk * w) / 4, q1 = (A.nchunk * (w + 1)) / 4;
#pragma unroll 8
      for (int c = q0; c < q1; c++) {
        double2 p = A.part[(size_t)c * A.ktot + i];
        sr += p.x;
        si += p.y;
      }
    }
    qs[w][lane] = make_double2(sr, si);
  }
  __syncthreads();
  if (w == 0 && i < A.ktot) {
    double sr = ((qs[0][lane].x + qs[1][lane].x) + qs[2][lane].x) + qs[3][lane].x;
    double si = ((qs[0][lane].y + qs[1][lane].y) + qs[2][lane].y) + qs[3][lane].y;
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
    d.cr = sr * 2. * preg;
    d.ci = si * 2. * preg;
    A.kd[i] = d;
  }
  block_reduce_store<NB_NE_EW>(acc, A.epart + bid, A.nblocks);
}
NB_KERNEL_PAIR(k_ewkred, EwKredArgs, 256)

// k_ewforce: tiles of whole rows (<= NB_EW_KT entries, one per value of m)
// are staged in LDS; a range is a run of tiles (rtile[r] .. rtile[r+1]);
// tile t covers rows trow[t] .. trow[t+1].  Each lane takes NB_EWF_Q
// charges.  In a row (ka, kb fixed) the k vector is A + kc Gc with
// A = 2 pi (ka h^-T_a + kb h^-T_b) and Gc = 2 pi h^-T_c, and the force sums
// sum_k k tt_k with tt_k = Im(conj(K_k) e^{i k.r}), K_k = 2 preg S(k).  With
// w = e^{i (ka theta_a + kb theta_b)}, z = e^{i theta_c} and, per m = |kc|,
// E = conj(K_{+m}) + conj(K_{-m}), D = conj(K_{+m}) - conj(K_{-m}):
//   sum_row tt    = Im(w P),  P = sum_m (E Re z^m + i D Im z^m)
//   sum_row kc tt = Im(w Q),  Q = sum_m m (D Re z^m + i E Im z^m)
// so a row costs one step of z^m per value of m for two k vectors (the
// three-term recurrence of cos(m theta_c) and sin(m theta_c), as in
// k_ewsfac).  An LDS entry holds {Er, Ei, -Di, Dr, m Dr, -m Ei, m Di, m Er}.
struct EwForceArgs {
  const EwRow *row; const int4 *ent; const int *rtile, *trow; const KData *kd;
  const double4 *chg, *stp; int nq, nqb; double *fpart; Box b;
};
#define EWF_Q NB_EWF_Q
static __device__ __forceinline__ void k_ewforce_body(const EwForceArgs &A, int bid) {
#pragma clang fp contract(fast)
  __shared__ double2 kl[NB_EW_KT][4];
  const int qb = bid % A.nqb, r = bid / A.nqb;
  const int cbase = qb * blockDim.x * EWF_Q + threadIdx.x;
  double ta[EWF_Q], tb[EWF_Q], tc[EWF_Q], zr[EWF_Q], zi[EWF_Q], t2[EWF_Q];
  double fx[EWF_Q], fy[EWF_Q], fz[EWF_Q];
  double wr[EWF_Q], wi[EWF_Q];  // e^{i (ka theta_a + kb theta_b)} of the last row
  double2 eb[EWF_Q];            // e^{i theta_b}: from row kb to row kb + 1
#pragma unroll
  for (int u = 0; u < EWF_Q; u++) {
    const int c = cbase + u * blockDim.x;
    double4 q4 = make_double4(0., 0., 0., 0.), st = make_double4(1., 0., 1., 0.);
    if (c < A.nq) { q4 = A.chg[c]; st = A.stp[c]; }
    ta[u] = q4.x; tb[u] = q4.y; tc[u] = q4.z; zr[u] = st.x; zi[u] = st.y; t2[u] = 2. * st.x;
    fx[u] = fy[u] = fz[u] = 0.;
    wr[u] = 1.; wi[u] = 0.;
    nb_sincos(q4.y, &eb[u].y, &eb[u].x);
  }
  const double tpi = 2. * M_PI, *hi = A.b.hi;
  const double gcx = tpi * hi[6], gcy = tpi * hi[7], gcz = tpi * hi[8];
  cptr4<EwRow> row = as_const4(A.row);
  cptr4<int> rt = as_const4(A.rtile), tr = as_const4(A.trow);
  const int t0 = rt[r], t1 = rt[r + 1], gfirst = tr[t0];
  for (int t = t0; t < t1; t++) {
    const int g0 = tr[t], g1 = tr[t + 1];
    const int e0 = row[g0].e0, ne = row[g1 - 1].e0 + (row[g1 - 1].mhi - row[g1 - 1].mlo + 1) - e0;
    __syncthreads();
    for (int u = threadIdx.x; u < ne; u += blockDim.x) {
      const int4 en = A.ent[e0 + u];
      const KData p = A.kd[en.x >= 0 ? en.x : 0], q = A.kd[en.y >= 0 ? en.y : 0];
      const double pr = en.x >= 0 ? p.cr : 0., pi = en.x >= 0 ? p.ci : 0.;
      const double mr = en.y >= 0 ? q.cr : 0., mi = en.y >= 0 ? q.ci : 0.;
      const double er = pr + mr, ei = -(pi + mi), dr = pr - mr, di = mi - pi, fm = en.z;
      kl[u][0] = make_double2(er, ei);
      kl[u][1] = make_double2(-di, dr);
      kl[u][2] = make_double2(fm * dr, -(fm * ei));
      kl[u][3] = make_double2(fm * di, fm * er);
    }
    __syncthreads();
    for (int g = g0; g < g1; g++) {
      const int ika = row[g].ka, ikb = row[g].kb, mlo = row[g].mlo, mhi = row[g].mhi;
      const double ka = ika, kb = ikb;
      if (g != gfirst && row[g].chain) {
        // the next kb of the same ka: one step of e^{i theta_b}
#pragma unroll
        for (int u = 0; u < EWF_Q; u++) {
          const double c2 = wr[u] * eb[u].x - wi[u] * eb[u].y;
          wi[u] = wi[u] * eb[u].x + wr[u] * eb[u].y;
          wr[u] = c2;
        }
      } else {
#pragma unroll
        for (int u = 0; u < EWF_Q; u++) nb_sincos(ka * ta[u] + kb * tb[u], &wi[u], &wr[u]);
      }
      // z^m from m = mlo (0 or 1 in cm3d's k lists) and z^{m-1}; then the
      // three-term recurrence z^{m+1} = 2 cos(theta_c) z^m - z^{m-1}
      double cs[EWF_Q], sn[EWF_Q], cp[EWF_Q], sp[EWF_Q];
#pragma unroll
      for (int u = 0; u < EWF_Q; u++) {
        if (mlo == 0) { cs[u] = 1.; sn[u] = 0.; cp[u] = zr[u]; sp[u] = -zi[u]; }
        else if (mlo == 1) { cs[u] = zr[u]; sn[u] = zi[u]; cp[u] = 1.; sp[u] = 0.; }
        else {
          nb_sincos(mlo * tc[u], &sn[u], &cs[u]);
          cp[u] = cs[u] * zr[u] + sn[u] * zi[u];
          sp[u] = sn[u] * zr[u] - cs[u] * zi[u];
        }
      }
      double pr[EWF_Q], pi[EWF_Q], qr[EWF_Q], qi[EWF_Q];
#pragma unroll
      for (int u = 0; u < EWF_Q; u++) { pr[u] = 0.; pi[u] = 0.; qr[u] = 0.; qi[u] = 0.; }
      const double2 (*K)[4] = kl + (row[g].e0 - e0);
      const int cnt = mhi - mlo + 1;
#pragma unroll 2
      for (int j = 0; j < cnt; j++) {
        const double2 k0 = K[j][0], k1 = K[j][1], k2 = K[j][2], k3 = K[j][3];
#pragma unroll
        for (int u = 0; u < EWF_Q; u++) {
          pr[u] = __builtin_fma(k1.x, sn[u], __builtin_fma(k0.x, cs[u], pr[u]));
          pi[u] = __builtin_fma(k1.y, sn[u], __builtin_fma(k0.y, cs[u], pi[u]));
          qr[u] = __builtin_fma(k2.y, sn[u], __builtin_fma(k2.x, cs[u], qr[u]));
          qi[u] = __builtin_fma(k3.y, sn[u], __builtin_fma(k3.x, cs[u], qi[u]));
          const double cn = __builtin_fma(t2[u], cs[u], -cp[u]);
          const double snn = __builtin_fma(t2[u], sn[u], -sp[u]);
          cp[u] = cs[u]; sp[u] = sn[u];
          cs[u] = cn; sn[u] = snn;
        }
      }
      const double ax = tpi * (ka * hi[0] + kb * hi[3]);
      const double ay = tpi * (ka * hi[1] + kb * hi[4]);
      const double az = tpi * (ka * hi[2] + kb * hi[5]);
#pragma unroll
      for (int u = 0; u < EWF_Q; u++) {
        const double s0 = wr[u] * pi[u] + wi[u] * pr[u];  // Im(w P)
        const double s1 = wr[u] * qi[u] + wi[u] * qr[u];  // Im(w Q)
        fx[u] += ax * s0 + gcx * s1;
        fy[u] += ay * s0 + gcy * s1;
        fz[u] += az * s0 + gcz * s1;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < EWF_Q; u++) {
    const int c = cbase + u * blockDim.x;
    if (c < A.nq) {
      double *o = A.fpart + (size_t)(3 * r) * A.nq + c;
      o[0] = fx[u];
      o[A.nq] = fy[u];
      o[2 * (size_t)A.nq] = fz[u];
    }
  }
}
#ifndef NB_EWF_WPE
#define NB_EWF_WPE 1
#endif
NB_KERNEL_PAIR_W(k_ewforce, EwForceArgs, 256, NB_EWF_WPE)

// k_ewfimg: the force pass's coefficient image, a thread per (tile, m,
// row) entry: the row's coefficients of cos / sin of m theta_c in Re P
// (Er, -Di), Im P (Ei, Dr), Re Q (m Dr, -m Ei) and Im Q (m Di, m Er), from
// S(k) of k = +m and -m, stored [m][coefficient][row] so that a tile is one
// block of 2 nks KB that k_ewforcem copies into LDS as it is
struct EwFimgArgs { const int4 *tent; const KData *kd; int nent; double *img; };
static __device__ __forceinline__ void k_ewfimg_body(const EwFimgArgs &A, int bid) {
  const int e = bid * 256 + threadIdx.x;
  if (e >= A.nent) return;
  const int4 en = A.tent[e];
  const KData p = A.kd[en.x >= 0 ? en.x : 0], q = A.kd[en.y >= 0 ? en.y : 0];
  const double pr = en.x >= 0 ? p.cr : 0., pi = en.x >= 0 ? p.ci : 0.;
  const double mr = en.y >= 0 ? q.cr : 0., mi = en.y >= 0 ? q.ci : 0.;
  const double er = pr + mr, ei = -(pi + mi), dr = pr - mr, di = mi - pi, fm = en.z;
  double *d = A.img + (size_t)(e >> 4) * 128 + (e & 15);  // entry e = 16 m + row
  d[0] = er; d[16] = -di; d[32] = ei; d[48] = dr;
  d[64] = fm * dr; d[80] = -(fm * ei); d[96] = fm * di; d[112] = fm * er;
}
NB_KERNEL_PAIR(k_ewfimg, EwFimgArgs, 256)

// k_ewforcem: the reciprocal force pass on the FP64 matrix cores (see
// k_ewsfacm for the v_mfma_f64_16x16x4f64 lane layout).  For a tile of 16
// k rows and 16 charges, P and Q of k_ewforce for every (row, charge) are
// four products D = A B over the K index (m, c / s), 2 values of m a step:
// A[row][(m, c)] is the row's coefficient of cos (c = 0) or sin (c = 1) of
// m theta_c in Re P, Im P, Re Q or Im Q (k_ewfimg's image of the tile,
// copied HBM -> LDS by direct loads into one of two buffers while the
// previous tile's products run); B[(m, c)][charge] is cos / sin m theta_c,
// which each lane steps in registers by the three-term recurrence in steps
// of 2 in m (z = e^{i theta_c} and z^2 start it).  The
// result gives a lane 4 rows of one run of consecutive kb for one charge,
// so w = e^{i (ka theta_a + kb theta_b)} takes one sine / cosine per run
// and steps of e^{i theta_b}; the lane sums ka Im(w P), kb Im(w P) and
// Im(w Q) over its rows for its charge, and the force A Im(w P) + Gc Im(w Q)
// (A = 2 pi (ka h^-T_a + kb h^-T_b)) follows from the 4 lane groups'
// combined sums at the end.
struct EwFmArgs {
  const EwFTile *tile; const double *img; const int *rtile;
  const double4 *chg, *stp; int nq, ncb; double *fpart; Box b;
};
// per charge tile of a lane: phases and steps, and the sums over rows of
// ka Im(w P), kb Im(w P) and Im(w Q) (the force is 2 pi (h^-T_a sa +
// h^-T_b sb) + Gc s1q)
struct EfmState {
  double ta[NB_EFM_CT], tb[NB_EFM_CT], zr[NB_EFM_CT], zi[NB_EFM_CT];
  double tz[NB_EFM_CT], z2i[NB_EFM_CT];  // 2 Re z^2, Im z^2
  double ebr[NB_EFM_CT], ebi[NB_EFM_CT];
  double sa[NB_EFM_CT], sb[NB_EFM_CT], s1q[NB_EFM_CT];
};
// a tile's fields as a lane uses them (its run's ka, kb0), as loaded: the
// uniform ones are moved to scalar registers (efm_uni) only where used, so
// that nothing waits for the load where it is issued
struct EfmDesc { int ent_off, mt0, nks, ka, kb0; };
static __device__ __forceinline__ EfmDesc efm_desc(const EwFmArgs &A, int t, int kq) {
  const EwFTile &T = A.tile[t];
  return EfmDesc{T.ent_off, T.mt0, T.nks, T.ka[kq], T.kb0[kq]};
}
static __device__ __forceinline__ EfmDesc efm_uni(const EfmDesc &d) {
  return EfmDesc{__builtin_amdgcn_readfirstlane(d.ent_off), __builtin_amdgcn_readfirstlane(d.mt0),
                 __builtin_amdgcn_readfirstlane(d.nks), d.ka, d.kb0};
}
// a tile's image (2 nks KB from entry ent_off) into an LDS buffer: 1 KB per
// wave-load, 16 B a lane, waited for by the caller
static __device__ __forceinline__ void efm_load(const EwFmArgs &A, double *dst, int ent_off,
                                                int nks) {
  const char *src = (const char *)(A.img + (size_t)ent_off * 8);
  const int lane = lane_id();
  for (int q = threadIdx.x >> 6; q < 2 * nks; q += 4)
    __builtin_amdgcn_global_load_lds(src + q * 1024 + lane * 16,
                                     (sfm_lptr)((char *)dst + q * 1024), 16, 0, 0);
}
// the products and row sums of one staged tile (m from mt0 in nks steps of
// 2; this lane's run starts at (ka, kb0))
static __device__ __forceinline__ void efm_tile(const EwFmArgs &A, const double *kl, int mt0,
                                                int nks, double ka, double kb0, int cw,
                                                EfmState &S) {
#pragma clang fp contract(fast)
  constexpr int CT = NB_EFM_CT;
  const int lane = lane_id(), kq = lane >> 4, j16 = lane & 15;
  const bool sn = kq & 1;
  // f(m) and f(m - 2) of this lane's function (cos or sin of m theta_c) for
  // m = mt0 + (kq >> 1): from 1, z, z^2 when mt0 <= 1, else two sines /
  // cosines of the charge's theta_c
  double fa[CT], fb[CT];
  const int m0 = mt0 + (kq >> 1);
#pragma unroll
  for (int ct = 0; ct < CT; ct++) {
    const double z2r = 0.5 * S.tz[ct];
    if (mt0 <= 1) {
      if (sn) {
        fa[ct] = m0 == 0 ? 0. : (m0 == 1 ? S.zi[ct] : S.z2i[ct]);
        fb[ct] = m0 == 0 ? -S.z2i[ct] : (m0 == 1 ? -S.zi[ct] : 0.);
      } else {
        fa[ct] = m0 == 0 ? 1. : (m0 == 1 ? S.zr[ct] : z2r);
        fb[ct] = m0 == 0 ? z2r : (m0 == 1 ? S.zr[ct] : 1.);
      }
    } else {
      const int c = (cw * CT + ct) * 16 + j16;
      const double th = c < A.nq ? A.chg[c].z : 0.;
      double s0, c0, s2, c2;
      nb_sincos(m0 * th, &s0, &c0);
      nb_sincos((m0 - 2) * th, &s2, &c2);
      fa[ct] = sn ? s0 : c0;
      fb[ct] = sn ? s2 : c2;
    }
  }
  // image [m][coefficient][row]: this lane's m is 2 ks + (kq >> 1), its
  // coefficient 2 p + (kq & 1) for product p, its row j16
  const double *ab = kl + ((kq >> 1) * 8 + (kq & 1)) * 16 + j16;
  sfm_d4 acc[CT][4];
  // B of step ks is f(m), stepped as f(m + 2) = 2 cos 2theta f(m) - f(m - 2)
  // (one FMA): even steps read fa and step fb, odd steps the other way
  {  // the first K step starts the sums (no zeroing)
    const double a0 = ab[0], a1 = ab[32], a2 = ab[64], a3 = ab[96];
    const sfm_d4 zero = {0., 0., 0., 0.};
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
      acc[ct][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, fa[ct], zero, 0, 0, 0);
      acc[ct][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, fa[ct], zero, 0, 0, 0);
      acc[ct][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, fa[ct], zero, 0, 0, 0);
      acc[ct][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a3, fa[ct], zero, 0, 0, 0);
      fb[ct] = S.tz[ct] * fa[ct] - fb[ct];
    }
  }
  int ks = 1;
#pragma unroll 1
  for (; ks + 1 < nks; ks += 2) {
    {
      const double *a = ab + ks * 256;
      const double a0 = a[0], a1 = a[32], a2 = a[64], a3 = a[96];
#pragma unroll
      for (int ct = 0; ct < CT; ct++) {
        acc[ct][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, fb[ct], acc[ct][0], 0, 0, 0);
        acc[ct][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, fb[ct], acc[ct][1], 0, 0, 0);
        acc[ct][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, fb[ct], acc[ct][2], 0, 0, 0);
        acc[ct][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a3, fb[ct], acc[ct][3], 0, 0, 0);
        fa[ct] = S.tz[ct] * fb[ct] - fa[ct];
      }
    }
    {
      const double *a = ab + (ks + 1) * 256;
      const double a0 = a[0], a1 = a[32], a2 = a[64], a3 = a[96];
#pragma unroll
      for (int ct = 0; ct < CT; ct++) {
        acc[ct][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, fa[ct], acc[ct][0], 0, 0, 0);
        acc[ct][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, fa[ct], acc[ct][1], 0, 0, 0);
        acc[ct][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, fa[ct], acc[ct][2], 0, 0, 0);
        acc[ct][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a3, fa[ct], acc[ct][3], 0, 0, 0);
        fb[ct] = S.tz[ct] * fa[ct] - fb[ct];
      }
    }
  }
  if (ks < nks) {
    const double *a = ab + ks * 256;
    const double a0 = a[0], a1 = a[32], a2 = a[64], a3 = a[96];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
      acc[ct][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, fb[ct], acc[ct][0], 0, 0, 0);
      acc[ct][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, fb[ct], acc[ct][1], 0, 0, 0);
      acc[ct][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, fb[ct], acc[ct][2], 0, 0, 0);
      acc[ct][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a3, fb[ct], acc[ct][3], 0, 0, 0);
    }
  }
  // this lane's rows: run kq, rows rr = 0..3 (kb = kb0 + rr), result
  // double rr of each product; w of each charge from one sine / cosine,
  // then steps of e^{i theta_b}
  double wr[CT], wi[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ct++) nb_sincos(ka * S.ta[ct] + kb0 * S.tb[ct], &wi[ct], &wr[ct]);
#pragma unroll
  for (int ct = 0; ct < CT; ct++) {
    double p0 = 0., pb = 0.;  // sum of Im(w P), of kb Im(w P) over the run
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const double s0 = wr[ct] * acc[ct][1][rr] + wi[ct] * acc[ct][0][rr];  // Im(w P)
      const double s1 = wr[ct] * acc[ct][3][rr] + wi[ct] * acc[ct][2][rr];  // Im(w Q)
      p0 += s0;
      pb += (kb0 + rr) * s0;
      S.s1q[ct] += s1;
      if (rr < 3) {
        const double w2 = wr[ct] * S.ebr[ct] - wi[ct] * S.ebi[ct];
        wi[ct] = wr[ct] * S.ebi[ct] + wi[ct] * S.ebr[ct];
        wr[ct] = w2;
      }
    }
    S.sa[ct] += ka * p0;
    S.sb[ct] += pb;
  }
}
static __device__ __forceinline__ void k_ewforcem_body(const EwFmArgs &A, int bid) {
#pragma clang fp contract(fast)
  constexpr int CT = NB_EFM_CT;
  __shared__ double efm_b0[NB_EFM_MS * 128], efm_b1[NB_EFM_MS * 128];
  const int r = bid / A.ncb, cb = bid % A.ncb;
  const int lane = lane_id(), kq = lane >> 4, j16 = lane & 15, w = threadIdx.x >> 6;
  const int cw = cb * 4 + w;  // this wave's charge tiles: cw * CT .. + CT - 1
  // a lane's B operand is cos (kq even) or sin (kq odd) of m theta_c
  EfmState S;
#pragma unroll
  for (int ct = 0; ct < CT; ct++) {
    const int c = (cw * CT + ct) * 16 + j16;
    double4 q4 = make_double4(0., 0., 0., 0.), st = make_double4(1., 0., 1., 0.);
    if (c < A.nq) { q4 = A.chg[c]; st = A.stp[c]; }
    S.ta[ct] = q4.x; S.tb[ct] = q4.y;
    S.zr[ct] = st.x; S.zi[ct] = st.y;
    S.tz[ct] = 2. * (st.x * st.x - st.y * st.y);
    S.z2i[ct] = 2. * st.x * st.y;
    nb_sincos(q4.y, &S.ebi[ct], &S.ebr[ct]);
    S.sa[ct] = S.sb[ct] = S.s1q[ct] = 0.;
  }
  // a tile's fields are read one tile ahead, so that the next tile's copy
  // is issued at once and the products wait for no load
  const int t0 = A.rtile[r], t1 = A.rtile[r + 1];
  EfmDesc cur = {0, 0, 0, 0, 0}, nxt = {0, 0, 0, 0, 0};
  if (t0 < t1) {
    cur = efm_uni(efm_desc(A, t0, kq));
    efm_load(A, efm_b0, cur.ent_off, cur.nks);
  }
  if (t0 + 1 < t1) nxt = efm_desc(A, t0 + 1, kq);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  // two tiles a trip so that each buffer is one LDS object at every access
  for (int t = t0; t < t1; t += 2) {
    {
      cur = efm_uni(cur);
      if (t + 1 < t1) {
        const EfmDesc u = efm_uni(nxt);
        efm_load(A, efm_b1, u.ent_off, u.nks);
      }
      EfmDesc nn = {0, 0, 0, 0, 0};
      if (t + 2 < t1) nn = efm_desc(A, t + 2, kq);
      efm_tile(A, efm_b0, cur.mt0, cur.nks, cur.ka, cur.kb0, cw, S);
      cur = nxt;
      nxt = nn;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t + 1 >= t1) break;
    {
      cur = efm_uni(cur);
      if (t + 2 < t1) {
        const EfmDesc u = efm_uni(nxt);
        efm_load(A, efm_b0, u.ent_off, u.nks);
      }
      EfmDesc nn = {0, 0, 0, 0, 0};
      if (t + 3 < t1) nn = efm_desc(A, t + 3, kq);
      efm_tile(A, efm_b1, cur.mt0, cur.nks, cur.ka, cur.kb0, cw, S);
      cur = nxt;
      nxt = nn;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // the 4 lane groups hold the same 16 charges' sums over different rows
  const double tpi = 2. * M_PI, *hi = A.b.hi;
  const double gcx = tpi * hi[6], gcy = tpi * hi[7], gcz = tpi * hi[8];
#pragma unroll
  for (int ct = 0; ct < CT; ct++) {
    double sa = S.sa[ct], sb = S.sb[ct], s1q = S.s1q[ct];
    sa += __shfl_xor(sa, 16); sb += __shfl_xor(sb, 16); s1q += __shfl_xor(s1q, 16);
    sa += __shfl_xor(sa, 32); sb += __shfl_xor(sb, 32); s1q += __shfl_xor(s1q, 32);
    const int c = (cw * CT + ct) * 16 + j16;
    if (kq == 0 && c < A.nq) {
      double *o = A.fpart + (size_t)(3 * r) * A.nq + c;
      o[0] = tpi * (hi[0] * sa + hi[3] * sb) + gcx * s1q;
      o[A.nq] = tpi * (hi[1] * sa + hi[4] * sb) + gcy * s1q;
      o[2 * (size_t)A.nq] = tpi * (hi[2] * sa + hi[5] * sb) + gcz * s1q;
    }
  }
}
NB_KERNEL_PAIR_W(k_ewforcem, EwFmArgs, 256, 2)

// k_ewfin, a workgroup per 64 charges (lane = charge): the four waves sum
// interleaved quarters of the range partials (ranges w, w+4, ...), combined
// in LDS in a fixed order; wave 0 then does ecorr (md_ewald.c:385-453) over
// the atom's excluded charged partners (both directions, so energies and
// virials are halved later), on the caller's unwrapped positions with no
// minimum image, and adds it and q times the reciprocal force to fl.  The
// ecorr energy / virial partials go to a column-major table, one row per
// workgroup.
struct EwFinArgs {
  const double *fpart; const double4 *chg; const int *qidx; int nq, n, nrange, ew, ec;
  const int *eoff, *elist; const double *pos, *q; double kappa; double *fl;
  int nblocks; double *epart;
};
static __device__ __forceinline__ void k_ewfin_body(const EwFinArgs &A, int bid) {
  __shared__ double rs[3][3][NB_WAVE];
  double acc[NB_NE_ECORR];
#pragma unroll
  for (int k = 0; k < NB_NE_ECORR; k++) acc[k] = 0.;
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const int c = bid * NB_WAVE + lane;
  double fx = 0., fy = 0., fz = 0.;
  if (A.ew && c < A.nq) {
    const size_t nq = A.nq;
    for (int r = w; r < A.nrange; r += 4) {
      const double *p = A.fpart + 3 * (size_t)r * nq + c;
      fx += p[0];
      fy += p[nq];
      fz += p[2 * nq];
    }
  }
  if (w > 0) { rs[w - 1][0][lane] = fx; rs[w - 1][1][lane] = fy; rs[w - 1][2][lane] = fz; }
  __syncthreads();
  if (w == 0 && c < A.nq) {
    const int a = A.qidx[c], n = A.n;
    double gx = A.fl[a], gy = A.fl[n + a], gz = A.fl[2 * n + a];
    if (A.ec && A.eoff[a + 1] > A.eoff[a]) {
      const double spi = 1. / sqrt(M_PI), kappa = A.kappa, ka2 = kappa * kappa;
      const double *pos = A.pos, *q = A.q;
      double ex = 0., ey = 0., ez = 0.;
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
        ex += dx * dv; ey += dy * dv; ez += dz * dv;
      }
      gx += ex; gy += ey; gz += ez;
    }
    if (A.ew) {
      fx = (fx + rs[0][0][lane]) + (rs[1][0][lane] + rs[2][0][lane]);
      fy = (fy + rs[0][1][lane]) + (rs[1][1][lane] + rs[2][1][lane]);
      fz = (fz + rs[0][2][lane]) + (rs[1][2][lane] + rs[2][2][lane]);
      const double qc = A.chg[c].w;
      gx += qc * fx; gy += qc * fy; gz += qc * fz;
    }
    A.fl[a] = gx; A.fl[n + a] = gy; A.fl[2 * n + a] = gz;
  }
  if (A.ec) block_reduce_store<NB_NE_ECORR>(acc, A.epart + bid, A.nblocks);
}
NB_KERNEL_PAIR(k_ewfin, EwFinArgs, 256)

static __device__ __forceinline__ void k_red_body(const RedSet &R, int bid) { reduce_col(R, bid); }
NB_KERNEL_PAIR(k_red, RedSet, 256)

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
  CellArgs ca = {c->d_pos, n, blocks(n, 256), c->ncells, b, c->ncx, c->ncy, c->ncz, c->d_cell,
                 c->d_cwork, c->d_ccount, c->d_cstart, c->d_cfill, c->d_flags, cond};
  NB_GO(c, K_CELL, k_cell, blocks(n, 256), 256, 0, ca, s);
  ScatArgs sc = {c->d_cell, n, c->d_cfill, c->d_tmp, cond};
  NB_GO(c, K_SCAT, k_scat, blocks(n, 256), 256, 0, sc, s);
  // the sorted cell id of each sorted slot overwrites d_cell
  OrderArgs oa = {c->d_ccount, c->d_cstart, c->d_tmp, c->d_perm, c->d_iperm, c->d_cell,
                  c->d_pos, c->d_qch, c->d_itype, (double4 *)c->d_rec64, c->d_stype,
                  c->d_posb, n, c->d_flags + 6,
                  c->prec == NB_PREC_MIXED ? (uint4 *)c->d_recf : nullptr, b, cond,
                  c->ncells, c->ncells >= 64};
  NB_GO(c, K_ORDER, k_order, oa.packed ? (c->ncells + 15) / 16 : c->ncells, 256, 0, oa, s);
  ListArgs A;
  A.n = n; A.ntypes = c->ntypes; A.cap = c->cap;
  A.ncx = c->ncx; A.ncy = c->ncy; A.ncz = c->ncz; A.rx = c->crx; A.ry = c->cry; A.rz = c->crz;
  A.pcut_lds = c->ntypes * c->ntypes <= NB_PCUT_LDS;
  // column pruning needs cells that are boxes in Cartesian space
  A.prune = b.ortho && c->ncx > 1 && c->ncy > 1 && c->ncz > 1;
  A.rprune2 = c->rprune2;
  A.wcx = b.h[0] / c->ncx; A.wcy = b.h[4] / c->ncy; A.wcz = b.h[8] / c->ncz;
  A.rec = (const double4 *)c->d_rec64; A.stype = c->d_stype; A.perm = c->d_perm;
  A.iperm = c->d_iperm;
  A.scell = c->d_cell; A.cstart = c->d_cstart; A.ccount = c->d_ccount;
  A.xoff = c->d_xoff; A.xlist = c->d_xlist; A.pcut = (const double2 *)c->d_pcut;
  A.rmaxs_skin2_e = c->rmaxs_skin2_e; A.rmaxl_skin2_e = c->rmaxl_skin2_e;
  A.b = b; A.nbr = c->d_nbr; A.nnb = c->d_nnb; A.flags = c->d_flags;
  A.ibits = c->ibits; A.etype = c->etype;
  A.hcnt = (int4 *)c->d_hcnt; A.cond = cond;
  NB_GO(c, K_LIST, k_list, blocks(n, 4), 256, 0, A, s);
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

// The energy-column reductions of force(...,1): which tables, how many rows.
RedSet nb_red_set(const nb_ctx *c, int iflg, bool pair, bool quirk) {
  RedSet R;
  memset(&R, 0, sizeof(R));
  const bool ec = iflg == 1 && c->iperd == 3 && c->necorr > 0;
  const bool ew = iflg == 1 && c->iperd == 3 && c->ktot > 0 && c->ncharge > 0;
  // no pair energy pass ran: its columns reduce over no rows, i.e. to zero
  R.part[0] = c->d_epair; R.rows[0] = pair ? c->npair_blocks : 0;
  R.cols[0] = quirk ? NB_NE_PAIRQ : NB_NE_PAIR;
  R.out[0] = c->d_eout;
  R.part[1] = c->d_eec; R.rows[1] = c->nfin_blocks; R.cols[1] = ec ? NB_NE_ECORR : 0;
  R.out[1] = c->d_eout + NB_EOUT_ECORR;
  R.part[2] = c->d_eew; R.rows[2] = c->new_blocks; R.cols[2] = ew ? NB_NE_EW : 0;
  R.out[2] = c->d_eout + NB_EOUT_EW;
  R.ncols = R.cols[0] + R.cols[1] + R.cols[2];
  return R;
}

// force(...,1) / force(...,0).  flags: NB_FORCE_PACKED -- the sorted records
// and the per-charge Ewald data are already current (the MD step writes
// them); NB_FORCE_NORED -- leave the energy columns unreduced (the MD kick
// reduces them); NB_FORCE_NOENER -- forces only; NB_FORCE_EONLY -- energies
// and virials only.  Pair energies and virials are always evaluated in FP64
// (k_pair64_1f / 1t / 1e): a mixed context computes its forces with the FP32
// kernel and, when energies are asked for, runs the FP64 energies-only
// kernel beside it.
int nb_launch_force(nb_ctx *c, int iflg, const Box &b, double *d_fs, double *d_fl,
                    hipStream_t s, int flags) {
  const int n = c->natoms;
  const bool packed = flags & NB_FORCE_PACKED;
  const bool eonly = flags & NB_FORCE_EONLY, noener = (flags & NB_FORCE_NOENER) && !eonly;
  const bool mixed = c->prec == NB_PREC_MIXED;
  const bool reuse = eonly && (flags & NB_FORCE_REUSE_RECIP);
  c->recip_owner = nullptr;  // the caller sets it again when the columns are an MD state's
  nb_time_mark(c, 0, s);
  if (flags & NB_FORCE_NO_PAIR) {
    nb_time_mark(c, 1, s);
  } else {
    if (mixed && !eonly) {
      nb_time_mark(c, 1, s);
      int rc = nb_launch_pair_mixed(c, iflg, b, d_fs, d_fl, s, packed);
      if (rc) return rc;
    }
    if (!mixed || (iflg == 1 && !noener)) {
      // a mixed context's step keeps only the FP32 records current
      Pack64Args pa = {c->d_pos, c->d_qch, c->d_perm, n, (double4 *)c->d_rec64};
      NB_GO(c, K_PACK64, k_pack64, packed && !mixed ? 0 : blocks(n, 256), 256, 0, pa, s);
      if (!mixed) nb_time_mark(c, 1, s);
      PairArgs P;
      P.n = n; P.ntypes = c->ntypes; P.ninter = c->ninter; P.ntable = c->ntable;
      P.cap = c->cap; P.tab_lds = c->tab_lds_bytes > 0; P.ctab_off = c->ctab_off;
      P.nblocks = c->npair_blocks;
      P.nbr = c->d_nbr; P.nnb = c->d_nnb; P.stype = c->d_stype; P.perm = c->d_perm;
      P.map = c->d_map; P.rec = (const double4 *)c->d_rec64; P.ip = c->d_ip;
      P.vtab3 = c->d_vtab3; P.ctab3 = c->d_ctab3;
      P.fcoef = c->d_fcoef; P.fc_lds = c->fcoef_lds;
      P.imask = c->imask; P.ibits = c->ibits; P.etype = c->etype;
      P.rmins_coul2 = c->rmins_coul2; P.dx2coul = c->dx2coul; P.rmaxl_coul2 = c->rmaxl_coul2;
      P.clast = c->clast;
      P.b = b; P.fl = d_fl; P.fs = d_fs; P.epart = c->d_epair;
      const int g = c->npair_blocks, lds = c->tab_lds_bytes;
      if (eonly && (flags & NB_FORCE_CM3D_EONLY))
        NB_GO(c, K_PAIR64_1Q, k_pair64_1q, g, 256, lds, P, s);
      else if (mixed || eonly)
        NB_GO(c, K_PAIR64_1E, k_pair64_1e, g, 256, lds, P, s);
      else if (iflg == 0)
        NB_GO(c, K_PAIR64_0T, k_pair64_0t, g, 256, lds, P, s);
      else if (d_fs)
        NB_GO(c, K_PAIR64_1T, k_pair64_1t, g, 256, lds, P, s);
      else if (noener)
        NB_GO(c, K_PAIR64_1N, k_pair64_1n, g, 256, c->fcoef_lds, P, s);
      else
        NB_GO(c, K_PAIR64_1F, k_pair64_1f, g, 256, lds, P, s);
      if (!c->rec) LCHK(hipGetLastError());
    }
  }
  nb_time_mark(c, 2, s);
  const bool ec = iflg == 1 && c->iperd == 3 && c->necorr > 0;
  const bool ew = iflg == 1 && c->iperd == 3 && c->ktot > 0 && c->ncharge > 0;
  const bool ewf = ew && !(flags & NB_FORCE_EONLY);  // reciprocal-space forces
  const bool rp = !(flags & NB_FORCE_NO_RECIP) && !reuse, fin = !(flags & NB_FORCE_NO_FIN);
  const int nq = c->ncharge;
  EwChgArgs ch = {c->d_pos, c->d_qch, c->d_qidx, nq, n, b, (double4 *)c->d_chg,
                  (double4 *)c->d_stp};
  NB_GO(c, K_EWCHG, k_ewchg, rp && ew && !packed ? blocks(nq, 256) : 0, 256, 0, ch, s);
  nb_time_mark(c, 3, s);
  EwSfacArgs sf = {c->d_sym, (const double4 *)c->d_chg, (const double4 *)c->d_stp, nq, c->qc,
                   c->nsym, c->nsw, c->ktot, (double2 *)c->d_spart};
  // the structure-factor kernels are always recorded, the unused ones with
  // no workgroups, so that systems of any lane kind record the same step (a
  // batch launches each kernel once for the systems that give it work)
  const int nsfac = rp && ew && !c->sfm ? c->nsw * c->nchunk : 0;
  NB_GO(c, K_EWSFAC, k_ewsfac, c->sfac_rows == 1 ? nsfac : 0, 256, 0, sf, s);
  NB_GO(c, K_EWSFACR, k_ewsfacr, c->sfac_rows > 1 ? nsfac : 0, 256, 0, sf, s);
  const int nqp = NB_SFM_NQ * ((nq + NB_SFM_NQ - 1) / NB_SFM_NQ);
  EwSfmTabArgs st = {(const double4 *)c->d_chg, nq, nqp, c->sfm_la, c->sfm_lb, c->sfm_lz,
                     c->d_sfmtab};
  const long long nte = (long long)nqp * SFM_NE_MAX;
  NB_GO(c, K_EWSFMTAB, k_ewsfmtab, rp && ew && c->sfm ? (int)((nte + 255) / 256) : 0, 256, 0,
        st, s);
  EwSfmArgs sm = {c->d_rowm, c->d_sfmw, c->d_sfmtab, nq, c->qc, c->sfm_nrwg, c->nchunk,
                  c->sfm_la, c->sfm_lb, c->sfm_lz, c->ktot, (double2 *)c->d_spart};
  NB_GO(c, K_EWSFACM, k_ewsfacm, rp && ew && c->sfm ? 8 * c->sfm_nrwg * ((c->nchunk + 7) / 8) : 0,
        256, 0, sm, s);
  nb_time_mark(c, 4, s);
  double det = b.h[0] * (b.h[4] * b.h[8] - b.h[5] * b.h[7]) +
               b.h[1] * (b.h[5] * b.h[6] - b.h[3] * b.h[8]) +
               b.h[2] * (b.h[3] * b.h[7] - b.h[4] * b.h[6]);
  double falp2 = c->kappa > 0 ? 1. / (4. * c->kappa * c->kappa) : 0.;
  // one workgroup per 64 k
  EwKredArgs kr = {c->d_kabc, c->ktot, (const double2 *)c->d_spart, c->nchunk, c->new_blocks,
                   b, 1. / det, falp2, (KData *)c->d_kd, c->d_eew};
  NB_GO(c, K_EWKRED, k_ewkred, rp && ew ? c->new_blocks : 0, 256, 0, kr, s);
  // the matrix-core force pass's coefficient image (from the k data just
  // reduced) is timed with the k reduction, so that the force phase is
  // k_ewforcem's own
  EwFimgArgs fi = {c->d_ftent, (const KData *)c->d_kd, c->efm_nent, c->d_fimg};
  NB_GO(c, K_EWFIMG, k_ewfimg, rp && ewf && c->efm ? blocks(c->efm_nent, 256) : 0, 256, 0, fi, s);
  nb_time_mark(c, 5, s);
  EwForceArgs ef = {c->d_row, c->d_ent, c->d_rtile, c->d_trow, (const KData *)c->d_kd,
                    (const double4 *)c->d_chg, (const double4 *)c->d_stp, nq, c->nqb, c->d_fpart,
                    b};
  NB_GO(c, K_EWFORCE, k_ewforce, rp && ewf && !c->efm ? c->nqb * c->nrange : 0, 256, 0, ef, s);
  EwFmArgs em = {c->d_ftile, c->d_fimg, c->d_frtile, (const double4 *)c->d_chg,
                 (const double4 *)c->d_stp, nq, c->efm_ncb, c->d_fpart, b};
  NB_GO(c, K_EWFORCEM, k_ewforcem, rp && ewf && c->efm ? c->efm_ncb * c->nrange : 0, 256, 0, em,
        s);
  nb_time_mark(c, 6, s);
  EwFinArgs fa = {c->d_fpart, (const double4 *)c->d_chg, c->d_qidx, nq, n, c->nrange, ewf, ec,
                  c->d_eoff, c->d_elist, c->d_pos, c->d_qch, c->kappa, d_fl, c->nfin_blocks,
                  c->d_eec};
  NB_GO(c, K_EWFIN, k_ewfin, fin && !reuse && (ewf || ec) ? c->nfin_blocks : 0, 256, 0, fa, s);
  if (fin && !(flags & NB_FORCE_NORED)) {
    // the pair energy columns were written by an FP64 pair kernel with
    // energies (0t counts clamps; a mixed context's only one is 1e)
    const bool pcols = !(flags & NB_FORCE_NO_PAIR) &&
                       (mixed ? iflg == 1 && !noener : iflg == 0 || !noener || d_fs);
    RedSet R = nb_red_set(c, iflg, pcols, eonly && (flags & NB_FORCE_CM3D_EONLY));
    NB_GO(c, K_RED, k_red, R.ncols, 256, 0, R, s);
  }
  if (!c->rec) LCHK(hipGetLastError());
  nb_time_mark(c, 7, s);
  if (c->timing && !c->rec) c->trec++;
  return NB_OK;
}

size_t nb_kdata_bytes() { return sizeof(KData); }

static struct RegisterKernels {
  RegisterKernels() {
    g_batch_fn[K_ZERO] = k_zero_bl;
    g_batch_fn[K_CELL] = k_cell_bl;
    g_batch_fn[K_SCAT] = k_scat_bl;
    g_batch_fn[K_ORDER] = k_order_bl;
    g_batch_fn[K_PACK64] = k_pack64_bl;
    g_batch_fn[K_LIST] = k_list_bl;
    g_batch_fn[K_CHECK] = k_check_bl;
    g_batch_fn[K_PAIR64_0T] = k_pair64_0t_bl;
    g_batch_fn[K_PAIR64_1T] = k_pair64_1t_bl;
    g_batch_fn[K_PAIR64_1F] = k_pair64_1f_bl;
    g_batch_fn[K_PAIR64_1N] = k_pair64_1n_bl;
    g_batch_fn[K_PAIR64_1E] = k_pair64_1e_bl;
    g_batch_fn[K_PAIR64_1Q] = k_pair64_1q_bl;
    g_batch_fn[K_RED] = k_red_bl;
    g_batch_fn[K_EWCHG] = k_ewchg_bl;
    g_batch_fn[K_EWSFAC] = k_ewsfac_bl;
    g_batch_fn[K_EWKRED] = k_ewkred_bl;
    g_batch_fn[K_EWFORCE] = k_ewforce_bl;
    g_batch_fn[K_EWSFACR] = k_ewsfacr_bl;
    g_batch_fn[K_EWSFACM] = k_ewsfacm_bl;
    g_batch_fn[K_EWSFMTAB] = k_ewsfmtab_bl;
    g_batch_fn[K_EWFORCEM] = k_ewforcem_bl;
    g_batch_fn[K_EWFIMG] = k_ewfimg_bl;
    g_batch_fn[K_EWFIN] = k_ewfin_bl;
  }
} g_register_kernels;
