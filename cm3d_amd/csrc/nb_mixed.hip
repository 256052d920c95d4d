// nb_mixed.hip -- the mixed-precision pair kernel (NB_PREC_MIXED).
//
// Same lists, tables and semantics as k_pair64 (force_npol md_fnopol.c:46,
// force_e finter_e.c:42), different arithmetic:
//   - positions are packed per evaluation as 32-bit fixed-point fractional
//     coordinates u = frac(h^-1 r) * 2^32; the displacement of a pair is the
//     wrapping int32 difference, which *is* the minimum image in fractional
//     space, converted to FP32 and mapped back by h (FP32).  Its relative
//     precision is 2^-24 of the pair distance whatever the box size.
//   - table interpolation, forces and per-pair energies in FP32 from FP32
//     copies of the tables (LDS-resident);
//   - each lane sums its few pair forces for an atom in FP32, the lane sums
//     are converted to 64-bit fixed point (2^-24 K/A resolution) and reduced
//     across the wavefront in integer arithmetic, so the per-atom force does
//     not depend on the order in which the lanes' partial sums meet;
//   - no energies or virials: a mixed context's pair energies and virials
//     are evaluated in FP64 by the energy pass of nb_launch_force when they
//     are asked for (the MD step never needs them), so the FP32 kernel is
//     forces only.
#include "nb_device.h"

#define NB_FX_SCALE 16777216.0  // 2^24 fixed-point units per K/A


static __device__ __forceinline__ long long wave_sum_ll(long long v) { return nb_wave_sum(v); }

// Sorted-order records for the mixed kernel (nb_mixed_record).
struct PackMixArgs { const double *pos, *q; const int *perm; int n; Box b; uint4 *rec; };
static __device__ __forceinline__ void k_packmix_body(const PackMixArgs &A, int bid) {
  int s = bid * blockDim.x + threadIdx.x;
  if (s >= A.n) return;
  const int n = A.n, a = A.perm[s];
  A.rec[s] = nb_mixed_record(A.pos[a], A.pos[n + a], A.pos[2 * n + a], A.q[a], A.b.hi);
}
NB_KERNEL_PAIR(k_packmix, PackMixArgs, 256)

#define NB_TPF_LDS 256  // type pairs whose FP32 parameters are staged in LDS

struct MixArgs {
  int n, ntypes, ninter, ntable, cap, ctab_off, tab_lds, nblocks;
  int g16;  // 16 lanes per atom (four atoms a wavefront) instead of 64
  const uint32_t *nbr;
  const int *nnb, *stype, *perm;
  const uint4 *rec;
  const TPairF *tpf;
  const float *vtab3, *ctab3;
  float h[9];
  float rmins_coul2, dx2coul, rmaxl_coul2;
  double *fl, *fs;
  // forces-only kernel (k_mix_1f)
  const float4 *fcf;  // force-column coefficients {c0, c1, c2, 0}: vdW tables, then coulomb
  int fc_lds;         // their LDS bytes (0: read from HBM)
  int fc_coff;        // node offset of the coulomb table
  int ortho;          // h is diagonal
  float hs[9];        // h * 2^-32: maps a fixed-point difference to a displacement
  uint32_t imask;     // list entry: sorted index bits
  int ibits, etype;   // ... and the partner's type above them (when etype)
};

static __device__ __forceinline__ float interp3f(const float *t, int k, int c, float p) {
#pragma clang fp contract(fast)
  float fm1 = t[3 * (k - 1) + c], f0 = t[3 * k + c], f1 = t[3 * (k + 1) + c];
  return f0 + .5f * p * (f1 - fm1 + p * (fm1 - 2.f * f0 + f1));
}

static __device__ __forceinline__ long long to_fix(double v) {
  return __double2ll_rn(v * NB_FX_SCALE);
}

// G lanes per sorted atom (G = 64: one atom per wavefront; G = 16: four
// atoms per wavefront, one per DPP row), each lane taking every G-th list
// entry; a wave takes several atoms (groups of 64 / G) in turn.  A lane's
// pair forces for the atom are summed in FP32 (a handful of terms), then go
// to fixed point and are reduced across the G lanes in integers, so the sum
// does not depend on the order.  List entries are loaded two chunks ahead
// of the evaluation and partner records one chunk ahead.  G = 16 suits
// short lists (dense LJ-like mixtures, ~100 entries): the last, partly
// filled chunk wastes fewer lanes and one reduction serves four atoms.
// TLDS: tables (vt / ct) and type-pair parameters (stp) are all in LDS, so
// their reads compile to LDS reads, not generic (flat) ones.
template <int G>
static __device__ __forceinline__ long long group_sum_ll(long long v) {
  if constexpr (G == NB_WAVE) return wave_sum_ll(v);
  else return nb_row_sum(v);
}
template <int IFLG, bool WANT_FS, bool TLDS, int G>
static __device__ __forceinline__ void pair_mixed_run(const MixArgs &A, int bid, const float *vt,
                                                      const float *ct, const TPairF *stp,
                                                      bool tp_lds) {
#pragma clang fp contract(fast)
  static_assert(G == 16 || G == NB_WAVE, "16 or 64 lanes per atom");
  constexpr int APW = NB_WAVE / G;  // atoms a wavefront evaluates together
  const int lane = lane_id(), gl = lane % G, sub = lane / G;
  const int nw = A.nblocks * 4;
  constexpr unsigned fv = IFLG ? NB_FLAG_VT : NB_FLAG_VS;
  constexpr unsigned fe = IFLG ? NB_FLAG_ET : NB_FLAG_ES;
  const float s32 = 2.3283064365386963e-10f;  // 2^-32
  for (int s0 = (bid * 4 + (threadIdx.x >> 6)) * APW; s0 < A.n; s0 += nw * APW) {
    const int s = s0 + sub;
    const bool live = APW == 1 || s < A.n;
    const int sa = live ? s : s0;  // a dead group reads its wave's first atom
    const uint4 ri = A.rec[sa];
    const float qi = __uint_as_float(ri.w);
    const int ti = A.stype[sa];
    const int nb = live ? A.nnb[sa] : 0;
    const uint32_t *lst = A.nbr + (size_t)sa * A.cap;
    const TPairF *tprow = (TLDS ? stp : (tp_lds ? stp : A.tpf)) + ti * A.ntypes;
    float fx = 0.f, fy = 0.f, fz = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
    // list entries run two chunks ahead of the evaluation and partner
    // records one chunk ahead, so a record gather never waits on the list
    // load it depends on
    uint32_t ec = gl < nb ? lst[gl] : 0u;
    uint32_t en = G + gl < nb ? lst[G + gl] : 0u;
    uint4 rjc = make_uint4(0u, 0u, 0u, 0u);
    int tjc = 0;
    if ((ec >> 28) & (fv | fe)) {
      const int j = (int)(ec & A.imask);
      rjc = A.rec[j];
      tjc = A.stype[j];
    }
    // trip bound: the longest list of the wave's atoms, read from lane 0 of
    // each group, so the compiler sees a wave-uniform loop
    int nbw = nb;
    if constexpr (APW == 4)
      nbw = max(max(__builtin_amdgcn_readlane(nb, 0), __builtin_amdgcn_readlane(nb, 16)),
                max(__builtin_amdgcn_readlane(nb, 32), __builtin_amdgcn_readlane(nb, 48)));
    for (int k0 = 0; k0 < nbw; k0 += G) {
      const int k2 = k0 + 2 * G + gl;
      const uint32_t e2 = k2 < nb ? lst[k2] : 0u;
      uint4 rjn = make_uint4(0u, 0u, 0u, 0u);
      int tjn = 0;
      if ((en >> 28) & (fv | fe)) {
        const int j = (int)(en & A.imask);
        rjn = A.rec[j];
        tjn = A.stype[j];
      }
      const unsigned fl = ec >> 28;
      if (fl & (fv | fe)) {
        const uint4 rj = rjc;
        float ux = (float)(int)(ri.x - rj.x) * s32;
        float uy = (float)(int)(ri.y - rj.y) * s32;
        float uz = (float)(int)(ri.z - rj.z) * s32;
        float dx = ux * A.h[0] + uy * A.h[1] + uz * A.h[2];
        float dy = ux * A.h[3] + uy * A.h[4] + uz * A.h[5];
        float dz = ux * A.h[6] + uy * A.h[7] + uz * A.h[8];
        float d2 = dx * dx + dy * dy + dz * dz;
        if (fl & fv) {
          const TPairF P = tprow[tjc];
          const float rmax2 = IFLG ? P.rmaxl2 : P.rmaxs2;
          if (d2 < rmax2) {
            float p = (d2 - P.rmins2) * P.dx2;
            int kk = (int)p;
            if (kk < 1) { kk = 1; p = 1.f; }
            if (kk > A.ntable - 3) kk = A.ntable - 3;
            p = p - (float)kk;
            const float *T = vt + (size_t)P.in * A.ntable * 3;
            float f = interp3f(T, kk, 0, p);
            float swv = 1.f;
            if ((WANT_FS || !IFLG) && d2 > P.rminl2) swv = interp3f(T, kk, 1, p);
            if (IFLG) {
              float gx = dx * f, gy = dy * f, gz = dz * f;
              fx -= gx; fy -= gy; fz -= gz;
              if (WANT_FS) { sx -= gx * swv; sy -= gy * swv; sz -= gz * swv; }
            } else {
              f *= swv;
              sx -= dx * f; sy -= dy * f; sz -= dz * f;
            }
          }
        }
        if ((fl & fe) && d2 < A.rmaxl_coul2) {
          float p = (d2 - A.rmins_coul2) * A.dx2coul;
          int kk = (int)p;
          if (kk < 1) { kk = 1; p = 1.f; }
          if (kk > A.ntable - 3) kk = A.ntable - 3;
          p = p - (float)kk;
          float qij = qi * __uint_as_float(rj.w);
          float fcc = qij * interp3f(ct, kk, 0, p);
          float gx = dx * fcc, gy = dy * fcc, gz = dz * fcc;
          if (IFLG) { fx -= gx; fy -= gy; fz -= gz; }
          if (WANT_FS || !IFLG) { sx -= gx; sy -= gy; sz -= gz; }
        }
      }
      ec = en;
      en = e2;
      rjc = rjn;
      tjc = tjn;
    }
    const int ai = A.perm[sa];
    const int n = A.n;
    const bool wr = gl == 0 && live;
    if (IFLG) {
      long long X = group_sum_ll<G>(to_fix(fx)), Y = group_sum_ll<G>(to_fix(fy));
      long long Z = group_sum_ll<G>(to_fix(fz));
      if (wr) {
        A.fl[ai] = (double)X / NB_FX_SCALE;
        A.fl[n + ai] = (double)Y / NB_FX_SCALE;
        A.fl[2 * n + ai] = (double)Z / NB_FX_SCALE;
      }
    }
    if (WANT_FS || !IFLG) {
      long long X = group_sum_ll<G>(to_fix(sx)), Y = group_sum_ll<G>(to_fix(sy));
      long long Z = group_sum_ll<G>(to_fix(sz));
      if (wr) {
        A.fs[ai] = (double)X / NB_FX_SCALE;
        A.fs[n + ai] = (double)Y / NB_FX_SCALE;
        A.fs[2 * n + ai] = (double)Z / NB_FX_SCALE;
      }
    }
  }
}

template <int IFLG, bool WANT_FS, int G>
static __device__ __forceinline__ void pair_mixed_tabs(const MixArgs &A, int bid, const float *tabsf,
                                                       const TPairF *stp, bool tp_lds) {
  if (A.tab_lds && tp_lds)
    pair_mixed_run<IFLG, WANT_FS, true, G>(A, bid, tabsf, tabsf + A.ctab_off, stp, tp_lds);
  else if (A.tab_lds)
    pair_mixed_run<IFLG, WANT_FS, false, G>(A, bid, tabsf, tabsf + A.ctab_off, stp, tp_lds);
  else
    pair_mixed_run<IFLG, WANT_FS, false, G>(A, bid, A.vtab3, A.ctab3, stp, tp_lds);
}

template <int IFLG, bool WANT_FS>
static __device__ __forceinline__ void pair_mixed_body(const MixArgs &A, int bid) {
  extern __shared__ float tabsf[];
  __shared__ TPairF stp[NB_TPF_LDS];
  if (A.tab_lds) {
    int nv = A.ninter * A.ntable * 3, nc = A.ntable * 3;
    for (int t = threadIdx.x; t < nv; t += blockDim.x) tabsf[t] = A.vtab3[t];
    for (int t = threadIdx.x; t < nc; t += blockDim.x) tabsf[A.ctab_off + t] = A.ctab3[t];
  }
  const int nt2 = A.ntypes * A.ntypes;
  const bool tp_lds = nt2 <= NB_TPF_LDS;
  if (tp_lds)
    for (int t = threadIdx.x; t < nt2; t += blockDim.x) stp[t] = A.tpf[t];
  __syncthreads();
  if (A.g16)
    pair_mixed_tabs<IFLG, WANT_FS, 16>(A, bid, tabsf, stp, tp_lds);
  else
    pair_mixed_tabs<IFLG, WANT_FS, NB_WAVE>(A, bid, tabsf, stp, tp_lds);
}
// ---------------------------------------------------------------------------
// Forces-only mixed kernel of the MD step (k_mix_1f): fl over the t / te
// lists.  Same pairs and FP32 arithmetic as pair_mixed_run, in fewer
// instructions per pair:
//   - a chunk of entries is loaded when it starts inside the wave's longest
//     row (a wave-uniform test); its entries past a row's end read the
//     slots that follow (the neighbour array ends in NB_LIST_PAD spare
//     slots) and a select zeroes them; every lane gathers its partner's
//     record (entry 0 reads atom 0's), so no exec-mask change guards a
//     load; partner types are gathered only for chunks that hold a vdW
//     entry;
//   - the coulomb term is evaluated on every lane and kept by a select (zero
//     for an entry not in te or a pair beyond the cut-off, whose table node
//     is clamped into the table); the vdW term runs only where an entry
//     carries the vdW flag;
//   - one table read per term: the node's coefficients {c0, c1, c2}
//     (computed in FP64 from cm3d's table), c0 + p (c1 + p c2);
//   - the vdW and coulomb F/r of a pair are added before the force update;
//   - an orthorhombic box takes the diagonal of h * 2^-32.
// ---------------------------------------------------------------------------
// record / type of sorted atom j (j < NB_MAX_ATOMS: a 32-bit byte offset)
static __device__ __forceinline__ uint4 recf_at(const uint4 *rec, uint32_t j) {
  return *(const uint4 *)((const char *)rec + (j << 4));
}
static __device__ __forceinline__ int typef_at(const int *t, uint32_t j) {
  return *(const int *)((const char *)t + (j << 2));
}

template <bool TLDS, bool ORTHO, int G>
static __device__ __forceinline__ void mix_fonly_run(const MixArgs &A, int bid, const float4 *fc,
                                                     const TPairF *stp) {
#pragma clang fp contract(fast)
  static_assert(G == 16 || G == NB_WAVE, "16 or 64 lanes per atom");
  constexpr int APW = NB_WAVE / G;
  const int lane = lane_id(), gl = lane % G, sub = lane / G;
  const int nw = A.nblocks * 4;
  const float4 *cc = fc + A.fc_coff;
  const int kmax = A.ntable - 3;
  const float pc0 = A.rmins_coul2 * A.dx2coul;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  for (int s0 = (bid * 4 + wv) * APW; s0 < A.n; s0 += nw * APW) {
    const int s = s0 + sub;
    const bool live = APW == 1 || s < A.n;
    const int sa = live ? s : s0;  // a dead group reads its wave's first atom
    const uint4 ri = A.rec[sa];
    const float qi = __uint_as_float(ri.w);
    const int ti = A.stype[sa];
    const int nb = live ? A.nnb[sa] : 0;
    const uint32_t *lst = A.nbr + (size_t)sa * A.cap + gl;
    const TPairF *tprow = (TLDS ? stp : A.tpf) + ti * A.ntypes;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    // entries two chunks ahead of the evaluation, partner records and types
    // one chunk ahead
    int nbw = nb;
    if constexpr (APW == 4)
      nbw = max(max(__builtin_amdgcn_readlane(nb, 0), __builtin_amdgcn_readlane(nb, 16)),
                max(__builtin_amdgcn_readlane(nb, 32), __builtin_amdgcn_readlane(nb, 48)));
    // a chunk that starts past the longest row of the wave is not loaded (a
    // wave-uniform branch); entries past a row's end are zeroed
    uint32_t ec = lst[0], en = 0u;
    ec = gl < nb ? ec : 0u;
    if (G < nbw) {
      en = lst[G];
      en = G + gl < nb ? en : 0u;
    }
    const uint32_t im = A.imask;
    uint4 rjc = recf_at(A.rec, ec & im);
    int tjc = 0;
    if (!A.etype && __any((ec >> 28) & NB_FLAG_VT)) tjc = typef_at(A.stype, ec & im);
#pragma unroll 2
    for (int k0 = 0; k0 < nbw; k0 += G) {
      // loaded raw; the select that zeroes entries past the row's end is
      // applied a chunk later, when the load has long completed
      uint32_t e2 = 0u;
      if (k0 + 2 * G < nbw) e2 = lst[k0 + 2 * G];
      const uint4 rjn = recf_at(A.rec, en & im);
      // partner types from the entry, else gathered only for chunks holding
      // vdW entries (the gathers, one cache line a lane, bound this kernel)
      int tjn = 0;
      if (!A.etype && __any((en >> 28) & NB_FLAG_VT)) tjn = typef_at(A.stype, en & im);
      const unsigned fl = ec >> 28;
      const float ux = (float)(int)(ri.x - rjc.x);
      const float uy = (float)(int)(ri.y - rjc.y);
      const float uz = (float)(int)(ri.z - rjc.z);
      float dx, dy, dz;
      if (ORTHO) {
        dx = ux * A.hs[0];
        dy = uy * A.hs[4];
        dz = uz * A.hs[8];
      } else {
        dx = ux * A.hs[0] + uy * A.hs[1] + uz * A.hs[2];
        dy = ux * A.hs[3] + uy * A.hs[4] + uz * A.hs[5];
        dz = ux * A.hs[6] + uy * A.hs[7] + uz * A.hs[8];
      }
      const float d2 = dx * dx + dy * dy + dz * dz;
      float fr;  // F/r of the pair
      {
        float p = d2 * A.dx2coul - pc0;
        int kk = (int)p;
        const bool lo = kk < 1;
        p = lo ? 1.f : p;
        kk = lo ? 1 : min(kk, kmax);
        p -= (float)kk;
        const float4 T = cc[kk];
        const bool on = (fl & NB_FLAG_ET) && d2 < A.rmaxl_coul2;
        const float f = T.x + p * (T.y + p * T.z);
        fr = on ? qi * __uint_as_float(rjc.w) * f : 0.f;
      }
      if (fl & NB_FLAG_VT) {
        const int tj = A.etype ? (int)((ec & NB_IDX_MASK) >> A.ibits) : tjc;
        const TPairF P = tprow[tj];
        if (d2 < P.rmaxl2) {
          float p = (d2 - P.rmins2) * P.dx2;
          int kk = (int)p;
          if (kk < 1) { kk = 1; p = 1.f; }
          kk = min(kk, kmax);
          p -= (float)kk;
          const float4 T = fc[P.in * A.ntable + kk];
          fr += T.x + p * (T.y + p * T.z);
        }
      }
      fx -= dx * fr;
      fy -= dy * fr;
      fz -= dz * fr;
      ec = en;
      en = k0 + 2 * G + gl < nb ? e2 : 0u;
      rjc = rjn;
      tjc = tjn;
    }
    const long long X = group_sum_ll<G>(to_fix(fx)), Y = group_sum_ll<G>(to_fix(fy));
    const long long Z = group_sum_ll<G>(to_fix(fz));
    if (gl == 0 && live) {
      const int ai = A.perm[sa], n = A.n;
      A.fl[ai] = (double)X / NB_FX_SCALE;
      A.fl[n + ai] = (double)Y / NB_FX_SCALE;
      A.fl[2 * n + ai] = (double)Z / NB_FX_SCALE;
    }
  }
}

template <bool TLDS, int G>
static __device__ __forceinline__ void mix_fonly_box(const MixArgs &A, int bid, const float4 *fc,
                                                     const TPairF *stp) {
  if (A.ortho) mix_fonly_run<TLDS, true, G>(A, bid, fc, stp);
  else mix_fonly_run<TLDS, false, G>(A, bid, fc, stp);
}

static __device__ __forceinline__ void mix_fonly_body(const MixArgs &A, int bid) {
  extern __shared__ float4 tabs4[];
  __shared__ TPairF stp[NB_TPF_LDS];
  const int nt2 = A.ntypes * A.ntypes;
  const bool lds = A.fc_lds && nt2 <= NB_TPF_LDS;
  if (lds) {
    const int nn = A.fc_lds / (int)sizeof(float4);
    for (int t = threadIdx.x; t < nn; t += blockDim.x) tabs4[t] = A.fcf[t];
    for (int t = threadIdx.x; t < nt2; t += blockDim.x) stp[t] = A.tpf[t];
  }
  __syncthreads();
  if (lds) {
    if (A.g16) mix_fonly_box<true, 16>(A, bid, tabs4, stp);
    else mix_fonly_box<true, NB_WAVE>(A, bid, tabs4, stp);
  } else {
    if (A.g16) mix_fonly_box<false, 16>(A, bid, A.fcf, stp);
    else mix_fonly_box<false, NB_WAVE>(A, bid, A.fcf, stp);
  }
}

#define k_mix_0t_body pair_mixed_body<0, true>
#define k_mix_1t_body pair_mixed_body<1, true>
#define k_mix_1f_body mix_fonly_body
NB_KERNEL_PAIR(k_mix_0t, MixArgs, 256)
NB_KERNEL_PAIR(k_mix_1t, MixArgs, 256)
NB_KERNEL_PAIR(k_mix_1f, MixArgs, 256)

int nb_launch_pair_mixed(nb_ctx *c, int iflg, const Box &b, double *d_fs, double *d_fl,
                         hipStream_t s, bool packed) {
  const int n = c->natoms;
  PackMixArgs pk = {c->d_pos, c->d_qch, c->d_perm, n, b, (uint4 *)c->d_recf};
  NB_GO(c, K_PACKMIX, k_packmix, packed ? 0 : (n + 255) / 256, 256, 0, pk, s);
  MixArgs M;
  M.n = n; M.ntypes = c->ntypes; M.ninter = c->ninter; M.ntable = c->ntable;
  M.cap = c->cap; M.ctab_off = c->ctab_off; M.tab_lds = c->tab_lds_bytes_f > 0;
  M.nblocks = c->npair_blocks;
  M.g16 = c->mix_g16;
  M.nbr = c->d_nbr; M.nnb = c->d_nnb; M.stype = c->d_stype; M.perm = c->d_perm;
  M.rec = (const uint4 *)c->d_recf; M.tpf = (const TPairF *)c->d_tpf;
  M.vtab3 = c->d_vtab3f; M.ctab3 = c->d_ctab3f;
  for (int k = 0; k < 9; k++) M.h[k] = (float)b.h[k];
  M.rmins_coul2 = (float)c->rmins_coul2; M.dx2coul = (float)c->dx2coul;
  M.rmaxl_coul2 = (float)c->rmaxl_coul2;
  M.fl = d_fl; M.fs = d_fs;
  M.fcf = c->d_fcoeff; M.fc_lds = c->fcoef_lds_f;
  M.fc_coff = c->ctab_off / 3;
  M.ortho = b.ortho;
  M.imask = c->imask; M.ibits = c->ibits; M.etype = c->etype;
  const float s32 = 2.3283064365386963e-10f;  // 2^-32
  for (int k = 0; k < 9; k++) M.hs[k] = M.h[k] * s32;
  const int g = c->npair_blocks, lds = c->tab_lds_bytes_f;
  if (iflg == 0)
    NB_GO(c, K_MIX_0T, k_mix_0t, g, 256, lds, M, s);
  else if (d_fs)
    NB_GO(c, K_MIX_1T, k_mix_1t, g, 256, lds, M, s);
  else
    NB_GO(c, K_MIX_1F, k_mix_1f, g, 256, c->fcoef_lds_f, M, s);
  if (!c->rec) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      nb_set_error("k_pair_mixed: %s", hipGetErrorString(e));
      return NB_ERR_HIP;
    }
  }
  return NB_OK;
}

static struct RegisterMixed {
  RegisterMixed() {
    g_batch_fn[K_PACKMIX] = k_packmix_bl;
    g_batch_fn[K_MIX_0T] = k_mix_0t_bl;
    g_batch_fn[K_MIX_1T] = k_mix_1t_bl;
    g_batch_fn[K_MIX_1F] = k_mix_1f_bl;
  }
} g_register_mixed;
