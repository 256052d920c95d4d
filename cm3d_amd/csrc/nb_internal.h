// nb_internal.h -- device-side data layout and helpers of the gfx950
// nonbonded engine (not part of the C ABI).
//
// HBM layout per context (N atoms, Q charges, K k-vectors, cap = list slots
// per atom):
//   pos      3N f64   caller's positions, cm3d SoA (x[N] y[N] z[N])
//   posb     3N f64   positions at the last list build (rebuild trigger)
//   fs, fl   3N f64   outputs, cm3d SoA, before rmass
//   perm/iperm N i32  cell-sorted order <-> atom index
//   rec64    N x {x,y,z,q} f64 in sorted order, refreshed every evaluation
//   stype    N i32    atom type in sorted order
//   nbr      N*cap u32 full (both directions) neighbor lists in sorted
//                     order: low 28 bits the sorted index j, high 4 bits
//                     the list flags (vdW t, vdW s, coul te, coul se)
//   nnb      N i32    list lengths
//   chg      Q x {2pi s_a, 2pi s_b, 2pi s_c, q sqrt(4pi)} f64, stp Q x {cos,
//                     sin}(2pi s_c), {cos, sin}(2pi NB_EW_SYM s_c):
//                     per-evaluation Ewald charge records
//   kd       K x {2 preg Re S, 2 preg Im S} f64
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/cm3d_nb.h"

struct LaunchRec;

#define NB_WAVE 64
#define NB_FLAG_VT 1u  // vdW t list  (rmaxl+skin)
#define NB_FLAG_VS 2u  // vdW s list  (rmaxs+skin)
#define NB_FLAG_ET 4u  // coulomb te  (rcute_max+skin)
#define NB_FLAG_ES 8u  // coulomb se  (rcute_resp+skin)
#define NB_IDX_MASK 0x0FFFFFFFu  // index and type bits of a list entry
#define NB_LIST_PAD 256  // slots after the last row: the forces-only pair kernels read
                         // up to 63 entries past a row's end unguarded
#define NB_MAX_ATOMS (1 << 27)  // also keeps a 32-byte record's byte offset in 32 bits
#ifndef NB_EW_SYM
#define NB_EW_SYM 16     // values of m = |kc| per structure-factor lane (2 k vectors each)
#endif
#define NB_EW_KT 512     // force-pass entries (values of m) per LDS tile
#ifndef NB_EWF_Q
#define NB_EWF_Q 3       // charges per lane in the Ewald force pass
#endif
#ifndef NB_EWF_WG
#define NB_EWF_WG 2048   // workgroups the Ewald force pass aims for
#endif
#ifndef NB_EWF_MIN_ENT
#define NB_EWF_MIN_ENT 64  // fewest entries (values of |kc| over rows) in a force-pass range
#endif
#ifndef NB_SFAC_WG
#define NB_SFAC_WG 4096  // workgroups the structure-factor pass aims for
#endif
#ifndef NB_CELL_REACH_MAX
#define NB_CELL_REACH_MAX 3  // list stencil: cells of >= rlist / reach, reach <= 3
#endif
#ifndef NB_CELL3_MIN_ATOMS
#define NB_CELL3_MIN_ATOMS 2.0  // mean atoms per cell needed for reach-3 cells
#endif
#ifndef NB_SFAC_PAIR
#define NB_SFAC_PAIR 1   // structure factor: two charges per trip of the charge loop
#endif
#ifndef NB_SFAC_ROWS_MIN_Q
#define NB_SFAC_ROWS_MIN_Q 16384  // charges from which a lone system's structure-factor lanes take several rows
#endif
#ifndef NB_SFAC_MIN_Q
#define NB_SFAC_MIN_Q 64  // fewest charges in a structure-factor chunk
#endif
#ifndef NB_EWF_WORK
#define NB_EWF_WORK 65536  // k x charge products per force workgroup in a shared context
#endif
#ifndef NB_SFAC_WORK
#define NB_SFAC_WORK 65536  // the same for a structure-factor workgroup
#endif
#ifndef NB_SFM_NT
#define NB_SFM_NT 5  // MFMA structure factor: m tiles of 8 a row tile may need (m < 40)
#endif
#define NB_SFM_NQ 16    // ... charges per staged tile (4 matrix-core K steps of 4)
#define NB_SFM_R 4      // ... row tiles (8 rows) per wave (fewer for waves of 4 or 5 m tiles)
// row tiles a wave of nt m tiles takes: the accumulators and two K steps'
// operands stay in 256 registers
__host__ __device__ static constexpr int nb_sfm_wave_tiles(int nt) { return nt >= 5 ? 2 : (nt == 4 ? 3 : 4); }
#define NB_SFM_TAB 40   // ... |ka|, |kb| below this (phase-table length)
#ifndef NB_SFM_WG
#define NB_SFM_WG 1024  // ... workgroups it aims for
#endif
#ifndef NB_SFM_MIN_WORK
#define NB_SFM_MIN_WORK (1LL << 30)  // ... charges x k vectors from which it is used
#endif
#ifndef NB_MIX_G16_MAX
#define NB_MIX_G16_MAX 256  // mixed pair kernel: 16 lanes per atom below this mean list length
#endif
#ifndef NB_MIX_ATOMS_WG
#define NB_MIX_ATOMS_WG 8    // atoms per workgroup of the mixed pair kernel
#endif
#ifndef NB_MIX_MAX_WG
#define NB_MIX_MAX_WG 8192  // workgroups of the mixed pair kernel at most
#endif
#ifndef NB_MIX_ATOMS_WG_SHARED
#define NB_MIX_ATOMS_WG_SHARED 64  // the same for a context with sharing >= 4
#endif
#ifndef NB_PAIR64_MAX_WG
#define NB_PAIR64_MAX_WG 2048  // workgroup cap of the FP64 pair kernel
#endif
#define NB_NE_PAIR 19    // pair energy/virial accumulators + clamp count
#define NB_NE_PAIRQ 27   // the same + cm3d's energy-only coulomb sums (see k_pair64_1q)
#define NB_EOUT_ECORR 28 // reduced energy columns: pair [0, 27), ecorr [28, 39),
#define NB_EOUT_EW 40    //   Ewald [40, 51)
#define NB_NE_ECORR 11   // ecorr energy/virial accumulators
#define NB_NE_EW 11      // Ewald accumulators

struct InterPar {   // one vdW interaction (cm3d INTER arrays at index in)
  double rmins2, rmaxs2, rminl2, rmaxl2, dx2, rmaxs_skin2, rmaxl_skin2, vlast;
};

struct TPairF {     // one type pair for the FP32 kernel (null pair: in -1, cutoffs -1)
  float rmins2, rminl2, rmaxs2, rmaxl2, dx2, vlast;
  int in, pad;
};

// The k vectors as rows: maximal runs of consecutive kc at one (ka, kb) in
// k order (cm3d's enumeration gives kc = -n..n, or 1..n for ka = kb = 0).
// Both reciprocal passes walk a row by m = |kc|, pairing kc with -kc: the
// phases e^{+-i m theta_c} share one rotation.
struct EwRow {
  int k_lo;          // k index of kc = kc_lo
  int kc_lo, kc_hi, ka, kb;
  int mlo, mhi;      // values of m the row covers
  int e0;            // its first force-pass entry (one per m)
  int chain;         // the previous row is (ka, kb - 1): its phase follows by one kb step
};
// A row of the MFMA structure factor, in its tile order
struct EwRowM { int ka, kb, k_lo, kc_lo, kc_hi, mhi, tmhi, pad; };  // tmhi: its tile's longest m
// Per-charge tables of the MFMA structure factor, in doubles: q e^{i ka
// theta_a} as {im, re, -im} for ka = 0..la-1 (padded to an even count),
// then e^{i kb theta_b} for kb = 0..lb-1, then e^{i m theta_c} for m =
// 0..lz-1; the stride odd in 16-byte units so that consecutive charges
// start in different LDS banks
__host__ __device__ static inline int nb_sfm_tb(int la) { return (3 * la + 1) & ~1; }
__host__ __device__ static inline int nb_sfm_tz(int la, int lb) { return nb_sfm_tb(la) + 2 * lb; }
__host__ __device__ static inline int nb_sfm_stride(int la, int lb, int lz) {
  int s = nb_sfm_tz(la, lb) + 2 * lz;
  if ((s / 2) % 2 == 0) s += 2;
  return s;
}
// A 16-row tile of the MFMA reciprocal force pass: 4 runs of up to 4
// consecutive kb in one ka plane (run q's row r is tile row q + 4 r), its
// m range from mt0 in nks steps of 2, its coefficient entries at ent_off
// (2 nks x 16, {k of +m, k of -m, m, 0}, -1: absent)
struct EwFTile { int ent_off, mt0, nks, pad; int ka[4], kb0[4]; };
#define NB_EFM_CT 3    // MFMA force pass: charge tiles (16 charges) per wave
#ifndef NB_EFM_MIN_WORK
#define NB_EFM_MIN_WORK (1LL << 24)  // ... charges x k vectors from which a lone system uses it
#endif
#define NB_EFM_MS 40   // ... values of m a tile may span (its image: 1 KB per value of m)
// A wave of the MFMA structure factor: its row tiles (an empty one past
// the last real tile where it has fewer) and their m-tile count
struct EwSfmWave { int t[4]; int nt, nr, pad0, pad1; };
struct EwSym {       // a structure-factor lane: nr rows (ka, kb .. kb + nr - 1), cnt values of m from m0
  int ka, kb, m0, cnt, nr, pad0, pad1, pad2;
  int k_lo[4], kc_lo[4], kc_hi[4];  // per row: k index of kc_lo, its kc range
};

// cm3d's -DLINUX anint: (int)(x +- 0.5), truncating (md_macros.h:31-52).
__device__ __forceinline__ double nb_anint(double a) {
  // (int)(a +- .5) as cm3d writes it, kept in FP64: a +- .5 is -(|a| + .5)
  // for a < 0, exactly, so the truncated magnitude of |a| + .5 with a's sign
  // is the same integer (|a| < 2^31 for any fractional coordinate), in three
  // instructions (add with an abs operand, trunc, sign insert) instead of a
  // compare, select and two conversions; only the sign of a zero differs
  return __builtin_copysign(__builtin_trunc(__builtin_fabs(a) + .5), a);
}

// Minimum image in cm3d's operation order (md_period.new.c:39-51).  The
// library is compiled with -ffp-contract=off so no FMA is formed here and
// the result is bitwise the reference's.
__device__ __forceinline__ void nb_minimg(double &dx, double &dy, double &dz,
                                          const double *h, const double *hi) {
  double sx = dx * hi[0] + dy * hi[1] + dz * hi[2];
  double sy = dx * hi[3] + dy * hi[4] + dz * hi[5];
  double sz = dx * hi[6] + dy * hi[7] + dz * hi[8];
  sx -= nb_anint(sx);
  sy -= nb_anint(sy);
  sz -= nb_anint(sz);
  dx = sx * h[0] + sy * h[1] + sz * h[2];
  dy = sx * h[3] + sy * h[4] + sz * h[5];
  dz = sx * h[6] + sy * h[7] + sz * h[8];
}

// 3-point quadratic interpolation (md_fnopol.c:146) on an interleaved table
// t[3*k + c]: c = component (0 force, 1 switch, 2 energy).
__device__ __forceinline__ double nb_interp3(const double *t, int k, int c,
                                             double p) {
  double fm1 = t[3 * (k - 1) + c], f0 = t[3 * k + c], f1 = t[3 * (k + 1) + c];
  return f0 + .5 * p * (f1 - fm1 + p * (fm1 - 2. * f0 + f1));
}

// Wavefront sums with DPP lane moves (VALU, no LDS round trips): swaps
// within quads, then rotations by 4 and 8 lanes within each row of 16,
// leave a row total in every lane of a row; the totals of rows 0..3 are then
// read from lanes 0, 16, 32, 48 and added in a fixed order.  Deterministic,
// and every lane returns the same value.
template <int CTRL>
__device__ __forceinline__ int nb_dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ long long nb_dpp(long long v) {
  const int lo = nb_dpp_i<CTRL>((int)v), hi = nb_dpp_i<CTRL>((int)(v >> 32));
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ int nb_dpp(int v) { return nb_dpp_i<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ double nb_dpp(double v) {
  return __longlong_as_double(nb_dpp<CTRL>((long long)__double_as_longlong(v)));
}
template <int CTRL>
__device__ __forceinline__ float nb_dpp(float v) {
  return __int_as_float(nb_dpp_i<CTRL>(__float_as_int(v)));
}
__device__ __forceinline__ int nb_readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ long long nb_readlane(long long v, int l) {
  const int lo = __builtin_amdgcn_readlane((int)v, l), hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double nb_readlane(double v, int l) {
  return __longlong_as_double(nb_readlane((long long)__double_as_longlong(v), l));
}
__device__ __forceinline__ float nb_readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
// sum over the 16 lanes of a DPP row; every lane of the row gets it
template <typename T>
__device__ __forceinline__ T nb_row_sum(T v) {
  v += nb_dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += nb_dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += nb_dpp<0x124>(v);  // row_ror:4
  v += nb_dpp<0x128>(v);  // row_ror:8
  return v;
}
template <typename T>
__device__ __forceinline__ T nb_wave_sum(T v) {
  v += nb_dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += nb_dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += nb_dpp<0x124>(v);  // row_ror:4
  v += nb_dpp<0x128>(v);  // row_ror:8
  return (nb_readlane(v, 0) + nb_readlane(v, 16)) + (nb_readlane(v, 32) + nb_readlane(v, 48));
}

struct Box {
  double h[9], hi[9];
  int ortho;  // off-diagonal entries of h (and so of h^-1) are all zero
};

// nb_minimg for a Box: with an orthorhombic box the off-diagonal products
// are exact zeros, x + 0 == x, so dropping them gives the same bits (up to
// the sign of a zero) at a third of the cost.
__device__ __forceinline__ void nb_minimg_box(double &dx, double &dy, double &dz,
                                              const Box &b) {
  if (b.ortho) {
    double sx = dx * b.hi[0], sy = dy * b.hi[4], sz = dz * b.hi[8];
    sx -= nb_anint(sx);
    sy -= nb_anint(sy);
    sz -= nb_anint(sz);
    dx = sx * b.h[0];
    dy = sy * b.h[4];
    dz = sz * b.h[8];
  } else {
    nb_minimg(dx, dy, dz, b.h, b.hi);
  }
}

struct nb_ctx {
  int dev;
  hipStream_t stream;
  hipStream_t stream2;        // the reciprocal sum's branch of the MD step
  hipEvent_t fork_ev[2];      // fork / join of that branch
  int natoms, ntypes, ninter, ntable, iperd, prec;
  int ncharge, ktot, nrow, nsym, necorr;
  double skin, kappa, ewald_self, ewald_bkg;
  double dx2coul, rmins_coul2, rmaxs_coul2, rminl_coul2, rmaxl_coul2;
  double rmaxs_skin2_e, rmaxl_skin2_e, rlist, clast;
  // vinter_e's per-pair table values: the coulomb table and its derivative
  // interpolated at its clamp (k = 1, p = -1; finter_e.c:330-352)
  double cq0, dq0;
  double rprune2;  // every list cutoff^2, widened by 2e-6: the column-pruning bound
  // parameters on device
  int *d_itype, *d_map;
  InterPar *d_ip;
  double *d_vtab3, *d_ctab3;
  float *d_vtab3f, *d_ctab3f;
  double *d_qch;
  int *d_xoff, *d_xlist;       // exclusions, both directions, flag bits
  int *d_eoff, *d_elist;       // ecorr partners, both directions
  int *d_qidx;                 // charge -> atom
  int *d_qpos;                 // atom -> charge (-1: uncharged)
  EwRow *d_row;                // k order
  EwSym *d_sym;                // structure-factor lanes, longest first
  int4 *d_ent;                 // force-pass entries: {k of +m, k of -m (-1: absent), m, 0}
  int *d_trow, *d_rtile;       // force pass: tiles (row bounds), ranges (tile bounds)
  int *d_kabc;                 // ka,kb,kc per k (3K)
  // state
  double *d_pos, *d_posb, *d_fs, *d_fl;
  int *d_perm, *d_iperm, *d_cell, *d_ccount, *d_cstart, *d_cfill, *d_cwork, *d_tmp;
  int *d_hcnt;                 // per sorted atom: half-pair counts s,t,se,te
  double *d_pcut;              // per type pair: (rmaxl+skin)^2, (rmaxs+skin)^2; -1 null
  double *d_rec64;
  int *d_stype;
  uint32_t *d_nbr;
  int *d_nnb;
  int cap;
  // a list entry: 4 flag bits, then (when ntypes fits) the partner's type
  // above its sorted index: [flags:4 | type : 28 - ibits | index : ibits]
  int ibits;                   // bits of the sorted index (enough for natoms)
  int etype;                   // entries carry the partner's type
  uint32_t imask;              // (1 << ibits) - 1
  int *d_flags;                // [0] overflow max count of the last build,
                               // [1] rebuild flag, [2] overflow max since the
                               // host last cleared it (sticky over a batch
                               // tick), [6] builds done, [7] k_cell's ticket
  // Bumped whenever a buffer or a launch shape that a recorded step (a
  // captured graph, a batch's argument arrays) refers to is reallocated or
  // changed: the neighbour slots, the cell arrays or the cell grid.
  unsigned gen;
  // The MD state whose positions the reciprocal-sum and ecorr energy
  // columns (d_eew, d_eec) were last computed at, by its step's force call;
  // null after any other force evaluation on this context
  const void *recip_owner;
  int ncx, ncy, ncz, ncells, cells_alloc;
  int crx, cry, crz;           // cell reach of the list stencil per axis (0..3)
  double *d_chg, *d_stp, *d_spart, *d_kd, *d_fpart;
  int nchunk, qc, nsw;         // structure factor: charge chunks, waves of lanes
  int sfac_rows;               // ... rows per lane (1: k_ewsfac; 2 or 4: k_ewsfacr)
  int sfm;                     // ... as matrix products (k_ewsfacm) instead
  int sfm_la, sfm_lb, sfm_lz;  // ... its phase-table lengths
  int sfm_nrwg;                // ... workgroups per chunk
  EwRowM *d_rowm;              // ... its rows, sorted, in tiles of 8
  EwSfmWave *d_sfmw;           // ... its waves' tiles (4 a workgroup)
  double *d_sfmtab;            // ... per-charge phase tables (charges padded to whole tiles)
  int efm, efm_ncb;            // reciprocal force as matrix products (k_ewforcem), charge blocks
  long long sfm_mfma, efm_mfma;  // matrix-core instructions a structure-factor / force launch issues
  EwFTile *d_ftile;            // ... its row tiles
  int4 *d_ftent;               // ... their coefficient entries
  int efm_nent;                // ... how many (2 nks x 16 a tile)
  double *d_fimg;              // ... the tiles' coefficients as staged in LDS (8 per entry)
  int *d_frtile;               // ... ranges (tile bounds)
  int nrange, nqb, ntile;      // force: tile ranges, charge workgroups, tiles
  double *d_epair, *d_eew, *d_eec, *d_eout;
  int npair_blocks, new_blocks, nfin_blocks;
  int mix_g16;     // mixed pair kernel: 16 lanes per atom (short lists)
  long long *d_fxp;            // fixed-point force accumulators (mixed)
  float4 *d_recf;              // mixed: {sx, sy, sz (u32 bits), q}
  TPairF *d_tpf;               // mixed: per type pair parameters
  // forces-only pair kernels: the force column of every table as per-node
  // coefficients {f0, (f1 - f-1)/2, (f-1 - 2 f0 + f1)/2}, the vdW
  // interactions' tables first, then the coulomb table's; FP64 (3 doubles
  // a node) and, in a mixed context, FP32 ({f0, c1, c2, 0})
  double *d_fcoef;
  float4 *d_fcoeff;
  int fcoef_lds, fcoef_lds_f;  // their LDS bytes (0: too large, read from HBM)
  Box box_built;
  int have_lists;
  size_t dev_bytes;            // device memory held by this context (and its MD state)
  int tab_lds_bytes, tab_lds_bytes_f, ctab_off;
  // host mirrors
  int *h_flags;
  // optional per-phase timing: a ring of event records, one per force call
  int timing, trec;
  hipEvent_t *tev;            // NB_TIME_REC x NB_TIME_EV events
  // when set, launch sites append to this list instead of launching
  std::vector<LaunchRec> *rec;
};

// Doubles of the MFMA structure factor's per-charge tables (charges padded
// to whole staged tiles)
static inline size_t nb_sfm_table_doubles(const nb_ctx *c) {
  const size_t nqp = NB_SFM_NQ * (((size_t)c->ncharge + NB_SFM_NQ - 1) / NB_SFM_NQ);
  return nqp * (size_t)nb_sfm_stride(c->sfm_la, c->sfm_lb, c->sfm_lz);
}
#define NB_TIME_REC 2048
#define NB_TIME_EV 8  // start | pack | pair | charge records | structure factor |
                      // k reduction | reciprocal force | ecorr + forces + reductions (end)

// kernel launchers (nb_kernels.hip)
int nb_launch_build(nb_ctx *c, const Box &b, hipStream_t s);
int nb_launch_build_cond(nb_ctx *c, const Box &b, hipStream_t s);
int nb_build_now(nb_ctx *c, const double hmat[9], hipStream_t s);
#define NB_FORCE_PACKED 1  // sorted records / Ewald charge data already current
#define NB_FORCE_NORED 2   // leave the energy columns unreduced
#define NB_FORCE_EONLY 4   // energies and virials only: no reciprocal-space force pass
#define NB_FORCE_NO_PAIR 8    // leave out the pair part (pack + pair kernel)
#define NB_FORCE_NO_RECIP 16  // leave out the reciprocal sum (charge records .. k_ewforce)
#define NB_FORCE_NO_FIN 32    // leave out k_ewfin (ecorr + reciprocal forces into fl) and k_red
#define NB_FORCE_NOENER 64    // forces only: no pair energies / virials (the MD step's
                              // force evaluation; its energies are evaluated on request)
#define NB_FORCE_REUSE_RECIP 256  // with EONLY: the reciprocal-sum and ecorr energy columns
                                  // of the last force(...,1) are of these same positions:
                                  // reduce them as they are (no charge records, structure
                                  // factor, k reduction or k_ewfin)
#define NB_FORCE_CM3D_EONLY 128  // with EONLY: also the sums cm3d's energy-only
                                 // coulomb routine (vinter_e) reduces to
int nb_launch_force(nb_ctx *c, int iflg, const Box &b, double *d_fs,
                    double *d_fl, hipStream_t s, int flags);
int nb_launch_rebuild_check(nb_ctx *c, const double *d_pos, hipStream_t s);
// nb_energy from the reduced columns E of a force(...,iflg) evaluation
// cm3d: the coulomb fields take vinter_e's values (the reduction ran with
// NB_FORCE_CM3D_EONLY), the *_phys fields the force path's
void nb_fill_energy(nb_ctx *c, int iflg, const double *E, const double hmat[9], nb_energy *e,
                    bool cm3d = false);
// det(h), cm3d's get_deth
static inline double nb_det9(const double h[9]) {
  return h[0] * (h[4] * h[8] - h[5] * h[7]) + h[1] * (h[5] * h[6] - h[3] * h[8]) +
         h[2] * (h[3] * h[7] - h[4] * h[6]);
}
void nb_time_mark(nb_ctx *c, int ev, hipStream_t s);
int nb_launch_pair_mixed(nb_ctx *c, int iflg, const Box &b, double *d_fs, double *d_fl,
                         hipStream_t s, bool packed);
struct RedSet;
// the energy-column reductions of a force evaluation; pair: the pair
// kernel's columns were written (not a forces-only evaluation)
RedSet nb_red_set(const nb_ctx *c, int iflg, bool pair = true, bool quirk = false);
void nb_set_error(const char *fmt, ...);
