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
//   - each lane sums its pairs' FP32 forces in FP64 (exact for a handful of
//     terms), the lane sums are converted to 64-bit fixed point (2^-24 K/A
//     resolution) and reduced across the wavefront in integer arithmetic, so
//     the per-atom force does not depend on how pairs are spread over lanes
//     or in which order partial sums meet;
//   - energies and virials accumulate in FP64.
#include "nb_launch.h"

#define NB_FX_SCALE 16777216.0  // 2^24 fixed-point units per K/A

static __device__ __forceinline__ int lane_idm() { return threadIdx.x & (NB_WAVE - 1); }

static __device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, NB_WAVE);
  return v;
}

// Sorted-order records for the mixed kernel: {u_x, u_y, u_z, q as f32 bits}.
struct PackMixArgs { const double *pos, *q; const int *perm; int n; Box b; uint4 *rec; };
static __device__ __forceinline__ void k_packmix_body(const PackMixArgs &A, int bid) {
  int s = bid * blockDim.x + threadIdx.x;
  if (s >= A.n) return;
  const int n = A.n, a = A.perm[s];
  const double *pos = A.pos;
  double x = pos[a], y = pos[n + a], z = pos[2 * n + a];
  const double *hi = A.b.hi;
  double sx = x * hi[0] + y * hi[1] + z * hi[2];
  double sy = x * hi[3] + y * hi[4] + z * hi[5];
  double sz = x * hi[6] + y * hi[7] + z * hi[8];
  sx -= floor(sx);
  sy -= floor(sy);
  sz -= floor(sz);
  const double two32 = 4294967296.0;
  uint4 r;
  r.x = (unsigned)(unsigned long long)(sx * two32);
  r.y = (unsigned)(unsigned long long)(sy * two32);
  r.z = (unsigned)(unsigned long long)(sz * two32);
  r.w = __float_as_uint((float)A.q[a]);
  A.rec[s] = r;
}
NB_KERNEL_PAIR(k_packmix, PackMixArgs, 256)

struct MixArgs {
  int n, ntypes, ninter, ntable, cap, ctab_off, tab_lds, nblocks;
  const uint32_t *nbr;
  const int *nnb, *stype, *perm, *map;
  const uint4 *rec;
  const InterPar *ip;
  const float *vtab3, *ctab3;
  float h[9];
  float rmins_coul2, dx2coul, rmaxl_coul2, clast;
  double *fl, *fs;
  double *epart;
};

enum {
  M_SPOTL = 0, M_SPOTHL, M_SPRSL, M_SL0, M_SL1, M_SL2, M_SL4, M_SL5, M_SL8,
  M_SPOTE, M_SPOTHE, M_SPRSE, M_SE0, M_SE1, M_SE2, M_SE4, M_SE5, M_SE8, M_CLAMP
};

static __device__ __forceinline__ float interp3f(const float *t, int k, int c, float p) {
#pragma clang fp contract(fast)
  float fm1 = t[3 * (k - 1) + c], f0 = t[3 * k + c], f1 = t[3 * (k + 1) + c];
  return f0 + .5f * p * (f1 - fm1 + p * (fm1 - 2.f * f0 + f1));
}

// block totals of NACC accumulators stored column-major (out[k * stride])
template <int NACC>
static __device__ __forceinline__ void block_reduce_store_m(const double *acc, double *out,
                                                            int stride) {
  __shared__ double red[4][NACC];
  const int w = threadIdx.x >> 6, lane = lane_idm();
#pragma unroll
  for (int k = 0; k < NACC; k++) {
    double v = nb_wave_sum(acc[k]);
    if (lane == 0) red[w][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NACC) {
    double v = red[0][threadIdx.x] + red[1][threadIdx.x];
    v += red[2][threadIdx.x] + red[3][threadIdx.x];
    out[(size_t)threadIdx.x * stride] = v;
  }
}

static __device__ __forceinline__ long long to_fix(double v) {
  return __double2ll_rn(v * NB_FX_SCALE);
}

template <int IFLG, bool WANT_FS>
static __device__ __forceinline__ void pair_mixed_body(const MixArgs &A, int bid) {
  // FP32 pair arithmetic may contract to FMA (the FP64 path may not)
#pragma clang fp contract(fast)
  extern __shared__ float tabsf[];
  const float *vt = A.vtab3, *ct = A.ctab3;
  if (A.tab_lds) {
    int nv = A.ninter * A.ntable * 3, nc = A.ntable * 3;
    for (int t = threadIdx.x; t < nv; t += blockDim.x) tabsf[t] = A.vtab3[t];
    for (int t = threadIdx.x; t < nc; t += blockDim.x) tabsf[A.ctab_off + t] = A.ctab3[t];
    __syncthreads();
    vt = tabsf;
    ct = tabsf + A.ctab_off;
  }
  double acc[NB_NE_PAIR];
#pragma unroll
  for (int k = 0; k < NB_NE_PAIR; k++) acc[k] = 0.;
  const int lane = lane_idm();
  const int nw = A.nblocks * 4;
  constexpr unsigned fv = IFLG ? NB_FLAG_VT : NB_FLAG_VS;
  constexpr unsigned fe = IFLG ? NB_FLAG_ET : NB_FLAG_ES;
  const int ntm4 = A.ntable - 4;
  const float s32 = 2.3283064365386963e-10f;  // 2^-32
  for (int s = bid * 4 + (threadIdx.x >> 6); s < A.n; s += nw) {
    const uint4 ri = A.rec[s];
    const float qi = __uint_as_float(ri.w);
    const int ti = A.stype[s];
    const int nb = A.nnb[s];
    const uint32_t *lst = A.nbr + (size_t)s * A.cap;
    double fx = 0., fy = 0., fz = 0., sx = 0., sy = 0., sz = 0.;
    for (int k = lane; k < nb; k += NB_WAVE) {
      uint32_t e = lst[k];
      unsigned fl = e >> 28;
      if (!(fl & (fv | fe))) continue;
      int j = (int)(e & NB_IDX_MASK);
      uint4 rj = A.rec[j];
      float ux = (float)(int)(ri.x - rj.x) * s32;
      float uy = (float)(int)(ri.y - rj.y) * s32;
      float uz = (float)(int)(ri.z - rj.z) * s32;
      float dx = ux * A.h[0] + uy * A.h[1] + uz * A.h[2];
      float dy = ux * A.h[3] + uy * A.h[4] + uz * A.h[5];
      float dz = ux * A.h[6] + uy * A.h[7] + uz * A.h[8];
      float d2 = dx * dx + dy * dy + dz * dz;
      if (fl & fv) {
        int in = A.map[ti * A.ntypes + A.stype[j]];
        const InterPar &P = A.ip[in];
        float rmax2 = (float)(IFLG ? P.rmaxl2 : P.rmaxs2);
        if (d2 < rmax2) {
          float p = (d2 - (float)P.rmins2) * (float)P.dx2;
          int kk = (int)p;
          if (kk < 1) { acc[M_CLAMP] += 1.; kk = 1; p = 1.f; }
          if (kk > A.ntable - 3) kk = A.ntable - 3;
          p = p - (float)kk;
          const float *T = vt + (size_t)in * A.ntable * 3;
          float f = interp3f(T, kk, 0, p);
          float swv = 1.f;
          if ((WANT_FS || !IFLG) && d2 > (float)P.rminl2) swv = interp3f(T, kk, 1, p);
          if (IFLG) {
            acc[M_SPRSL] -= (double)(d2 * f);
            acc[M_SL0] -= (double)(dx * dx * f);
            acc[M_SL1] -= (double)(dy * dx * f);
            acc[M_SL2] -= (double)(dz * dx * f);
            acc[M_SL4] -= (double)(dy * dy * f);
            acc[M_SL5] -= (double)(dz * dy * f);
            acc[M_SL8] -= (double)(dz * dz * f);
            float gx = dx * f, gy = dy * f, gz = dz * f;
            fx -= (double)gx; fy -= (double)gy; fz -= (double)gz;
            if (WANT_FS) { sx -= (double)(gx * swv); sy -= (double)(gy * swv); sz -= (double)(gz * swv); }
            float v = interp3f(T, kk, 2, p);
            acc[M_SPOTL] += (double)v;
            acc[M_SPOTHL] += (double)(v - T[3 * ntm4 + 2]);
          } else {
            f *= swv;
            sx -= (double)(dx * f); sy -= (double)(dy * f); sz -= (double)(dz * f);
          }
        }
      }
      if ((fl & fe) && d2 < A.rmaxl_coul2) {
        float p = (d2 - A.rmins_coul2) * A.dx2coul;
        int kk = (int)p;
        if (kk < 1) { acc[M_CLAMP] += 1.; kk = 1; p = 1.f; }
        if (kk > A.ntable - 3) kk = A.ntable - 3;
        p = p - (float)kk;
        float qij = qi * __uint_as_float(rj.w);
        float fcc = qij * interp3f(ct, kk, 0, p);
        float gx = dx * fcc, gy = dy * fcc, gz = dz * fcc;
        if (IFLG) {
          acc[M_SPRSE] -= (double)(d2 * fcc);
          acc[M_SE0] -= (double)(dx * dx * fcc);
          acc[M_SE1] -= (double)(dy * dx * fcc);
          acc[M_SE2] -= (double)(dz * dx * fcc);
          acc[M_SE4] -= (double)(dy * dy * fcc);
          acc[M_SE5] -= (double)(dz * dy * fcc);
          acc[M_SE8] -= (double)(dz * dz * fcc);
          fx -= (double)gx; fy -= (double)gy; fz -= (double)gz;
          float v = interp3f(ct, kk, 2, p);
          acc[M_SPOTHE] += (double)(qij * (v - A.clast));
          acc[M_SPOTE] += (double)(qij * v);
        }
        if (WANT_FS || !IFLG) { sx -= (double)gx; sy -= (double)gy; sz -= (double)gz; }
      }
    }
    const int ai = A.perm[s];
    const int n = A.n;
    if (IFLG) {
      long long X = wave_sum_ll(to_fix(fx)), Y = wave_sum_ll(to_fix(fy)), Z = wave_sum_ll(to_fix(fz));
      if (lane == 0) {
        A.fl[ai] = (double)X / NB_FX_SCALE;
        A.fl[n + ai] = (double)Y / NB_FX_SCALE;
        A.fl[2 * n + ai] = (double)Z / NB_FX_SCALE;
      }
    }
    if (WANT_FS || !IFLG) {
      long long X = wave_sum_ll(to_fix(sx)), Y = wave_sum_ll(to_fix(sy)), Z = wave_sum_ll(to_fix(sz));
      if (lane == 0) {
        A.fs[ai] = (double)X / NB_FX_SCALE;
        A.fs[n + ai] = (double)Y / NB_FX_SCALE;
        A.fs[2 * n + ai] = (double)Z / NB_FX_SCALE;
      }
    }
  }
  block_reduce_store_m<NB_NE_PAIR>(acc, A.epart + bid, A.nblocks);
}
#define k_mix_0t_body pair_mixed_body<0, true>
#define k_mix_1t_body pair_mixed_body<1, true>
#define k_mix_1f_body pair_mixed_body<1, false>
NB_KERNEL_PAIR(k_mix_0t, MixArgs, 256)
NB_KERNEL_PAIR(k_mix_1t, MixArgs, 256)
NB_KERNEL_PAIR(k_mix_1f, MixArgs, 256)

int nb_launch_pair_mixed(nb_ctx *c, int iflg, const Box &b, double *d_fs, double *d_fl,
                         hipStream_t s) {
  const int n = c->natoms;
  PackMixArgs pk = {c->d_pos, c->d_qch, c->d_perm, n, b, (uint4 *)c->d_recf};
  NB_GO(c, K_PACKMIX, k_packmix, (n + 255) / 256, 256, 0, pk, s);
  MixArgs M;
  M.n = n; M.ntypes = c->ntypes; M.ninter = c->ninter; M.ntable = c->ntable;
  M.cap = c->cap; M.ctab_off = c->ctab_off; M.tab_lds = c->tab_lds_bytes_f > 0;
  M.nblocks = c->npair_blocks;
  M.nbr = c->d_nbr; M.nnb = c->d_nnb; M.stype = c->d_stype; M.perm = c->d_perm;
  M.map = c->d_map; M.rec = (const uint4 *)c->d_recf; M.ip = c->d_ip;
  M.vtab3 = c->d_vtab3f; M.ctab3 = c->d_ctab3f;
  for (int k = 0; k < 9; k++) M.h[k] = (float)b.h[k];
  M.rmins_coul2 = (float)c->rmins_coul2; M.dx2coul = (float)c->dx2coul;
  M.rmaxl_coul2 = (float)c->rmaxl_coul2; M.clast = (float)c->clast;
  M.fl = d_fl; M.fs = d_fs; M.epart = c->d_epair;
  const int g = c->npair_blocks, lds = c->tab_lds_bytes_f;
  if (iflg == 0)
    NB_GO(c, K_MIX_0T, k_mix_0t, g, 256, lds, M, s);
  else if (d_fs)
    NB_GO(c, K_MIX_1T, k_mix_1t, g, 256, lds, M, s);
  else
    NB_GO(c, K_MIX_1F, k_mix_1f, g, 256, lds, M, s);
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
