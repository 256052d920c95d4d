// nb_device.h -- device helpers shared by the kernel files: block
// reductions into column-major partial tables, the column-sum body used by
// k_red and the MD kick, and the per-atom record packers used by both the
// force path and the fused MD step.
#pragma once
#include "nb_launch.h"
#include "nb_sincos.h"

static __device__ __forceinline__ int lane_id() { return threadIdx.x & (NB_WAVE - 1); }

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

// Column sums of the column-major partial tables: block b reduces column b
// of the concatenation [pair | ecorr | Ewald] with a fixed assignment of rows
// to threads and a fixed tree (deterministic).  The same body runs inside the
// MD kick kernel.
struct RedSet {
  const double *part[3]; int rows[3], cols[3]; double *out[3]; int ncols;
};
static __device__ __forceinline__ void reduce_col(const RedSet &R, int col) {
  int t = 0;
  while (t < 2 && col >= R.cols[t]) { col -= R.cols[t]; t++; }
  const double *p = R.part[t] + (size_t)col * R.rows[t];
  double acc = 0.;
  for (int r = threadIdx.x; r < R.rows[t]; r += blockDim.x) acc += p[r];
  block_reduce_store<1>(&acc, R.out[t] + col, 1);
}
// Sorted-order record of the mixed kernel: {u_x, u_y, u_z, q as f32 bits},
// u = frac(h^-1 r) * 2^32 (the wrapping int32 difference of two records is
// the minimum image in fractional space).
static __device__ __forceinline__ uint4 nb_mixed_record(double x, double y, double z, double q,
                                                        const double *hi) {
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
  r.w = __float_as_uint((float)q);
  return r;
}

// Per-charge Ewald data: 2 pi s with s = h^-1 r wrapped to [-1/2,1/2) (exact
// for e^{2 pi i k.s} with integer k), the charge times sqrt(4 pi)
// (md_ewald.c:220-224), and e^{2 pi i s_c} and e^{2 pi i NB_EW_SYM s_c}: the
// c-axis step of the recursion and the length of a structure-factor lane.
static __device__ __forceinline__ void nb_charge_record(double x, double y, double z, double q,
                                                        const double *hi, double4 *chg,
                                                        double4 *stp) {
  double sa = x * hi[0] + y * hi[1] + z * hi[2];
  double sb = x * hi[3] + y * hi[4] + z * hi[5];
  double sc = x * hi[6] + y * hi[7] + z * hi[8];
  sa -= rint(sa);
  sb -= rint(sb);
  sc -= rint(sc);
  const double tpi = 2. * M_PI;
  *chg = make_double4(tpi * sa, tpi * sb, tpi * sc, q * sqrt(4.0 * M_PI));
  double sn, cs, sn16, cs16;
  nb_sincos(tpi * sc, &sn, &cs);
  nb_sincos((double)NB_EW_SYM * (tpi * sc), &sn16, &cs16);
  *stp = make_double4(cs, sn, cs16, sn16);
}
