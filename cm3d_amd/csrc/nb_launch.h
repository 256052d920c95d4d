// nb_launch.h -- one launch path for single systems and batches.
//
// Every kernel has a __device__ body taking (args, block index) and two
// entry points: the single-system __global__ (args by value, blockIdx.x) and
// the batched one, which reads its system's args from an array in HBM and
// its (system, local block) from a block map.  A launch site either launches
// the single-system kernel or, when the context is recording, appends a
// LaunchRec; a batch of systems records the same step for each of them and
// then issues ONE grid per kernel covering every system's blocks, so the
// launch count per step does not grow with the number of systems.
#pragma once
#include <vector>

#include "nb_internal.h"

enum KernelId {
  K_ZERO, K_CELL, K_SCAT, K_ORDER, K_PACK64, K_LIST, K_CHECK,
  K_PAIR64_0T, K_PAIR64_1T, K_PAIR64_1F, K_PAIR64_1N, K_PAIR64_1E, K_PAIR64_1Q, K_RED, K_EWCHG, K_EWSFAC, K_EWKRED,
  K_EWFORCE, K_EWFIN, K_PACKMIX, K_MIX_0T, K_MIX_1T, K_MIX_1F, K_MDINNER, K_MDKICK,
  K_MDNHC, K_MDSAVE, K_OVFCHK, K_EWSFACR, K_EWSFACM, K_EWSFMTAB, K_EWFORCEM, K_EWFIMG, K_COUNT
};

struct PairArgs {
  int n, ntypes, ninter, ntable, cap, tab_lds, ctab_off, nblocks;
  const uint32_t *nbr;
  const int *nnb, *stype, *perm, *map;
  const double4 *rec;
  const InterPar *ip;
  const double *vtab3, *ctab3;
  const double *fcoef;  // forces-only kernel: force-column coefficients (vdW tables, then coulomb)
  int fc_lds;           // ... staged in LDS (bytes > 0) or read from HBM
  uint32_t imask;       // list entry: sorted index bits
  int ibits, etype;     // ... and the partner's type above them (when etype)
  double rmins_coul2, dx2coul, rmaxl_coul2, clast;
  Box b;
  double *fl, *fs;
  double *epart;
};


struct LaunchRec {
  int kid, nblk, tpb, lds;
  std::vector<char> args;
};

typedef void (*BatchFn)(const void *args, const int2 *bmap, int nblk, int tpb, int lds,
                        hipStream_t s);
extern BatchFn g_batch_fn[K_COUNT];

#define NB_KERNEL_PAIR(name, ArgsT, TPB) NB_KERNEL_PAIR_W(name, ArgsT, TPB, 1)
// WPE: waves per SIMD the register allocation must leave room for
#define NB_KERNEL_PAIR_W(name, ArgsT, TPB, WPE)                                      \
  __global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(WPE)))        \
  void name(ArgsT A) { name##_body(A, blockIdx.x); }                                 \
  __global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(WPE)))        \
  void name##_b(const ArgsT *__restrict__ AA, const int2 *__restrict__ bm) {         \
    const int2 m = bm[blockIdx.x];                                                   \
    const int sys = __builtin_amdgcn_readfirstlane(m.x);                             \
    const int lb = __builtin_amdgcn_readfirstlane(m.y);                              \
    name##_body(AA[sys], lb);                                                        \
  }                                                                                  \
  static void name##_bl(const void *a, const int2 *bm, int nb, int tpb, int lds,     \
                        hipStream_t s) {                                             \
    name##_b<<<nb, tpb, lds, s>>>((const ArgsT *)a, bm);                             \
  }

static inline void nb_record(nb_ctx *c, int kid, int nblk, int tpb, int lds, const void *a,
                             size_t bytes) {
  LaunchRec r;
  r.kid = kid; r.nblk = nblk; r.tpb = tpb; r.lds = lds;
  r.args.assign((const char *)a, (const char *)a + bytes);
  c->rec->push_back(r);
}

// Launch (or record) kernel KERN with id KID.
#define NB_GO(c, KID, KERN, NBLK, TPB, LDS, ARGS, S)                              \
  do {                                                                            \
    if ((c)->rec)                                                                 \
      nb_record((c), (KID), (NBLK), (TPB), (LDS), &(ARGS), sizeof(ARGS));         \
    else if ((NBLK) > 0)                                                          \
      KERN<<<(NBLK), (TPB), (LDS), (S)>>>(ARGS);                                  \
  } while (0)
