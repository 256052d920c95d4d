// nb_api.hip -- C ABI entry points (include/cm3d_nb.h): context set-up,
// list build, rebuild trigger, force evaluation and energy readback.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include <cstdlib>

#include "nb_internal.h"

static thread_local char g_err[512];

void nb_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char *nb_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// The static scratch area of nb_force_device_noalloc: one per GPU, held by
// the loaded code object until the process ends.  Host code never uses the
// symbol's name as an address: each device's address is looked up once with
// that device current (at nb_create, else on the entry point's first call
// there) and kept below.
#define NB_SCRATCH_BYTES (32u << 20)
#define NB_MAX_DEVICES 64
static __device__ __align__(256) char scratch[NB_SCRATCH_BYTES];
static char *g_scratch_ptr[NB_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_ptr(int dev, char **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= NB_MAX_DEVICES) {
    nb_set_error("scratch: device %d out of range", dev);
    return NB_ERR_ARG;
  }
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != dev) hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      nb_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev,
                   hipGetErrorString(e != hipSuccess ? e : hipErrorInvalidSymbol));
      return NB_ERR_HIP;
    }
    g_scratch_ptr[dev] = (char *)p;
  }
  *out = g_scratch_ptr[dev];
  return NB_OK;
}

#define HCHK(x)                                                        \
  do {                                                                 \
    hipError_t _e = (x);                                               \
    if (_e != hipSuccess) {                                            \
      nb_set_error("%s: %s", #x, hipGetErrorString(_e));               \
      return NB_ERR_HIP;                                               \
    }                                                                  \
  } while (0)

template <typename T>
static int dalloc(nb_ctx *c, T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    nb_set_error("hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    return NB_ERR_HIP;
  }
  c->dev_bytes += count * sizeof(T);
  return NB_OK;
}

template <typename T>
static int upload(nb_ctx *c, T **p, const T *src, size_t count) {
  int rc = dalloc(c, p, count);
  if (rc) return rc;
  if (count && src) HCHK(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
  return NB_OK;
}

#define RC(x)                      \
  do {                             \
    int _r = (x);                  \
    if (_r != NB_OK) return _r;    \
  } while (0)

// gethinv9 (utils/md_mathlib.c:30-50): cofactor inverse, same operation order.
static Box make_box(const double h[9]) {
  Box b;
  for (int k = 0; k < 9; k++) b.h[k] = h[k];
  double rvol = 1. / (h[0] * (h[4] * h[8] - h[5] * h[7]) + h[1] * (h[5] * h[6] - h[3] * h[8]) +
                      h[2] * (h[3] * h[7] - h[4] * h[6]));
  b.hi[0] = (h[4] * h[8] - h[5] * h[7]) * rvol;
  b.hi[1] = (h[2] * h[7] - h[1] * h[8]) * rvol;
  b.hi[2] = (h[1] * h[5] - h[2] * h[4]) * rvol;
  b.hi[3] = (h[5] * h[6] - h[3] * h[8]) * rvol;
  b.hi[4] = (h[0] * h[8] - h[2] * h[6]) * rvol;
  b.hi[5] = (h[2] * h[3] - h[0] * h[5]) * rvol;
  b.hi[6] = (h[3] * h[7] - h[4] * h[6]) * rvol;
  b.hi[7] = (h[1] * h[6] - h[0] * h[7]) * rvol;
  b.hi[8] = (h[0] * h[4] - h[1] * h[3]) * rvol;
  b.ortho = h[1] == 0. && h[2] == 0. && h[3] == 0. && h[5] == 0. && h[6] == 0. && h[7] == 0.;
  return b;
}

static double det9(const double h[9]) { return nb_det9(h); }

extern "C" void nb_destroy(nb_ctx *c) {
  if (!c) return;
  hipSetDevice(c->dev);
  hipDeviceSynchronize();
  void *ptrs[] = {c->d_itype, c->d_map, c->d_ip, c->d_vtab3, c->d_ctab3, c->d_vtab3f,
                  c->d_ctab3f, c->d_qch, c->d_xoff, c->d_xlist, c->d_eoff, c->d_elist,
                  c->d_qidx, c->d_qpos, c->d_row, c->d_sym, c->d_rowm, c->d_sfmw, c->d_sfmtab, c->d_ftile, c->d_ftent, c->d_fimg, c->d_frtile, c->d_ent, c->d_trow, c->d_rtile, c->d_fpart, c->d_kabc, c->d_pos, c->d_posb, c->d_fs, c->d_fl,
                  c->d_perm, c->d_iperm, c->d_cell, c->d_ccount, c->d_cstart, c->d_cfill, c->d_cwork,
                  c->d_hcnt, c->d_pcut,
                  c->d_tmp, c->d_rec64, c->d_stype, c->d_nbr, c->d_nnb, c->d_flags,
                  c->d_chg, c->d_stp, c->d_spart, c->d_kd, c->d_epair, c->d_eew, c->d_eec, c->d_eout,
                  c->d_fxp, c->d_recf, c->d_tpf, c->d_fcoef, c->d_fcoeff};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (c->h_flags) hipHostFree(c->h_flags);
  if (c->tev) {
    for (int k = 0; k < NB_TIME_REC * NB_TIME_EV; k++) hipEventDestroy(c->tev[k]);
    delete[] c->tev;
  }
  if (c->stream) hipStreamDestroy(c->stream);
  if (c->stream2) hipStreamDestroy(c->stream2);
  for (int k = 0; k < 2; k++)
    if (c->fork_ev[k]) hipEventDestroy(c->fork_ev[k]);
  delete c;
}

size_t nb_kdata_bytes();

extern "C" int nb_create(const nb_params *p, nb_ctx **out) {
  if (!p || !out) { nb_set_error("nb_create: null argument"); return NB_ERR_ARG; }
  *out = nullptr;
  const int n = p->natoms;
  if (n <= 0 || n >= NB_MAX_ATOMS || p->ntypes <= 0 || p->ninter < 0 || p->ntable < 4 ||
      !p->itype || !p->map || !p->qch) {
    nb_set_error("nb_create: bad sizes or missing arrays");
    return NB_ERR_ARG;
  }
  if (p->iperd != 3) {
    nb_set_error("nb_create: only iperd==3 (3-D periodic) is supported");
    return NB_ERR_UNSUPPORTED;
  }
  if (p->precision != NB_PREC_FP64 && p->precision != NB_PREC_MIXED) {
    nb_set_error("nb_create: unknown precision %d", p->precision);
    return NB_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    nb_set_error("nb_create: no HIP device");
    return NB_ERR_NODEVICE;
  }
  if (p->device < 0 || p->device >= ndev) {
    nb_set_error("nb_create: device %d out of range", p->device);
    return NB_ERR_ARG;
  }
  // host-side validation of the type map and interaction indices
  for (int i = 0; i < n; i++)
    if (p->itype[i] < 0 || p->itype[i] >= p->ntypes) {
      nb_set_error("nb_create: itype[%d]=%d out of range", i, p->itype[i]);
      return NB_ERR_ARG;
    }
  for (int k = 0; k < p->ntypes * p->ntypes; k++)
    if (p->map[k] < -1 || p->map[k] >= p->ninter) {
      nb_set_error("nb_create: map entry %d out of range", p->map[k]);
      return NB_ERR_ARG;
    }
  nb_ctx *c = new nb_ctx();
  memset(c, 0, sizeof(*c));
  c->dev = p->device;
  if (hipSetDevice(c->dev) != hipSuccess) { delete c; nb_set_error("hipSetDevice failed"); return NB_ERR_HIP; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    nb_set_error("hipStreamCreate failed");
    return NB_ERR_HIP;
  }
  int rc = NB_OK;
  auto fail = [&](int r) { nb_destroy(c); return r; };
  // default-priority second stream and the two events of the MD step's fork
  if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork_ev[1], hipEventDisableTiming) != hipSuccess) {
    nb_set_error("hipStreamCreate / hipEventCreate failed");
    return fail(NB_ERR_HIP);
  }
  c->natoms = n; c->ntypes = p->ntypes; c->ninter = p->ninter; c->ntable = p->ntable;
  c->ibits = 1;
  while ((1 << c->ibits) < n) c->ibits++;
  c->imask = (1u << c->ibits) - 1u;
  c->etype = c->ibits < 28 && p->ntypes <= (1 << (28 - c->ibits));
  // CM3D_NB_ENTRY_TYPE=0 keeps the types out of the entries, as when they
  // do not fit, so that the kernels' gather of partner types can be tested
  // on small systems
  if (const char *ev = getenv("CM3D_NB_ENTRY_TYPE"))
    if (ev[0] == '0') c->etype = 0;
  c->iperd = p->iperd; c->prec = p->precision; c->skin = p->skin;
  c->kappa = p->kappa; c->ewald_self = p->ewald_self; c->ewald_bkg = p->ewald_bkg;
  c->dx2coul = p->dx2coul; c->rmins_coul2 = p->rmins_coul2; c->rmaxs_coul2 = p->rmaxs_coul2;
  c->rminl_coul2 = p->rminl_coul2; c->rmaxl_coul2 = p->rmaxl_coul2;
  c->rmaxs_skin2_e = p->rmaxs_skin2_e; c->rmaxl_skin2_e = p->rmaxl_skin2_e;

  const int nt = p->ntable, ni = std::max(p->ninter, 1);
  // interleaved tables {dV/r, switch, V} per node
  std::vector<double> vt3((size_t)ni * nt * 3, 0.), ct3((size_t)nt * 3, 0.);
  std::vector<InterPar> ip(ni);
  memset(ip.data(), 0, sizeof(InterPar) * ni);
  for (int k = 0; k < p->ninter; k++) {
    for (int t = 0; t < nt; t++) {
      vt3[((size_t)k * nt + t) * 3 + 0] = p->dvtab[(size_t)k * nt + t];
      vt3[((size_t)k * nt + t) * 3 + 1] = p->switchs[(size_t)k * nt + t];
      vt3[((size_t)k * nt + t) * 3 + 2] = p->vtab[(size_t)k * nt + t];
    }
    ip[k].rmins2 = p->rmins_cut2[k]; ip[k].rmaxs2 = p->rmaxs_cut2[k];
    ip[k].rminl2 = p->rminl_cut2[k]; ip[k].rmaxl2 = p->rmaxl_cut2[k];
    ip[k].dx2 = p->dx2tab[k]; ip[k].rmaxs_skin2 = p->rmaxs_skin2[k];
    ip[k].rmaxl_skin2 = p->rmaxl_skin2[k]; ip[k].vlast = p->vtab[(size_t)k * nt + nt - 4];
  }
  int ncharge = 0;
  std::vector<int> qidx, qpos(n, -1);
  for (int i = 0; i < n; i++)
    if (p->qch[i] != 0.) { qpos[i] = ncharge; qidx.push_back(i); ncharge++; }
  c->ncharge = ncharge;
  if (ncharge > 0) {
    if (!p->coultab || !p->dcoultab || !p->switchse) {
      nb_set_error("nb_create: charges present but coulomb tables missing");
      return fail(NB_ERR_ARG);
    }
    for (int t = 0; t < nt; t++) {
      ct3[(size_t)t * 3 + 0] = p->dcoultab[t];
      ct3[(size_t)t * 3 + 1] = p->switchse[t];
      ct3[(size_t)t * 3 + 2] = p->coultab[t];
    }
    c->clast = p->coultab[nt - 4];
    // vinter_e's clamp: k = 1, p = 0 - 1 (finter_e.c:325-352)
    auto at_clamp = [](const double *t) {
      const double pp = -1., fm1 = t[0], f0 = t[1], f1 = t[2];
      return f0 + .5 * pp * (f1 - fm1 + pp * (fm1 - 2. * f0 + f1));
    };
    c->cq0 = at_clamp(p->coultab);
    c->dq0 = at_clamp(p->dcoultab);
  }
  // list cutoffs per type pair: null pairs get -1 (no r^2 is below it)
  std::vector<double> pcut(2 * (size_t)p->ntypes * p->ntypes, -1.);
  for (int a = 0; a < p->ntypes; a++)
    for (int b = 0; b < p->ntypes; b++) {
      int k = p->map[a * p->ntypes + b];
      if (k < 0) continue;
      if (k >= p->ninter) { nb_set_error("nb_create: map entry %d out of range", k); return fail(NB_ERR_ARG); }
      pcut[2 * ((size_t)a * p->ntypes + b)] = p->rmaxl_skin2[k];
      pcut[2 * ((size_t)a * p->ntypes + b) + 1] = p->rmaxs_skin2[k];
    }
  // list cutoff: the longest (r+skin) of any interaction / the coulomb list
  double r2 = 0.;
  for (int k = 0; k < p->ninter; k++) r2 = std::max(r2, p->rmaxl_skin2[k]);
  if (ncharge > 0) r2 = std::max(r2, p->rmaxl_skin2_e);
  c->rlist = sqrt(r2);
  {
    double rp2 = r2;
    for (int k = 0; k < p->ninter; k++) rp2 = std::max(rp2, p->rmaxs_skin2[k]);
    if (ncharge > 0) rp2 = std::max(rp2, p->rmaxs_skin2_e);
    c->rprune2 = rp2 * (1. + 2e-6);
  }

  // exclusions, both directions: partner | bit30 (vdW) | bit31 (coulomb)
  std::vector<std::vector<unsigned>> ex(n);
  auto add_ex = [&](int a, int b, unsigned bit) {
    for (auto &v : ex[a])
      if ((int)(v & 0x3FFFFFFFu) == b) { v |= bit; return; }
    ex[a].push_back((unsigned)b | bit);
  };
  {
    size_t off = 0;
    for (int i = 0; i < n; i++) {
      int m = p->nexclude ? p->nexclude[i] : 0;
      for (int k = 0; k < m; k++) {
        int j = p->exclude[off + k];
        if (j < 0 || j >= n) { nb_set_error("nb_create: exclusion out of range"); return fail(NB_ERR_ARG); }
        add_ex(i, j, 1u << 30);
        add_ex(j, i, 1u << 30);
      }
      off += m;
    }
    off = 0;
    for (int i = 0; i < n; i++) {
      int m = p->cnexclude ? p->cnexclude[i] : 0;
      for (int k = 0; k < m; k++) {
        int j = p->cexclude[off + k];
        if (j < 0 || j >= n) { nb_set_error("nb_create: exclusion out of range"); return fail(NB_ERR_ARG); }
        add_ex(i, j, 1u << 31);
        add_ex(j, i, 1u << 31);
      }
      off += m;
    }
  }
  std::vector<int> xoff(n + 1, 0), xlist;
  for (int i = 0; i < n; i++) {
    std::sort(ex[i].begin(), ex[i].end(), [](unsigned a, unsigned b) {
      return (a & 0x3FFFFFFFu) < (b & 0x3FFFFFFFu);
    });
    for (unsigned v : ex[i]) xlist.push_back((int)v);
    xoff[i + 1] = (int)xlist.size();
  }
  // ecorr partners, both directions
  c->necorr = p->necorr;
  std::vector<std::vector<int>> ec(n);
  for (int k = 0; k < p->necorr; k++) {
    int a = p->iecorr[k], b = p->jecorr[k];
    if (a < 0 || a >= n || b < 0 || b >= n) { nb_set_error("nb_create: ecorr pair out of range"); return fail(NB_ERR_ARG); }
    ec[a].push_back(b);
    ec[b].push_back(a);
  }
  std::vector<int> eoff(n + 1, 0), elist;
  for (int i = 0; i < n; i++) {
    for (int v : ec[i]) elist.push_back(v);
    eoff[i + 1] = (int)elist.size();
  }
  // Ewald k vectors as rows (maximal runs of consecutive kc at one (ka, kb)
  // in k order), each walked by m = |kc| with kc and -kc paired
  c->ktot = (ncharge > 0) ? p->ktot : 0;
  std::vector<EwRow> row;
  std::vector<int> kabc(3 * (size_t)std::max(c->ktot, 1), 0);
  for (int k = 0; k < c->ktot; k++) {
    kabc[3 * k] = p->ka[k]; kabc[3 * k + 1] = p->kb[k]; kabc[3 * k + 2] = p->kc[k];
    if (!row.empty() && row.back().ka == p->ka[k] && row.back().kb == p->kb[k] &&
        row.back().kc_hi + 1 == p->kc[k]) {
      row.back().kc_hi++;
    } else {
      EwRow r = {k, p->kc[k], p->kc[k], p->ka[k], p->kb[k], 0, 0, 0, 0};
      row.push_back(r);
    }
  }
  c->nrow = (int)row.size();
  std::vector<int4> ent;  // per m of each row: k of +m, k of -m (-1: absent), m
  std::vector<EwSym> sym;
  for (int g = 0; g < c->nrow; g++) {
    EwRow &r = row[g];
    r.mlo = r.kc_lo > 0 ? r.kc_lo : (r.kc_hi < 0 ? -r.kc_hi : 0);
    r.mhi = std::max(std::abs(r.kc_lo), std::abs(r.kc_hi));
    r.e0 = (int)ent.size();
    r.chain = g > 0 && row[g - 1].ka == r.ka && row[g - 1].kb + 1 == r.kb;
    if (r.mhi - r.mlo + 1 > NB_EW_KT) {
      nb_set_error("nb_create: a row of %d k vectors along c exceeds %d", r.mhi - r.mlo + 1,
                   NB_EW_KT);
      return fail(NB_ERR_ARG);
    }
    for (int m = r.mlo; m <= r.mhi; m++) {
      const int kp = (m >= r.kc_lo && m <= r.kc_hi) ? r.k_lo + m - r.kc_lo : -1;
      const int km = (m > 0 && -m >= r.kc_lo && -m <= r.kc_hi) ? r.k_lo - m - r.kc_lo : -1;
      ent.push_back(make_int4(kp, km, m, 0));
    }
  }
  // structure-factor lanes.  When every row is short (<= 8 or <= 4 values
  // of m, the small k spheres of small systems), a lane takes 2 or 4 rows of
  // one ka plane with consecutive kb, m from 0 (a row whose kc range starts
  // above m = 0 only leaves its first slots unstored): the rows share the
  // c-axis recurrence and each row's phase follows from the previous row's
  // by one step of e^{i theta_b}, so the lane's sine / cosine per charge is
  // spread over 2 or 4 rows.  Otherwise each lane is a run of <=
  // NB_EW_SYM values of m of one row.  One row count for the whole system
  // keeps the lanes in as few 64-lane groups as possible.  Fewer lanes also
  // mean fewer workgroups, which a small system alone on the GPU needs to
  // fill it, so only a context that shares the GPU (a batch member) or has
  // NB_SFAC_ROWS_MIN_Q charges or more takes several rows a lane.
  {
    int mmax = 0;
    for (int g = 0; g < c->nrow; g++) mmax = std::max(mmax, row[g].mhi + 1);
    const bool full = p->sharing > 1 || ncharge >= NB_SFAC_ROWS_MIN_Q;
    const int nr = !full ? 1 : (mmax <= NB_EW_SYM / 4 ? 4 : (mmax <= NB_EW_SYM / 2 ? 2 : 1));
    c->sfac_rows = nr;
    for (int g = 0; g < c->nrow;) {
      EwSym y;
      memset(&y, 0, sizeof(y));
      y.ka = row[g].ka; y.kb = row[g].kb; y.nr = nr;
      for (int q = 0; q < 4; q++) { y.k_lo[q] = 0; y.kc_lo[q] = 1; y.kc_hi[q] = 0; }
      if (nr == 1) {
        const EwRow &r = row[g];
        y.k_lo[0] = r.k_lo; y.kc_lo[0] = r.kc_lo; y.kc_hi[0] = r.kc_hi;
        for (int m0 = r.mlo; m0 <= r.mhi; m0 += NB_EW_SYM) {
          y.m0 = m0;
          y.cnt = std::min(NB_EW_SYM, r.mhi - m0 + 1);
          sym.push_back(y);
        }
        g++;
        continue;
      }
      // rows g, g + 1, ... while they continue the plane by one kb
      int q = 0;
      for (; q < nr && g + q < c->nrow && (q == 0 || row[g + q].chain); q++) {
        const EwRow &r = row[g + q];
        y.k_lo[q] = r.k_lo; y.kc_lo[q] = r.kc_lo; y.kc_hi[q] = r.kc_hi;
        y.cnt = std::max(y.cnt, r.mhi + 1);
      }
      sym.push_back(y);
      g += q;
    }
  }
  const int nent = (int)ent.size();
  // structure factor: lanes (above) longest first (a wave's trip count is
  // its first lane's), then by m0 (lanes of one wave start their phases
  // alike); a workgroup of 4 waves takes 64
  // lanes and one chunk of charges (>= NB_SFAC_MIN_Q); chunks sized for ~NB_SFAC_WG
  // workgroups, ~16 a CU, which the long charge loops need to hide FP64
  // latency; a system sharing the GPU with others gets a proportionally
  // coarser split
  const int share = std::max(1, p->sharing);
  const int sfac_wg = std::max(1, NB_SFAC_WG / share);
  const int ewf_wg = std::max(1, NB_EWF_WG / share);
  std::stable_sort(sym.begin(), sym.end(), [](const EwSym &a, const EwSym &b) {
    return a.cnt != b.cnt ? a.cnt > b.cnt : a.m0 < b.m0;
  });
  c->nsym = (int)sym.size();
  c->nsw = std::max(1, (c->nsym + NB_WAVE - 1) / NB_WAVE);
  {
    int maxch = std::max(1, (ncharge + NB_SFAC_MIN_Q - 1) / NB_SFAC_MIN_Q);  // charges a workgroup
    int nch = std::max(1, (sfac_wg + c->nsw - 1) / c->nsw);
    // in a batch, workgroups of every system should take about the same
    // time: cap a workgroup's k x charge products
    if (share > 1)
      nch = std::max(nch, (int)(((long long)c->ktot * ncharge + (long long)NB_SFAC_WORK * c->nsw - 1) /
                                ((long long)NB_SFAC_WORK * c->nsw)));
    nch = std::min(nch, maxch);
    c->qc = std::max(1, (ncharge + nch - 1) / nch);
    c->nchunk = std::max(1, (ncharge + c->qc - 1) / c->qc);
  }
  // The structure factor as matrix products on the FP64 matrix cores
  // (k_ewsfacm), for systems whose lanes hold one row: per ka plane it is
  // [Re w; Im w] (rows x charges) times [cos m theta_c, sin m theta_c]
  // (charges x m), with w = q e^{i (ka theta_a + kb theta_b)} from per-charge
  // phase tables, so neither the per-lane sine / cosine nor the per-entry
  // recurrence of the VALU kernel remains.  Rows in tiles of 8; a wave
  // takes NB_SFM_R tiles and every charge of its chunk.  CM3D_NB_SFAC_MFMA=0 / 1 turns it off / on for any system it
  // can serve (for tests and A/B timing).
  std::vector<EwRowM> rowm;
  std::vector<EwSfmWave> sfmw;
  {
    int mmax = -1, kamin = 0, kamax = 0, kbmax = 0;
    for (const EwRow &r : row) {
      mmax = std::max(mmax, r.mhi);
      kamin = std::min(kamin, r.ka); kamax = std::max(kamax, r.ka);
      kbmax = std::max(kbmax, std::abs(r.kb));
    }
    const bool can = c->sfac_rows == 1 && c->nrow > 0 && ncharge > 0 && kamin >= 0 &&
                     mmax < 8 * NB_SFM_NT && kamax < NB_SFM_TAB && kbmax < NB_SFM_TAB;
    bool use = can && (long long)ncharge * c->ktot >= NB_SFM_MIN_WORK;
    if (const char *ev = getenv("CM3D_NB_SFAC_MFMA")) use = can && ev[0] != '0';
    c->sfm = use;
    if (use) {
      // tiles: runs of <= 8 rows, each row the previous one's (ka, kb + 1),
      // so that a tile's 8 rows read one a-table entry and 8 consecutive
      // b-table entries of a charge (no LDS bank conflicts); tiles sorted by
      // their longest m range, longest first
      std::vector<std::vector<int>> tl;
      for (int g = 0; g < c->nrow; g++) {
        if (tl.empty() || !row[g].chain || tl.back().size() == 8) tl.emplace_back();
        tl.back().push_back(g);
      }
      std::vector<int> tmhi(tl.size());
      for (size_t t = 0; t < tl.size(); t++) {
        tmhi[t] = -1;
        for (int g : tl[t]) tmhi[t] = std::max(tmhi[t], row[g].mhi);
      }
      std::vector<int> ord(tl.size());
      for (size_t t = 0; t < tl.size(); t++) ord[t] = (int)t;
      std::stable_sort(ord.begin(), ord.end(),
                       [&](int a, int b) { return (tmhi[a] >> 3) > (tmhi[b] >> 3); });
      // waves: NB_SFM_R tiles each, fewer where the first (longest) tile
      // needs 4 or 5 m tiles (the accumulators then fit the same
      // registers); the empty tile past the last real one fills the rest
      const int ntile = (int)tl.size();
      for (int t = 0; t < ntile;) {
        EwSfmWave w = {{ntile, ntile, ntile, ntile}, (tmhi[ord[t]] + 8) >> 3, 0, 0, 0};
        w.nr = nb_sfm_wave_tiles(w.nt);
        for (int q = 0; q < w.nr && t < ntile; q++, t++) w.t[q] = t;
        sfmw.push_back(w);
      }
      while (sfmw.size() % 4) sfmw.push_back(EwSfmWave{{ntile, ntile, ntile, ntile}, 0, NB_SFM_R, 0, 0});
      c->sfm_nrwg = (int)sfmw.size() / 4;
      rowm.assign((size_t)(ntile + 1) * 8, EwRowM{0, 0, 0, 1, 0, -1, -1, 0});  // no k
      for (int t = 0; t < ntile; t++) {
        const std::vector<int> &rs = tl[ord[t]];
        for (size_t q = 0; q < rs.size(); q++) {
          const EwRow &r = row[rs[q]];
          rowm[(size_t)t * 8 + q] = EwRowM{r.ka, r.kb, r.k_lo, r.kc_lo, r.kc_hi, r.mhi,
                                           tmhi[ord[t]], 0};
        }
      }
      c->sfm_la = kamax + 1;
      c->sfm_lb = kbmax + 1;
      c->sfm_lz = 8 * ((mmax + 8) / 8);
      // chunks of whole staged tiles for ~NB_SFM_WG / share workgroups
      const int nq16 = (ncharge + NB_SFM_NQ - 1) / NB_SFM_NQ;
      int nch = std::max(1, std::max(1, NB_SFM_WG / share) / c->sfm_nrwg);
      nch = std::min(nch, std::max(1, nq16 / 4));  // at least 4 tiles a chunk
      c->qc = NB_SFM_NQ * ((nq16 + nch - 1) / nch);
      c->nchunk = std::max(1, (ncharge + c->qc - 1) / c->qc);
      // matrix-core instructions a launch issues: each wave R x NT a K step
      // of 4 charges, for every 16-charge tile of its chunk
      long long per_step = 0, steps = 0;
      for (const EwSfmWave &w : sfmw) per_step += (long long)w.nr * w.nt;
      for (int ch = 0; ch < c->nchunk; ch++) {
        const int nc = std::min(ncharge, (ch + 1) * c->qc) - ch * c->qc;
        steps += 4LL * ((nc + NB_SFM_NQ - 1) / NB_SFM_NQ);
      }
      c->sfm_mfma = per_step * steps;
    }
  }
  // force: NB_EWF_Q charges per lane, a workgroup per (256 NB_EWF_Q
  // charges, tile range); tiles are runs of whole rows small enough for LDS
  // and for enough ranges to give ~NB_EWF_WG workgroups (each range's
  // partial forces are 24 B a charge to write and to read back in k_ewfin);
  // ranges balanced by entry count
  c->nqb = std::max(1, (ncharge + 256 * NB_EWF_Q - 1) / (256 * NB_EWF_Q));
  std::vector<int> trow(1, 0), rtile(1, 0);
  {
    int nrt = std::max(1, ewf_wg / c->nqb);  // at most ~NB_EWF_WG: one round of waves
    if (share > 1)  // as for the structure factor: a bounded k x charge product per workgroup
      nrt = std::max(nrt, (int)(((long long)c->ktot * 256 * NB_EWF_Q + NB_EWF_WORK - 1) / NB_EWF_WORK));
    // each range's partial forces cost 24 B a charge each way: no range of
    // fewer than NB_EWF_MIN_ENT entries
    nrt = std::min(nrt, std::max(1, std::min(c->nrow, nent / NB_EWF_MIN_ENT)));
    int rmax = 1;
    for (const EwRow &r : row) rmax = std::max(rmax, r.mhi - r.mlo + 1);
    const int lim = std::min(NB_EW_KT, std::max(rmax, (nent + nrt - 1) / std::max(1, nrt)));
    int ne = 0;
    std::vector<int> te;  // entries per tile
    for (int g = 0; g < c->nrow; g++) {
      const int cnt = row[g].mhi - row[g].mlo + 1;
      if (ne > 0 && ne + cnt > lim) { trow.push_back(g); te.push_back(ne); ne = 0; }
      ne += cnt;
    }
    trow.push_back(c->nrow);
    te.push_back(ne);
    c->ntile = (int)trow.size() - 1;
    // close range r after the tile where the running entry count reaches
    // r/nr of the total: every range gets at least one tile, and uneven tiles
    // only ever give fewer ranges, never one range for everything
    const int nr = std::min(nrt, c->ntile);
    long long cum = 0;
    for (int t = 0; t + 1 < c->ntile; t++) {
      cum += te[t];
      const int r = (int)rtile.size();
      if (r < nr && cum * nr >= (long long)r * nent) rtile.push_back(t + 1);
    }
    rtile.push_back(c->ntile);
    c->nrange = (int)rtile.size() - 1;
  }
  // The reciprocal force as matrix products (k_ewforcem), for contexts
  // that share the GPU and systems of NB_EFM_MIN_WORK charge x k products
  // or more (CM3D_NB_EWF_MFMA=0 / 1 turns it off / on for any system it can
  // serve; the structure factor's setting when only that one is given):
  // per 16-row tile and 16
  // charges, P and Q of k_ewforce for every (row, charge) from four
  // products per 2 values of m, the rows as 4 runs of 4 consecutive kb
  // (run q's row r is tile row q + 4 r, so that a lane's 4 rows of the
  // result are one run: one sine / cosine per run and charge)
  std::vector<EwFTile> ftile;
  std::vector<int4> ftent;
  std::vector<int> frtile(1, 0);
  {
    int mmax = -1;
    for (const EwRow &r : row) mmax = std::max(mmax, r.mhi);
    const bool can = c->nrow > 0 && ncharge > 0 && mmax < NB_EFM_MS;
    bool use = can && (share > 1 || (long long)ncharge * c->ktot >= NB_EFM_MIN_WORK);
    if (const char *ev = getenv("CM3D_NB_SFAC_MFMA")) use = can && ev[0] != '0';
    if (const char *ev = getenv("CM3D_NB_EWF_MFMA")) use = can && ev[0] != '0';
    c->efm = use;
    if (use) {
      std::vector<std::vector<int>> runs;
      for (int g = 0; g < c->nrow; g++) {
        if (runs.empty() || !row[g].chain || runs.back().size() == 4) runs.emplace_back();
        runs.back().push_back(g);
      }
      std::vector<int> rmhi(runs.size()), ord(runs.size());
      for (size_t q = 0; q < runs.size(); q++) {
        rmhi[q] = -1;
        for (int g : runs[q]) rmhi[q] = std::max(rmhi[q], row[g].mhi);
        ord[q] = (int)q;
      }
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return rmhi[a] > rmhi[b]; });
      std::vector<long long> work;
      for (size_t q0 = 0; q0 < ord.size(); q0 += 4) {
        EwFTile t;
        memset(&t, 0, sizeof(t));
        int mt0 = 1 << 30, mt1 = -1;
        for (int q = 0; q < 4 && q0 + q < ord.size(); q++) {
          const std::vector<int> &rs = runs[ord[q0 + q]];
          t.ka[q] = row[rs[0]].ka; t.kb0[q] = row[rs[0]].kb;
          for (int g : rs) { mt0 = std::min(mt0, row[g].mlo); mt1 = std::max(mt1, row[g].mhi); }
        }
        t.mt0 = mt0;
        t.nks = (mt1 - mt0 + 2) / 2;
        t.ent_off = (int)ftent.size();
        for (int mm = 0; mm < 2 * t.nks; mm++) {
          const int m = mt0 + mm;
          for (int i = 0; i < 16; i++) {
            int4 e = make_int4(-1, -1, m, 0);
            const int q = i % 4, rr = i / 4;
            if (q0 + q < ord.size() && rr < (int)runs[ord[q0 + q]].size()) {
              const EwRow &r = row[runs[ord[q0 + q]][rr]];
              if (m >= r.mlo && m <= r.mhi) {
                e.x = (m >= r.kc_lo && m <= r.kc_hi) ? r.k_lo + m - r.kc_lo : -1;
                e.y = (m > 0 && -m >= r.kc_lo && -m <= r.kc_hi) ? r.k_lo - m - r.kc_lo : -1;
              }
            }
            ftent.push_back(e);
          }
        }
        ftile.push_back(t);
        work.push_back(t.nks);
      }
      // charge blocks of 4 waves x NB_EFM_CT x 16 charges; ranges of tiles
      // of about equal work for ~NB_EWF_WG / share workgroups in all
      c->efm_ncb = (ncharge + 64 * NB_EFM_CT - 1) / (64 * NB_EFM_CT);
      const int nt = (int)ftile.size();
      const int nr = std::max(1, std::min(nt, std::max(1, ewf_wg / c->efm_ncb)));
      long long tot = 0, cum = 0;
      for (long long w : work) tot += w;
      for (int t = 0; t + 1 < nt; t++) {
        cum += work[t];
        const int r = (int)frtile.size();
        if (r < nr && cum * nr >= (long long)r * tot) frtile.push_back(t + 1);
      }
      frtile.push_back(nt);
      c->nrange = (int)frtile.size() - 1;
      c->efm_nent = (int)ftent.size();
      // each wave: NB_EFM_CT charge tiles x 4 products a K step of every tile
      c->efm_mfma = (long long)c->efm_ncb * 4 * NB_EFM_CT * 4 * tot;
    }
  }

  // device uploads
  if ((rc = upload(c, &c->d_itype, p->itype, n))) return fail(rc);
  if ((rc = upload(c, &c->d_map, p->map, (size_t)p->ntypes * p->ntypes))) return fail(rc);
  if ((rc = upload(c, &c->d_ip, ip.data(), ni))) return fail(rc);
  if ((rc = upload(c, &c->d_pcut, pcut.data(), pcut.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_vtab3, vt3.data(), vt3.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_ctab3, ct3.data(), ct3.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_qch, p->qch, n))) return fail(rc);
  if ((rc = upload(c, &c->d_xoff, xoff.data(), n + 1))) return fail(rc);
  if ((rc = upload(c, &c->d_xlist, xlist.data(), xlist.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_eoff, eoff.data(), n + 1))) return fail(rc);
  if ((rc = upload(c, &c->d_elist, elist.data(), elist.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_qidx, qidx.data(), qidx.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_qpos, qpos.data(), qpos.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_row, row.data(), row.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_sym, sym.data(), sym.size()))) return fail(rc);
  if (c->sfm && (rc = upload(c, &c->d_rowm, rowm.data(), rowm.size()))) return fail(rc);
  if (c->sfm && (rc = upload(c, &c->d_sfmw, sfmw.data(), sfmw.size()))) return fail(rc);
  if (c->efm && (rc = upload(c, &c->d_ftile, ftile.data(), ftile.size()))) return fail(rc);
  if (c->efm && (rc = upload(c, &c->d_ftent, ftent.data(), ftent.size()))) return fail(rc);
  if (c->efm && (rc = upload(c, &c->d_frtile, frtile.data(), frtile.size()))) return fail(rc);
  if (c->efm && (rc = dalloc(c, &c->d_fimg, 8 * (size_t)c->efm_nent))) return fail(rc);
  if (c->sfm && (rc = dalloc(c, &c->d_sfmtab, nb_sfm_table_doubles(c)))) return fail(rc);
  if ((rc = upload(c, &c->d_ent, ent.data(), ent.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_trow, trow.data(), trow.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_rtile, rtile.data(), rtile.size()))) return fail(rc);
  if ((rc = upload(c, &c->d_kabc, kabc.data(), kabc.size()))) return fail(rc);
  if ((rc = dalloc(c, &c->d_pos, 3 * (size_t)n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_posb, 3 * (size_t)n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_fs, 3 * (size_t)n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_fl, 3 * (size_t)n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_perm, n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_iperm, n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_cell, n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_tmp, n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_rec64, 4 * (size_t)n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_stype, n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_nnb, n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_hcnt, 4 * (size_t)n))) return fail(rc);
  if ((rc = dalloc(c, &c->d_flags, 8))) return fail(rc);
  if ((rc = dalloc(c, &c->d_chg, 4 * (size_t)std::max(ncharge, 1)))) return fail(rc);
  if ((rc = dalloc(c, &c->d_stp, 4 * (size_t)std::max(ncharge, 1)))) return fail(rc);
  if ((rc = dalloc(c, &c->d_spart, 2 * (size_t)c->nchunk * std::max(c->ktot, 1)))) return fail(rc);
  if ((rc = dalloc(c, (char **)&c->d_kd, nb_kdata_bytes() * std::max(c->ktot, 1)))) return fail(rc);
  if ((rc = dalloc(c, &c->d_fpart, 3 * (size_t)c->nrange * std::max(ncharge, 1)))) return fail(rc);
  // pair kernels: a wave per atom in turn; the FP32 kernel takes 8 atoms a
  // workgroup (2 a wave: enough waves in flight to cover its gathers), 64
  // in a context that shares the GPU (measured on the 64-system batch:
  // 0.68 ms a step against 0.77 with 16 and 0.88 with 8)
  const int apw = share >= 4 ? NB_MIX_ATOMS_WG_SHARED : NB_MIX_ATOMS_WG;
  c->npair_blocks = p->precision == NB_PREC_MIXED ? std::max(1, std::min((n + apw - 1) / apw, NB_MIX_MAX_WG))
                                                  : std::min((n + 3) / 4, NB_PAIR64_MAX_WG);
  c->new_blocks = std::max(1, (c->ktot + NB_WAVE - 1) / NB_WAVE);
  c->nfin_blocks = std::max(1, (ncharge + NB_WAVE - 1) / NB_WAVE);  // 64 charges a workgroup
  if ((rc = dalloc(c, &c->d_epair, (size_t)c->npair_blocks * NB_NE_PAIRQ))) return fail(rc);
  if ((rc = dalloc(c, &c->d_eec, (size_t)c->nfin_blocks * NB_NE_ECORR))) return fail(rc);
  if ((rc = dalloc(c, &c->d_eew, (size_t)c->new_blocks * NB_NE_EW))) return fail(rc);
  if ((rc = dalloc(c, &c->d_eout, 64))) return fail(rc);
  if (hipHostMalloc((void **)&c->h_flags, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    nb_set_error("hipHostMalloc failed");
    return fail(NB_ERR_HIP);
  }
  // tables in LDS when they fit the 64 KiB a workgroup may claim
  size_t tab_bytes = (vt3.size() + ct3.size()) * sizeof(double);
  c->ctab_off = (int)vt3.size();
  c->tab_lds_bytes = tab_bytes <= 60 * 1024 ? (int)tab_bytes : 0;
  c->tab_lds_bytes_f = tab_bytes / 2 <= 60 * 1024 ? (int)(tab_bytes / 2) : 0;
  // force-column coefficients of the forces-only kernels: node t holds
  // {f_t, (f_t+1 - f_t-1)/2, (f_t-1 - 2 f_t + f_t+1)/2}, so cm3d's quadratic
  // interpolation f0 + .5 p (f1 - f-1 + p (f-1 - 2 f0 + f1)) is
  // c0 + p (c1 + p c2); the differences are formed in cm3d's order, and only
  // nodes 1 .. ntable-2 (those a pair can select) are filled
  {
    const int ntab = ni + 1;
    std::vector<double> fc((size_t)ntab * nt * 3, 0.);
    for (int k = 0; k < ntab; k++) {
      const double *col = k < ni ? vt3.data() + (size_t)k * nt * 3 : ct3.data();
      for (int t = 1; t + 1 < nt; t++) {
        const double fm1 = col[3 * (t - 1)], f0 = col[3 * t], f1 = col[3 * (t + 1)];
        double *o = fc.data() + ((size_t)k * nt + t) * 3;
        o[0] = f0;
        o[1] = .5 * (f1 - fm1);
        o[2] = .5 * (fm1 - 2. * f0 + f1);
      }
    }
    if ((rc = upload(c, &c->d_fcoef, fc.data(), fc.size()))) return fail(rc);
    const size_t fb = fc.size() * sizeof(double);
    c->fcoef_lds = fb <= 60 * 1024 ? (int)fb : 0;
    if (p->precision == NB_PREC_MIXED) {
      std::vector<float4> ff((size_t)ntab * nt);
      for (size_t t = 0; t < ff.size(); t++)
        ff[t] = make_float4((float)fc[3 * t], (float)fc[3 * t + 1], (float)fc[3 * t + 2], 0.f);
      if ((rc = upload(c, &c->d_fcoeff, ff.data(), ff.size()))) return fail(rc);
      const size_t fbf = ff.size() * sizeof(float4);
      c->fcoef_lds_f = fbf <= 60 * 1024 ? (int)fbf : 0;
    }
  }
  if (c->prec == NB_PREC_MIXED) {
    std::vector<float> vf(vt3.begin(), vt3.end()), cf(ct3.begin(), ct3.end());
    if ((rc = upload(c, &c->d_vtab3f, vf.data(), vf.size()))) return fail(rc);
    if ((rc = upload(c, &c->d_ctab3f, cf.data(), cf.size()))) return fail(rc);
    if ((rc = dalloc(c, &c->d_recf, (size_t)n))) return fail(rc);
    std::vector<TPairF> tpf((size_t)p->ntypes * p->ntypes);
    for (int a = 0; a < p->ntypes; a++)
      for (int b = 0; b < p->ntypes; b++) {
        TPairF &t = tpf[(size_t)a * p->ntypes + b];
        const int k = p->map[a * p->ntypes + b];
        if (k < 0) {
          t.rmins2 = t.rminl2 = t.rmaxs2 = t.rmaxl2 = -1.f; t.dx2 = 0.f; t.vlast = 0.f;
          t.in = -1;
        } else {
          t.rmins2 = (float)ip[k].rmins2; t.rminl2 = (float)ip[k].rminl2;
          t.rmaxs2 = (float)ip[k].rmaxs2; t.rmaxl2 = (float)ip[k].rmaxl2;
          t.dx2 = (float)ip[k].dx2; t.vlast = (float)ip[k].vlast; t.in = k;
        }
        t.pad = 0;
      }
    if ((rc = upload(c, &c->d_tpf, tpf.data(), tpf.size()))) return fail(rc);
  }
  // neighbour capacity from the density, grown on overflow at build time
  {
    double vol_per_atom = 1e300;
    (void)vol_per_atom;
    c->cap = 0;
  }
  HCHK(hipMemset(c->d_eout, 0, 64 * sizeof(double)));
  HCHK(hipMemset(c->d_flags, 0, 8 * sizeof(int)));
  {
    char *sp;
    if ((rc = scratch_ptr(c->dev, &sp))) return fail(rc);
  }
  *out = c;
  return NB_OK;
}

// Grid of cells for the list stencil, per box axis: the perpendicular width
// of the box along a is V/|b x c| (box vectors are hmat's columns).
static void cell_grid(nb_ctx *c, const Box &b, int *nx, int *ny, int *nz) {
  const double *h = b.h;
  double a[3] = {h[0], h[3], h[6]}, bb[3] = {h[1], h[4], h[7]}, cc[3] = {h[2], h[5], h[8]};
  auto cross = [](const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  double w[3], vol = fabs(det9(h));
  double widths[3];
  cross(bb, cc, w); widths[0] = vol / sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  cross(cc, a, w);  widths[1] = vol / sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  cross(a, bb, w);  widths[2] = vol / sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  int ncs[3], reach[3];
  // third-cutoff cells only where they hold a few atoms each on average:
  // sparser systems pay more in per-cell work than they save in candidates
  bool third = NB_CELL_REACH_MAX >= 3 && c->rlist > 0;
  double cells3 = 1.;
  for (int d = 0; d < 3 && third; d++) {
    const int m3 = (int)floor(3. * widths[d] / (c->rlist * (1. + 1e-9)));
    third = m3 >= 7;
    cells3 *= std::min(m3, 128);
  }
  third = third && c->natoms >= NB_CELL3_MIN_ATOMS * cells3;
  for (int d = 0; d < 3; d++) {
    double r = c->rlist > 0 ? c->rlist * (1. + 1e-9) : widths[d];
    // third-cutoff cells with a reach of 3 when at least 7 fit on every
    // axis (so the 7 cells around are distinct; 7 x 7 columns still fit one
    // wavefront's lanes) and the cells are dense enough (above), else
    // half-cutoff cells with a reach of 2 when at least 5 fit,
    // else cutoff cells with a reach of 1 when at least 3 fit, else the
    // whole axis as one cell.  Finer cells hug the list sphere more closely:
    // fewer candidates per listed pair.
    int m3 = (int)floor(3. * widths[d] / r);
    int m2 = (int)floor(2. * widths[d] / r), m1 = (int)floor(widths[d] / r);
    if (third) {
      ncs[d] = std::min(m3, 128); reach[d] = 3;
    } else if (m2 >= 5) {
      ncs[d] = std::min(m2, 128); reach[d] = 2;
    } else if (m1 >= 3) {
      ncs[d] = m1; reach[d] = 1;
    } else {
      ncs[d] = 1; reach[d] = 0;
    }
  }
  *nx = ncs[0]; *ny = ncs[1]; *nz = ncs[2];
  c->crx = reach[0]; c->cry = reach[1]; c->crz = reach[2];
}

static int ensure_capacity(nb_ctx *c, const Box &b, int want_cap) {
  if (want_cap <= c->cap && c->d_nbr) return NB_OK;
  if (c->d_nbr) {
    // a recorded step of this context may still be running on any stream
    HCHK(hipDeviceSynchronize());
    hipFree(c->d_nbr);
    c->d_nbr = nullptr;
    c->dev_bytes -= ((size_t)c->cap * c->natoms + NB_LIST_PAD) * sizeof(uint32_t);
  }
  c->cap = (want_cap + 63) & ~63;
  c->gen++;
  return dalloc(c, &c->d_nbr, (size_t)c->cap * c->natoms + NB_LIST_PAD);
}

int nb_grow_capacity(nb_ctx *c, int want) {
  Box b;
  memset(&b, 0, sizeof(b));
  return ensure_capacity(c, b, want);
}

extern "C" int nb_list_reserve(nb_ctx *c, int slots) {
  if (!c || slots <= 0) { nb_set_error("nb_list_reserve: bad argument"); return NB_ERR_ARG; }
  HCHK(hipSetDevice(c->dev));
  HCHK(hipDeviceSynchronize());  // recorded steps may still read the slots
  if (c->d_nbr) {
    hipFree(c->d_nbr);
    c->d_nbr = nullptr;
    c->dev_bytes -= ((size_t)c->cap * c->natoms + NB_LIST_PAD) * sizeof(uint32_t);
  }
  c->cap = (slots + 63) & ~63;
  c->gen++;
  RC(dalloc(c, &c->d_nbr, (size_t)c->cap * c->natoms + NB_LIST_PAD));
  // the old rows are gone: empty rows until the next build, which the
  // rebuild flag makes the next device step do first
  HCHK(hipMemset(c->d_nnb, 0, sizeof(int) * c->natoms));
  const int one = 1;
  HCHK(hipMemcpy(c->d_flags + 1, &one, sizeof(int), hipMemcpyHostToDevice));
  c->have_lists = 0;
  return NB_OK;
}

int nb_build_now(nb_ctx *c, const double hmat[9], hipStream_t s) {
  Box b = make_box(hmat);
  int nx, ny, nz;
  const int rx = c->crx, ry = c->cry, rz = c->crz;
  cell_grid(c, b, &nx, &ny, &nz);
  if (nx != c->ncx || ny != c->ncy || nz != c->ncz || rx != c->crx || ry != c->cry ||
      rz != c->crz)
    c->gen++;
  c->ncx = nx; c->ncy = ny; c->ncz = nz;
  c->ncells = nx * ny * nz;
  if (c->ncells > c->cells_alloc) {
    if (c->d_ccount) HCHK(hipDeviceSynchronize());
    if (c->d_ccount) c->dev_bytes -= 4 * (size_t)c->cells_alloc * sizeof(int);
    if (c->d_ccount) hipFree(c->d_ccount);
    if (c->d_cstart) hipFree(c->d_cstart);
    if (c->d_cfill) hipFree(c->d_cfill);
    if (c->d_cwork) hipFree(c->d_cwork);
    c->d_ccount = c->d_cstart = c->d_cfill = c->d_cwork = nullptr;
    RC(dalloc(c, &c->d_ccount, c->ncells));
    RC(dalloc(c, &c->d_cstart, c->ncells));
    RC(dalloc(c, &c->d_cfill, c->ncells));
    RC(dalloc(c, &c->d_cwork, c->ncells));
    HCHK(hipMemsetAsync(c->d_cwork, 0, c->ncells * sizeof(int), s));
    c->cells_alloc = c->ncells;
    c->gen++;
  }
  {
    // expected full-list length from the mean density: the initial slots
    // (with a margin) and the mixed kernel's lanes per atom
    double vol = fabs(det9(hmat));
    double nexp = 4.18879 * c->rlist * c->rlist * c->rlist * c->natoms / vol;
    const int g16 = nexp < NB_MIX_G16_MAX;
    if (g16 != c->mix_g16) c->gen++;
    c->mix_g16 = g16;
    if (c->cap == 0) {
      int want = (int)(1.6 * nexp) + 64;
      if (want > c->natoms) want = c->natoms + 63;
      RC(ensure_capacity(c, b, want));
    }
  }
  // a host build grows the slots itself: start the sticky overflow record
  // afresh (device-side builds from here on are what it reports)
  HCHK(hipMemsetAsync(c->d_flags + 2, 0, sizeof(int), s));
  for (int attempt = 0; attempt < 3; attempt++) {
    RC(nb_launch_build(c, b, s));
    HCHK(hipMemcpyAsync(c->h_flags, c->d_flags, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    if (c->h_flags[0] <= c->cap) {
      HCHK(hipMemsetAsync(c->d_flags + 2, 0, sizeof(int), s));
      c->box_built = b;
      c->have_lists = 1;
      return NB_OK;
    }
    RC(ensure_capacity(c, b, (int)(c->h_flags[0] * 1.25) + 64));
  }
  nb_set_error("neighbour list capacity could not be satisfied");
  return NB_ERR_CAPACITY;
}

extern "C" int nb_build_lists(nb_ctx *c, const double *pxyz, const double hmat[9]) {
  if (!c || !pxyz || !hmat) { nb_set_error("nb_build_lists: null argument"); return NB_ERR_ARG; }
  HCHK(hipSetDevice(c->dev));
  HCHK(hipMemcpyAsync(c->d_pos, pxyz, sizeof(double) * 3 * c->natoms, hipMemcpyHostToDevice,
                      c->stream));
  return nb_build_now(c, hmat, c->stream);
}

extern "C" int nb_build_lists_device(nb_ctx *c, const double *d_pxyz, const double hmat[9],
                                     void *stream) {
  if (!c || !d_pxyz || !hmat) { nb_set_error("nb_build_lists_device: null argument"); return NB_ERR_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  HCHK(hipSetDevice(c->dev));
  if (d_pxyz != c->d_pos)
    HCHK(hipMemcpyAsync(c->d_pos, d_pxyz, sizeof(double) * 3 * c->natoms,
                        hipMemcpyDeviceToDevice, s));
  return nb_build_now(c, hmat, s);
}

extern "C" int nb_need_rebuild(nb_ctx *c, const double *pxyz, int *flag) {
  if (!c || !pxyz || !flag) { nb_set_error("nb_need_rebuild: null argument"); return NB_ERR_ARG; }
  if (!c->have_lists) { *flag = 1; return NB_OK; }
  HCHK(hipSetDevice(c->dev));
  // positions go to the scratch force buffer so the build copy stays intact
  HCHK(hipMemcpyAsync(c->d_fs, pxyz, sizeof(double) * 3 * c->natoms, hipMemcpyHostToDevice,
                      c->stream));
  RC(nb_launch_rebuild_check(c, c->d_fs, c->stream));
  HCHK(hipMemcpyAsync(c->h_flags + 1, c->d_flags + 1, sizeof(int), hipMemcpyDeviceToHost,
                      c->stream));
  HCHK(hipStreamSynchronize(c->stream));
  *flag = c->h_flags[1] ? 1 : 0;
  return NB_OK;
}

void nb_fill_energy(nb_ctx *c, int iflg, const double *E, const double hmat[9],
                    nb_energy *e, bool cm3d) {
  memset(e, 0, sizeof(*e));
  e->nclamp = (int)(E[NB_NE_PAIR - 1] * .5 + .25);
  if (iflg != 1) return;
  const double hf = .5;
  e->spotl = hf * E[0]; e->spothl = hf * E[1]; e->sprsl = hf * E[2];
  e->sprslt[0] = hf * E[3]; e->sprslt[1] = hf * E[4]; e->sprslt[2] = hf * E[5];
  e->sprslt[4] = hf * E[6]; e->sprslt[5] = hf * E[7]; e->sprslt[8] = hf * E[8];
  e->sprslt[3] = e->sprslt[1]; e->sprslt[6] = e->sprslt[2]; e->sprslt[7] = e->sprslt[5];
  e->spote = hf * E[9]; e->spothe = hf * E[10]; e->sprse = hf * E[11];
  e->sprset[0] = hf * E[12]; e->sprset[1] = hf * E[13]; e->sprset[2] = hf * E[14];
  e->sprset[4] = hf * E[15]; e->sprset[5] = hf * E[16]; e->sprset[8] = hf * E[17];
  e->sprset[3] = e->sprset[1]; e->sprset[6] = e->sprset[2]; e->sprset[7] = e->sprset[5];
  e->spote_phys = e->spote; e->spothe_phys = e->spothe; e->sprse_phys = e->sprse;
  for (int k = 0; k < 9; k++) e->sprset_phys[k] = e->sprset[k];
  if (cm3d && c->ncharge > 0) {
    // getvirter_e -> vinter_e (finter_e.c:260-352): every te pair inside the
    // cutoff takes the table at its clamp, so the sums are closed forms in
    // sum q_i q_j (E[19]) and sum q_i q_j r r (E[20..26])
    const double sq = hf * E[19];
    e->spote = c->cq0 * sq;
    e->spothe = (c->cq0 - c->clast) * sq;
    e->sprse = -c->dq0 * hf * E[20];
    const int map6[6] = {0, 1, 2, 4, 5, 8};
    for (int k = 0; k < 9; k++) e->sprset[k] = 0.;
    for (int k = 0; k < 6; k++) e->sprset[map6[k]] = -c->dq0 * hf * E[21 + k];
    e->sprset[3] = e->sprset[1]; e->sprset[6] = e->sprset[2]; e->sprset[7] = e->sprset[5];
  }
  if (c->iperd == 3 && c->necorr > 0) {
    const double *X = E + NB_EOUT_ECORR;
    e->pot_ecorr = hf * X[0];
    e->prs_ecorr = hf * X[1];
    for (int k = 0; k < 9; k++) e->prs_ecorr_tensor[k] = hf * X[2 + k];
  }
  if (c->iperd == 3 && c->ktot > 0) {
    const double *W = E + NB_EOUT_EW;
    double bk = c->ewald_bkg / det9(hmat);
    e->vk_ewald = W[0] + bk + c->ewald_self;
    e->vp_ewald = W[1] + 3. * (W[0] + bk);
    for (int k = 0; k < 9; k++) e->vp_ewald_tensor[k] = W[2 + k];
    e->vp_ewald_tensor[0] += W[0] + bk;
    e->vp_ewald_tensor[4] += W[0] + bk;
    e->vp_ewald_tensor[8] += W[0] + bk;
  } else if (c->iperd == 3) {
    // cm3d still runs fk_ewald with no k vectors: background + self only
    double bk = c->ewald_bkg / det9(hmat);
    e->vk_ewald = bk + c->ewald_self;
    e->vp_ewald = 3. * bk;
    e->vp_ewald_tensor[0] = e->vp_ewald_tensor[4] = e->vp_ewald_tensor[8] = bk;
  }
}

// want_e: the caller reads the energies (else a forces-only evaluation)
static int force_on(nb_ctx *c, int iflg, const double hmat[9], double *d_fs, double *d_fl,
                    nb_energy *e, hipStream_t s, bool want_e) {
  if (iflg != 0 && iflg != 1) { nb_set_error("nb_force: iflg must be 0 or 1"); return NB_ERR_ARG; }
  if (!c->have_lists) { nb_set_error("nb_force: lists not built (call nb_build_lists)"); return NB_ERR_ARG; }
  Box b = make_box(hmat);
  RC(nb_launch_force(c, iflg, b, d_fs, d_fl, s, want_e ? 0 : NB_FORCE_NOENER));
  if (e) {
    double E[64];
    HCHK(hipMemcpyAsync(E, c->d_eout, sizeof(E), hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    nb_fill_energy(c, iflg, E, hmat, e);
  }
  return NB_OK;
}

extern "C" int nb_force(nb_ctx *c, int iflg, const double *pxyz, const double hmat[9],
                        double *fs, double *fl, nb_energy *e) {
  if (!c || !pxyz || !hmat || !fs || (iflg == 1 && !fl)) {
    nb_set_error("nb_force: null argument");
    return NB_ERR_ARG;
  }
  HCHK(hipSetDevice(c->dev));
  const size_t bytes = sizeof(double) * 3 * c->natoms;
  HCHK(hipMemcpyAsync(c->d_pos, pxyz, bytes, hipMemcpyHostToDevice, c->stream));
  RC(force_on(c, iflg, hmat, c->d_fs, iflg ? c->d_fl : nullptr, nullptr, c->stream, e != nullptr));
  HCHK(hipMemcpyAsync(fs, c->d_fs, bytes, hipMemcpyDeviceToHost, c->stream));
  if (iflg == 1) HCHK(hipMemcpyAsync(fl, c->d_fl, bytes, hipMemcpyDeviceToHost, c->stream));
  if (e) {
    double E[64];
    HCHK(hipMemcpyAsync(E, c->d_eout, sizeof(E), hipMemcpyDeviceToHost, c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    nb_fill_energy(c, iflg, E, hmat, e);
  } else {
    HCHK(hipStreamSynchronize(c->stream));
  }
  return NB_OK;
}

extern "C" int nb_force_device(nb_ctx *c, int iflg, const double *d_pxyz, const double hmat[9],
                               double *d_fs, double *d_fl, nb_energy *e, void *stream) {
  if (!c || !d_pxyz || !hmat || (iflg == 1 && !d_fl) || (iflg == 0 && !d_fs)) {
    nb_set_error("nb_force_device: null argument");
    return NB_ERR_ARG;
  }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  HCHK(hipSetDevice(c->dev));
  if (d_pxyz != c->d_pos)
    HCHK(hipMemcpyAsync(c->d_pos, d_pxyz, sizeof(double) * 3 * c->natoms,
                        hipMemcpyDeviceToDevice, s));
  return force_on(c, iflg, hmat, d_fs, iflg ? d_fl : nullptr, e, s, e != nullptr);
}

// nb_force_device with the per-evaluation scratch (Ewald charge records,
// structure-factor partials, per-k data, reciprocal-force partials, energy
// columns) carved from the static scratch area instead of the context's own
// buffers; allocates nothing.  Every kernel overwrites the scratch it reads
// before reading it, so nothing in it has to start at zero.
static size_t scratch_layout(const nb_ctx *c, size_t off[12]) {
  const size_t nq = std::max(c->ncharge, 1), kt = std::max(c->ktot, 1);
  const size_t bytes[11] = {32 * nq, 32 * nq, 16 * (size_t)c->nchunk * kt, nb_kdata_bytes() * kt,
                            24 * (size_t)c->nrange * nq,
                            8 * (size_t)c->npair_blocks * NB_NE_PAIRQ,
                            8 * (size_t)c->nfin_blocks * NB_NE_ECORR,
                            8 * (size_t)c->new_blocks * NB_NE_EW, 8 * 64,
                            c->sfm ? 8 * nb_sfm_table_doubles(c) : 0,
                            c->efm ? 64 * (size_t)c->efm_nent : 0};
  size_t o = 0;
  for (int k = 0; k < 11; k++) {
    off[k] = o;
    o += (bytes[k] + 255) & ~(size_t)255;
  }
  off[11] = o;
  return o;
}

extern "C" int nb_force_device_noalloc(nb_ctx *c, int iflg, const double *d_pxyz,
                                       const double hmat[9], double *d_fs, double *d_fl,
                                       nb_energy *e, void *stream) {
  if (!c || !d_pxyz || !hmat || (iflg == 1 && !d_fl) || (iflg == 0 && !d_fs)) {
    nb_set_error("nb_force_device_noalloc: null argument");
    return NB_ERR_ARG;
  }
  if (iflg != 0 && iflg != 1) {
    nb_set_error("nb_force_device_noalloc: iflg must be 0 or 1");
    return NB_ERR_ARG;
  }
  size_t off[12];
  const size_t need = scratch_layout(c, off);
  if (need > NB_SCRATCH_BYTES) {
    nb_set_error("nb_force_device_noalloc: this system needs %zu bytes of scratch, the static "
                 "area holds %u (use nb_force_device)", need, NB_SCRATCH_BYTES);
    return NB_ERR_CAPACITY;
  }
  if (!c->have_lists) {
    nb_set_error("nb_force_device_noalloc: lists not built (call nb_build_lists)");
    return NB_ERR_ARG;
  }
  char *base;
  RC(scratch_ptr(c->dev, &base));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  HCHK(hipSetDevice(c->dev));
  if (d_pxyz != c->d_pos)
    HCHK(hipMemcpyAsync(c->d_pos, d_pxyz, sizeof(double) * 3 * c->natoms,
                        hipMemcpyDeviceToDevice, s));
  // the context's tables, lists and records; the scratch from the static area
  nb_ctx v = *c;
  v.d_chg = (double *)(base + off[0]);
  v.d_stp = (double *)(base + off[1]);
  v.d_spart = (double *)(base + off[2]);
  v.d_kd = (double *)(base + off[3]);
  v.d_fpart = (double *)(base + off[4]);
  v.d_epair = (double *)(base + off[5]);
  v.d_eec = (double *)(base + off[6]);
  v.d_eew = (double *)(base + off[7]);
  v.d_eout = (double *)(base + off[8]);
  if (c->sfm) v.d_sfmtab = (double *)(base + off[9]);
  if (c->efm) v.d_fimg = (double *)(base + off[10]);
  v.timing = 0;
  v.rec = nullptr;
  Box b = make_box(hmat);
  RC(nb_launch_force(&v, iflg, b, d_fs, iflg ? d_fl : nullptr, s, e ? 0 : NB_FORCE_NOENER));
  if (e) {
    double E[64];
    HCHK(hipMemcpyAsync(E, v.d_eout, sizeof(E), hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    nb_fill_energy(c, iflg, E, hmat, e);
  }
  return NB_OK;
}

// getvireal(...,0)'s pair sums (getvirter / getvirter_e over the t and te
// lists, md_finter.c:314-477, md_finter_e.c:235-366) plus the reciprocal
// and ecorr energies and virials, with no force output.  The pair kernels
// are the force path's (their force stores go to scratch: 24 B per atom
// beside 4 B per listed pair they gather), so the values are bit for bit
// force(...,1)'s; the reciprocal sum skips its force pass.
extern "C" int nb_energy_only(nb_ctx *c, const double *pxyz, const double hmat[9],
                              nb_energy *e) {
  if (!c || !pxyz || !hmat || !e) {
    nb_set_error("nb_energy_only: null argument");
    return NB_ERR_ARG;
  }
  if (!c->have_lists) {
    nb_set_error("nb_energy_only: lists not built (call nb_build_lists)");
    return NB_ERR_ARG;
  }
  HCHK(hipSetDevice(c->dev));
  HCHK(hipMemcpyAsync(c->d_pos, pxyz, sizeof(double) * 3 * c->natoms, hipMemcpyHostToDevice,
                      c->stream));
  Box b = make_box(hmat);
  RC(nb_launch_force(c, 1, b, nullptr, c->d_fs, c->stream, NB_FORCE_EONLY | NB_FORCE_CM3D_EONLY));
  double E[64];
  HCHK(hipMemcpyAsync(E, c->d_eout, sizeof(E), hipMemcpyDeviceToHost, c->stream));
  HCHK(hipStreamSynchronize(c->stream));
  nb_fill_energy(c, 1, E, hmat, e, true);
  return NB_OK;
}

extern "C" int nb_recip_info(nb_ctx *c, long long out[4]) {
  if (!c || !out) { nb_set_error("nb_recip_info: null argument"); return NB_ERR_ARG; }
  out[0] = (c->sfm ? 1 : 0) | (c->efm ? 2 : 0);
  out[1] = c->sfm ? c->sfm_mfma : 0;
  out[2] = c->efm ? c->efm_mfma : 0;
  out[3] = 0;
  return NB_OK;
}

extern "C" int nb_footprint(nb_ctx *c, size_t *bytes) {
  if (!c || !bytes) { nb_set_error("nb_footprint: null argument"); return NB_ERR_ARG; }
  *bytes = c->dev_bytes;
  return NB_OK;
}

extern "C" int nb_list_counts(nb_ctx *c, int counts[4]) {
  if (!c || !counts) { nb_set_error("nb_list_counts: null argument"); return NB_ERR_ARG; }
  if (!c->have_lists) { nb_set_error("nb_list_counts: lists not built"); return NB_ERR_ARG; }
  HCHK(hipSetDevice(c->dev));
  HCHK(hipStreamSynchronize(c->stream));
  std::vector<int> h(4 * (size_t)c->natoms);
  HCHK(hipMemcpy(h.data(), c->d_hcnt, h.size() * sizeof(int), hipMemcpyDeviceToHost));
  long long t[4] = {0, 0, 0, 0};
  for (int i = 0; i < c->natoms; i++)
    for (int k = 0; k < 4; k++) t[k] += h[4 * (size_t)i + k];
  for (int k = 0; k < 4; k++) counts[k] = (int)t[k];
  return NB_OK;
}

extern "C" int nb_get_pairs(nb_ctx *c, int which, int *ip, int *jp, int cap, int *nout) {
  if (!c || !nout || which < 0 || which > 3) { nb_set_error("nb_get_pairs: bad argument"); return NB_ERR_ARG; }
  if (!c->have_lists) { nb_set_error("nb_get_pairs: lists not built"); return NB_ERR_ARG; }
  HCHK(hipSetDevice(c->dev));
  const int n = c->natoms;
  std::vector<uint32_t> nbr((size_t)n * c->cap);
  std::vector<int> nnb(n), perm(n);
  HCHK(hipStreamSynchronize(c->stream));
  HCHK(hipMemcpy(nbr.data(), c->d_nbr, nbr.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HCHK(hipMemcpy(nnb.data(), c->d_nnb, n * sizeof(int), hipMemcpyDeviceToHost));
  HCHK(hipMemcpy(perm.data(), c->d_perm, n * sizeof(int), hipMemcpyDeviceToHost));
  static const unsigned bit[4] = {NB_FLAG_VS, NB_FLAG_VT, NB_FLAG_ES, NB_FLAG_ET};
  std::vector<std::pair<int, int>> pr;
  for (int s = 0; s < n; s++) {
    int a = perm[s];
    for (int k = 0; k < nnb[s]; k++) {
      uint32_t e = nbr[(size_t)s * c->cap + k];
      if (!((e >> 28) & bit[which])) continue;
      int b = perm[e & c->imask];
      if (a < b) pr.emplace_back(a, b);
    }
  }
  std::sort(pr.begin(), pr.end());
  *nout = (int)pr.size();
  if ((int)pr.size() > cap) { nb_set_error("nb_get_pairs: capacity %d < %zu", cap, pr.size()); return NB_ERR_CAPACITY; }
  for (size_t k = 0; k < pr.size(); k++) { ip[k] = pr[k].first; jp[k] = pr[k].second; }
  return NB_OK;
}

void nb_time_mark(nb_ctx *c, int ev, hipStream_t s) {
  if (!c->timing || c->trec >= NB_TIME_REC) return;
  hipEventRecord(c->tev[c->trec * NB_TIME_EV + ev], s);
}

extern "C" int nb_timing(nb_ctx *c, int mode, double out[NB_TIMING_N]) {
  if (!c) { nb_set_error("nb_timing: null context"); return NB_ERR_ARG; }
  HCHK(hipSetDevice(c->dev));
  if (mode == 1) {  // start / restart recording
    if (!c->tev) {
      c->tev = new hipEvent_t[NB_TIME_REC * NB_TIME_EV];
      for (int k = 0; k < NB_TIME_REC * NB_TIME_EV; k++) HCHK(hipEventCreate(&c->tev[k]));
    }
    c->timing = 1;
    c->trec = 0;
    return NB_OK;
  }
  if (mode == 0) { c->timing = 0; return NB_OK; }
  // mode 2: averages over the recorded force calls, milliseconds:
  // out = {calls, pack, pair kernel, charge records, structure factor,
  //        k reduction, reciprocal force, ecorr + force sum + reductions, total}
  if (!out || !c->tev) { nb_set_error("nb_timing: nothing recorded"); return NB_ERR_ARG; }
  int nrec = c->trec < NB_TIME_REC ? c->trec : NB_TIME_REC;
  for (int k = 0; k < NB_TIMING_N; k++) out[k] = 0.;
  out[0] = nrec;
  for (int r = 0; r < nrec; r++) {
    hipEvent_t *e = c->tev + r * NB_TIME_EV;
    HCHK(hipEventSynchronize(e[NB_TIME_EV - 1]));
    float ms;
    for (int k = 1; k < NB_TIME_EV; k++) {
      HCHK(hipEventElapsedTime(&ms, e[k - 1], e[k]));
      out[k] += ms;
    }
    HCHK(hipEventElapsedTime(&ms, e[0], e[NB_TIME_EV - 1]));
    out[NB_TIME_EV] += ms;
  }
  if (nrec > 0)
    for (int k = 1; k < NB_TIMING_N; k++) out[k] /= nrec;
  return NB_OK;
}
