// nb_batch.hip -- many independent systems advanced together.
//
// cm3d runs one system per process (main/md_main.new.c:61-134) and keeps its
// per-system state in file statics; here each system is an nb_ctx + nb_md
// with all of that state in its own HBM arrays, and a batch launches every
// kernel of the step ONCE for all systems: the step is recorded per system
// (same kernel sequence, per-system arguments and block counts), the
// arguments are uploaded as one array per kernel, and a block map sends each
// workgroup of the batched grid to (system, local block).  The per-step
// launch count is that of one system (~11), independent of the batch size,
// and the step can be captured into one hipGraph.
//
// A tick (one nb_batch_step call) is bracketed by two more batched
// launches: a snapshot of every system's dynamic state before the steps,
// and a check of every system's sticky overflow flag after them.  A system
// whose device-side list build wanted more neighbour slots than it has
// (cm3d instead regrows its lists by counting, ngbr_list.c:331-335) is
// restored from the snapshot, given more slots, rebuilt and stepped again
// alone -- what the batch computes for a system is bitwise what it computes
// alone -- so no system ever continues on truncated lists.  The recording
// is redone whenever a member's buffers or launch shapes changed since it
// was made (the generation counters of its context and MD state).
#include <string.h>

#include <vector>

#include "nb_launch.h"

int md_enqueue_step(nb_md *md, hipStream_t s);
int md_enqueue_save(nb_md *md, hipStream_t s, int restore);
int md_enqueue_ovfchk(nb_md *md, hipStream_t s, int *d_out);
int md_ensure_save(nb_md *md);
int md_recover(nb_md *md, int nsteps, int want, hipStream_t s);
void nb_md_gens(nb_md *md, unsigned *cgen, unsigned *tgen);
nb_ctx *nb_md_ctx(nb_md *md);

struct BatchKernel {
  int kid, tpb, lds, total;
  char *d_args;
  int2 *d_bmap;
};

struct nb_batch {
  int dev, nsys;
  std::vector<nb_md *> mds;
  std::vector<BatchKernel> ks;     // one step
  std::vector<BatchKernel> pre;    // the tick's snapshot
  std::vector<BatchKernel> post;   // the tick's overflow check
  std::vector<unsigned> cgen, tgen;  // members' generations at recording
  int *d_ovf, *h_ovf;              // per system: overflowed row length, or 0
  long long recoveries;            // systems redone after an overflow
  hipGraph_t graph;
  hipGraphExec_t gexec;
  int have_graph;
  hipStream_t stream;
};

#define BCHK(x)                                                        \
  do {                                                                 \
    hipError_t _e = (x);                                               \
    if (_e != hipSuccess) {                                            \
      nb_set_error("%s: %s", #x, hipGetErrorString(_e));               \
      return NB_ERR_HIP;                                               \
    }                                                                  \
  } while (0)

static void free_kernels(std::vector<BatchKernel> &ks) {
  for (auto &k : ks) {
    if (k.d_args) hipFree(k.d_args);
    if (k.d_bmap) hipFree(k.d_bmap);
  }
  ks.clear();
}

static void drop_recording(nb_batch *b) {
  if (b->have_graph) {
    hipGraphExecDestroy(b->gexec);
    hipGraphDestroy(b->graph);
    b->have_graph = 0;
  }
  free_kernels(b->ks);
  free_kernels(b->pre);
  free_kernels(b->post);
}

extern "C" void nb_batch_destroy(nb_batch *b) {
  if (!b) return;
  hipSetDevice(b->dev);
  hipDeviceSynchronize();  // the batch's graph may still be running
  drop_recording(b);
  if (b->d_ovf) hipFree(b->d_ovf);
  if (b->h_ovf) hipHostFree(b->h_ovf);
  delete b;
}

// Merge per-system recordings (same kernel sequence) into batched launches.
static int merge(nb_batch *b, std::vector<std::vector<LaunchRec>> &recs,
                 std::vector<BatchKernel> &out) {
  const int nsys = b->nsys;
  const size_t nk = recs[0].size();
  for (int i = 1; i < nsys; i++) {
    if (recs[i].size() != nk) {
      nb_set_error("nb_batch: system %d records a different step (precision or options differ)", i);
      return NB_ERR_UNSUPPORTED;
    }
    for (size_t k = 0; k < nk; k++)
      if (recs[i][k].kid != recs[0][k].kid || recs[i][k].tpb != recs[0][k].tpb ||
          recs[i][k].args.size() != recs[0][k].args.size()) {
        nb_set_error("nb_batch: system %d step differs at launch %zu", i, k);
        return NB_ERR_UNSUPPORTED;
      }
  }
  for (size_t k = 0; k < nk; k++) {
    BatchKernel bk;
    bk.kid = recs[0][k].kid;
    bk.tpb = recs[0][k].tpb;
    bk.lds = 0;
    bk.total = 0;
    bk.d_args = nullptr;
    bk.d_bmap = nullptr;
    const size_t asz = recs[0][k].args.size();
    std::vector<char> args(asz * nsys);
    std::vector<int2> bmap;
    for (int i = 0; i < nsys; i++) {
      const LaunchRec &r = recs[i][k];
      memcpy(args.data() + asz * i, r.args.data(), asz);
      if (r.lds > bk.lds) bk.lds = r.lds;
      for (int q = 0; q < r.nblk; q++) bmap.push_back(make_int2(i, q));
    }
    bk.total = (int)bmap.size();
    out.push_back(bk);
    BatchKernel &o = out.back();
    if (o.total > 0) {
      if (hipMalloc((void **)&o.d_args, args.size()) != hipSuccess ||
          hipMalloc((void **)&o.d_bmap, bmap.size() * sizeof(int2)) != hipSuccess) {
        nb_set_error("nb_batch: out of device memory");
        return NB_ERR_HIP;
      }
      BCHK(hipMemcpy(o.d_args, args.data(), args.size(), hipMemcpyHostToDevice));
      BCHK(hipMemcpy(o.d_bmap, bmap.data(), bmap.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    if (!g_batch_fn[o.kid]) {
      nb_set_error("nb_batch: kernel %d has no batched entry", o.kid);
      return NB_ERR_UNSUPPORTED;
    }
  }
  return NB_OK;
}

// Record one step, the snapshot and the overflow check of every system.
static int record(nb_batch *b) {
  const int nsys = b->nsys;
  std::vector<std::vector<LaunchRec>> st(nsys), sv(nsys), ck(nsys);
  for (int i = 0; i < nsys; i++) {
    nb_md *md = b->mds[i];
    nb_ctx *c = nb_md_ctx(md);
    int rc = md_ensure_save(md);
    if (rc) return rc;
    c->rec = &st[i];
    rc = md_enqueue_step(md, c->stream);
    if (!rc) { c->rec = &sv[i]; rc = md_enqueue_save(md, c->stream, 0); }
    if (!rc) { c->rec = &ck[i]; rc = md_enqueue_ovfchk(md, c->stream, b->d_ovf + i); }
    c->rec = nullptr;
    if (rc) return rc;
  }
  int rc = merge(b, st, b->ks);
  if (!rc) rc = merge(b, sv, b->pre);
  if (!rc) rc = merge(b, ck, b->post);
  if (rc) return rc;
  b->cgen.resize(nsys);
  b->tgen.resize(nsys);
  for (int i = 0; i < nsys; i++) nb_md_gens(b->mds[i], &b->cgen[i], &b->tgen[i]);
  return NB_OK;
}

static bool stale(nb_batch *b) {
  for (int i = 0; i < b->nsys; i++) {
    unsigned cg, tg;
    nb_md_gens(b->mds[i], &cg, &tg);
    if (cg != b->cgen[i] || tg != b->tgen[i]) return true;
  }
  return false;
}

extern "C" int nb_batch_create(nb_md **mds, int nsys, nb_batch **out) {
  if (!mds || nsys <= 0 || !out) { nb_set_error("nb_batch_create: bad argument"); return NB_ERR_ARG; }
  *out = nullptr;
  nb_batch *b = new nb_batch();
  b->dev = nb_md_ctx(mds[0])->dev;
  b->nsys = nsys;
  b->have_graph = 0;
  b->recoveries = 0;
  b->d_ovf = nullptr;
  b->h_ovf = nullptr;
  b->stream = nb_md_ctx(mds[0])->stream;
  b->mds.assign(mds, mds + nsys);
  for (int i = 0; i < nsys; i++)
    if (nb_md_ctx(mds[i])->dev != b->dev) {
      delete b;
      nb_set_error("nb_batch_create: systems on different devices");
      return NB_ERR_ARG;
    }
  hipSetDevice(b->dev);
  if (hipMalloc((void **)&b->d_ovf, nsys * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void **)&b->h_ovf, nsys * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    nb_batch_destroy(b);
    nb_set_error("nb_batch_create: out of memory");
    return NB_ERR_HIP;
  }
  int rc = record(b);
  if (rc) {
    nb_batch_destroy(b);
    return rc;
  }
  *out = b;
  return NB_OK;
}

static void launch_all(const std::vector<BatchKernel> &ks, hipStream_t s) {
  for (auto &k : ks)
    if (k.total > 0) g_batch_fn[k.kid](k.d_args, k.d_bmap, k.total, k.tpb, k.lds, s);
}

extern "C" int nb_batch_step(nb_batch *b, int nsteps, void *stream, int use_graph) {
  if (!b || nsteps < 0) { nb_set_error("nb_batch_step: bad argument"); return NB_ERR_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : b->stream;
  BCHK(hipSetDevice(b->dev));
  if (stale(b)) {
    // a member was regrown, re-gridded or re-thermostatted since recording
    BCHK(hipDeviceSynchronize());
    drop_recording(b);
    int rc = record(b);
    if (rc) return rc;
  }
  if (use_graph && !b->have_graph) {
    BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_all(b->ks, s);
    hipGraph_t g;
    hipError_t e = hipStreamEndCapture(s, &g);
    BCHK(e);
    BCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->graph = g;
    b->have_graph = 1;
  }
  if (nsteps == 0) return NB_OK;  // prepared only
  launch_all(b->pre, s);  // snapshot + clear the sticky overflow flags
  for (int k = 0; k < nsteps; k++) {
    if (use_graph) BCHK(hipGraphLaunch(b->gexec, s));
    else launch_all(b->ks, s);
  }
  launch_all(b->post, s);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(b->h_ovf, b->d_ovf, b->nsys * sizeof(int), hipMemcpyDeviceToHost, s));
  BCHK(hipStreamSynchronize(s));
  for (int i = 0; i < b->nsys; i++) {
    // each member's last force(...,1) is at the positions the tick leaves
    nb_md_ctx(b->mds[i])->recip_owner = b->mds[i];
    if (b->h_ovf[i] <= 0) continue;
    int rc = md_recover(b->mds[i], nsteps, b->h_ovf[i], s);
    if (rc) return rc;
    b->recoveries++;
  }
  return NB_OK;
}

extern "C" int nb_batch_launches(nb_batch *b, int *nkernels) {
  if (!b || !nkernels) { nb_set_error("nb_batch_launches: null argument"); return NB_ERR_ARG; }
  int n = 0;
  for (auto &k : b->ks) n += k.total > 0;
  *nkernels = n;
  return NB_OK;
}

extern "C" int nb_batch_recoveries(nb_batch *b, long long *count) {
  if (!b || !count) { nb_set_error("nb_batch_recoveries: null argument"); return NB_ERR_ARG; }
  *count = b->recoveries;
  return NB_OK;
}
