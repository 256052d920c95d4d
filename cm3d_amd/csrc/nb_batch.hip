// nb_batch.hip -- many independent systems advanced together.
//
// cm3d runs one system per process (main/md_main.new.c:61-134) and keeps its
// per-system state in file statics; here each system is an nb_ctx + nb_md
// with all of that state in its own HBM arrays, and a batch launches every
// kernel of the step ONCE for all systems: the step is recorded per system
// (same kernel sequence, per-system arguments and block counts), the
// arguments are uploaded as one array per kernel, and a block map sends each
// workgroup of the batched grid to (system, local block).  The per-step
// launch count is that of one system (~25), independent of the batch size,
// and the step can be captured into one hipGraph.
#include <string.h>

#include <vector>

#include "nb_launch.h"

int md_enqueue_step(nb_md *md, hipStream_t s);
nb_ctx *nb_md_ctx(nb_md *md);

struct BatchKernel {
  int kid, tpb, lds, total;
  char *d_args;
  int2 *d_bmap;
};

struct nb_batch {
  int dev, nsys;
  std::vector<nb_md *> mds;
  std::vector<BatchKernel> ks;
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

extern "C" void nb_batch_destroy(nb_batch *b) {
  if (!b) return;
  hipSetDevice(b->dev);
  hipDeviceSynchronize();  // the batch's graph may still be running
  if (b->have_graph) { hipGraphExecDestroy(b->gexec); hipGraphDestroy(b->graph); }
  for (auto &k : b->ks) {
    if (k.d_args) hipFree(k.d_args);
    if (k.d_bmap) hipFree(k.d_bmap);
  }
  delete b;
}

extern "C" int nb_batch_create(nb_md **mds, int nsys, nb_batch **out) {
  if (!mds || nsys <= 0 || !out) { nb_set_error("nb_batch_create: bad argument"); return NB_ERR_ARG; }
  *out = nullptr;
  nb_batch *b = new nb_batch();
  b->dev = nb_md_ctx(mds[0])->dev;
  b->nsys = nsys;
  b->have_graph = 0;
  b->stream = nb_md_ctx(mds[0])->stream;
  b->mds.assign(mds, mds + nsys);
  hipSetDevice(b->dev);
  // record one step of every system
  std::vector<std::vector<LaunchRec>> recs(nsys);
  for (int i = 0; i < nsys; i++) {
    nb_ctx *c = nb_md_ctx(mds[i]);
    if (c->dev != b->dev) { delete b; nb_set_error("nb_batch_create: systems on different devices"); return NB_ERR_ARG; }
    c->rec = &recs[i];
    int rc = md_enqueue_step(mds[i], c->stream);
    c->rec = nullptr;
    if (rc) { delete b; return rc; }
  }
  const size_t nk = recs[0].size();
  for (int i = 1; i < nsys; i++) {
    if (recs[i].size() != nk) {
      delete b;
      nb_set_error("nb_batch_create: system %d records a different step (precision or options differ)", i);
      return NB_ERR_UNSUPPORTED;
    }
    for (size_t k = 0; k < nk; k++)
      if (recs[i][k].kid != recs[0][k].kid || recs[i][k].tpb != recs[0][k].tpb ||
          recs[i][k].args.size() != recs[0][k].args.size()) {
        delete b;
        nb_set_error("nb_batch_create: system %d step differs at launch %zu", i, k);
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
    if (bk.total > 0) {
      if (hipMalloc((void **)&bk.d_args, args.size()) != hipSuccess ||
          hipMalloc((void **)&bk.d_bmap, bmap.size() * sizeof(int2)) != hipSuccess) {
        b->ks.push_back(bk);
        nb_batch_destroy(b);
        nb_set_error("nb_batch_create: out of device memory");
        return NB_ERR_HIP;
      }
      hipMemcpy(bk.d_args, args.data(), args.size(), hipMemcpyHostToDevice);
      hipMemcpy(bk.d_bmap, bmap.data(), bmap.size() * sizeof(int2), hipMemcpyHostToDevice);
    }
    if (!g_batch_fn[bk.kid]) {
      b->ks.push_back(bk);
      nb_batch_destroy(b);
      nb_set_error("nb_batch_create: kernel %d has no batched entry", bk.kid);
      return NB_ERR_UNSUPPORTED;
    }
    b->ks.push_back(bk);
  }
  *out = b;
  return NB_OK;
}

static int batch_enqueue(nb_batch *b, hipStream_t s) {
  for (auto &k : b->ks) {
    if (k.total == 0) continue;
    g_batch_fn[k.kid](k.d_args, k.d_bmap, k.total, k.tpb, k.lds, s);
  }
  BCHK(hipGetLastError());
  return NB_OK;
}

extern "C" int nb_batch_step(nb_batch *b, int nsteps, void *stream, int use_graph) {
  if (!b || nsteps < 0) { nb_set_error("nb_batch_step: bad argument"); return NB_ERR_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : b->stream;
  BCHK(hipSetDevice(b->dev));
  if (use_graph && !b->have_graph) {
    BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, s);
    hipGraph_t g;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    BCHK(e);
    BCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->graph = g;
    b->have_graph = 1;
  }
  for (int k = 0; k < nsteps; k++) {
    if (use_graph) {
      BCHK(hipGraphLaunch(b->gexec, s));
    } else {
      int rc = batch_enqueue(b, s);
      if (rc) return rc;
    }
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
