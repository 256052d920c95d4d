"""Benchmark of the MI355X nonbonded engine (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload (value): the 98,304-atom flexible water -- 32,768 3-site
waters in a cubic box of 99.3 A at 1.0 g/cc, O-O Lennard-Jones and erfc
real-space Coulomb tabulated with 500 nodes, rc 12 A, skin 1 A, Ewald alpha
0.30/A with kmax 32 (68,532 half-space k vectors), harmonic O-H bonds and
H-O-H bends, RESPA with 4 x 0.5 fs inner and 2 fs outer steps, 300 K, FP32
pair math with 64-bit fixed-point force sums and FP64 Ewald / energy /
virial sums and integrator; K = 500 steps (the configuration's run length).
A step is one outer MD step on the device, replayed from a hipGraph: the
inner bonded RESPA loop, cm3d's nrefsh rebuild test (and the list build
when it fires), force(...,1) -- pairs, Ewald, ecorr -- and the kicks.  The
timed loop also evaluates cm3d's energies every --energy-every steps (the
decks' iwr_inst cadence): the FP64 energy pass and its read-back.  value =
atom-steps/s of that system.  Reported beside it: the same run in plain
FP64, the mixed form's force error against FP64, energy drift per ns of
both, p50/p99 step latency, the per-kernel roofline from HIP events on the
kernels' stream, pair interactions/s (pairs within the cutoffs), and the
CPU baseline -- the clean-room C oracle of cm3d's CPU path timed on this
host on a bounded sample of the same system.

Secondary blocks: the 12,288-atom water (4096 molecules, rc 10, alpha 0.35,
kmax 9) in both precisions with the pair kernel's roofline and a 40,000-step
drift fit; the 64-system batch of LJ + charge mixtures (N log-uniform in
500-20,000) batched in mixed precision against the same 64 systems stepped
one at a time on the plain FP64 path (the batched-vs-plain ratio, like for
like), with force errors and fitted drifts for both precisions; the serving
stream of 2000 jobs with Poisson arrivals at 20/s.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
VALU_ISSUE_PER_S = 256 * 2.4e9  # wave instructions: one per CU per cycle
DT_PS = 0.002
# algorithmic FP64 work of the reciprocal sum per (k vector, charge):
# structure factor 10 (complex recursion 6 + accumulate 4), force pass 16
EW_SFAC_FLOP, EW_FORCE_FLOP = 10.0, 16.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="timed steps of the headline run")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--energy-every", type=int, default=10,
                    help="energies evaluated (and read back) every this many timed steps")
    ap.add_argument("--nside", type=int, default=32, help="waters per box edge (headline)")
    ap.add_argument("--equil", type=int, default=200,
                    help="untimed steps with velocity rescaling to 300 K before timing")
    ap.add_argument("--c3-drift-steps", type=int, default=2000)
    ap.add_argument("--lat-steps", type=int, default=50,
                    help="single synchronised steps timed for the latency percentiles")
    ap.add_argument("--c2-steps", type=int, default=300)
    ap.add_argument("--c2-equil", type=int, default=600)
    ap.add_argument("--drift-steps", type=int, default=40000, help="drift run of the 4096 waters")
    ap.add_argument("--nsys", type=int, default=64, help="systems in the batched form")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--batch-drift-steps", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--serve-jobs", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="mean arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=64)
    for blk in ("cpu", "c2", "batch", "serve"):
        ap.add_argument(f"--no-{blk}", action="store_true")
    ap.add_argument("--traffic-probe", choices=["c2", "c2m", "c3"], default=None,
                    help="run only a 256 MiB device copy (counter calibration) and 5 force(...,1) "
                         "calls as the MD step makes them (forces only) on the 4096-water "
                         "system in FP64 (c2) or mixed (c2m), or the 32768-water system mixed "
                         "(c3), for rocprofv3 --pmc")
    ap.add_argument("--only-batch", action="store_true",
                    help="run the batched form alone (profiling); prints its own JSON")
    return ap.parse_args()


def water_c3(S, nside=32):
    return S.water(nside=nside, rc=12.0, alpha=0.30, kmax=32)


def _bonded(sysd):
    return dict(ib=sysd["bond_i"], jb=sysd["bond_j"], bk=sysd["bond_k"], beq=sysd["bond_r0"],
                ia=sysd["bend_i"], ja=sysd["bend_j"], ka=sysd["bend_k"], ak=sysd["bend_fk"],
                aeq=sysd["bend_th0"])


def _first_k(sysd, keep):
    """The same system with only its first `keep` k vectors."""
    t = dict(sysd)
    t["ka"], t["kb"], t["kc"] = (np.asarray(sysd[k])[:keep].copy() for k in ("ka", "kb", "kc"))
    kup = np.asarray(sysd["kup"])[:keep + 1].copy()
    if keep:
        kup[keep] = 1
    t["kup"] = kup
    t["ktot"] = keep
    return t


def cpu_baseline(sysd, budget_s, rebuild_rate):
    """The oracle (clean-room restatement of cm3d's CPU path, gcc -O2, one
    core) on a bounded sample of the same system: one list build, the
    real-space force(...,1) (every pair, ecorr) and the four bonded
    evaluations of a RESPA step in full, and the reciprocal sum over the
    first k vectors, scaled to all of them (its cost is linear in K)."""
    from oracle import oracle as O
    pos, h = sysd["pos"], sysd["hmat"]
    n = int(sysd["natoms"])
    ktot = int(sysd["ktot"])
    t0 = time.perf_counter()
    L = O.build_lists(sysd, pos, h)
    t_list = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.force(_first_k(sysd, 0), L, pos, h, 1)
    t_real = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(4):
        O.bonded_force(n, _bonded(sysd), pos)
    t_bond = time.perf_counter() - t0
    left = max(1.0, budget_s - t_list - t_real - t_bond)
    # calibrate the reciprocal sum on a small k set, then size the sample to
    # the remaining budget
    k1 = min(ktot, 64)
    t0 = time.perf_counter()
    O.force(_first_k(sysd, k1), L, pos, h, 1)
    t_k1 = max(1e-9, time.perf_counter() - t0 - t_real)
    ks = int(min(ktot, max(k1, k1 * left / t_k1 * 0.8)))
    t0 = time.perf_counter()
    O.force(_first_k(sysd, ks), L, pos, h, 1)
    t_ks = max(1e-9, time.perf_counter() - t0 - t_real)
    t_ew = t_ks * ktot / ks
    t_step = t_real + t_ew + t_bond + rebuild_rate * t_list
    sample = (f"{n}-atom water: one list build ({t_list:.2f} s), the real-space force(...,1) "
              f"({t_real:.2f} s) and 4 bonded evaluations ({t_bond:.3f} s) in full, the "
              f"reciprocal sum over the first {ks} of {ktot} k vectors ({t_ks:.2f} s) scaled "
              f"by {ktot / ks:.1f}; per step = real + reciprocal + bonded + list builds at the "
              f"GPU run's rate ({rebuild_rate:.3f}/step)")
    return n / t_step, sample


def _sync_steps(torch, step, nsteps):
    """Per-step latency as a caller sees it: launch one step, wait for it."""
    lat = []
    for _ in range(nsteps):
        t = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append(1e3 * (time.perf_counter() - t))
    return np.array(lat)


def _timed_md(torch, md, steps, energy_every, sp):
    """steps graph-replayed steps, energies evaluated and read back every
    energy_every steps (0: never); returns (wall seconds, last energies)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, e = 0, None
    while done < steps:
        k = min(energy_every if energy_every > 0 else steps, steps - done)
        md.step(k, stream=sp, graph=True)
        done += k
        if energy_every > 0:
            e = md.energies()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, e


def _drift(mds, syss, nsteps, stepper, nsamp=20):
    """Slope (K/atom/ns) of a linear fit to the mean over systems of Ham/N
    sampled nsamp+1 times over nsteps NVE steps, its standard error, the
    RMS about the fit and the initial value."""
    every = max(1, nsteps // nsamp)
    t_ns = np.arange(nsamp + 1) * every * DT_PS * 1e-3
    ns = [int(s["natoms"]) for s in syss]

    def sample():
        return float(np.mean([m.energies()["ham"] / n for m, n in zip(mds, ns)]))
    ham = [sample()]
    for _ in range(nsamp):
        stepper(every)
        ham.append(sample())
    ham = np.array(ham)
    fit = np.polyfit(t_ns, ham, 1)
    resid = np.polyval(fit, t_ns) - ham
    se = float(np.sqrt(np.sum(resid ** 2) / (len(t_ns) - 2) / np.sum((t_ns - t_ns.mean()) ** 2)))
    return {"K_per_atom_per_ns": round(float(fit[0]), 6), "stderr": round(se, 6),
            "rms_about_fit_K_per_atom": round(float(np.sqrt(np.mean(resid ** 2))), 6),
            "ham0_K_per_atom": round(float(ham[0]), 6),
            "run": f"{nsamp * every} NVE steps, Ham sampled every {every}"}


def pairs_within_rc(nb, s, pos):
    """Pairs inside the interaction cutoffs (not the list radius) at pos:
    the list counts of a context whose list radii carry no skin.  For these
    systems every vdW pair is also a Coulomb pair or vice versa, so the
    larger count is the number of distinct interacting pairs."""
    t = dict(s)
    t["rmaxl_skin2"] = np.asarray(s["rmaxl_cut2"]).copy()
    t["rmaxs_skin2"] = np.asarray(s["rmaxs_cut2"]).copy()
    if int(np.count_nonzero(s["qch"])):
        t["rmaxl_skin2_e"] = float(np.asarray(s["rmaxl_coul2"]).reshape(-1)[0])
        t["rmaxs_skin2_e"] = float(np.asarray(s["rmaxs_coul2"]).reshape(-1)[0])
    c = nb.NBContext(t, precision="fp64")
    c.build_lists(pos, s["hmat"])
    cnt = c.list_counts()
    c.close()
    return max(cnt["nprt"], cnt["nprte"]), cnt


def force_err(nb, s, pos, prec_a="mixed", prec_b="fp64", sharing=1):
    out = []
    for prec in (prec_a, prec_b):
        c = nb.NBContext(s, precision=prec, sharing=sharing)
        c.build_lists(pos, s["hmat"])
        out.append(c.force(pos, s["hmat"], iflg=1)[1])
        c.close()
    return float(np.sqrt(np.mean((out[0] - out[1]) ** 2) / np.mean(out[1] ** 2)))


def kernel_phases(torch, ctx, s, pos, sp, ncalls=10):
    """force(...,1) as the MD step evaluates it (forces only) on pos, with
    HIP events around each phase on the kernels' stream: ms per call."""
    ctx.build_lists(pos, s["hmat"])
    d_pos = torch.from_numpy(pos).cuda()
    d_fl = torch.zeros_like(d_pos)
    for _ in range(2):
        ctx.force_device(1, d_pos.data_ptr(), s["hmat"], 0, d_fl.data_ptr(), stream=sp)
    torch.cuda.synchronize()
    ctx.timing_start()
    for _ in range(ncalls):
        ctx.force_device(1, d_pos.data_ptr(), s["hmat"], 0, d_fl.data_ptr(), stream=sp)
    torch.cuda.synchronize()
    return ctx.timing_read()


def _profile_json(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def run_c3(nb, S, a, torch):
    s = water_c3(S, a.nside)
    n = int(s["natoms"])
    sp = torch.cuda.current_stream().cuda_stream
    ctx = nb.NBContext(s, precision="mixed")
    md = nb.NBMD(ctx, s, DT_PS, nres_intra=4)
    md.set_state(s["pos"], s["vel"], s["hmat"])
    md.init()
    # the lattice start relaxes and heats: bring it to 300 K first (untimed)
    S.equilibrate([md], [s], 300.0, a.equil, every=10)
    pos_eq, vel_eq = md.get_state()
    print(f"[bench] c3 equilibrated ({a.equil} steps)", file=sys.stderr, flush=True)
    md.step(a.warmup, stream=sp, graph=True)
    r0 = md.rebuilds()
    wall, _ = _timed_md(torch, md, a.steps, a.energy_every, sp)
    nreb = md.rebuilds() - r0
    lat = _sync_steps(torch, lambda: md.step(1, stream=sp, graph=True), a.lat_steps)
    pos_now, _ = md.get_state()
    tm = kernel_phases(torch, ctx, s, pos_now, sp)
    print(f"[bench] c3 mixed {1e3 * wall / a.steps:.3f} ms/step", file=sys.stderr, flush=True)
    # the plain FP64 path on the same system from the same state
    ctx64 = nb.NBContext(s, precision="fp64")
    md64 = nb.NBMD(ctx64, s, DT_PS, nres_intra=4)
    md64.set_state(pos_eq, vel_eq, s["hmat"])
    md64.init()
    md64.step(a.warmup, stream=sp, graph=True)
    wall64, _ = _timed_md(torch, md64, a.steps, a.energy_every, sp)
    tm64 = kernel_phases(torch, ctx64, s, pos_now, sp)
    print(f"[bench] c3 fp64 {1e3 * wall64 / a.steps:.3f} ms/step", file=sys.stderr, flush=True)
    ferr = force_err(nb, s, pos_eq)
    drift = {}
    for name, m in (("mixed", md), ("fp64", md64)):
        m.set_state(pos_eq, vel_eq, s["hmat"])
        m.init()
        drift[name] = _drift([m], [s], a.c3_drift_steps, lambda k, m=m: m.step(k, graph=True))
    npairs, cnt_rc = pairs_within_rc(nb, s, pos_eq)
    counts = ctx.list_counts()
    info = ctx.recip_info()
    md.close(); md64.close(); ctx.close(); ctx64.close()
    return s, dict(n=n, wall=wall, wall64=wall64, lat=lat, rebuilds=nreb, tm=tm, tm64=tm64,
                   ferr=ferr, drift=drift, npairs=npairs, cnt_rc=cnt_rc, counts=counts,
                   info=info)


def ewald_roofline(name, tm_ms, flop, mfma_insts=0):
    """A reciprocal-sum kernel against the FP64 peak (78.6 TFLOP/s, vector
    and matrix cores alike on MI355X).  achieved: the algorithmic FLOP per
    (k vector, charge) -- 10 structure factor, 16 force pass, cm3d's
    recursion and sums -- times K x Q, over the launch time.  The kernels
    pair kc = +m and -m and (on the matrix cores) share each charge's
    cos / sin of m theta_c across rows, so they execute fewer FP64
    operations than that count; a matrix-core kernel's executed FLOP (2048
    a v_mfma_f64_16x16x4f64, the count the library reports) are given
    beside it."""
    tf = flop / (tm_ms * 1e-3) / 1e12
    out = {"kernel": name, "bound": "mfma-fp64" if mfma_insts else "valu-fp64",
           "achieved": round(tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": round(tf / FP64_PEAK_TFLOPS, 5), "ms_per_launch": round(tm_ms, 5),
           "flop_per_launch": flop,
           "flop_basis": "algorithmic: per (k vector, charge) 10 FLOP structure factor, 16 FLOP "
                         "force pass (cm3d's c-axis recursion); the kernels serve kc = +m and -m "
                         "with one phase step (and, on the matrix cores, one cos/sin of m "
                         "theta_c per charge for all rows), so they execute fewer operations: "
                         "executed_* (matrix cores) or valu_issue_frac (VALU) give that work "
                         "against the peak"}
    if mfma_insts:
        ex = 2048.0 * mfma_insts / (tm_ms * 1e-3) / 1e12
        out.update(mfma_insts_per_launch=int(mfma_insts), executed_tflops=round(ex, 3),
                   executed_frac=round(ex / FP64_PEAK_TFLOPS, 5))
    return out


def run_c2(nb, S, a, torch):
    s = S.water(nside=16)
    n = int(s["natoms"])
    sp = torch.cuda.current_stream().cuda_stream
    res = {}
    eq = None
    for prec in ("fp64", "mixed"):
        ctx = nb.NBContext(s, precision=prec)
        md = nb.NBMD(ctx, s, DT_PS, nres_intra=4)
        if eq is None:
            md.set_state(s["pos"], s["vel"], s["hmat"])
            md.init()
            S.equilibrate([md], [s], 300.0, a.c2_equil, every=10)
            eq = md.get_state()
        else:
            md.set_state(eq[0], eq[1], s["hmat"])
            md.init()
        md.step(a.warmup, stream=sp, graph=True)
        r0 = md.rebuilds()
        wall, _ = _timed_md(torch, md, a.c2_steps, a.energy_every, sp)
        nreb = md.rebuilds() - r0
        tm = kernel_phases(torch, ctx, s, md.get_state()[0], sp, ncalls=20)
        md.set_state(eq[0], eq[1], s["hmat"])
        md.init()
        drift = _drift([md], [s], a.drift_steps, lambda k, m=md: m.step(k, graph=True))
        res[prec] = dict(wall=wall, tm=tm, drift=drift, rebuilds=nreb, counts=ctx.list_counts())
        md.close()
        ctx.close()
    npairs, _ = pairs_within_rc(nb, s, eq[0])
    return s, dict(n=n, res=res, ferr=force_err(nb, s, eq[0]), npairs=npairs)


def _c4_systems(S, nsys):
    rng = np.random.default_rng(2026)
    sizes = np.exp(rng.uniform(math.log(500), math.log(20000), nsys)).astype(int)
    return [S.mixture(int(nat), seed=100 + k, rc=9.0, kmax=7) for k, nat in enumerate(sizes)]


def _mds(nb, syss, prec, sharing, states=None):
    ctxs, mds = [], []
    for k, s in enumerate(syss):
        c = nb.NBContext(s, precision=prec, sharing=sharing)
        m = nb.NBMD(c, s, DT_PS, nres_intra=1)
        p, v = states[k] if states else (s["pos"], s["vel"])
        m.set_state(p, v, s["hmat"])
        m.init()
        ctxs.append(c)
        mds.append(m)
    return ctxs, mds


def _close(ctxs, mds, b=None):
    if b is not None:
        b.close()
    for m in mds:
        m.close()
    for c in ctxs:
        c.close()


def run_batch(nb, S, a, torch):
    syss = _c4_systems(S, a.nsys)
    ntot = int(sum(int(s["natoms"]) for s in syss))
    sp = torch.cuda.current_stream().cuda_stream
    # batched, mixed precision (the serving form)
    ctxs, mds = _mds(nb, syss, "mixed", a.nsys)
    b = nb.NBBatch(mds)
    S.equilibrate(mds, syss, 300.0, a.equil, every=10,
                  stepper=lambda k: b.step(k, stream=sp, graph=True))
    states = [m.get_state() for m in mds]
    b.step(10, stream=sp, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(a.batch_steps, stream=sp, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = _sync_steps(torch, lambda: b.step(1, stream=sp, graph=True), a.lat_steps)
    _close(ctxs, mds, b)
    print(f"[bench] batch mixed {1e3 * wall / a.batch_steps:.3f} ms/step", file=sys.stderr,
          flush=True)
    # the same systems from the same states, one at a time, plain FP64
    ctxs, mds = _mds(nb, syss, "fp64", 1, states)
    for m in mds:
        m.step(2, stream=sp, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in mds:
        m.step(a.batch_steps, stream=sp, graph=True)
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t0
    # the same FP64 contexts batched (what batching alone gains)
    b64 = nb.NBBatch(mds)
    b64.step(2, stream=sp, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b64.step(a.batch_steps, stream=sp, graph=True)
    torch.cuda.synchronize()
    wall_b64 = time.perf_counter() - t0
    print(f"[bench] batch fp64 one-at-a-time {1e3 * wall_seq / a.batch_steps:.3f} ms/step, "
          f"batched {1e3 * wall_b64 / a.batch_steps:.3f}", file=sys.stderr, flush=True)
    for m, st, s in zip(mds, states, syss):
        m.set_state(st[0], st[1], s["hmat"])
        m.init()
    drift64 = _drift(mds, syss, a.batch_drift_steps,
                     lambda k: b64.step(k, stream=sp, graph=True))
    _close(ctxs, mds, b64)
    ctxs, mds = _mds(nb, syss, "mixed", a.nsys, states)
    b = nb.NBBatch(mds)
    driftm = _drift(mds, syss, a.batch_drift_steps, lambda k: b.step(k, stream=sp, graph=True))
    _close(ctxs, mds, b)
    errs = [force_err(nb, s, st[0], sharing=a.nsys) for s, st in zip(syss, states)]
    return dict(natoms_total=ntot, nsys=a.nsys, wall=wall, wall_seq=wall_seq, wall_b64=wall_b64,
                lat=lat, drift_mixed=driftm, drift_fp64=drift64, errs=errs,
                sizes=[int(s["natoms"]) for s in syss])


def run_serve(a, torch):
    from cm3d_amd import serve
    jobs = serve.job_stream(a.serve_jobs, a.serve_rate, seed=2027)
    free, _ = torch.cuda.mem_get_info()
    budget = int(0.8 * free)
    srv = serve.Server(budget_bytes=budget, max_slots=a.serve_slots, tick_steps=50,
                       precision="mixed", keep_results=False)
    last = [time.perf_counter()]

    def log(msg):  # a progress line every ~20 s of the stream
        if time.perf_counter() - last[0] > 20.0:
            last[0] = time.perf_counter()
            print(f"[bench] serve {msg}", file=sys.stderr, flush=True)
    st = srv.run(jobs, log=log)
    st["budget_bytes"] = budget
    st["max_slots"] = a.serve_slots
    return st


def traffic_probe(nb, S, a, torch):
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):  # reads 256 MiB, writes 256 MiB per launch
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    if a.traffic_probe == "c3":
        s, prec = water_c3(S), "mixed"
    else:
        s, prec = S.water(nside=16), "mixed" if a.traffic_probe == "c2m" else "fp64"
    ctx = nb.NBContext(s, precision=prec)
    ctx.build_lists(s["pos"], s["hmat"])
    d_pos = torch.from_numpy(s["pos"]).cuda()
    d_fl = torch.zeros_like(d_pos)
    for _ in range(5):
        ctx.force_device(1, d_pos.data_ptr(), s["hmat"], 0, d_fl.data_ptr())
    torch.cuda.synchronize()
    ctx.close()
    print(json.dumps({"traffic_probe": a.traffic_probe, "neg_bytes_each_way": 256 << 20}),
          flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from cm3d_amd import nb, systems as S

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU (libcm3d_nb.so has no CPU path)")
    if a.traffic_probe:
        traffic_probe(nb, S, a, torch)
        return
    if a.only_batch:
        b = run_batch(nb, S, a, torch)
        print(json.dumps({"batched_atom_steps_per_s": b["natoms_total"] * a.batch_steps / b["wall"],
                          "ms_per_step": 1e3 * b["wall"] / a.batch_steps}), flush=True)
        return
    s, r = run_c3(nb, S, a, torch)
    n, wall = r["n"], r["wall"]
    atom_steps = n * a.steps / wall
    ns_day = (a.steps / wall) * DT_PS * 1e-3 * 86400.0
    ktot = int(s["ktot"])
    nq = int((s["qch"] != 0).sum())
    tm, tm64 = r["tm"], r["tm64"]
    # the dominant kernel of the mixed step: the longest phase of force(...,1)
    kq = float(ktot) * nq
    info = r["info"]
    # the structure-factor phase on the matrix cores is k_ewsfmtab (phase
    # tables) + k_ewsfacm
    ksf = "k_ewsfacm" if info["sfac_mfma"] else "k_ewsfac"
    kfo = "k_ewforcem" if info["force_mfma"] else "k_ewforce"
    phases = {ksf: (tm["sfac"], EW_SFAC_FLOP * kq, info["sfac_mfma_insts"]),
              kfo: (tm["ewforce"], EW_FORCE_FLOP * kq, info["force_mfma_insts"])}
    dom = max(phases, key=lambda k: phases[k][0])
    roof = ewald_roofline(dom, *phases[dom])
    pj = _profile_json(f"{dom}_water32768_traffic.json")
    roof["traffic"] = round(float(pj["traffic_bytes_per_launch"]), 0) if pj else None
    roof["traffic_source"] = (f"profiles/ewald_water32768_counters.md (rocprofv3 --pmc "
                              f"FETCH_SIZE / WRITE_SIZE, copy-calibrated)") if pj else None
    other = kfo if dom == ksf else ksf
    roof_o = ewald_roofline(other, *phases[other])
    for rf in (roof, roof_o):
        vj = _profile_json(f"{rf['kernel']}_water32768_valu.json")
        if vj:  # executed wave-instructions against the VALU issue peak
            insts = float(vj["valu_insts_per_launch"])
            rf["valu_insts_per_launch"] = insts
            rf["valu_issue_frac"] = round(insts / (VALU_ISSUE_PER_S * rf["ms_per_launch"] * 1e-3), 4)
    out = {
        "metric": "atom-steps/s", "value": round(atom_steps, 1), "unit": "atom-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 * wall / a.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pairs / int64 fixed-point forces / f64 Ewald, energies, integrator",
        "data": "synthetic",
        "config": {"workload": f"water{n // 3}: {n}-atom flexible 3-site water, box "
                               f"{s['hmat'][0]:.2f} A, rc 12 A, skin 1 A, Ewald alpha 0.30 kmax 32 "
                               f"({ktot} k), RESPA 4x0.5/2 fs, mixed precision, device-resident "
                               f"MD step (hipGraph), energies every {a.energy_every} steps",
                   "natoms": n, "ktot": ktot, "dt_fs": 2.0, "nres_intra": 4,
                   "energy_every": a.energy_every},
        "ns_per_day": round(ns_day, 3),
        "p50_step_ms": round(float(np.percentile(r["lat"], 50)), 5),
        "p99_step_ms": round(float(np.percentile(r["lat"], 99)), 5),
        "latency_basis": f"{a.lat_steps} single steps, each launched and waited for",
        "force_rel_rms_vs_fp64": r["ferr"],
        "fp64": {"atom_steps_per_s": round(n * a.steps / r["wall64"], 1),
                 "ms_per_step": round(1e3 * r["wall64"] / a.steps, 5),
                 "force_phase_ms": {k: round(v, 5) for k, v in tm64.items() if k != "calls"}},
        "mixed_vs_fp64": round(r["wall64"] / wall, 4),
        "energy_drift": {"mixed": r["drift"]["mixed"], "fp64": r["drift"]["fp64"]},
        "list_rebuilds": r["rebuilds"],
        "pairs_listed_half": r["counts"],
        "pairs_within_rc": r["npairs"],
        "pair_interactions_per_s": round(r["npairs"] * a.steps / wall, 1),
        "force_phase_ms": {k: round(v, 5) for k, v in tm.items() if k != "calls"},
        "roofline": roof,
        "roofline_other": {other: roof_o},
    }
    print(f"[bench] headline {atom_steps:.4g} atom-steps/s", file=sys.stderr, flush=True)
    if not a.no_c2:
        s2, r2 = run_c2(nb, S, a, torch)
        n2 = r2["n"]
        blk = {"workload": f"water{n2 // 3}: {n2}-atom flexible water, rc 10 A, skin 1 A, Ewald "
                           f"alpha 0.35 kmax 9 ({int(s2['ktot'])} k), RESPA 4x0.5/2 fs, energies "
                           f"every {a.energy_every} steps",
               "force_rel_rms_mixed_vs_fp64": r2["ferr"], "pairs_within_rc": r2["npairs"]}
        for prec, rr in r2["res"].items():
            w = rr["wall"]
            tmp = rr["tm"]
            pb = 4.0 * max(rr["counts"]["nprte"], rr["counts"]["nprt"]) + 44.0 * n2
            blk[prec] = {"atom_steps_per_s": round(n2 * a.c2_steps / w, 1),
                         "ms_per_step": round(1e3 * w / a.c2_steps, 5),
                         "ns_per_day": round(a.c2_steps / w * DT_PS * 1e-3 * 86400.0, 2),
                         "pair_interactions_per_s": round(r2["npairs"] * a.c2_steps / w, 1),
                         "force_phase_ms": {k: round(v, 5) for k, v in tmp.items() if k != "calls"},
                         "energy_drift": rr["drift"], "list_rebuilds": rr["rebuilds"]}
            kname = "k_pair64_1n" if prec == "fp64" else "k_mix_1f"
            gbs = pb / (tmp["pair"] * 1e-3) / 1e9
            roofp = {"kernel": kname, "bound": "hbm", "achieved": round(gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                     "ms_per_launch": round(tmp["pair"], 5), "bytes_per_launch": pb}
            pj = _profile_json(f"{kname}_water4096_traffic.json")
            if pj:
                roofp["traffic"] = round(float(pj["traffic_bytes_per_launch"]), 0)
            vj = _profile_json(f"{kname}_water4096_valu.json")
            if vj:
                insts = float(vj["valu_insts_per_launch"])
                roofp["valu_insts_per_launch"] = insts
                roofp["valu_issue_frac"] = round(insts / (VALU_ISSUE_PER_S * tmp["pair"] * 1e-3), 4)
            blk[prec]["roofline_pair"] = roofp
        out["water4096"] = blk
        print("[bench] water4096 done", file=sys.stderr, flush=True)
    if not a.no_batch:
        b = run_batch(nb, S, a, torch)
        nt = b["natoms_total"]
        bas = nt * a.batch_steps / b["wall"]
        seq = nt * a.batch_steps / b["wall_seq"]
        out["batched"] = {
            "workload": f"{b['nsys']} LJ+charge mixtures, N log-uniform 500-20000 ({nt} atoms), "
                        f"rc 9 A kmax 7",
            "atom_steps_per_s": round(bas, 1),
            "ms_per_step": round(1e3 * b["wall"] / a.batch_steps, 5),
            "dtype": "fp32 pairs / int64 fixed-point forces / f64 Ewald+energies",
            "fp64_one_at_a_time_atom_steps_per_s": round(seq, 1),
            "vs_fp64_one_at_a_time": round(bas / seq, 3),
            "fp64_batched_atom_steps_per_s": round(nt * a.batch_steps / b["wall_b64"], 1),
            "ratio_basis": "the same 64 systems from the same states: batched mixed (one set of "
                           "launches per step for all) vs each stepped alone in turn on the plain "
                           "FP64 path (its own graph per system)",
            "p50_step_ms": round(float(np.percentile(b["lat"], 50)), 5),
            "p99_step_ms": round(float(np.percentile(b["lat"], 99)), 5),
            "force_rel_rms_vs_fp64": {"max": max(b["errs"]), "mean": float(np.mean(b["errs"]))},
            "energy_drift": {"mixed": b["drift_mixed"], "fp64": b["drift_fp64"]},
        }
        print("[bench] batch done", file=sys.stderr, flush=True)
    if not a.no_serve:
        sv = run_serve(a, torch)
        out["serving"] = {
            "workload": f"{sv['jobs']} MD jobs, Poisson arrivals at {a.serve_rate}/s, each a "
                        f"random instance: half LJ+charge mixtures (N log-uniform 500-20000, rc "
                        f"9 A, kmax 7), half flexible water (n^3 molecules, n = 7..20: 343-8000, "
                        f"RESPA 4x0.5/2 fs), 200-5000 steps each, mixed precision, admitted by "
                        f"slot ({sv['max_slots']}) and HBM footprint",
            "atom_steps_per_s": round(sv["atom_steps_per_s"], 1),
            "wall_s": round(sv["wall_s"], 3),
            "p50_step_ms": round(sv["p50_step_ms"], 4), "p99_step_ms": round(sv["p99_step_ms"], 4),
            "p50_job_s": round(sv["p50_job_s"], 3), "p99_job_s": round(sv["p99_job_s"], 3),
            "mean_batch_systems": round(sv["mean_batch"], 2), "peak_slots": sv["peak_slots"],
            "batch_rebuilds": sv["batch_rebuilds"], "peak_bytes": sv["peak_bytes"],
            "budget_bytes": sv["budget_bytes"], "overflow_recoveries": sv["overflow_recoveries"],
            "failed_jobs": len(sv["failed"]),
        }
        print("[bench] serving done", file=sys.stderr, flush=True)
    if not a.no_cpu:
        rate, sample = cpu_baseline(s, a.cpu_seconds, r["rebuilds"] / a.steps)
        out["cpu_baseline"] = {"value": round(rate, 2), "unit": "atom-steps/s", "cores": 1,
                               "kind": "port", "sample": sample,
                               "host_cores_total": os.cpu_count()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
