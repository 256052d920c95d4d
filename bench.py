"""Benchmark of the MI355X nonbonded engine (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload (the configuration the metric is quoted on): 4096 flexible
3-site waters (12,288 atoms, cubic box 49.67 A, 1.0 g/cc), O-O Lennard-Jones
and erfc real-space Coulomb tabulated with 500 nodes, rc 10 A, skin 1 A,
Ewald alpha 0.35/A, kmax 9 (1535 half-space k vectors), harmonic O-H bonds
and H-O-H bends, RESPA with 4 x 0.5 fs inner and 2 fs outer steps, 300 K,
plain FP64.  A step is one outer MD step, entirely on the device and
replayed from a hipGraph: the inner bonded RESPA loop, cm3d's nrefsh
rebuild test (and the list build when it fires), force(...,1) -- LJ +
real-space Coulomb pairs, ecorr, Ewald, energies and virials -- and the
velocity kicks.  value = atom-steps/s of that single FP64 system.

Beside it, the batched form: 64 independent LJ + charge binary mixtures
(N log-uniform in [500, 20000] from the seed, rc 9 A, kmax 7, mixed
precision: FP32 pairs, 64-bit fixed-point forces, FP64 Ewald/energies)
stepped by one set of batched graph-replayed launches; its aggregate
atom-steps/s and its ratio to the FP64 single-system value are reported.

Also printed: ns/day, p50/p99 step latency, energy drift per ns for both
forms, the roofline of the dominant kernel from HIP events on the kernels'
stream, and the CPU baseline -- the clean-room C oracle of cm3d's CPU path
timed on this host on a bounded sample of the same workload.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nside", type=int, default=16, help="waters per box edge")
    ap.add_argument("--nsys", type=int, default=64, help="systems in the batched form")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--equil", type=int, default=600,
                    help="untimed steps with velocity rescaling to 300 K before timing")
    ap.add_argument("--drift-steps", type=int, default=40000)
    ap.add_argument("--lat-steps", type=int, default=100,
                    help="single synchronised steps timed for the latency percentiles")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=100,
                    help="jobs in the serving stream (the full stream is 2000)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="mean arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the full-size (98,304-atom) reciprocal-sum timing")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only a 256 MiB device copy (counter calibration) and 5 FP64 "
                         "force(...,1) calls on the water system, for rocprofv3 --pmc")
    ap.add_argument("--only-batch", action="store_true",
                    help="run the batched form alone (profiling); prints its own JSON")
    return ap.parse_args()


def cpu_baseline(sysd, budget_s):
    """The oracle (clean-room restatement of cm3d's CPU path, gcc -O2, one
    core) on the same system: list build + force(...,1) + bonded forces,
    timed until about budget_s seconds have run."""
    from oracle import oracle as O
    pos = sysd["pos"]
    h = sysd["hmat"]
    bonded = dict(ib=sysd["bond_i"], jb=sysd["bond_j"], bk=sysd["bond_k"], beq=sysd["bond_r0"],
                  ia=sysd["bend_i"], ja=sysd["bend_j"], ka=sysd["bend_k"], ak=sysd["bend_fk"],
                  aeq=sysd["bend_th0"])
    t0 = time.perf_counter()
    L = O.build_lists(sysd, pos, h)
    t_list = time.perf_counter() - t0
    nf = 0
    t0 = time.perf_counter()
    while True:
        O.force(sysd, L, pos, h, 1)
        for _ in range(4):  # the four inner RESPA bonded evaluations
            O.bonded_force(int(sysd["natoms"]), bonded, pos)
        nf += 1
        if time.perf_counter() - t0 > budget_s or nf >= 200:
            break
    return t_list, (time.perf_counter() - t0) / nf, nf


def _sync_steps(torch, step, nsteps):
    """Per-step latency as a caller sees it: launch one step, wait for it."""
    lat = []
    for _ in range(nsteps):
        t = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append(1e3 * (time.perf_counter() - t))
    return np.array(lat)


def run_single(nb, S, a, torch):
    s = S.water(nside=a.nside)
    n = int(s["natoms"])
    ctx = nb.NBContext(s, precision="fp64")
    md = nb.NBMD(ctx, s, DT_PS, nres_intra=4)
    md.set_state(s["pos"], s["vel"], s["hmat"])
    md.init()
    # the lattice start relaxes and heats: bring it to 300 K first (untimed)
    S.equilibrate([md], [s], 300.0, a.equil, every=10)
    pos_eq, vel_eq = md.get_state()
    sp = torch.cuda.current_stream().cuda_stream
    md.step(a.warmup, stream=sp, graph=True)
    torch.cuda.synchronize()
    r0 = md.rebuilds()
    t0 = time.perf_counter()
    md.step(a.steps, stream=sp, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nreb = md.rebuilds() - r0
    lat = _sync_steps(torch, lambda: md.step(1, stream=sp, graph=True), a.lat_steps)
    # kernel durations: force(...,1) on the current state with HIP events on
    # the stream the kernels run on
    pos, _ = md.get_state()
    ctx.build_lists(pos, s["hmat"])
    d_pos = torch.from_numpy(pos).cuda()
    d_fl = torch.zeros_like(d_pos)
    torch.cuda.synchronize()
    ctx.timing_start()
    for _ in range(20):  # force(...,1) as the MD step calls it: fl only (k_pair64_1f)
        ctx.force_device(1, d_pos.data_ptr(), s["hmat"], 0, d_fl.data_ptr(), stream=sp)
    torch.cuda.synchronize()
    tm = ctx.timing_read()
    counts = ctx.list_counts()
    # energy drift of both precisions from the same equilibrated state, and
    # the mixed form's force error against FP64 there
    ctxm = nb.NBContext(s, precision="mixed")
    mdm = nb.NBMD(ctxm, s, DT_PS, nres_intra=4)
    drift = {}
    nsamp = 20
    every = max(1, a.drift_steps // nsamp)
    t_ns = np.arange(nsamp + 1) * every * DT_PS * 1e-3
    for name, c, m in (("fp64", ctx, md), ("mixed", ctxm, mdm)):
        m.set_state(pos_eq, vel_eq, s["hmat"])
        m.init()
        ham = [m.energies()["ham"]]
        for _ in range(nsamp):
            m.step(every, graph=True)
            ham.append(m.energies()["ham"])
        ham = np.array(ham) / n
        fit = np.polyfit(t_ns, ham, 1)
        drift[name] = float(fit[0])
        resid = np.polyval(fit, t_ns) - ham
        drift[name + "_rms_about_fit_K_per_atom"] = float(np.sqrt(np.mean(resid ** 2)))
        # standard error of the slope: the two precisions' drifts are equal
        # when they agree within it
        drift[name + "_stderr"] = float(np.sqrt(np.sum(resid ** 2) / (len(t_ns) - 2)
                                                / np.sum((t_ns - t_ns.mean()) ** 2)))
        drift[name + "_ham0_K_per_atom"] = float(ham[0])
    t0 = time.perf_counter()
    mdm.step(a.steps, graph=True)
    torch.cuda.synchronize()
    wall_m = time.perf_counter() - t0
    ctx.build_lists(pos_eq, s["hmat"])
    _, fl64, _ = ctx.force(pos_eq, s["hmat"], iflg=1)
    ctxm.build_lists(pos_eq, s["hmat"])
    _, flm, _ = ctxm.force(pos_eq, s["hmat"], iflg=1)
    ferr = float(np.sqrt(((flm - fl64) ** 2).mean() / (fl64 ** 2).mean()))
    md.close(); mdm.close(); ctx.close(); ctxm.close()
    return s, dict(n=n, wall=wall, lat=lat, rebuilds=nreb, tm=tm, counts=counts,
                   drift=drift, wall_mixed=wall_m, mixed_force_err=ferr)


def run_batch(nb, S, a, torch):
    rng = np.random.default_rng(2026)
    sizes = np.exp(rng.uniform(math.log(500), math.log(20000), a.nsys)).astype(int)
    ctxs, mds, syss = [], [], []
    for k, nat in enumerate(sizes):
        s = S.mixture(int(nat), seed=100 + k, rc=9.0, kmax=7)
        c = nb.NBContext(s, precision="mixed", sharing=a.nsys)
        m = nb.NBMD(c, s, DT_PS, nres_intra=1)
        m.set_state(s["pos"], s["vel"], s["hmat"])
        m.init()
        ctxs.append(c); mds.append(m); syss.append(s)
    b = nb.NBBatch(mds)
    sp = torch.cuda.current_stream().cuda_stream
    S.equilibrate(mds, syss, 300.0, a.equil, every=10,
                  stepper=lambda k: b.step(k, stream=sp, graph=True))
    b.step(10, stream=sp, graph=True)
    torch.cuda.synchronize()
    e0 = [m.energies()["ham"] for m in mds]
    t0 = time.perf_counter()
    b.step(a.batch_steps, stream=sp, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e1 = [m.energies()["ham"] for m in mds]
    lat = _sync_steps(torch, lambda: b.step(1, stream=sp, graph=True), a.lat_steps)
    ntot = int(sum(int(s["natoms"]) for s in syss))
    drift = float(np.mean([(y - x) / int(s["natoms"]) for x, y, s in zip(e0, e1, syss)]))
    out = dict(natoms_total=ntot, nsys=a.nsys, wall=wall, lat=lat, launches=b.launches(),
               drift_K_per_atom=drift, sizes=[int(x) for x in sizes])
    b.close()
    for m in mds:
        m.close()
    for c in ctxs:
        c.close()
    return out


def run_c3_ewald(nb, S, torch):
    """The reciprocal sum at full size (98,304-atom water, alpha 0.30, kmax
    32: 68,532 k): force(...,1) phases timed with HIP events on the
    kernels' stream, the Ewald kernels against the FP64 peak."""
    s = S.water(nside=32, rc=12.0, alpha=0.30, kmax=32)
    ctx = nb.NBContext(s, precision="fp64")
    ctx.build_lists(s["pos"], s["hmat"])
    d_pos = torch.from_numpy(s["pos"]).cuda()
    d_fl = torch.zeros_like(d_pos)
    sp = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        ctx.force_device(1, d_pos.data_ptr(), s["hmat"], 0, d_fl.data_ptr(), stream=sp)
    torch.cuda.synchronize()
    ctx.timing_start()
    for _ in range(5):
        ctx.force_device(1, d_pos.data_ptr(), s["hmat"], 0, d_fl.data_ptr(), stream=sp)
    torch.cuda.synchronize()
    tm = ctx.timing_read()
    ctx.close()
    nq = int((s["qch"] != 0).sum())
    flop = 26.0 * int(s["ktot"]) * nq
    tf = flop / (tm["ewald"] * 1e-3) / 1e12
    return {"workload": f"{int(s['natoms'])}-atom water, alpha 0.30, kmax 32 ({int(s['ktot'])} k), "
                        f"FP64 force(...,1)",
            "kernels": "k_ewchg+k_ewsfac+k_ewkred+k_red+k_ewforce+k_ewfin", "bound": "valu-fp64",
            "achieved": round(tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / FP64_PEAK_TFLOPS, 5), "ms_per_launch": round(tm["ewald"], 5),
            "flop_per_launch": flop, "pair_ms": round(tm["pair"], 5),
            "force_ms": round(tm["total"], 5)}


def run_serve(a, torch):
    from cm3d_amd import serve
    jobs = serve.job_stream(a.serve_jobs, a.serve_rate, seed=2027)
    free, _ = torch.cuda.mem_get_info()
    budget = int(0.8 * free)
    srv = serve.Server(budget_bytes=budget, max_slots=a.serve_slots, tick_steps=50,
                       precision="mixed")
    st = srv.run(jobs)
    st["budget_bytes"] = budget
    st["max_slots"] = a.serve_slots
    return st


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
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):  # reads 256 MiB, writes 256 MiB per launch
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        s = S.water(nside=a.nside)
        ctx = nb.NBContext(s, precision="fp64")
        ctx.build_lists(s["pos"], s["hmat"])
        d_pos = torch.from_numpy(s["pos"]).cuda()
        d_fl = torch.zeros_like(d_pos)
        for _ in range(5):
            ctx.force_device(1, d_pos.data_ptr(), s["hmat"], 0, d_fl.data_ptr())
        torch.cuda.synchronize()
        ctx.close()
        print(json.dumps({"traffic_probe": "done", "neg_bytes_each_way": 256 << 20}), flush=True)
        return
    if a.only_batch:
        b = run_batch(nb, S, a, torch)
        print(json.dumps({"batched_atom_steps_per_s": b["natoms_total"] * a.batch_steps / b["wall"],
                          "ms_per_step": 1e3 * b["wall"] / a.batch_steps}), flush=True)
        return
    s, r = run_single(nb, S, a, torch)
    n = r["n"]
    wall = r["wall"]
    atom_steps = n * a.steps / wall
    ns_day = (a.steps / wall) * DT_PS * 1e-3 * 86400.0
    tm = r["tm"]
    counts = r["counts"]
    # dominant kernel roofline.  Pair kernel: 4 B per listed half pair (the
    # merged vdW/coulomb list holds the coulomb te pairs, which include the
    # O-O vdW pairs) + 44 B per atom.  Ewald: 26 FLOP per (k, charge).
    half_pairs = max(counts["nprte"], counts["nprt"])
    pair_bytes = 4.0 * half_pairs + 44.0 * n
    ktot = int(s["ktot"])
    nq = int((s["qch"] != 0).sum())
    ew_flop = 26.0 * ktot * nq
    pair_gbs = pair_bytes / (tm["pair"] * 1e-3) / 1e9
    ew_tfs = ew_flop / (tm["ewald"] * 1e-3) / 1e12
    # the pair kernel is the longest single kernel of the step and is bound
    # by HBM (a gather over the list); the Ewald kernels together are bound
    # by FP64 VALU throughput and are reported beside it against that roof
    traffic, tsrc = None, None
    tpath = os.path.join(ROOT, "profiles", "pair_kernel_water4096_traffic.json")
    if os.path.exists(tpath) and a.nside == 16:
        with open(tpath) as f:
            tj = json.load(f)
        traffic = round(float(tj["traffic_bytes_per_launch"]), 0)
        tsrc = "profiles/pair_kernel_water4096_traffic.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)"
    roof = {"kernel": "k_pair64_1f", "bound": "hbm", "achieved": round(pair_gbs, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(pair_gbs / HBM_PEAK_GBS, 5),
            "traffic": traffic, "traffic_source": tsrc,
            "ms_per_launch": round(tm["pair"], 5), "bytes_per_launch": pair_bytes}
    # what does bound the kernel: vector instruction issue.  The launch's
    # VALU instruction count (rocprofv3 SQ_INSTS_VALU, committed beside the
    # traffic) over what 256 CUs issue at one wave instruction per cycle at
    # 2.4 GHz in the launch time measured here
    vpath = os.path.join(ROOT, "profiles", "pair_kernel_water4096_valu.json")
    if os.path.exists(vpath) and a.nside == 16:
        with open(vpath) as f:
            vj = json.load(f)
        insts = float(vj["valu_insts_per_launch"])
        roof["valu_insts_per_launch"] = insts
        roof["valu_issue_frac"] = round(insts / (VALU_ISSUE_PER_S * tm["pair"] * 1e-3), 4)
        roof["valu_source"] = "profiles/pair_kernel_water4096_valu.md (rocprofv3 --pmc SQ_INSTS_VALU)"
    roof_ew = {"kernels": "k_ewchg+k_ewsfac+k_ewkred+k_red+k_ewforce+k_ewfin",
               "bound": "valu-fp64", "achieved": round(ew_tfs, 3), "peak": FP64_PEAK_TFLOPS,
               "unit": "TFLOP/s", "frac": round(ew_tfs / FP64_PEAK_TFLOPS, 5),
               "ms_per_launch": round(tm["ewald"], 5), "flop_per_launch": ew_flop}
    out = {
        "metric": "atom-steps/s", "value": round(atom_steps, 1), "unit": "atom-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 * wall / a.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"water{n // 3}: {n}-atom flexible 3-site water, rc 10 A, skin 1 A, "
                               f"Ewald alpha 0.35 kmax 9 ({ktot} k), RESPA 4x0.5/2 fs, FP64, "
                               f"device-resident MD step (hipGraph)",
                   "natoms": n, "ktot": ktot, "dt_fs": 2.0, "nres_intra": 4},
        "ns_per_day": round(ns_day, 2),
        "p50_step_ms": round(float(np.percentile(r["lat"], 50)), 5),
        "p99_step_ms": round(float(np.percentile(r["lat"], 99)), 5),
        "latency_basis": f"{a.lat_steps} single steps, each launched and waited for",
        "energy_drift_K_per_atom_per_ns": {k: round(v, 6) for k, v in r["drift"].items()},
        "drift_run": f"slope of a linear fit to Ham/N sampled every {max(1, a.drift_steps // 20)} "
                     f"steps over {a.drift_steps} NVE steps from the same 300 K state, each precision; "
                     f"*_stderr: the slope's standard error",
        "mixed_single": {"atom_steps_per_s": round(n * a.steps / r["wall_mixed"], 1),
                         "force_rel_rms_vs_fp64": r["mixed_force_err"]},
        "list_rebuilds": r["rebuilds"],
        "pairs_half": counts,
        "force_phase_ms": {k: round(v, 5) for k, v in tm.items() if k != "calls"},
        "pair_interactions_per_s": round(counts["nprte"] * a.steps / wall, 1),
        "roofline": roof,
        "roofline_ewald": roof_ew,
        "pair_kernel_gbs": round(pair_gbs, 2),
        "ewald_tflops": round(ew_tfs, 4),
    }
    if not a.no_c3:
        out["roofline_ewald_c3"] = run_c3_ewald(nb, S, torch)
    if not a.no_batch:
        b = run_batch(nb, S, a, torch)
        bas = b["natoms_total"] * a.batch_steps / b["wall"]
        bt_ns = a.batch_steps * DT_PS * 1e-3
        out["batched"] = {
            "workload": f"{b['nsys']} LJ+charge mixtures, N log-uniform 500-20000 "
                        f"({b['natoms_total']} atoms), rc 9 A kmax 7, mixed precision, "
                        f"one batched graph per step",
            "atom_steps_per_s": round(bas, 1),
            "vs_fp64_single": round(bas / atom_steps, 3),
            "ms_per_step": round(1e3 * b["wall"] / a.batch_steps, 5),
            "p50_step_ms": round(float(np.percentile(b["lat"], 50)), 5),
            "p99_step_ms": round(float(np.percentile(b["lat"], 99)), 5),
            "launches_per_step": b["launches"],
            "energy_drift_K_per_atom_per_ns": round(b["drift_K_per_atom"] / bt_ns, 6),
            "dtype": "fp32 pairs / int64 fixed-point forces / f64 Ewald+energies",
        }
    if not a.no_serve:
        sv = run_serve(a, torch)
        out["serving"] = {
            "workload": f"{sv['jobs']} MD jobs, Poisson arrivals at {a.serve_rate}/s, half LJ+charge "
                        f"mixtures (N log-uniform 500-20000, rc 9 A, kmax 7), half flexible water "
                        f"(343-8000 molecules: n^3 for n = 7..20, within 256-8192; RESPA 4x0.5/2 fs), 200-5000 steps each, mixed "
                        f"precision, admitted by slot ({sv['max_slots']}) and HBM footprint",
            "atom_steps_per_s": round(sv["atom_steps_per_s"], 1),
            "wall_s": round(sv["wall_s"], 3),
            "p50_step_ms": round(sv["p50_step_ms"], 4), "p99_step_ms": round(sv["p99_step_ms"], 4),
            "p50_job_s": round(sv["p50_job_s"], 3), "p99_job_s": round(sv["p99_job_s"], 3),
            "mean_batch_systems": round(sv["mean_batch"], 2), "peak_slots": sv["peak_slots"],
            "batch_rebuilds": sv["batch_rebuilds"], "peak_bytes": sv["peak_bytes"],
            "budget_bytes": sv["budget_bytes"],
        }
    if not a.no_cpu:
        t_list, t_step, nf = cpu_baseline(s, a.cpu_seconds)
        rate = r["rebuilds"] / a.steps
        t = t_step + rate * t_list
        out["cpu_baseline"] = {"value": round(n / t, 1), "unit": "atom-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{nf} steps of the same {n}-atom water: force(...,1) "
                                         f"+ 4 bonded evaluations each, plus list builds at the "
                                         f"GPU run's measured rate ({rate:.3f}/step)",
                               "host_cores_total": os.cpu_count()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
