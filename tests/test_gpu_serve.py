"""The continuous-batching server: jobs that arrive while others run are
admitted as slots and memory free up, and every job's final state equals
the state the same job reaches alone (bit for bit: a system's arithmetic is
the same in and out of a batch, and each job runs exactly its own steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs():
    from cm3d_amd import serve, systems as S
    specs = [("mix", 600, 30), ("water", 4, 45), ("mix", 1100, 20), ("water", 5, 60),
             ("mix", 800, 35), ("mix", 1500, 25), ("water", 4, 30)]
    jobs = []
    for i, (kind, size, nst) in enumerate(specs):
        if kind == "mix":
            s = S.mixture(size, seed=50 + i, rc=9.0, kmax=5)
            nri = 1
        else:
            s = S.water(nside=size, kmax=6, seed=60 + i)
            nri = 4
        vel = S.maxwell_velocities(s["amass"], 300.0, 70 + i)
        jobs.append(serve.Job(i, s, nst, arrival=0.0 if i < 3 else 0.001 * i, nres_intra=nri,
                              vel=vel))
    return jobs


def _alone(job, precision):
    from cm3d_amd import nb
    ctx = nb.NBContext(job.sys, precision=precision, sharing=job.sharing)
    md = nb.NBMD(ctx, job.sys, job.dt_ps, nres_intra=job.nres_intra)
    md.set_state(job.sys["pos"], job.vel, job.sys["hmat"])
    md.init()
    md.step(job.nsteps, graph=True)
    out = md.get_state()
    md.close()
    ctx.close()
    return out


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_serve_results_equal_alone(prec):
    from cm3d_amd import serve
    jobs = _jobs()
    print(f"[serve] {prec}: {len(jobs)} jobs, 3 slots", flush=True)
    # 3 slots and ticks of at most 10 steps: jobs queue, join running
    # batches and leave them at different steps
    srv = serve.Server(budget_bytes=1 << 34, max_slots=3, tick_steps=10, precision=prec)
    st = srv.run(jobs)
    assert st["batch_rebuilds"] >= 3 and st["peak_slots"] == 3
    for j in jobs:
        assert j.done == j.nsteps and j.result is not None
        pos, vel = _alone(j, prec)
        assert np.array_equal(j.result[0], pos), j.id
        assert np.array_equal(j.result[1], vel), j.id


def test_serve_admission_by_memory():
    """With a budget that holds about two of these jobs at a time, the server
    never exceeds it (beyond the one job it must admit to make progress)."""
    from cm3d_amd import serve
    jobs = _jobs()
    budget = int(2.2 * max(j.est_bytes for j in jobs))
    srv = serve.Server(budget_bytes=budget, max_slots=8, tick_steps=10)
    st = srv.run(jobs)
    assert st["peak_bytes"] <= budget
    assert all(j.done == j.nsteps for j in jobs)
    assert all(j.bytes <= j.est_bytes for j in jobs), [(j.bytes, j.est_bytes) for j in jobs]


@pytest.mark.parametrize("sharing", [1, 16])
def test_footprint_estimate_bounds_large_jobs(sharing):
    """Admission relies on estimate_bytes being an upper bound of what a job
    allocates: checked at the sizes of the serving stream's largest jobs
    (a 20,000-atom mixture at kmax 7, 8000 waters), in both precisions, and
    not more than 1.5x + 8 MiB of it."""
    from cm3d_amd import nb, serve, systems as S
    cases = [S.mixture(20000, seed=5, rc=9.0, kmax=7), S.water(nside=20, kmax=9, seed=6)]
    for s in cases:
        for prec in ("mixed", "fp64"):
            est = serve.estimate_bytes(s, sharing)
            ctx = nb.NBContext(s, precision=prec, sharing=sharing)
            md = nb.NBMD(ctx, s, 0.002, nres_intra=1)
            md.set_state(s["pos"], s["vel"], s["hmat"])
            md.init()
            got = ctx.footprint()
            md.close()
            ctx.close()
            print(f"  N={int(s['natoms'])} {prec} sharing={sharing}: {got} <= {est}", flush=True)
            assert got <= est <= 1.5 * got + (8 << 20), (int(s["natoms"]), prec, got, est)
