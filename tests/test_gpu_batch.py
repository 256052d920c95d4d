"""Batches: several heterogeneous systems stepped by one set of batched
launches give, system by system, the same trajectory as each system stepped
alone -- bit for bit (every system's arithmetic and summation order is the
same in and out of the batch)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems():
    from cm3d_amd import systems as S
    return [S.mixture(500, seed=3), S.water(nside=5, kmax=6), S.mixture(1400, seed=4),
            S.water(nside=6, kmax=7, seed=9)]


def _run_alone(s, prec, nsteps, graph):
    from cm3d_amd import nb
    ctx = nb.NBContext(s, precision=prec)
    md = nb.NBMD(ctx, s, 0.002, nres_intra=4 if "mol_start" in s else 1)
    md.set_state(s["pos"], s["vel"], s["hmat"])
    md.init()
    md.step(nsteps, graph=graph)
    out = md.get_state()
    md.close()
    ctx.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_equals_alone(prec):
    from cm3d_amd import nb
    nsteps = 12
    syss = _systems()
    print(f"[batch] {prec}: sizes {[s['natoms'] for s in syss]}", flush=True)
    alone = [_run_alone(s, prec, nsteps, graph=False) for s in syss]
    ctxs, mds = [], []
    for s in syss:
        c = nb.NBContext(s, precision=prec)
        m = nb.NBMD(c, s, 0.002, nres_intra=4 if "mol_start" in s else 1)
        m.set_state(s["pos"], s["vel"], s["hmat"])
        m.init()
        ctxs.append(c)
        mds.append(m)
    b = nb.NBBatch(mds)
    assert b.launches() < 40
    b.step(nsteps // 2, graph=False)
    b.step(nsteps - nsteps // 2, graph=True)
    for k, (m, ref) in enumerate(zip(mds, alone)):
        pos, vel = m.get_state()
        assert np.array_equal(pos, ref[0]), (prec, k)
        assert np.array_equal(vel, ref[1]), (prec, k)
    b.close()
    for m in mds:
        m.close()
    for c in ctxs:
        c.close()


def test_batch_equals_alone_across_rebuilds():
    """Long enough that every system rebuilds its lists on the device, each
    at its own steps: the batched step still reproduces each system alone
    bit for bit."""
    from cm3d_amd import nb
    nsteps = 160
    syss = _systems()
    print(f"[batch-long] fp64 {nsteps} steps", flush=True)
    alone = [_run_alone(s, "fp64", nsteps, graph=True) for s in syss]
    ctxs, mds = [], []
    for s in syss:
        c = nb.NBContext(s, precision="fp64")
        m = nb.NBMD(c, s, 0.002, nres_intra=4 if "mol_start" in s else 1)
        m.set_state(s["pos"], s["vel"], s["hmat"])
        m.init()
        ctxs.append(c)
        mds.append(m)
    b = nb.NBBatch(mds)
    b.step(nsteps, graph=True)
    for k, (m, ref) in enumerate(zip(mds, alone)):
        pos, vel = m.get_state()
        assert m.rebuilds() > 0, k
        assert np.array_equal(pos, ref[0]), k
        assert np.array_equal(vel, ref[1]), k
    b.close()
    for m in mds:
        m.close()
    for c in ctxs:
        c.close()
