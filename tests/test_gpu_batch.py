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


@pytest.mark.parametrize("mfma", [None, "1"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_equals_alone(prec, mfma, monkeypatch):
    """mfma "1": every member's structure factor on the matrix cores (its
    batched table and product kernels)."""
    from cm3d_amd import nb
    if mfma:
        monkeypatch.setenv("CM3D_NB_SFAC_MFMA", mfma)
    nsteps = 12
    syss = _systems()
    print(f"[batch] {prec} mfma={mfma}: sizes {[s['natoms'] for s in syss]}", flush=True)
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


def _make(s, prec="fp64"):
    from cm3d_amd import nb
    c = nb.NBContext(s, precision=prec)
    m = nb.NBMD(c, s, 0.002, nres_intra=4 if "mol_start" in s else 1)
    m.set_state(s["pos"], s["vel"], s["hmat"])
    m.init()
    return c, m


def test_batch_recovers_from_list_overflow():
    """A member whose neighbour slots are reserved far too small (64 a row
    for a water whose rows hold ~500) overflows at its first device-side
    rebuild.  The tick detects it (sticky flag, read after the tick),
    restores the member's snapshot, regrows its slots, rebuilds and redoes
    its steps alone: the batch reports one recovery, the member's trajectory
    equals the same system run alone with ample slots (to rounding: its
    lists were built at other positions, so its pair sums run in another
    order), and the other member is untouched, bit for bit.  The condition
    is forced from the host; every kernel stays within its slots."""
    from cm3d_amd import nb, systems as S
    syss = [S.mixture(900, seed=21), S.water(nside=6, kmax=7, seed=22)]
    nsteps = 10
    alone = [_run_alone(s, "fp64", nsteps, graph=False) for s in syss]
    made = [_make(s) for s in syss]
    made[1][0].list_reserve(64)
    b = nb.NBBatch([m for _, m in made])
    print("[batch-overflow] stepping", flush=True)
    b.step(nsteps, graph=True)
    assert b.recoveries() == 1
    p0, v0 = made[0][1].get_state()
    assert np.array_equal(p0, alone[0][0]) and np.array_equal(v0, alone[0][1])
    p1, v1 = made[1][1].get_state()
    assert np.max(np.abs(p1 - alone[1][0])) < 1e-8
    assert np.max(np.abs(v1 - alone[1][1])) < 1e-6 * np.max(np.abs(alone[1][1]))
    made[1][1].rebuilds()  # raises on an unrecovered overflow
    assert made[1][1].energies()["overflow"] == 0
    # the regrown member is re-recorded into the batch and steps on
    b.step(4, graph=True)
    assert b.recoveries() == 1
    b.close()
    for c, m in made:
        m.close()
        c.close()


def _nhc(m, n, nchain=3):
    gkt = np.full(nchain, 300.0)
    gkt[0] *= 3 * n
    return dict(nchain=nchain, gkt=gkt, mnh=gkt * 0.05 ** 2, veta0=np.zeros(nchain),
                ithm=np.zeros(n, np.int32))


def test_batch_rerecords_after_members_change():
    """Setting a thermostat on the members of an existing batch changes
    their kernel sequence and reallocates their chain buffers: the batch
    records its step again before the next tick (and then matches the same
    history run alone, bit for bit); with only one member changed the
    members' steps differ and the tick is refused on the host (nothing is
    launched)."""
    from cm3d_amd import nb, systems as S
    syss = [S.mixture(700, seed=31), S.mixture(1300, seed=32)]
    alone = []
    for s in syss:
        c, m = _make(s)
        m.step(5, graph=False)
        m.set_thermostat(**_nhc(m, int(s["natoms"])))
        m.step(5, graph=False)
        alone.append(m.get_state())
        m.close()
        c.close()
    made = [_make(s) for s in syss]
    b = nb.NBBatch([m for _, m in made])
    b.step(5, graph=True)
    for (c, m), s in zip(made, syss):
        m.set_thermostat(**_nhc(m, int(s["natoms"])))
    b.step(5, graph=True)
    for (c, m), ref in zip(made, alone):
        pos, vel = m.get_state()
        assert np.array_equal(pos, ref[0]) and np.array_equal(vel, ref[1])
    made[0][1].set_thermostat(0, [], [], [], np.zeros(0, np.int32))
    with pytest.raises(nb.NBError):
        b.step(1, graph=True)
    b.close()
    for c, m in made:
        m.close()
        c.close()


def test_energies_of_a_batch_member():
    """energies() on a member right after a batched tick (no get_state in
    between) sees the finished step: it equals the same system's energies
    after the same steps alone, bit for bit."""
    from cm3d_amd import nb, systems as S
    syss = [S.water(nside=5, kmax=6, seed=41), S.mixture(800, seed=42)]
    ref = []
    for s in syss:
        c, m = _make(s)
        m.step(8, graph=True)
        ref.append(m.energies())
        m.close()
        c.close()
    made = [_make(s) for s in syss]
    b = nb.NBBatch([m for _, m in made])
    b.step(8, graph=True)
    for k in (1, 0):
        e = made[k][1].energies()
        for key in ("poter", "poth", "potra", "zke", "prsi", "ham"):
            assert e[key] == ref[k][key], (k, key, e[key], ref[k][key])
    b.close()
    for c, m in made:
        m.close()
        c.close()
