"""GPU parity of the device-resident RESPA step against cm3d's own step()
trajectories (the fixtures' traj_pos / traj_vel, written by the compiled
reference from the same initial state).  The nonbonded force sums run in a
different order than cm3d's, so positions agree to rounding that grows over
the few steps compared: tolerance 1e-9 A on positions and 1e-9 relative RMS
on velocities."""
import numpy as np
import pytest

from conftest import load_fixture, rel_rms

pytestmark = pytest.mark.gpu

TRAJ_CASES = ["ref_lj2", "ref_lj256", "ar864", "mix600", "water512", "water512_tri",
              "mix600_tri"]


def _system(name):
    from cm3d_amd import systems as S
    d = load_fixture(name)
    if name.startswith("water512"):
        w = S.water(nside=8)
        for k in ("mol_start", "bond_i", "bond_j", "bond_k", "bond_r0", "bend_i", "bend_j",
                  "bend_k", "bend_fk", "bend_th0"):
            d[k] = w[k]
    return d


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("name", TRAJ_CASES)
def test_md_trajectory_matches_cm3d_step(name, graph):
    from cm3d_amd import nb
    d = _system(name)
    print(f"[md] {name} graph={graph}", flush=True)
    ctx = nb.NBContext(d, precision="fp64")
    md = nb.NBMD(ctx, d, float(d["dt"][0]), nres_intra=int(d["ninnra"][0]))
    md.set_state(d["traj_pos0"], d["traj_vel0"], d["hmat"])
    md.init()
    for s in range(1, int(d["traj_nsteps"][0]) + 1):
        md.step(1, graph=graph)
        pos, vel = md.get_state()
        assert np.abs(pos - d[f"traj_pos{s}"]).max() < 1e-9, (name, s)
        assert rel_rms(vel, d[f"traj_vel{s}"]) < 1e-9, (name, s)
    e = md.energies()
    assert abs(e["poter"] - d[f"traj_poter{s}"][0]) <= 1e-9 * max(1.0, abs(d[f"traj_poter{s}"][0]))
    assert abs(e["zke"] - d[f"traj_zke{s}"][0]) <= 1e-9 * abs(d[f"traj_zke{s}"][0])
    assert abs(e["potra"] - d[f"traj_potra{s}"][0]) <= 1e-9 * max(1.0, abs(d[f"traj_potra{s}"][0]))
    # geteng's other outputs: poth (shifted pair energy), pressure, temperature
    md.ndof = int(d["ndof"][0])
    e = md.energies()
    for k in ("poth", "prsi", "tiout"):
        ref = d[f"traj_{k}{s}"][0]
        assert abs(e[k] - ref) <= 1e-9 * max(1.0, abs(ref)), (name, k, e[k], ref)
    assert md.rebuilds() >= 0
    md.close()
    ctx.close()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("name", ["water512", "water512_tri"])
def test_md_trajectory_matrix_core_reciprocal(name, graph, monkeypatch):
    """The same trajectories with both reciprocal passes forced onto the
    matrix cores, as large systems run them: the step then launches the
    list build, pair kernel and reciprocal sum on one stream instead of
    forking the reciprocal sum onto the context's second stream."""
    from cm3d_amd import nb
    d = _system(name)
    print(f"[md] {name} matrix-core reciprocal graph={graph}", flush=True)
    monkeypatch.setenv("CM3D_NB_SFAC_MFMA", "1")
    ctx = nb.NBContext(d, precision="fp64")
    monkeypatch.delenv("CM3D_NB_SFAC_MFMA")
    assert ctx.recip_info()["sfac_mfma"] and ctx.recip_info()["force_mfma"]
    md = nb.NBMD(ctx, d, float(d["dt"][0]), nres_intra=int(d["ninnra"][0]))
    md.set_state(d["traj_pos0"], d["traj_vel0"], d["hmat"])
    md.init()
    for s in range(1, int(d["traj_nsteps"][0]) + 1):
        md.step(1, graph=graph)
        pos, vel = md.get_state()
        assert np.abs(pos - d[f"traj_pos{s}"]).max() < 1e-9, (name, s)
        assert rel_rms(vel, d[f"traj_vel{s}"]) < 1e-9, (name, s)
    e = md.energies()
    assert abs(e["poter"] - d[f"traj_poter{s}"][0]) <= 1e-9 * max(1.0, abs(d[f"traj_poter{s}"][0]))
    md.close()
    ctx.close()


def _long_run(sysd, prec, nsteps, graph, nres_intra=4):
    from cm3d_amd import nb
    ctx = nb.NBContext(sysd, precision=prec)
    md = nb.NBMD(ctx, sysd, 0.002, nres_intra=nres_intra)
    md.set_state(sysd["pos"], sysd["vel"], sysd["hmat"])
    md.init()
    e0 = md.energies()
    md.step(nsteps, graph=graph)
    pos, vel = md.get_state()
    e1 = md.energies()
    nreb = md.rebuilds()
    md.close()
    ctx.close()
    return pos, vel, e0, e1, nreb


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_long_run_rebuilds_lists_on_device(prec):
    """Several hundred steps of 648-atom water: the device-side nrefsh test
    must trigger list rebuilds, after which the list the integrator holds
    gives the same nonbonded energies as lists built from scratch by the
    oracle on the final positions (a stale list would miss pairs that moved
    inside rc+skin).  Energies come from the FP64 energy pass in either
    precision: 1e-10 relative."""
    from cm3d_amd import systems as S
    from oracle import oracle as O
    s = S.water(nside=6, kmax=7, seed=5)
    nsteps = 300
    print(f"[md-long] water648 {prec} {nsteps} steps", flush=True)
    pos, vel, e0, e1, nreb = _long_run(s, prec, nsteps, graph=True)
    assert nreb >= 3, nreb
    assert e1["overflow"] == 0
    L = O.build_lists(s, pos, s["hmat"])
    _, _, oe = O.force(s, L, pos, s["hmat"], 1)
    assert abs(e1["vinter"] - oe["spotl"]) <= 1e-10 * abs(oe["spotl"]), (e1["vinter"], oe["spotl"])
    assert abs(e1["velec"] - oe["spote"]) <= 1e-10 * abs(oe["spote"]), (e1["velec"], oe["spote"])


def test_md_mixed_conserves_ham_like_fp64():
    """Energy conservation of the two precisions from the same state: the
    648-atom water after 1000 NVE steps from its lattice start (where it
    relaxes and heats), then 600 more steps in each precision with Ham
    sampled every 20.  Ham (in units of the kinetic energy) shows no
    systematic drift in either run -- the slope of a linear fit within four
    standard errors of zero -- the two slopes agree within three standard
    errors of their difference, and the fluctuation about the fit (the
    2 fs RESPA step's own, about 1e-3 here) is no larger in the mixed run
    than 1.5 times the FP64 one."""
    from cm3d_amd import nb, systems as S
    s = S.water(nside=6, kmax=7, seed=5)
    pos, vel, _, _, _ = _long_run(s, "fp64", 1000, graph=True)
    fit = {}
    for prec in ("fp64", "mixed"):
        ctx = nb.NBContext(s, precision=prec)
        md = nb.NBMD(ctx, s, 0.002, nres_intra=4)
        md.set_state(pos, vel, s["hmat"])
        md.init()
        t, h = [], []
        for k in range(31):
            e = md.energies()
            t.append(20 * k)
            h.append(e["ham"])
            zke = e["zke"] if k == 0 else zke
            if k < 30:
                md.step(20, graph=True)
        md.close()
        ctx.close()
        t, h = np.array(t, float), (np.array(h) - h[0]) / abs(zke)
        A = np.vstack([t, np.ones_like(t)]).T
        coef, res, *_ = np.linalg.lstsq(A, h, rcond=None)
        sig2 = float(res[0]) / (len(t) - 2) if len(res) else 0.
        se = np.sqrt(sig2 / np.sum((t - t.mean()) ** 2))
        fit[prec] = (coef[0], se, float(np.sqrt(sig2)))
        print(f"[md-ham] {prec}: slope {coef[0]:.3e} +- {se:.1e} /step (of zke), "
              f"rms about the fit {fit[prec][2]:.2e}", flush=True)
    (sf, ef, rf), (sm, em, rm) = fit["fp64"], fit["mixed"]
    assert abs(sf) <= 4. * ef and abs(sm) <= 4. * em, fit
    assert abs(sm - sf) <= 3. * np.hypot(ef, em), fit
    assert rm <= 1.5 * rf, fit


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_energies_reuse_step_reciprocal_columns(prec):
    """After a step, nb_md_energies reduces the reciprocal-sum and ecorr
    energy columns of the step's own force(...,1) instead of recomputing
    them; after any other force evaluation on the context it recomputes
    them.  Both give the same energies bit for bit."""
    from cm3d_amd import nb, systems as S
    s = S.water(nside=6, kmax=7, seed=5)
    ctx = nb.NBContext(s, precision=prec)
    md = nb.NBMD(ctx, s, 0.002, nres_intra=4)
    md.set_state(s["pos"], s["vel"], s["hmat"])
    md.init()
    md.step(7, graph=True)
    e_reuse = md.energies()
    pos, _ = md.get_state()
    # a force evaluation through the API on the same context and positions
    # (same lists): the columns are no longer marked as the state's
    ctx.force(pos, s["hmat"], iflg=1)
    e_full = md.energies()
    md.close()
    ctx.close()
    for k, v in e_full.items():
        assert np.array_equal(np.asarray(e_reuse[k]), np.asarray(v)), (k, e_reuse[k], v)
