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

TRAJ_CASES = ["ref_lj2", "ref_lj256", "ar864", "mix600", "water512"]


def _system(name):
    from cm3d_amd import systems as S
    d = load_fixture(name)
    if name == "water512":
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
    assert md.rebuilds() >= 0
    md.close()
    ctx.close()
