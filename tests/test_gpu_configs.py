"""GPU parity on the benchmark configurations at their own sizes:

- the 4096-molecule water (12,288 atoms, box 49.7 A, rc 10 A, skin 1 A,
  Ewald alpha 0.35 with kmax 9: 1535 k-vectors): Verlet lists bit-exact
  against the oracle's, FP64 forces within 1e-11 relative RMS and energies /
  virials within 1e-10 of the oracle, the mixed form within the 1e-4 force
  and 1e-5 energy bars;
- the 64-system batch of LJ + charge mixtures (N log-uniform in 500-20,000,
  rc 9 A, kmax 7), as the benchmark builds it: each system's mixed forces
  within 1e-4 of its own FP64 forces, and a few batched steps of all 64
  equal to each system stepped alone, bit for bit;
- a slice of the serving stream (jobs drawn from the same families and
  ranges: waters of 343-8000 molecules, mixtures of 500-20,000 atoms, each
  its own random instance, arriving while others run): every job's final
  state equals the job run alone, bit for bit, no list overflow, no failed
  job."""
import math

import numpy as np
import pytest

from conftest import rel_rms
from make_fixtures import pair_digest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
MIXED_FORCE_TOL = 1e-4
MIXED_ENERGY_TOL = 1e-5


@pytest.fixture(scope="module")
def c2():
    from cm3d_amd import systems as S
    return S.water(nside=16)


@pytest.fixture(scope="module")
def c2_oracle(c2):
    from oracle import oracle as O
    L = O.build_lists(c2, c2["pos"], c2["hmat"])
    fs, fl, e = O.force(c2, L, c2["pos"], c2["hmat"], 1)
    return L, fs, fl, e


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_c2_lists_bit_exact(c2, c2_oracle):
    from cm3d_amd import nb
    print("[c2] lists", flush=True)
    L = c2_oracle[0]
    c = nb.NBContext(c2, precision="fp64")
    c.build_lists(c2["pos"], c2["hmat"])
    cnt = c.list_counts()
    for t, (a, b) in (("t", ("inpt", "jnpt")), ("s", ("inps", "jnps")),
                      ("te", ("inpte", "jnpte")), ("se", ("inpse", "jnpse"))):
        i, j = c.get_pairs(t)
        assert len(i) == len(L[a]) == cnt["npr" + t], (t, len(i), len(L[a]))
        assert pair_digest(i, j) == pair_digest(L[a], L[b]), t
    print(f"  counts {cnt}", flush=True)
    c.close()


def test_c2_fp64_forces_and_energies(c2, c2_oracle):
    from cm3d_amd import nb
    print("[c2] fp64 force(...,1) and force(...,0)", flush=True)
    _, ofs, ofl, oe = c2_oracle
    c = nb.NBContext(c2, precision="fp64")
    c.build_lists(c2["pos"], c2["hmat"])
    fs, fl, e = c.force(c2["pos"], c2["hmat"], iflg=1)
    assert rel_rms(fl, ofl) < FP64_TOL, rel_rms(fl, ofl)
    assert rel_rms(fs, ofs) < FP64_TOL, rel_rms(fs, ofs)
    for k in ("spotl", "spothl", "spote", "spothe", "vk_ewald", "pot_ecorr", "sprsl",
              "sprse", "vp_ewald", "prs_ecorr"):
        assert _rel(e[k], oe[k]) < 1e-10, (k, e[k], oe[k])
    for k in ("sprslt", "sprset", "vp_ewald_tensor", "prs_ecorr_tensor"):
        scale = np.max(np.abs(oe[k]))
        assert np.max(np.abs(np.asarray(e[k]) - oe[k])) < 1e-10 * scale, k
    c.close()


def test_c2_mixed_within_bars(c2, c2_oracle):
    from cm3d_amd import nb
    print("[c2] mixed", flush=True)
    _, _, ofl, oe = c2_oracle
    c = nb.NBContext(c2, precision="mixed")
    c.build_lists(c2["pos"], c2["hmat"])
    _, fl, e = c.force(c2["pos"], c2["hmat"], iflg=1)
    err = rel_rms(fl, ofl)
    print(f"  force rel RMS {err:.2e}", flush=True)
    assert err < MIXED_FORCE_TOL
    for k in ("spotl", "spote", "sprsl", "sprse"):  # the FP64 energy pass
        r = abs(e[k] - oe[k]) / max(abs(oe[k]), 1.0)
        print(f"  {k} rel {r:.2e}", flush=True)
        assert r < min(MIXED_ENERGY_TOL, 1e-10), (k, e[k], oe[k])
    assert _rel(e["vk_ewald"], oe["vk_ewald"]) < 1e-10
    c.close()


def _c4_systems(nsys=64, seed=2026):
    """The benchmark's batch: N log-uniform in [500, 20000] from the seed."""
    from cm3d_amd import systems as S
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(500), math.log(20000), nsys)).astype(int)
    return [S.mixture(int(n), seed=100 + k, rc=9.0, kmax=7) for k, n in enumerate(sizes)]


def test_c4_batch_mixed_forces_and_batch_equals_alone():
    from cm3d_amd import nb
    syss = _c4_systems()
    nsteps = 6
    print(f"[c4] 64 systems, {sum(int(s['natoms']) for s in syss)} atoms", flush=True)
    worst = 0.
    alone = []
    for k, s in enumerate(syss):
        out = {}
        for prec in ("fp64", "mixed"):
            c = nb.NBContext(s, precision=prec, sharing=64)
            c.build_lists(s["pos"], s["hmat"])
            out[prec] = c.force(s["pos"], s["hmat"], iflg=1)
            c.close()
        err = rel_rms(out["mixed"][1], out["fp64"][1])
        worst = max(worst, err)
        assert err < MIXED_FORCE_TOL, (k, int(s["natoms"]), err)
        for key in ("spotl", "spote"):
            ref = out["fp64"][2][key]
            assert abs(out["mixed"][2][key] - ref) <= MIXED_ENERGY_TOL * max(abs(ref), 1.), (k, key)
        c = nb.NBContext(s, precision="mixed", sharing=64)
        m = nb.NBMD(c, s, 0.002)
        m.set_state(s["pos"], s["vel"], s["hmat"])
        m.init()
        m.step(nsteps, graph=True)
        alone.append(m.get_state())
        m.close()
        c.close()
    print(f"  worst mixed force rel RMS {worst:.2e}", flush=True)
    made = []
    for s in syss:
        c = nb.NBContext(s, precision="mixed", sharing=64)
        m = nb.NBMD(c, s, 0.002)
        m.set_state(s["pos"], s["vel"], s["hmat"])
        m.init()
        made.append((c, m))
    b = nb.NBBatch([m for _, m in made])
    b.step(nsteps, graph=True)
    assert b.recoveries() == 0
    for k, ((c, m), ref) in enumerate(zip(made, alone)):
        pos, vel = m.get_state()
        assert np.array_equal(pos, ref[0]) and np.array_equal(vel, ref[1]), k
    b.close()
    for c, m in made:
        m.close()
        c.close()


def test_c5_stream_slice_served_equals_alone():
    from cm3d_amd import nb, serve
    # 20 jobs of the serving families at the full size ranges, all arriving
    # within a fraction of a millisecond so that they queue behind 6 slots
    # and join running batches as slots free up; short runs (the test checks
    # the mix, not the run length)
    jobs = serve.job_stream(20, 1e5, seed=77, steps=(10, 40))
    for j in jobs:
        j.prepare()
    sizes = sorted(j.natoms for j in jobs)
    print(f"[c5] 20 jobs, atoms {sizes[0]}..{sizes[-1]}", flush=True)
    srv = serve.Server(budget_bytes=1 << 36, max_slots=6, tick_steps=8, precision="mixed")
    st = srv.run(jobs)
    assert st["overflow_recoveries"] == 0 and not st["failed"]
    assert st["peak_slots"] == 6 and st["batch_rebuilds"] >= 4
    for j in jobs:
        assert j.done == j.nsteps and j.result is not None
        s = j.make()
        ctx = nb.NBContext(s, precision="mixed", sharing=j.sharing)
        md = nb.NBMD(ctx, s, j.dt_ps, nres_intra=j.nres_intra)
        md.set_state(s["pos"], j.vel, s["hmat"])
        md.init()
        md.step(j.nsteps, graph=True)
        pos, vel = md.get_state()
        md.close()
        ctx.close()
        assert np.array_equal(j.result[0], pos) and np.array_equal(j.result[1], vel), j.id
