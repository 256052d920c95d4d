"""GPU parity at full size on the 98,304-atom water (32^3 molecules, box
99.3 A, rc 12 A, skin 1 A, Ewald alpha 0.30 with kmax 32: 68,532 k-vectors,
all 98,304 atoms charged):

- the Verlet lists are bit-exact against the oracle's (sorted pair sets);
- with the reciprocal sum off, FP64 forces match the oracle within 1e-11
  relative RMS and energies / virials within 1e-10 (every real-space pair,
  ecorr, self and background terms at full N);
- the reciprocal sum over the first two k planes (ka <= 1: every charge
  against ~6k k-vectors) matches the oracle within 1e-11;
- the full 68,532-vector sum is pinned by physics: the total electrostatic
  energy and forces do not depend on the Ewald splitting (alpha 0.30 / kmax
  32 against alpha 0.27 / kmax 30, both converged to ~1e-5): forces within
  1e-4 relative RMS, energies within 3e-5 of the size of the Ewald pieces;
- the mixed form (FP32 pair math, 64-bit fixed-point forces) is within the
  1e-4 relative RMS force bar of the FP64 result, and a short device-resident
  MD run in each precision conserves cm3d's Ham."""
import numpy as np
import pytest

from conftest import rel_rms
from make_fixtures import pair_digest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    from cm3d_amd import systems as S
    return S.water(nside=32, rc=12.0, alpha=0.30, kmax=32)


@pytest.fixture(scope="module")
def c3_lists(c3):
    from oracle import oracle as O
    return O.build_lists(c3, c3["pos"], c3["hmat"])


def _without_kspace(s, keep=0):
    """The same system with only the first `keep` k-vectors."""
    t = dict(s)
    t["ka"], t["kb"], t["kc"] = (np.asarray(s[k])[:keep].copy() for k in ("ka", "kb", "kc"))
    kup = np.asarray(s["kup"])[:keep + 1].copy()
    if keep:
        kup[keep] = 1
    t["kup"] = kup
    t["ktot"] = keep
    return t


def test_c3_lists_bit_exact(c3, c3_lists):
    from cm3d_amd import nb
    print("[c3] lists", flush=True)
    L = c3_lists
    c = nb.NBContext(c3, precision="fp64")
    c.build_lists(c3["pos"], c3["hmat"])
    cnt = c.list_counts()
    for t, (a, b) in (("t", ("inpt", "jnpt")), ("s", ("inps", "jnps")),
                      ("te", ("inpte", "jnpte")), ("se", ("inpse", "jnpse"))):
        i, j = c.get_pairs(t)
        assert len(i) == len(L[a]) == cnt["npr" + t], (t, len(i), len(L[a]))
        assert pair_digest(i, j) == pair_digest(L[a], L[b]), t
    print(f"  counts {cnt}", flush=True)
    c.close()


def test_c3_real_space_matches_oracle(c3, c3_lists):
    from cm3d_amd import nb
    from oracle import oracle as O
    print("[c3] real space", flush=True)
    s = _without_kspace(c3)
    L = c3_lists
    ofs, ofl, oe = O.force(s, L, s["pos"], s["hmat"], 1)
    c = nb.NBContext(s, precision="fp64")
    c.build_lists(s["pos"], s["hmat"])
    fs, fl, e = c.force(s["pos"], s["hmat"], iflg=1)
    assert rel_rms(fl, ofl) < 1e-11, rel_rms(fl, ofl)
    assert rel_rms(fs, ofs) < 1e-11, rel_rms(fs, ofs)
    # sums of 5M (vdW) and 45M (coulomb) pair terms in different orders:
    # 1e-10 relative (the virials cancel large terms of both signs)
    for k in ("spotl", "spothl", "sprsl", "spote", "sprse", "vk_ewald", "pot_ecorr", "prs_ecorr"):
        assert abs(e[k] - oe[k]) <= 1e-10 * max(abs(oe[k]), 1.0), (k, e[k], oe[k])
    c.close()


def test_c3_kspace_planes_match_oracle(c3, c3_lists):
    from cm3d_amd import nb
    from oracle import oracle as O
    keep = int(np.count_nonzero(np.asarray(c3["ka"]) <= 1))
    print(f"[c3] k-space, first {keep} k-vectors", flush=True)
    s = _without_kspace(c3, keep)
    L = c3_lists
    _, ofl, oe = O.force(s, L, s["pos"], s["hmat"], 1)
    c = nb.NBContext(s, precision="fp64")
    c.build_lists(s["pos"], s["hmat"])
    _, fl, e = c.force(s["pos"], s["hmat"], iflg=1)
    assert rel_rms(fl, ofl) < 1e-11, rel_rms(fl, ofl)
    for k in ("vk_ewald", "vp_ewald"):
        assert abs(e[k] - oe[k]) <= 1e-11 * abs(oe[k]), (k, e[k], oe[k])
    assert np.allclose(e["vp_ewald_tensor"], oe["vp_ewald_tensor"], rtol=1e-10, atol=1e-6)
    c.close()


def _coulomb(s):
    """Total electrostatic energy (real space + reciprocal + self + background
    + ecorr), the magnitude of its pieces, the vdW energy and the full force
    of one FP64 evaluation."""
    from cm3d_amd import nb
    c = nb.NBContext(s, precision="fp64")
    c.build_lists(s["pos"], s["hmat"])
    _, fl, e = c.force(s["pos"], s["hmat"], iflg=1)
    c.close()
    self_e = float(s["ewald_self"])
    pieces = (e["spote"], e["vk_ewald"] - self_e, self_e, e["pot_ecorr"])
    return sum(pieces), sum(abs(p) for p in pieces), e["spotl"], fl


def test_c3_full_kspace_independent_of_alpha(c3):
    """Ewald splitting: the total does not depend on alpha once both sums
    have converged.  Truncation estimates: real space erfc(alpha rc) =
    3.6e-7 / 5.9e-6 and reciprocal exp(-(pi kmax / alpha L)^2) = 1.1e-5 /
    4.3e-6 for (0.30, 32) / (0.27, 30)."""
    from cm3d_amd import systems as S
    print("[c3] alpha independence", flush=True)
    other = S.water(nside=32, rc=12.0, alpha=0.27, kmax=30)
    assert np.array_equal(other["pos"], c3["pos"])
    ea, sa, lja, fa = _coulomb(c3)
    eb, sb, ljb, fb = _coulomb(other)
    assert lja == ljb
    # the total is a small difference of large pieces (self ~ -alpha/sqrt(pi)
    # sum q^2, real space, reciprocal, ecorr): the truncation errors are
    # relative to the pieces
    rel_e = abs(ea - eb) / sa
    rel_f = rel_rms(fa, fb)
    print(f"  E_coul {ea:.3f} vs {eb:.3f} K (pieces {sa:.4e} K): rel {rel_e:.2e}; "
          f"force rel RMS {rel_f:.2e}", flush=True)
    assert rel_e < 3e-5
    assert rel_f < 1e-4


def test_c3_mixed_forces_within_tolerance(c3):
    from cm3d_amd import nb
    print("[c3] mixed vs fp64", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        c = nb.NBContext(c3, precision=prec)
        c.build_lists(c3["pos"], c3["hmat"])
        out[prec] = c.force(c3["pos"], c3["hmat"], iflg=1)
        c.close()
    err = rel_rms(out["mixed"][1], out["fp64"][1])
    print(f"  mixed force rel RMS {err:.3e}", flush=True)
    assert err < 1e-4
    assert out["mixed"][2]["vk_ewald"] == pytest.approx(out["fp64"][2]["vk_ewald"], rel=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c3_short_md_conserves_ham(c3, prec):
    """After 300 steps of velocity rescaling to 300 K (the lattice start is
    far from equilibrium), 200 NVE RESPA steps (4 x 0.5 fs inner, 2 fs
    outer) on the device: cm3d's Ham (poth + potra + zke) moves by less than
    1% of the kinetic energy, with lists rebuilt on the device as needed."""
    from cm3d_amd import nb, systems as S
    print(f"[c3] md {prec}", flush=True)
    ctx = nb.NBContext(c3, precision=prec)
    md = nb.NBMD(ctx, c3, 0.002, nres_intra=4)
    md.set_state(c3["pos"], c3["vel"], c3["hmat"])
    md.init()
    S.equilibrate([md], [c3], 300.0, 300, every=10)
    e0 = md.energies()
    md.step(200, graph=True)
    e1 = md.energies()
    nreb = md.rebuilds()
    drift = abs(e1["ham"] - e0["ham"]) / e0["zke"]
    print(f"  ham drift / zke {drift:.2e}, rebuilds {nreb}", flush=True)
    assert e1["overflow"] == 0
    assert drift < 1e-2
    md.close()
    ctx.close()
