"""Pin the clean-room oracle to the compiled cm3d reference (the golden
fixtures written by tests/golden/make_fixtures.py) and to the reference's own
printed golden outputs.  CPU only."""
import numpy as np
import pytest

from conftest import LIST_KEYS, load_fixture
from make_fixtures import order_digest, pair_digest
from oracle import oracle as O


def _lists(d):
    return {k: d[k] for _, a, b in LIST_KEYS for k in (a, b) if k in d}


def test_lists_match_reference_in_order(fixture):
    name, d = fixture
    L = O.build_lists(d, d["pos"], d["hmat"], use_cells=True)
    for t, a, b in LIST_KEYS:
        assert len(L[a]) == int(d[f"npr_{t}"]), (name, t)
        assert order_digest(L[a], L[b]) == str(d[f"digest_ord_{t}"]), (name, t)
        if a in d:
            assert np.array_equal(L[a], d[a]) and np.array_equal(L[b], d[b])


def test_cell_lists_equal_bruteforce_pair_set(fixture):
    """ngbr_list over link cells vs the O(N^2) walk: same (min,max) pair set."""
    name, d = fixture
    L = O.build_lists(d, d["pos"], d["hmat"], use_cells=False)
    for t, a, b in LIST_KEYS:
        assert pair_digest(L[a], L[b]) == str(d[f"digest_set_{t}"]), (name, t)


def test_forces_and_energies_bitwise(fixture):
    name, d = fixture
    L = O.build_lists(d, d["pos"], d["hmat"], use_cells=True)
    fs, fl, e = O.force(d, L, d["pos"], d["hmat"], 1)
    assert np.array_equal(fl, d["fxl"]), name
    assert np.array_equal(fs, d["fxs"]), name
    assert e["spotl"] == d["e_vinter"][0]
    assert e["spothl"] == d["e_vhinter"][0]
    assert e["sprsl"] == d["e_winter"][0]
    assert np.array_equal(e["sprslt"], d["e_wintertensor"])
    assert e["spote"] == d["e_velec"][0]
    assert e["sprse"] == d["e_welec"][0]
    assert np.array_equal(e["sprset"], d["e_welectensor"])
    assert e["vk_ewald"] == d["vk_ewald"][0]
    assert e["vp_ewald"] == d["vp_ewald"][0]
    assert np.array_equal(e["vp_ewald_tensor"], d["vp_ewald_tensor"])
    assert e["pot_ecorr"] == d["pot_ecorr"][0]
    assert np.array_equal(e["prs_ecorr_tensor"], d["prs_ecorr_tensor"])


def test_two_atom_closed_form():
    """test/lj/lj.out_nve2.t: two LJ atoms 2 sigma apart, <inter> = -0.061523 K
    = 4 (2^-12 - 2^-6) within the table's interpolation error."""
    d = load_fixture("ref_lj2")
    L = O.build_lists(d, d["pos"], d["hmat"])
    _, fl, e = O.force(d, L, d["pos"], d["hmat"], 1)
    exact = 4 * (2.0 ** -12 - 2.0 ** -6)
    assert abs(e["spotl"] - exact) < 1e-6
    assert round(e["spotl"], 6) == -0.061523
    # F = -dV/dr along x at r = 2: 4 (12 r^-13 - 6 r^-7)
    fexact = 4 * (12 * 2.0 ** -13 - 6 * 2.0 ** -7)
    assert abs(fl[0] - (-fexact)) < 1e-5 and fl[0] == -fl[1]


def test_lj256_pair_counts_match_printed_golden():
    """test/lj/lj.out_nve.t prints nprs=2334, nprt=7851 for the 256-atom deck."""
    d = load_fixture("ref_lj256")
    L = O.build_lists(d, d["pos"], d["hmat"])
    assert len(L["inps"]) == 2334 and len(L["inpt"]) == 7851


def test_abutane_setup_matches_printed_golden():
    """The abutane deck sets up 7073 k-vectors and 64 ecorr pairs (cm3d log)."""
    d = load_fixture("ref_abutane")
    ka, kb, kc, kup = O.kvectors(15)
    assert len(ka) == 7073 == int(d["ktot"][0])
    assert np.array_equal(ka, d["ka"]) and np.array_equal(kc, d["kc"])
    assert np.array_equal(kup[:-1], d["kup"][:-1])
    assert int(d["necorr"][0]) == 64


# ---------------------------------------------------------------------------
# set-up constants and the bonded energies, against compiled cm3d

def _synthetic(name):
    from cm3d_amd import systems as S
    return {"ar864": S.lj_fcc, "water512": lambda: S.water(nside=8),
            "mix600": lambda: S.mixture(600)}[name]()


@pytest.mark.parametrize("name", ["water512", "mix600", "ref_abutane", "water512_nvt"])
def test_ewald_constants_bitwise(name):
    """orc_ewald_consts gives ewald_setup's self energy and neutralising
    background (md_ewald.c:59-79) bit for bit.  mix600 and abutane carry a
    rounding-size net charge, so their background (~1e-22, ~1e-25) pins the
    -pi (sum q)^2 / (2 alpha^2) form as well."""
    d = load_fixture(name)
    self_, bkg = O.ewald_consts(d["qch"], float(d["kappa"][0]))
    assert self_ == d["ewald_self"][0] and self_ != 0.
    assert bkg == d["ewald_bkg"][0]


@pytest.mark.parametrize("name", ["ar864", "water512", "mix600"])
def test_lrc_totals_bitwise(name):
    """orc_lrc_total over the per-interaction tails of orc_table_pair equals
    simparms->vlrc / wlrc as search_ter_base summed them
    (md_searchbase.c:212-235), bit for bit (one, one and three LJ
    interactions; water's O-H and H-H pairs are null)."""
    d = load_fixture(name)
    s = _synthetic(name)
    nt = int(d["ntable"][0])
    vl, wl = [], []
    for prm in s["pair_params"]:
        t = O.table_pair(O.POT_LJ, prm[:2], prm[2], prm[3], prm[4], float(d["rheal"][0]),
                         float(d["skin"][0]), nt)
        vl.append(t["vlrc"])
        wl.append(t["wlrc"])
    v, w = O.lrc_total(d["itype"], int(d["ntypes"][0]), d["map"], np.array(vl), np.array(wl))
    assert v == d["vlrc"][0] and w == d["wlrc"][0]
    assert v < 0. < w


@pytest.mark.parametrize("name", ["water512", "water512_nvt"])
def test_bonded_energies_bitwise(name):
    """orc_bonded_force's energies are getvbond / getvbend's (V = k/2 dx^2,
    md_fbond.c:131-200, md_fbend.c:262-340): vbond and vbend at the snapshot,
    and potra along cm3d's own RESPA trajectory, where the bonds are
    stretched by ~10^4-10^5 K, bit for bit.  A k-for-k/2 slip doubles them."""
    d = load_fixture(name)
    s = _synthetic("water512")
    b = dict(ib=s["bond_i"], jb=s["bond_j"], bk=s["bond_k"], beq=s["bond_r0"],
             ia=s["bend_i"], ja=s["bend_j"], ka=s["bend_k"], ak=s["bend_fk"],
             aeq=s["bend_th0"])
    n = int(d["natoms"][0])
    _, _, (vb, va) = O.bonded_force(n, b, d["pos"], parts=True)
    assert vb == d["e_vbond"][0] and va == d["e_vbend"][0]
    for st in range(int(d["traj_nsteps"][0]) + 1):
        v, _, (vb, va) = O.bonded_force(n, b, d[f"traj_pos{st}"], parts=True)
        assert v == d[f"traj_potra{st}"][0], st
        if st > 0:
            assert vb > 1e3 and va > 1e2, (st, vb, va)


def test_c1_fixture_is_thermalised():
    """The 864-atom argon fixture is taken after 200 of cm3d's own steps
    from the FCC deck at 120 K: its forces are O(10^2) K/A, so force
    comparisons on it test the pair kernel rather than lattice rounding."""
    d = load_fixture("ar864")
    rms = float(np.sqrt(np.mean(d["fxl"] ** 2)))
    assert 50. < rms < 5e3, rms
    assert int(d["npre"][0]) == 200
    assert not np.array_equal(d["pos"], d["pos_deck"])
