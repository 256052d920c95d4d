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
