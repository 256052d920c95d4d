"""GPU parity: the gfx950 path against the compiled cm3d reference's outputs
(golden fixtures) and the oracle, through the C ABI.  Tolerances:
  lists       bit-exact as sorted (min,max) pair sets and counts
  FP64 path   per-atom forces within 1e-11 relative RMS of cm3d's (only the
              summation order differs), energies / virials within 1e-11
              relative
"""
import numpy as np
import pytest

from conftest import FIXTURES, LIST_KEYS, load_fixture, rel_rms
from make_fixtures import pair_digest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11


def _ctx(d, precision="fp64"):
    from cm3d_amd import nb
    return nb.NBContext(d, precision=precision)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("name", FIXTURES)
def test_lists_bit_exact(name):
    d = load_fixture(name)
    print(f"[lists] {name}", flush=True)
    c = _ctx(d)
    c.build_lists(d["pos"], d["hmat"])
    cnt = c.list_counts()
    assert cnt["nprt"] == int(d["npr_t"]) and cnt["nprs"] == int(d["npr_s"])
    assert cnt["nprte"] == int(d["npr_te"]) and cnt["nprse"] == int(d["npr_se"])
    for t, a, b in LIST_KEYS:
        i, j = c.get_pairs(t)
        assert pair_digest(i, j) == str(d[f"digest_set_{t}"]), (name, t)
    c.close()


@pytest.mark.parametrize("name", FIXTURES)
def test_fp64_forces_energies(name):
    d = load_fixture(name)
    print(f"[fp64] {name}", flush=True)
    c = _ctx(d)
    c.build_lists(d["pos"], d["hmat"])
    fs, fl, e = c.force(d["pos"], d["hmat"], iflg=1)
    assert rel_rms(fl, d["fxl"]) < FP64_TOL, rel_rms(fl, d["fxl"])
    assert rel_rms(fs, d["fxs"]) < FP64_TOL, rel_rms(fs, d["fxs"])
    for mine, ref in (("spotl", "e_vinter"), ("spothl", "e_vhinter"), ("sprsl", "e_winter"),
                      ("spote", "e_velec"), ("sprse", "e_welec")):
        assert _rel(e[mine], d[ref][0]) < FP64_TOL or abs(e[mine] - d[ref][0]) < 1e-9, (mine, e[mine], d[ref][0])
    assert np.allclose(e["sprslt"], d["e_wintertensor"], rtol=FP64_TOL, atol=1e-9)
    assert np.allclose(e["sprset"], d["e_welectensor"], rtol=FP64_TOL, atol=1e-9)
    assert _rel(e["vk_ewald"], d["vk_ewald"][0]) < FP64_TOL or abs(e["vk_ewald"]) < 1e-12
    assert np.allclose(e["vp_ewald_tensor"], d["vp_ewald_tensor"], rtol=1e-10, atol=1e-8)
    assert np.allclose(e["prs_ecorr_tensor"], d["prs_ecorr_tensor"], rtol=1e-10, atol=1e-8)
    assert _rel(e["pot_ecorr"], d["pot_ecorr"][0]) < FP64_TOL or d["pot_ecorr"][0] == 0
    c.close()


@pytest.mark.parametrize("name", ["ar864", "water512", "mix600", "ref_abutane", "water512_tri",
                                  "mix600_tri"])
def test_fp64_short_level_matches_oracle(name):
    """force(...,0): the RESPA short level over the s lists."""
    from oracle import oracle as O
    d = load_fixture(name)
    L = O.build_lists(d, d["pos"], d["hmat"])
    ofs, _, _ = O.force(d, L, d["pos"], d["hmat"], 0)
    c = _ctx(d)
    c.build_lists(d["pos"], d["hmat"])
    fs, _, _ = c.force(d["pos"], d["hmat"], iflg=0)
    assert rel_rms(fs, ofs) < FP64_TOL
    c.close()


def _irregular_klist(d):
    """cm3d's kmax 7 half sphere cut into rows that start away from kc = 0
    (kc >= 3 only, kc <= -2 only), a row of kc 2..20 and the rest whole:
    the reciprocal passes then start a row's phase at m = |kc| > 1 (sine /
    cosine) and split long rows into lanes that start at m = 18.  The
    oracle evaluates every k directly (kup = 1)."""
    from cm3d_amd import nb
    ka, kb, kc = nb.kvectors(7)
    keep = []
    for a, b, c in zip(ka, kb, kc):
        r = (int(a) + int(b)) % 3
        if (a, b) == (0, 1):
            continue
        if (r == 0 and c >= 3) or (r == 1 and c <= -2) or r == 2:
            keep.append((a, b, c))
    keep += [(0, 1, c) for c in range(2, 21)]
    k = np.array(sorted(keep, key=lambda t: (t[0], t[1], t[2])), np.int32)
    e = dict(d)
    e.update(ktot=len(k), ka=k[:, 0].copy(), kb=k[:, 1].copy(), kc=k[:, 2].copy(),
             kup=np.ones(len(k) + 1, np.int32))
    return e


@pytest.mark.parametrize("mfma", [None, "1"])
@pytest.mark.parametrize("name", ["water512", "water512_tri"])
def test_fp64_irregular_k_rows_match_oracle(name, mfma, monkeypatch):
    """mfma "1": the structure factor on the matrix cores, whose row tiles
    are runs of consecutive kb (broken here by the missing (0, 1) row)."""
    from oracle import oracle as O
    if mfma:
        monkeypatch.setenv("CM3D_NB_SFAC_MFMA", mfma)
    d = _irregular_klist(load_fixture(name))
    print(f"[irregular k] {name} mfma={mfma}: {int(d['ktot'])} k", flush=True)
    L = O.build_lists(d, d["pos"], d["hmat"])
    _, ofl, oe = O.force(d, L, d["pos"], d["hmat"], 1)
    c = _ctx(d)
    c.build_lists(d["pos"], d["hmat"])
    _, fl, e = c.force(d["pos"], d["hmat"], iflg=1)
    assert rel_rms(fl, ofl) < FP64_TOL, rel_rms(fl, ofl)
    assert _rel(e["vk_ewald"], oe["vk_ewald"]) < 1e-11
    assert np.allclose(e["vp_ewald_tensor"], oe["vp_ewald_tensor"], rtol=1e-10, atol=1e-8)
    c.close()


def test_fp64_reproducible_run_twice():
    d = load_fixture("water512")
    c = _ctx(d)
    c.build_lists(d["pos"], d["hmat"])
    a = c.force(d["pos"], d["hmat"], 1)
    c.build_lists(d["pos"], d["hmat"])
    b = c.force(d["pos"], d["hmat"], 1)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])
    assert a[2]["spote"] == b[2]["spote"] and a[2]["vk_ewald"] == b[2]["vk_ewald"]
    c.close()


def test_need_rebuild_criterion():
    from oracle import oracle as O
    d = load_fixture("water512")
    c = _ctx(d)
    c.build_lists(d["pos"], d["hmat"])
    assert not c.need_rebuild(d["pos"])
    n = int(d["natoms"][0])
    skin = float(d["skin"][0])
    p = d["pos"].copy()
    p[7] += 0.49 * skin       # just inside skin/2
    assert c.need_rebuild(p) == O.need_rebuild(n, skin, p, d["pos"]) == False
    p[n + 11] += 0.51 * skin  # one atom beyond skin/2
    assert c.need_rebuild(p) == O.need_rebuild(n, skin, p, d["pos"]) == True
    c.close()


MIXED_FORCE_TOL = 1e-4      # force RMS error bar vs the FP64 result
MIXED_ENERGY_TOL = 1e-5     # energy / virial relative error bar vs the FP64 result


@pytest.mark.parametrize("name", ["ar864", "water512", "mix600", "ref_lj256", "water512_tri",
                                  "mix600_tri"])
def test_mixed_forces_within_tolerance(name):
    """FP32 pair math + 64-bit fixed-point forces vs cm3d's FP64 forces."""
    d = load_fixture(name)
    print(f"[mixed] {name}", flush=True)
    c = _ctx(d, "mixed")
    c.build_lists(d["pos"], d["hmat"])
    fs, fl, e = c.force(d["pos"], d["hmat"], iflg=1)
    err = rel_rms(fl, d["fxl"])
    print(f"  mixed force rel RMS error {err:.3e}")
    assert err < MIXED_FORCE_TOL
    # a mixed context's pair energies and virials come from the FP64 energy
    # pass (the FP32 kernel computes forces only): cm3d's values to the FP64
    # bar, well inside MIXED_ENERGY_TOL
    for mine, ref in (("spotl", "e_vinter"), ("spote", "e_velec"), ("sprsl", "e_winter"),
                      ("sprse", "e_welec")):
        err = abs(e[mine] - d[ref][0]) / max(abs(d[ref][0]), 1.0)
        print(f"  mixed {mine} rel error {err:.2e}")
        assert err <= min(MIXED_ENERGY_TOL, 1e-10), (mine, e[mine], d[ref][0])
    # Ewald and ecorr stay FP64
    assert _rel(e["vk_ewald"], d["vk_ewald"][0]) < 1e-10 or abs(e["vk_ewald"]) < 1e-12
    c.close()


def test_mixed_bitwise_reproducible():
    d = load_fixture("water512")
    c = _ctx(d, "mixed")
    c.build_lists(d["pos"], d["hmat"])
    a = c.force(d["pos"], d["hmat"], 1)
    b = c.force(d["pos"], d["hmat"], 1)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["ar864", "water512", "mix600", "ref_abutane", "water512_tri"])
def test_energy_only_equals_force_energies(name, prec):
    """nb_energy_only returns what cm3d's getvireal(...,0) sums: the vdW
    fields (vinter_npol) equal force(...,1)'s bit for bit and cm3d's within
    the FP64 bar; the coulomb fields are cm3d's vinter_e values -- every te
    pair inside the cutoff clamped to the table origin -- equal to the
    fixture's dumped e0_velec / e0_welec / tensor within the FP64 bar, and
    the force path's coulomb values are in the *_phys fields, bit for bit
    force(...,1)'s; Ewald and ecorr bit for bit."""
    d = load_fixture(name)
    print(f"[eonly] {name} {prec}", flush=True)
    c = _ctx(d, prec)
    c.build_lists(d["pos"], d["hmat"])
    _, _, ef = c.force(d["pos"], d["hmat"], iflg=1)
    eo = c.energy_only(d["pos"], d["hmat"])
    coul = ("spote", "spothe", "sprse", "sprset")
    for k, v in ef.items():
        if k in coul:
            continue
        if k.endswith("_phys"):
            assert np.array_equal(np.asarray(eo[k]), np.asarray(ef[k[:-5]])), (k, eo[k])
        else:
            assert np.array_equal(np.asarray(eo[k]), np.asarray(v)), (k, eo[k], v)
    for mine, ref in (("spotl", "e0_vinter"), ("spothl", "e0_vhinter"), ("spote", "e0_velec"),
                      ("sprse", "e0_welec")):
        assert _rel(eo[mine], d[ref][0]) < FP64_TOL or abs(eo[mine] - d[ref][0]) < 1e-9, \
            (mine, eo[mine], d[ref][0])
    assert np.allclose(eo["sprset"], d["e0_welectensor"], rtol=FP64_TOL, atol=1e-9)
    if float(np.abs(d["qch"]).max()) > 0:
        # cm3d's energy-only coulomb value is not the physical one
        assert abs(eo["spote"] - eo["spote_phys"]) > 1e-6 * abs(eo["spote_phys"])
    c.close()


def test_sharing_changes_only_the_work_split():
    """A context expecting to share the GPU (sharing=64) splits the
    reciprocal sum into fewer, larger pieces; the forces and energies equal
    the alone context's to rounding and cm3d's to the FP64 bar."""
    d = load_fixture("water512")
    from cm3d_amd import nb
    out = []
    for share in (1, 64):
        c = nb.NBContext(d, precision="fp64", sharing=share)
        c.build_lists(d["pos"], d["hmat"])
        out.append(c.force(d["pos"], d["hmat"], iflg=1))
        c.close()
    assert rel_rms(out[1][1], out[0][1]) < 1e-13
    assert rel_rms(out[1][1], d["fxl"]) < FP64_TOL
    assert _rel(out[1][2]["vk_ewald"], out[0][2]["vk_ewald"]) < 1e-13


@pytest.mark.parametrize("name", ["water512", "mix600", "ref_abutane"])
def test_partner_types_gathered_when_entries_lack_them(name, monkeypatch):
    """List entries carry the partner's type above its index when the type
    count fits the spare bits; otherwise the forces-only kernels of the MD
    step gather the type.  Forcing the second layout on small systems
    (CM3D_NB_ENTRY_TYPE=0) gives the same lists and bitwise the same
    forces-only forces (nb_force_device without energies) in both
    precisions."""
    from conftest import DevBuf, hip_runtime
    from cm3d_amd import nb
    d = load_fixture(name)
    h = hip_runtime()
    pos = np.ascontiguousarray(d["pos"], dtype=np.float64)
    for prec in ("fp64", "mixed"):
        got = []
        for flag in ("1", "0"):
            monkeypatch.setenv("CM3D_NB_ENTRY_TYPE", flag)
            c = nb.NBContext(d, precision=prec)
            c.build_lists(d["pos"], d["hmat"])
            pairs = [c.get_pairs(t) for t, _, _ in LIST_KEYS]
            d_pos, fl = DevBuf(h, pos), DevBuf(h, np.full_like(pos, np.nan))
            c.force_device(1, d_pos.p.value, d["hmat"], 0, fl.p.value)
            got.append((pairs, fl.get()))
            d_pos.free()
            fl.free()
            c.close()
        monkeypatch.delenv("CM3D_NB_ENTRY_TYPE")
        for (i0, j0), (i1, j1) in zip(got[0][0], got[1][0]):
            assert np.array_equal(i0, i1) and np.array_equal(j0, j1), (name, prec)
        assert np.isfinite(got[0][1]).all()
        assert np.array_equal(got[0][1], got[1][1]), (name, prec)


@pytest.mark.parametrize("name", ["mix600", "mix600_tri"])
def test_multirow_structure_factor_lanes(name):
    """With every k row short (kmax 7: <= 8 values of |kc|) a context that
    shares the GPU gives each structure-factor lane two rows of consecutive
    kb, the second row's phase stepped from the first's by e^{i theta_b};
    the reciprocal energy, virial and forces stay within the FP64 bar of
    cm3d's, and within rounding of the one-row lanes of a lone context."""
    d = load_fixture(name)
    from cm3d_amd import nb
    out = []
    for share in (1, 16):
        c = nb.NBContext(d, precision="fp64", sharing=share)
        c.build_lists(d["pos"], d["hmat"])
        out.append(c.force(d["pos"], d["hmat"], iflg=1))
        c.close()
    fs, fl, e = out[1]
    assert rel_rms(fl, d["fxl"]) < FP64_TOL, rel_rms(fl, d["fxl"])
    assert _rel(e["vk_ewald"], d["vk_ewald"][0]) < FP64_TOL
    assert np.allclose(e["vp_ewald_tensor"], d["vp_ewald_tensor"], rtol=1e-10, atol=1e-8)
    assert rel_rms(fl, out[0][1]) < 1e-13
    assert _rel(e["vk_ewald"], out[0][2]["vk_ewald"]) < 1e-13


@pytest.mark.parametrize("name", ["water512", "water512_tri", "mix600", "mix600_tri",
                                  "ref_abutane"])
def test_matrix_core_structure_factor(name, monkeypatch):
    """The reciprocal sum as FP64 matrix-core products (k_ewsfacm and
    k_ewforcem, used by default for large k spheres and charge counts)
    forced on the fixtures (CM3D_NB_SFAC_MFMA=1, which the force pass
    follows): the reciprocal energy, virial and forces stay within the FP64
    bar of cm3d's, and within rounding of the VALU kernels'
    (CM3D_NB_SFAC_MFMA=0), in both precisions."""
    d = load_fixture(name)
    from cm3d_amd import nb
    for prec in ("fp64", "mixed"):
        out = {}
        for flag in ("1", "0"):
            monkeypatch.setenv("CM3D_NB_SFAC_MFMA", flag)
            c = nb.NBContext(d, precision=prec)
            c.build_lists(d["pos"], d["hmat"])
            out[flag] = c.force(d["pos"], d["hmat"], iflg=1)
            c.close()
        monkeypatch.delenv("CM3D_NB_SFAC_MFMA")
        fs, fl, e = out["1"]
        assert np.isfinite(fl).all()
        assert rel_rms(fl, out["0"][1]) < 1e-13, (name, prec, rel_rms(fl, out["0"][1]))
        assert _rel(e["vk_ewald"], out["0"][2]["vk_ewald"]) < 1e-13
        assert np.allclose(e["vp_ewald_tensor"], out["0"][2]["vp_ewald_tensor"], rtol=1e-12,
                           atol=1e-9)
        if prec == "fp64":
            assert rel_rms(fl, d["fxl"]) < FP64_TOL, rel_rms(fl, d["fxl"])
            assert _rel(e["vk_ewald"], d["vk_ewald"][0]) < FP64_TOL


@pytest.mark.parametrize("sfac,ewf", [("1", "0"), ("0", "1")])
@pytest.mark.parametrize("name", ["water512", "water512_tri", "ref_abutane", "irregular"])
def test_matrix_core_reciprocal_kernels_each(name, sfac, ewf, monkeypatch):
    """Each matrix-core reciprocal kernel on its own (CM3D_NB_SFAC_MFMA,
    CM3D_NB_EWF_MFMA) against both VALU kernels: forces and reciprocal
    energy to rounding, FP64.  "irregular": water512 with the cut k rows
    (rows from |kc| = 3 up, so that force tiles start above m = 1)."""
    d = _irregular_klist(load_fixture("water512")) if name == "irregular" else load_fixture(name)
    from cm3d_amd import nb
    out = {}
    for flags in ((sfac, ewf), ("0", "0")):
        monkeypatch.setenv("CM3D_NB_SFAC_MFMA", flags[0])
        monkeypatch.setenv("CM3D_NB_EWF_MFMA", flags[1])
        c = nb.NBContext(d, precision="fp64")
        c.build_lists(d["pos"], d["hmat"])
        out[flags] = c.force(d["pos"], d["hmat"], iflg=1)
        c.close()
    a, b = out[(sfac, ewf)], out[("0", "0")]
    assert np.isfinite(a[1]).all()
    assert rel_rms(a[1], b[1]) < 1e-13, (name, sfac, ewf, rel_rms(a[1], b[1]))
    assert _rel(a[2]["vk_ewald"], b[2]["vk_ewald"]) < 1e-13


def test_matrix_core_structure_factor_full_size(monkeypatch):
    """The 98,304-atom water of kmax 32 (68,532 k vectors), where both
    matrix-core reciprocal kernels are the default (nb_recip_info says so):
    the same forces and reciprocal energy as the VALU kernels to
    rounding."""
    from cm3d_amd import nb, systems as S
    s = S.water(nside=32, rc=12.0, alpha=0.30, kmax=32)
    out = {}
    for flag in (None, "0"):
        if flag:
            monkeypatch.setenv("CM3D_NB_SFAC_MFMA", flag)
        c = nb.NBContext(s, precision="mixed")
        info = c.recip_info()
        if flag is None:  # the default for this size: both passes on the matrix cores
            assert info["sfac_mfma"] and info["force_mfma"], info
            assert info["sfac_mfma_insts"] > 0 and info["force_mfma_insts"] > 0, info
        else:
            assert not info["sfac_mfma"] and not info["force_mfma"], info
            assert info["sfac_mfma_insts"] == 0 and info["force_mfma_insts"] == 0, info
        c.build_lists(s["pos"], s["hmat"])
        out[flag] = c.force(s["pos"], s["hmat"], iflg=1)
        c.close()
    fl, e = out[None][1], out[None][2]
    assert np.isfinite(fl).all()
    assert rel_rms(fl, out["0"][1]) < 1e-13, rel_rms(fl, out["0"][1])
    assert _rel(e["vk_ewald"], out["0"][2]["vk_ewald"]) < 1e-13


@pytest.mark.parametrize("sharing", [1, 16])
@pytest.mark.parametrize("natoms", [1001, 1002, 1003])
def test_mixed_short_lists_partial_groups(natoms, sharing):
    """LJ + charge mixtures have ~100-entry lists, which the mixed kernel
    evaluates four atoms to a wavefront (16 lanes an atom).  With N not a
    multiple of 4 the last wavefront has idle groups; with sharing >= 4 a
    workgroup takes 64 atoms instead of 16.  Both force kinds (fl at iflg
    1, fs at iflg 0) stay within the mixed-precision bar of the FP64
    forces."""
    from cm3d_amd import nb, systems as S
    s = S.mixture(natoms, seed=300 + natoms, rc=9.0, kmax=5)
    out = {}
    for prec in ("fp64", "mixed"):
        c = nb.NBContext(s, precision=prec, sharing=sharing)
        c.build_lists(s["pos"], s["hmat"])
        _, fl, _ = c.force(s["pos"], s["hmat"], iflg=1)
        fs, _, _ = c.force(s["pos"], s["hmat"], iflg=0)
        out[prec] = (fl, fs)
        c.close()
    for k in range(2):
        err = rel_rms(out["mixed"][k], out["fp64"][k])
        assert err < MIXED_FORCE_TOL, (natoms, sharing, k, err)
