"""nb_force_device_noalloc (scratch in the library's static per-GPU area, no
device allocation) against nb_force_device, through the C ABI: forces and
every energy field bit for bit, on two calls in a row, the second on a
smaller system, in both precisions and at both RESPA levels, with either
structure-factor kernel; a system whose
scratch exceeds the static area is refused with an error on the host."""
import numpy as np
import pytest

from conftest import DevBuf as _Dev, hip_runtime as _hip, load_fixture

pytestmark = pytest.mark.gpu

def _both(d, prec, iflg):
    from cm3d_amd import nb
    h = _hip()
    ctx = nb.NBContext(d, precision=prec)
    ctx.build_lists(d["pos"], d["hmat"])
    pos = np.ascontiguousarray(d["pos"], dtype=np.float64)
    d_pos = _Dev(h, pos)
    out = []
    for fn in (ctx.force_device, ctx.force_device_noalloc):
        fs = _Dev(h, np.full_like(pos, np.nan))
        fl = _Dev(h, np.full_like(pos, np.nan))
        e = fn(iflg, d_pos.p.value, d["hmat"], fs.p.value, fl.p.value if iflg else 0,
               energy=True)
        out.append((fs.get(), fl.get(), e))
        fs.free()
        fl.free()
    d_pos.free()
    ctx.close()
    return out


@pytest.mark.parametrize("mfma", [None, "1"])
@pytest.mark.parametrize("iflg", [1, 0])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_equals_force_device(prec, iflg, mfma, monkeypatch):
    # mfma "1": the structure factor on the matrix cores, its phase tables in
    # the static area too
    if mfma:
        monkeypatch.setenv("CM3D_NB_SFAC_MFMA", mfma)
    for name in ("water512", "mix600"):  # the second call on a smaller system
        print(f"[noalloc] {name} {prec} iflg={iflg} mfma={mfma}", flush=True)
        d = load_fixture(name)
        (fs0, fl0, e0), (fs1, fl1, e1) = _both(d, prec, iflg)
        assert np.array_equal(fs0, fs1), name
        if iflg:
            assert np.array_equal(fl0, fl1), name
            assert np.isfinite(fl1).all()
        for k, v in e0.items():
            assert np.array_equal(np.asarray(e1[k]), np.asarray(v)), (name, k)


def test_noalloc_refuses_oversized_scratch():
    from cm3d_amd import nb, systems as S
    h = _hip()
    s = S.water(nside=32, rc=12.0, alpha=0.30, kmax=32)  # 98,304 charges, 68,532 k
    ctx = nb.NBContext(s, precision="fp64")
    pos = np.ascontiguousarray(s["pos"])
    d_pos, fs, fl = _Dev(h, pos), _Dev(h, pos), _Dev(h, pos)
    with pytest.raises(nb.NBError, match="bytes of scratch"):  # refused on the host
        ctx.force_device_noalloc(1, d_pos.p.value, s["hmat"], fs.p.value, fl.p.value)
    for b in (d_pos, fs, fl):
        b.free()
    ctx.close()
