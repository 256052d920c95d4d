"""ctypes binding of the clean-room C oracle (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline -- never by the shipped package.

A *system* is a dict of numpy arrays keyed like cm3d's INTER / COUL / COORDS
fields (the same keys oracle/ref_dump.c writes), e.g. ``natoms``, ``itype``,
``map``, ``vtab``, ``rmaxl_skin2``, ``qch``, ``coultab``, ``ka`` ...  Each
function names the cm3d routine it restates (see cm3d_oracle.c).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcSys(C.Structure):
    _fields_ = [
        ("natoms", C.c_int), ("ntypes", C.c_int), ("ninter", C.c_int),
        ("ntable", C.c_int), ("iperd", C.c_int), ("ncell", C.c_int),
        ("nlen", C.c_int),
        ("itype", _ip), ("map", _ip),
        ("vtab", _dp), ("dvtab", _dp), ("switchs", _dp),
        ("rmins_cut2", _dp), ("rmaxs_cut2", _dp), ("rminl_cut2", _dp),
        ("rmaxl_cut2", _dp), ("dx2tab", _dp), ("rmaxs_skin2", _dp),
        ("rmaxl_skin2", _dp), ("skin", C.c_double),
        ("nexclude", _ip), ("exclude", _ip),
        ("qch", _dp),
        ("coultab", _dp), ("dcoultab", _dp), ("switchse", _dp),
        ("dx2coul", C.c_double), ("rmins_coul2", C.c_double),
        ("rmaxs_coul2", C.c_double), ("rminl_coul2", C.c_double),
        ("rmaxl_coul2", C.c_double), ("rmaxs_skin2_e", C.c_double),
        ("rmaxl_skin2_e", C.c_double),
        ("cnexclude", _ip), ("cexclude", _ip),
        ("necorr", C.c_int), ("iecorr", _ip), ("jecorr", _ip),
        ("ktot", C.c_int), ("ka", _ip), ("kb", _ip), ("kc", _ip), ("kup", _ip),
        ("kappa", C.c_double), ("ewald_self", C.c_double),
        ("ewald_bkg", C.c_double),
    ]


class OrcLists(C.Structure):
    _fields_ = [("nprt", C.c_int), ("nprs", C.c_int), ("nprte", C.c_int),
                ("nprse", C.c_int)] + [
        (n, _ip) for n in ("inpt", "jnpt", "inps", "jnps", "inpte", "jnpte",
                           "inpse", "jnpse")]


class OrcEnergy(C.Structure):
    _fields_ = [("spotl", C.c_double), ("spothl", C.c_double),
                ("sprsl", C.c_double), ("sprslt", C.c_double * 9),
                ("spote", C.c_double), ("spothe", C.c_double),
                ("sprse", C.c_double), ("sprset", C.c_double * 9),
                ("vk_ewald", C.c_double), ("vp_ewald", C.c_double),
                ("vp_ewald_tensor", C.c_double * 9),
                ("pot_ecorr", C.c_double), ("prs_ecorr", C.c_double),
                ("prs_ecorr_tensor", C.c_double * 9), ("nclamp", C.c_int)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(v[:]) if hasattr(v, "__len__") else v
        return out


class OrcBonded(C.Structure):
    _fields_ = [("nbond", C.c_int), ("nbend", C.c_int),
                ("ib", _ip), ("jb", _ip), ("bk", _dp), ("beq", _dp),
                ("ia", _ip), ("ja", _ip), ("ka", _ip), ("ak", _dp),
                ("aeq", _dp)]


def build():
    """Compile liboracle.so (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_build_lists.argtypes = [C.POINTER(OrcSys), _dp, _dp, C.c_int,
                                      C.POINTER(OrcLists)]
        L.orc_free_lists.argtypes = [C.POINTER(OrcLists)]
        L.orc_force.argtypes = [C.POINTER(OrcSys), C.POINTER(OrcLists), C.c_int,
                                _dp, _dp, _dp, _dp, C.POINTER(OrcEnergy)]
        L.orc_need_rebuild.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.orc_need_rebuild.restype = C.c_int
        L.orc_table_pair.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int, _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp]
        L.orc_table_shift.argtypes = [C.c_int, _dp]
        L.orc_table_coul.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int, _dp, _dp, _dp, _dp,
                                     _dp]
        L.orc_kvectors.argtypes = [C.c_int, _ip, _ip, _ip, _ip]
        L.orc_kvectors.restype = C.c_int
        L.orc_ewald_consts.argtypes = [C.c_int, _dp, C.c_double, _dp, _dp]
        L.orc_lrc_total.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp,
                                    _dp]
        L.orc_ewald_recip_direct.argtypes = [C.POINTER(OrcSys), _dp, _dp]
        L.orc_ewald_recip_direct.restype = C.c_double
        L.orc_coul_real_direct.argtypes = [C.POINTER(OrcSys), _dp, _dp,
                                           C.c_double]
        L.orc_coul_real_direct.restype = C.c_double
        L.orc_lj_direct.argtypes = [C.POINTER(OrcSys), _dp, _dp, _dp, _dp, _dp]
        L.orc_lj_direct.restype = C.c_double
        L.orc_bonded_force.argtypes = [C.c_int, C.POINTER(OrcBonded), _dp, _dp, _dp]
        L.orc_bonded_force.restype = C.c_double
        _LIB = L
    return _LIB


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _scalar(sysd, key, default=0):
    v = sysd.get(key, default)
    return np.asarray(v).reshape(-1)[0].item() if np.size(v) else default


class _Bound:
    """An OrcSys whose pointers stay alive with their numpy owners."""

    def __init__(self, sysd):
        self.keep = []
        s = OrcSys()
        for name, typ in OrcSys._fields_:
            if typ is C.c_int:
                setattr(s, name, int(_scalar(sysd, name, 0)))
            elif typ is C.c_double:
                setattr(s, name, float(_scalar(sysd, name, 0.0)))
            elif name in sysd and np.size(sysd[name]) > 0:
                arr, ptr = (_d if typ is _dp else _i)(sysd[name])
                self.keep.append(arr)
                setattr(s, name, ptr)
        if s.ncell == 0:
            s.ncell = 7
        if s.nlen == 0:
            s.nlen = 500
        self.s = s


def build_lists(sysd, pos, hmat, use_cells=True):
    """get_ngbr + get_ngbr_e: dict of int32 arrays in cm3d's list order."""
    b = _Bound(sysd)
    p, pp = _d(pos)
    h, hp = _d(hmat)
    L = OrcLists()
    lib().orc_build_lists(C.byref(b.s), pp, hp, int(use_cells), C.byref(L))
    out = {}
    for cnt, a, bname in (("nprt", "inpt", "jnpt"), ("nprs", "inps", "jnps"),
                          ("nprte", "inpte", "jnpte"),
                          ("nprse", "inpse", "jnpse")):
        n = getattr(L, cnt)
        out[a] = np.ctypeslib.as_array(getattr(L, a), (n,)).copy() if n else np.zeros(0, np.int32)
        out[bname] = np.ctypeslib.as_array(getattr(L, bname), (n,)).copy() if n else np.zeros(0, np.int32)
    lib().orc_free_lists(C.byref(L))
    return out


def _lists_struct(lists):
    L = OrcLists()
    keep = []
    for cnt, a, bname in (("nprt", "inpt", "jnpt"), ("nprs", "inps", "jnps"),
                          ("nprte", "inpte", "jnpte"),
                          ("nprse", "inpse", "jnpse")):
        ai, ap = _i(lists.get(a, np.zeros(0, np.int32)))
        bi, bp = _i(lists.get(bname, np.zeros(0, np.int32)))
        keep += [ai, bi]
        setattr(L, cnt, len(ai))
        setattr(L, a, ap)
        setattr(L, bname, bp)
    return L, keep


def force(sysd, lists, pos, hmat, iflg=1):
    """force(...,iflg) before rmass: returns (fs, fl, energy-dict)."""
    b = _Bound(sysd)
    L, keep = _lists_struct(lists)
    p, pp = _d(pos)
    h, hp = _d(hmat)
    n3 = 3 * int(_scalar(sysd, "natoms"))
    fs = np.zeros(n3)
    fl = np.zeros(n3)
    e = OrcEnergy()
    lib().orc_force(C.byref(b.s), C.byref(L), int(iflg), pp, hp,
                    fs.ctypes.data_as(_dp), fl.ctypes.data_as(_dp), C.byref(e))
    return fs, fl, e.as_dict()


def need_rebuild(natoms, skin, pos, pos_build):
    p, pp = _d(pos)
    q, qp = _d(pos_build)
    return bool(lib().orc_need_rebuild(int(natoms), float(skin), pp, qp))


def ewald_recip_direct(sysd, pos, hmat):
    b = _Bound(sysd)
    p, pp = _d(pos)
    h, hp = _d(hmat)
    return lib().orc_ewald_recip_direct(C.byref(b.s), pp, hp)


def coul_real_direct(sysd, pos, hmat, rcut):
    b = _Bound(sysd)
    p, pp = _d(pos)
    h, hp = _d(hmat)
    return lib().orc_coul_real_direct(C.byref(b.s), pp, hp, float(rcut))


def lj_direct(sysd, pos, hmat, eps, sig):
    b = _Bound(sysd)
    p, pp = _d(pos)
    h, hp = _d(hmat)
    e, ep = _d(eps)
    s, sp = _d(sig)
    f = np.zeros(3 * int(_scalar(sysd, "natoms")))
    v = lib().orc_lj_direct(C.byref(b.s), pp, hp, ep, sp, f.ctypes.data_as(_dp))
    return v, f


def bonded_force(natoms, bonded, pos, parts=False):
    """Harmonic bond + bend forces (added to zero) and energy; with parts,
    also (vbond, vbend) as cm3d's getvbond / getvbend sum them."""
    B = OrcBonded()
    keep = []
    B.nbond = len(bonded.get("ib", []))
    B.nbend = len(bonded.get("ia", []))
    for name, typ in OrcBonded._fields_[2:]:
        arr, ptr = (_d if typ is _dp else _i)(bonded.get(name, np.zeros(0)))
        keep.append(arr)
        setattr(B, name, ptr)
    p, pp = _d(pos)
    f = np.zeros(3 * natoms)
    vp = np.zeros(2)
    v = lib().orc_bonded_force(int(natoms), C.byref(B), pp, f.ctypes.data_as(_dp),
                               vp.ctypes.data_as(_dp))
    return (v, f, (vp[0], vp[1])) if parts else (v, f)


# ---------------------------------------------------------------- set-up
POT_LJ, POT_WILLIAMS = 0, 1


def table_pair(pot, prm, rmin, rmax, rres, rheal, skin, ntable, scaleeps=1.0):
    prm_a, pp = _d(prm)
    v = np.zeros(ntable); dv = np.zeros(ntable); sw = np.zeros(ntable)
    cut = np.zeros(7); dx2 = np.zeros(1); vl = np.zeros(1); wl = np.zeros(1)
    lib().orc_table_pair(int(pot), pp, rmin, rmax, rres, rheal, skin, scaleeps,
                         int(ntable), *[a.ctypes.data_as(_dp) for a in
                                        (v, dv, sw, cut, dx2, vl, wl)])
    return dict(vtab=v, dvtab=dv, switchs=sw, cut2=cut, dx2tab=dx2[0],
                vlrc=vl[0], wlrc=wl[0])


def table_coul(ntable, iperd, alpha, rmin, rresp, rmax, rheal, skin, ishift=0):
    t = np.zeros(ntable); dt = np.zeros(ntable); sw = np.zeros(ntable)
    cut = np.zeros(6); dx2 = np.zeros(1)
    lib().orc_table_coul(int(ntable), int(iperd), alpha, rmin, rresp, rmax,
                         rheal, skin, int(ishift),
                         *[a.ctypes.data_as(_dp) for a in (t, dt, sw, cut, dx2)])
    return dict(coultab=t, dcoultab=dt, switchse=sw, cut2=cut, dx2coul=dx2[0])


def kvectors(kmax):
    n = lib().orc_kvectors(int(kmax), None, None, None, None)
    ka = np.zeros(n, np.int32); kb = np.zeros(n, np.int32)
    kc = np.zeros(n, np.int32); kup = np.zeros(n + 1, np.int32)
    lib().orc_kvectors(int(kmax), *[a.ctypes.data_as(_ip) for a in (ka, kb, kc, kup)])
    return ka, kb, kc, kup


def ewald_consts(qch, alpha):
    q, qp = _d(qch)
    s = np.zeros(1); b = np.zeros(1)
    lib().orc_ewald_consts(len(q), qp, alpha, s.ctypes.data_as(_dp),
                           b.ctypes.data_as(_dp))
    return s[0], b[0]


def lrc_total(itype, ntypes, mapping, vlrc, wlrc):
    it, itp = _i(itype)
    m, mp = _i(mapping)
    v, vp = _d(vlrc)
    w, wp = _d(wlrc)
    vo = np.zeros(1); wo = np.zeros(1)
    lib().orc_lrc_total(len(it), int(ntypes), itp, mp, vp, wp,
                        vo.ctypes.data_as(_dp), wo.ctypes.data_as(_dp))
    return vo[0], wo[0]
