"""ctypes binding of libcm3d_nb.so (the C ABI in include/cm3d_nb.h).

This is the host-side mirror of cm3d's inter-force interface for Python
callers (tests, bench, the serving loop): ``NBContext`` wraps one device
context and exposes the cm3d entry points it replaces --

    build_lists   get_ngbr + get_ngbr_e  (energy/control/ngbr_list.c:297)
    need_rebuild  nrefsh's criterion     (energy/control/md_finter.c:43-73)
    force         force(...,iflg) minus rmass (energy/control/md_fcntr.c:44)

A *system* is a dict of numpy arrays keyed like cm3d's INTER / COUL /
SIMPARMS fields (``natoms``, ``itype``, ``map``, ``vtab`` ...), as built by
:mod:`cm3d_amd.systems`.  The library must be built (``python -m
cm3d_amd.build``); there is no fallback: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcm3d_nb.so")

NB_OK = 0
NB_PREC_FP64 = 0
NB_PREC_MIXED = 1
PRECISIONS = {"fp64": NB_PREC_FP64, "mixed": NB_PREC_MIXED}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class NBParams(C.Structure):
    _fields_ = [
        ("natoms", C.c_int), ("ntypes", C.c_int), ("ninter", C.c_int),
        ("ntable", C.c_int), ("iperd", C.c_int),
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
        ("ktot", C.c_int), ("ka", _ip), ("kb", _ip), ("kc", _ip),
        ("kappa", C.c_double), ("ewald_self", C.c_double),
        ("ewald_bkg", C.c_double),
        ("precision", C.c_int), ("device", C.c_int), ("sharing", C.c_int),
    ]


class NBEnergy(C.Structure):
    _fields_ = [("spotl", C.c_double), ("spothl", C.c_double),
                ("sprsl", C.c_double), ("sprslt", C.c_double * 9),
                ("spote", C.c_double), ("spothe", C.c_double),
                ("sprse", C.c_double), ("sprset", C.c_double * 9),
                ("vk_ewald", C.c_double), ("vp_ewald", C.c_double),
                ("vp_ewald_tensor", C.c_double * 9),
                ("pot_ecorr", C.c_double), ("prs_ecorr", C.c_double),
                ("prs_ecorr_tensor", C.c_double * 9), ("nclamp", C.c_int),
                ("spote_phys", C.c_double), ("spothe_phys", C.c_double),
                ("sprse_phys", C.c_double), ("sprset_phys", C.c_double * 9)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(v[:]) if hasattr(v, "__len__") else v
        return out


class NBMDParams(C.Structure):
    _fields_ = [("amass", _dp), ("nmol", C.c_int), ("mol_start", _ip),
                ("nbond", C.c_int), ("bond_i", _ip), ("bond_j", _ip),
                ("bond_k", _dp), ("bond_r0", _dp),
                ("nbend", C.c_int), ("bend_i", _ip), ("bend_j", _ip), ("bend_k", _ip),
                ("bend_fk", _dp), ("bend_th0", _dp),
                ("dt", C.c_double), ("nres_intra", C.c_int), ("nres_lrf", C.c_int)]


class NBMDEnergy(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("vinter", "vhinter", "velec", "vewald", "vbond",
                                          "vbend", "zke", "poter", "poth", "potra", "potn",
                                          "zken", "ham", "winter", "welec", "wewald", "wbond",
                                          "prsi")] + \
               [("prs_tensor", C.c_double * 9), ("tiout", C.c_double), ("overflow", C.c_int)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = np.array(v[:]) if hasattr(v, "__len__") else v
        return out


# every symbol include/cm3d_nb.h declares
EXPORTS = ["nb_create", "nb_destroy", "nb_last_error", "nb_build_lists",
           "nb_need_rebuild", "nb_force", "nb_force_device", "nb_energy_only",
           "nb_build_lists_device", "nb_list_counts", "nb_get_pairs", "nb_timing",
           "nb_footprint", "nb_recip_info", "nb_list_reserve", "nb_force_device_noalloc",
           "nb_md_create", "nb_md_destroy", "nb_md_set_state", "nb_md_get_state",
           "nb_md_init", "nb_md_step", "nb_md_energies", "nb_md_rebuilds", "nb_md_get_forces",
           "nb_md_set_thermostat", "nb_md_get_thermostat",
           "nb_batch_create", "nb_batch_destroy", "nb_batch_step", "nb_batch_launches",
           "nb_batch_recoveries", "nb_table_pair", "nb_table_coul", "nb_kvectors", "nb_ewald_consts"]

_LIB = None


class NBError(RuntimeError):
    pass


def lib():
    """Load libcm3d_nb.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NBError("libcm3d_nb.so is not built (python -m cm3d_amd.build)")
    L = C.CDLL(LIB_PATH)
    L.nb_create.argtypes = [C.POINTER(NBParams), C.POINTER(C.c_void_p)]
    L.nb_destroy.argtypes = [C.c_void_p]
    L.nb_destroy.restype = None
    L.nb_last_error.restype = C.c_char_p
    L.nb_build_lists.argtypes = [C.c_void_p, _dp, _dp]
    L.nb_need_rebuild.argtypes = [C.c_void_p, _dp, _ip]
    L.nb_force.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp,
                           C.POINTER(NBEnergy)]
    L.nb_force_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, _dp,
                                  C.c_void_p, C.c_void_p, C.POINTER(NBEnergy),
                                  C.c_void_p]
    L.nb_build_lists_device.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_void_p]
    L.nb_force_device_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, _dp,
                                          C.c_void_p, C.c_void_p, C.POINTER(NBEnergy),
                                          C.c_void_p]
    L.nb_energy_only.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(NBEnergy)]
    L.nb_list_counts.argtypes = [C.c_void_p, _ip]
    L.nb_footprint.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    L.nb_recip_info.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.nb_get_pairs.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_int, _ip]
    L.nb_timing.argtypes = [C.c_void_p, C.c_int, _dp]
    L.nb_md_create.argtypes = [C.c_void_p, C.POINTER(NBMDParams), C.POINTER(C.c_void_p)]
    L.nb_md_destroy.argtypes = [C.c_void_p]
    L.nb_md_destroy.restype = None
    L.nb_md_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.nb_md_get_state.argtypes = [C.c_void_p, _dp, _dp]
    L.nb_md_get_forces.argtypes = [C.c_void_p, _dp, _dp]
    L.nb_md_init.argtypes = [C.c_void_p]
    L.nb_md_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    L.nb_md_energies.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int,
                                  C.POINTER(NBMDEnergy)]
    L.nb_md_rebuilds.argtypes = [C.c_void_p, _ip]
    L.nb_md_set_thermostat.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _ip]
    L.nb_md_get_thermostat.argtypes = [C.c_void_p, _dp, _dp]
    L.nb_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
    L.nb_batch_destroy.argtypes = [C.c_void_p]
    L.nb_batch_destroy.restype = None
    L.nb_batch_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    L.nb_batch_launches.argtypes = [C.c_void_p, _ip]
    L.nb_batch_recoveries.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.nb_list_reserve.argtypes = [C.c_void_p, C.c_int]
    L.nb_table_pair.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.nb_table_coul.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.nb_kvectors.argtypes = [C.c_int, _ip, _ip, _ip]
    L.nb_ewald_consts.argtypes = [C.c_int, _dp, C.c_double, _dp, _dp]
    _LIB = L
    return L


def check(rc, what="nb call"):
    if rc != NB_OK:
        msg = lib().nb_last_error().decode(errors="replace")
        raise NBError(f"{what} failed ({rc}): {msg}")


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _scalar(sysd, key, default=0):
    v = sysd.get(key, default)
    return np.asarray(v).reshape(-1)[0].item() if np.size(v) else default


def make_params(sysd, precision="fp64", device=0, sharing=1):
    """NBParams from a system dict; returns (params, keep-alive list)."""
    p = NBParams()
    keep = []
    for name, typ in NBParams._fields_:
        if name in ("precision", "device", "sharing"):
            continue
        if typ is C.c_int:
            setattr(p, name, int(_scalar(sysd, name, 0)))
        elif typ is C.c_double:
            setattr(p, name, float(_scalar(sysd, name, 0.0)))
        elif name in sysd and np.size(sysd[name]) > 0:
            arr, ptr = (_d if typ is _dp else _i)(sysd[name])
            keep.append(arr)
            setattr(p, name, ptr)
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p.device = int(device)
    p.sharing = int(sharing)
    return p, keep


class NBContext:
    """One device context of the nonbonded engine (a cm3d system)."""

    def __init__(self, sysd, precision="fp64", device=0, sharing=1):
        """sharing: systems expected to run on the GPU together with this one
        (a batch's or a server's size); it only sizes the reciprocal sum's
        work split."""
        self.natoms = int(_scalar(sysd, "natoms"))
        params, keep = make_params(sysd, precision, device, sharing)
        h = C.c_void_p()
        check(lib().nb_create(C.byref(params), C.byref(h)), "nb_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().nb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def build_lists(self, pos, hmat):
        p, pp = _d(pos)
        h, hp = _d(hmat)
        check(lib().nb_build_lists(self._h, pp, hp), "nb_build_lists")

    def need_rebuild(self, pos):
        p, pp = _d(pos)
        f = C.c_int(0)
        check(lib().nb_need_rebuild(self._h, pp, C.byref(f)), "nb_need_rebuild")
        return bool(f.value)

    def force(self, pos, hmat, iflg=1):
        """(fs, fl, energy dict) -- cm3d's force(...,iflg) before rmass."""
        p, pp = _d(pos)
        h, hp = _d(hmat)
        n3 = 3 * self.natoms
        fs = np.zeros(n3)
        fl = np.zeros(n3)
        e = NBEnergy()
        check(lib().nb_force(self._h, int(iflg), pp, hp, fs.ctypes.data_as(_dp),
                             fl.ctypes.data_as(_dp), C.byref(e)), "nb_force")
        return fs, fl, e.as_dict()

    def energy_only(self, pos, hmat):
        """Energies and virials of force(...,1) without the forces (the
        getvireal(...,0) pair sums plus the Ewald and ecorr terms)."""
        p, pp = _d(pos)
        h, hp = _d(hmat)
        e = NBEnergy()
        check(lib().nb_energy_only(self._h, pp, hp, C.byref(e)), "nb_energy_only")
        return e.as_dict()

    def force_device(self, iflg, d_pos, hmat, d_fs, d_fl, stream=None, energy=False):
        """Device-pointer variant (ints from torch.Tensor.data_ptr())."""
        h, hp = _d(hmat)
        e = NBEnergy() if energy else None
        check(lib().nb_force_device(self._h, int(iflg), C.c_void_p(d_pos), hp,
                                    C.c_void_p(d_fs) if d_fs else None,
                                    C.c_void_p(d_fl) if d_fl else None,
                                    C.byref(e) if e is not None else None,
                                    C.c_void_p(stream) if stream else None),
              "nb_force_device")
        return e.as_dict() if e is not None else None

    def force_device_noalloc(self, iflg, d_pos, hmat, d_fs, d_fl, stream=None, energy=False):
        """force_device with its scratch in the library's static area: no
        device allocation (nb_force_device_noalloc)."""
        h, hp = _d(hmat)
        e = NBEnergy() if energy else None
        check(lib().nb_force_device_noalloc(self._h, int(iflg), C.c_void_p(d_pos), hp,
                                            C.c_void_p(d_fs) if d_fs else None,
                                            C.c_void_p(d_fl) if d_fl else None,
                                            C.byref(e) if e is not None else None,
                                            C.c_void_p(stream) if stream else None),
              "nb_force_device_noalloc")
        return e.as_dict() if e is not None else None

    def build_lists_device(self, d_pos, hmat, stream=None):
        h, hp = _d(hmat)
        check(lib().nb_build_lists_device(self._h, C.c_void_p(d_pos), hp,
                                          C.c_void_p(stream) if stream else None),
              "nb_build_lists_device")

    def timing_start(self):
        check(lib().nb_timing(self._h, 1, None), "nb_timing")

    def timing_read(self):
        """Average ms per force call by phase: pack, pair, ewchg (charge
        records), sfac (structure factor), kred (k reduction, with the
        matrix-core force pass's coefficient image), ewforce (reciprocal
        force), fin (ecorr + force sum + energy reductions),
        ewald (ewchg .. fin), total."""
        out = np.zeros(9)
        check(lib().nb_timing(self._h, 2, out.ctypes.data_as(_dp)), "nb_timing")
        check(lib().nb_timing(self._h, 0, None), "nb_timing")
        d = dict(calls=int(out[0]), pack=out[1], pair=out[2], ewchg=out[3], sfac=out[4],
                 kred=out[5], ewforce=out[6], fin=out[7], total=out[8])
        d["ewald"] = d["ewchg"] + d["sfac"] + d["kred"] + d["ewforce"] + d["fin"]
        return d

    def list_counts(self):
        """Half-pair counts as cm3d reports them: nprs, nprt, nprse, nprte."""
        c = (C.c_int * 4)()
        check(lib().nb_list_counts(self._h, c), "nb_list_counts")
        return dict(nprs=c[0], nprt=c[1], nprse=c[2], nprte=c[3])

    def footprint(self):
        """Device bytes held by this context and its MD state."""
        b = C.c_size_t(0)
        check(lib().nb_footprint(self._h, C.byref(b)), "nb_footprint")
        return int(b.value)

    def recip_info(self):
        """Which reciprocal kernels this context launches (nb_recip_info):
        {"sfac_mfma", "force_mfma"} flags and the matrix-core instructions
        one launch of each issues."""
        o = (C.c_longlong * 4)()
        check(lib().nb_recip_info(self._h, o), "nb_recip_info")
        return {"sfac_mfma": bool(o[0] & 1), "force_mfma": bool(o[0] & 2),
                "sfac_mfma_insts": int(o[1]), "force_mfma_insts": int(o[2])}

    def list_reserve(self, slots):
        """Neighbour slots per atom for the next build (nb_list_reserve);
        drops the current lists."""
        check(lib().nb_list_reserve(self._h, int(slots)), "nb_list_reserve")

    def get_pairs(self, which):
        """Sorted (i<j) pair arrays of list 's', 't', 'se' or 'te'."""
        w = {"s": 0, "t": 1, "se": 2, "te": 3}[which]
        n = C.c_int(0)
        lib().nb_get_pairs(self._h, w, None, None, 0, C.byref(n))
        ip = np.zeros(max(n.value, 1), np.int32)
        jp = np.zeros(max(n.value, 1), np.int32)
        check(lib().nb_get_pairs(self._h, w, ip.ctypes.data_as(_ip),
                                 jp.ctypes.data_as(_ip), len(ip), C.byref(n)),
              "nb_get_pairs")
        return ip[:n.value].copy(), jp[:n.value].copy()


# ---------------------------------------------------------------- set-up
POT_LJ, POT_WILLIAMS = 0, 1


def table_pair(pot, prm, rmin, rmax, rres, rheal, skin, ntable, scaleeps=1.0):
    """get_ter_parm's table for one interaction (host)."""
    prm_a, pp = _d(prm)
    v = np.zeros(ntable); dv = np.zeros(ntable); sw = np.zeros(ntable)
    cut = np.zeros(7); dx2 = np.zeros(1); vl = np.zeros(1); wl = np.zeros(1)
    check(lib().nb_table_pair(int(pot), pp, rmin, rmax, rres, rheal, skin, scaleeps,
                              int(ntable), *[a.ctypes.data_as(_dp) for a in
                                             (v, dv, sw, cut, dx2, vl, wl)]),
          "nb_table_pair")
    return dict(vtab=v, dvtab=dv, switchs=sw, cut2=cut, dx2tab=dx2[0],
                vlrc=vl[0], wlrc=wl[0])


def table_coul(ntable, iperd, alpha, rmin, rresp, rmax, rheal, skin, ishift=0):
    """set_charges' erfc(a r)/r (or 1/r) table (host)."""
    t = np.zeros(ntable); dt = np.zeros(ntable); sw = np.zeros(ntable)
    cut = np.zeros(6); dx2 = np.zeros(1)
    check(lib().nb_table_coul(int(ntable), int(iperd), alpha, rmin, rresp, rmax,
                              rheal, skin, int(ishift),
                              *[a.ctypes.data_as(_dp) for a in (t, dt, sw, cut, dx2)]),
          "nb_table_coul")
    return dict(coultab=t, dcoultab=dt, switchse=sw, cut2=cut, dx2coul=dx2[0])


def kvectors(kmax):
    n = lib().nb_kvectors(int(kmax), None, None, None)
    ka = np.zeros(n, np.int32); kb = np.zeros(n, np.int32); kc = np.zeros(n, np.int32)
    lib().nb_kvectors(int(kmax), *[a.ctypes.data_as(_ip) for a in (ka, kb, kc)])
    return ka, kb, kc


def ewald_consts(qch, alpha):
    q, qp = _d(qch)
    s = np.zeros(1); b = np.zeros(1)
    check(lib().nb_ewald_consts(len(q), qp, alpha, s.ctypes.data_as(_dp),
                                b.ctypes.data_as(_dp)), "nb_ewald_consts")
    return s[0], b[0]


class NBMD:
    """Device-resident RESPA MD of one system around an NBContext (cm3d's
    step_init + step(), integrate/md_step.c:32-159, nres_lrf = 1).

    sysd needs ``amass`` and, for molecules, ``mol_start`` plus the bonded
    arrays ``bond_i/j/k/r0`` and ``bend_i/j/k/fk/th0`` (water from
    cm3d_amd.systems has them); without ``mol_start`` every atom is its own
    molecule."""

    def __init__(self, ctx, sysd, dt_ps, nres_intra=1):
        self.ctx = ctx
        n = ctx.natoms
        self.n = n
        self._keep = []
        p = NBMDParams()

        def put(name, arr, conv):
            a, ptr = conv(arr)
            self._keep.append(a)
            setattr(p, name, ptr)
            return len(a)

        put("amass", sysd["amass"], _d)
        ms = sysd.get("mol_start")
        if ms is None:
            ms = np.arange(n + 1, dtype=np.int32)
        p.nmol = put("mol_start", ms, _i) - 1
        p.nbond = len(sysd.get("bond_i", []))
        p.nbend = len(sysd.get("bend_i", []))
        for name in ("bond_i", "bond_j", "bend_i", "bend_j", "bend_k"):
            put(name, sysd.get(name, np.zeros(0, np.int32)), _i)
        for name, key in (("bond_k", "bond_k"), ("bond_r0", "bond_r0"),
                          ("bend_fk", "bend_fk"), ("bend_th0", "bend_th0")):
            put(name, sysd.get(key, np.zeros(0)), _d)
        p.dt = float(dt_ps)
        p.nres_intra = int(nres_intra)
        p.nres_lrf = 1
        h = C.c_void_p()
        check(lib().nb_md_create(ctx.handle, C.byref(p), C.byref(h)), "nb_md_create")
        self._h = h
        self.vlrc = float(np.asarray(sysd.get("vlrc", 0.0)).reshape(-1)[0])
        self.wlrc = float(np.asarray(sysd.get("wlrc", 0.0)).reshape(-1)[0])
        # cm3d's ndof for NVE with periodic boundaries (3N - 3); an NVT run
        # sets 3N (set_thermostat does)
        self.ndof = 3 * self.n - 3

    def close(self):
        if getattr(self, "_h", None):
            lib().nb_md_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, pos, vel, hmat):
        p, pp = _d(pos)
        v, vp = _d(vel)
        h, hp = _d(hmat)
        check(lib().nb_md_set_state(self._h, pp, vp, hp), "nb_md_set_state")

    def get_state(self):
        pos = np.zeros(3 * self.n)
        vel = np.zeros(3 * self.n)
        check(lib().nb_md_get_state(self._h, pos.ctypes.data_as(_dp),
                                    vel.ctypes.data_as(_dp)), "nb_md_get_state")
        return pos, vel

    def get_forces(self):
        """(fl, aa): the nonbonded force before rmass and the bonded
        accelerations of the current state (what save_force prints)."""
        fl = np.zeros(3 * self.n)
        aa = np.zeros(3 * self.n)
        check(lib().nb_md_get_forces(self._h, fl.ctypes.data_as(_dp), aa.ctypes.data_as(_dp)),
              "nb_md_get_forces")
        return fl, aa

    def init(self):
        check(lib().nb_md_init(self._h), "nb_md_init")

    def step(self, nsteps=1, stream=None, graph=True):
        check(lib().nb_md_step(self._h, int(nsteps), C.c_void_p(stream) if stream else None,
                               int(bool(graph))), "nb_md_step")

    def energies(self):
        e = NBMDEnergy()
        check(lib().nb_md_energies(self._h, self.vlrc, self.wlrc, int(self.ndof), C.byref(e)),
              "nb_md_energies")
        return e.as_dict()

    def set_thermostat(self, nchain, gkt, mnh, veta0, ithm, eta0=None):
        """One Nose-Hoover chain (cm3d's global thermostat) over atoms with
        ithm == 0; nchain 0 returns to NVE.  eta0 None starts eta at 0."""
        g, gp = _d(gkt)
        m, mp = _d(mnh)
        v, vp = _d(veta0)
        t, tp = _i(ithm)
        e, ep = _d(eta0) if eta0 is not None else (None, None)
        check(lib().nb_md_set_thermostat(self._h, int(nchain), gp, mp, vp, ep, tp),
              "nb_md_set_thermostat")
        self.nchain = int(nchain)
        self.ndof = 3 * self.n if nchain else 3 * self.n - 3

    def get_thermostat(self):
        """(veta, eta) of the chain now."""
        nc = getattr(self, "nchain", 0)
        veta, eta = np.zeros(nc), np.zeros(nc)
        if nc:
            check(lib().nb_md_get_thermostat(self._h, veta.ctypes.data_as(_dp),
                                             eta.ctypes.data_as(_dp)), "nb_md_get_thermostat")
        return veta, eta

    def rebuilds(self):
        c = C.c_int(0)
        check(lib().nb_md_rebuilds(self._h, C.byref(c)), "nb_md_rebuilds")
        return c.value


class NBBatch:
    """Several NBMD systems advanced together: every kernel of the step is
    launched once for all of them (nb_batch_*).  The systems must share the
    precision; their sizes, boxes, force fields and cutoffs may differ."""

    def __init__(self, mds):
        self.mds = list(mds)
        arr = (C.c_void_p * len(self.mds))(*[m._h for m in self.mds])
        h = C.c_void_p()
        check(lib().nb_batch_create(arr, len(self.mds), C.byref(h)), "nb_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().nb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1, stream=None, graph=True):
        """One tick of nsteps steps; returns when it is done (systems whose
        list build overflowed are redone with more slots).  nsteps=0 only
        re-records / captures the step, so that this can happen on the
        caller's thread between ticks run elsewhere."""
        check(lib().nb_batch_step(self._h, int(nsteps), C.c_void_p(stream) if stream else None,
                                  int(bool(graph))), "nb_batch_step")

    def recoveries(self):
        """Systems redone after a neighbour-list overflow so far."""
        c = C.c_longlong(0)
        check(lib().nb_batch_recoveries(self._h, C.byref(c)), "nb_batch_recoveries")
        return c.value

    def launches(self):
        c = C.c_int(0)
        check(lib().nb_batch_launches(self._h, C.byref(c)), "nb_batch_launches")
        return c.value
