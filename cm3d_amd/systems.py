"""Synthetic cm3d systems for the nonbonded engine (seeded, deterministic).

Each builder returns a *system dict* with cm3d's field names (the arrays
nb_create takes) plus positions ``pos`` (3N SoA), box ``hmat``, masses
``amass`` (cm3d units, mass x MCONV), initial velocities ``vel`` and, for
water, the bonded terms of the flexible model.  Tables, cut^2 values and
k-vectors come from the library's host set-up helpers (restating
get_ter_parm / set_charges / ewald_setup), so a system built here is what
cm3d's input layer would build from the equivalent deck.

Workloads (units: Angstrom, Kelvin, e before the sqrt(CCONV) scaling):
  lj_fcc     864 argon atoms, FCC 6x6x6 at 1.374 g/cc, rc 10, skin 1
  water      flexible 3-site water, O-O LJ, erfc real space + Ewald
  mixture    50/50 LJ + charge binary mixture (three interactions)
"""
import math

import numpy as np

from . import nb

# include/md_numbers.h
KCAL = 503.22289                    # kcal/mol -> K
CCONV = 167102.4                    # e^2/A -> K
BOLTS = 1.38065812e-23
DCONV = BOLTS * 1.e23
MCONV = 10. / (6.022045 * DCONV)    # g/mol -> K ps^2 / A^2
AVOGADRO = 6.022045e23
SQRT_CCONV = math.sqrt(CCONV)


def _base(natoms, ntypes, ntable, skin, rheal):
    return dict(natoms=natoms, ntypes=ntypes, ntable=ntable, iperd=3, skin=skin,
                rheal=rheal, ncell=7, nlen=500)


def _pair_tables(sysd, inter, ntable, skin, rheal):
    """inter: list of (pot, prm, rmin, rmax, rres) in interaction index order."""
    ni = len(inter)
    vt, dvt, sw = [], [], []
    cuts = np.zeros((ni, 7))
    dx2 = np.zeros(ni)
    vl = np.zeros(ni)
    wl = np.zeros(ni)
    for k, (pot, prm, rmin, rmax, rres) in enumerate(inter):
        t = nb.table_pair(pot, prm, rmin, rmax, rres, rheal, skin, ntable)
        vt.append(t["vtab"]); dvt.append(t["dvtab"]); sw.append(t["switchs"])
        cuts[k] = t["cut2"]; dx2[k] = t["dx2tab"]; vl[k] = t["vlrc"]; wl[k] = t["wlrc"]
    sysd.update(ninter=ni, vtab=np.concatenate(vt), dvtab=np.concatenate(dvt),
                switchs=np.concatenate(sw), rmins_cut2=cuts[:, 0].copy(),
                rmaxs_cut2=cuts[:, 1].copy(), rminl_cut2=cuts[:, 2].copy(),
                rmaxl_cut2=cuts[:, 3].copy(), rmaxs_skin2=cuts[:, 4].copy(),
                rmaxl_skin2=cuts[:, 5].copy(), dx2tab=dx2, vlrc_pair=vl,
                wlrc_pair=wl,
                pair_params=np.array([[p[1][0], p[1][1], p[2], p[3], p[4]] for p in inter]))


def _lrc(sysd):
    """Tail-correction totals over type pairs (input/md_searchbase.c:212-224)."""
    it = sysd["itype"]
    nt = int(sysd["ntypes"])
    m = sysd["map"].reshape(nt, nt)
    cnt = np.array([(it == t).sum() for t in range(nt)], dtype=np.float64)
    v = w = 0.
    for i in range(nt):
        for j in range(nt):
            l = m[i, j]
            if l != -1:
                v += cnt[i] * cnt[j] * math.pi * 2. * sysd["vlrc_pair"][l]
                w += cnt[i] * cnt[j] * math.pi * 2. * sysd["wlrc_pair"][l] / 3.
    sysd["vlrc"] = v
    sysd["wlrc"] = w


def _coulomb(sysd, ntable, alpha, rmin, rresp, rmax, rheal, skin, kmax):
    t = nb.table_coul(ntable, 3, alpha, rmin, rresp, rmax, rheal, skin)
    c = t["cut2"]
    sysd.update(coultab=t["coultab"], dcoultab=t["dcoultab"], switchse=t["switchse"],
                dx2coul=t["dx2coul"], rmins_coul2=c[0], rmaxs_coul2=c[1],
                rminl_coul2=c[2], rmaxl_coul2=c[3], rmaxs_skin2_e=c[4],
                rmaxl_skin2_e=c[5])
    ka, kb, kc = nb.kvectors(kmax)
    kup = np.zeros(len(ka) + 1, np.int32)
    kup[0] = kup[-1] = 1      # row starts; the sentinel as ewald_setup leaves it
    if len(ka) > 1:
        kup[1:len(ka)] = ((ka[1:] != ka[:-1]) | (kb[1:] != kb[:-1])).astype(np.int32)
    self_, bkg = nb.ewald_consts(sysd["qch"], alpha)
    sysd.update(ktot=len(ka), ka=ka, kb=kb, kc=kc, kup=kup, kappa=alpha,
                ewald_self=self_, ewald_bkg=bkg, kmax=kmax)


def _exclusions(sysd, pairs):
    """Sorted exclusion rows (partners j > i) from unordered pairs; the
    coulomb rows keep the charged partners of charged atoms and the ecorr
    list every excluded charged pair (energy/control/elec_ngbr.c:139-218)."""
    n = int(sysd["natoms"])
    rows = [[] for _ in range(n)]
    for a, b in pairs:
        i, j = (a, b) if a < b else (b, a)
        if j not in rows[i]:
            rows[i].append(j)
    for r in rows:
        r.sort()
    q = sysd["qch"]
    crows = [[j for j in rows[i] if q[j] != 0.] if q[i] != 0. else [] for i in range(n)]
    sysd["nexclude"] = np.array([len(r) for r in rows], np.int32)
    sysd["exclude"] = np.array([j for r in rows for j in r], np.int32)
    sysd["cnexclude"] = np.array([len(r) for r in crows], np.int32)
    sysd["cexclude"] = np.array([j for r in crows for j in r], np.int32)
    ec = [(i, j) for i in range(n) for j in crows[i]]
    sysd["necorr"] = len(ec)
    sysd["iecorr"] = np.array([a for a, _ in ec], np.int32)
    sysd["jecorr"] = np.array([b for _, b in ec], np.int32)


def maxwell_velocities(amass, temp, seed):
    """Gaussian velocities at temperature temp (K), zero total momentum,
    rescaled to exactly temp over 3N-3 degrees of freedom (the effect of
    cm3d's samvel, integrate/md_zeroam.c:33-46, with numpy's RNG)."""
    rng = np.random.default_rng(seed)
    n = len(amass)
    v = rng.standard_normal((3, n)) * np.sqrt(temp / amass)[None, :]
    p = (v * amass[None, :]).sum(axis=1) / amass.sum()
    v -= p[:, None]
    ke2 = (amass[None, :] * v * v).sum()
    v *= math.sqrt(temp * (3 * n - 3) / ke2)
    return v.reshape(-1)


def lj_fcc(ncells=6, density=1.374, eps_kcal=0.238, sig=3.405, mass=39.95,
           rmin=2.5, rc=10.0, skin=1.0, rheal=0.5, ntable=500, temp=120.0,
           seed=12345):
    """Argon on an FCC lattice, positions centred on the origin."""
    n = 4 * ncells ** 3
    vol = n * mass / (AVOGADRO * density) * 1e24
    L = vol ** (1. / 3.)
    a = L / ncells
    base = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    g = np.stack(np.meshgrid(np.arange(ncells), np.arange(ncells), np.arange(ncells),
                             indexing="ij"), -1).reshape(-1, 3)
    xyz = ((g[:, None, :] + base[None, :, :]).reshape(-1, 3) + .25) * a - L / 2.
    s = _base(n, 1, ntable, skin, rheal)
    s.update(itype=np.zeros(n, np.int32), map=np.zeros(1, np.int32),
             qch=np.zeros(n), hmat=np.diag([L, L, L]).reshape(-1).astype(np.float64),
             pos=xyz.T.reshape(-1).copy(), amass=np.full(n, mass * MCONV))
    _pair_tables(s, [(nb.POT_LJ, [eps_kcal * KCAL, sig], rmin, rc, rc)], ntable, skin, rheal)
    _lrc(s)
    _exclusions(s, [])
    s.update(ktot=0, necorr=0, kappa=0.0, ewald_self=0.0, ewald_bkg=0.0)
    s["vel"] = maxwell_velocities(s["amass"], temp, seed)
    s["name"] = f"lj_fcc_{n}"
    return s


def _rand_rotations(rng, m):
    q = rng.standard_normal((m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((m, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water(nside=16, density=1.0, rc=10.0, alpha=0.35, kmax=9, skin=1.0, rheal=0.5,
          ntable=500, rcute_min=0.5, seed=2024, temp=300.0, qo=-0.8476,
          sig=3.166, eps_kcal=0.155, roh=1.0, theta=109.47, kb_kcal=1059.,
          ka_kcal=75.9, ljmin=1.5):
    """nside^3 flexible 3-site waters: O on a cubic lattice, random rigid
    orientations (the configs' 'random non-overlapping placement')."""
    nmol = nside ** 3
    mw = 15.9994 + 2 * 1.008
    L = (nmol * mw / (AVOGADRO * density) * 1e24) ** (1. / 3.)
    a = L / nside
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(nside), np.arange(nside), np.arange(nside),
                             indexing="ij"), -1).reshape(-1, 3)
    O = (g + .5) * a - L / 2.
    th = math.radians(theta)
    loc = np.array([[0., 0., 0.], [roh, 0., 0.], [roh * math.cos(th), roh * math.sin(th), 0.]])
    R = _rand_rotations(rng, nmol)
    xyz = O[:, None, :] + np.einsum("mij,kj->mki", R, loc)
    xyz = xyz.reshape(-1, 3)
    n = 3 * nmol
    itype = np.tile(np.array([0, 1, 1], np.int32), nmol)
    q = np.tile(np.array([qo, -qo / 2., -qo / 2.]), nmol) * SQRT_CCONV
    s = _base(n, 2, ntable, skin, rheal)
    s.update(itype=itype, map=np.array([0, -1, -1, -1], np.int32), qch=q,
             hmat=np.diag([L, L, L]).reshape(-1).astype(np.float64),
             pos=xyz.T.reshape(-1).copy(),
             amass=np.tile(np.array([15.9994, 1.008, 1.008]) * MCONV, nmol))
    _pair_tables(s, [(nb.POT_LJ, [eps_kcal * KCAL, sig], ljmin, rc, rc)], ntable, skin, rheal)
    _lrc(s)
    m3 = 3 * np.arange(nmol)
    pairs = list(zip(m3, m3 + 1)) + list(zip(m3, m3 + 2)) + list(zip(m3 + 1, m3 + 2))
    _exclusions(s, pairs)
    _coulomb(s, ntable, alpha, rcute_min, rc, rc, rheal, skin, kmax)
    s.update(bond_i=np.concatenate([m3 + 1, m3 + 2]).astype(np.int32),
             bond_j=np.concatenate([m3, m3]).astype(np.int32),
             bond_k=np.full(2 * nmol, kb_kcal * KCAL), bond_r0=np.full(2 * nmol, roh),
             bend_i=(m3 + 1).astype(np.int32), bend_j=m3.astype(np.int32),
             bend_k=(m3 + 2).astype(np.int32), bend_fk=np.full(nmol, ka_kcal * KCAL),
             bend_th0=np.full(nmol, math.radians(theta)))
    s["mol_start"] = (3 * np.arange(nmol + 1)).astype(np.int32)
    s["vel"] = maxwell_velocities(s["amass"], temp, seed + 1)
    s["name"] = f"water_{nmol}"
    return s


def mixture(natoms=2000, seed=7, rc=9.0, alpha=0.35, kmax=7, skin=1.0, rheal=0.5,
            ntable=500, sig=(3.0, 3.4), eps_kcal=(0.2, 0.3), charge=0.5,
            rho_reduced=0.8, sigbar=3.2, mass=(39.95, 39.95), temp=300.0,
            ljmin=1.5):
    """50/50 LJ + charge binary mixture (Lorentz-Berthelot), neutral, on a
    jittered simple-cubic lattice at rho = 0.8 sigbar^-3."""
    natoms = 2 * (natoms // 2)
    rng = np.random.default_rng(seed)
    rho = rho_reduced / sigbar ** 3
    L = (natoms / rho) ** (1. / 3.)
    m = int(math.ceil(natoms ** (1. / 3.)))
    a = L / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"),
                 -1).reshape(-1, 3)
    sites = g[rng.permutation(len(g))[:natoms]]
    xyz = (sites + .5) * a - L / 2. + rng.uniform(-.1, .1, (natoms, 3)) * a
    itype = np.zeros(natoms, np.int32)
    itype[natoms // 2:] = 1     # species-major, as cm3d orders molecules
    q = np.where(itype == 0, charge, -charge) * SQRT_CCONV
    s = _base(natoms, 2, ntable, skin, rheal)
    s.update(itype=itype, map=np.array([0, 1, 1, 2], np.int32), qch=q,
             hmat=np.diag([L, L, L]).reshape(-1).astype(np.float64),
             pos=xyz.T.reshape(-1).copy(),
             amass=np.where(itype == 0, mass[0], mass[1]) * MCONV)
    inter = []
    for (i, j) in ((0, 0), (0, 1), (1, 1)):
        sg = .5 * (sig[i] + sig[j])
        ep = math.sqrt(eps_kcal[i] * eps_kcal[j]) * KCAL
        inter.append((nb.POT_LJ, [ep, sg], ljmin, rc, rc))
    _pair_tables(s, inter, ntable, skin, rheal)
    _lrc(s)
    _exclusions(s, [])
    _coulomb(s, ntable, alpha, 0.5, rc, rc, rheal, skin, kmax)
    s["vel"] = maxwell_velocities(s["amass"], temp, seed + 1)
    s["name"] = f"mixture_{natoms}"
    return s


def kinetic_temperature(vel, amass):
    """Instantaneous temperature (K) over 3N-3 degrees of freedom (kB = 1
    in cm3d's units: sum m v^2 is in K)."""
    n = len(amass)
    v = np.asarray(vel).reshape(3, n)
    return float((amass[None, :] * v * v).sum() / (3 * n - 3))


def equilibrate(mds, systems, temp, nsteps, every=10, stepper=None):
    """Bring lattice starts to temperature: advance nsteps outer steps,
    rescaling every system's velocities to temp every `every` steps (a set-up
    device, not one of cm3d's thermostats).  mds: NBMD objects (already
    init()ed); stepper(k) advances all of them by k steps (default: each
    md.step(k))."""
    if stepper is None:
        def stepper(k):
            for m in mds:
                m.step(k, graph=True)
    done = 0
    while done < nsteps:
        k = min(every, nsteps - done)
        stepper(k)
        done += k
        for m, s in zip(mds, systems):
            pos, vel = m.get_state()
            t = kinetic_temperature(vel, s["amass"])
            if t > 0:
                m.set_state(pos, vel * math.sqrt(temp / t), s["hmat"])


def rocksalt(ncell=4, a=5.64, alpha=0.35, kmax=10, rc=10.0, skin=1.0, rheal=0.5,
             ntable=5000, mass=(22.99, 35.45), seed=5):
    """NaCl rock-salt crystal: ncell^3 conventional cells (8 ions each,
    charges +-1 e) with lattice constant a, centred on the origin.  LJ with a
    negligible well keeps the vdW path on; the electrostatic energy of the
    perfect lattice is -M N/2 e^2/(a/2), M = 1.747565 (Madelung)."""
    m = 2 * ncell
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"),
                 -1).reshape(-1, 3)
    n = len(g)
    L = ncell * a
    xyz = g * (a / 2.) - L / 2.
    itype = ((g.sum(axis=1)) % 2).astype(np.int32)   # 0: Na+, 1: Cl-
    q = np.where(itype == 0, 1.0, -1.0) * SQRT_CCONV
    s = _base(n, 2, ntable, skin, rheal)
    s.update(itype=itype, map=np.array([0, 0, 0, 0], np.int32), qch=q,
             hmat=np.diag([L, L, L]).reshape(-1).astype(np.float64),
             pos=xyz.T.reshape(-1).copy(),
             amass=np.where(itype == 0, mass[0], mass[1]) * MCONV)
    _pair_tables(s, [(nb.POT_LJ, [1e-12, 2.5], 1.0, rc, rc)], ntable, skin, rheal)
    _lrc(s)
    _exclusions(s, [])
    _coulomb(s, ntable, alpha, 0.5, rc, rc, rheal, skin, kmax)
    s["vel"] = maxwell_velocities(s["amass"], 10.0, seed)
    s["name"] = f"nacl_{n}"
    return s
