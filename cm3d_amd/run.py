"""Run a cm3d MD deck on the device: read the deck and the files it names
(cm3d_amd.decks), take the velocities from a cold start on cm3d's own random
stream or from a restart file, integrate with the device-resident step
(NBMD, NVE or NVT), and produce what cm3d's MD control loop produces
(main/md_mdcntr.new.c:83-260): the screen report every iwr_screen steps
(output/md_writescreen.c), with cm3d's running averages (accumulate,
integrate/md_averages.new.c:95-136) beside the instantaneous values, the
.ham / .eng records every iwr_inst steps (output/md_saveinst.c), the restart
file every iwr_dump steps and at the end, and the final coordinates as
<initfile>_N (output/md_writestart.c).

    python -m cm3d_amd.run deck [--precision fp64|mixed] [--steps N] [--no-files]
                                [--no-averages]

cm3d evaluates its energies after every step; the averages do the same (one
energy evaluation per step).  --no-averages advances between report steps
in one hipGraph-replayed burst and reports the instantaneous values only.
"""
import argparse
import math
import os
import sys

import numpy as np

from . import decks, nb

PCONV_ATM = decks.PCONV_ATM


def setup(deck, precision="fp64"):
    """(sysd, run, ctx, md) for a deck, initialised (step_init done)."""
    sysd, run = decks.load(deck)
    ctx = nb.NBContext(sysd, precision=precision)
    md = nb.NBMD(ctx, sysd, run["dt_ps"], nres_intra=run["nres_intra"])
    if run["ensemble"] == "nvt":
        md.set_thermostat(run["nchain"], run["gkt"], run["mnh"], run["veta"], run["ithm"],
                          eta0=run["eta"])
    md.ndof = run["ndof"]
    md.set_state(sysd["pos"], run["vel"], sysd["hmat"])
    md.init()
    return sysd, run, ctx, md


class Averages:
    """cm3d's running sums (zero_averages / accumulate,
    md_averages.new.c:54-136), optionally continued from a restart file."""

    def __init__(self, rst=None, econi=0.0):
        self.s = {k: 0.0 for row in decks.AVERAGE_KEYS for k in row}
        self.s["avprs_tensor"] = np.zeros(9)
        if rst is not None:
            for row in decks.AVERAGE_KEYS:
                for k in row:
                    self.s[k] = rst[k]
            self.s["avprs_tensor"] = np.array(rst["avprs_tensor"], float)
        self.estep = 0
        self.deltae2 = 0.0
        self.econi = econi

    def add(self, e, vol):
        s = self.s
        ham = e["ham"]
        s["avpotra"] += e["potra"]
        s["avpoter"] += e["poter"]
        s["avzke"] += e["zke"]
        s["avpotn"] += e["potn"]
        s["avzken"] += e["zken"]
        s["avham"] += ham
        s["avprs"] += e["prsi"]
        s["avprs_tensor"] = s["avprs_tensor"] + np.asarray(e["prs_tensor"])
        s["avtemp"] += e["tiout"]
        s["avvol"] += vol
        s["avpot_inter"] += e["vinter"]
        s["avpot_bond"] += e["vbond"]
        s["avpot_bend"] += e["vbend"]
        s["avpot_elec"] += e["velec"]
        s["avpot_recip"] += e["vewald"]
        s["econv"] += ham
        s["econ2"] += ham * ham
        self.estep += 1
        d = (self.econi - ham)
        self.deltae2 += (d * d) / (self.econi * self.econi)


def report(e, natoms, nchain=0, ntherm=0, av=None, step=1):
    """cm3d's screen quantities (md_writescreen.c:83-390) as name ->
    (average, instantaneous); the average is 0 without running sums."""
    s = av.s if av is not None else None
    st = float(step) if step else 1.0

    def a(key, scale=1.0):
        return (s[key] * scale / st) if s is not None else 0.0

    rn = 1.0 / natoms
    r = {}
    if e["potra"]:
        r["Ura/N"] = (a("avpotra", rn), e["potra"] * rn)
    r["Uer/N"] = (a("avpoter", rn), e["poter"] * rn)
    r["Ke /N"] = (a("avzke", rn), e["zke"] * rn)
    r["inter"] = (a("avpot_inter"), e["vinter"])
    if ntherm:
        r["Ueta/NC"] = (a("avpotn") / (nchain * ntherm) if s else 0.0,
                        e["potn"] / (nchain * ntherm))
        r["Keta/NC"] = ((s["avzken"] * 2. / (st * (nchain * (ntherm + 1)))) if s else 0.0,
                        e["zken"] * 2. / (nchain * (ntherm + 1.)))
    r["Ham/N"] = (a("avham", rn), e["ham"] * rn)
    r["temp"] = (a("avtemp"), e["tiout"])
    r["pres"] = (a("avprs"), e["prsi"])
    r["pres_atm"] = (a("avprs", PCONV_ATM), e["prsi"] * PCONV_ATM)
    r["ptensor_atm"] = ((s["avprs_tensor"] * PCONV_ATM / st) if s else np.zeros(9),
                        np.asarray(e["prs_tensor"]) * PCONV_ATM)
    if s is not None and step > 1 and av.estep:
        sd = abs((st * s["econ2"] - s["econv"] * s["econv"]) / ((st - 1.) * st))
        r["dE"] = (math.sqrt(sd) * st / abs(s["econv"]), math.sqrt(av.deltae2) / av.estep)
    return r


def _vol(h):
    return float(np.linalg.det(np.asarray(h, float).reshape(3, 3)))


def run_deck(deck, precision="fp64", nsteps=None, files=True, out=None, graph=True,
             averages=True):
    """Integrate a deck; returns [(istep, report dict, energies dict)] for
    the initial report and every iwr_screen steps.  With files, the .ham /
    .eng records, the restart file (dpo_file) and <initfile>_N go next to
    the deck (or into `out`)."""
    sysd, run, ctx, md = setup(deck, precision)
    ctl = run["ctl"]
    n = int(sysd["natoms"])
    nsteps = run["nsteps"] if nsteps is None else int(nsteps)
    # istep of the first step: 1 (istart 1, 2) or the stored step + 1
    first = run["istep0"] + 1 if run["istart"] == 3 else 1
    last = first + nsteps - 1
    every = run["iwr_screen"]
    inst = run["iwr_inst"] if files else 0
    dump = run["iwr_dump"] if files else 0
    nchain = run.get("nchain", 0)
    ntherm = 1 if run["ensemble"] == "nvt" else 0
    d = out or run["base"]
    fham = feng = None
    mode = "a" if run["istart"] == 3 else "w"  # a continued run appends (md_headers.c:70-84)
    if inst > 0:
        fham = open(os.path.join(d, ctl["ham_file"] or "md.ham"), mode)
        feng = open(os.path.join(d, ctl["eng_file"] or "md.eng"), mode)
        if run["istart"] != 3:
            decks.write_inst_headers(fham, feng, n, ntherm, nchain, inst, run["dt_ps"])
    # .conf / .vel / .force every iwr_confp / iwr_confv / iwr_conff steps
    traj = {}
    for kind, key, fkey, hdr in (("conf", "iwr_confp", "pos_file", "iwr_confp"),
                                 ("vel", "iwr_confv", "vel_file", "iwr_confv"),
                                 ("force", "iwr_conff", "force_file", "iwr_confv")):
        k = run[key] if files else 0
        if k > 0:
            fh = open(os.path.join(d, ctl[fkey] or f"md.{kind}"), mode)
            if run["istart"] != 3:
                decks.write_traj_header(fh, n, run[hdr], run["dt_ps"])
            traj[kind] = (k, fh)

    def save_traj(istep):
        """cm3d's record order after a step: conf, vel, force
        (main/md_mdcntr.new.c:222-231)."""
        want = [kind for kind in ("conf", "vel", "force")
                if kind in traj and istep % traj[kind][0] == 0]
        if not want:
            return
        pos, vel = md.get_state()
        for kind in want:
            fh = traj[kind][1]
            if kind == "conf":
                decks.write_conf(fh, pos, sysd["hmat"])
            elif kind == "vel":
                decks.write_vel(fh, vel)
            else:
                fl, aa = md.get_forces()
                decks.write_force(fh, fl, aa, sysd["amass"])

    def restart(istep, av):
        pos, vel = md.get_state()
        veta, eta = md.get_thermostat()
        decks.write_restart(os.path.join(d, ctl["dpo_file"] or "md.save"), istep, n, ntherm,
                            nchain, run["temp"], run["pext"], run["dt_ps"],
                            av.s if av is not None else {}, pos, vel, sysd["hmat"],
                            [eta] if ntherm else None, [veta] if ntherm else None)
        return pos

    try:
        e = md.energies()
        av = Averages(run["restart"] if run["istart"] == 3 else None, econi=e["ham"]) \
            if averages else None
        # the initial report is printed with the step counter at `first`
        recs = [(first - 1, report(e, n, nchain, ntherm, av, first if av else 1), e)]
        vol = _vol(sysd["hmat"])
        if averages:
            for istep in range(first, last + 1):
                md.step(1, graph=graph)
                e = md.energies()
                if e["overflow"]:
                    raise nb.NBError(f"neighbour list overflow at step {istep}")
                av.add(e, vol)
                if every and istep % every == 0:
                    recs.append((istep, report(e, n, nchain, ntherm, av, istep), e))
                if dump and istep % dump == 0:
                    restart(istep, av)
                save_traj(istep)
                if inst and istep % inst == 0:
                    decks.write_inst(fham, feng, istep, run["dt_ps"], e)
        else:
            marks = set()
            for k in (every, inst, dump) + tuple(v[0] for v in traj.values()):
                if k:
                    marks |= {m for m in range(first, last + 1) if m % k == 0}
            marks = sorted(marks | {last})
            done = first - 1
            for istep in marks:
                md.step(istep - done, graph=graph)
                done = istep
                e = md.energies()
                if e["overflow"]:
                    raise nb.NBError(f"neighbour list overflow by step {istep}")
                if every and istep % every == 0:
                    recs.append((istep, report(e, n, nchain, ntherm), e))
                if dump and istep % dump == 0:
                    restart(istep, None)
                save_traj(istep)
                if inst and istep % inst == 0:
                    decks.write_inst(fham, feng, istep, run["dt_ps"], e)
        if files:  # finish(): restart + final coordinates
            pos = restart(last, av)
            decks.write_coord(os.path.join(d, os.path.basename(ctl["dpi_file"]) + "_N"), pos,
                              sysd["hmat"])
    finally:
        if fham:
            fham.close()
            feng.close()
        for _, fh in traj.values():
            fh.close()
        md.close()
        ctx.close()
    return recs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("deck")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--no-files", action="store_true")
    ap.add_argument("--no-averages", action="store_true")
    a = ap.parse_args(argv)
    for k, r, _ in run_deck(a.deck, a.precision, a.steps, files=not a.no_files,
                            averages=not a.no_averages):
        print(f"step {k}")
        for name, (avg, inst) in r.items():
            if name == "ptensor_atm":
                for i in range(3):
                    print("%10f %10f %10f ~ %10f %10f %10f atm" % (tuple(avg[3 * i:3 * i + 3]) +
                                                                 tuple(inst[3 * i:3 * i + 3])))
                continue
            unit = "atm" if name == "pres_atm" else ("K/A^3" if name == "pres" else "K")
            print(f"<{name:8s}>\t= {avg:15.6f}\t~ {inst:15.6f} {unit}")
    sys.stdout.flush()


if __name__ == "__main__":
    main()
