"""Run a cm3d MD deck on the device: read the deck and the files it names
(cm3d_amd.decks), cold-start the velocities (and the chain) from cm3d's own
random stream, integrate with the device-resident step (NBMD, NVE or NVT),
and report what cm3d reports every iwr_screen steps -- the instantaneous
column of its screen output (output/md_writescreen.c:83-330) -- and write
the .ham / .eng records every iwr_inst steps (output/md_saveinst.c).

    python -m cm3d_amd.run deck [--precision fp64|mixed] [--steps N] [--no-files]

cm3d's running averages (the left column) need the energies of every step;
this runner evaluates them only at the report steps, so it prints the
instantaneous values alone.
"""
import argparse
import os
import sys


from . import decks, nb

PCONV_ATM = 138.0622 / 1.01325  # K/A^3 -> atm (include/md_numbers.h:33)


def setup(deck, precision="fp64"):
    """(sysd, run, ctx, md) for a deck, initialised (step_init done)."""
    sysd, run = decks.load(deck)
    ctx = nb.NBContext(sysd, precision=precision)
    md = nb.NBMD(ctx, sysd, run["dt_ps"], nres_intra=run["nres_intra"])
    if run["ensemble"] == "nvt":
        md.set_thermostat(run["nchain"], run["gkt"], run["mnh"], run["veta"], run["ithm"])
    md.ndof = run["ndof"]
    md.set_state(sysd["pos"], run["vel"], sysd["hmat"])
    md.init()
    return sysd, run, ctx, md


def report(e, natoms, nchain=0, ntherm=0):
    """cm3d's instantaneous screen quantities (md_writescreen.c) from an
    NBMD.energies() dict, as a name -> value dict in cm3d's order."""
    r = {}
    if e["potra"]:
        r["Ura/N"] = e["potra"] / natoms
    r["Uer/N"] = e["poter"] / natoms
    r["Ke /N"] = e["zke"] / natoms
    r["inter"] = e["vinter"]
    if ntherm:
        r["Ueta/NC"] = e["potn"] / (nchain * ntherm)
        r["Keta/NC"] = e["zken"] * 2. / (nchain * (ntherm + 1.))
    r["Ham/N"] = e["ham"] / natoms
    r["temp"] = e["tiout"]
    r["pres"] = e["prsi"]
    r["pres_atm"] = e["prsi"] * PCONV_ATM
    return r


def run_deck(deck, precision="fp64", nsteps=None, files=True, out=None, graph=True):
    """Integrate a deck; returns [(step, report dict, energies dict)] for step
    0 (the initial report) and every iwr_screen steps.  With files, the .ham
    and .eng records go next to the deck (or into `out`)."""
    sysd, run, ctx, md = setup(deck, precision)
    n = int(sysd["natoms"])
    nsteps = run["nsteps"] if nsteps is None else int(nsteps)
    every = run["iwr_screen"] if run["iwr_screen"] > 0 else nsteps
    inst = run["iwr_inst"] if files else 0
    nchain = run.get("nchain", 0)
    ntherm = 1 if run["ensemble"] == "nvt" else 0
    fham = feng = None
    if inst > 0:
        ctl = run["ctl"]
        d = out or run["base"]
        fham = open(os.path.join(d, ctl["ham_file"] or "md.ham"), "w")
        feng = open(os.path.join(d, ctl["eng_file"] or "md.eng"), "w")
        decks.write_inst_headers(fham, feng, n, ntherm, nchain, inst, run["dt_ps"])
    try:
        e = md.energies()
        recs = [(0, report(e, n, nchain, ntherm), e)]
        marks = sorted(set(range(every, nsteps + 1, every)) |
                       (set(range(inst, nsteps + 1, inst)) if inst > 0 else set()))
        done = 0
        for k in marks:
            md.step(k - done, graph=graph)
            done = k
            e = md.energies()
            if e["overflow"]:
                raise nb.NBError(f"neighbour list overflow by step {k}")
            if inst > 0 and k % inst == 0:
                decks.write_inst(fham, feng, k, run["dt_ps"], e)
            if k % every == 0:
                recs.append((k, report(e, n, nchain, ntherm), e))
    finally:
        if fham:
            fham.close()
            feng.close()
        md.close()
        ctx.close()
    return recs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("deck")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--no-files", action="store_true")
    a = ap.parse_args(argv)
    for k, r, _ in run_deck(a.deck, a.precision, a.steps, files=not a.no_files):
        print(f"step {k}")
        for name, v in r.items():
            unit = "atm" if name == "pres_atm" else ("K/A^3" if name == "pres" else "K")
            print(f"  <{name:8s}> ~ {v:15.6f} {unit}")
    sys.stdout.flush()


if __name__ == "__main__":
    main()
