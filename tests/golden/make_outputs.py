"""Golden trajectory output files (.conf / .vel / .force: save_conf, save_vel
and save_force, output/md_saveout.c:30-90) written by the compiled cm3d
program itself (oracle/_ref/cm3d, built by `make -C oracle ref-exe` from the
reference sources where they lie).

Run in the build container only (the program is built there):
    python tests/golden/make_outputs.py

Cases, each a deck directory under tests/golden/ with cm3d's outputs in its
cm3d/ subdirectory:
  outputs_lj     the reference's test/lj NVE deck (256 LJ atoms), cut to 100
                 steps, writing positions and velocities every 10 steps (as
                 the deck does) and forces every 20
  outputs_water  64 flexible waters (erfc real space + Ewald kmax 5, RESPA
                 4 x 0.5 / 2 fs), 10 steps, positions and forces every 2
                 steps, velocities every 5
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from cm3d_amd import systems as S  # noqa: E402
from decks import write_deck  # noqa: E402
from make_fixtures import coul_keys  # noqa: E402

CM3D = os.path.join(ROOT, "oracle", "_ref", "cm3d")


def run_cm3d(workdir, deck):
    r = subprocess.run([CM3D, deck], cwd=workdir, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])


def keep(case, workdir, deck_files, outputs):
    dst = os.path.join(HERE, case)
    shutil.rmtree(dst, ignore_errors=True)
    os.makedirs(os.path.join(dst, "cm3d"))
    for f in deck_files:
        shutil.copy(os.path.join(workdir, f), os.path.join(dst, f))
    for f in outputs:
        shutil.copy(os.path.join(workdir, f), os.path.join(dst, "cm3d", f))
    print(case, {f: os.path.getsize(os.path.join(dst, "cm3d", f)) for f in outputs})


def lj_case():
    with tempfile.TemporaryDirectory() as tmp:
        for f in os.listdir(os.path.join(HERE, "cm3d_lj")):
            shutil.copy(os.path.join(HERE, "cm3d_lj", f), tmp)
        text = open(os.path.join(tmp, "lj.input_ljnve")).read()
        text = text.replace(r"\ntime{1000}", r"\ntime{100}")
        text = re.sub(r"\\iwr_dump\{\d+\}", r"\\iwr_dump{1000}", text)
        text += "\\iwr_conff{20} \\force_file{lj.force}\n"
        open(os.path.join(tmp, "lj.input_out"), "w").write(text)
        run_cm3d(tmp, "lj.input_out")
        keep("outputs_lj", tmp, ["lj.input_out", "lj.init", "lj.set", "lj.top", "data_base"],
             ["lj.conf", "lj.vel", "lj.force"])


def water_case():
    w = S.water(nside=4, kmax=5, seed=11)
    extra = coul_keys(w) + [r"\iwr_confp{2}", r"\iwr_confv{5}", r"\iwr_conff{2}",
                            r"\pos_file{sys.conf}", r"\vel_file{sys.vel}",
                            r"\force_file{sys.force}"]
    with tempfile.TemporaryDirectory() as tmp:
        write_deck(tmp, w, "water", nstep=10, dt=2.0, nres_intra=4, temp=300.0, extra=extra)
        run_cm3d(tmp, "sys.input")
        files = [f for f in os.listdir(tmp) if f in ("sys.input", "sys.init", "sys.set",
                                                     "data_base") or f.endswith(".top")]
        keep("outputs_water", tmp, files, ["sys.conf", "sys.vel", "sys.force"])


def main():
    if not os.path.exists(CM3D):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref-exe"], check=True)
    lj_case()
    water_case()


if __name__ == "__main__":
    main()
