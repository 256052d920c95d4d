"""cm3d's trajectory files from the device-resident run: the deck runner
(cm3d_amd/run.py) writes .conf / .vel / .force records (save_conf, save_vel,
save_force, output/md_saveout.c:30-90) at the deck's iwr_confp / iwr_confv /
iwr_conff intervals, compared with the files the compiled cm3d program wrote
for the same decks (tests/golden/make_outputs.py): the same headers and
lines, every number equal at the printed precision (%g, 6 significant
digits; the trajectories agree to ~1e-9, so a value may differ only where
it falls within that of a rounding boundary of the last printed digit)."""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _compare(mine, ref):
    a = open(mine).read().splitlines()
    b = open(ref).read().splitlines()
    assert len(a) == len(b), (mine, len(a), len(b))
    vals = np.array([float(t) for line in b if not line.startswith("#") for t in line.split()])
    scale = float(np.sqrt(np.mean(vals ** 2))) if len(vals) else 1.0
    tot = same = 0
    for la, lb in zip(a, b):
        if lb.startswith("#"):
            assert la == lb, (mine, la, lb)
            continue
        ta, tb = la.split(), lb.split()
        assert len(ta) == len(tb), (mine, la, lb)
        for u, v in zip(ta, tb):
            tot += 1
            same += u == v
            x, y = float(u), float(v)
            assert abs(x - y) <= 1.5e-5 * abs(y) + 1e-6 * scale, (mine, la, lb)
    assert same >= 0.99 * tot, (mine, same, tot)
    return tot, same


@pytest.mark.parametrize("case,deck,files", [
    ("outputs_lj", "lj.input_out", ("lj.conf", "lj.vel", "lj.force")),
    ("outputs_water", "sys.input", ("sys.conf", "sys.vel", "sys.force")),
])
def test_trajectory_files_match_cm3d(case, deck, files, tmp_path):
    from cm3d_amd import run as R
    src = os.path.join(GOLDEN, case)
    w = tmp_path / "w"
    shutil.copytree(src, w, ignore=shutil.ignore_patterns("cm3d"))
    print(f"[outputs] {case}", flush=True)
    R.run_deck(str(w / deck), "fp64", files=True, averages=False)
    for f in files:
        tot, same = _compare(str(w / f), os.path.join(src, "cm3d", f))
        print(f"  {f}: {same}/{tot} numbers identical as printed", flush=True)
