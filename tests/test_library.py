"""The C-ABI library loads, exports every symbol include/cm3d_nb.h declares,
and its host set-up helpers rebuild cm3d's own tables.  CPU only (no kernel
is launched)."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_fixture
from cm3d_amd import nb, systems as S
from oracle import oracle as O


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "cm3d_nb.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(nb_\w+)\s*\(", hdr, re.M))
    assert declared >= {"nb_create", "nb_force", "nb_build_lists"}
    L = nb.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(nb.EXPORTS)


@pytest.mark.parametrize("name", ["ar864", "water512", "mix600"])
def test_tables_equal_cm3d_setup(name):
    """nb_table_pair / nb_table_coul / nb_kvectors rebuild, bit for bit, what
    cm3d's input layer built from the equivalent deck (the fixture)."""
    d = load_fixture(name)
    builder = {"ar864": lambda: S.lj_fcc(), "water512": lambda: S.water(nside=8),
               "mix600": lambda: S.mixture(600)}[name]
    s = builder()
    for k in ("vtab", "dvtab", "switchs", "rmins_cut2", "rmaxs_cut2", "rminl_cut2",
              "rmaxl_cut2", "rmaxs_skin2", "rmaxl_skin2", "dx2tab", "itype", "map",
              "nexclude", "exclude"):
        assert np.array_equal(np.asarray(s[k]).ravel(), d[k].ravel()), k
    assert np.array_equal(s["qch"], d["qch"])
    assert s["vlrc"] == d["vlrc"][0] and s["wlrc"] == d["wlrc"][0]
    if int(d["ktot"][0]) > 0:
        for k in ("coultab", "dcoultab", "switchse"):
            assert np.array_equal(s[k], d[k]), k
        for k in ("dx2coul", "rmins_coul2", "rmaxl_coul2", "rmaxs_skin2_e", "rmaxl_skin2_e",
                  "ewald_self", "ewald_bkg"):
            assert float(s[k]) == float(d[k][0]), k
        for k in ("ka", "kb", "kc", "kup", "cnexclude", "cexclude", "iecorr", "jecorr"):
            ref = d.get(k, np.zeros(0, np.int32))
            assert np.array_equal(np.asarray(s[k]), ref), k


def test_tables_equal_oracle_restatement():
    """The product's and the oracle's restatements agree bit for bit (LJ and
    Williams pair tables, the erfc table, the k sphere)."""
    for pot, prm in ((nb.POT_LJ, [77.99954795, 3.166]),
                     (nb.POT_WILLIAMS, [42084530., 3.60, 285830.6, 0.0, 0.0])):
        a = nb.table_pair(pot, prm, 1.0, 9.0, 7.0, 0.5, 1.0, 800)
        b = O.table_pair(pot, prm, 1.0, 9.0, 7.0, 0.5, 1.0, 800)
        for k in ("vtab", "dvtab", "switchs", "cut2"):
            assert np.array_equal(a[k], b[k]), k
        assert a["dx2tab"] == b["dx2tab"] and a["vlrc"] == b["vlrc"] and a["wlrc"] == b["wlrc"]
    a = nb.table_coul(600, 3, 0.3, 0.5, 8.0, 12.0, 0.5, 1.0)
    b = O.table_coul(600, 3, 0.3, 0.5, 8.0, 12.0, 0.5, 1.0)
    for k in ("coultab", "dcoultab", "switchse", "cut2"):
        assert np.array_equal(a[k], b[k]), k
    for kmax in (0, 1, 7, 9, 15):
        ka, kb, kc = nb.kvectors(kmax)
        oa, ob, oc, _ = O.kvectors(kmax)
        assert np.array_equal(ka, oa) and np.array_equal(kb, ob) and np.array_equal(kc, oc)


def test_create_without_gpu_fails_loudly():
    """No device: nb_create returns an error (never a CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(nb.NBError):
        nb.NBContext(S.lj_fcc(ncells=3))


def test_header_is_plain_c(tmp_path):
    """cm3d is C: the header compiles as C99 with warnings as errors, every
    declared entry point links against the built library, and a program
    that calls nb_create without a GPU gets an error code and a message
    (no GPU is touched beyond that failing call)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    hdr = open(os.path.join(ROOT, "include", "cm3d_nb.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(nb_\w+)\s*\(", hdr, re.M)))
    src = tmp_path / "use_header.c"
    src.write_text(
        '#include <stdio.h>\n#include <string.h>\n#include "cm3d_nb.h"\n'
        "int main(void) {\n"
        "  void *fns[] = {" + ", ".join(f"(void *){n}" for n in names) + "};\n"
        "  nb_params p; memset(&p, 0, sizeof p);\n"
        "  nb_ctx *c = 0;\n"
        "  int rc = nb_create(&p, &c);\n"
        '  printf("%d %d %s\\n", (int)(sizeof fns / sizeof fns[0]), rc, rc ? nb_last_error() : "");\n'
        "  return rc == 0;\n}\n")
    exe = tmp_path / "use_header"
    libdir = os.path.join(ROOT, "cm3d_amd")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", libdir,
                        "-lcm3d_nb", f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and not os.path.exists("/dev/kfd"):
        out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
        n, rc, *msg = out.stdout.split(None, 2)
        assert int(n) == len(names) and int(rc) != 0 and msg, out.stdout


def test_integration_shim_compiles_against_cm3d_headers(tmp_path):
    """The binding INTEGRATION.md documents (md_nbgpu.c, the first C block of
    its section 2) compiles against cm3d's own md.h and this header, so the
    documented drop-in cannot drift from either side.  Skipped where the
    reference tree is absent."""
    import shutil
    import subprocess
    ref_inc = "/root/reference/include"
    if not os.path.isdir(ref_inc) or shutil.which("gcc") is None:
        pytest.skip("reference headers or gcc not available")
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = doc[doc.index("## 2."):]
    code = sec[sec.index("```c") + 4:]
    code = code[:code.index("```")]
    assert "nbgpu_setup" in code and "nb_create" in code
    src = tmp_path / "md_nbgpu.c"
    src.write_text(code)
    r = subprocess.run(["gcc", "-std=gnu99", "-fsyntax-only", "-DLINUX", "-I", ref_inc,
                        "-I", "/root/reference/minimize", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
