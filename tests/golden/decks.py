"""Write a cm3d deck (control file, set, topology, data_base, init file) for a
synthetic system dict from cm3d_amd.systems, in the formats cm3d's input layer
reads (input/md_readsim.c, input/md_readset.c, input/md_getmols.c,
input/md_readparm.c, input/md_readpos.c).  Used only to generate the golden
fixtures with the compiled reference.
"""
import math
import os

import numpy as np

from cm3d_amd import systems as S


def _init_file(path, sysd):
    n = int(sysd["natoms"])
    p = sysd["pos"].reshape(3, n)
    with open(path, "w") as f:
        f.write(f"{n} 1\n")
        for i in range(n):
            f.write("%.17g %.17g %.17g\n" % (p[0, i], p[1, i], p[2, i]))
        h = sysd["hmat"]
        for r in range(3):
            f.write("%.17g %.17g %.17g\n" % (h[3 * r], h[3 * r + 1], h[3 * r + 2]))


def write_deck(d, sysd, kind, nstep=1, dt=2.0, nres_intra=1, temp=300.0, seed=12345,
               extra=None, nhc=0, tau_nhc=1000.0, therm=None):
    """kind: 'lj' (one species, one atom), 'water' (3-site flexible water) or
    'mixture' (two one-atom species, species-major).  nhc > 0 makes it an
    NVT deck with an nhc-link Nose-Hoover chain (tau_nhc in fs); therm gives
    each species' mol_therm_opt ('global' or 'none', default all global
    under NVT).  Returns the deck path."""
    os.makedirs(d, exist_ok=True)
    name = "sys"
    ntable = int(sysd["ntable"])
    ens = "nvt" if nhc else "nve"
    th = list(therm) if therm else ["global" if nhc else "none"] * 2
    keys = [r"\calculation{md}", r"\istart{1}", rf"\ntime{{{nstep}}}", rf"\ensemble{{{ens}}}",
            r"\nres_lrf{1}", rf"\nres_intra{{{nres_intra}}}", r"\nres_tors{1}",
            rf"\len_nhc{{{nhc}}}", rf"\tau_nhc{{{tau_nhc!r}}}", r"\nbar{0}",
            rf"\t_ext{{{temp}}}", r"\p_ext{0.0}",
            rf"\dt{{{dt}}}", r"\iperd{3}", r"\iwr_screen{1000000}", r"\iwr_dump{1000000}",
            r"\iwr_inst{1000000}", rf"\dpi_file{{{name}.init}}", rf"\dpo_file{{{name}.save}}",
            rf"\ham_file{{{name}.ham}}", rf"\eng_file{{{name}.eng}}",
            rf"\eval_file{{{name}.eval}}", rf"\mol_set_file{{{name}.set}}",
            rf"\skin{{{float(sysd['skin'])!r}}}", rf"\rheal_res{{{float(sysd['rheal'])!r}}}",
            rf"\iseed{{{seed}}}", rf"\ntable{{{ntable}}}",
            r"\inter_file{data_base}", r"\bond_file{data_base}", r"\bend_file{data_base}",
            r"\torsion_file{data_base}", r"\onefour_file{data_base}"]
    if extra:
        keys += extra
    with open(os.path.join(d, f"{name}.input"), "w") as f:
        f.write("\n".join(keys) + "\n")
    _init_file(os.path.join(d, f"{name}.init"), sysd)
    qe = sysd["qch"] / S.SQRT_CCONV
    m = sysd["amass"] / S.MCONV
    with open(os.path.join(d, f"{name}.set"), "w") as fset, \
            open(os.path.join(d, "data_base"), "w") as fdb:
        if kind == "lj":
            n = int(sysd["natoms"])
            fset.write(rf"~mol_def[\mol_parm_file{{a.top}} \mol_therm_opt{{{th[0]}}} \nmol{{{n}}}\mol_index{{1}}]" + "\n")
            with open(os.path.join(d, "a.top"), "w") as t:
                t.write(r"~mol_name_def[\mol_name{atom}\natom{1}\nbond{0}]" + "\n")
                t.write(rf"~atom_def[\atom_typ{{a}}\atom_ind{{1}}\mass{{{float(m[0])!r}}} \charge{{{float(qe[0])!r}}}]" + "\n")
            types = ["a"]
        elif kind == "mixture":
            n = int(sysd["natoms"])
            h = n // 2
            fset.write(rf"~mol_def[\mol_parm_file{{a.top}} \mol_therm_opt{{{th[0]}}} \nmol{{{h}}}\mol_index{{1}}]" + "\n")
            fset.write(rf"~mol_def[\mol_parm_file{{b.top}} \mol_therm_opt{{{th[1]}}} \nmol{{{n - h}}}\mol_index{{2}}]" + "\n")
            for nm, idx in (("a", 0), ("b", h)):
                with open(os.path.join(d, f"{nm}.top"), "w") as t:
                    t.write(rf"~mol_name_def[\mol_name{{{nm}}}\natom{{1}}\nbond{{0}}]" + "\n")
                    t.write(rf"~atom_def[\atom_typ{{{nm}}}\atom_ind{{1}}\mass{{{float(m[idx])!r}}} \charge{{{float(qe[idx])!r}}}]" + "\n")
            types = ["a", "b"]
        elif kind == "water":
            nmol = int(sysd["natoms"]) // 3
            fset.write(rf"~mol_def[\mol_parm_file{{w.top}} \mol_therm_opt{{{th[0]}}} \nmol{{{nmol}}}\mol_index{{1}}]" + "\n")
            with open(os.path.join(d, "w.top"), "w") as t:
                t.write(r"~mol_name_def[\mol_name{water}\natom{3}\nbond{2}\nbend{1}]" + "\n")
                t.write(rf"~atom_def[\atom_typ{{ow}}\atom_ind{{1}}\mass{{{float(m[0])!r}}} \charge{{{float(qe[0])!r}}}]" + "\n")
                t.write(rf"~atom_def[\atom_typ{{hw}}\atom_ind{{2}}\mass{{{float(m[1])!r}}} \charge{{{float(qe[1])!r}}}]" + "\n")
                t.write(rf"~atom_def[\atom_typ{{hw}}\atom_ind{{3}}\mass{{{float(m[2])!r}}} \charge{{{float(qe[2])!r}}}]" + "\n")
                t.write(r"~bond_def[\atom1{2}\atom2{1}]" + "\n")
                t.write(r"~bond_def[\atom1{3}\atom2{1}]" + "\n")
                t.write(r"~bend_def[\atom1{2}\atom2{1}\atom3{3}]" + "\n")
            types = ["ow", "hw"]
            fdb.write(rf"~bond_parm[\atom1{{ow}}\atom2{{hw}}\pot_type{{harm}}\fk{{{float(sysd['bond_k'][0])!r}}}\eq{{{float(sysd['bond_r0'][0])!r}}}]" + "\n")
            fdb.write(rf"~bend_parm[\atom1{{hw}}\atom2{{ow}}\atom3{{hw}}\pot_type{{harm}}\fk{{{float(sysd['bend_fk'][0])!r}}}\eq{{{float(math.degrees(sysd['bend_th0'][0]))!r}}}]" + "\n")
        else:
            raise ValueError(kind)
        nt = len(types)
        mp = sysd["map"].reshape(nt, nt)
        for i in range(nt):
            for j in range(i, nt):
                k = mp[i, j]
                if k == -1:
                    fdb.write(rf"~inter_parm[\atom1{{{types[i]}}}\atom2{{{types[j]}}}\pot_type{{null}}]" + "\n")
                    continue
                p = sysd["pair_params"][k]
                fdb.write(rf"~inter_parm[\atom1{{{types[i]}}}\atom2{{{types[j]}}}\pot_type{{lennard-jones}}"
                          rf"\eps{{{float(p[0])!r}}}\sig{{{float(p[1])!r}}}\min_dist{{{float(p[2])!r}}}\max_dist{{{float(p[3])!r}}}\res_dist{{{float(p[4])!r}}}]" + "\n")
    return os.path.join(d, f"{name}.input")
