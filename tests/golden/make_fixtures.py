"""Write the golden fixtures of the parity tests (run here, where the reference is mounted).

1. kat_logs.json — known-answer values printed by the reference itself in its archived run
   logs (exp/old_2020/log/{time,old}/*.log, printed by precomp.h:39-45,59-67 and
   main.c:238): for every grid size the values at ind(bord+1,bord+1,bord+1) of
   ch1dxx..ch1dyz, vsv, vpz, v2pz, v2sz, v2px, v2pn and the recommended dt.
2. ref_<form>_<variant>_n<nx>_a<absorb>_s<steps>.npz — pc, qc, pp, qp after the reference
   C/OpenMP program (oracle/_ref/libref_<variant>.so, built from the unmodified
   sources by oracle/Makefile) ran `FORM nx nx nx absorb 12.5 12.5 12.5 0.001 tmax`.

Usage: python tests/golden/make_fixtures.py   (needs oracle/_ref; rebuild with `make -C oracle ref`)
"""
import glob
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

LOGS = "/root/reference/exp/old_2020/log"

# (form, variant, nx, absorb, steps)
RUNS = [(f, v, 8, 4, s) for f in ("ISO", "VTI", "TTI") for v in ("original", "der1der1") for s in (10, 30)]
RUNS += [("TTI", "original", 24, 16, 10), ("TTI", "der1der1lm", 8, 4, 30)]


def fixture_name(form, variant, nx, absorb, steps):
    return "ref_%s_%s_n%d_a%d_s%d.npz" % (form, variant, nx, absorb, steps)


def parse_logs():
    rx = {
        "grid": re.compile(r"Grid size is \((\d+),(\d+),(\d+)\)"),
        "ch": re.compile(r"ch1dxx=([-\d.e+]+); ch1dyy=([-\d.e+]+); ch1dzz=([-\d.e+]+); "
                         r"ch1dxy=([-\d.e+]+); ch1dxz=([-\d.e+]+); ch1dyz=([-\d.e+]+)"),
        "v1": re.compile(r"vsv=([-\d.e+]+); vpz=([-\d.e+]+), v2pz=([-\d.e+]+)"),
        "v2": re.compile(r"v2sz=([-\d.e+]+); v2pz=([-\d.e+]+), v2px=([-\d.e+]+), v2pn=([-\d.e+]+)"),
        "dt": re.compile(r"Recomended maximum time step is ([\d.]+)"),
        "abs": re.compile(r"extended by (\d+) absortion points and (\d+) border"),
    }
    table = {}
    for path in sorted(glob.glob(os.path.join(LOGS, "*", "*.log"))):
        cur = {}
        with open(path, errors="replace") as fh:
            lines = fh.readlines()
        for line in lines:
            m = rx["grid"].search(line)
            if m:
                cur = {"nx": int(m.group(1)), "ny": int(m.group(2)), "nz": int(m.group(3))}
                continue
            for key in ("abs", "dt", "ch", "v1", "v2"):
                m = rx[key].search(line)
                if m and cur:
                    cur[key] = [float(g) for g in m.groups()]
            if all(k in cur for k in ("abs", "ch", "v1", "v2", "dt")):
                key = "%d_%d" % (cur["nx"], int(cur["abs"][0]))
                entry = {"nx": cur["nx"], "absorb": int(cur["abs"][0]), "recdt": cur["dt"][0],
                         "ch1dxx": cur["ch"][0], "ch1dyy": cur["ch"][1], "ch1dzz": cur["ch"][2],
                         "ch1dxy": cur["ch"][3], "ch1dxz": cur["ch"][4], "ch1dyz": cur["ch"][5],
                         "vsv": cur["v1"][0], "vpz": cur["v1"][1], "v2pz": cur["v1"][2],
                         "v2sz": cur["v2"][0], "v2px": cur["v2"][2], "v2pn": cur["v2"][3]}
                table.setdefault(key, [])
                if entry not in table[key]:
                    table[key].append(entry)
                cur = {}
    return table


def main():
    kat = parse_logs()
    with open(os.path.join(HERE, "kat_logs.json"), "w") as f:
        json.dump({"source": "reference exp/old_2020/log/*/*.log (TTI runs, dx=12.5)",
                   "sizes": kat}, f, indent=1, sort_keys=True)
    print("kat sizes:", sorted(kat, key=lambda k: int(k.split("_")[0])))

    import oracle as O
    for form, variant, nx, absorb, steps in RUNS:
        R = O.Ref(variant)
        tmax = steps * 0.001
        cap = R.main([form, nx, nx, nx, absorb, 12.5, 12.5, 12.5, 0.001, "%.3f" % tmax])
        assert cap["st"] == steps, (cap["st"], steps)
        keep = ("pc", "qc", "pp", "qp") if nx <= 8 else ("pc", "qc")  # keep the larger case small
        out = {k: cap[k] for k in keep}
        out["isource"] = np.int64(cap["isource"])
        name = fixture_name(form, variant, nx, absorb, steps)
        np.savez_compressed(os.path.join(HERE, name), **out)
        print(name, "max|pc| = %.4e" % np.abs(cap["pc"]).max())


if __name__ == "__main__":
    main()
