"""The `fletcher` command-line front-end on the device with FL_RSF set, against the snapshot
file that the reference der1der1 front-end (its OpenMP build in oracle/_ref) writes for the
same invocation: pc at the start and after every 0.01 s of simulated time
(der1der1/main.c:273-300, model.c:97-102) in the layout of utils.c:71-205.  The header text
must be the same bytes; the binary must hold the same number of frames, the first (the zero
initial field) bitwise equal and each later frame within relative L2 1e-5 of the reference's.
Report.csv (model.c:135-158) has the reference's lines and keys.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

ARGS = ["TTI", 24, 20, 16, 8, 12.5, 12.5, 12.5, 0.001, 0.025]


@pytest.fixture(scope="module")
def ref_dump(oracle, tmp_path_factory):
    if not oracle.ref_available("der1der1"):
        pytest.skip("reference build oracle/_ref absent")
    d = tmp_path_factory.mktemp("ref")
    cap = oracle.Ref("der1der1").main(ARGS, workdir=str(d))
    return cap, d


@pytest.mark.parametrize("variant", ["fused", "twopass"])
def test_cli_rsf_snapshots_match_reference_dump(fl, ref_dump, variant, tmp_path):
    cap, ref_dir = ref_dump
    env = dict(os.environ, FL_RSF="TTI", FL_VARIANT=variant, FL_PRECISION="fp32")
    print("cli: %s %s" % (variant, " ".join(map(str, ARGS))), flush=True)
    r = subprocess.run([fl.CLI_PATH] + [str(a) for a in ARGS], cwd=str(tmp_path), env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "MSamples/s" in r.stdout
    head_ref = (ref_dir / "TTI.rsf").read_bytes()
    assert (tmp_path / "TTI.rsf").read_bytes() == head_ref
    sz, sy, sx = cap["shape"]
    ref = np.fromfile(str(ref_dir / "TTI.rsf@"), np.float32).reshape(-1, sz, sy, sx)
    got = np.fromfile(str(tmp_path / "TTI.rsf@"), np.float32).reshape(-1, sz, sy, sx)
    assert ref.shape[0] == got.shape[0] == 3  # t = 0, after steps 10 and 20
    assert not ref[0].any() and not got[0].any()
    for k in range(1, ref.shape[0]):
        assert np.abs(ref[k]).max() > 0
        err = rel_l2(got[k], ref[k])
        print("%s frame %d: relL2 vs reference %.3e" % (variant, k, err))
        assert err <= 1e-5
    # Report.csv (model.c:135-158): the same problem-size line, the same keys on the metrics line
    mine = (tmp_path / "Report.csv").read_text().splitlines()
    theirs = (ref_dir / "Report.csv").read_text().splitlines()
    assert mine[0] == theirs[0]
    keys = lambda line: [f.strip() for f in line.split(";")][0::2]
    assert keys(mine[1]) == keys(theirs[1])
    assert float(mine[1].split(";")[3]) > 0  # MSamples/s
    assert "Memory High Water Mark is" in r.stdout
