"""GPU: the launch-shape autotuner (on by default for sectors of at least
DMED_TUNE_MIN_DIM states; DMED_TUNE_FILE keeps the choices).

The tuner only chooses among launch shapes that leave each element's sum
unchanged -- rows per lane (workgroup lanes) and workgroups per CU of the
column kernel, LDS tile and lanes of the tile kernel -- so H*v with tuned
shapes must equal the fixed-rule H*v bit for bit, in FP64 and FP32.  The
column kernel's Lanczos-mode partial sums are one per column and the tile
kernel keeps its fixed rule in Lanczos mode, so chain alpha/beta are
bit-identical as well.  Checked in a child process that tunes every sector
(DMED_TUNE_MIN_DIM=0), since the settings are read once per process; this
process keeps the fixed rules (both sectors are below the default minimum).
Shapes: workload 1 (D_up = 70, tile kernel) and workload 2 (D_up = 924,
column kernel).  A second child loads the recorded choices instead of timing
again.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import dmed
from dmed import configs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "dmft-ed_amd")

CHILD = r'''
import sys
import numpy as np
sys.path.insert(0, {pkg!r})
import dmed
from dmed import configs
out = {{}}
dmed.init(0)
for cfg in (1, 2):
    c = configs.config(cfg)
    dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"], c["ust"],
                   c["jh"], 0.0, 0.0, c["xmu"], True, False)
    s = dmed.Sector(*c["sector"])
    x = np.load({xdir!r} + f"/x{{cfg}}.npy")
    out[f"y{{cfg}}"] = s.hxv(x)
    out[f"z{{cfg}}"] = s.hxv(x, prec=dmed.FP32)
    sh = s.hxv_shape()
    out[f"shape{{cfg}}"] = np.array([sh["lanes"], sh["groups_per_cu"], sh["columns_per_group"]])
    a, b, n = dmed.chain(s, s.vec(x), 20)
    out[f"a{{cfg}}"], out[f"b{{cfg}}"] = a, b
np.savez({xdir!r} + "/child.npz", **out)
print("child ok", flush=True)
'''


def _child(tmp, tune_file):
    script = os.path.join(tmp, "child.py")
    with open(script, "w") as f:
        f.write(CHILD.format(pkg=PKG, xdir=tmp))
    env = dict(os.environ, DMED_TUNE="1", DMED_TUNE_MIN_DIM="0", DMED_TUNE_FILE=tune_file)
    r = subprocess.run([sys.executable, "-u", script], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return dict(np.load(os.path.join(tmp, "child.npz")))


def test_tuned_launch_shapes_give_identical_hxv(tmp_path):
    dmed.init(0)
    ref = {}
    for cfg in (1, 2):
        c = configs.config(cfg)
        dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"], c["ust"],
                       c["jh"], 0.0, 0.0, c["xmu"], True, False)
        s = dmed.Sector(*c["sector"])
        x = np.random.default_rng(cfg).standard_normal(s.dim)
        np.save(tmp_path / f"x{cfg}.npy", x)
        ref[f"y{cfg}"] = s.hxv(x)
        ref[f"z{cfg}"] = s.hxv(x, prec=dmed.FP32)
        a, b, n = dmed.chain(s, s.vec(x), 20)
        ref[f"a{cfg}"], ref[f"b{cfg}"] = a, b
    tune_file = str(tmp_path / "tune.txt")
    for run in (1, 2):
        print(f"[autotune] child run {run}", flush=True)
        got = _child(str(tmp_path), tune_file)
        for cfg in (1, 2):
            assert np.array_equal(got[f"y{cfg}"], ref[f"y{cfg}"])
            assert np.array_equal(got[f"z{cfg}"], ref[f"z{cfg}"])
            assert np.array_equal(got[f"a{cfg}"], ref[f"a{cfg}"])
            assert np.array_equal(got[f"b{cfg}"], ref[f"b{cfg}"])
        print(f"[autotune] tuned shapes (lanes, per CU, columns): {got['shape1']}, "
              f"{got['shape2']}", flush=True)
        lines = open(tune_file).read().split("\n")
        lines = [l for l in lines if l.strip()]
        # one line per (shape, precision); the second run only reads them
        assert len(lines) == 4, lines
