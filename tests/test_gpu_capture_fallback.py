"""GPU: a Lanczos chain whose step-graph capture fails falls back to eager
steps from the right vectors.

Chains of 30 steps or more record three steps into a HIP graph and replay
it.  If the capture fails part-way, nothing recorded has run, so the eager
loop has to resume from the buffers as they were before the capture.
DMED_TEST_CAPTURE_FAIL=q forces step q of the captured window to fail before
it is recorded; alpha/beta must then equal the all-eager chain
(DMED_GRAPHS=0) bit for bit, for q = 0, 1, 2, in FP64 and in the mixed form.
Each setting runs in a child process, since both are read once per process.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "dmft-ed_amd")

CHILD = r'''
import sys
import numpy as np
sys.path.insert(0, {pkg!r})
import dmed
from dmed import configs
dmed.init(0)
c = configs.config(2)
dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"], c["ust"],
               c["jh"], 0.0, 0.0, c["xmu"], True, False)
s = dmed.Sector(*c["sector"])
x = np.random.default_rng(21).standard_normal(s.dim)
out = {{}}
for name, prec in (("fp64", dmed.FP64), ("mixed", dmed.MIXED)):
    a, b, n = dmed.chain(s, s.vec(x), 64, prec=prec)
    out["a_" + name], out["b_" + name] = a, b
np.savez({out!r}, **out)
print("child ok", flush=True)
'''


def _child(tmp, tag, env_extra):
    script = os.path.join(tmp, "child.py")
    out = os.path.join(tmp, f"{tag}.npz")
    with open(script, "w") as f:
        f.write(CHILD.format(pkg=PKG, out=out))
    env = dict(os.environ, **env_extra)
    print(f"[capture] child {tag} {env_extra}", flush=True)
    r = subprocess.run([sys.executable, "-u", script], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return dict(np.load(out)), r.stderr


def test_failed_capture_resumes_from_the_right_vectors(tmp_path):
    ref, _ = _child(str(tmp_path), "eager", {"DMED_GRAPHS": "0"})
    graphed, _ = _child(str(tmp_path), "graphs", {})
    for k in ref:
        assert np.array_equal(graphed[k], ref[k]), k
    for q in (0, 1, 2):
        got, err = _child(str(tmp_path), f"fail{q}", {"DMED_TEST_CAPTURE_FAIL": str(q)})
        assert "capture unavailable" in err
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (q, k)
