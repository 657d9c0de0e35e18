"""GPU: the two-half form of the column H*v kernel gives the one-pass result.

Where a column fits twice in LDS but the one-pass kernel's registers allow
only one 1024-lane workgroup per CU, the fixed launch rule runs the form in
which each lane takes its rows in two halves, two workgroups per CU.  The
terms of every row are summed in the same order, so y and the Lanczos
alpha/beta must equal those of the one-pass form (DMED_COL_HALVES=0) bit for
bit.  Sectors: Ns=16 8/8 (D_up 12,870; FP32 columns of 51 KB) and Ns=16 6/8
(D_up 8,008; FP64 columns of 63 KB), plain H*v and 40-step chains in FP32,
FP64 and the mixed form.  Each setting runs in a child process (the switch
is read once per process), with the autotuner off so that only the form
differs; results are compared through SHA-256 digests of their bytes.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "dmft-ed_amd")

CHILD = r'''
import hashlib, json, sys
import numpy as np
sys.path.insert(0, {pkg!r})
import dmed
from dmed import configs
dmed.init(0)
c = configs.config(5)
dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"], c["ust"],
               c["jh"], 0.0, 0.0, c["xmu"], True, False)
h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
out = {{}}
for (nu, nd), precs in (((8, 8), ("fp32",)), ((6, 8), ("fp64", "fp32"))):
    s = dmed.Sector(nu, nd)
    x = np.random.default_rng(nu * 10 + nd).standard_normal(s.dim)
    for nm in precs:
        prec = dmed.FP64 if nm == "fp64" else dmed.FP32
        key = f"{{nu}}/{{nd}} {{nm}}"
        out[key + " shape"] = s.hxv_shape(prec)
        out[key + " y"] = h(s.hxv(x, prec=prec))
        cprec = prec if nm == "fp64" else dmed.MIXED
        a, b, n = dmed.chain(s, s.vec(x), 40, prec=cprec)
        out[key + " chain"] = h(a) + h(b) + str(n)
        print("[halves child]", key, out[key + " shape"], flush=True)
    s.free()
print("RESULT " + json.dumps(out), flush=True)
'''


def _child(tmp, tag, env_extra):
    script = os.path.join(tmp, "child.py")
    with open(script, "w") as f:
        f.write(CHILD.format(pkg=PKG))
    env = dict(os.environ, DMED_TUNE="0", **env_extra)
    print(f"[halves] child {tag} {env_extra}", flush=True)
    r = subprocess.run([sys.executable, "-u", script], env=env, capture_output=True, text=True,
                       timeout=280)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    return json.loads(line[len("RESULT "):])


def test_two_half_column_kernel_matches_one_pass(tmp_path):
    one = _child(str(tmp_path), "one-pass", {"DMED_COL_HALVES": "0"})
    two = _child(str(tmp_path), "default", {})
    halves_used = [k for k in two if k.endswith("shape") and two[k]["row_passes"] == 2]
    print("[halves] two-half form on:", halves_used, flush=True)
    assert "8/8 fp32 shape" in halves_used
    for k in one:
        if k.endswith("shape"):
            assert one[k]["row_passes"] == 1, k
            continue
        assert one[k] == two[k], k
