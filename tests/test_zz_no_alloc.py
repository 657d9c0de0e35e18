"""GPU: dmed_chain_noalloc, the Lanczos chain that allocates no device memory,
against dmed_chain through the C ABI.

The no-alloc form runs the same kernels on vectors and partial sums in the
library's static device workspace, with the autotuner's timing runs and the
step graphs switched off; plain-mode shapes and per-column Lanczos partial
sums make neither matter, so alpha, beta and the step count must equal
dmed_chain's bit for bit (the parity bar for these integer-indexed kernels).
Two calls in a row, the second on a smaller sector (the workspace keeps the
first call's contents), in FP64 and in the mixed form; a sector above
DMED_NOALLOC_MAX_DIM is refused with DMED_EINVAL and a message.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import dmed
from dmed import configs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def set_cfg(n):
    c = configs.config(n)
    dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"], c["ust"],
                   c["jh"], 0.0, 0.0, c["xmu"], True, False)
    return c


def noalloc_chain(sector, v0, nlanc, prec, threshold=1e-13):
    a = np.zeros(nlanc)
    b = np.zeros(nlanc)
    ns = C.c_int32()
    rc = dmed.lib().dmed_chain_noalloc(sector.h, v0.h, nlanc, prec, threshold,
                                      a.ctypes.data_as(C.POINTER(C.c_double)),
                                      b.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ns))
    return rc, a, b, ns.value


def test_noalloc_chain_matches_dmed_chain():
    dmed.init(0)
    cases = [(2, (7, 6), 120), (1, (5, 4), 60)]      # dims 731,808 then 3,920
    for cfg, (nu, nd), nlanc in cases:
        set_cfg(cfg)
        s = dmed.Sector(nu, nd)
        assert s.dim <= 2097152
        x = np.random.default_rng(100 + cfg).standard_normal(s.dim)
        v = s.vec(x)
        for prec in (dmed.FP64, dmed.MIXED):
            say(f"[no-alloc] config {cfg} sector {nu},{nd} dim {s.dim} prec {prec}")
            a0, b0, n0 = dmed.chain(s, v, nlanc, prec=prec)
            rc, a1, b1, n1 = noalloc_chain(s, v, nlanc, prec)
            assert rc == 0, dmed.lib().dmed_last_error()
            assert n1 == n0
            assert np.array_equal(a1, a0)
            assert np.array_equal(b1, b0)
        v.free()
        s.free()


def test_noalloc_chain_refuses_large_sectors():
    dmed.init(0)
    set_cfg(3)
    s = dmed.Sector(7, 7)                            # dim 11,778,624
    assert s.dim > 2097152
    v = dmed.Vec(s, dmed.FP64)
    say(f"[no-alloc] refusal at dim {s.dim}")
    rc, _, _, _ = noalloc_chain(s, v, 10, dmed.FP64)
    assert rc == -1                                  # DMED_EINVAL
    assert b"DMED_NOALLOC_MAX_DIM" in dmed.lib().dmed_last_error()
    v.free()
    s.free()
