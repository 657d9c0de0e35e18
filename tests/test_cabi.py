"""CPU-side checks of the C ABI: the library loads, exports every entry point
declared in include/dmed.h, and its host-only paths behave (no GPU work)."""
import math
import os
import re

import numpy as np
import pytest

import dmed
from dmed import configs
from conftest import ROOT


def header_functions():
    txt = open(os.path.join(ROOT, "include", "dmed.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dmed_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    names = header_functions()
    bound = sorted(n for n, _, _ in dmed.SIGNATURES)
    assert names == bound


def test_library_exports_every_header_symbol():
    L = dmed.lib()
    for name in header_functions():
        assert hasattr(L, name), name


def test_calls_before_init_fail_cleanly():
    L = dmed.lib()
    m = dmed.Model()
    rc = L.dmed_set_model(m)
    assert rc == -2
    assert b"dmed_init" in L.dmed_last_error()


def test_binomial_reference_rounding():
    L = dmed.lib()
    for n in range(0, 31):
        for k in range(-1, n + 1):
            assert L.dmed_binomial(n, k) == (math.comb(n, k) if k >= 0 else 0)


def test_splitmix64_known_values():
    # first outputs of SplitMix64 seeded with 0 (Vigna's reference sequence)
    r = configs.SplitMix64(0)
    assert r.next_u64() == 0xE220A8397B1DCDAF
    assert r.next_u64() == 0x6E789E6AA1B965F4
    assert r.next_u64() == 0x06C45D188009454F


def test_configs_shapes_and_ranges():
    for n in range(1, 6):
        c = configs.config(n)
        nb = c["nspin"] * c["norb"] * c["nbath"]
        assert c["bath"].shape == (2 * nb,)
        e, v = c["bath"][:nb], c["bath"][nb:]
        if n == 1:
            assert np.all(np.abs(e) <= 1.0) and np.all(np.abs(v - 0.5) < 0.5)
        if n in (2, 4):
            assert np.allclose(np.sort(e), -np.sort(e)[::-1])   # symmetric levels
        if n in (3, 5):
            assert np.all(np.abs(e) <= 2.0) and np.all((v >= 0.1) & (v <= 0.5))


def test_restart_file_roundtrip(tmp_path):
    c = configs.config(3)
    p = tmp_path / "hamiltonian.restart"
    configs.write_restart(p, c)
    lines = p.read_text().splitlines()
    assert lines[0].startswith(" ") or lines[0].startswith("#")
    rows = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    assert rows.shape == (c["nbath"], 2 * c["norb"] * c["nspin"])
    nb = c["nbath"]
    st = c["norb"] * nb
    for k in range(nb):
        for a in range(c["norb"]):
            assert np.isclose(rows[k, 2 * a], c["bath"][k + a * nb], atol=1e-11)
            assert np.isclose(rows[k, 2 * a + 1], c["bath"][st + k + a * nb], atol=1e-11)
