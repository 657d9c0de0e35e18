"""GPU parity at the benchmarked sizes and in the benchmarked forms.

  - Ns=16, U=2.5, sector 8/8 (dim 1.66e8): FP64 and FP32 H*v against the
    oracle on row blocks (first and last columns, whole columns in between,
    and the row tails past 12*1024 of many columns: the 13th row of each lane
    of the column kernel), bars 1e-13 and 2e-6 of max|y_ref|; the Lanczos
    epilogue of the same kernel (first alpha/beta) against the plain product.
  - Workload 4 (Ns=12, all sectors, beta=50) in its batched + mixed form
    against the serial FP64 solve: |dE0| <= 1e-10 relative, max|dSigma(iw)|
    <= 1e-6; dn, dd reported.
  - Workload 5: 500-step mixed-precision chains against FP64 chains from the
    same ground state, max|dSigma(iw)| <= 1e-6.
  - Workload 3: the particle chain of orbital 1 past its first steps, by the
    Krylov moments <v|H^k|v>, k <= 39, which 20 Lanczos steps fix exactly;
    the oracle computes them with 20 of its own H*v products.
"""
import os
import sys

import numpy as np
import pytest

import dmed
import edref
from dmed import configs

pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    edref.build()
    dmed.init(0)
    yield


def set_cfg(c):
    args = (c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"], c["ust"],
            c["jh"], 0.0, 0.0, c["xmu"], True, False)
    edref.set_model(*args)
    return dmed.set_model(*args)


def test_ns16_hxv_vs_oracle_row_blocks():
    c = configs.config(5)
    set_cfg(c)
    say("[config 5] sector 8,8: GPU H*v")
    s = dmed.Sector(8, 8)
    dup, ddw = s.dim_up, s.dim_dw
    assert (dup, ddw) == (12870, 12870)
    rng = np.random.default_rng(55)
    x = rng.standard_normal(s.dim)
    y = s.hxv(x)
    y32 = s.hxv(x, prec=dmed.FP32)
    # the chain's first alpha/beta come from the kernel's Lanczos epilogue
    a, b, n = dmed.chain(s, s.vec(x), 2)
    s.free()
    nx2 = float(np.dot(x, x))
    a0 = float(np.dot(x, y)) / nx2
    r = y / np.sqrt(nx2) - a0 * x / np.sqrt(nx2)
    assert abs(a[0] - a0) <= 1e-12 * max(1.0, abs(a0))
    assert abs(b[1] - np.linalg.norm(r)) <= 1e-11 * np.linalg.norm(r)
    say("[config 5] oracle row blocks")
    m = edref.build_sector(8, 8)
    blocks = [(c_ * dup, (c_ + 1) * dup) for c_ in (0, 1, 2, 4095, 6434, 6435, 9999, ddw - 2, ddw - 1)]
    blocks += [(c_ * dup + 12 * 1024, (c_ + 1) * dup) for c_ in range(3, ddw - 2, 37)]
    xc = x.astype(np.complex128)
    del x
    err64 = err32 = scale = 0.0
    for first, last in blocks:
        yr = edref.hxv_block(m, xc, first, last, nthreads=NTH)
        assert np.abs(yr.imag).max() == 0.0
        yr = yr.real
        scale = max(scale, np.abs(yr).max())
        err64 = max(err64, np.abs(y[first:last] - yr).max())
        err32 = max(err32, np.abs(y32[first:last] - yr).max())
    nrows = sum(l - f for f, l in blocks)
    say(f"[config 5] {len(blocks)} blocks, {nrows} rows: FP64 err {err64 / scale:.2e}, "
        f"FP32 err {err32 / scale:.2e} (relative to max|y|)")
    assert err64 <= 1e-13 * scale
    assert err32 <= 2e-6 * scale


def _solve4(extra):
    c = configs.config(4)
    dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"],
                   c["ust"], c["jh"], 0.0, 0.0, c["xmu"], True, False)
    kw = dict(beta=c["beta"], lmats=c["lmats"], lreal=200, lanc_ngfiter=200,
              lanc_nstates_sector=4, lanc_nstates_total=16)
    return dmed.solve(1, 1, 12, dmed.default_params(**kw, **extra))


def test_workload4_batched_mixed_vs_serial_fp64():
    say("[config 4] serial FP64")
    r64 = _solve4(dict(batched=0, prec_gf=dmed.FP64))
    say("[config 4] batched + mixed")
    rmx = _solve4(dict(batched=1, prec_gf=dmed.MIXED))
    assert r64["nstates"] == rmx["nstates"] == 16
    de = abs(rmx["egs"] - r64["egs"]) / abs(r64["egs"])
    dn = float(np.abs(rmx["dens"] - r64["dens"]).max())
    dd = float(np.abs(rmx["docc"] - r64["docc"]).max())
    ds = float(np.abs(rmx["smats"] - r64["smats"]).max())
    dg = float(np.abs(rmx["gmats"] - r64["gmats"]).max())
    say(f"[config 4] dE0 {de:.2e} dn {dn:.2e} dd {dd:.2e} dG {dg:.2e} dSigma {ds:.2e}")
    assert de <= 1e-10
    assert ds <= 1e-6
    assert dn <= 1e-9 and dd <= 1e-9


def test_workload5_mixed_chains_vs_fp64_chains():
    c = configs.config(5)
    set_cfg(c)
    s, sp, sh = dmed.Sector(8, 8), dmed.Sector(9, 8), dmed.Sector(7, 8)
    say("[config 5] ground state")
    e0, psi, r0, rc = s.ground_state(prec=dmed.FP64, tol=1e-12, maxit=6000)
    assert rc == 0 and r0 <= 1e-12 * max(1.0, abs(e0))
    wm = np.pi / c["beta"] * (2 * np.arange(c["lmats"]) + 1)
    sig = {}
    for prec in (dmed.FP64, dmed.MIXED):
        say(f"[config 5] 500-step chains, prec {prec}")
        g = np.zeros(c["lmats"], complex)
        gr = np.zeros(1, complex)
        for dst, dag, sgn in ((sp, 1, 1), (sh, 0, -1)):
            v, n2 = dmed.apply_op(psi, 1, dag, dst)
            a, b, n = dmed.chain(dst, v, c["nlanc"], prec=prec)
            assert n == c["nlanc"]
            dmed.gf_add_chain(n2, e0, a, b, sgn, wm, np.zeros(1), 0.01, g, gr)
            v.free()
        sig[prec] = (g, dmed.sigma(0, 0, 1j * wm, g))
    psi.free()
    for v in (s, sp, sh):
        v.free()
    dg = np.abs(sig[dmed.MIXED][0] - sig[dmed.FP64][0]).max()
    ds = np.abs(sig[dmed.MIXED][1] - sig[dmed.FP64][1]).max()
    say(f"[config 5] mixed vs FP64 chains: max|dG| {dg:.2e}, max|dSigma| {ds:.2e}")
    assert ds <= 1e-6


def test_config3_particle_chain_krylov_moments():
    c = configs.config(3)
    set_cfg(c)
    nu, nd = c["sector"]
    s = dmed.Sector(nu, nd)
    say("[config 3] GS")
    ev, vecs, res, rc = s.eigh(1, tol=1e-12, dense_threshold=0)
    assert rc == 0
    t = dmed.Sector(nu + 1, nd)
    v, n2 = dmed.apply_op(vecs[0], 1, True, t)
    x = v.download()
    nsteps = 20
    say("[config 3] GPU chain")
    a, b, n = dmed.chain(t, v, nsteps)
    assert n == nsteps
    T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    # moments of the unit start vector, scaled by sc^k to stay O(1)
    sc = np.abs(a).max() + 2 * np.abs(b).max()
    Ts = T / sc
    mu_gpu = np.zeros(2 * nsteps + 1)
    e0v = np.zeros(nsteps)
    e0v[0] = 1.0
    w = e0v.copy()
    for k in range(2 * nsteps + 1):
        mu_gpu[k] = w[0]
        w = Ts @ w
    say("[config 3] oracle moments")
    mJ = edref.build_sector(nu + 1, nd)
    q = x / np.linalg.norm(x)
    mu_ref = np.zeros(2 * nsteps + 1)
    p = q.astype(np.complex128)
    for j in range(nsteps + 1):
        mu_ref[2 * j] = np.vdot(p, p).real
        if j == nsteps:
            break
        hp = edref.hxv(mJ, p, nthreads=NTH) / sc
        mu_ref[2 * j + 1] = np.vdot(p, hp).real
        p = hp
    # an n-step tridiagonal matrix fixes the moments k <= 2n - 1
    kk = 2 * nsteps
    err = np.abs(mu_gpu[:kk] - mu_ref[:kk]) / np.maximum(np.abs(mu_ref[:kk]), 1e-3)
    say(f"[config 3] moments k<={kk - 1}: max relative error {err.max():.2e}")
    assert err.max() <= 1e-9
