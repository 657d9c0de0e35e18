"""Pin the oracle (oracle/edref.c + oracle/edref.py) before trusting it.

Each check compares the oracle with something it does not share code with:
the independent Jordan-Wigner fixtures (tests/golden/make_golden.py), library
routines (math.comb, numpy.searchsorted, scipy eigh_tridiagonal) or closed
forms (Toeplitz spectrum, U=0 and V=0 limits, Krylov moments).
"""
import math
import os

import numpy as np
import pytest

import edref
from conftest import GOLDEN

NAMES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def bath_type(g):
    return str(g["bath_type"]) if "bath_type" in g else "normal"


def set_model_from(g):
    return edref.set_model(int(g["norb"]), int(g["nbath"]), int(g["nspin"]), g["bath"],
                           g["hloc_f"], g["uloc"], float(g["ust"]), float(g["jh"]),
                           float(g["jx"]), float(g["jp"]), float(g["xmu"]),
                           int(g["hfmode"]), int(g["jhflag"]), bath_type(g))


@pytest.fixture(scope="module", autouse=True)
def _build():
    edref.build()


def test_binomial_matches_math_comb():
    for n in range(0, 31):
        for k in range(-1, n + 1):
            assert edref.binomial(n, k) == (math.comb(n, k) if k >= 0 else 0)


@pytest.mark.parametrize("name", NAMES)
def test_sector_maps_and_hxv_match_jordan_wigner(name):
    g = load(name)
    ns = set_model_from(g)
    offs = g["sector_offsets"]
    k = 0
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            m = edref.build_sector(nu, nd)
            ref = g["maps"][offs[k]:offs[k + 1]]
            assert np.array_equal(m, ref), (nu, nd)
            x = g["x"][offs[k]:offs[k + 1]]
            y = edref.hxv(m, x)
            hx = g["hx"][offs[k]:offs[k + 1]]
            assert np.abs(y.imag).max(initial=0) == 0.0
            scale = max(1.0, np.abs(hx).max(initial=0))
            assert np.abs(y.real - hx).max(initial=0) <= 1e-13 * scale, (nu, nd)
            # complex input: H is real symmetric, so H(x + i x') = Hx + i Hx'
            if len(m) > 1:
                yc = edref.hxv(m, x + 1j * x[::-1])
                assert np.allclose(yc.imag, edref.hxv(m, x[::-1]).real, atol=1e-13 * scale)
            k += 1


def test_openmp_row_blocks_match_serial():
    g = load("m3_norb2_nb2_jh")
    set_model_from(g)
    m = edref.build_sector(3, 3)
    x = np.random.default_rng(1).standard_normal(len(m)) + 0j
    assert np.array_equal(edref.hxv(m, x, nthreads=3), edref.hxv(m, x))


def test_hxv_block_matches_full_product():
    """ref_hxv_block (rows [first, last) into their own array) against the
    rows of the full product, on blocks at the start, inside and at the end
    of a sector, for 1 and 3 threads."""
    g = load("m3_norb2_nb2_jh")
    set_model_from(g)
    m = edref.build_sector(3, 3)
    x = np.random.default_rng(2).standard_normal(len(m)) + 0j
    y = edref.hxv(m, x)
    n = len(m)
    for first, last in ((0, 7), (5, n - 3), (n - 11, n), (17, 17)):
        for t in (1, 3):
            assert np.array_equal(edref.hxv_block(m, x, first, last, nthreads=t), y[first:last])


def test_c_cdg_pauli_and_sign():
    # sign = (-1)^(occupied positions below pos); Pauli violations raise
    out, s = edref.op_c(3, 0b0111)
    assert out == 0b0011 and s == 1.0
    out, s = edref.op_c(2, 0b0111)
    assert out == 0b0101 and s == -1.0
    out, s = edref.op_c(4, 0b0111, dagger=True)
    assert out == 0b1111 and s == -1.0
    with pytest.raises(ValueError):
        edref.op_c(4, 0b0111)
    with pytest.raises(ValueError):
        edref.op_c(1, 0b0111, dagger=True)


def test_binary_search_matches_searchsorted():
    rng = np.random.default_rng(7)
    for n in (0, 1, 2, 3, 10, 257, 1000):
        a = np.unique(rng.integers(0, 10 * n + 5, size=n))
        for v in list(a) + [-1, 10 * n + 7]:
            j = edref.binary_search(a, int(v))
            pos = np.searchsorted(a, v)
            expect = pos + 1 if pos < len(a) and a[pos] == v else 0
            assert j == expect


def test_tql2_toeplitz_closed_form_and_vectors():
    for n in (1, 2, 5, 40, 200):
        a, b = 0.3, -0.7
        d = np.full(n, a)
        e = np.zeros(n)
        e[1:] = b
        w, z, ierr = edref.tql2(d, e)
        assert ierr == 0
        k = np.arange(1, n + 1)
        exact = np.sort(a + 2 * b * np.cos(k * np.pi / (n + 1)))
        assert np.allclose(w, exact, atol=1e-12)
        T = np.diag(d) + np.diag(e[1:], 1) + np.diag(e[1:], -1)
        assert np.allclose(T @ z, z * w, atol=1e-11)
        assert np.allclose(z.T @ z, np.eye(n), atol=1e-11)


def test_tql2_matches_scipy_random():
    from scipy.linalg import eigh_tridiagonal
    rng = np.random.default_rng(3)
    for n in (3, 17, 300):
        d = rng.standard_normal(n)
        off = rng.uniform(0.01, 1.0, n - 1)
        e = np.concatenate([[0.0], off])
        w, z, ierr = edref.tql2(d, e)
        ws, zs = eigh_tridiagonal(d, off)
        assert ierr == 0
        assert np.allclose(w, ws, atol=1e-12)
        # first components squared are the pole weights used by the GF
        assert np.allclose(z[0] ** 2, zs[0] ** 2, atol=1e-10)


def test_pythag():
    for a, b in ((3.0, 4.0), (0.0, 0.0), (1e-300, 1e-300), (1e200, 1e200), (-5.0, 12.0)):
        assert np.isclose(edref.lib().ref_pythag(a, b), math.hypot(a, b), rtol=1e-14)


def test_lanczos_chain_reproduces_krylov_moments():
    g = load("m4_norb3_nb1")
    set_model_from(g)
    m = edref.build_sector(3, 3)
    H = edref.dense_h(m)
    v = np.random.default_rng(11).standard_normal(len(m))
    v /= np.linalg.norm(v)
    n = 12
    a, b, steps = edref.lanc_tridiag(m, v, n)
    assert steps == n
    T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    for k in range(0, 2 * n - 2):
        mom_h = v @ np.linalg.matrix_power(H, k) @ v
        mom_t = np.linalg.matrix_power(T, k)[0, 0]
        assert np.isclose(mom_h, mom_t, rtol=1e-9, atol=1e-9 * np.abs(H).max() ** k)


def test_lanczos_chain_early_exit_on_invariant_subspace():
    g = load("m1_norb1_nb3")
    set_model_from(g)
    m = edref.build_sector(2, 2)
    H = edref.dense_h(m)
    w, U = np.linalg.eigh(H)
    v = U[:, 0] + U[:, 1]            # lives in a 2-dimensional invariant subspace
    a, b, steps = edref.lanc_tridiag(m, v / np.linalg.norm(v), 10)
    assert steps == 2
    assert np.allclose(np.sort(np.linalg.eigvalsh(np.diag(a[:2]) + np.diag(b[1:2], 1)
                                                   + np.diag(b[1:2], -1))), w[:2])
    assert np.all(a[2:] == 0) and np.all(b[3:] == 0)


@pytest.mark.parametrize("name", ["m1_norb1_nb3", "m3_norb2_nb2_jh", "m4_norb3_nb1"])
def test_plain_lanczos_ground_state(name):
    g = load(name)
    ns = set_model_from(g)
    offs, eoffs = g["sector_offsets"], g["eig_offsets"]
    k = 0
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            m = edref.build_sector(nu, nd)
            e0 = g["eigs"][eoffs[k]]
            if len(m) > 30:
                v0 = np.random.default_rng(k).standard_normal(len(m))
                eg, vec = edref.lanczos_gs(m, v0)
                assert abs(eg[0] - e0) <= 1e-10 * max(1.0, abs(e0))
                r = edref.hxv(m, vec) - eg[0] * vec
                assert np.linalg.norm(r) < 1e-6
            k += 1


def _solve_golden(g, **pkw):
    wm, wr = g["wm"], g["wr"]
    p = edref.default_params(beta=float(g["beta"]), lmats=len(wm), lreal=len(wr),
                             wini=float(wr[0]), wfin=float(wr[-1]), eps=float(g["eps"]),
                             **pkw)
    norb = int(g["norb"])
    return edref.ed_solve(norb, int(g["nbath"]), int(g["nspin"]), g["bath"], g["hloc_f"],
                          g["uloc"][:norb], float(g["ust"]), float(g["jh"]),
                          float(g["xmu"]), bool(g["hfmode"]), params=p, jx=float(g["jx"]),
                          jp=float(g["jp"]), jhflag=bool(g["jhflag"]), bath_type=bath_type(g))


@pytest.mark.parametrize("name", ["m1_norb1_nb3", "m2_norb2_nb1", "m3_norb2_nb2_jh", "m4_norb3_nb1",
                                  "m6_u0"])
def test_ed_solve_zero_temperature_matches_lehmann(name):
    g = load(name)
    r = _solve_golden(g, lanc_ngfiter=400)
    assert abs(r["egs"] - g["egs"]) <= 1e-10 * abs(g["egs"])
    assert len(r["states"]) == int(g["nstates"])
    assert np.allclose(r["dens"], g["dens"], atol=1e-10)
    assert np.allclose(r["docc"], g["docc"], atol=1e-10)
    nsp = int(g["nspin"])
    assert np.abs(r["gmats"] - g["gmats"][:nsp]).max() < 1e-8
    # Real axis: the plain (non-reorthogonalised) chain of the reference loses
    # orthogonality once it runs over a whole small sector, so poles of weight
    # ~1e-6 are replaced by ghosts; at eps=0.05 that moves G(w+i eps) by ~1e-4
    # while every Matsubara moment stays exact.
    assert np.abs(r["greal"] - g["greal"][:nsp]).max() < 5e-4


def test_ed_solve_finite_temperature_matches_lehmann():
    g = load("m5_norb1_nb5_T")
    r = _solve_golden(g, lanc_nstates_sector=400, lanc_nstates_total=216, lanc_ngfiter=400)
    assert abs(r["egs"] - g["egs"]) <= 1e-10 * abs(g["egs"])
    assert np.allclose(r["dens"], g["dens"], atol=1e-8)
    assert np.allclose(r["docc"], g["docc"], atol=1e-8)
    assert np.abs(r["gmats"] - g["gmats"]).max() < 1e-7


def test_u0_closed_form():
    """U=0: E0 = 2 * sum of negative single-particle levels and G = 1/(iw+mu-ed-Delta),
    so Sigma vanishes."""
    g = load("m6_u0")
    nb = int(g["nbath"])
    e, v = g["bath"][:nb], g["bath"][nb:]
    h1 = np.zeros((nb + 1, nb + 1))
    h1[0, 0] = -float(g["xmu"])
    h1[0, 1:] = h1[1:, 0] = v
    h1[1:, 1:] = np.diag(e)
    eps = np.linalg.eigvalsh(h1)
    assert np.isclose(2 * eps[eps < 0].sum(), g["egs"], atol=1e-12)
    r = _solve_golden(g, lanc_ngfiter=400)
    z = 1j * g["wm"]
    g0 = 1.0 / (z + float(g["xmu"]) - (v ** 2 / (z[:, None] - e)).sum(axis=1))
    assert np.abs(r["gmats"][0, 0] - g0).max() < 1e-9
    assert np.abs(r["smats"][0, 0]).max() < 1e-7


def test_atomic_limit_closed_form():
    """V=0, hfmode, half filling: G = 1/2 [1/(iw-U/2) + 1/(iw+U/2)],
    Sigma = U^2/(4 iw) (the bath levels decouple)."""
    g = load("m7_atomic")
    U = float(g["uloc"][0])
    z = 1j * g["wm"]
    gat = 0.5 * (1 / (z - U / 2) + 1 / (z + U / 2))
    assert np.abs(g["gmats"][0, 0] - gat).max() < 1e-12
    r = _solve_golden(g, lanc_nstates_sector=64, lanc_nstates_total=64, lanc_ngfiter=64)
    assert np.abs(r["gmats"][0, 0] - gat).max() < 1e-9
    assert np.abs(r["smats"][0, 0] - U * U / (4 * z)).max() < 1e-7


# columns of the per-state observables in the fixtures (make_golden.OBS_COLS)
_COLS = (["dens_up%d" % a for a in range(3)] + ["dens_dw%d" % a for a in range(3)] +
         ["docc%d" % a for a in range(3)] +
         ["sz2_%d%d" % (a, b) for a in range(3) for b in range(3)] +
         ["n2_%d%d" % (a, b) for a in range(3) for b in range(3)] +
         ["s2tot", "eknot_diag", "eknot_off", "epot_dd", "ehartree", "dust", "dund",
          "dse", "dph"])


def _golden_local_obs(g):
    """The reference's accumulation of the fixture's per-state expectation
    values: gs_weight terms weighted by peso; off-diagonal Hloc and the
    spin-exchange / pair-hopping terms summed without it (as the reference's
    local_energy_impurity does, ED_OBSERVABLES.f90:536-543, 597-647)."""
    w = g["obs_peso"]
    V = g["obs_vals"]
    col = {k: V[:, i] for i, k in enumerate(_COLS)}
    ws = lambda k: float(w @ col[k])
    us = lambda k: float(col[k].sum())
    norb = int(g["norb"])
    d = {}
    d["dens_up"] = np.array([ws("dens_up%d" % a) for a in range(norb)])
    d["dens_dw"] = np.array([ws("dens_dw%d" % a) for a in range(norb)])
    d["dens"] = d["dens_up"] + d["dens_dw"]
    d["magz"] = d["dens_up"] - d["dens_dw"]
    d["docc"] = np.array([ws("docc%d" % a) for a in range(norb)])
    d["sz2"] = np.array([[ws("sz2_%d%d" % (a, b)) for b in range(norb)] for a in range(norb)])
    d["n2"] = np.array([[ws("n2_%d%d" % (a, b)) for b in range(norb)] for a in range(norb)])
    d["s2tot"] = ws("s2tot")
    d["eknot"] = ws("eknot_diag") + us("eknot_off")
    d["ehartree"] = ws("ehartree")
    d["dust"] = ws("dust")
    d["dund"] = ws("dund")
    d["dse"] = us("dse")
    d["dph"] = us("dph")
    d["epot"] = (ws("epot_dd") + d["ehartree"] + float(g["jx"]) * d["dse"] +
                 float(g["jp"]) * d["dph"])
    d["eint"] = 0.0
    return d


@pytest.mark.parametrize("name", NAMES)
def test_local_observables_match_jordan_wigner(name):
    """ref_local_obs over the fixture's kept states (recomputed from the
    oracle's dense sector matrices) against the operator expectation values
    of the Jordan-Wigner construction: n, n n, S S, the Hloc, interaction,
    Hartree-shift, spin-exchange and pair-hopping parts of H."""
    g = load(name)
    ns = set_model_from(g)
    norb = int(g["norb"])
    finite = bool(int(g["finite_t"]))
    beta = float(g["beta"])
    egs = float(g["egs"])
    states = []
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            m = edref.build_sector(nu, nd)
            w, U = np.linalg.eigh(edref.dense_h(m))
            for k in range(len(w)):
                keep = (w[k] - egs < float(g["cutoff_states"])) if finite else \
                    abs(w[k] - egs) <= 1e-9
                if keep:
                    states.append((w[k], m, U[:, k]))
    assert len(states) == len(g["obs_peso"])
    zeta = sum(np.exp(-beta * (e - egs)) for e, _, _ in states) if finite else len(states)
    acc = np.zeros(edref.NOBS)
    for e, m, vec in states:
        peso = (np.exp(-beta * (e - egs)) if finite else 1.0) / zeta
        edref.local_obs(m, vec, peso, acc)
    r = edref.obs_dict(acc, norb)
    ref = _golden_local_obs(g)
    for k, v in ref.items():
        assert np.allclose(r[k], v, atol=1e-10, rtol=0), (k, r[k], v)


def _spin_chi_solve(name, **pkw):
    g = load(name)
    return g, _solve_golden(g, chiflag=True, **pkw)


def test_spin_chi_atomic_limit_closed_form():
    """V=0, half filling, T=0 (lanc_nstates_total=1): every kept state has the
    impurity singly occupied, so S_z psi = +-psi/2 is an eigenvector and the
    only pole sits at dE = 0.  With the reference's weights norm^2/zeta and
    its final division by zeta (build_chi_spin.f90:14-16, 190),
    chi(iv_0) = beta/(2 zeta), chi(iv_n>0) = 0 and chi(tau) = 1/(2 zeta) for
    every tau -- the Curie values beta/4 and 1/4 when zeta = 2."""
    g, r = _spin_chi_solve("m7_atomic", lanc_ngfiter=64)
    beta, zeta = float(g["beta"]), r["zeta"]
    assert zeta >= 2
    assert abs(r["spin_chi_iv"][0, 0] - beta / (2 * zeta)) < 1e-12
    assert np.abs(r["spin_chi_iv"][0, 1:]).max() < 1e-12
    assert np.abs(r["spin_chi_tau"][0] - 1.0 / (2 * zeta)).max() < 1e-12
    assert np.abs(r["spin_chi_w"][0]).max() < 1e-12


@pytest.mark.parametrize("name", ["m1_norb1_nb3", "m2_norb2_nb1"])
def test_spin_chi_consistency(name):
    """T=0 spin susceptibility, per orbital and (Norb=2) total:
    - chi(tau) = chi(beta - tau) (particle and hole terms mirror each other);
    - chi(iv_n) is real (S_z is hermitian);
    - chi(tau = 0) = sum over kept states of (<S_z^2> + <S_z>^2)/zeta^2, from
      the state vectors directly (the hole terms at tau = 0 add only the
      dE = 0 pole; the others are below exp(-beta*gap));
    - chi(iv_n) = int_0^beta exp(i v_n tau) chi(tau) dtau (trapezoid on the
      tau grid), which ties the two formulas' signs and factors."""
    g, r = _spin_chi_solve(name, lanc_ngfiter=400)
    norb, beta, zeta = int(g["norb"]), float(g["beta"]), r["zeta"]
    ns = (int(g["nbath"]) + 1) * norb
    rows = [[a] for a in range(norb)] + ([list(range(norb))] if norb > 1 else [])
    tau, vm = r["tau"], r["vm"]
    for row, orbs in enumerate(rows):
        ct, civ = r["spin_chi_tau"][row], r["spin_chi_iv"][row]
        assert np.abs(ct - ct[::-1]).max() <= 1e-10 * np.abs(ct).max()
        assert np.abs(civ.imag).max() <= 1e-12 * max(1.0, np.abs(civ.real).max())
        exp0 = 0.0
        for (e, (nu, nd)), vec in zip(r["states"], r["vectors"]):
            m = np.asarray(edref.build_sector(nu, nd), dtype=np.int64)
            sz = sum(0.5 * (((m >> a) & 1) - ((m >> (a + ns)) & 1)) for a in orbs)
            p2 = np.abs(vec) ** 2
            exp0 += (np.dot(p2, sz * sz) + np.dot(p2, sz) ** 2) / zeta ** 2
        assert abs(ct[0] - exp0) <= 1e-9 * max(1.0, exp0)
        for n in (0, 1, 2):
            quad = np.trapezoid(np.exp(1j * vm[n] * tau) * ct, tau)
            assert abs(quad - civ[n]) <= 5e-3 * abs(civ[n]) + 1e-12


def test_dens_chi_atomic_limit_closed_form():
    """V=0, half filling, T=0: every kept state has the impurity singly
    occupied, so N psi = psi and the one pole sits at dE = 0.  The particle
    and hole terms then give chi(tau) = 2*sum_states 1/zeta^2 = 2/zeta for
    every tau (weights norm^2/zeta, then the final division by zeta,
    build_chi_dens.f90:36-38, 521), while the reference's iv = 0 terms of the
    two branches (-peso*beta and +peso*beta, :540 and :561) cancel and every
    other frequency vanishes with exp(-beta*dE) - 1 = 0.  With Norb = 1 the
    mix and total arrays stay zero (:10-35)."""
    g = load("m7_atomic")
    r = _solve_golden(g, chiflag=True, lanc_ngfiter=64)
    zeta = r["zeta"]
    assert zeta >= 2
    assert np.abs(r["dens_chi_tau"][0, 0] - 2.0 / zeta).max() < 1e-12
    assert np.abs(r["dens_chi_iv"][0, 0]).max() < 1e-12
    assert np.abs(r["dens_chi_w"][0, 0]).max() < 1e-12
    for k in ("dens_chi_mix_iv", "dens_chi_mix_tau", "dens_chi_mix_w", "dens_chi_tot_iv",
              "dens_chi_tot_tau", "dens_chi_tot_w"):
        assert not np.any(r[k]), k


def _quad(ct, tau, vm, n):
    return np.trapezoid(np.exp(1j * vm[n] * tau) * ct, tau)


@pytest.mark.parametrize("name", ["m1_norb1_nb3", "m2_norb2_nb1", "m8_norb2_nb2_nohf"])
def test_dens_chi_diag_tot_mix_consistency(name):
    """T=0 charge susceptibility, diagonal, total and mix:
    - chi(iv_0) = 0 for the diagonal and total arrays: one chain feeds both
      branches and the reference's iv = 0 terms have opposite signs
      (build_chi_dens.f90:539-542 against 560-563);
    - chi(iv_n) = int_0^beta exp(i v_n tau) chi(tau) dtau for n >= 1 (trapezoid),
      which ties the tau and frequency formulas' signs and factors;
    - the diagonal chi(tau) at tau = 0 is sum over kept states of
      (<N_a^2> + <N_a>^2)/zeta^2 from the state vectors (the hole term at
      tau = 0 keeps only the dE = 0 pole, the rest is below exp(-beta*gap));
    - mix(b, a)(tau) = mix(a, b)(beta - tau): the particle chain of (a, b) is
      the hole chain of (b, a) (:339-490)."""
    g = load(name)
    r = _solve_golden(g, chiflag=True, lanc_ngfiter=400)
    norb, zeta = int(g["norb"]), r["zeta"]
    ns = (int(g["nbath"]) + 1) * norb
    tau, vm = r["tau"], r["vm"]
    arrays = [(r["dens_chi_tau"][a, a], r["dens_chi_iv"][a, a]) for a in range(norb)]
    if norb > 1:
        arrays.append((r["dens_chi_tot_tau"], r["dens_chi_tot_iv"]))
    for ct, civ in arrays:
        assert abs(civ[0]) <= 1e-12 * np.abs(civ).max()
        for n in (1, 2, 3):
            assert abs(_quad(ct, tau, vm, n) - civ[n]) <= 5e-3 * abs(civ[n]) + 1e-12
    for a in range(norb):
        exp0 = 0.0
        for (e, (nu, nd)), vec in zip(r["states"], r["vectors"]):
            m = np.asarray(edref.build_sector(nu, nd), dtype=np.int64)
            na = ((m >> a) & 1) + ((m >> (a + ns)) & 1)
            p2 = np.abs(vec) ** 2
            exp0 += (np.dot(p2, na * na) + np.dot(p2, na) ** 2) / zeta ** 2
        ct = r["dens_chi_tau"][a, a]
        assert abs(ct[0] - exp0) <= 1e-9 * max(1.0, exp0)
    if norb > 1:
        mt, miv = r["dens_chi_mix_tau"], r["dens_chi_mix_iv"]
        assert np.abs(mt).max() > 0
        for a in range(norb):
            for b in range(norb):
                assert np.abs(mt[b, a] - mt[a, b][::-1]).max() <= 1e-10 * np.abs(mt).max()
                for n in (1, 2):
                    q = _quad(mt[a, b], tau, vm, n)
                    assert abs(q - miv[a, b, n]) <= 5e-3 * abs(miv[a, b, n]) + 1e-9


@pytest.mark.parametrize("name", ["m2_norb2_nb1", "m8_norb2_nb2_nohf"])
def test_dens_chi_offdiag_against_total_and_diagonal(name):
    """Norb = 2, T=0, chains run to the end of each small sector.  The
    (N_0 + N_1) chain of the off-diagonal element is the total chain; the two
    complex chains (N_0 -/+ i N_1, weight i*norm^2, build_chi_dens.f90:238-
    304) carry, under the real H, the spectral measure of N_0 psi plus that of
    N_1 psi, i.e. the two diagonal chains.  So, with dens divided by zeta
    twice and tot once:
      dens_tau(0,1) = tot_tau/zeta                      (tau takes Re only)
      dens_iv(0,1)  = tot_iv/zeta + i (dens_iv(0,0) + dens_iv(1,1))
      dens_w(0,1)   = (tot_w/zeta - dens_w(0,0) - dens_w(1,1))/2
    the last after the (1+i) combination of :19-23, in which the complex
    chains cancel and the cross term <N_0; N_1> remains."""
    g = load(name)
    r = _solve_golden(g, chiflag=True, lanc_ngfiter=400)
    zeta = r["zeta"]
    dt, div, dw = r["dens_chi_tau"], r["dens_chi_iv"], r["dens_chi_w"]
    tt, tiv, tw = r["dens_chi_tot_tau"], r["dens_chi_tot_iv"], r["dens_chi_tot_w"]
    scale = np.abs(div).max()
    assert np.abs(dt[0, 1] - tt / zeta).max() <= 1e-10 * np.abs(tt).max()
    assert np.abs(div[0, 1] - (tiv / zeta + 1j * (div[0, 0] + div[1, 1]))).max() <= 1e-8 * scale
    want = 0.5 * (tw / zeta - dw[0, 0] - dw[1, 1])
    assert np.abs(dw[0, 1] - want).max() <= 1e-8 * np.abs(dw).max()
    assert not np.any(dt[1, 0]) and not np.any(div[1, 0]) and not np.any(dw[1, 0])


def test_dens_chi_mix_tau0_sum_rule():
    """mix(a,b)(tau=0) at T=0: the particle chain of c^+_b c_a psi adds its
    full weight sum ||c^+_b c_a psi||^2/zeta; the hole chain of c^+_a c_b psi
    adds exp(-beta*dE) times its weights, which keeps only the overlaps with
    the sector's ground states: |<gs'|c^+_a c_b|psi>|^2/zeta, from a dense
    eigen-decomposition of the sector.  c and c^+ act through their matrices
    built here from the Jordan-Wigner signs on the state codes."""
    g = load("m2_norb2_nb1")
    r = _solve_golden(g, chiflag=True, lanc_ngfiter=400)
    norb, zeta, egs = int(g["norb"]), r["zeta"], r["egs"]
    ns = (int(g["nbath"]) + 1) * norb

    def op(m_from, m_to, pos, dagger):
        idx = {int(x): i for i, x in enumerate(m_to)}
        M = np.zeros((len(m_to), len(m_from)))
        for j, x in enumerate(m_from):
            x = int(x)
            occ = (x >> pos) & 1
            if occ == (0 if dagger else 1):
                y = x ^ (1 << pos)
                M[idx[y], j] = (-1) ** bin(x & ((1 << pos) - 1)).count("1")
        return M

    for a in range(norb):
        for b in range(norb):
            want = 0.0
            for (e, (nu, nd)), vec in zip(r["states"], r["vectors"]):
                m = edref.build_sector(nu, nd)
                mm = edref.build_sector(nu - 1, nd)
                w, U = np.linalg.eigh(edref.dense_h(m))
                gsv = U[:, np.abs(w - egs) < 1e-9]
                fwd = op(mm, m, b, True) @ op(m, mm, a, False) @ vec
                bwd = op(mm, m, a, True) @ op(m, mm, b, False) @ vec
                want += (fwd @ fwd + np.sum((gsv.T @ bwd) ** 2)) / zeta
            got = r["dens_chi_mix_tau"][a, b, 0]
            assert abs(got - want) <= 1e-8 * max(1.0, abs(want)), (a, b, got, want)


@pytest.mark.parametrize("name", ["m9_hybrid_norb2_nb3", "m10_hybrid_norb3_nb2_spin2"])
def test_ed_solve_hybrid_matches_lehmann(name):
    """Hybrid bath (Nbath levels shared by the two orbitals, Ns = Nbath+Norb,
    off-diagonal Hloc): E0, n, d and the full orbital matrix G_ab(iw_n) of the
    oracle's solve -- diagonal chains plus the (c_a + c_b), (c_a -/+ i c_b)
    chains and the (1-i) combination of build_gf_normal.f90:18-55 -- against
    the Jordan-Wigner Lehmann sum of the fixture.  The real axis carries the
    ghost-pole error of the plain chain (see the Lehmann test above), compared
    in relative L2 over all elements and points."""
    g = load(name)
    r = _solve_golden(g, lanc_ngfiter=400)
    assert abs(r["egs"] - g["egs"]) <= 1e-10 * abs(g["egs"])
    assert np.allclose(r["dens"], g["dens"], atol=1e-10)
    assert np.allclose(r["docc"], g["docc"], atol=1e-10)
    nsp = int(g["nspin"])
    assert np.abs(r["gmats_full"] - g["gmats_full"][:nsp]).max() < 1e-8
    assert np.abs(g["gmats_full"][0, 0, 1]).max() > 0.05
    x, y = r["greal_full"], g["greal_full"][:nsp]
    assert np.linalg.norm(x - y) / np.linalg.norm(y) < 1e-3


def test_hybrid_sigma_vanishes_without_interaction():
    """U = Ust = Jh = 0 on the hybrid bath: G is the non-interacting G0, so the
    matrix Sigma = G0^-1 - G^-1 (get_sigma_normal.f90:40-56, with Delta_ab
    and the off-diagonal Hloc in G0^-1) must vanish in every orbital element."""
    g = load("m9_hybrid_norb2_nb3")
    g.update(uloc=np.zeros(3), ust=np.array(0.0), jh=np.array(0.0))
    r = _solve_golden(g, lanc_ngfiter=400)
    assert np.abs(r["gmats_full"][0, 0, 1]).max() > 0.01
    assert np.abs(r["smats_full"]).max() < 1e-8
