"""oracle/edref.py -- ctypes front end and solve orchestration of the oracle.

TEST INFRASTRUCTURE ONLY (the GPU results are tested against it): imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
product library.  The arithmetic lives in oracle/edref.c (a C restatement of
the reference); this file only sequences it the way the reference's
ed_solve_single does (ED_MAIN.f90:251-289):

  diagonalize_impurity  -> ed_diag_c      ED_DIAG.f90:93-228
                        -> ed_analysis    ED_DIAG.f90:248-398
  observables_impurity  (dens/docc)      ED_OBSERVABLES.f90:74-168
  local observables and energies         ED_OBSERVABLES.f90:74-150, 450-669
  buildgf_impurity      -> build_gf_normal / lanc_build_gf_normal_c
                                          build_gf_normal.f90:4-180
                        -> get_sigma_normal  get_sigma_normal.f90:1-38

The per-sector eigensolver of the reference is SciFortran's sp_eigh (ARPACK,
ED_DIAG.f90:179-181) for dim > lanc_dim_threshold and LAPACK eigh on the
dense sector matrix otherwise (:185-193).  Here the dense path is numpy's
eigh on the matrix that ref_dense_h builds column by column from the oracle
H*v, and the sparse path is scipy's eigsh -- the same ARPACK family -- driven
by the oracle H*v.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libedref.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def build(force=False):
    """Compile oracle/edref.c into oracle/libedref.so with gcc (OpenMP)."""
    src = os.path.join(HERE, "edref.c")
    if not force and os.path.exists(LIB_PATH) and \
            os.path.getmtime(LIB_PATH) >= os.path.getmtime(src):
        return LIB_PATH
    import subprocess
    subprocess.check_call(["gcc", "-O2", "-fopenmp", "-shared", "-fPIC", "-Wall",
                           "-o", LIB_PATH, src, "-lm"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ref_set_model.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_int]
        L.ref_set_bath_type.argtypes = [C.c_int]
        L.ref_delta_ab.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp]
        L.ref_binomial.restype = C.c_longlong
        L.ref_binomial.argtypes = [C.c_int, C.c_int]
        L.ref_sector_dim.restype = C.c_longlong
        L.ref_build_sector.restype = C.c_longlong
        L.ref_build_sector.argtypes = [C.c_int, C.c_int, _i64p]
        L.ref_c.argtypes = [C.c_int, C.c_int64, _i64p, _dp]
        L.ref_cdg.argtypes = [C.c_int, C.c_int64, _i64p, _dp]
        L.ref_binary_search.restype = C.c_longlong
        L.ref_binary_search.argtypes = [_i64p, C.c_longlong, C.c_int64]
        L.ref_hxv.argtypes = [_i64p, C.c_longlong, _dp, _dp]
        L.ref_hxv_omp.argtypes = [_i64p, C.c_longlong, _dp, _dp, C.c_int]
        L.ref_hxv_block.argtypes = [_i64p, C.c_longlong, _dp, _dp, C.c_longlong, C.c_longlong,
                                    C.c_int]
        L.ref_pythag.restype = C.c_double
        L.ref_pythag.argtypes = [C.c_double, C.c_double]
        L.ref_tql2.argtypes = [C.c_int, _dp, _dp, _dp]
        L.ref_lanc_tridiag.argtypes = [_i64p, C.c_longlong, _dp, C.c_int, _dp, _dp,
                                       C.c_double, C.c_int, C.POINTER(C.c_int)]
        L.ref_lanczos_gs.argtypes = [_i64p, C.c_longlong, _dp, C.c_int, C.c_double,
                                     C.c_int, _dp, C.c_int]
        L.ref_apply_op.argtypes = [_i64p, C.c_longlong, _dp, _i64p, C.c_longlong,
                                   C.c_int, C.c_int, _dp]
        L.ref_add_to_gf.argtypes = [C.c_double, C.c_double, _dp, _dp, C.c_int, C.c_int,
                                    _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, _dp]
        L.ref_invg0.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp]
        L.ref_delta.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp]
        L.ref_sigma.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp]
        L.ref_observables.argtypes = [_i64p, C.c_longlong, _dp, C.c_double, _dp, _dp]
        L.ref_dense_h.argtypes = [_i64p, C.c_longlong, _dp]
        L.ref_local_obs.argtypes = [_i64p, C.c_longlong, _dp, C.c_double, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_i64p)


# ------------------------------------------------------------ wrappers
def set_model(norb, nbath, nspin, bath, hloc_f, uloc, ust=0.0, jh=0.0, jx=0.0, jp=0.0,
              xmu=0.0, hfmode=True, jhflag=False, bath_type="normal"):
    """bath_type "normal" or "hybrid" (Nbath levels shared by all orbitals);
    returns Ns."""
    bath = np.ascontiguousarray(bath, dtype=np.float64)
    hloc_f = np.ascontiguousarray(hloc_f, dtype=np.float64)
    u = np.zeros(3)
    u[:len(uloc)] = uloc
    hyb = bath_type == "hybrid"
    lib().ref_set_bath_type(int(hyb))
    rc = lib().ref_set_model(norb, nbath, nspin, _d(bath), _d(hloc_f), _d(u), ust, jh,
                             jx, jp, xmu, int(bool(hfmode)), int(bool(jhflag)))
    if rc != 0:
        raise ValueError(f"ref_set_model failed ({rc})")
    return nbath + norb if hyb else (nbath + 1) * norb


def binomial(n1, n2):
    return lib().ref_binomial(n1, n2)


def build_sector(nup, ndw):
    L = lib()
    dim = L.ref_build_sector(nup, ndw, None)
    m = np.empty(dim, dtype=np.int64)
    L.ref_build_sector(nup, ndw, _i(m))
    return m


def op_c(pos, state, dagger=False):
    out = C.c_int64(0)
    s = C.c_double(0)
    fn = lib().ref_cdg if dagger else lib().ref_c
    rc = fn(pos, state, C.byref(out), C.byref(s))
    if rc != 0:
        raise ValueError("Pauli violation")
    return out.value, s.value


def binary_search(a, value):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return lib().ref_binary_search(_i(a), len(a), value)


def hxv(smap, x, nthreads=1):
    """Complex(8) on-the-fly H*v of the reference (directMatVec_cc)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    if nthreads > 1:
        rc = lib().ref_hxv_omp(_i(smap), len(smap), _d(x.view(np.float64)),
                               _d(y.view(np.float64)), nthreads)
    else:
        rc = lib().ref_hxv(_i(smap), len(smap), _d(x.view(np.float64)), _d(y.view(np.float64)))
    if rc != 0:
        raise RuntimeError(f"ref_hxv failed ({rc})")
    return y


def hxv_block(smap, x, first, last, nthreads=1):
    """Rows [first, last) of the complex(8) on-the-fly H*v (directMatVec_cc)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty(last - first, dtype=np.complex128)
    rc = lib().ref_hxv_block(_i(smap), len(smap), _d(x.view(np.float64)),
                             _d(y.view(np.float64)), first, last, nthreads)
    if rc != 0:
        raise RuntimeError(f"ref_hxv_block failed ({rc})")
    return y


def tql2(d, e, z=None):
    """EISPACK tql2; e has the sub-diagonal in e[1:] (e[0] ignored)."""
    n = len(d)
    d = np.array(d, dtype=np.float64)
    e = np.array(e, dtype=np.float64)
    z = np.asfortranarray(np.eye(n) if z is None else np.array(z, dtype=np.float64))
    ierr = lib().ref_tql2(n, _d(d), _d(e), z.ctypes.data_as(_dp))
    return d, z, ierr


def lanc_tridiag(smap, vin, nlanc, threshold=1e-13, nthreads=1):
    v = np.array(vin, dtype=np.complex128)
    a = np.zeros(nlanc)
    b = np.zeros(nlanc)
    ns = C.c_int(0)
    rc = lib().ref_lanc_tridiag(_i(smap), len(smap), _d(v.view(np.float64)), nlanc, _d(a),
                                _d(b), threshold, nthreads, C.byref(ns))
    if rc != 0:
        raise RuntimeError(f"ref_lanc_tridiag failed ({rc})")
    return a, b, ns.value


def lanczos_gs(smap, v0, nitermax=512, threshold=1e-13, ncheck=10, nthreads=1):
    v = np.array(v0, dtype=np.complex128)
    egs = np.zeros(nitermax)
    n = lib().ref_lanczos_gs(_i(smap), len(smap), _d(v.view(np.float64)), nitermax,
                             threshold, ncheck, _d(egs), nthreads)
    if n <= 0:
        raise RuntimeError(f"ref_lanczos_gs failed ({n})")
    return egs[:n], v


def apply_op(mapI, psi, mapJ, isite, dagger):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros(len(mapJ), dtype=np.complex128)
    rc = lib().ref_apply_op(_i(mapI), len(mapI), _d(psi.view(np.float64)), _i(mapJ),
                            len(mapJ), isite, int(bool(dagger)), _d(out.view(np.float64)))
    if rc != 0:
        raise RuntimeError("ref_apply_op: target state absent")
    return out


def add_to_gf(peso_bz, ei, alanc, blanc, isign, wm, wr, eps, gm, gr):
    alanc = np.ascontiguousarray(alanc, dtype=np.float64)
    blanc = np.ascontiguousarray(blanc, dtype=np.float64)
    wm = np.ascontiguousarray(wm, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    return lib().ref_add_to_gf(peso_bz, ei, _d(alanc), _d(blanc), len(alanc), isign,
                               _d(wm), len(wm), _d(wr), len(wr), eps,
                               _d(gm.view(np.float64)), _d(gr.view(np.float64)))


def invg0(ispin, iorb, z):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.empty_like(z)
    lib().ref_invg0(ispin, iorb, _d(z.view(np.float64)), len(z), _d(out.view(np.float64)))
    return out


def delta(ispin, iorb, z):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.empty_like(z)
    lib().ref_delta(ispin, iorb, _d(z.view(np.float64)), len(z), _d(out.view(np.float64)))
    return out


def delta_ab(ispin, a, b, z):
    """Delta_ab(z) of the current bath (delta_bath_mats, ED_BATH_FUNCTIONS.f90:221-310)."""
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.empty_like(z)
    lib().ref_delta_ab(ispin, a, b, _d(z.view(np.float64)), len(z), _d(out.view(np.float64)))
    return out


def sigma_matrix(ispin, z, g, hloc, xmu):
    """get_sigma_normal for the hybrid bath (get_sigma_normal.f90:40-56):
    Sigma = G0^-1 - G^-1 as Norb x Norb matrices at every point, with
    G0^-1 = (z + xmu) 1 - Hloc - Delta (ED_BATH_FUNCTIONS.f90:1836-1848).
    g: (Norb, Norb, L); hloc: (Norb, Norb) real."""
    norb, L = g.shape[0], g.shape[2]
    D = np.array([[delta_ab(ispin, a, b, z) for b in range(norb)] for a in range(norb)])
    out = np.empty((norb, norb, L), dtype=np.complex128)
    for i in range(L):
        g0inv = (z[i] + xmu) * np.eye(norb) - hloc - D[:, :, i]
        out[:, :, i] = g0inv - np.linalg.inv(g[:, :, i])
    return out


def sigma(ispin, iorb, z, g):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    g = np.ascontiguousarray(g, dtype=np.complex128)
    out = np.empty_like(z)
    lib().ref_sigma(ispin, iorb, _d(z.view(np.float64)), _d(g.view(np.float64)), len(z),
                    _d(out.view(np.float64)))
    return out


def observables(smap, psi, peso, norb):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    dens = np.zeros(norb)
    docc = np.zeros(norb)
    lib().ref_observables(_i(smap), len(smap), _d(psi.view(np.float64)), peso, _d(dens),
                          _d(docc))
    return dens, docc


OBS_FIELDS = (("dens", 0, 3), ("dens_up", 3, 3), ("dens_dw", 6, 3), ("docc", 9, 3),
              ("magz", 12, 3), ("sz2", 15, 9), ("n2", 24, 9), ("s2tot", 33, 1),
              ("eknot", 34, 1), ("epot", 35, 1), ("ehartree", 36, 1), ("dust", 37, 1),
              ("dund", 38, 1), ("dse", 39, 1), ("dph", 40, 1), ("eint", 41, 1))
NOBS = 42


def local_obs(smap, psi, peso, acc=None):
    """Accumulate observables_impurity / local_energy_impurity of one state
    into the flat 42-vector (layout OBS_FIELDS)."""
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros(NOBS) if acc is None else acc
    lib().ref_local_obs(_i(smap), len(smap), _d(psi.view(np.float64)), peso, _d(out))
    return out


def obs_dict(v, norb):
    """Flat 42-vector -> dict; per-orbital arrays trimmed to norb, pairs to norb x norb."""
    d = {}
    for k, o, n in OBS_FIELDS:
        if n == 1:
            d[k] = float(v[o])
        elif n == 3:
            d[k] = np.array(v[o:o + norb])
        else:
            d[k] = np.array(v[o:o + 9]).reshape(3, 3)[:norb, :norb]
    return d


def dense_h(smap):
    n = len(smap)
    H = np.zeros((n, n), order="F")
    rc = lib().ref_dense_h(_i(smap), n, H.ctypes.data_as(_dp))
    if rc != 0:
        raise RuntimeError("ref_dense_h failed")
    return H


# ------------------------------------------------------ solve sequence
class StateList:
    """Energy-ordered state list with the reference's insert/pop rules
    (ED_EIGENSPACE.f90:126-299): a new state goes before existing states of
    equal energy; pop removes the last (highest) state; emax is only ever
    raised by inserts (es_pop_state leaves it stale, :284-290)."""

    def __init__(self):
        self.items = []            # (E, sector, vec)
        self.emax = -np.inf
        self.emin = np.inf

    def free(self):
        self.items = []
        self.emax = -np.inf
        self.emin = np.inf

    def insert(self, e, sector, vec):
        k = 0
        while k < len(self.items) and not (e <= self.items[k][0]):
            k += 1
        self.items.insert(k, (e, sector, vec))
        self.emax = max(self.emax, e)
        self.emin = min(self.emin, e)

    def add(self, e, sector, vec, size=None):
        if size is None or len(self.items) < size:
            self.insert(e, sector, vec)
        elif e < self.items[-1][0]:
            self.items.pop()
            self.insert(e, sector, vec)

    def pop(self):
        self.items.pop()


def default_params(**kw):
    p = dict(lanc_nstates_sector=6, lanc_nstates_total=1, lanc_nstates_step=2,
             lanc_niter=512, lanc_ngfiter=200, lanc_tolerance=1e-12,
             lanc_dim_threshold=256, lanc_ncv_factor=3, lanc_ncv_add=5,
             gs_threshold=1e-9, cutoff=1e-9, beta=1000.0, lmats=5000, lreal=5000,
             wini=-5.0, wfin=5.0, eps=0.01, chain_threshold=1e-13, chiflag=False, ltau=1000)
    p.update(kw)
    return p


def spin_chi_poles(peso_f, peso_bz, ei, lam, w2, beta, vm, tau, wr, eps, chi_iv, chi_tau,
                   chi_w):
    """add_to_lanczos_spinChi (build_chi_spin.f90:186-250) for isign = +1 and -1
    of one chain with poles lam and weights w2 = Z(1,j)^2 (pure-Python loops
    over the poles, numpy over the frequencies)."""
    for isign in (1, -1):
        for j in range(len(lam)):
            de = lam[j] - ei
            peso = peso_f * w2[j] * peso_bz
            ex = np.exp(-beta * de)
            chi_iv[0] += peso * beta if beta * de < 1e-1 else peso * (1.0 - ex) / de
            if isign == 1:
                chi_iv[1:] += peso * (ex - 1.0) / (1j * vm[1:] - de)
                chi_tau += peso * np.exp(-tau * de)
                chi_w += peso * (ex - 1.0) / (wr + 1j * eps - de)
            else:
                chi_iv[1:] += peso * (1.0 - ex) / (1j * vm[1:] + de)
                chi_tau += peso * np.exp(-(beta - tau) * de)
                chi_w += peso * (1.0 - ex) / (wr + 1j * eps + de)


def sz_vector(smap, vec, ns, orbs):
    """S_z |psi> with S_z = (1/2) sum_{a in orbs} (n_a,up - n_a,dw), from the
    state codes (build_chi_spin.f90:55-60, 131-136; up bits 0..ns-1, down
    bits ns..2ns-1)."""
    smap = np.asarray(smap, dtype=np.int64)
    sgn = np.zeros(len(smap))
    for a in orbs:
        sgn += ((smap >> a) & 1) - ((smap >> (a + ns)) & 1)
    return 0.5 * sgn * np.asarray(vec)


def dens_chi_branch(peso, peso_bz, ei, lam, w2, isign, beta, vm, tau, wr, eps, chi_iv,
                    chi_tau, chi_w):
    """One isign branch of add_to_lanczos_densChi / _densChi_mix / _densChi_tot
    (build_chi_dens.f90:505-733) for a chain with poles lam and weights
    w2 = Z(1,j)^2; peso may be complex (the xi*norm0 weights of :263, 300).
    chi_tau is a real array in the reference, so it takes the real part.  The
    iv = 0 term of isign = +1 is -peso*beta or peso*(exp(-beta dE)-1)/dE, as
    the reference writes it (:539-542)."""
    for j in range(len(lam)):
        de = lam[j] - ei
        pj = peso * w2[j] * peso_bz
        ex = np.exp(-beta * de)
        if isign == 1:
            chi_iv[0] += -pj * beta if beta * de < 1e-1 else pj * (ex - 1.0) / de
            chi_iv[1:] += pj * (ex - 1.0) / (1j * vm[1:] - de)
            chi_tau += np.real(pj * np.exp(-tau * de))
            chi_w += pj * (ex - 1.0) / (wr + 1j * eps - de)
        elif isign == -1:
            chi_iv[0] += pj * beta if beta * de < 1e-1 else pj * (1.0 - ex) / de
            chi_iv[1:] += pj * (1.0 - ex) / (1j * vm[1:] + de)
            chi_tau += np.real(pj * np.exp(-(beta - tau) * de))
            chi_w += pj * (1.0 - ex) / (wr + 1j * eps + de)
        else:
            raise ValueError("isign not in {-1, 1}")


def n_vector(smap, vec, ns, orbs):
    """N |psi> with N = sum_{a in orbs} (n_a,up + n_a,dw), from the state codes
    (build_chi_dens.f90:69-75, 143-149, 214-221)."""
    smap = np.asarray(smap, dtype=np.int64)
    sgn = np.zeros(len(smap))
    for a in orbs:
        sgn += ((smap >> a) & 1) + ((smap >> (a + ns)) & 1)
    return sgn * np.asarray(vec)


def chi_chain(smap, v, p):
    """sp_lanc_tridiag of the normalised start vector v/|v| and the tql2 poles
    of add_to_lanczos_*Chi: (norm2, lam, w2), or None for a zero vector (the
    reference divides by its norm there)."""
    v = np.asarray(v, dtype=np.complex128)
    norm2 = float(np.vdot(v, v).real)
    if norm2 == 0.0:
        return None
    nlanc = min(len(smap), p["lanc_ngfiter"])
    al, bl, _ = lanc_tridiag(smap, v / np.sqrt(norm2), nlanc, p["chain_threshold"])
    dd, z, ierr = tql2(al, np.concatenate(([0.0], bl[1:])))
    if ierr:
        raise RuntimeError("tql2 did not converge")
    return norm2, dd, z[0, :] ** 2


def dens_chi(items, ns, norb, nspin, zeta, egs, finite_t, p, vm, tau, wr):
    """build_chi_dens (build_chi_dens.f90:4-41) over the kept states
    items = [(e, (nu, nd), vec)]: diag (:49-110), offdiag (:188-323, complex
    start vectors N_a -/+ i N_b), the real-frequency combination (:19-23),
    mix (:339-490) and tot (:118-178); dens divided by zeta once more (:36-38).
    A mix chain whose c_{a,s} empties an empty spin species is skipped (the
    reference has no valid vector there)."""
    beta = p["beta"]
    nv, nt, lr = len(vm), len(tau), len(wr)
    d_iv = np.zeros((norb, norb, nv), dtype=np.complex128)
    d_tau = np.zeros((norb, norb, nt))
    d_w = np.zeros((norb, norb, lr), dtype=np.complex128)
    m_iv, m_tau, m_w = np.zeros_like(d_iv), np.zeros_like(d_tau), np.zeros_like(d_w)
    t_iv = np.zeros(nv, dtype=np.complex128)
    t_tau = np.zeros(nt)
    t_w = np.zeros(lr, dtype=np.complex128)

    def add(ch, e, scale, isigns, civ, ctau, cw):
        if ch is None:
            return
        norm2, lam, w2 = ch
        pbz = np.exp(-beta * (e - egs)) if finite_t else 1.0
        for isign in isigns:
            dens_chi_branch(scale * norm2 / zeta, pbz, e, lam, w2, isign, beta, vm, tau, wr,
                            p["eps"], civ, ctau, cw)

    for a in range(norb):
        for e, (nu, nd), vec in items:
            m = build_sector(nu, nd)
            add(chi_chain(m, n_vector(m, vec, ns, [a]), p), e, 1.0, (1, -1),
                d_iv[a, a], d_tau[a, a], d_w[a, a])
    if norb > 1:
        for a in range(norb):
            for b in range(a + 1, norb):
                for e, (nu, nd), vec in items:
                    m = build_sector(nu, nd)
                    add(chi_chain(m, n_vector(m, vec, ns, [a, b]), p), e, 1.0, (1, -1),
                        d_iv[a, b], d_tau[a, b], d_w[a, b])
                    va = n_vector(m, vec, ns, [a])
                    vb = n_vector(m, vec, ns, [b])
                    add(chi_chain(m, va - 1j * vb, p), e, 1j, (1,), d_iv[a, b], d_tau[a, b],
                        d_w[a, b])
                    add(chi_chain(m, va + 1j * vb, p), e, 1j, (-1,), d_iv[a, b], d_tau[a, b],
                        d_w[a, b])
        for a in range(norb):
            for b in range(a + 1, norb):
                d_w[a, b] = 0.5 * (d_w[a, b] - (1 + 1j) * d_w[a, a] - (1 + 1j) * d_w[b, b])
        for a in range(norb):
            for b in range(norb):
                for e, (nu, nd), vec in items:
                    m = build_sector(nu, nd)
                    for (i, j, isign) in ((a, b, 1), (b, a, -1)):
                        for s in range(nspin):
                            mu, md = nu - (s == 0), nd - (s == 1)
                            if mu < 0 or md < 0:
                                continue
                            mm = build_sector(mu, md)
                            t = apply_op(m, vec, mm, i + 1 + s * ns, False)
                            v = apply_op(mm, t, m, j + 1 + s * ns, True)
                            add(chi_chain(m, v, p), e, 1.0, (isign,), m_iv[a, b], m_tau[a, b],
                                m_w[a, b])
        for e, (nu, nd), vec in items:
            m = build_sector(nu, nd)
            add(chi_chain(m, n_vector(m, vec, ns, range(norb)), p), e, 1.0, (1, -1), t_iv, t_tau,
                t_w)
    return dict(dens_chi_iv=d_iv / zeta, dens_chi_tau=d_tau / zeta, dens_chi_w=d_w / zeta,
                dens_chi_mix_iv=m_iv, dens_chi_mix_tau=m_tau, dens_chi_mix_w=m_w,
                dens_chi_tot_iv=t_iv, dens_chi_tot_tau=t_tau, dens_chi_tot_w=t_w)


def hybrid_offdiag_gf(sl, ns, norb, nspin, zeta, egs, finite_t, p, wm, wr, gm, gr):
    """build_gf_normal for bath_type /= normal (build_gf_normal.f90:18-55):
    for a < b, lanc_build_gf_normal_mix_c (:192-420) adds, per kept state,
    the chains of (c^+_a + c^+_b) psi (+1) and (c_a + c_b) psi (-1) with
    weight norm^2, then (c^+_a + i c^+_b) psi (+1) and (c_a - i c_b) psi (-1)
    with weight -i*norm^2; then G_ab = (G_ab - (1-i) G_aa - (1-i) G_bb)/2 and
    G_ba = G_ab.  Returns the full (Nspin, Norb, Norb, L) matrices with the
    diagonal from gm, gr."""
    beta = p["beta"]
    gmf = np.zeros((nspin, norb, norb, len(wm)), dtype=np.complex128)
    grf = np.zeros((nspin, norb, norb, len(wr)), dtype=np.complex128)
    for s in range(nspin):
        for a in range(norb):
            gmf[s, a, a] = gm[s, a]
            grf[s, a, a] = gr[s, a]
    for s in range(nspin):
        for a in range(norb):
            for b in range(a + 1, norb):
                for e, (nu, nd), vec in sl.items:
                    mapI = build_sector(nu, nd)
                    if finite_t:
                        f = np.exp(-beta * (e - egs)) / zeta if beta * (e - egs) < 200 else 0.0
                    else:
                        f = 1.0 / zeta
                    for k, (dagger, isign, cb) in enumerate(((True, 1, 1.0), (False, -1, 1.0),
                                                             (True, 1, 1j), (False, -1, -1j))):
                        ju = nu + (1 if dagger else -1) * (s == 0)
                        jd = nd + (1 if dagger else -1) * (s == 1)
                        if ju < 0 or jd < 0 or ju > ns or jd > ns:
                            continue
                        mapJ = build_sector(ju, jd)
                        va = apply_op(mapI, vec.astype(np.complex128), mapJ, a + 1 + s * ns, dagger)
                        vb = apply_op(mapI, vec.astype(np.complex128), mapJ, b + 1 + s * ns, dagger)
                        v = va + cb * vb
                        norm2 = float(np.vdot(v, v).real)
                        if norm2 == 0.0:
                            continue
                        nlanc = min(len(mapJ), p["lanc_ngfiter"])
                        al, bl, _ = lanc_tridiag(mapJ, v / np.sqrt(norm2), nlanc,
                                                 p["chain_threshold"])
                        if k < 2:
                            add_to_gf(norm2 * f, e, al, bl, isign, wm, wr, p["eps"],
                                      gmf[s, a, b], grf[s, a, b])
                        else:
                            tm = np.zeros(len(wm), dtype=np.complex128)
                            tr = np.zeros(len(wr), dtype=np.complex128)
                            add_to_gf(-norm2 * f, e, al, bl, isign, wm, wr, p["eps"], tm, tr)
                            gmf[s, a, b] += 1j * tm
                            grf[s, a, b] += 1j * tr
        for a in range(norb):
            for b in range(a + 1, norb):
                gmf[s, a, b] = 0.5 * (gmf[s, a, b] - (1 - 1j) * gmf[s, a, a] - (1 - 1j) * gmf[s, b, b])
                grf[s, a, b] = 0.5 * (grf[s, a, b] - (1 - 1j) * grf[s, a, a] - (1 - 1j) * grf[s, b, b])
                gmf[s, b, a] = gmf[s, a, b]
                grf[s, b, a] = grf[s, a, b]
    return gmf, grf


def _sector_eigs(smap, neigen, p, seed=1234567):
    dim = len(smap)
    lanc = not (neigen == dim or dim <= p["lanc_dim_threshold"])
    if not lanc:
        H = dense_h(smap)
        w, U = np.linalg.eigh(H)
        return w[:neigen], U[:, :neigen]
    from scipy.sparse.linalg import LinearOperator, eigsh
    op = LinearOperator((dim, dim), matvec=lambda x: hxv(smap, x.astype(np.complex128)).real,
                        dtype=np.float64)
    rng = np.random.default_rng(seed)
    ncv = min(dim, max(p["lanc_ncv_factor"] * neigen + p["lanc_ncv_add"], 2 * neigen + 1))
    w, U = eigsh(op, k=neigen, which="SA", tol=p["lanc_tolerance"], ncv=ncv,
                 v0=rng.uniform(-1, 1, dim), maxiter=10000)
    order = np.argsort(w)
    return w[order], U[:, order]


def ed_solve(norb, nbath, nspin, bath, hloc_f, uloc, ust=0.0, jh=0.0, xmu=0.0,
             hfmode=True, params=None, sectors=None, jx=0.0, jp=0.0, jhflag=False,
             bath_type="normal"):
    """The reference's ed_solve_single sequence on the oracle (small Ns only).
    Returns a dict with egs, zeta, states, dens, docc, gmats, greal, smats, sreal
    (arrays indexed [ispin, iorb, freq])."""
    p = default_params() if params is None else params
    ns = set_model(norb, nbath, nspin, bath, hloc_f, uloc, ust, jh, jx, jp, xmu, hfmode, jhflag,
                   bath_type)
    finite_t = p["lanc_nstates_total"] != 1
    nstates_sector = p["lanc_nstates_sector"]
    nstates_total = p["lanc_nstates_total"]
    if finite_t:
        nstates_sector += nstates_sector % 2
        nstates_total += nstates_total % 2
    sl = StateList()
    oldzero = 1000.0
    all_sectors = [(nu, nd) for nu in range(ns + 1) for nd in range(ns + 1)]
    for (nu, nd) in (all_sectors if sectors is None else sectors):
        smap = build_sector(nu, nd)
        dim = len(smap)
        neigen = min(dim, min(dim, nstates_sector))
        w, U = _sector_eigs(smap, neigen, p)
        for i in range(neigen):
            e = float(w[i])
            if finite_t:
                sl.add(e, (nu, nd), U[:, i], size=nstates_total)
            else:
                if e < oldzero - 10.0 * p["gs_threshold"]:
                    oldzero = e
                    sl.free()
                    sl.add(e, (nu, nd), U[:, i])
                elif abs(e - oldzero) <= p["gs_threshold"]:
                    oldzero = min(oldzero, e)
                    sl.add(e, (nu, nd), U[:, i])
    beta = p["beta"]
    egs = sl.emin
    if finite_t:
        zeta = sum(np.exp(-beta * (e - egs)) for e, _, _ in sl.items)
        ec = sl.emax
        if not (np.exp(-beta * (ec - egs)) > p["cutoff"]):
            while np.exp(-beta * (sl.items[-1][0] - egs)) <= p["cutoff"]:
                sl.pop()
    else:
        zeta = float(len(sl.items))
    # observables
    dens = np.zeros(norb)
    docc = np.zeros(norb)
    acc = np.zeros(NOBS)
    for e, (nu, nd), vec in sl.items:
        peso = (np.exp(-beta * (e - egs)) if finite_t else 1.0) / zeta
        m = build_sector(nu, nd)
        dd, oo = observables(m, vec, peso, norb)
        dens += dd
        docc += oo
        local_obs(m, vec, peso, acc)
    # Green's functions
    wm = np.pi / beta * (2 * np.arange(1, p["lmats"] + 1) - 1)
    wr = np.linspace(p["wini"], p["wfin"], p["lreal"])
    gm = np.zeros((nspin, norb, p["lmats"]), dtype=np.complex128)
    gr = np.zeros((nspin, norb, p["lreal"]), dtype=np.complex128)
    for s in range(nspin):
        for a in range(norb):
            isite = a + 1 + s * ns
            for e, (nu, nd), vec in sl.items:
                mapI = build_sector(nu, nd)
                for dagger, isign in ((True, 1), (False, -1)):
                    ju = nu + (1 if dagger else -1) * (s == 0)
                    jd = nd + (1 if dagger else -1) * (s == 1)
                    if ju < 0 or jd < 0 or ju > ns or jd > ns:
                        continue
                    mapJ = build_sector(ju, jd)
                    v = apply_op(mapI, vec.astype(np.complex128), mapJ, isite, dagger)
                    norm2 = float(np.vdot(v, v).real)
                    if norm2 == 0.0:
                        # c^+|psi> (or c|psi>) vanishes: zero spectral weight.  The
                        # reference divides by sqrt(0) here (build_gf_normal.f90:115).
                        continue
                    v = v / np.sqrt(norm2)
                    nlanc = min(len(mapJ), p["lanc_ngfiter"])
                    al, bl, _ = lanc_tridiag(mapJ, v, nlanc, p["chain_threshold"])
                    if finite_t:
                        if beta * (e - egs) < 200:
                            peso = norm2 * np.exp(-beta * (e - egs)) / zeta
                        else:
                            peso = 0.0
                    else:
                        peso = norm2 / zeta
                    add_to_gf(peso, e, al, bl, isign, wm, wr, p["eps"], gm[s, a], gr[s, a])
    sm = np.zeros_like(gm)
    sr = np.zeros_like(gr)
    hybrid = bath_type == "hybrid"
    if hybrid:
        gmf, grf = hybrid_offdiag_gf(sl, ns, norb, nspin, zeta, egs, finite_t, p, wm, wr,
                                     gm, gr)
        hl = np.asarray(hloc_f, dtype=np.float64).reshape((nspin, nspin, norb, norb), order="F")
        smf = np.zeros_like(gmf)
        srf = np.zeros_like(grf)
        for s in range(nspin):
            smf[s] = sigma_matrix(s, 1j * wm, gmf[s], hl[s, s], xmu)
            srf[s] = sigma_matrix(s, wr + 1j * p["eps"], grf[s], hl[s, s], xmu)
            for a in range(norb):
                sm[s, a] = smf[s, a, a]
                sr[s, a] = srf[s, a, a]
    else:
        for s in range(nspin):
            for a in range(norb):
                sm[s, a] = sigma(s, a, 1j * wm, gm[s, a])
                sr[s, a] = sigma(s, a, wr + 1j * p["eps"], gr[s, a])
    out = dict(egs=egs, zeta=zeta, states=[(e, sec) for e, sec, _ in sl.items],
               vectors=[v for _, _, v in sl.items], dens=dens, docc=docc,
               obs=obs_dict(acc, norb), gmats=gm,
               greal=gr, smats=sm, sreal=sr, wm=wm, wr=wr)
    if hybrid:
        out.update(gmats_full=gmf, greal_full=grf, smats_full=smf, sreal_full=srf)
    if p["chiflag"]:
        # build_chi_spin (build_chi_spin.f90:4-18): rows 0..Norb-1 per orbital,
        # row Norb the total when Norb > 1; weights norm^2/zeta, then a second
        # division by zeta at the end, as the reference does
        ltau = max(int(beta), p["ltau"])
        vm = np.pi / beta * 2.0 * np.arange(p["lmats"] + 1)
        tau = np.linspace(0.0, beta, ltau + 1)
        civ = np.zeros((norb + 1, p["lmats"] + 1), dtype=np.complex128)
        ctau = np.zeros((norb + 1, ltau + 1))
        cw = np.zeros((norb + 1, p["lreal"]), dtype=np.complex128)
        rows = [[a] for a in range(norb)] + ([list(range(norb))] if norb > 1 else [])
        for row, orbs in enumerate(rows):
            for e, (nu, nd), vec in sl.items:
                m = build_sector(nu, nd)
                v = sz_vector(m, vec, ns, orbs)
                norm2 = float(np.dot(v, v))
                if norm2 == 0.0:
                    continue
                v = (v / np.sqrt(norm2)).astype(np.complex128)
                nlanc = min(len(m), p["lanc_ngfiter"])
                al, bl, _ = lanc_tridiag(m, v, nlanc, p["chain_threshold"])
                dd, z, ierr = tql2(al, np.concatenate(([0.0], bl[1:])))
                if ierr:
                    raise RuntimeError("tql2 did not converge")
                pbz = np.exp(-beta * (e - egs)) if finite_t else 1.0
                spin_chi_poles(norm2 / zeta, pbz, e, dd, z[0, :] ** 2, beta, vm, tau, wr,
                               p["eps"], civ[row], ctau[row], cw[row])
        out.update(spin_chi_iv=civ / zeta, spin_chi_tau=ctau / zeta, spin_chi_w=cw / zeta,
                   vm=vm, tau=tau)
        out.update(dens_chi(sl.items, ns, norb, nspin, zeta, egs, finite_t, p, vm, tau, wr))
    return out
