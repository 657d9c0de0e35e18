"""ctypes binding of libdmed.so (the C ABI declared in include/dmed.h).

Used by the tests and bench.py.  The drop-in interface for DMFT codes is the
Fortran module DMFT_ED in dmft-ed_amd/fortran/; this module mirrors the same
C entry points for Python callers.  There is no CPU fallback: if the library
is missing or no GPU is present, calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libdmed.so")

FP64, FP32, MIXED = 0, 1, 2   # MIXED: chains only (FP64 head steps, then FP32)
ERRORS = {-1: "EINVAL", -2: "ESTATE", -3: "EHIP", -4: "ENOMEM", -5: "EUNSUPP",
          -6: "ENOCONV", -7: "ENORM"}

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class DmedError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Model(C.Structure):
    _fields_ = [("norb", C.c_int32), ("nbath", C.c_int32), ("nspin", C.c_int32),
                ("bath", _dp), ("hloc", _dp), ("uloc", C.c_double * 3),
                ("ust", C.c_double), ("jh", C.c_double), ("jx", C.c_double),
                ("jp", C.c_double), ("xmu", C.c_double), ("hfmode", C.c_int32),
                ("jhflag", C.c_int32), ("bath_type", C.c_int32), ("pad_", C.c_int32)]


class SolveParams(C.Structure):
    _fields_ = [("lanc_nstates_sector", C.c_int32), ("lanc_nstates_total", C.c_int32),
                ("lanc_nstates_step", C.c_int32), ("lanc_niter", C.c_int32),
                ("lanc_ngfiter", C.c_int32), ("lanc_ncv_factor", C.c_int32),
                ("lanc_ncv_add", C.c_int32), ("lanc_dim_threshold", C.c_int32),
                ("lanc_tolerance", C.c_double), ("gs_threshold", C.c_double),
                ("cutoff", C.c_double), ("beta", C.c_double), ("eps", C.c_double),
                ("wini", C.c_double), ("wfin", C.c_double), ("lmats", C.c_int32),
                ("lreal", C.c_int32), ("prec_gf", C.c_int32), ("batched", C.c_int32),
                ("verbose", C.c_int32), ("gf_fp64_head", C.c_int32),
                ("prec_gs", C.c_int32), ("chi_spin", C.c_int32), ("ltau", C.c_int32),
                ("chi_dens", C.c_int32)]


class LocalObs(C.Structure):
    _fields_ = [("dens", C.c_double * 3), ("dens_up", C.c_double * 3),
                ("dens_dw", C.c_double * 3), ("docc", C.c_double * 3), ("magz", C.c_double * 3),
                ("sz2", C.c_double * 9), ("n2", C.c_double * 9), ("s2tot", C.c_double),
                ("eknot", C.c_double), ("epot", C.c_double), ("ehartree", C.c_double),
                ("dust", C.c_double), ("dund", C.c_double), ("dse", C.c_double),
                ("dph", C.c_double), ("eint", C.c_double)]

    def as_dict(self, norb):
        d = {}
        for k, t in self._fields_:
            v = getattr(self, k)
            if t is C.c_double:
                d[k] = float(v)
            elif len(v) == 3:
                d[k] = np.array(v[:norb])
            else:
                d[k] = np.array(v[:]).reshape(3, 3)[:norb, :norb]
        return d


class SolveOut(C.Structure):
    _fields_ = [("egs", C.c_double), ("zeta", C.c_double), ("nstates", C.c_int32),
                ("numgs", C.c_int32), ("dens", _dp), ("docc", _dp), ("gmats", _dp),
                ("greal", _dp), ("smats", _dp), ("sreal", _dp),
                ("neigen_sector", _i32p), ("lanc_nstates_total", _i32p),
                ("t_diag", C.c_double), ("t_gf", C.c_double),
                ("obs", C.POINTER(LocalObs)), ("spin_chi_iv", _dp), ("spin_chi_tau", _dp),
                ("spin_chi_w", _dp), ("ltau_used", C.c_int32), ("pad2_", C.c_int32),
                ("dens_chi_iv", _dp), ("dens_chi_tau", _dp), ("dens_chi_w", _dp),
                ("dens_chi_mix_iv", _dp), ("dens_chi_mix_tau", _dp), ("dens_chi_mix_w", _dp),
                ("dens_chi_tot_iv", _dp), ("dens_chi_tot_tau", _dp), ("dens_chi_tot_w", _dp)]


_lib = None

# (name, restype, argtypes) of every entry point of include/dmed.h
SIGNATURES = [
    ("dmed_init", C.c_int, [C.c_int]),
    ("dmed_finalize", C.c_int, []),
    ("dmed_last_error", C.c_char_p, []),
    ("dmed_set_model", C.c_int, [C.POINTER(Model)]),
    ("dmed_get_ns", C.c_int, [_i32p]),
    ("dmed_binomial", C.c_int64, [C.c_int32, C.c_int32]),
    ("dmed_sector_create", C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("dmed_sector_free", C.c_int, [C.c_void_p]),
    ("dmed_sector_dims", C.c_int, [C.c_void_p, _i64p, _i64p, _i64p]),
    ("dmed_sector_map", C.c_int, [C.c_void_p, _i64p]),
    ("dmed_vec_create", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("dmed_vec_free", C.c_int, [C.c_void_p]),
    ("dmed_vec_upload", C.c_int, [C.c_void_p, _dp]),
    ("dmed_vec_download", C.c_int, [C.c_void_p, _dp]),
    ("dmed_vec_data", C.c_void_p, [C.c_void_p]),
    ("dmed_vec_copy", C.c_int, [C.c_void_p, C.c_void_p]),
    ("dmed_local_observables", C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
    ("dmed_hxv", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    ("dmed_hxv_host", C.c_int, [C.c_void_p, _dp, _dp, C.c_int32]),
    ("dmed_hxv_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    ("dmed_sync", C.c_int, []),
    ("dmed_hxv_timed", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, _dp]),
    ("dmed_hxv_shape_of", C.c_int, [C.c_void_p, C.c_int32, _i32p]),
    ("dmed_eigh", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                            C.c_int32, _dp, C.POINTER(C.c_void_p), _dp]),
    ("dmed_ground_state", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, _dp,
                                    C.POINTER(C.c_void_p), _dp]),
    ("dmed_apply_op", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, _dp]),
    ("dmed_chain", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double, _dp,
                             _dp, _i32p]),
    ("dmed_chain_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                     _dp, _dp, _i32p]),
    ("dmed_gf_add_chain", C.c_int, [C.c_double, C.c_double, _dp, _dp, C.c_int32, C.c_int32,
                                    _dp, C.c_int32, _dp, C.c_int32, C.c_double, _dp, _dp]),
    ("dmed_apply_sz", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, _dp]),
    ("dmed_apply_occ", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, _dp]),
    ("dmed_chi_add_chain", C.c_int, [C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int32,
                                     C.c_double, _dp, C.c_int32, _dp, C.c_int32, _dp, C.c_int32,
                                     C.c_double, _dp, _dp, _dp]),
    ("dmed_observables", C.c_int, [C.c_void_p, C.c_double, _dp, _dp]),
    ("dmed_solve_params_default", None, [C.POINTER(SolveParams)]),
    ("dmed_solve", C.c_int, [C.POINTER(SolveParams), C.POINTER(SolveOut)]),
    ("dmed_solve_sites", C.c_int, [C.c_int32, C.POINTER(Model), C.POINTER(SolveParams),
                                   C.POINTER(SolveOut)]),
    ("dmed_sigma", C.c_int, [C.c_int32, C.c_int32, _dp, _dp, C.c_int32, _dp]),
    ("dmed_delta", C.c_int, [C.c_int32, C.c_int32, _dp, C.c_int32, _dp]),
    ("dmed_chi2_fit_hybrid", C.c_int, [C.c_int32, C.c_int32, _dp, C.c_int32, C.c_double,
                                       C.c_double, _dp, C.c_int32, C.c_int32, _dp, _dp,
                                       C.c_int32, C.c_double, _dp, _i32p]),
    ("dmed_chi2_fit", C.c_int, [C.c_int32, C.c_int32, _dp, C.c_int32, C.c_double, C.c_double,
                                C.c_double, C.c_int32, _dp, _dp, C.c_int32, C.c_double, _dp,
                                _i32p]),
]


def lib():
    """Load libdmed.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built: run python dmft-ed_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise DmedError(rc, lib().dmed_last_error().decode())
    return rc


def _d(a):
    return a.ctypes.data_as(_dp)


def init(device=0):
    _check(lib().dmed_init(device))


def finalize():
    _check(lib().dmed_finalize())


BATH_TYPES = {"normal": 0, "hybrid": 1}


def nsites(norb, nbath, bath_type="normal"):
    """Ns of the model: (Nbath+1)*Norb, or Nbath+Norb for the hybrid bath."""
    return nbath + norb if bath_type == "hybrid" else (nbath + 1) * norb


def make_model(norb, nbath, nspin, bath, hloc_f, uloc, ust=0.0, jh=0.0, jx=0.0, jp=0.0,
               xmu=0.0, hfmode=True, jhflag=False, bath_type="normal"):
    """A dmed_model struct and the arrays it points to (keep them alive).
    bath: set_dmft_bath layout of the bath type; hloc_f: real
    Hloc(Nspin,Nspin,Norb,Norb) in Fortran order (flattened) or a 4-d array."""
    bath = np.ascontiguousarray(bath, dtype=np.float64)
    h = np.asarray(hloc_f, dtype=np.float64)
    if h.ndim == 4:
        h = h.flatten(order="F")
    h = np.ascontiguousarray(h)
    m = Model()
    m.norb, m.nbath, m.nspin = norb, nbath, nspin
    m.bath, m.hloc = _d(bath), _d(h)
    u = list(uloc) + [0.0] * (3 - len(uloc))
    for i in range(3):
        m.uloc[i] = u[i]
    m.ust, m.jh, m.jx, m.jp, m.xmu = ust, jh, jx, jp, xmu
    m.hfmode, m.jhflag = int(bool(hfmode)), int(bool(jhflag))
    m.bath_type = BATH_TYPES[bath_type]
    return m, (bath, h)


def set_model(norb, nbath, nspin, bath, hloc_f, uloc, ust=0.0, jh=0.0, jx=0.0, jp=0.0,
              xmu=0.0, hfmode=True, jhflag=False, bath_type="normal"):
    """dmed_set_model; returns Ns."""
    global _keep
    m, _keep = make_model(norb, nbath, nspin, bath, hloc_f, uloc, ust, jh, jx, jp, xmu, hfmode,
                          jhflag, bath_type)
    _check(lib().dmed_set_model(C.byref(m)))
    return nsites(norb, nbath, bath_type)


_keep = None


class Sector:
    def __init__(self, nup, ndw):
        self.h = C.c_void_p()
        _check(lib().dmed_sector_create(nup, ndw, C.byref(self.h)))
        d, du, dd = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().dmed_sector_dims(self.h, C.byref(d), C.byref(du), C.byref(dd)))
        self.nup, self.ndw = nup, ndw
        self.dim, self.dim_up, self.dim_dw = d.value, du.value, dd.value

    def free(self):
        if self.h:
            lib().dmed_sector_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def map(self):
        out = np.empty(self.dim, dtype=np.int64)
        _check(lib().dmed_sector_map(self.h, out.ctypes.data_as(_i64p)))
        return out

    def hxv(self, x, prec=FP64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().dmed_hxv_host(self.h, _d(x), _d(y), prec))
        return y

    def hxv_shape(self, prec=FP64):
        """Launch shape of this sector's plain H*v (after the autotuner ran
        for it): kernel ("tile" or "column"), lanes, rows_per_lane,
        columns_per_group, groups_per_cu, grid, tuned, row_passes."""
        v = np.zeros(8, dtype=np.int32)
        _check(lib().dmed_hxv_shape_of(self.h, prec, v.ctypes.data_as(_i32p)))
        keys = ("kernel", "lanes", "rows_per_lane", "columns_per_group", "groups_per_cu", "grid",
                "tuned", "row_passes")
        out = dict(zip(keys, (int(a) for a in v)))
        out["kernel"] = {1: "tile", 2: "column"}[out["kernel"]]
        out["tuned"] = bool(out["tuned"])
        return out

    def eigh(self, neig=1, ncv=0, maxit=0, tol=1e-12, dense_threshold=256, vectors=True):
        ev = np.zeros(neig)
        res = np.zeros(neig)
        vecs = (C.c_void_p * neig)() if vectors else None
        rc = lib().dmed_eigh(self.h, neig, ncv, maxit, tol, dense_threshold, _d(ev),
                             vecs, _d(res))
        if rc not in (0, -6):
            _check(rc)
        out = [Vec.wrap(self, C.c_void_p(v), FP64) for v in vecs] if vectors else None
        return ev, out, res, rc

    def ground_state(self, prec=FP64, maxit=0, tol=1e-12):
        """Plain-Lanczos ground state (FP64 or MIXED): (E0, Vec, residual, rc)."""
        e = C.c_double()
        r = C.c_double()
        h = C.c_void_p()
        rc = lib().dmed_ground_state(self.h, prec, maxit, tol, C.byref(e), C.byref(h), C.byref(r))
        if rc not in (0, -6):
            _check(rc)
        return e.value, Vec.wrap(self, h, FP64), r.value, rc

    def vec(self, data=None, prec=FP64):
        v = Vec(self, prec)
        if data is not None:
            v.upload(data)
        return v


class Vec:
    def __init__(self, sector, prec=FP64, handle=None):
        self.sector = sector
        self.prec = prec
        if handle is None:
            self.h = C.c_void_p()
            _check(lib().dmed_vec_create(sector.h, prec, C.byref(self.h)))
        else:
            self.h = handle

    @classmethod
    def wrap(cls, sector, handle, prec):
        return cls(sector, prec, handle)

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.shape == (self.sector.dim,)
        _check(lib().dmed_vec_upload(self.h, _d(x)))

    def download(self):
        out = np.empty(self.sector.dim)
        _check(lib().dmed_vec_download(self.h, _d(out)))
        return out

    def copy_from(self, src):
        """self = src on the device (precision converted)."""
        _check(lib().dmed_vec_copy(self.h, src.h))

    def ptr(self):
        return lib().dmed_vec_data(self.h)

    def free(self):
        if self.h:
            lib().dmed_vec_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def apply_op(psi, isite, dagger, dst):
    out = Vec(dst, psi.prec)
    n2 = C.c_double()
    _check(lib().dmed_apply_op(psi.h, isite, int(bool(dagger)), dst.h, out.h, C.byref(n2)))
    return out, n2.value


def chain(sector, v0, nlanc, prec=FP64, threshold=1e-13):
    a = np.zeros(nlanc)
    b = np.zeros(nlanc)
    ns = C.c_int32()
    _check(lib().dmed_chain(sector.h, v0.h, nlanc, prec, threshold, _d(a), _d(b), C.byref(ns)))
    return a, b, ns.value


def gf_add_chain(peso, ei, alpha, beta, isign, wm, wr, eps, gm, gr):
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    wm = np.ascontiguousarray(wm, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    _check(lib().dmed_gf_add_chain(peso, ei, _d(alpha), _d(beta), len(alpha), isign, _d(wm),
                                   len(wm), _d(wr), len(wr), eps, _d(gm.view(np.float64)),
                                   _d(gr.view(np.float64))))


def observables(psi, peso, norb):
    dens = np.zeros(norb)
    docc = np.zeros(norb)
    _check(lib().dmed_observables(psi.h, peso, _d(dens), _d(docc)))
    return dens, docc


def sigma(ispin, iorb, z, g):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    g = np.ascontiguousarray(g, dtype=np.complex128)
    out = np.empty_like(z)
    _check(lib().dmed_sigma(ispin, iorb, _d(z.view(np.float64)), _d(g.view(np.float64)),
                            len(z), _d(out.view(np.float64))))
    return out


def default_params(**kw):
    p = SolveParams()
    lib().dmed_solve_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _solve_out(norb, nspin, ns, p):
    """A SolveOut wired to fresh result arrays (the reference's layouts)."""
    g = {}
    for k, L in (("gmats", p.lmats), ("greal", p.lreal), ("smats", p.lmats), ("sreal", p.lreal)):
        g[k] = np.zeros((nspin, nspin, norb, norb, L), dtype=np.complex128, order="F")
    r = dict(dens=np.zeros(norb), docc=np.zeros(norb),
             neigen_sector=np.zeros((ns + 1) ** 2, dtype=np.int32), **g)
    ntot = C.c_int32(p.lanc_nstates_total)
    o = SolveOut()
    for k in ("dens", "docc", "gmats", "greal", "smats", "sreal"):
        setattr(o, k, r[k].ctypes.data_as(_dp))
    o.neigen_sector = r["neigen_sector"].ctypes.data_as(_i32p)
    o.lanc_nstates_total = C.pointer(ntot)
    obs = LocalObs()
    o.obs = C.pointer(obs)
    ltau = max(int(p.beta), p.ltau)     # Ltau raised to beta
    if p.chi_spin:
        # spinChi_* (Norb+1, 0:L) in Fortran order
        r["spin_chi_iv"] = np.zeros((norb + 1, p.lmats + 1), dtype=np.complex128, order="F")
        r["spin_chi_tau"] = np.zeros((norb + 1, ltau + 1), order="F")
        r["spin_chi_w"] = np.zeros((norb + 1, p.lreal), dtype=np.complex128, order="F")
    if p.chi_dens:
        # densChi_*, densChi_mix_* (Norb, Norb, 0:L) and densChi_tot_* (0:L), Fortran order
        for pre in ("dens_chi_", "dens_chi_mix_"):
            r[pre + "iv"] = np.zeros((norb, norb, p.lmats + 1), dtype=np.complex128, order="F")
            r[pre + "tau"] = np.zeros((norb, norb, ltau + 1), order="F")
            r[pre + "w"] = np.zeros((norb, norb, p.lreal), dtype=np.complex128, order="F")
        r["dens_chi_tot_iv"] = np.zeros(p.lmats + 1, dtype=np.complex128)
        r["dens_chi_tot_tau"] = np.zeros(ltau + 1)
        r["dens_chi_tot_w"] = np.zeros(p.lreal, dtype=np.complex128)
    for k in r:
        if k.startswith(("spin_chi_", "dens_chi_")):
            setattr(o, k, r[k].ctypes.data_as(_dp))
    return o, r, ntot, obs


def _solve_result(o, r, ntot, obs, norb):
    return dict(egs=o.egs, zeta=o.zeta, nstates=o.nstates, numgs=o.numgs,
                lanc_nstates_total=ntot.value, t_diag=o.t_diag, t_gf=o.t_gf,
                obs=obs.as_dict(norb), **r)


def solve(norb, nspin, ns, params):
    """ed_solve on the current model.  Returns a dict with arrays shaped like
    the reference's (Nspin,Nspin,Norb,Norb,L) complex outputs."""
    o, r, ntot, obs = _solve_out(norb, nspin, ns, params)
    _check(lib().dmed_solve(C.byref(params), C.byref(o)))
    return _solve_result(o, r, ntot, obs, norb)


def solve_sites(sites, params):
    """ed_solve_lattice: one dmed_solve per site, sites[i] the keyword
    arguments of make_model; concurrent on the GPU when params.batched.
    Returns a list of solve() dicts."""
    ms = (Model * len(sites))()
    keep, outs, res = [], (SolveOut * len(sites))(), []
    for i, kw in enumerate(sites):
        m, k = make_model(**kw)
        ms[i] = m
        keep.append(k)
        norb, nspin = kw["norb"], kw["nspin"]
        ns = nsites(norb, kw["nbath"], kw.get("bath_type", "normal"))
        o, r, ntot, obs = _solve_out(norb, nspin, ns, params)
        outs[i] = o
        res.append((r, ntot, obs, norb))
    _check(lib().dmed_solve_sites(len(sites), ms, C.byref(params), outs))
    return [_solve_result(outs[i], r, ntot, obs, norb) for i, (r, ntot, obs, norb) in enumerate(res)]


def apply_sz(psi, iorb):
    """S_z psi (iorb = -1: all impurity orbitals) as a new vector, and its norm^2."""
    out = Vec(psi.sector, psi.prec)
    n2 = C.c_double()
    _check(lib().dmed_apply_sz(psi.h, iorb, out.h, C.byref(n2)))
    return out, n2.value


def apply_occ(psi, orbitals):
    """(sum over a in orbitals of n_a,up + n_a,dw) psi as a new vector, and its norm^2."""
    out = Vec(psi.sector, psi.prec)
    n2 = C.c_double()
    mask = 0
    for a in orbitals:
        mask |= 1 << a
    _check(lib().dmed_apply_occ(psi.h, mask, out.h, C.byref(n2)))
    return out, n2.value


def chi_add_chain(peso_f, peso_bz, ei, alpha, beta_l, beta, vm, tau, wr, eps, chi_iv, chi_tau,
                  chi_w):
    """add_to_lanczos_spinChi of one chain into chi_iv (complex), chi_tau, chi_w (complex)."""
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    beta_l = np.ascontiguousarray(beta_l, dtype=np.float64)
    vm = np.ascontiguousarray(vm, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    _check(lib().dmed_chi_add_chain(peso_f, peso_bz, ei, _d(alpha), _d(beta_l), len(alpha), beta,
                                    _d(vm), len(vm), _d(tau), len(tau), _d(wr), len(wr), eps,
                                    _d(chi_iv.view(np.float64)), _d(chi_tau),
                                    _d(chi_w.view(np.float64))))


def local_observables(psi, peso, acc=None):
    """Accumulate one normalised state's local observables into acc (LocalObs)."""
    acc = LocalObs() if acc is None else acc
    _check(lib().dmed_local_observables(psi.h, peso, C.byref(acc)))
    return acc


def chi2_fit(fg, beta, e, v, scheme="weiss", weight=0, xmu=0.0, hloc=0.0, niter=500,
             ftol=1e-5):
    """Fit one orbital's bath (e, v) to fg(iw_n), n < len(fg).  Host-only."""
    fg = np.ascontiguousarray(fg, dtype=np.complex128)
    e = np.array(e, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    chi = C.c_double()
    it = C.c_int32()
    _check(lib().dmed_chi2_fit(0 if scheme == "weiss" else 1, weight, _d(fg.view(np.float64)),
                               len(fg), beta, xmu, hloc, len(e), _d(e), _d(v), niter, ftol,
                               C.byref(chi), C.byref(it)))
    return e, v, chi.value, it.value


def chi2_fit_hybrid(fg, beta, e, v, hloc, scheme="weiss", weight=0, xmu=0.0, niter=500,
                    ftol=1e-5):
    """Fit one spin of a hybrid bath: fg (Norb, Norb, L) complex, e (Nbath,),
    v (Norb, Nbath), hloc (Norb, Norb) real.  Host-only."""
    fg = np.asfortranarray(fg, dtype=np.complex128)
    norb, nb = v.shape
    e = np.array(e, dtype=np.float64)
    vv = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(norb * nb))
    h = np.asfortranarray(hloc, dtype=np.float64)
    chi = C.c_double()
    it = C.c_int32()
    _check(lib().dmed_chi2_fit_hybrid(0 if scheme == "weiss" else 1, weight,
                                      fg.ctypes.data_as(_dp), fg.shape[2], beta, xmu,
                                      h.ctypes.data_as(_dp), norb, nb, _d(e), _d(vv), niter, ftol,
                                      C.byref(chi), C.byref(it)))
    return e, vv.reshape(norb, nb), chi.value, it.value
