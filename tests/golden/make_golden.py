#!/usr/bin/env python3
"""Write the small golden fixtures under tests/golden/ (run once, committed).

The reference ships no tests, fixtures or golden outputs, so every vector here
is produced by an *independent* construction that shares no code and no
bit-twiddling formula with the oracle (oracle/edref.c) or with the HIP library:

* Fock space of 2*Ns fermionic modes built from Jordan-Wigner Kronecker
  products  c_p = I x ... x I x a x Z x ... x Z  (a on mode p, Z on the modes
  below p).  Mode p is bit p of the reference's combined state code
  ``iup + idw*2**Ns`` (ED_SETUP.f90:889), so mode p<Ns is up orbital p+1 and
  mode p>=Ns is down orbital p-Ns+1 (ED_SETUP.f90:1055-1081 conventions).
* The Hamiltonian is written from the physics, as operator products
  (c^+ c, n n, c^+ c^+ c c), with the coefficient conventions of
  HxVimp.f90, HxVint.f90, HxVbath.f90 and HxVimp_bath.f90.
* Sector bases are listed by brute force and ordered the way build_sector
  orders them (ED_SETUP.f90:872-889).
* Spectra come from dense numpy.linalg.eigh, Green's functions from the exact
  Lehmann sum over all eigenstates of the N+1 / N-1 sectors, with the
  reference's T=0 / finite-T weights (ED_DIAG.f90:202-213, build_gf_normal.f90:457-466).
* Local observables and energies of every kept state are expectation values
  of operator matrices (n, n n, S S, the Hloc, interaction, Hartree-shift,
  spin-exchange and pair-hopping pieces of H), stored per state with its
  weight; the pieces are checked to add up to H.

Everything is seeded (numpy default_rng); the output files are plain .npz
(no pickles).
"""
import os
import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------- Fock space
def fock_ops(nmodes):
    """Annihilators c_p (sparse, 2**nmodes square) in the integer-code basis."""
    a = sp.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))   # <0|a|1> = 1
    z = sp.csr_matrix(np.diag([1.0, -1.0]))
    eye = sp.identity(2, format="csr")
    ops = []
    for p in range(nmodes):
        # kron order: leftmost factor = most significant bit (mode nmodes-1)
        m = sp.identity(1, format="csr")
        for q in range(nmodes - 1, -1, -1):
            f = a if q == p else (z if q < p else eye)
            m = sp.kron(m, f, format="csr")
        ops.append(m)
    return ops


def nsites(mdl):
    """Ns: Norb*(Nbath+1) for the normal bath, Norb+Nbath for the hybrid bath
    (ED_SETUP.f90:55-59)."""
    if mdl.get("bath_type", "normal") == "hybrid":
        return mdl["nbath"] + mdl["norb"]
    return (mdl["nbath"] + 1) * mdl["norb"]


def bath_levels(mdl):
    """[(site, e_up, e_dw)] and [(orbital, site, v_up, v_dw)]: a normal bath
    has Nbath levels per orbital at sites Norb + a*Nbath + k, each coupled to
    its orbital only; a hybrid bath has Nbath levels at sites Norb + k, each
    coupled to every orbital (getBathStride, ED_SETUP.f90:446-455)."""
    norb, nbath, nspin = mdl["norb"], mdl["nbath"], mdl["nspin"]
    e, v, sd = mdl["e"], mdl["v"], nspin - 1
    if mdl.get("bath_type", "normal") == "hybrid":
        lev = [(norb + k, e[0][k], e[sd][k]) for k in range(nbath)]
        hyb = [(a, norb + k, v[0][a][k], v[sd][a][k]) for a in range(norb) for k in range(nbath)]
    else:
        lev = [(norb + a * nbath + k, e[0][a][k], e[sd][a][k])
               for a in range(norb) for k in range(nbath)]
        hyb = [(a, norb + a * nbath + k, v[0][a][k], v[sd][a][k])
               for a in range(norb) for k in range(nbath)]
    return lev, hyb


def model_hamiltonian(mdl):
    """Full Fock-space H (sparse) of the Anderson model (normal or hybrid bath)."""
    norb, nbath, nspin = mdl["norb"], mdl["nbath"], mdl["nspin"]
    ns = nsites(mdl)
    c = fock_ops(2 * ns)
    cd = [op.T.tocsr() for op in c]
    n = [cd[p] @ c[p] for p in range(2 * ns)]
    dim = 2 ** (2 * ns)
    H = sp.csr_matrix((dim, dim))
    e, v = mdl["e"], mdl["v"]            # [nspin][norb][nbath]
    hloc = mdl["hloc"]                   # [nspin][norb][norb] (spin-diagonal)
    U, Ust, Jh, Jx, Jp = mdl["uloc"], mdl["ust"], mdl["jh"], mdl["jx"], mdl["jp"]
    xmu, hf, jhflag = mdl["xmu"], mdl["hfmode"], mdl["jhflag"]

    def up(a):
        return a

    def dw(a):
        return a + ns

    sd = nspin - 1
    for a in range(norb):
        H = H - xmu * (n[up(a)] + n[dw(a)])
        for b in range(norb):
            # Hloc_{ab} c^+_a c_b for each spin (the diagonal a==b is n_a)
            if hloc[0][a][b] != 0.0:
                H = H + hloc[0][a][b] * (cd[up(a)] @ c[up(b)])
            if hloc[sd][a][b] != 0.0:
                H = H + hloc[sd][a][b] * (cd[dw(a)] @ c[dw(b)])
    for a in range(norb):
        H = H + U[a] * (n[up(a)] @ n[dw(a)])
    if norb > 1:
        for a in range(norb):
            for b in range(a + 1, norb):
                H = H + Ust * (n[up(a)] @ n[dw(b)] + n[up(b)] @ n[dw(a)])
                H = H + (Ust - Jh) * (n[up(a)] @ n[up(b)] + n[dw(a)] @ n[dw(b)])
    ident = sp.identity(dim, format="csr")
    if hf:
        for a in range(norb):
            H = H - 0.5 * U[a] * (n[up(a)] + n[dw(a)]) + 0.25 * U[a] * ident
        if norb > 1:
            for a in range(norb):
                for b in range(a + 1, norb):
                    ntot = n[up(a)] + n[dw(a)] + n[up(b)] + n[dw(b)]
                    H = H - 0.5 * Ust * ntot + 0.25 * Ust * ident
                    H = H - 0.5 * (Ust - Jh) * ntot + 0.25 * (Ust - Jh) * ident
    if norb > 1 and jhflag:
        for a in range(norb):
            for b in range(norb):
                if a == b:
                    continue
                # spin exchange  Jx c^+_{a up} c^+_{b dw} c_{a dw} c_{b up}
                H = H + Jx * (cd[up(a)] @ cd[dw(b)] @ c[dw(a)] @ c[up(b)])
                # pair hopping   Jp c^+_{a up} c^+_{a dw} c_{b dw} c_{b up}
                H = H + Jp * (cd[up(a)] @ cd[dw(a)] @ c[dw(b)] @ c[up(b)])
    lev, hyb = bath_levels(mdl)
    for m, eu, ed in lev:
        H = H + eu * n[up(m)] + ed * n[dw(m)]
    for a, m, vu, vd in hyb:
        H = H + vu * (cd[up(a)] @ c[up(m)] + cd[up(m)] @ c[up(a)])
        H = H + vd * (cd[dw(a)] @ c[dw(m)] + cd[dw(m)] @ c[dw(a)])
    return H.tocsr(), c, ns


OBS_COLS = (["dens_up%d" % a for a in range(3)] + ["dens_dw%d" % a for a in range(3)] +
            ["docc%d" % a for a in range(3)] +
            ["sz2_%d%d" % (a, b) for a in range(3) for b in range(3)] +
            ["n2_%d%d" % (a, b) for a in range(3) for b in range(3)] +
            ["s2tot", "eknot_diag", "eknot_off", "epot_dd", "ehartree", "dust", "dund",
             "dse", "dph"])


def local_operators(mdl, c, ns):
    """The pieces of H as separate Fock-space operators, and the local
    observables as operators, all from Jordan-Wigner products."""
    norb, nbath, nspin = mdl["norb"], mdl["nbath"], mdl["nspin"]
    cd = [op.T.tocsr() for op in c]
    n = [cd[p] @ c[p] for p in range(2 * ns)]
    dim = c[0].shape[0]
    zero = sp.csr_matrix((dim, dim))
    ident = sp.identity(dim, format="csr")
    e, v, hloc = mdl["e"], mdl["v"], mdl["hloc"]
    U, Ust, Jh = mdl["uloc"], mdl["ust"], mdl["jh"]
    sd = nspin - 1
    up = lambda a: a
    dw = lambda a: a + ns
    ops = {}
    ops["hloc_diag"] = sum((hloc[0][a][a] * n[up(a)] + hloc[sd][a][a] * n[dw(a)]
                            for a in range(norb)), zero)
    ops["hloc_off"] = sum((hloc[0][a][b] * (cd[up(a)] @ c[up(b)]) +
                           hloc[sd][a][b] * (cd[dw(a)] @ c[dw(b)])
                           for a in range(norb) for b in range(norb) if a != b), zero)
    ops["mu"] = -mdl["xmu"] * sum((n[up(a)] + n[dw(a)] for a in range(norb)), zero)
    dd = sum((U[a] * (n[up(a)] @ n[dw(a)]) for a in range(norb)), zero)
    dust = zero
    dund = zero
    hf = zero
    for a in range(norb):
        for b in range(a + 1, norb):
            dust = dust + n[up(a)] @ n[dw(b)] + n[up(b)] @ n[dw(a)]
            dund = dund + n[up(a)] @ n[up(b)] + n[dw(a)] @ n[dw(b)]
    dd = dd + Ust * dust + (Ust - Jh) * dund
    if mdl["hfmode"]:
        for a in range(norb):
            hf = hf - 0.5 * U[a] * (n[up(a)] + n[dw(a)]) + 0.25 * U[a] * ident
            for b in range(a + 1, norb):
                nt = n[up(a)] + n[dw(a)] + n[up(b)] + n[dw(b)]
                hf = hf - 0.5 * Ust * nt + 0.25 * Ust * ident
                hf = hf - 0.5 * (Ust - Jh) * nt + 0.25 * (Ust - Jh) * ident
    ops["epot_dd"], ops["ehartree"], ops["dust"], ops["dund"] = dd, hf, dust, dund
    se = zero
    ph = zero
    if norb > 1 and mdl["jhflag"]:
        for a in range(norb):
            for b in range(norb):
                if a != b:
                    se = se + cd[up(a)] @ cd[dw(b)] @ c[dw(a)] @ c[up(b)]
                    ph = ph + cd[up(a)] @ cd[dw(a)] @ c[dw(b)] @ c[up(b)]
    ops["dse"], ops["dph"] = se, ph
    lev, hyb = bath_levels(mdl)
    ops["bath"] = sum((eu * n[up(m)] + ed * n[dw(m)] for m, eu, ed in lev), zero)
    ops["hyb"] = sum((vu * (cd[up(a)] @ c[up(m)] + cd[up(m)] @ c[up(a)]) +
                      vd * (cd[dw(a)] @ c[dw(m)] + cd[dw(m)] @ c[dw(a)])
                      for a, m, vu, vd in hyb), zero)
    ops["n_up"] = [n[up(a)] for a in range(norb)]
    ops["n_dw"] = [n[dw(a)] for a in range(norb)]
    return ops


def state_observables(ops, norb, full):
    """<full| O |full> for the columns of OBS_COLS (orbital slots >= norb are 0)."""
    ev = lambda O: float(full @ (O @ full))
    out = dict.fromkeys(OBS_COLS, 0.0)
    nu, nd = ops["n_up"], ops["n_dw"]
    szo = [0.5 * (nu[a] - nd[a]) for a in range(norb)]
    nto = [nu[a] + nd[a] for a in range(norb)]
    for a in range(norb):
        out["dens_up%d" % a] = ev(nu[a])
        out["dens_dw%d" % a] = ev(nd[a])
        out["docc%d" % a] = ev(nu[a] @ nd[a])
        for b in range(norb):
            out["sz2_%d%d" % (a, b)] = ev(szo[a] @ szo[b])
            out["n2_%d%d" % (a, b)] = ev(nto[a] @ nto[b])
    stot = sum(szo[1:], szo[0])
    out["s2tot"] = ev(stot @ stot)
    for k in ("eknot_diag", "eknot_off", "epot_dd", "ehartree", "dust", "dund", "dse", "dph"):
        key = {"eknot_diag": "hloc_diag", "eknot_off": "hloc_off"}.get(k, k)
        out[k] = ev(ops[key])
    return np.array([out[k] for k in OBS_COLS])


def sector_codes(ns, nup, ndw):
    """Reference ordering: idw ascending (outer), iup ascending (inner)."""
    ups = [s for s in range(2 ** ns) if bin(s).count("1") == nup]
    dws = [s for s in range(2 ** ns) if bin(s).count("1") == ndw]
    return np.array([iu + (idn << ns) for idn in dws for iu in ups], dtype=np.int64)


def bath_vector(mdl):
    """Flat bath in the set_dmft_bath layout (dmft_aux.f90:489-512; hybrid:
    e(ispin,1,i) then v(ispin,iorb,i), :572-590)."""
    norb, nbath, nspin = mdl["norb"], mdl["nbath"], mdl["nspin"]
    if mdl.get("bath_type", "normal") == "hybrid":
        out = np.zeros(nspin * nbath + nspin * norb * nbath)
        for s in range(nspin):
            for k in range(nbath):
                out[k + s * nbath] = mdl["e"][s][k]
            for a in range(norb):
                for k in range(nbath):
                    out[nspin * nbath + k + a * nbath + s * norb * nbath] = mdl["v"][s][a][k]
        return out
    out = np.zeros(2 * nspin * norb * nbath)
    st = nspin * norb * nbath
    for s in range(nspin):
        for a in range(norb):
            for k in range(nbath):
                io = k + a * nbath + s * nbath * norb
                out[io] = mdl["e"][s][a][k]
                out[st + io] = mdl["v"][s][a][k]
    return out


def hloc_fortran(mdl):
    """Real Hloc(Nspin,Nspin,Norb,Norb) flattened in Fortran order."""
    nspin, norb = mdl["nspin"], mdl["norb"]
    h = np.zeros((nspin, nspin, norb, norb))
    for s in range(nspin):
        h[s, s] = np.array(mdl["hloc"][s])
    return h.flatten(order="F")


def make_model(rng, norb, nbath, nspin, uloc, ust=0.0, jh=0.0, jx=0.0, jp=0.0,
               xmu=0.0, hfmode=1, jhflag=0, offdiag=0.0, crystal=None, vzero=False):
    e = rng.uniform(-1.5, 1.5, size=(nspin, norb, nbath))
    v = 0.0 * e if vzero else rng.uniform(0.2, 0.7, size=(nspin, norb, nbath))
    hloc = np.zeros((nspin, norb, norb))
    for s in range(nspin):
        if crystal is not None:
            hloc[s] += np.diag(crystal)
        if offdiag != 0.0 and norb > 1:
            for a in range(norb):
                for b in range(a + 1, norb):
                    t = offdiag * (1 + 0.1 * (a + b))
                    hloc[s, a, b] = hloc[s, b, a] = t
    u = np.zeros(3)
    u[:len(uloc)] = uloc
    return dict(norb=norb, nbath=nbath, nspin=nspin, e=e.tolist(), v=v.tolist(),
                hloc=hloc.tolist(), uloc=u.tolist(), ust=ust, jh=jh, jx=jx, jp=jp,
                xmu=xmu, hfmode=hfmode, jhflag=jhflag)


def lehmann_gf(H, c, ns, mdl, beta, wm, wr, eps, finite_t, gs_threshold=1e-9,
               cutoff_states=None):
    """Exact G_{a sigma}(iw_n), G(w+i eps), the full orbital matrix G_{ab sigma}
    and E per sector via dense eigh."""
    norb, nspin = mdl["norb"], mdl["nspin"]
    spectra = {}
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            idx = sector_codes(ns, nu, nd)
            Hs = H[idx][:, idx].toarray()
            w, U = np.linalg.eigh(Hs)
            spectra[(nu, nd)] = (idx, w, U)
    egs = min(s[1][0] for s in spectra.values())
    states = []       # (E, sector, vector in sector basis)
    for key, (idx, w, U) in spectra.items():
        for k in range(len(w)):
            if finite_t:
                if cutoff_states is None or w[k] - egs < cutoff_states:
                    states.append((w[k], key, U[:, k]))
            elif abs(w[k] - egs) <= gs_threshold:
                states.append((w[k], key, U[:, k]))
    if finite_t:
        zeta = sum(np.exp(-beta * (E - egs)) for E, _, _ in states)
    else:
        zeta = float(len(states))
    gm = np.zeros((nspin, norb, len(wm)), dtype=complex)
    gr = np.zeros((nspin, norb, len(wr)), dtype=complex)
    # full orbital matrix G_ab (used for the hybrid bath)
    gmf = np.zeros((nspin, norb, norb, len(wm)), dtype=complex)
    grf = np.zeros((nspin, norb, norb, len(wr)), dtype=complex)
    dens = np.zeros(norb)
    docc = np.zeros(norb)
    ops = local_operators(mdl, c, ns)
    # the pieces add up to H (checks the decomposition used below)
    parts = (ops["hloc_diag"] + ops["hloc_off"] + ops["mu"] + ops["epot_dd"] + ops["ehartree"] +
             mdl["jx"] * ops["dse"] + mdl["jp"] * ops["dph"] + ops["bath"] + ops["hyb"])
    assert abs(parts - H).max() < 1e-12
    obs_peso, obs_vals = [], []
    for E, (nu, nd), vec in states:
        peso = (np.exp(-beta * (E - egs)) if finite_t else 1.0) / zeta
        idx = spectra[(nu, nd)][0]
        for a in range(norb):
            nua = (idx >> a) & 1
            nda = (idx >> (a + ns)) & 1
            dens[a] += peso * np.sum(np.abs(vec) ** 2 * (nua + nda))
            docc[a] += peso * np.sum(np.abs(vec) ** 2 * nua * nda)
        full = np.zeros(H.shape[0])
        full[idx] = vec
        obs_peso.append(peso)
        obs_vals.append(state_observables(ops, norb, full))
        for s in range(nspin):
            tgt = (nu + 1, nd) if s == 0 else (nu, nd + 1)
            if tgt in spectra:
                jdx, wj, Uj = spectra[tgt]
                amp = [Uj.T @ (c[a + s * ns].T @ full)[jdx] for a in range(norb)]
                for a in range(norb):
                    for b in range(norb):
                        wt = peso * amp[a] * amp[b]
                        gmf[s, a, b] += (wt[None, :] / (1j * wm[:, None] - (wj - E)[None, :])).sum(1)
                        grf[s, a, b] += (wt[None, :] / (wr[:, None] + 1j * eps - (wj - E)[None, :])).sum(1)
            tgt = (nu - 1, nd) if s == 0 else (nu, nd - 1)
            if tgt in spectra:
                jdx, wj, Uj = spectra[tgt]
                amp = [Uj.T @ (c[a + s * ns] @ full)[jdx] for a in range(norb)]
                for a in range(norb):
                    for b in range(norb):
                        wt = peso * amp[a] * amp[b]
                        gmf[s, a, b] += (wt[None, :] / (1j * wm[:, None] + (wj - E)[None, :])).sum(1)
                        grf[s, a, b] += (wt[None, :] / (wr[:, None] + 1j * eps + (wj - E)[None, :])).sum(1)
            for a in range(norb):
                p = a + s * ns
                # particle: c^+_p |E> into sector (nu+1,nd) or (nu,nd+1)
                tgt = (nu + 1, nd) if s == 0 else (nu, nd + 1)
                if tgt in spectra:
                    jdx, wj, Uj = spectra[tgt]
                    amp = Uj.T @ (c[p].T @ full)[jdx]
                    for k in range(len(wj)):
                        wt = peso * amp[k] ** 2
                        gm[s, a] += wt / (1j * wm - (wj[k] - E))
                        gr[s, a] += wt / (wr + 1j * eps - (wj[k] - E))
                tgt = (nu - 1, nd) if s == 0 else (nu, nd - 1)
                if tgt in spectra:
                    jdx, wj, Uj = spectra[tgt]
                    amp = Uj.T @ (c[p] @ full)[jdx]
                    for k in range(len(wj)):
                        wt = peso * amp[k] ** 2
                        gm[s, a] += wt / (1j * wm + (wj[k] - E))
                        gr[s, a] += wt / (wr + 1j * eps + (wj[k] - E))
    return (egs, zeta, len(states), gm, gr, dens, docc, spectra, np.array(obs_peso),
            np.array(obs_vals), gmf, grf)


def write_model(name, mdl, rng, beta=50.0, lmats=64, lreal=48, eps=0.05,
                finite_t=False, cutoff_states=None):
    H, c, ns = model_hamiltonian(mdl)
    wm = np.pi / beta * (2 * np.arange(1, lmats + 1) - 1)
    wr = np.linspace(-4.0, 4.0, lreal)
    egs, zeta, nstates, gm, gr, dens, docc, spectra, opeso, ovals, gmf, grf = lehmann_gf(
        H, c, ns, mdl, beta, wm, wr, eps, finite_t, cutoff_states=cutoff_states)
    out = dict(norb=mdl["norb"], nbath=mdl["nbath"], nspin=mdl["nspin"],
               bath=bath_vector(mdl), hloc_f=hloc_fortran(mdl),
               uloc=np.array(mdl["uloc"]), ust=mdl["ust"], jh=mdl["jh"],
               jx=mdl["jx"], jp=mdl["jp"], xmu=mdl["xmu"], hfmode=mdl["hfmode"],
               jhflag=mdl["jhflag"], beta=beta, eps=eps, wm=wm, wr=wr,
               egs=egs, zeta=zeta, nstates=nstates, gmats=gm, greal=gr,
               dens=dens, docc=docc, finite_t=int(finite_t),
               obs_peso=opeso, obs_vals=ovals)
    if cutoff_states is not None:
        out["cutoff_states"] = cutoff_states
    if mdl.get("bath_type", "normal") == "hybrid":
        out.update(bath_type="hybrid", gmats_full=gmf, greal_full=grf)
    # per-sector map, seeded H*x, lowest eigenvalues
    nsec = (ns + 1) ** 2
    maps, xs, hxs, evs = [], [], [], []
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            idx, w, _ = spectra[(nu, nd)]
            x = rng.standard_normal(len(idx))
            hx = H[idx][:, idx] @ x
            maps.append(idx)
            xs.append(x)
            hxs.append(hx)
            evs.append(w[:8])
    offs = np.cumsum([0] + [len(m) for m in maps])
    eoffs = np.cumsum([0] + [len(w) for w in evs])
    out.update(sector_offsets=offs, maps=np.concatenate(maps),
               x=np.concatenate(xs), hx=np.concatenate(hxs),
               eig_offsets=eoffs, eigs=np.concatenate(evs), nsectors=nsec)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: Ns={ns} sectors={nsec} egs={egs:.12f} nstates={nstates}")


def make_hybrid_model(rng, norb, nbath, nspin, uloc, ust=0.0, jh=0.0, xmu=0.0, hfmode=1,
                      offdiag=0.0, crystal=None):
    """Hybrid bath: Nbath levels e[s][k] shared by all orbitals, couplings v[s][a][k]."""
    e = rng.uniform(-1.5, 1.5, size=(nspin, nbath))
    v = rng.uniform(0.2, 0.7, size=(nspin, norb, nbath))
    mdl = make_model(rng, norb, 1, nspin, uloc, ust=ust, jh=jh, xmu=xmu, hfmode=hfmode,
                     offdiag=offdiag, crystal=crystal)
    mdl.update(nbath=nbath, e=e.tolist(), v=v.tolist(), bath_type="hybrid")
    return mdl


def main_hybrid():
    """Hybrid-bath fixtures (own seed, so the fixtures above are unchanged)."""
    rng = np.random.default_rng(20261203)
    # two orbitals sharing Nbath=3 levels (Ns=5), Hloc off-diagonal, T=0
    write_model("m9_hybrid_norb2_nb3", make_hybrid_model(rng, 2, 3, 1, [2.5, 2.0], ust=1.5,
                                                         jh=0.25, offdiag=0.1,
                                                         crystal=[-0.1, 0.1]), rng)
    # three orbitals sharing Nbath=2 levels (Ns=5), two spin species with their
    # own bath, no Hartree shift, xmu != 0, T=0
    write_model("m10_hybrid_norb3_nb2_spin2",
                make_hybrid_model(rng, 3, 2, 2, [2.0, 2.2, 1.8], ust=1.2, jh=0.2, xmu=0.3,
                                  hfmode=0, offdiag=0.08, crystal=[-0.1, 0.0, 0.1]), rng)


def main():
    rng = np.random.default_rng(20261202)
    # single orbital, Nbath=3 (Ns=4), hfmode, xmu != 0, T=0
    write_model("m1_norb1_nb3", make_model(rng, 1, 3, 1, [2.0], xmu=0.1), rng)
    # two orbitals, Nbath=1 (Ns=4): Hloc off-diagonal hops, Ust/Jh, crystal field
    write_model("m2_norb2_nb1", make_model(rng, 2, 1, 1, [3.0, 2.5], ust=2.0, jh=0.3,
                                            offdiag=0.15, crystal=[-0.05, 0.05]), rng)
    # two orbitals, Nbath=2 (Ns=6), Nspin=2, spin-exchange/pair-hopping, no hfmode
    write_model("m3_norb2_nb2_jh", make_model(rng, 2, 2, 2, [2.0, 2.0], ust=1.5, jh=0.25,
                                               jx=0.25, jp=0.2, xmu=0.3, hfmode=0,
                                               jhflag=1), rng)
    # three orbitals, Nbath=1 (Ns=6)
    write_model("m4_norb3_nb1", make_model(rng, 3, 1, 1, [2.0, 2.2, 2.4], ust=1.6, jh=0.2,
                                            crystal=[-0.1, 0.0, 0.1]), rng)
    # single orbital, Nbath=5 (Ns=6), Nspin=2 (spin-dependent bath), finite T
    write_model("m5_norb1_nb5_T", make_model(rng, 1, 5, 2, [2.5], xmu=-0.2), rng,
                beta=10.0, finite_t=True, cutoff_states=3.0)
    # U=0 (closed form) and V=0 (atomic limit) single-orbital models
    write_model("m6_u0", make_model(rng, 1, 4, 1, [0.0], xmu=0.0), rng)
    write_model("m7_atomic", make_model(rng, 1, 2, 1, [2.0], vzero=True), rng, beta=10.0,
                finite_t=True, cutoff_states=50.0)
    # two orbitals, Nbath=2 (Ns=6), no Hartree shift (hfmode=F), no spin-flip terms
    write_model("m8_norb2_nb2_nohf", make_model(rng, 2, 2, 1, [2.2, 1.8], ust=1.4, jh=0.2,
                                                 xmu=0.4, hfmode=0), rng)


if __name__ == "__main__":
    import sys
    if "--hybrid-only" in sys.argv:
        main_hybrid()
    else:
        main()
        main_hybrid()
