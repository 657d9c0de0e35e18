"""CPU tests of the bath fits (dmed_chi2_fit, dmed_chi2_fit_hybrid) and of the
Fortran layer build.

The reference's minimiser (SciFortran fmin_cg) is not available, so the fit's
parity is unpinned; these tests pin what the mathematics fixes: a bath whose
own function is the target is recovered to chi^2 ~ 0 from a perturbed start,
for both fit schemes."""
import os
import subprocess

import numpy as np
import pytest

import dmed
from conftest import ROOT


def delta(z, e, v):
    return (v[None, :] ** 2 / (z[:, None] - e[None, :])).sum(axis=1)


@pytest.mark.parametrize("scheme", ["delta", "weiss"])
@pytest.mark.parametrize("weight", [0, 1, 2, 3])
def test_fit_recovers_exact_bath(scheme, weight):
    beta = 50.0
    L = 400
    wm = np.pi / beta * (2 * np.arange(1, L + 1) - 1)
    e0 = np.array([-1.1, -0.4, 0.05, 0.5, 1.3])
    v0 = np.array([0.3, 0.45, 0.2, 0.5, 0.35])
    xmu, hloc = 0.1, -0.05
    d = delta(1j * wm, e0, v0)
    fg = d if scheme == "delta" else 1.0 / (1j * wm + xmu - hloc - d)
    rng = np.random.default_rng(weight)
    e_start = e0 + 0.05 * rng.standard_normal(5)
    v_start = v0 * (1 + 0.05 * rng.standard_normal(5))
    e, v, chi, it = dmed.chi2_fit(fg, beta, e_start, v_start, scheme=scheme, weight=weight,
                                  xmu=xmu, hloc=hloc, niter=2000, ftol=1e-14)
    fit = delta(1j * wm, e, v)
    fit = fit if scheme == "delta" else 1.0 / (1j * wm + xmu - hloc - fit)
    assert np.abs(fit - fg).max() < 2e-4, (chi, it)
    assert chi < 1e-6


def hybrid_target(z, e, v, hloc, xmu, scheme):
    """Delta_ab(z) = sum_k V_ak V_bk/(z - e_k), or G0 = [(z + xmu) 1 - Hloc - Delta]^-1."""
    D = np.einsum("ak,bk,nk->abn", v, v, 1.0 / (z[:, None] - e[None, :]))
    if scheme == "delta":
        return D
    no = v.shape[0]
    out = np.empty_like(D)
    for n in range(len(z)):
        out[:, :, n] = np.linalg.inv((z[n] + xmu) * np.eye(no) - hloc - D[:, :, n])
    return out


@pytest.mark.parametrize("scheme", ["delta", "weiss"])
@pytest.mark.parametrize("weight,norb", [(0, 2), (3, 2), (1, 3)])
def test_hybrid_fit_recovers_exact_bath(scheme, weight, norb):
    """dmed_chi2_fit_hybrid (chi2_fitgf_hybrid_normal): two orbitals sharing
    four levels, off-diagonal Hloc; the target is the bath's own Delta or G0
    matrix, the start a perturbed bath.  The fit must return to chi^2 ~ 0 and
    reproduce every orbital element, which needs the analytic gradient of
    both schemes (including G0 dDelta G0) to be right."""
    beta, L = 50.0, 300
    wm = np.pi / beta * (2 * np.arange(1, L + 1) - 1)
    e0 = np.array([-1.0, -0.3, 0.35, 1.2])
    v0 = np.array([[0.4, 0.25, 0.3, 0.5], [0.2, 0.45, 0.35, 0.3], [0.3, 0.2, 0.5, 0.25]])[:norb]
    hloc = np.array([[-0.1, 0.08, 0.02], [0.08, 0.05, -0.03], [0.02, -0.03, 0.12]])[:norb, :norb]
    xmu = 0.1
    fg = hybrid_target(1j * wm, e0, v0, hloc, xmu, scheme)
    rng = np.random.default_rng(10 + weight)
    e_start = e0 + 0.05 * rng.standard_normal(4)
    v_start = v0 * (1 + 0.05 * rng.standard_normal(v0.shape))
    e, v, chi, it = dmed.chi2_fit_hybrid(fg, beta, e_start, v_start, hloc, scheme=scheme,
                                         weight=weight, xmu=xmu, niter=3000, ftol=1e-14)
    fit = hybrid_target(1j * wm, e, v, hloc, xmu, scheme)
    assert np.abs(fit - fg).max() < 2e-4, (chi, it)
    assert chi < 1e-6


def test_fit_rejects_bad_arguments():
    with pytest.raises(dmed.DmedError):
        dmed.chi2_fit(np.zeros(0, complex), 10.0, [0.0], [0.1])


def test_fortran_driver_and_modules_built():
    """The Fortran DMFT_ED layer (ISO_C_BINDING over libdmed.so) is built by
    __graft_entry__.build(); its driver links against the library and its
    module exposes the reference's public names."""
    b = os.path.join(ROOT, "dmft-ed_amd", "fortran", "build")
    exe = os.path.join(b, "ed_hm_bethe")
    if not os.path.exists(exe):
        pytest.skip("Fortran layer not built (run __graft_entry__.build())")
    for e in (exe, os.path.join(b, "ed_hm_bethe_ineq")):
        out = subprocess.run(["ldd", e], capture_output=True, text=True).stdout
        assert "libdmed.so" in out and "not found" not in out.split("libdmed.so")[1].split("\n")[0]
    mod = open(os.path.join(b, "dmft_ed.mod"), "rb").read().decode("latin-1").lower()
    for name in ("ed_read_input", "ed_init_solver", "ed_solve", "ed_get_sigma_matsubara",
                 "ed_chi2_fitgf", "get_bath_dimension"):
        assert name in mod, name
    # the operator-level interface (cc_sparse_HxV shape) of module ED_HXV_GPU
    shim = open(os.path.join(b, "ed_hxv_gpu.mod"), "rb").read().decode("latin-1").lower()
    for name in ("setup_hv_sector", "delete_hv_sector", "directmatvec_cc", "cc_sparse_hxv"):
        assert name in shim, name


UTIL_CHECK = """program util_check
  use DMFT_ED
  implicit none
  real(8),allocatable :: b(:), bl(:,:)
  real(8)    :: hso(4,4), hnn(2,2,2,2), hl(8,8), hnnn(2,2,2,2,2)
  integer    :: i, nb
  call ed_read_input("inputED.in")
  nb = get_bath_dimension()
  allocate(b(nb), bl(2,nb))
  b = [(0.1d0*i, i=1,nb)]
  call spin_symmetrize_bath(b,save=.false.)
  write(*,"(A,100ES24.15)")"spin_sym ",b
  b = [(0.1d0*i, i=1,nb)]
  call break_symmetry_bath(b,0.05d0,1d0,save=.false.)
  write(*,"(A,100ES24.15)")"break ",b
  b = [(0.1d0*i, i=1,nb)]
  call ph_symmetrize_bath(b,save=.false.)
  write(*,"(A,100ES24.15)")"ph_sym ",b
  b = [(0.1d0*i, i=1,nb)]
  call ph_trans_bath(b,save=.false.)
  write(*,"(A,100ES24.15)")"ph_trans ",b
  bl(1,:) = [(0.1d0*i, i=1,nb)]; bl(2,:) = -bl(1,:)
  call spin_symmetrize_bath(bl,save=.false.)
  write(*,"(A,100ES24.15)")"spin_sym_lat ",bl(2,:)
  hso = reshape([(dble(i), i=1,16)],[4,4])
  hnn = so2nn_reshape(hso,2,2)
  write(*,"(A,100ES24.15)")"so2nn ",hnn
  write(*,"(A,100ES24.15)")"nn2so ",nn2so_reshape(hnn,2,2)
  hl = reshape([(dble(i), i=1,64)],[8,8])
  hnnn = lso2nnn_reshape(hl,2,2,2)
  write(*,"(A,100ES24.15)")"lso2nnn ",hnnn
  write(*,"(A,100ES24.15)")"nnn2lso ",nnn2lso_reshape(hnnn,2,2,2)
end program util_check
"""


def test_fortran_bath_utilities(tmp_path):
    """The host-side helpers of the Fortran API (no GPU call): the bath
    symmetry operations of ED_BATH/user_aux.f90:1071-1270 (site and lattice
    forms) and the spin/orbital/site reshapes of ED_AUX_FUNX.f90:177-351,
    compiled against the module and checked against the index rules written
    out here."""
    b = os.path.join(ROOT, "dmft-ed_amd", "fortran", "build")
    if not os.path.exists(os.path.join(b, "dmft_ed.o")):
        pytest.skip("Fortran layer not built (run __graft_entry__.build())")
    import shutil
    flang = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"
    (tmp_path / "check.f90").write_text(UTIL_CHECK)
    lib = os.path.join(ROOT, "dmft-ed_amd")
    subprocess.run([flang, "-O0", "-I", b, "check.f90", os.path.join(b, "dmed_cabi.o"),
                    os.path.join(b, "dmft_ed.o"), "-L" + lib, "-ldmed", "-Wl,-rpath," + lib,
                    "-o", "check"], cwd=str(tmp_path), check=True, capture_output=True)
    (tmp_path / "inputED.in").write_text("NORB=2\nNBATH=3\nNSPIN=2\nNREAD=0.6\nNDELTA=0.1\n")
    out = subprocess.run(["./check"], cwd=str(tmp_path), capture_output=True, text=True,
                         timeout=60, check=True).stdout
    res = {}
    for line in out.splitlines():
        parts = line.split()
        if parts and parts[0] in ("spin_sym", "break", "ph_sym", "ph_trans", "spin_sym_lat",
                                  "so2nn", "nn2so", "lso2nnn", "nnn2lso"):
            res[parts[0]] = parts[1:]
    nspin, norb, nb = 2, 2, 3
    st = nspin * norb * nb
    b0 = 0.1 * np.arange(1, 2 * st + 1)

    def ev(x):      # e(s,a,k), v(s,a,k) of the set_dmft_bath layout
        e = np.zeros((nspin, norb, nb))
        v = np.zeros_like(e)
        for s_ in range(nspin):
            for a in range(norb):
                for k in range(nb):
                    io = k + a * nb + s_ * nb * norb
                    e[s_, a, k], v[s_, a, k] = x[io], x[st + io]
        return e, v

    def flat(e, v):
        x = np.zeros(2 * st)
        for s_ in range(nspin):
            for a in range(norb):
                for k in range(nb):
                    io = k + a * nb + s_ * nb * norb
                    x[io], x[st + io] = e[s_, a, k], v[s_, a, k]
        return x

    got = {k: np.array(list(map(float, v))) for k, v in res.items()}
    e, v = ev(b0)
    e1, v1 = e.copy(), v.copy()
    e1[1], v1[1] = e[0], v[0]
    assert np.allclose(got["spin_sym"], flat(e1, v1), rtol=0, atol=1e-14)
    e1 = e.copy()
    e1[0] += 0.05
    e1[1] -= 0.05
    assert np.allclose(got["break"], flat(e1, v), rtol=0, atol=1e-14)
    e1, v1 = e.copy(), v.copy()
    e1[:, :, 2], v1[:, :, 2], e1[:, :, 1] = -e[:, :, 0], v[:, :, 0], 0.0
    assert np.allclose(got["ph_sym"], flat(e1, v1), rtol=0, atol=1e-14)
    e1, v1 = e.copy(), v.copy()
    e1[:, 0], e1[:, 1] = -e[:, 1], -e[:, 0]
    v1[:, 0], v1[:, 1] = v[:, 1], v[:, 0]
    assert np.allclose(got["ph_trans"], flat(e1, v1), rtol=0, atol=1e-14)
    e2, v2 = ev(-b0)
    e2[1], v2[1] = e2[0], v2[0]
    assert np.allclose(got["spin_sym_lat"], flat(e2, v2), rtol=0, atol=1e-14)
    hso = np.arange(1, 17, dtype=float).reshape(4, 4, order="F")
    hnn = np.zeros((2, 2, 2, 2))
    for s1 in range(2):
        for s2 in range(2):
            for a in range(2):
                for b_ in range(2):
                    hnn[s1, s2, a, b_] = hso[a + 2 * s1, b_ + 2 * s2]
    assert np.array_equal(got["so2nn"], hnn.flatten(order="F"))
    assert np.array_equal(got["nn2so"], hso.flatten(order="F"))
    hl = np.arange(1, 65, dtype=float).reshape(8, 8, order="F")
    hnnn = np.zeros((2, 2, 2, 2, 2))
    back = np.zeros((8, 8))
    for l in range(2):
        for s1 in range(2):
            for s2 in range(2):
                for a in range(2):
                    for b_ in range(2):
                        i, j = a + 2 * s1 + 4 * l, b_ + 2 * s2 + 4 * l
                        hnnn[l, s1, s2, a, b_] = hl[i, j]
                        back[i, j] = hl[i, j]
    assert np.array_equal(got["lso2nnn"], hnnn.flatten(order="F"))
    assert np.array_equal(got["nnn2lso"], back.flatten(order="F"))



COMP_CHECK = """program comp_check
  use DMFT_ED
  implicit none
  real(8),allocatable    :: b(:), b2(:), a3(:,:,:), a2(:,:), a1(:)
  complex(8)             :: m(4,4), r(4,4), back(4,4)
  complex(8),allocatable :: f5(:,:,:,:,:), f3(:,:,:), f1(:), fl(:,:)
  real(8)                :: phi(2), p1
  integer                :: n, i, j, d3(3), d2(2)
  character(len=7)       :: bt
  read(*,*) bt
  Norb = 2; Nspin = 2; Nbath = 3; bath_type = bt
  n = get_bath_dimension()
  allocate(b(n), b2(n), a1(Nbath))
  b  = [(0.1d0*i, i=1,n)]
  b2 = [(-1d0*i, i=1,n)]
  d3 = get_component_bath_dimension(1)
  write(*,"(A,3I4)") "dim_e", d3
  allocate(a3(d3(1),d3(2),d3(3)))
  call get_component_bath(a3,b,1)
  write(*,"(A,100ES24.15)") "get_e", a3
  deallocate(a3)
  d3 = get_component_bath_dimension(2)
  write(*,"(A,3I4)") "dim_v", d3
  allocate(a3(d3(1),d3(2),d3(3)))
  call get_component_bath(a3,b,2)
  write(*,"(A,100ES24.15)") "get_v", a3
  d2 = get_spin_component_bath_dimension(2)
  allocate(a2(d2(1),d2(2)))
  call get_component_bath(a2,b,2,2)
  write(*,"(A,100ES24.15)") "get_v_s2", a2
  call get_component_bath(a1,b,1,2,2)
  write(*,"(A,100ES24.15)") "get_e_s2a2", a1
  a1 = [7d0,8d0,9d0]
  call set_component_bath(a1,b,2,1,2)
  write(*,"(A,100ES24.15)") "set_v_s1a2", b
  call copy_component_bath(b2,1,b,2)
  write(*,"(A,100ES24.15)") "copy_s1_s2", b
  call copy_component_bath(b2,2,1,b,1,2,1)
  write(*,"(A,100ES24.15)") "copy_e_s2a1_s1a2", b
  call copy_component_bath(b2,b,2)
  write(*,"(A,100ES24.15)") "copy_v_full", b
  a3 = 100d0 + a3
  call set_component_bath(a3,b,2)
  write(*,"(A,100ES24.15)") "set_v_full", b
  do i=1,4
     do j=1,4
        m(i,j) = cmplx(10*i+j, -j, kind=8)
     enddo
  enddo
  r = so2os_reshape(m,2,2)
  back = os2so_reshape(r,2,2)
  write(*,"(A,100ES24.15)") "so2os", ((real(r(i,j)),j=1,4),i=1,4)
  write(*,"(A,ES24.15)") "roundtrip", maxval(abs(back-m))
  allocate(f5(2,2,2,2,5), f3(4,4,5), f1(5), fl(3,5))
  f5 = (1d0,1d0); f3 = (1d0,1d0); f1 = (1d0,1d0); fl = (1d0,1d0); phi = 1d0; p1 = 1d0
  call ed_get_self_matsubara(f5)
  call ed_get_self_real(f3)
  call ed_get_fimp_matsubara(f1,1,2,2,1)
  call ed_get_fimp_real(fl,3,1,1,1,1)
  call ed_get_phisc(phi)
  call ed_get_phisc(p1,2)
  write(*,"(A,5ES24.15)") "anom", maxval(abs(f5)), maxval(abs(f3)), maxval(abs(f1)), &
       maxval(abs(fl)), maxval(abs(phi))+abs(p1)
end program comp_check
"""


def test_fortran_bath_components(tmp_path):
    """get/set/copy_component_bath and the component dimensions
    (ED_BATH/user_aux.f90:96-894) for the normal and the hybrid bath,
    so2os/os2so_reshape (ED_AUX_FUNX.f90:354-404) and the anomalous getters
    (zero in the normal mode), against the index rules of the flat bath
    vector written out here: normal e(s,a,k) at k + a*Nb + s*Nb*Norb and v
    after Nspin*Norb*Nb; hybrid e(s,k) at k + s*Nb and v(s,a,k) at
    Nspin*Nb + k + a*Nb + s*Norb*Nb (0-based)."""
    b_ = os.path.join(ROOT, "dmft-ed_amd", "fortran", "build")
    if not os.path.exists(os.path.join(b_, "dmft_ed.o")):
        pytest.skip("Fortran layer not built (run __graft_entry__.build())")
    import shutil
    flang = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"
    (tmp_path / "check.f90").write_text(COMP_CHECK)
    lib = os.path.join(ROOT, "dmft-ed_amd")
    subprocess.run([flang, "-O0", "-I", b_, "check.f90", os.path.join(b_, "dmed_cabi.o"),
                    os.path.join(b_, "dmft_ed.o"), "-L" + lib, "-ldmed", "-Wl,-rpath," + lib,
                    "-o", "check"], cwd=str(tmp_path), check=True, capture_output=True)
    ns, no, nb = 2, 2, 3
    for bt in ("normal", "hybrid"):
        out = subprocess.run(["./check"], cwd=str(tmp_path), input=bt + "\n",
                             capture_output=True, text=True, timeout=60, check=True).stdout
        res = {l.split()[0]: np.array([float(x) for x in l.split()[1:]])
               for l in out.splitlines() if l.strip()}
        ne = 1 if bt == "hybrid" else no

        def split(x):
            e = np.zeros((ns, ne, nb))
            v = np.zeros((ns, no, nb))
            st = ns * ne * nb
            for s in range(ns):
                for a in range(ne):
                    for k in range(nb):
                        e[s, a, k] = x[k + a * nb + s * ne * nb]
                for a in range(no):
                    for k in range(nb):
                        v[s, a, k] = x[st + k + a * nb + s * no * nb]
            return e, v

        def join(e, v):
            x = np.zeros(ns * ne * nb + ns * no * nb)
            st = ns * ne * nb
            for s in range(ns):
                for a in range(ne):
                    for k in range(nb):
                        x[k + a * nb + s * ne * nb] = e[s, a, k]
                for a in range(no):
                    for k in range(nb):
                        x[st + k + a * nb + s * no * nb] = v[s, a, k]
            return x

        n = ns * ne * nb + ns * no * nb
        b = 0.1 * np.arange(1, n + 1)
        b2 = -1.0 * np.arange(1, n + 1)
        assert np.array_equal(join(*split(b)), b)
        e, v = split(b)
        e2, v2 = split(b2)
        assert list(res["dim_e"]) == [ns, ne, nb] and list(res["dim_v"]) == [ns, no, nb]
        assert np.allclose(res["get_e"], e.ravel(order="F"), rtol=0, atol=1e-15)
        assert np.allclose(res["get_v"], v.ravel(order="F"), rtol=0, atol=1e-15)
        assert np.allclose(res["get_v_s2"], v[1].ravel(order="F"), rtol=0, atol=1e-15)
        assert np.allclose(res["get_e_s2a2"], e[1, min(1, ne - 1)], rtol=0, atol=1e-15)
        v[0, 1] = [7, 8, 9]
        assert np.allclose(res["set_v_s1a2"], join(e, v), rtol=0, atol=1e-15)
        e[1], v[1] = e2[0], v2[0]
        assert np.allclose(res["copy_s1_s2"], join(e, v), rtol=0, atol=1e-15)
        e[0, min(1, ne - 1)] = e2[1, 0]
        assert np.allclose(res["copy_e_s2a1_s1a2"], join(e, v), rtol=0, atol=1e-15)
        v = v2.copy()
        assert np.allclose(res["copy_v_full"], join(e, v), rtol=0, atol=1e-15)
        v = 100.0 + split(0.1 * np.arange(1, n + 1))[1]
        assert np.allclose(res["set_v_full"], join(e, v), rtol=0, atol=1e-13)
    # so2os: g(a + s*Norb, b + s'*Norb) = fg(s + a*Nspin, s' + b*Nspin), 0-based
    m = np.array([[10 * (i + 1) + (j + 1) for j in range(4)] for i in range(4)], float)
    g = np.zeros((4, 4))
    for s1 in range(2):
        for s2 in range(2):
            for a in range(2):
                for bb in range(2):
                    g[a + s1 * 2, bb + s2 * 2] = m[s1 + a * 2, s2 + bb * 2]
    assert np.array_equal(res["so2os"], g.ravel())
    assert res["roundtrip"][0] == 0.0
    assert np.all(res["anom"] == 0.0)
