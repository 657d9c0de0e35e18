"""Wall time of one ed_solve with and without the susceptibilities, and of a
hybrid-bath solve against a normal-bath solve of the same size.

    python dmft-ed_amd/dmed/chi_bench.py [--lmats 1024] [--lreal 1000] [--reps 3]

Cases:
  - workload 2 (Ns = 12, one orbital, T ~ 0): no chi, spin chi, spin and
    charge chi;
  - two orbitals, Ns = 12 (Nbath = 5 per orbital, normal bath) and the same
    Ns as a hybrid bath (Nbath = 10 levels shared by both orbitals), with
    Hloc off-diagonal so that the hybrid solve runs its mixed chains.
Prints one line per case: the best wall time of --reps solves, the
diagonalisation and Green's-function times reported by the solve, and E0.
"""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import dmed                      # noqa: E402
from dmed import configs         # noqa: E402


def _timed(norb, nspin, ns, p, reps):
    best, r = None, None
    for _ in range(reps):
        t = time.time()
        r = dmed.solve(norb, nspin, ns, p)
        dt = time.time() - t
        best = dt if best is None else min(best, dt)
    return best, r


def _line(name, best, r):
    print(f"{name:<44s} wall {best:7.3f} s  diag {r['t_diag']:7.3f} s  gf+chi {r['t_gf']:7.3f} s  "
          f"E0 {r['egs']:.12f}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lmats", type=int, default=1024)
    ap.add_argument("--lreal", type=int, default=1000)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    dmed.init(0)
    kw = dict(lmats=a.lmats, lreal=a.lreal, lanc_ngfiter=200, ltau=1000)

    c = configs.config(2)
    ns = dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"],
                        c["ust"], c["jh"], 0.0, 0.0, c["xmu"])
    for name, extra in (("ns12 norb1: G only", {}),
                        ("ns12 norb1: G + spin chi", dict(chi_spin=1)),
                        ("ns12 norb1: G + spin + charge chi", dict(chi_spin=1, chi_dens=1))):
        p = dmed.default_params(beta=c["beta"], **kw, **extra)
        dmed.solve(c["norb"], c["nspin"], ns, p)              # warm-up (sector builds, tuning)
        _line(name, *_timed(c["norb"], c["nspin"], ns, p, a.reps))

    rng = np.random.default_rng(5)
    norb, nspin = 2, 1
    hloc = np.zeros((nspin, nspin, norb, norb))
    hloc[0, 0, 0, 1] = hloc[0, 0, 1, 0] = 0.1
    hloc[0, 0, 0, 0], hloc[0, 0, 1, 1] = -0.1, 0.1
    uloc, ust, jh = [2.5, 2.0], 1.5, 0.25
    p = dmed.default_params(beta=100.0, **kw)
    # normal bath: e(a,k), v(a,k) per orbital, Nbath = 5
    nb = 5
    e = rng.uniform(-1.5, 1.5, size=(norb, nb))
    v = rng.uniform(0.2, 0.7, size=(norb, nb))
    bath = np.concatenate([e.ravel(), v.ravel()])
    ns = dmed.set_model(norb, nb, nspin, bath, hloc, uloc, ust, jh)
    dmed.solve(norb, nspin, ns, p)
    _line("ns12 norb2: normal bath (Nbath 5)", *_timed(norb, nspin, ns, p, a.reps))
    # hybrid bath: e(k) shared, v(a,k), Nbath = 10
    nb = 10
    e = rng.uniform(-1.5, 1.5, size=nb)
    v = rng.uniform(0.2, 0.7, size=(norb, nb))
    bath = np.concatenate([e, v.ravel()])
    ns = dmed.set_model(norb, nb, nspin, bath, hloc, uloc, ust, jh, bath_type="hybrid")
    dmed.solve(norb, nspin, ns, p)
    _line("ns12 norb2: hybrid bath (Nbath 10)", *_timed(norb, nspin, ns, p, a.reps))


if __name__ == "__main__":
    main()
