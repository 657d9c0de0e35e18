"""Seeded synthetic impurity models of the five benchmark workloads.

Every bath is drawn from a SplitMix64 stream (seed 1001..1005) mapped to
U[0,1) doubles, with Box-Muller for normal deviates, so any implementation
can regenerate the identical bath.  `write_restart` emits the bath in the
reference's ``hamiltonian.restart`` text format (ED_BATH/dmft_aux.f90:158-177:
a comment line, then one line per bath level k holding
(e(ispin,iorb,k), v(ispin,iorb,k), iorb=1..Norb), ispin=1..Nspin).
"""
import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def bath_vector(e, v):
    """e, v: arrays [nspin][norb][nbath] -> flat set_dmft_bath layout."""
    e = np.asarray(e, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    nspin, norb, nbath = e.shape
    out = np.zeros(2 * nspin * norb * nbath)
    st = nspin * norb * nbath
    for s in range(nspin):
        for a in range(norb):
            for k in range(nbath):
                io = k + a * nbath + s * nbath * norb
                out[io] = e[s, a, k]
                out[st + io] = v[s, a, k]
    return out


def hloc_fortran(nspin, norb, diag=None):
    h = np.zeros((nspin, nspin, norb, norb))
    if diag is not None:
        for s in range(nspin):
            for a in range(norb):
                h[s, s, a, a] = diag[a]
    return h.flatten(order="F")


def config(n):
    """Model dict of workload n (1..5)."""
    rng = SplitMix64(1000 + n)
    if n == 1:
        nb = 7
        e = [rng.uniform(-1.0, 1.0) for _ in range(nb)]
        v = [0.5 + 0.05 * rng.normal() for _ in range(nb)]
        return dict(name="ns8_norb1_half_filled_4x4", norb=1, nbath=nb, nspin=1,
                    bath=bath_vector([[e]], [[v]]), hloc=hloc_fortran(1, 1), uloc=[2.0],
                    ust=0.0, jh=0.0, xmu=0.0, beta=100.0, lmats=4096, sector=(4, 4),
                    nlanc=200)
    if n in (2, 4):
        eps = [rng.uniform(0.0, 2.0) for _ in range(5)]
        eps = [x if x > 0 else 1e-3 for x in eps]
        vv = [rng.uniform(0.2, 0.6) for _ in range(5)]
        v0 = rng.uniform(0.2, 0.6)
        u = [1.0, 2.0, 3.0][min(2, int(rng.uniform() * 3))]
        e = [-x for x in eps] + [0.0] + eps
        v = vv + [v0] + vv
        name = "ns12_norb1_half_filled_6x6" if n == 2 else "ns12_norb1_all_169_sectors"
        return dict(name=name, norb=1, nbath=11, nspin=1, bath=bath_vector([[e]], [[v]]),
                    hloc=hloc_fortran(1, 1), uloc=[u], ust=0.0, jh=0.0, xmu=0.0,
                    beta=200.0 if n == 2 else 50.0, lmats=4096, sector=(6, 6), nlanc=300)
    if n == 3:
        nb = 6
        e = [[rng.uniform(-2.0, 2.0) for _ in range(nb)] for _ in range(2)]
        v = [[rng.uniform(0.1, 0.5) for _ in range(nb)] for _ in range(2)]
        J = 0.25
        return dict(name="ns14_norb2_kanamori_7x7", norb=2, nbath=nb, nspin=1,
                    bath=bath_vector([e], [v]), hloc=hloc_fortran(1, 2, [-0.05, 0.05]),
                    uloc=[3.0, 3.0], ust=3.0 - 2 * J, jh=J, xmu=0.0, beta=100.0,
                    lmats=4096, sector=(7, 7), nlanc=400)
    if n == 5:
        nb = 15
        e = [rng.uniform(-2.0, 2.0) for _ in range(nb)]
        v = [rng.uniform(0.1, 0.5) for _ in range(nb)]
        return dict(name="ns16_norb1_half_filled_8x8", norb=1, nbath=nb, nspin=1,
                    bath=bath_vector([[e]], [[v]]), hloc=hloc_fortran(1, 1), uloc=[2.5],
                    ust=0.0, jh=0.0, xmu=0.0, beta=100.0, lmats=4096, sector=(8, 8),
                    nlanc=500)
    raise ValueError(n)


def write_restart(path, cfg):
    """Bath in the reference's hamiltonian.restart format (dmft_aux.f90:158-177)."""
    norb, nbath, nspin = cfg["norb"], cfg["nbath"], cfg["nspin"]
    b = cfg["bath"]
    st = nspin * norb * nbath
    with open(path, "w") as f:
        f.write(" ".join(f"{'#Ek_l%d_s%d' % (a + 1, s + 1):>21s} {'Vk_l%d_s%d' % (a + 1, s + 1):>21s}"
                         for s in range(nspin) for a in range(norb)) + "\n")
        for k in range(nbath):
            vals = []
            for s in range(nspin):
                for a in range(norb):
                    io = k + a * nbath + s * nbath * norb
                    vals += [b[io], b[st + io]]
            f.write(" ".join(f"{x:21.12f}" for x in vals) + "\n")
