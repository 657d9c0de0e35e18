#!/usr/bin/env python3
"""Benchmark of the impurity-solver hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--reps R]

A step is one FP64 sector H*v over the Ns=16 half-filled sector 8/8
(dim 12870 x 12870 = 165,636,900; Norb=1, Nbath=15, U=2.5, seeded bath),
inputs resident in HBM.  Prints ONE JSON line with the metric, the roofline
of the H*v kernel (algorithmic bytes 16*dim per launch: x read once, y written
once) measured with HIP events on the library stream, and the CPU baseline:
the oracle's complex(8) restatement of the reference H*v on the host cores,
on all logical CPUs and on one thread (a bounded row sample).

The second half of the metric, the solve time of the batched mixed-precision
form against the plain FP64 GPU path, is measured by default on two
workloads, each form repeated R times (speed-up = ratio of the medians):
  - the all-sector Ns=12 solve at beta=50 (169 sectors, 200-step chains);
  - the Ns=16 sector pipeline (ground state to 1e-12 and two 500-step chains).
Their errors against the FP64 path (E0, n, d, G(iw), Sigma(iw)) are reported
beside the speed-ups.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dmft-ed_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, samples):
    """Oracle H*v (combined map + binary search + complex(8), the reference
    algorithm) on rows [0, rows) of the sector, for each (rows, threads) in
    `samples` (rows may be a function of the previous samples' results);
    returns [(rows, threads, rows/s, seconds)]."""
    import ctypes as C
    import numpy as np
    import edref
    edref.build()
    edref.set_model(cfg["norb"], cfg["nbath"], cfg["nspin"], cfg["bath"], cfg["hloc"],
                    cfg["uloc"], cfg["ust"], cfg["jh"], 0.0, 0.0, cfg["xmu"], True, False)
    nu, nd = cfg["sector"]
    t0 = time.time()
    smap = edref.build_sector(nu, nd)
    log(f"cpu baseline: map of {len(smap)} codes built in {time.time() - t0:.1f}s")
    x = np.empty(len(smap), dtype=np.complex128)
    x.real = np.random.default_rng(1).standard_normal(len(smap))
    x.imag = 0.0
    L = edref.lib()
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int64)
    out = []
    for rows, threads in samples:
        if callable(rows):
            rows = rows(out)
        y = np.zeros(rows, dtype=np.complex128)
        t0 = time.perf_counter()
        rc = L.ref_hxv_block(smap.ctypes.data_as(ip), len(smap),
                             x.view(np.float64).ctypes.data_as(dp),
                             y.view(np.float64).ctypes.data_as(dp), 0, rows, threads)
        dt = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"oracle H*v failed ({rc})")
        out.append((rows, threads, rows / dt, dt))
    return out


def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


def solve_compare(reps=3):
    """Workload 4: Ns=12, all 169 sectors, beta=50, Boltzmann-weighted states.
    Plain FP64 (serial sectors and chains) against the batched mixed form
    (small sectors in one persistent work-queue launch, the rest on
    concurrent per-stream workers; chains of 32 FP64 steps, then FP32 with
    FP64 reductions).  The two forms alternate `reps` times; the speed-up is
    the ratio of the median wall times."""
    import numpy as np
    import dmed
    from dmed import configs
    c = configs.config(4)
    dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"],
                   c["ust"], c["jh"], 0.0, 0.0, c["xmu"], True, False)
    kw = dict(beta=c["beta"], lmats=c["lmats"], lreal=1000, lanc_ngfiter=200,
              lanc_nstates_sector=4, lanc_nstates_total=16)
    forms = (("fp64_plain", dict(batched=0, prec_gf=dmed.FP64)),
             ("mixed_batched", dict(batched=1, prec_gf=dmed.MIXED)))
    times = {n: [] for n, _ in forms}
    res = {}
    out = {}
    for rep in range(reps):
        for name, extra in forms:
            t0 = time.perf_counter()
            r = dmed.solve(1, 1, 12, dmed.default_params(**kw, **extra))
            dt = time.perf_counter() - t0
            res[name] = r
            times[name].append({"seconds": dt, "t_diag": r["t_diag"], "t_gf": r["t_gf"]})
            log(f"solve {name} #{rep}: {dt:.2f}s (diag {r['t_diag']:.2f}s, gf {r['t_gf']:.2f}s), "
                f"{r['nstates']} states")
    for name, _ in forms:
        out[name] = {"seconds_median": _median([t["seconds"] for t in times[name]]),
                     "t_diag_median": _median([t["t_diag"] for t in times[name]]),
                     "t_gf_median": _median([t["t_gf"] for t in times[name]]),
                     "runs": times[name], "nstates": res[name]["nstates"]}
    a64, amx = res["fp64_plain"], res["mixed_batched"]
    out["speedup"] = out["fp64_plain"]["seconds_median"] / out["mixed_batched"]["seconds_median"]
    out["errors_vs_fp64"] = {
        "E0_rel": abs(amx["egs"] - a64["egs"]) / abs(a64["egs"]),
        "dens": float(np.abs(amx["dens"] - a64["dens"]).max()),
        "docc": float(np.abs(amx["docc"] - a64["docc"]).max()),
        "G_iw": float(np.abs(amx["gmats"] - a64["gmats"]).max()),
        "Sigma_iw": float(np.abs(amx["smats"] - a64["smats"]).max()),
    }
    log(f"mixed/batched speedup {out['speedup']:.2f}x (median of {reps}), "
        f"errors {out['errors_vs_fp64']}")
    return out


def sector_pipeline_gpu(c):
    """Ground state of the workload's sector to residual 1e-12 (FP64), then
    the particle and hole chains of every orbital (FP64, c["nlanc"] steps):
    seconds, E0 and the chain lengths."""
    import dmed
    nu, nd = c["sector"]
    s = dmed.Sector(nu, nd)
    t0 = time.perf_counter()
    e0, psi, r0, rc = s.ground_state(prec=dmed.FP64, tol=1e-12, maxit=6000)
    steps = []
    for a in range(1, c["norb"] + 1):
        for dag, dn in ((1, 1), (0, -1)):
            dst = dmed.Sector(nu + dn, nd)
            v, n2 = dmed.apply_op(psi, a, dag, dst)
            al, bl, n = dmed.chain(dst, v, c["nlanc"], prec=dmed.FP64)
            steps.append(n)
            v.free()
            dst.free()
    dt = time.perf_counter() - t0
    psi.free()
    s.free()
    return dt, e0, steps


def sector_pipeline_oracle(c, threads):
    """The same pipeline on the oracle: the reference's plain Lanczos ground
    state (lanczos_plain_c stop rule) and sp_lanc_tridiag chains, on `threads`
    OpenMP threads."""
    import numpy as np
    import edref
    nu, nd = c["sector"]
    m = edref.build_sector(nu, nd)
    t0 = time.perf_counter()
    v0 = np.random.default_rng(7).standard_normal(len(m))
    egs, psi = edref.lanczos_gs(m, v0 / np.linalg.norm(v0), nitermax=1000, nthreads=threads)
    for a in range(1, c["norb"] + 1):
        for dag, dn in ((1, 1), (0, -1)):
            mj = edref.build_sector(nu + dn, nd)
            v = edref.apply_op(m, psi, mj, a, dag)
            edref.lanc_tridiag(mj, v / np.linalg.norm(v), c["nlanc"], nthreads=threads)
    return time.perf_counter() - t0, float(egs[0])


def configs_table(threads):
    """Workloads 1-3 beside the oracle on the host cores: GPU FP64 H*v of the
    ground-state sector (HIP events) against one full oracle H*v (the
    reference's map + binary-search algorithm, `threads` OpenMP threads); for
    workload 1 the whole ed_solve on both (T=0, Lmats=4096, 200-step chains);
    for workloads 2 and 3 the sector pipeline (ground state, then the particle
    and hole chains of every orbital) on the GPU, and for workload 2 on the
    oracle too (workload 3's would take about half an hour: its H*v count
    times the oracle H*v time gives the estimate)."""
    import ctypes as C
    import numpy as np
    import dmed
    import edref
    from dmed import configs
    edref.build()
    out = {}
    for n in (1, 2, 3):
        c = configs.config(n)
        args = (c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"], c["ust"],
                c["jh"], 0.0, 0.0, c["xmu"], True, False)
        dmed.set_model(*args)
        edref.set_model(*args)
        nu, nd = c["sector"]
        s = dmed.Sector(nu, nd)
        x = dmed.Vec(s, dmed.FP64)
        y = dmed.Vec(s, dmed.FP64)
        xh = np.random.default_rng(n).standard_normal(s.dim)
        x.upload(xh)
        L = dmed.lib()
        kms = C.c_double()
        dmed._check(L.dmed_hxv_timed(s.h, x.ptr(), y.ptr(), dmed.FP64, 3, C.byref(kms)))
        dmed._check(L.dmed_hxv_timed(s.h, x.ptr(), y.ptr(), dmed.FP64, 20, C.byref(kms)))
        m = edref.build_sector(nu, nd)
        t0 = time.perf_counter()
        edref.hxv(m, xh, nthreads=threads)
        tc = time.perf_counter() - t0
        row = {"workload": c["name"], "dim": s.dim, "gpu_hxv_ms": kms.value,
               "gpu_hxv_GBs": 16 * s.dim / kms.value / 1e6, "cpu_hxv_ms": 1e3 * tc,
               "cpu_threads": threads}
        x.free()
        y.free()
        s.free()
        if n == 1:
            kw = dict(beta=c["beta"], lmats=c["lmats"], lreal=200, lanc_ngfiter=200)
            t0 = time.perf_counter()
            r = dmed.solve(1, 1, 8, dmed.default_params(**kw))
            tg = time.perf_counter() - t0
            t0 = time.perf_counter()
            o = edref.ed_solve(1, c["nbath"], 1, c["bath"], c["hloc"], c["uloc"], 0.0, 0.0, 0.0,
                               True, params=edref.default_params(**kw))
            to = time.perf_counter() - t0
            row.update(gpu_solve_s=tg, cpu_solve_s=to, cpu_solve_kind="oracle ed_solve "
                       "(dense/eigsh sector eigensolves, C chains, 1 thread)",
                       solve_E0_rel=abs(r["egs"] - o["egs"]) / abs(o["egs"]),
                       solve_Sigma_iw=float(np.abs(r["smats"][0, 0, 0, 0] - o["smats"][0, 0]).max()))
        if n in (2, 3):
            dmed.set_model(*args)
            tg, e0, steps = sector_pipeline_gpu(c)
            row.update(gpu_pipeline_s=tg, gpu_pipeline_E0=e0, chain_steps=steps)
            if n == 2:
                edref.set_model(*args)
                to, eo = sector_pipeline_oracle(c, threads)
                row.update(cpu_pipeline_s=to, cpu_pipeline_threads=threads,
                           pipeline_E0_rel=abs(e0 - eo) / abs(eo))
        log(f"{c['name']}: GPU H*v {row['gpu_hxv_ms']:.4f} ms, oracle H*v {row['cpu_hxv_ms']:.1f} ms"
            + (f"; solve GPU {row['gpu_solve_s']:.2f}s vs oracle {row['cpu_solve_s']:.2f}s"
               if n == 1 else "")
            + (f"; sector pipeline GPU {row['gpu_pipeline_s']:.2f}s" if n in (2, 3) else "")
            + (f" vs oracle {row['cpu_pipeline_s']:.1f}s (dE0 {row['pipeline_E0_rel']:.1e})"
               if n == 2 else ""))
        out[f"workload{n}"] = row
    return out


def lanczos_step_rate(sec, D, prec, nsteps=40):
    """Wall time per plain Lanczos step (fused H*v epilogue + update + two
    one-block reductions) from a random start on the benchmark sector."""
    import numpy as np
    import dmed
    v0 = dmed.Vec(sec, dmed.FP64)
    v0.upload(np.random.default_rng(2).standard_normal(D))
    dmed.chain(sec, v0, 4, prec=prec)
    t0 = time.perf_counter()
    a, b, n = dmed.chain(sec, v0, nsteps, prec=prec)
    dt = (time.perf_counter() - t0) / max(n, 1)
    v0.free()
    esz = 8 if prec == dmed.FP64 else 4
    return {"ms_per_step": 1e3 * dt, "algorithmic_GBs": 6 * esz * D / dt / 1e9,
            "frac": 6 * esz * D / dt / 1e9 / HBM_PEAK_GBS, "steps": n}


def solve5_compare(reps=3):
    """Workload 5 sector pipeline (Ns=16, sector 8/8, dim 1.66e8): ground
    state to residual 1e-12, particle and hole chains of 500 steps from it,
    G(iw) and Sigma(iw) on 4096 Matsubara points at beta=100.  Plain FP64
    against the mixed form (ground state started on FP32 vectors and finished
    in FP64; chains of 32 FP64 steps, then FP32).  The forms alternate `reps`
    times after one untimed warm-up of each, which backs every work buffer the
    pipeline uses; the speed-up is the ratio of the median wall times."""
    import numpy as np
    import dmed
    from dmed import configs
    c = configs.config(5)
    dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"],
                   c["ust"], c["jh"], 0.0, 0.0, c["xmu"], True, False)
    s, sp, sh = dmed.Sector(8, 8), dmed.Sector(9, 8), dmed.Sector(7, 8)
    wm = np.pi / c["beta"] * (2 * np.arange(c["lmats"]) + 1)
    z = 1j * wm
    forms = (("fp64_plain", dmed.FP64, dmed.FP64), ("mixed", dmed.MIXED, dmed.MIXED),
             ("fp64_gs_mixed_chains", dmed.FP64, dmed.MIXED))

    def run(pgs, pgf, nlanc):
        t0 = time.perf_counter()
        e0, psi, r0, rc = s.ground_state(prec=pgs, tol=1e-12, maxit=6000)
        t_gs = time.perf_counter() - t0
        dens, docc = dmed.observables(psi, 1.0, 1)
        g = np.zeros(c["lmats"], complex)
        gr = np.zeros(1, complex)
        t1 = time.perf_counter()
        for dst, dag, sgn in ((sp, 1, 1), (sh, 0, -1)):
            v, n2 = dmed.apply_op(psi, 1, dag, dst)
            a, b, n = dmed.chain(dst, v, nlanc, prec=pgf)
            dmed.gf_add_chain(n2, e0, a, b, sgn, wm, np.zeros(1), 0.01, g, gr)
            v.free()
        t_gf = time.perf_counter() - t1
        psi.free()
        sig = dmed.sigma(0, 0, z, g)
        return dict(e0=e0, dens=dens, docc=docc, g=g, s=sig, r0=r0, rc=rc,
                    seconds=time.perf_counter() - t0, t_gs=t_gs, t_gf=t_gf)

    for _, pgs, pgf in forms[:2]:            # warm-up: back the buffers once
        run(pgs, pgf, 40)
    times = {n: [] for n, _, _ in forms}
    res = {}
    for rep in range(reps):
        for name, pgs, pgf in forms:
            if name == "fp64_gs_mixed_chains" and rep > 0:
                continue
            r = run(pgs, pgf, c["nlanc"])
            res[name] = r
            times[name].append({k: r[k] for k in ("seconds", "t_gs", "t_gf")})
            log(f"workload 5 {name} #{rep}: {r['seconds']:.2f}s (GS {r['t_gs']:.2f}s, chains "
                f"{r['t_gf']:.2f}s), E0 {r['e0']:.12f}, residual {r['r0']:.2e}")
    out = {}
    for name, _, _ in forms:
        out[name] = {"seconds_median": _median([t["seconds"] for t in times[name]]),
                     "t_gs_median": _median([t["t_gs"] for t in times[name]]),
                     "t_gf_median": _median([t["t_gf"] for t in times[name]]),
                     "runs": times[name], "E0": res[name]["e0"],
                     "gs_residual": res[name]["r0"], "gs_rc": res[name]["rc"]}
    a64, amx = res["fp64_plain"], res["mixed"]
    out["speedup"] = out["fp64_plain"]["seconds_median"] / out["mixed"]["seconds_median"]
    out["errors_vs_fp64"] = {
        "E0_rel": abs(amx["e0"] - a64["e0"]) / abs(a64["e0"]),
        "dens": float(np.abs(amx["dens"] - a64["dens"]).max()),
        "docc": float(np.abs(amx["docc"] - a64["docc"]).max()),
        "G_iw": float(np.abs(amx["g"] - a64["g"]).max()),
        "Sigma_iw": float(np.abs(amx["s"] - a64["s"]).max()),
    }
    b = res["fp64_gs_mixed_chains"]
    out["errors_vs_fp64_mixed_chains_only"] = {
        "G_iw": float(np.abs(b["g"] - a64["g"]).max()),
        "Sigma_iw": float(np.abs(b["s"] - a64["s"]).max()),
    }
    log(f"workload 5 mixed speedup {out['speedup']:.2f}x (median of {reps}), errors "
        f"{out['errors_vs_fp64']}; mixed chains alone {out['errors_vs_fp64_mixed_chains_only']}")
    for v in (s, sp, sh):
        v.free()
    return out


def chi_solves(reps=3, lmats=1024, lreal=1000):
    """Wall time of one ed_solve with and without the susceptibilities
    (workload 2: Ns = 12, one orbital), and of a hybrid-bath solve beside a
    normal-bath solve of the same size (two orbitals, Ns = 12, Hloc
    off-diagonal so that the hybrid solve runs its mixed chains).  Best of
    `reps` solves after a warm-up, with the diagonalisation and G/chi times."""
    import numpy as np
    import dmed
    from dmed import configs
    kw = dict(lmats=lmats, lreal=lreal, lanc_ngfiter=200, ltau=1000)
    out = {}

    def timed(name, norb, nspin, ns, p):
        dmed.solve(norb, nspin, ns, p)                        # warm-up: sector builds, tuning
        best, r = None, None
        for _ in range(reps):
            t = time.perf_counter()
            r = dmed.solve(norb, nspin, ns, p)
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        out[name] = {"wall_s": best, "t_diag": r["t_diag"], "t_gf_chi": r["t_gf"], "E0": r["egs"]}
        log(f"{name:<40s} wall {best:7.3f} s  diag {r['t_diag']:7.3f} s  "
            f"gf+chi {r['t_gf']:7.3f} s  E0 {r['egs']:.12f}")

    c = configs.config(2)
    ns = dmed.set_model(c["norb"], c["nbath"], c["nspin"], c["bath"], c["hloc"], c["uloc"],
                        c["ust"], c["jh"], 0.0, 0.0, c["xmu"])
    for name, extra in (("ns12_norb1_G", {}), ("ns12_norb1_G_spin_chi", dict(chi_spin=1)),
                        ("ns12_norb1_G_spin_charge_chi", dict(chi_spin=1, chi_dens=1))):
        timed(name, c["norb"], c["nspin"], ns, dmed.default_params(beta=c["beta"], **kw, **extra))
    rng = np.random.default_rng(5)
    norb, nspin = 2, 1
    hloc = np.zeros((nspin, nspin, norb, norb))
    hloc[0, 0, 0, 1] = hloc[0, 0, 1, 0] = 0.1
    hloc[0, 0, 0, 0], hloc[0, 0, 1, 1] = -0.1, 0.1
    uloc, ust, jh = [2.5, 2.0], 1.5, 0.25
    p = dmed.default_params(beta=100.0, **kw)
    nb = 5          # normal bath: e(a,k), v(a,k) per orbital
    bath = np.concatenate([rng.uniform(-1.5, 1.5, size=(norb, nb)).ravel(),
                           rng.uniform(0.2, 0.7, size=(norb, nb)).ravel()])
    ns = dmed.set_model(norb, nb, nspin, bath, hloc, uloc, ust, jh)
    timed("ns12_norb2_normal_bath_nbath5", norb, nspin, ns, p)
    nb = 10         # hybrid bath: e(k) shared, v(a,k)
    bath = np.concatenate([rng.uniform(-1.5, 1.5, size=nb),
                           rng.uniform(0.2, 0.7, size=(norb, nb)).ravel()])
    ns = dmed.set_model(norb, nb, nspin, bath, hloc, uloc, ust, jh, bath_type="hybrid")
    timed("ns12_norb2_hybrid_bath_nbath10", norb, nspin, ns, p)
    return out


def sites_solve(workload=1, nsites=16, lmats=1024):
    """Inequivalent-site solve (dmed_solve_sites): the sites one after another
    against all sites at once on the GPU.  Site i takes the workload's bath
    with its levels scaled by 1 + 0.02 i and U + 0.1 i."""
    import numpy as np
    import dmed
    from dmed import configs
    c = configs.config(workload)
    sites = []
    for i in range(nsites):
        b = np.array(c["bath"], dtype=float)
        b[:c["nbath"]] *= 1.0 + 0.02 * i
        sites.append(dict(norb=c["norb"], nbath=c["nbath"], nspin=c["nspin"], bath=b,
                          hloc_f=c["hloc"], uloc=[c["uloc"][0] + 0.1 * i], xmu=c["xmu"]))
    kw = dict(beta=c["beta"], lmats=lmats, lreal=100, lanc_ngfiter=200)
    dmed.solve_sites(sites[:2], dmed.default_params(batched=1, **kw))      # warm-up
    t = time.perf_counter()
    seq = dmed.solve_sites(sites, dmed.default_params(batched=0, **kw))
    t_seq = time.perf_counter() - t
    t = time.perf_counter()
    conc = dmed.solve_sites(sites, dmed.default_params(batched=1, **kw))
    t_conc = time.perf_counter() - t
    same = all(np.array_equal(x["smats"], y["smats"]) for x, y in zip(seq, conc))
    log(f"workload {workload} ({c['name']}): {nsites} sites  one after another {t_seq:.2f} s  "
        f"concurrent {t_conc:.2f} s  ratio {t_seq / t_conc:.2f}  identical={same}")
    return {"workload": c["name"], "sites": nsites, "sequential_s": t_seq,
            "concurrent_s": t_conc, "ratio": t_seq / t_conc, "sigma_bit_identical": same}


def cpu_model():
    """CPU model and logical CPU count of this host (lscpu)."""
    import subprocess
    model, n = None, os.cpu_count()
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in txt.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() == "Model name":
                model = v.strip()
            elif k.strip() == "CPU(s)":
                n = int(v.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return model, n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prec", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="rows of the oracle H*v sample (default: the whole sector)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the all-sector Ns=12 finite-T solve (plain FP64 vs batched "
                         "mixed precision)")
    ap.add_argument("--no-solve5", action="store_true",
                    help="skip the Ns=16 sector pipeline (ground state + 500-step chains, plain "
                         "FP64 vs mixed precision)")
    ap.add_argument("--reps", type=int, default=3,
                    help="repetitions of each solve form (the speed-ups use the medians)")
    ap.add_argument("--configs", action="store_true",
                    help="also time workloads 1-3 beside the oracle (H*v; workload-1 solve)")
    ap.add_argument("--chi", action="store_true",
                    help="also time solves with the susceptibilities and with a hybrid bath")
    ap.add_argument("--sites", default="",
                    help="W:S -- also time S inequivalent sites of workload W, one after "
                         "another and concurrently")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver runs the hot path on a single MI355X "
              "(no multi-GPU sharding); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import dmed
    from dmed import configs

    cfg = configs.config(5)
    dmed.init(0)
    dmed.set_model(cfg["norb"], cfg["nbath"], cfg["nspin"], cfg["bath"], cfg["hloc"], cfg["uloc"],
                   cfg["ust"], cfg["jh"], 0.0, 0.0, cfg["xmu"], True, False)
    nu, nd = cfg["sector"]
    sec = dmed.Sector(nu, nd)
    D = sec.dim
    prec = dmed.FP64 if a.prec == "fp64" else dmed.FP32
    esz = 8 if prec == dmed.FP64 else 4
    log(f"workload {cfg['name']}: dim {D} ({sec.dim_up} x {sec.dim_dw}), {a.prec}")
    x = dmed.Vec(sec, prec)
    y = dmed.Vec(sec, prec)
    x.upload(np.random.default_rng(0).standard_normal(D))
    L = dmed.lib()
    for _ in range(a.warmup):
        dmed._check(L.dmed_hxv_async(sec.h, x.ptr(), y.ptr(), prec))
    dmed._check(L.dmed_sync())
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dmed._check(L.dmed_hxv_async(sec.h, x.ptr(), y.ptr(), prec))
    dmed._check(L.dmed_sync())
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / a.steps
    import ctypes as C
    kms = C.c_double()
    dmed._check(L.dmed_hxv_timed(sec.h, x.ptr(), y.ptr(), prec, a.steps, C.byref(kms)))
    kernel_ms = kms.value
    alg_bytes = 2 * esz * D
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    value = a.steps / wall
    shape = sec.hxv_shape(prec)
    log(f"H*v launch shape: {shape}")
    log(f"H*v: {ms_step:.4f} ms/step wall, kernel {kernel_ms:.4f} ms, "
        f"{achieved:.1f} GB/s algorithmic ({100 * achieved / HBM_PEAK_GBS:.1f}% of HBM peak)")

    # a plain streaming copy of one vector: the practical HBM ceiling on this
    # box, and the calibration kernel for the HBM traffic counters
    extra = {}
    xc = dmed.Vec(sec, dmed.FP64)
    for _ in range(2):
        xc.copy_from(x)
    t0 = time.perf_counter()
    ncopy = 5
    for _ in range(ncopy):
        xc.copy_from(x)
    tc = (time.perf_counter() - t0) / ncopy
    extra["copy_fp64"] = {"ms_wall": 1e3 * tc, "GBs": 16 * D / tc / 1e9}
    log(f"copy of one FP64 vector: {1e3 * tc:.4f} ms wall, {16 * D / tc / 1e9:.1f} GB/s")
    xc.free()

    # the FP32 form of the same product (mixed-precision chains use it)
    if prec == dmed.FP64:
        x32 = dmed.Vec(sec, dmed.FP32)
        y32 = dmed.Vec(sec, dmed.FP32)
        dmed._check(L.dmed_hxv_timed(sec.h, x.ptr(), y.ptr(), prec, 1, C.byref(kms)))
        x32.upload(np.random.default_rng(0).standard_normal(D))
        k32 = C.c_double()
        dmed._check(L.dmed_hxv_timed(sec.h, x32.ptr(), y32.ptr(), dmed.FP32, 3, C.byref(k32)))
        dmed._check(L.dmed_hxv_timed(sec.h, x32.ptr(), y32.ptr(), dmed.FP32, a.steps, C.byref(k32)))
        ach32 = 8 * D / (k32.value * 1e-3) / 1e9
        extra["fp32_hxv"] = {"kernel_ms": k32.value, "hv_per_s": 1e3 / k32.value,
                             "achieved_GBs": ach32, "frac": ach32 / HBM_PEAK_GBS}
        log(f"FP32 H*v: kernel {k32.value:.4f} ms, {ach32:.1f} GB/s algorithmic")
        x32.free()
        y32.free()

    for p_, nm in ((dmed.FP64, "fp64"), (dmed.FP32, "fp32")):
        extra[f"lanczos_step_{nm}"] = st = lanczos_step_rate(sec, D, p_)
        log(f"Lanczos step {nm}: {st['ms_per_step']:.3f} ms, {st['algorithmic_GBs']:.1f} GB/s "
            f"algorithmic (48 B/element FP64, 24 FP32)")

    if a.configs:
        extra["workloads_1_3_vs_oracle"] = configs_table(a.cpu_threads or (os.cpu_count() or 1))
    if a.chi:
        extra["chi_and_hybrid_solves"] = chi_solves()
    if a.sites:
        w, n = (int(v) for v in a.sites.split(":"))
        extra["inequivalent_sites"] = sites_solve(w, n)
    dim_up, dim_dw = sec.dim_up, sec.dim_dw
    sec.free()
    x.free()
    y.free()
    solve4 = None if a.no_solve else solve_compare(a.reps)
    solve5 = None if a.no_solve5 else solve5_compare(a.reps)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hxv_ns16_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        key = f"{a.prec}_bytes_per_launch"
        traffic = t.get(key)

    cpu = None
    if not a.no_cpu:
        model, ncpu = cpu_model()
        threads = a.cpu_threads or ncpu or 1
        rows = min(a.cpu_rows, D) if a.cpu_rows > 0 else D
        # then one thread on a bounded row sample (about 10 s of work): the
        # reference itself is serial per MPI rank
        res = cpu_baseline(cfg, [(rows, threads),
                                 (lambda r: min(D, max(1 << 20, int(10.0 * r[0][2] / threads))), 1)])
        (_, _, rate, dt), (rows1, _, rate1, dt1) = res
        what = ("one full H*v of the dim-%d sector" % D if rows == D else
                f"rows [0,{rows}) of the dim-{D} sector H*v, extrapolated to the full sector")
        cpu = {"value": rate / D, "unit": "Hv/s", "cores": threads, "kind": "port",
               "cpu_model": model, "logical_cpus": ncpu,
               "sample": f"{what} ({dt:.1f}s) on {threads} OpenMP threads over row blocks: "
                         f"the oracle's complex(8) restatement of directMatVec_cc (state map "
                         f"+ binary search per hop)",
               "one_thread": {"value": rate1 / D, "unit": "Hv/s", "cores": 1,
                              "sample": f"rows [0,{rows1}) ({dt1:.1f}s), extrapolated to the "
                                        f"full sector"}}
        log(f"cpu baseline ({model}, {ncpu} logical CPUs): {rate / D:.4f} Hv/s on {threads} "
            f"threads, {rate1 / D:.5f} Hv/s on 1 thread")

    out = {
        "metric": "H*v/s at dim 1.66e8 (Ns=16 half-filled sector, FP64); batched mixed vs "
                  "FP64 solve time (solve_speedup)",
        "value": value,
        "unit": "Hv/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == dmed.FP64 else "f32",
        "data": "synthetic (seeded SplitMix64 bath, random FP64 vector)",
        "config": {"workload": cfg["name"], "dim": D, "dim_up": dim_up,
                   "dim_dw": dim_dw, "norb": 1, "nbath": 15, "U": 2.5,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "solve_speedup": {
            "workload4_all_sectors_mixed_batched_vs_fp64_plain":
                solve4["speedup"] if solve4 else None,
            "workload4_errors_vs_fp64": solve4["errors_vs_fp64"] if solve4 else None,
            "workload5_sector_pipeline_mixed_vs_fp64_plain":
                solve5["speedup"] if solve5 else None,
            "workload5_errors_vs_fp64": solve5["errors_vs_fp64"] if solve5 else None,
            "repetitions": a.reps, "statistic": "ratio of median wall times"},
        "solve_ns12_all_sectors": solve4,
        "solve_ns16_sector_pipeline": solve5,
        "kernel_ms": kernel_ms,
        "hxv_launch_shape": shape,
        "extra": extra,
        "algorithmic_bytes_per_launch": alg_bytes,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
