#!/usr/bin/env python3
"""Retarded BEM wavelength-sweep benchmark on one MI355X.

Workload: gold-Drude capped rod (10 nm diameter x 40 nm, curved triangles, 2,496 faces) in
eps = 1.77, plane wave polarised along the rod axis, 201 wavelengths over 500-1000 nm. One
step = the full sweep (assembly, factorisation, solve, far field, extinction for every
wavelength), `--batch` wavelengths per factorisation (default all 201 at once: about 150 GB
of HBM for either precision), inputs resident on the GPU.

value = wavelength solves/s of the mixed-precision path (complex64 LU chain whose GEMMs run
on the FP16 matrix cores with operands as scaled fp16 pairs, measured as accurate as the FP32
matrix cores, + FP64 iterative refinement on the exact operator); the complex128 LU path is
timed in the same run, at the same batch, as the reference point, and the GEMM forms with 3 and
2 bfloat16 terms are reported beside it. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s @2.5k faces mixed vs FP64 LU; 25k-face GMRES time-to-1e-6"
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP64_TFLOPS = 78.6
DRUDE = (10.0, 9.065, 0.0708)
ROD_N = [25, 11, 34]


def _gemm_kernel(tuning, split):
    """Name of the complex-float GEMM kernel that the mixed chain's wide launches run."""
    split = int(tuning.get("gemm_split", split))
    if split == 0:
        return "k_gemm_sub_c64v (complex-float GEMM, FP32 matrix cores)"
    tile = int(tuning.get("plane_tile", 11))
    mt, nt = {11: (1, 1), 111: (1, 1), 21: (2, 1), 12: (1, 2), 22: (2, 2)}[tile]
    if split == 16:
        if tile == 11:
            return "k_gemm_sub_c64p<2, true> (complex GEMM on the FP16 matrix cores, scaled fp16 pairs)"
        return "k_gemm_sub_c64q<2, {}, {}, true> (complex GEMM on the FP16 matrix cores, scaled fp16 pairs)".format(
            mt, nt)
    if tile == 11:
        return "k_gemm_sub_c64p<{}> (complex GEMM on the BF16 matrix cores, {} operand terms)".format(split, split)
    return "k_gemm_sub_c64q<{}, {}, {}> (complex GEMM on the BF16 matrix cores, {} operand terms)".format(
        split, mt, nt, split)


def _progress(msg):
    """One line on stderr per finished stage (the JSON result is the only stdout line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


# HBM bytes per k_gemm_sub_c64q<2, 1, 1, true> launch of this sweep (batch 201, 406 launches,
# scaled fp16 pairs, outer blocks of 256 by recursive halving) from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, each scaled by the counter's reading for a 256 MiB copy in the same run (FETCH x2.0,
# WRITE x1.0); the commands and raw averages are in the file named.
GEMM_TRAFFIC_BYTES = int((907156.5 * 2.0 + 1078164.0 * 1.0) * 1024)
GEMM_TRAFFIC_SOURCE = "profiles/gemm_pmc_rod2496.md"


def build_workload():
    import mnpbem_amd as M
    rod = M.trirod(10.0, 40.0, ROD_N, triangles=True)
    cp = M.ComParticle([M.EpsConst(1.77), M.EpsDrude(*DRUDE)], [rod], [[2, 1]], 1, interp="curv")
    exc = M.PlaneWaveRet([0.0, 0.0, 1.0], [1.0, 0.0, 0.0])
    return M, cp, exc


def _oracle_sweep(cp, lams, pol, direc, eps_out):
    """The oracle's per-wavelength loop (assembly + LU chain + solve + extinction) after its
    wavelength-independent set-up; returns seconds."""
    from oracle import bem_oracle as O
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    fp = np.asarray(cp.face_particle)
    closed = [(np.nonzero(fp == a)[0], np.nonzero(fp == a)[0], 1.0) for a in range(len(cp.p))]
    green = O.Green(s, fp, np.asarray(cp.inout), closed)
    pol, direc = np.asarray(pol, float), np.asarray(direc, float)
    t0 = time.perf_counter()
    for lam in lams:
        eps_k = [O.eps_const(lam, eps_out), O.eps_drude(lam, *DRUDE)]
        ks = [e[1] for e in eps_k]
        G1, G2, H1, H2 = O.matrices(green, ks)
        e1 = np.full(s.n, eps_k[1][0])
        e2 = np.full(s.n, eps_k[0][0])
        k0 = 2 * np.pi / lam
        bem = O.BEMRetOracle(G1, G2, H1, H2, s.nvec, e1, e2, k0)
        a2, a2p = O.planewave(s.pos, s.nvec, pol, direc, lam, eps_out, np.arange(s.n))
        z = np.zeros_like(a2)
        sol = bem.solve(*O.excitation(s.nvec, k0, e1, e2, z, z, a2, a2p))
        O.extinction(pol, direc, lam, ks[0], eps_out, s.pos, s.area, sol, np.zeros(0, int), np.arange(s.n))
    return time.perf_counter() - t0


def cpu_baseline(cp, nsample, nsample1=1):
    """The oracle (FP64 numpy/LAPACK restatement of the reference path) on the host cores:
    per-wavelength assembly + LU chain + solve + extinction, timed over `nsample` wavelengths
    with every BLAS thread and over `nsample1` wavelengths on one core, after the
    wavelength-independent set-up."""
    from threadpoolctl import threadpool_info, threadpool_limits
    lams = np.linspace(500, 1000, 201)[:: max(1, 201 // max(nsample, 1))][:nsample]
    affinity = len(os.sched_getaffinity(0))
    blas = max([int(i.get("num_threads", 1)) for i in threadpool_info()] or [1])
    dt = _oracle_sweep(cp, lams, [0, 0, 1.0], [1.0, 0, 0], 1.77)
    with threadpool_limits(limits=1):
        dt1 = _oracle_sweep(cp, lams[:nsample1], [0, 0, 1.0], [1.0, 0, 0], 1.77)
    return dict(value=round(len(lams) / dt, 4), unit="solves/s", cores=blas, kind="port",
                host_cpus=affinity, blas_threads=blas,
                single_core_value=round(nsample1 / dt1, 4),
                sample="{} of the 201 wavelengths at n={} with {} BLAS threads on {} host CPUs ({:.1f} s), and {} "
                       "on one core ({:.1f} s): assembly + LU chain + solve + extinction per wavelength, geometry "
                       "set-up excluded".format(len(lams), cp.nfaces, blas, affinity, dt, nsample1, dt1))


def cpu_other_configs():
    """CPU references for the 64-particle set and the 24.5k-face rod dimer, whose full runs take
    hours on the host, as a measured sample and a bound rather than an extrapolation:
      batched64: the oracle (all host threads) on the particles at the 0th, 20th and 40th
        percentile of the face counts, one wavelength (600 nm) each, per-wavelength loop after
        the wavelength-independent set-up; solves/s over the sample is an upper bound for the
        whole set (its other particles are larger).
      gmres25k: a lower bound for a dense direct solve at 24,544 faces on the host: the LU
        chain's 34.67 n^3 flops at the host's measured complex128 GEMM rate (numpy matmul of
        2048 x 2048 matrices); the reference's H-matrix GMRES at 25k faces has no published time
        (it timed out after > 5 min)."""
    parts = particles64()
    order = np.argsort([cp.nfaces for _, cp in parts])
    sample = [parts[i][1] for i in (order[0], order[len(order) // 5], order[2 * len(order) // 5])]
    ts = [_oracle_sweep(cp, [600.0], [1.0, 0, 0], [0, 0, 1.0], 1.33) for cp in sample]
    rng = np.random.default_rng(0)
    A = rng.standard_normal((2048, 2048)) + 1j * rng.standard_normal((2048, 2048))
    A @ A
    t0 = time.perf_counter()
    A @ A
    rate = 8.0 * 2048 ** 3 / (time.perf_counter() - t0)
    out = dict(batched64=dict(value=round(len(ts) / sum(ts), 4), unit="solves/s", kind="port", bound="upper",
                              sample_faces=[cp.nfaces for cp in sample], sample_s=[round(t, 3) for t in ts]),
               gmres25k=dict(value=round(34.67 * 24544.0 ** 3 / rate, 1), unit="s (dense direct solve)", kind="port",
                             bound="lower", host_zgemm_gflops=round(rate / 1e9, 1)))
    _progress("cpu other configs: {}".format(out))
    return out


def config1(skip_cpu=False):
    """The reference's own CPU-runnable case: trisphere(144, 20) gold-Drude sphere in vacuum,
    x-polarised plane wave along z, 41 wavelengths 400-800 nm. GPU: the complex128 sweep (one
    batch of 41) timed after a warm-up; CPU: the oracle's loop over all 41 wavelengths with every
    BLAS thread (the scope of validation/03_bemret/sphere/run_python.py:37-45)."""
    import torch
    import mnpbem_amd as M
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsDrude(*DRUDE)], [M.trisphere(144, 20)], [[2, 1]], 1, interp="curv")
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    lams = np.linspace(400, 800, 41)
    out = {"workload": "BEMRet gold-Drude sphere trisphere(144, 20), n={}, vacuum, 41 wavelengths 400-800 nm"
                       .format(cp.nfaces)}
    for prec in ("double", "mixed"):
        bem = M.BEMRet(cp, batch=41, precision=prec)
        bem.solve_spectrum(exc, lams)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = bem.solve_spectrum(exc, lams)
        torch.cuda.synchronize()
        out["gpu_{}_s".format(prec)] = round(time.perf_counter() - t0, 4)
        out["gpu_{}_solves_per_s".format(prec)] = round(41 / (time.perf_counter() - t0), 2)
        if prec == "double":
            ext_d = r["ext"][:, 0]
    if not skip_cpu:
        dt = _oracle_sweep(cp, lams, [1.0, 0, 0], [0, 0, 1.0], 1.0)
        out["cpu_s"] = round(dt, 3)
        out["cpu_solves_per_s"] = round(41 / dt, 3)
        out["cpu_kind"] = "port"
    out["ext_max"] = float(np.max(ext_d))
    return out


def gmres25k(preconditioner="particle"):
    """Rod dimer (two 20x80 nm capped rods end to end, 2 nm gap, ~24.5k faces), gold Drude in
    vacuum, 750 nm, axial polarisation: BEMRetIter with H-matrices (eta 2.5, leaf 64, ACA tol
    1e-6, kmax 100, low-rank factors complex64) and GMRES(50) to 1e-6. time_to_tol = H-matrix
    build (with the cluster tree) + preconditioner + GMRES; the wavelength-independent set-up
    (meshes, near-field pairs, closed-surface terms; for 'particle' also those of each rod alone)
    is reported separately. One untimed solve of a separate BEMRetIter first (the process's first
    use of the H-matrix, ACA and panel kernels and their workspaces, reported as
    time_to_tol_cold_s); the timed solve then builds its own H-matrix and preconditioner.
    preconditioner: 'auto' = the BEMRet chain of the whole dimer (the reference's choice),
    'particle' = the chains of the two rods alone as a block-diagonal preconditioner."""
    import torch
    import mnpbem_amd as M
    from mnpbem_amd.bem_iter import BEMRetIter
    t0 = time.perf_counter()
    cp = _dimer25k()
    g = M.CompGreenRet(cp, cp)
    g.geo
    exc = M.PlaneWaveRet([0, 0, 1.0], [1.0, 0, 0])
    e = exc(cp, 750.0)
    mk = lambda: BEMRetIter(cp, tol=1e-6, maxit=500, restart=50, hmatrix=True, htol=1e-6, kmax=100, cleaf=64,
                            eta=2.5, green=g, preconditioner=preconditioner)
    it = mk()
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    it.solve(e)
    torch.cuda.synchronize()
    cold = time.perf_counter() - t0
    del it
    it = mk()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sig, _ = it.solve(e)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    flag, relres, niter = it.info()
    st = it._hm.stats
    out = dict(workload="BEMRetIter gold-Drude rod dimer 2x(20x80 nm), 2 nm gap, n={}, 750 nm, axial; H-matrix "
                        "eta 2.5 leaf 64 htol 1e-6; GMRES(50) to 1e-6, {} (complex64)".format(
                            cp.nfaces, "BEMRet chain of each rod alone as block-diagonal preconditioner"
                            if preconditioner == "particle" else "BEMRet-chain preconditioner"),
               faces=cp.nfaces, time_to_tol_s=round(total, 3), time_to_tol_cold_s=round(cold, 3), hmatrix_build_s=round(it.timings.get("hmatrix", 0), 3),
               precond_s=round(it.timings.get("precond", 0), 3), gmres_s=round(it.timings.get("gmres", 0), 3),
               iterations=int(niter), relres=float(relres), converged=flag == 0, ext=float(exc.extinction(sig)),
               compression=[round(c, 4) for c in st["compression"]], rank_max=st.get("rank_max"),
               bytes_per_operator_apply=st["bytes_per_apply"], host_setup_s=round(setup, 3))
    del it, g
    torch.cuda.empty_cache()
    return out


def _dimer25k():
    import mnpbem_amd as M
    p1 = M.trirod(20, 80, [60, 14, 80], triangles=True).shift([0, 0, -41.0])
    p2 = M.trirod(20, 80, [60, 14, 80], triangles=True).shift([0, 0, 41.0])
    return M.ComParticle([M.EpsConst(1.0), M.EpsDrude(10, 9.065, 0.0708)], [p1, p2], [[2, 1], [2, 1]], 1, 2,
                         interp="curv")


def _training_particles(rng, count):
    """Seeded capped rods (diameter 8-20 nm, aspect 2.5-6) and prolate ellipsoids (semi-axes
    5-10 nm, aspect 1-4) alternating, 2.0k-2.9k curved triangles each, gold Drude in vacuum."""
    import mnpbem_amd as M
    drude = M.EpsDrude(10, 9.065, 0.0708)
    cps = []
    for i in range(count):
        if i % 2 == 0:
            d = float(rng.uniform(8, 20))
            L = float(rng.uniform(2.5, 6.0)) * d
            n = [int(rng.integers(24, 31)), 9, int(rng.integers(30, 35))]
            part = M.trirod(d, L, n, triangles=True)
        else:
            a = float(rng.uniform(5, 10))
            part = M.triellipsoid(int(rng.choice([1024, 1225, 1444])), [a, a, a * float(rng.uniform(1, 4))])
        cps.append(M.ComParticle([M.EpsConst(1.0), drude], [part], [[2, 1]], 1, interp="curv"))
    return cps


def warmstart25k(nlam=31, nparticles=40, lams_per_particle=100, seed=7):
    """Learned warm start on the rod dimer: the MLP is trained (forward, backward and Adam on the
    device) on nparticles x lams_per_particle direct solves (mixed precision device sweeps) of
    seeded rods and ellipsoids with 2.0k-2.9k faces at wavelengths drawn uniformly over
    450-950 nm; then GMRES(50) to 1e-6 with H-matrices runs at nlam wavelengths over 600-900 nm
    from the zero start, the previous wavelength's solution, the residual-minimising combination
    of the last three solutions ('extrapolate'), the learned guess, and the learned guess as a
    fourth candidate of that combination. The preconditioner (complex64 BEMRet chain) is factored
    once at 750 nm and applied through its explicit inverses. Times: GMRES including each start's
    operator applications, plus feature extraction and inference for the learned starts."""
    import torch
    import mnpbem_amd as M
    from mnpbem_amd import warmstart as W
    from mnpbem_amd.bem_iter import BEMRetIter
    pol, direc = [0, 0, 1.0], [1.0, 0, 0]
    rng = np.random.default_rng(seed)
    cps = _training_particles(rng, nparticles)
    lams_train = [np.sort(rng.uniform(450, 950, lams_per_particle)) for _ in cps]
    t0 = time.perf_counter()
    X, T = W.training_set(cps, lams_train, pol, direc)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    _progress("warm start: {} training solves in {:.1f} s".format(nparticles * lams_per_particle, t_data))
    net = W.WarmStartMLP(seed=0)
    t0 = time.perf_counter()
    hist = net.fit(X, T, steps=3000, batch=16384, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    _progress("warm start: trained in {:.1f} s, loss {:.4f} -> {:.4f}".format(t_train, hist[0], hist[-1]))
    cp = _dimer25k()
    g = M.CompGreenRet(cp, cp)
    exc = M.PlaneWaveRet(pol, direc)
    H = W.mean_curvature(cp)
    it = BEMRetIter(cp, tol=1e-6, maxit=400, restart=50, hmatrix=True, htol=1e-6, kmax=100, cleaf=64, eta=2.5,
                    green=g, precond_wavelength=750.0)
    learned = lambda lam: W.guess(cp, net.predict(W.features(cp, lam, pol, direc, curvature=H)), pol)
    modes = ("zero", "previous", "extrapolate", "learned", "extrapolate+learned")
    lams = np.linspace(600, 900, nlam)
    res = it.solve_spectrum(exc, lams, start=modes, history=3, warm_start=learned)
    precond_s = round(float(it.timings.get("precond_build", 0.0)), 3)
    out = {}
    for m in modes:
        r = res[m]
        out[m] = dict(iterations=r["iterations"], total_iterations=int(sum(r["iterations"])),
                      time_s=round(float(sum(r["time_s"])), 3),
                      initial_residual=[round(v, 4) for v in r["initial_residual"]],
                      converged=all(f == 0 for f in r["flag"]))
    ext_spread = float(max(np.max(np.abs(res[m]["ext"] - res["zero"]["ext"]) / np.abs(res["zero"]["ext"]))
                           for m in modes))
    del it, g
    torch.cuda.empty_cache()
    t = {m: out[m]["time_s"] for m in modes}
    return dict(workload="rod dimer n={}, {} wavelengths 600-900 nm, GMRES(50) to 1e-6, H-matrices, complex64 "
                         "BEMRet-chain preconditioner factored once at 750 nm, applied through explicit "
                         "inverses".format(cp.nfaces, nlam),
                training=dict(solves=nparticles * lams_per_particle, particles=nparticles,
                              faces=[int(min(c.nfaces for c in cps)), int(max(c.nfaces for c in cps))],
                              samples=int(X.shape[0]), data_s=round(t_data, 2), train_s=round(t_train, 2),
                              steps=3000, loss_first=round(hist[0], 4), loss_last=round(hist[-1], 4)),
                precond_once_s=precond_s, starts=out, ext_max_rel_spread_between_starts=ext_spread,
                speedup_learned_vs_zero=round(t["zero"] / max(t["learned"], 1e-9), 3),
                speedup_learned_vs_previous=round(t["previous"] / max(t["learned"], 1e-9), 3),
                speedup_extrapolate_vs_zero=round(t["zero"] / max(t["extrapolate"], 1e-9), 3),
                speedup_extrapolate_learned_vs_extrapolate=round(t["extrapolate"] /
                                                                 max(t["extrapolate+learned"], 1e-9), 3))


def sphere_mesh(nfaces, diameter):
    """Curved-triangle sphere whose face count is the even number closest to ``nfaces``: a
    latitude-longitude grid of m longitudes and t latitude bands (trispheresegment over the full
    sphere, poles merged) has exactly 2 m t faces; m / t is kept between 1.5 and 2.5."""
    import mnpbem_amd as M
    best = None
    for t in range(4, 80):
        for m in (int(np.floor(nfaces / (2 * t))), int(np.ceil(nfaces / (2 * t)))):
            if m < 6 or not 1.5 <= m / t <= 2.5:
                continue
            key = (abs(2 * m * t - nfaces), abs(m / t - 2.0))
            if best is None or key < best[0]:
                best = (key, m, t)
    _, m, t = best
    return M.trispheresegment(np.linspace(0, 2 * np.pi, m + 1), np.linspace(0, np.pi, t + 2), diameter,
                              triangles=True)


def particles64(seed=20261203):
    """64 seeded particles of uneven size, as (kind, ComParticle): spheres (20 nm), prolate
    ellipsoids (semi-axes 8, 8, 8 x aspect U[1, 4]) and gap dimers of two 20 nm spheres (gap
    U[1, 5] nm), face counts drawn uniform in 300-3000 and realised by the mesh to within a few
    faces (a dimer splits its count over its two spheres), gold Drude in eps 1.33."""
    import mnpbem_amd as M
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(64):
        kind = ("sphere", "ellipsoid", "dimer")[int(rng.integers(3))]
        nf = int(rng.integers(300, 3001))
        drude = M.EpsDrude(10, 9.065, 0.0708)
        if kind == "sphere":
            cp = M.ComParticle([M.EpsConst(1.33), drude], [sphere_mesh(nf, 20.0)], [[2, 1]], 1,
                               interp="curv")
        elif kind == "ellipsoid":
            asp = float(rng.uniform(1, 4))
            cp = M.ComParticle([M.EpsConst(1.33), drude], [sphere_mesh(nf, 2.0).scale([8, 8, 8 * asp])],
                               [[2, 1]], 1, interp="curv")
        else:
            gap = float(rng.uniform(1, 5))
            s1 = sphere_mesh(nf // 2, 20.0).shift([0, 0, -(10 + gap / 2)])
            s2 = sphere_mesh(nf // 2, 20.0).shift([0, 0, 10 + gap / 2])
            cp = M.ComParticle([M.EpsConst(1.33), drude], [s1, s2], [[2, 1], [2, 1]], 1, 2,
                               interp="curv")
        out.append((kind, cp))
    return out


def batched64(compare_unbatched=True):
    """64 particles of uneven size (particles64), 32 wavelengths each (450-900 nm, x-polarised
    plane wave along z), mixed precision: the particles are sorted by face count and cut into
    groups (padding at most 25 % of the LU work, matrices within the memory budget); each group
    is one batch of (particle, wavelength) matrices through assembly, LU chain, refinement and far
    field (launched eagerly; the same groups captured as HIP graphs, mnpbem_amd.batch.GraphedGroup,
    are timed beside it). solves/s over all 2048 solves; per-solve latency = the wall time of the group that
    holds the solve (all its solves complete together), p50/p99 over the 2048 solves; completion
    = time from the start of the set to the group's results. Host geometry set-up and a first
    (untimed) pass that runs the LU autotuner once per group shape are excluded. With
    compare_unbatched the previous scheme (each particle's 32 wavelengths as its own batch) is
    timed too."""
    import torch
    import mnpbem_amd as M
    from mnpbem_amd import batch as MB
    t0 = time.perf_counter()
    tagged = particles64()
    parts = [cp for _, cp in tagged]
    greens = [M.CompGreenRet(cp, cp) for cp in parts]
    for g in greens:
        g.geo
    setup = time.perf_counter() - t0
    lams = np.linspace(450, 900, 32)
    pol, direc = [1.0, 0, 0], [0, 0, 1.0]
    faces = [cp.nfaces for cp in parts]
    t_tune = time.perf_counter()
    MB.solve_particles(parts, lams, pol, direc, precision="mixed", greens=greens)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t_tune
    # the groups launched eagerly (the timed result)
    timings = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    out = MB.solve_particles(parts, lams, pol, direc, precision="mixed", greens=greens, timings=timings)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_all
    # the same groups captured as HIP graphs: each group's graphs are captured, instantiated and
    # replayed once, so this times capture + instantiation against the launches it saves
    torch.cuda.synchronize()
    t_g = time.perf_counter()
    out_g = MB.solve_particles(parts, lams, pol, direc, precision="mixed", greens=greens, graphs=True)
    torch.cuda.synchronize()
    t_g = time.perf_counter() - t_g
    lat, comp, elapsed = [], [], 0.0
    for grp, dt in timings:
        elapsed += dt
        lat += [dt] * (len(grp) * len(lams))
        comp += [elapsed] * (len(grp) * len(lams))
    res = dict(workload="64 seeded particles (spheres, prolate ellipsoids, gap dimers), faces {}-{} ({} distinct), "
                        "gold Drude in eps 1.33, 32 wavelengths 450-900 nm each, mixed precision, particles grouped "
                        "into padded batches".format(min(faces), max(faces), len(set(faces))),
               solves=len(parts) * len(lams), solves_per_s=round(len(parts) * len(lams) / total, 3),
               groups=len(timings), group_sizes=[len(g) for g, _ in timings],
               per_solve_latency_ms_p50=round(1e3 * float(np.percentile(lat, 50)), 3),
               per_solve_latency_ms_p99=round(1e3 * float(np.percentile(lat, 99)), 3),
               completion_ms_p50=round(1e3 * float(np.percentile(comp, 50)), 3),
               completion_ms_p99=round(1e3 * float(np.percentile(comp, 99)), 3),
               path_tflops=round(34.67 * len(lams) * sum(float(f) ** 3 for f in faces) / total / 1e12, 3),
               fp64_fallbacks=len(out["fallback"]), host_setup_s=round(setup, 3),
               first_pass_with_autotune_s=round(t_tune, 3),
               graphs_solves_per_s=round(len(parts) * len(lams) / t_g, 3),
               ext_rel_diff_graphs_vs_eager=float(np.max(np.abs(out_g["ext"] - out["ext"]) / np.abs(out["ext"]))))
    if compare_unbatched:
        exc = M.PlaneWaveRet(pol, direc)
        for cp, g in zip(parts, greens):
            M.BEMRet(cp, batch=32, precision="mixed", green=g, autotune=False)._get_solver(32, "mixed")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ref = np.zeros_like(out["ext"])
        for i, (cp, g) in enumerate(zip(parts, greens)):
            ref[i] = M.BEMRet(cp, batch=32, precision="mixed", green=g, autotune=False).solve_spectrum(
                exc, lams)["ext"][:, 0]
        torch.cuda.synchronize()
        t_un = time.perf_counter() - t1
        res["unbatched_solves_per_s"] = round(len(parts) * len(lams) / t_un, 3)
        res["ext_rel_diff_vs_unbatched"] = float(np.max(np.abs(out["ext"] - ref) / np.abs(ref)))
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0, help="wavelengths per factorisation (0: all 201 at once)")
    ap.add_argument("--gemm-split", type=int, default=16, choices=(0, 2, 3, 16),
                    help="operand form of the complex-float GEMMs: 16 scaled fp16 pairs (FP16 matrix cores), "
                         "2 or 3 bfloat16 terms (BF16 matrix cores), 0 the FP32 matrix cores")
    ap.add_argument("--no-mixed2", action="store_true", help="skip the extra mixed sweeps with the other GEMM forms")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="copy one 256 MiB buffer first (a kernel of known traffic for counter scaling)")
    ap.add_argument("--refine-tol", type=float, default=1e-12,
                    help="relative residual at which the FP64 refinement of the mixed sweep stops")
    ap.add_argument("--lu-tuning", type=str, default=None,
                    help="JSON dict of LU knobs (nb, nbo, lookahead, lu_recursive, plane_tile) for the mixed "
                         "sweep instead of the autotuner's choice (profiler runs, whose counters distort timing)")
    ap.add_argument("--nb", type=int, default=32)
    ap.add_argument("--cpu-sample", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--no-gmres25k", action="store_true")
    ap.add_argument("--no-batched64", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-warmstart", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the wavelength sweep would shard by giving each GPU its own "
              "wavelength batches (one process per GPU), which is not built; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from mnpbem_amd import _lib
    if args.pmc_calibration:
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    M, cp, exc = build_workload()
    n = cp.nfaces
    lams = np.linspace(500, 1000, 201)
    if args.batch <= 0:
        args.batch = len(lams)
    torch.cuda.set_device(0)
    g = M.CompGreenRet(cp, cp)
    _progress("workload n={} ready".format(n))

    def sweep(prec):
        bem = solvers[prec]
        return bem.solve_spectrum(exc, lams, precision=prec)

    solvers = {"mixed": M.BEMRet(cp, batch=args.batch, nb=args.nb, precision="mixed", green=g,
                                 gemm_split=args.gemm_split, refine_tol=args.refine_tol,
                                 lu_tuning=json.loads(args.lu_tuning) if args.lu_tuning else None)}
    if not args.skip_fp64:
        solvers["double"] = M.BEMRet(cp, batch=args.batch, nb=args.nb, precision="double", green=g)

    results = {}
    solvers_tuning = {}
    for prec in solvers:
        for _ in range(args.warmup):
            sweep(prec)
            _progress("{} warm-up sweep done".format(prec))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = sweep(prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        _progress("{} sweep: {:.1f} solves/s".format(prec, args.steps * len(lams) / dt))
        if prec == "double":
            # one more sweep, untimed, for the solutions themselves (copied to the host)
            out = solvers[prec].solve_spectrum(exc, lams, precision=prec, solutions=True)
        results[prec] = (dt, out)
        solvers_tuning[prec] = solvers[prec].tuning
        solvers[prec]._solver = None          # the next precision gets the memory (large batches)
        torch.cuda.empty_cache()

    # GEMM (dominant kernel) timing with HIP events over one extra mixed sweep
    # (its solutions are kept for the surface charge / current error against the FP64 sweep)
    _lib.call("mnp_prof_enable", 1)
    sol_mixed = solvers["mixed"].solve_spectrum(exc, lams, precision="mixed", solutions=True)["solutions"]
    torch.cuda.synchronize()
    prof = np.zeros(4)
    _lib.call("mnp_prof_collect", prof.ctypes.data)
    _lib.call("mnp_prof_enable", 0)
    gemm_ms, gemm_flops, gemm_launches, gemm_bytes = prof
    _progress("GEMM timing sweep: {:.1f} ms in {} launches".format(gemm_ms, int(gemm_launches)))
    # algorithmic work of the GEMM launches (8 m n k real flops per complex product, the
    # complex FP32-equivalent count) over their kernel time, against the peak of the matrix
    # cores the kernel runs on (BF16 for the split operands, FP32 otherwise)
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    # the split-BF16 kernel executes terms (terms + 1) / 2 BF16 products per real product
    products = {0: 1, 2: 3, 3: 6, 16: 3}[args.gemm_split]
    peak = PEAK_BF16_TFLOPS if args.gemm_split else PEAK_FP32_TFLOPS    # FP16 and BF16 dense peaks are equal

    dt_mixed, out_mixed = results["mixed"]
    value = args.steps * len(lams) / dt_mixed
    extra = {}
    if "double" in results:
        dt_d, out_d = results["double"]
        v_d = args.steps * len(lams) / dt_d
        rel = np.max(np.abs(out_mixed["ext"] - out_d["ext"]) / np.abs(out_d["ext"]))
        extra = dict(fp64_solves_per_s=round(v_d, 3), speedup_vs_fp64=round(value / v_d, 3),
                     ext_rel_err_vs_fp64=float(rel))
        # per wavelength |x_mixed - x_fp64| / |x_fp64| over each of sigma1, h1, sigma2, h2 (the
        # packed rows 0, 1-3, 4, 5-7), largest over the sweep
        xd = out_d["solutions"]
        errs = {}
        for name, rows in (("sig1", [0]), ("h1", [1, 2, 3]), ("sig2", [4]), ("h2", [5, 6, 7])):
            num = np.linalg.norm((sol_mixed[:, :, rows] - xd[:, :, rows]).reshape(len(lams), -1), axis=1)
            den = np.linalg.norm(xd[:, :, rows].reshape(len(lams), -1), axis=1)
            errs[name] = float(np.max(num / den))
        extra["surface_rel_err_vs_fp64"] = errs
    # the same sweep with the other GEMM operand forms: 2 bfloat16 terms (about 16 significant
    # bits) and scaled fp16 pairs (about 22 bits); the main solver's buffers go first (two would
    # not fit at batch 201)
    for other, tag in ((3, "mixed_3term"), (2, "mixed_2term"), (16, "mixed_f16pair")):
        if args.no_mixed2 or args.gemm_split == other or "double" not in results:
            continue
        solvers["mixed"]._solver = None
        torch.cuda.empty_cache()
        b2 = M.BEMRet(cp, batch=args.batch, nb=args.nb, precision="mixed", green=g, gemm_split=other,
                      refine_tol=args.refine_tol)
        b2.solve_spectrum(exc, lams, precision="mixed")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out2 = b2.solve_spectrum(exc, lams, precision="mixed")
        torch.cuda.synchronize()
        dt2 = time.perf_counter() - t0
        extra[tag + "_solves_per_s"] = round(len(lams) / dt2, 3)
        extra[tag + "_speedup_vs_fp64"] = round(len(lams) / dt2 / extra["fp64_solves_per_s"], 3)
        extra[tag + "_ext_rel_err_vs_fp64"] = float(np.max(np.abs(out2["ext"] - results["double"][1]["ext"]) /
                                                           np.abs(results["double"][1]["ext"])))
        extra[tag + "_refine_iterations_max"] = int(max(out2.get("refine_iterations", [0]) or [0]))
        extra[tag + "_fp64_fallbacks"] = len(out2.get("fallback", []))
        extra[tag + "_lu_tuning"] = b2.tuning
        b2._solver = None
        del b2
        torch.cuda.empty_cache()
        _progress("{} sweep: {:.1f} solves/s".format(tag, extra[tag + "_solves_per_s"]))
        torch.cuda.empty_cache()
    flops_per_solve = 34.67 * n ** 3
    extra["path_tflops_mixed"] = round(value * flops_per_solve / 1e12, 3)
    extra["refine_iterations_max"] = int(max(out_mixed.get("refine_iterations", [0]) or [0]))
    extra["refine_tol"] = args.refine_tol
    extra["refine_relres_history_max"] = [float("{:.3g}".format(v)) for v in out_mixed.get("refine_history", [])]
    extra["fp64_fallbacks"] = len(out_mixed.get("fallback", []))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cp, args.cpu_sample)
        _progress("cpu baseline: {}".format(cpu))
        if not (args.no_batched64 and args.no_gmres25k):
            cpu["other_configs"] = cpu_other_configs()

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt_mixed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 LU ({}) + c128 refinement".format(
            {0: "FP32 MFMA", 2: "2-term split-BF16 MFMA", 3: "3-term split-BF16 MFMA, FP32-equivalent",
             16: "scaled fp16-pair MFMA, FP32-equivalent"}[args.gemm_split]),
        "data": "synthetic",
        "config": {"workload": "BEMRet gold-Drude capped rod 10x40 nm, n={} curved triangles, eps_out=1.77, "
                               "axial plane wave, 201 wavelengths 500-1000 nm, batch {}".format(n, args.batch),
                   "faces": n, "wavelengths": len(lams), "batch": args.batch,
                   "lu_tuning_mixed": solvers_tuning.get("mixed"), "lu_tuning_fp64": solvers_tuning.get("double")},
        "roofline": {"bound": "mfma",
                     "kernel": _gemm_kernel(solvers_tuning.get("mixed") or {}, args.gemm_split),
                     "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4),
                     "work": "algorithmic 8 m n k real flops per launch (complex FP32-equivalent)",
                     "executed_16bit_tflops": round(achieved * products, 3) if args.gemm_split else None,
                     "executed_frac": round(achieved * products / peak, 4) if args.gemm_split else None,
                     "frac_of_fp32_peak": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": GEMM_TRAFFIC_BYTES, "traffic_source": GEMM_TRAFFIC_SOURCE,
                     "algorithmic_bytes_per_launch": round(gemm_bytes / max(gemm_launches, 1)),
                     "algorithmic_flops_per_launch": round(gemm_flops / max(gemm_launches, 1)),
                     "kernel_us_per_launch": round(1e3 * gemm_ms / max(gemm_launches, 1), 2),
                     "gemm_ms_per_sweep": round(gemm_ms, 3), "gemm_launches_per_sweep": int(gemm_launches)},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    solvers.clear()
    torch.cuda.empty_cache()
    if not args.no_config1:
        line["config1"] = config1(skip_cpu=args.no_cpu)
        _progress("config1 done")
    if not args.no_batched64:
        line["batched64"] = batched64()
        _progress("batched64 done")
    if not args.no_gmres25k:
        # time-to-1e-6 with the chain of each rod alone as preconditioner (the faster form), and
        # with the chain of the whole dimer (the reference's preconditioner for hmatrix=True)
        line["gmres25k"] = gmres25k("particle")
        _progress("gmres25k done")
        line["gmres25k_dimer_precond"] = gmres25k("auto")
        _progress("gmres25k with the whole-dimer preconditioner done")
    if not args.no_warmstart:
        line["warmstart25k"] = warmstart25k()
        _progress("warmstart25k done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
