#!/usr/bin/env python3
"""Pressure-correction solve benchmark on one MI355X.

Workload (one "step" = one pressure-correction solve of a time step):
standalone pCorr Poisson on a 1,000,000-cell unstructured triangular mesh of
[0,1]^2 (Delaunay of 512 points per edge + 498,977 uniform interior points,
seed 1), face coefficient rho*d lognormal sigma=1 (seed 2), right-hand side
i.i.d. N(0,1) projected to zero mean (seed 3), pCorr = 0 (FIXED) on the y-
edge, zero-gradient elsewhere.  Each step assembles fv::laplacian(gamma_f, p)
== b on the device from HBM-resident inputs, factors ILU(0) and solves to an
FP64 true relative residual of 1e-8 from a zero initial guess.

Prints ONE JSON line.  `value` is ms per step of the fastest method that
reaches the FP64 goal: plain FP64 BiCGStab+ILU(0) ("fp64"), FP32 BiCGStab+ILU(0)
in FP64 iterative refinement ("mixed"), either with the learned correction
("learned"), or the same with preconditioned CG in place of BiCGStab ("-cg":
Belos "CG", valid because the pressure matrix is symmetric).  Every method's
time is reported beside it, as are the SpMV bandwidths and the per-launch
kernel timings.
`roofline.traffic` is the HBM bytes per launch of the dominant kernel read from
the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, each scaled by
the same counter on a copy kernel of known size; `--copy-probe-mb` runs that
copy kernel for the counter pass).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "pCorr solve ms/step at 1e-8 rel. residual, learned+mixed vs plain; SpMV GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_sweeps_pcorr_poisson_1M.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (a trace name such as
    "phase::kLSolve<float, 3, 4>") in the committed counter summary's table:
    the "HBM bytes per launch" column of the row whose kernel or template
    column names it, or None."""
    want = "`%s`" % kernel.replace("phase::", "")
    try:
        col = None
        with open(TRAFFIC_FILE) as f:
            for line in f:
                cells = [c.strip() for c in line.split("|")]
                if "HBM bytes per launch" in cells:
                    col = cells.index("HBM bytes per launch")
                elif col is not None and len(cells) > col and want in cells[1:3]:
                    return int(cells[col].replace(",", ""))
    except (OSError, ValueError):
        return None
    return None


TRACE_FILE = os.path.join(ROOT, "profiles", "kernels_pcorr_poisson_1M.md")
# kernel names of a rocprofv3 trace -> entries of time_kernels() that time the same launches
TRACE_KEYS = {
    "phase::kLSolve<float, 3, 4>": "lsolve32cg_launch", "phase::kLUSolve<float, 3, 4>": "lusolve32cg_launch",
    "phase::kLSolve<float, 0, 4>": "lsolve32_launch", "phase::kLUSolve<float, 0, 4>": "lusolve32_launch",
    "phase::kUSolve<float, 4>": "usolve32_launch",
    "phase::kLSolve<double, 3, 4>": "lsolve64cg_launch", "phase::kLUSolve<double, 3, 4>": "lusolve64cg_launch",
    "phase::kLSolve<double, 0, 4>": "lsolve64_launch", "phase::kLUSolve<double, 0, 4>": "lusolve64_launch",
    "phase::kUSolve<double, 4>": "usolve64_launch", "phase::kSpmv<double>": "spmv64",
    "phase::kMlResid<float, double>": "mlresid32", "phase::kMlResid<double, double>": "mlresid64",
}


def dominant_from_trace():
    """Kernel name of the first data row of the first table in the committed
    kernel-trace summary of the headline method (rows sorted by total time)."""
    try:
        with open(TRACE_FILE) as f:
            for line in f:
                cells = [c.strip() for c in line.split("|")]
                if len(cells) > 2 and cells[1].startswith("`phase::"):
                    return cells[1].strip("`")
    except OSError:
        pass
    return None


def copy_probe(mib):
    """One plain copy of `mib` MiB (read once, written once; kCopy16) for the
    counter scale factors of the --pmc passes."""
    import phase_amd as P
    n = mib * (1 << 20) // 8
    a, b = P.DeviceArray.zeros(n), P.DeviceArray.zeros(n)
    b.copy_from(a)
    log("copy probe: %d bytes read, %d bytes written" % (8 * n, 8 * n))
    a.free()
    b.free()


def clean(v):
    """JSON-safe number: non-finite -> null, floats to 4 significant digits."""
    if isinstance(v, float):
        return float("%.4g" % v) if np.isfinite(v) else None
    return v


FALLBACK_RATIO = 0.5  # set from --fallback-ratio


def apply_setup(s, args):
    """The extra solver setup of --setup (key=value,...)."""
    for item in filter(None, args.setup.split(",")):
        key, value = item.split("=", 1)
        s.setup(key, value)


def solver_kw(mode):
    """Solver set-up of a method name "<precision>[-cg|-gmres][-mg]": precision
    fp64, mixed or learned; -cg preconditioned CG (Belos "CG") and -gmres
    flexible restarted GMRES (FP64) instead of BiCGStab; -mg the aggregation
    multigrid preconditioner instead of ILU(0) alone."""
    parts = mode.split("-")
    solver = "cg" if "cg" in parts[1:] else "gmres" if "gmres" in parts[1:] else "bicgstab"
    kw = dict(mode=parts[0], solver=solver, preconditioner="multigrid" if "mg" in parts[1:] else "ilu")
    if parts[0] == "learned":
        kw["fallbackRatio"] = FALLBACK_RATIO
    return kw


def build_problem(n_cells, k, sigma):
    from phase_amd import meshgen as G

    ni = G.delaunay_cells_for(n_cells, k)
    xy, cptr, cind = G.delaunay_square(k, ni, seed=1)
    return xy, cptr, cind


def cpu_baseline(xy, cptr, cind, bc, g, b, iters_target):
    """The FP64 oracle on a bounded sample of the same solve: assembly +
    ILU(0) + 10 and 60 BiCGStab iterations, ms/step extrapolated to
    `iters_target` iterations (the GPU FP64 count; natural-order ILU needs at
    least as many); timed single-threaded and with the SpMV, dot products and
    vector updates over all the host's cores (the ILU sweeps stay sequential).
    The faster of the two is the reported baseline."""
    from oracle import oracle as O

    O.build()
    t0 = time.perf_counter()
    m = O.Mesh(xy, cptr, cind)
    e = O.laplacian(m, 3, gamma_face=g, bc_type=bc, phi_face=np.zeros(m.n_faces))
    e.eq(b)  # stored rhs = -b  ->  A x = b
    rp, ci, v, rhs = e.get()
    t_asm = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    runs = {}
    for threads in (1, cores):
        times = {}
        for it in (10, 60):
            t0 = time.perf_counter()
            if threads == 1:
                O.bicgstab_ilu0(rp, ci, v, -rhs, tol=1e-30, max_iters=it)
            else:
                O.bicgstab_ilu0_omp(rp, ci, v, -rhs, tol=1e-30, max_iters=it, threads=threads)
            times[it] = time.perf_counter() - t0
        per_iter = (times[60] - times[10]) / 50.0
        setup = max(times[10] - 10 * per_iter, 0.0)
        runs[threads] = (1e3 * (t_asm + setup + per_iter * iters_target), setup, per_iter)
    best = min(runs, key=lambda k: runs[k][0])
    text = "; ".join("%d thread%s: %.1f ms/iteration, ILU setup %.2fs, %.0f ms/step" %
                     (k, "" if k == 1 else "s", 1e3 * runs[k][2], runs[k][1], runs[k][0]) for k in sorted(runs))
    return dict(value=round(runs[best][0], 1), unit="ms/step", cores=best, kind="port", extrapolated=True,
                single_thread_ms=round(runs[1][0], 1),
                sample="extrapolated: oracle FP64 BiCGStab+ILU(0), assembly %.2fs + 10 and 60 iterations measured, "
                       "scaled to %d iterations (the GPU plain FP64 count); %s" % (t_asm, iters_target, text))


def train_network(s, mesh, d_g, d_bc, d_pf, args, kw, max_pairs=256, capture_solves=None):
    """The learned part's setup, outside the timed steps: FP64 solves of earlier
    'time steps' (right-hand sides from seeds 100, 101, ...) emit (residual,
    correction) pairs; the network trains on them with the device kernels."""
    from phase_amd import meshgen as G

    t0 = time.perf_counter()
    nsolves = args.capture_solves if capture_solves is None else capture_solves
    # pairs come from the plain solver (FP64 BiCGStab + ILU(0)), whose long
    # iteration histories hold many of them
    s.setup("mode", "fp64")
    s.setup("solver", "bicgstab")
    s.setup("preconditioner", "ilu")
    s.setup("learnedObjective", args.learned_objective)
    for t in range(nsolves):
        s.assemble_laplacian_device(mesh, 3, gamma_face=d_g, bc_type=d_bc, phi_face=d_pf)
        if t == 0:
            s.learned_capture(args.capture_every, max_pairs)
        s.sub_rhs(G.zero_mean_normal(mesh.n_cells, seed=100 + t))
        s.zero_guess()
        s.solve(raise_on_noconv=False)
    s.learned_capture(0, max_pairs)
    pairs = s.learned_n_pairs()
    t1 = time.perf_counter()
    first = s.learned_train(50, lr=args.lr, seed=1)
    loss = s.learned_train(args.train_steps, lr=args.lr, seed=2)
    t2 = time.perf_counter()
    for k, v in kw.items():
        s.setup(k, v)
    log("learned: %d pairs from %d FP64 solves (%.1fs), %d Adam steps (%.1fs), loss %.3f -> %.3f"
        % (pairs, nsolves, t1 - t0, args.train_steps + 50, t2 - t1, first, loss))
    return dict(train_pairs=pairs, train_steps=args.train_steps + 50, objective=args.learned_objective,
                capture_s=round(t1 - t0, 2),
                train_s=round(t2 - t1, 2), loss_first=round(first, 4), loss_last=round(loss, 4))


def host_path(args, csr, method):
    """The same solve through the host-buffer boundary, as the reference's
    FiniteVolumeEquation::solve drives SparseMatrixSolver: set(rowPtr, colInd,
    vals) from host vectors (pattern once, values every step), setRhs, solve,
    x() back to the host.  Without cell centroids the rows get graph
    aggregates.  Timed per step including both PCIe copies."""
    from phase_amd import hip as P

    rp, ci, v, rhs = csr
    b = -rhs  # setRhs(-rhs_), FiniteVolumeEquation.tpp:73
    s = P.HipSparseMatrixSolver(0, tolerance=args.tol, maxIters=50000, innerTolerance=args.inner_tol,
                                blockRows=args.block_rows, **solver_kw(method))
    s.set(rp, ci, v)
    s.set_rhs(b)
    s.solve(raise_on_noconv=False)
    times, iters = [], 0
    for _ in range(2):
        t0 = time.perf_counter()
        s.set(rp, ci, v)
        s.set_rhs(b)
        s.zero_guess()
        rel = s.solve(raise_on_noconv=False)
        x = s.x()
        times.append(time.perf_counter() - t0)
        iters = s.stats()["inner_iters"]
    s.close()
    ms = 1e3 * min(times)
    log("host path %s: %.2f ms/step (PCIe copies included), relres %.2e, %d iterations" % (method, ms, rel, iters))
    return dict(method=method, ms_per_step=round(ms, 2), relres=float("%.4g" % rel), iters=iters,
                ordering="graph aggregates (no centroids at this boundary)", x_norm=float("%.6g" % np.linalg.norm(x)))


def spmv_size_point(args):
    """SpMV bandwidth far beyond the 256 MB last-level cache: the pressure
    matrix of an n x n quad mesh (`args.spmv_cells` cells, lognormal face
    coefficients), assembled on the device; FP64 and FP32 SpMV timed back to
    back and cache-cold.  No solve."""
    from phase_amd import hip as P
    from phase_amd import meshgen as G

    n = int(round(np.sqrt(args.spmv_cells)))
    t0 = time.perf_counter()
    xy, cptr, cind = G.rectilinear(n, n)
    mesh = P.HipMesh(xy, cptr, cind)
    log("spmv point: %d cells meshed (%.1fs)" % (mesh.n_cells, time.perf_counter() - t0))
    f = mesh.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy)
    bc = np.where(pid == 2, P.BC_FIXED, P.BC_NORMAL_GRADIENT).astype(np.int32)
    g = G.lognormal_face_coeffs(mesh.n_faces, 1.0, seed=2)
    s = P.HipSparseMatrixSolver(0, blockRows=args.block_rows)
    s.assemble_laplacian_device(mesh, 3, gamma_face=P.DeviceArray(g), bc_type=P.DeviceArray(bc),
                                phi_face=P.DeviceArray(np.zeros(mesh.n_faces)))
    log("spmv point: assembled and ordered (%.1fs)" % (time.perf_counter() - t0))
    km, kb = s.time_kernels(5), s.kernel_bytes()
    s.close()
    keys = ("spmv64", "spmv64_cold", "spmv32", "spmv32_cold")
    out = dict(workload="pcorr-poisson-quads-%d" % mesh.n_cells, cells=mesh.n_cells,
               bytes={k: round(kb[k]) for k in ("spmv64", "spmv32")},
               ms={k: round(km[k], 5) for k in keys},
               gbs={k: round(kb[k] / (km[k] * 1e-3) / 1e9, 1) for k in keys})
    log("spmv point: %s" % out["gbs"])
    return out


def learned_case(args):
    """The learned-preconditioner case: the standalone pressure Poisson on the
    same Delaunay generator at `args.c4_cells` cells with lognormal face
    coefficients of sigma = 2.  Plain FP64, mixed, and learned + mixed with a
    network trained on up to `args.c4_pairs` (residual, correction) pairs that
    FP64 solves of other right-hand sides emit (training outside the timed
    steps)."""
    from phase_amd import hip as P
    from phase_amd import meshgen as G

    k = int(round(512 * np.sqrt(args.c4_cells / 1e6)))
    xy, cptr, cind = build_problem(args.c4_cells, k, 2.0)
    mesh = P.HipMesh(xy, cptr, cind)
    f = mesh.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy)
    bc = np.where(pid == 2, P.BC_FIXED, P.BC_NORMAL_GRADIENT).astype(np.int32)
    g = G.lognormal_face_coeffs(mesh.n_faces, 2.0, seed=2)
    b = G.zero_mean_normal(mesh.n_cells, seed=3)
    d_g, d_bc, d_pf, d_b = (P.DeviceArray(g), P.DeviceArray(bc), P.DeviceArray(np.zeros(mesh.n_faces)),
                            P.DeviceArray(b))
    out = dict(workload="pcorr-poisson-delaunay-sigma2", cells=mesh.n_cells, edge_points=k, sigma=2.0)
    for mode in args.c4_modes.split(","):
        kw = solver_kw(mode)
        prec = kw["mode"]
        s = P.HipSparseMatrixSolver(0, tolerance=args.tol, maxIters=100000, innerTolerance=args.inner_tol,
                                    blockRows=args.block_rows, ordering=args.ordering, **kw)
        apply_setup(s, args)

        def step():
            s.assemble_laplacian_device(mesh, 3, gamma_face=d_g, bc_type=d_bc, phi_face=d_pf)
            s.sub_rhs_device(d_b)
            s.zero_guess()
            return s.solve(raise_on_noconv=False)

        info = None
        if prec == "learned":
            every = max(1, args.capture_every // 4)
            per_solve = max(1, out["fp64"]["iters"] // every)
            a2 = argparse.Namespace(**vars(args))
            a2.capture_every = every
            # one solve more than the count suggests: a solve's snapshots fall at
            # multiples of `every` and the last may not complete a pair
            info = train_network(s, mesh, d_g, d_bc, d_pf, a2, kw, max_pairs=args.c4_pairs,
                                 capture_solves=1 + max(1, -(-args.c4_pairs // per_solve)))
        step()
        t0 = time.perf_counter()
        rel = step()
        dt = time.perf_counter() - t0
        st = s.stats()
        out[mode] = dict(ms_per_step=round(1e3 * dt, 2), relres=float("%.4g" % rel), iters=st["inner_iters"],
                         fallbacks=st["fallbacks"])
        if info is not None:
            out[mode].update(info)
        log("sigma2 %s: %.2f ms/step, relres %.2e, %d iterations" % (mode, 1e3 * dt, rel, st["inner_iters"]))
        if mode == "fp64":
            km, kb = s.time_kernels(5), s.kernel_bytes()
            out["kernel_ms"] = {k: round(v, 5) for k, v in km.items()}
            out["spmv_gbs"] = {k: round(kb[k] / (km[k] * 1e-3) / 1e9, 1) for k in ("spmv64", "spmv64_cold")}
            out["ordering"] = s.ordering_info()
        s.close()
    return out


def bubble_case(args):
    """The C3 rising-bubble pressure system on a 1x2 quad mesh stored
    unstructured: a circular bubble of radius 0.25 centred at (0.5, 0.5), face
    density rho = 1000 outside, 1 inside (sharp indicator from the face
    centroids; FractionalStepMultiphase::updateProperties,
    FractionalStepMultiphase.cpp:175-185), pressure equation
    laplacian(dt / rho_f, p) == src::div(u) with dt = 1e-4
    (FractionalStepMultiphase.cpp:152-154), p = 0 on the top patch and zero
    gradient elsewhere; src::div(u) stands in as a zero-mean N(0,1) field."""
    from phase_amd import hip as P
    from phase_amd import meshgen as G

    nx = int(round(np.sqrt(args.bubble_cells / 2)))
    xy, cptr, cind = G.rectilinear(nx, 2 * nx, width=1.0, height=2.0)
    mesh = P.HipMesh(xy, cptr, cind)
    f = mesh.faces()
    _, _, fc, _ = mesh.geometry()
    a, bn = xy[f["lnode"]], xy[f["rnode"]]
    # the mesh's own top row of nodes (i * (2 / n) need not round to 2.0)
    ytop = xy[:, 1].max()
    top = (f["rcell"] < 0) & (a[:, 1] == ytop) & (bn[:, 1] == ytop)
    assert top.any(), "no top boundary faces"
    bc = np.where(top, P.BC_FIXED, P.BC_NORMAL_GRADIENT).astype(np.int32)
    inside = (fc[:, 0] - 0.5) ** 2 + (fc[:, 1] - 0.5) ** 2 < 0.25 ** 2
    rho_f = np.where(inside, 1.0, 1000.0)
    gamma = 1e-4 / rho_f
    rhs = G.zero_mean_normal(mesh.n_cells, seed=7)
    d_g, d_bc, d_pf, d_b = (P.DeviceArray(gamma), P.DeviceArray(bc), P.DeviceArray(np.zeros(mesh.n_faces)),
                            P.DeviceArray(rhs))
    out = dict(workload="rising-bubble-pressure-1x2-quads", cells=mesh.n_cells, density_ratio=1000.0)
    for mode in args.bubble_modes.split(","):
        s = P.HipSparseMatrixSolver(0, tolerance=args.tol, maxIters=100000, blockRows=args.block_rows,
                                    ordering=args.ordering, **solver_kw(mode))
        apply_setup(s, args)

        def step():
            s.assemble_laplacian_device(mesh, 3, gamma_face=d_g, bc_type=d_bc, phi_face=d_pf)
            s.sub_rhs_device(d_b)
            s.zero_guess()
            return s.solve(raise_on_noconv=False)

        step()
        t0 = time.perf_counter()
        rel = step()
        dt = time.perf_counter() - t0
        st = s.stats()
        out[mode] = dict(ms_per_step=round(1e3 * dt, 2), relres=float("%.4g" % rel), iters=st["inner_iters"],
                         fallbacks=st["fallbacks"])
        log("bubble %s: %.2f ms/step, relres %.2e, %d iterations" % (mode, 1e3 * dt, rel, st["inner_iters"]))
        s.close()
    return out


def multiphase_case(args):
    """Whole FractionalStepMultiphase time steps on the device driver, with
    the CELESTE surface tension, on the C3 rising bubble: 1x2 quad mesh
    stored unstructured, bubble (gamma = 1, rho2 = 1, mu2 = 0.1) of radius
    0.25 at (0.5, 0.5) in rho1 = 1000, mu1 = 10; sigma 24.5, g = (0, -0.98),
    dt = 1e-4; smoothing kernel radius 2.1 cells (the reference's
    RisingBubble example: 0.021 at h = 0.01), pow8; walls u = 0, p = 0 on the
    top.  Each step: volume-fraction solve, properties and surface force,
    then `args.multiphase_correctors` passes of momentum solve, pressure solve
    and corrections (1, the reference's fractional step: with two passes this
    step diverges, because dt = 1e-4 is above the capillary time-step limit of
    these 1e-3 cells, about 5.7e-5, and the pass iteration does not contract
    there)."""
    from phase_amd import hip as P
    from phase_amd import meshgen as G

    nx = int(round(np.sqrt(args.multiphase_cells / 2)))
    h = 1.0 / nx
    xy, cptr, cind = G.rectilinear(nx, 2 * nx, width=1.0, height=2.0)
    mesh = P.HipMesh(xy, cptr, cind)
    f = mesh.faces()
    _, cc, _, _ = mesh.geometry()
    a, bn = xy[f["lnode"]], xy[f["rnode"]]
    # the mesh's own top row of nodes (i * (2 / n) need not round to 2.0)
    ytop = xy[:, 1].max()
    top = (f["rcell"] < 0) & (a[:, 1] == ytop) & (bn[:, 1] == ytop)
    assert top.any(), "no top boundary faces"
    pbc = np.where(top, P.BC_FIXED, P.BC_NORMAL_GRADIENT).astype(np.int32)
    ubc = np.full(mesh.n_faces, P.BC_FIXED, np.int32)
    r = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 0.5)
    gamma = np.where(r < 0.25, 1.0, 0.0)
    t0 = time.perf_counter()
    d = P.HipMultiphase(mesh, 1000.0, 1.0, 10.0, 0.1, (0.0, -0.98), ubc, np.zeros(2 * mesh.n_faces), gamma,
                        p_bc=pbc, p_face0=np.zeros(mesh.n_faces))
    for k in ("gamma", "u", "p"):
        d.setup(k + ".tolerance", args.tol)
        d.setup(k + ".maxIters", 20000)
        d.setup(k + ".mode", args.multiphase_mode)
    d.setup("p.solver", "cg")
    d.setup("p.preconditioner", args.multiphase_precond)
    d.setup("numPressureCorrections", args.multiphase_correctors)
    d.surface_tension(24.5, 2.1 * h, "pow8")
    setup_s = time.perf_counter() - t0
    dt = args.multiphase_dt
    d.advance(dt, 2)
    t0 = time.perf_counter()
    hist = d.advance(dt, args.multiphase_steps)
    ms = 1e3 * (time.perf_counter() - t0) / args.multiphase_steps
    u, p, uf, g = d.fields()
    out = dict(workload="rising-bubble-multiphase-step-1x2-quads", cells=mesh.n_cells, ms_per_step=round(ms, 2),
               steps=args.multiphase_steps, dt=args.multiphase_dt, setup_s=round(setup_s, 2), mode=args.multiphase_mode,
               pressure_preconditioner=args.multiphase_precond, pressure_correctors=args.multiphase_correctors,
               iters_gamma_u_p=[float(x) for x in hist.mean(0)[:3]], max_div=float("%.3g" % hist[:, 3].max()),
               max_u=float("%.4g" % np.abs(u).max()))
    log("multiphase: %.2f ms/step (setup %.1f s), iterations gamma/u/p %s, max div %.2e" %
        (ms, setup_s, hist.mean(0)[:3], hist[:, 3].max()))
    d.close()
    return out


def ib_case(args):
    """The C5 pressure system with the ghost-cell immersed boundary: 30 x 20
    channel of quads stored unstructured, cylinder of diameter 1 at (10, 10);
    FractionalStepGCIB::solvePEqn's laplacian(dt / rho, p) + ib.bcs(p) ==
    src::div(u) (FractionalStepGCIB.cpp:44-46) with rho = 1, dt = 2e-3, the
    laplacian rows on the fluid cells only (remove(cell)), NORMAL_GRADIENT
    ghost rows on the cylinder (scaled by -4 dt h / rho, to a fluid row's
    diagonal, as the time-step driver scales them), p = 0 on the outlet
    (x = 30) and zero gradient elsewhere; src::div(u) stands in as a
    zero-mean N(0,1) field on the fluid cells.  The ghost rows are not
    symmetric: BiCGStab and flexible GMRES, plain FP64 against mixed
    precision, to 1e-8."""
    from phase_amd import hip as P
    from phase_amd import meshgen as G

    ny = int(round(np.sqrt(args.ib_cells / 1.5)))
    nx = int(round(1.5 * ny))
    xy, cptr, cind = G.rectilinear(nx, ny, width=30.0, height=20.0)
    mesh = P.HipMesh(xy, cptr, cind)
    f = mesh.faces()
    a, bn = xy[f["lnode"]], xy[f["rnode"]]
    # the mesh's own last column of nodes: i * (30 / nx) need not round to
    # 30.0 (at nx = 1732 it does not), and without outlet faces the pressure
    # system would be all-Neumann, singular
    xr = xy[:, 0].max()
    outlet = (f["rcell"] < 0) & (a[:, 0] == xr) & (bn[:, 0] == xr)
    assert outlet.any(), "no outlet faces"
    bc = np.where(outlet, P.BC_FIXED, P.BC_NORMAL_GRADIENT).astype(np.int32)
    ib = P.HipGhostCellIB(mesh, [(10.0, 10.0, 0.5)])
    types, _ = ib.cell_types()
    rhs = G.zero_mean_normal(mesh.n_cells, seed=11) * (types == 0)
    d_bc, d_pf, d_b = P.DeviceArray(bc), P.DeviceArray(np.zeros(mesh.n_faces)), P.DeviceArray(rhs)
    d_mask = P.DeviceArray.zeros(mesh.n_cells, np.int32)
    ib.cell_types_device(d_mask)
    eq = P.HipEquation(mesh, 1)
    ibq = P.HipEquation(mesh, 1)
    h = np.sqrt(30.0 * 20.0 / mesh.n_cells)  # the ghost cells' size (uniform quads)
    out = dict(workload="cylinder-ghost-cell-pressure-30x20-quads", cells=mesh.n_cells,
               ghost_cells=int((types == 1).sum()), solid_cells=int((types == 2).sum()))
    for mode in args.ib_modes.split(","):
        s = P.HipSparseMatrixSolver(0, tolerance=args.tol, maxIters=100000, **solver_kw(mode))

        def step():
            nonlocal eq
            eq.laplacian(1, gamma=2e-3, bc_type=d_bc, phi_face=d_pf)
            eq.remove_cells(d_mask)
            # the ghost rows scaled to a fluid row's diagonal in size and sign
            # (-4 dt h / rho; their coefficients are of order 1 / h), as the
            # time step does
            eq += ibq.gcib_bcs(ib, P.BC_NORMAL_GRADIENT, 0.0).scale(-4 * 2e-3 * h)
            eq.sub_vector(d_b)
            s.set_equation(eq)
            s.zero_guess()
            return s.solve(raise_on_noconv=False)

        step()
        t0 = time.perf_counter()
        rel = step()
        dt = time.perf_counter() - t0
        st = s.stats()
        out[mode] = dict(ms_per_step=round(1e3 * dt, 2), relres=float("%.4g" % rel), iters=st["inner_iters"],
                         fallbacks=st["fallbacks"])
        log("ib %s: %.2f ms/step, relres %.2e, %d iterations" % (mode, 1e3 * dt, rel, st["inner_iters"]))
        s.close()
    return out


def ib_flow_case(args):
    """Whole FractionalStepGCIB time steps (C5 geometry): 30 x 20 channel of
    quads stored unstructured, fixed cylinder of diameter 1 at (10, 10) with
    ghost-cell velocity rows and NORMAL_GRADIENT pressure rows, inflow u = (1, 0)
    at x = 0, walls at y = 0 and 20, outlet at x = 30 (u zero gradient, p = 0);
    rho = 1, Re = 200 (mu = 0.005), dt = 2e-3; BiCGStab to 1e-8 for both
    equations, each warm-started from the previous step,
    `args.ib_flow_correctors` passes per step.  Reports ms per step and the body force
    (applyHydrodynamicForce) after the last step."""
    from phase_amd import hip as P
    from phase_amd import meshgen as G

    ny = int(round(np.sqrt(args.ib_flow_cells / 1.5)))
    nx = int(round(1.5 * ny))
    xy, cptr, cind = G.rectilinear(nx, ny, width=30.0, height=20.0)
    mesh = P.HipMesh(xy, cptr, cind)
    f = mesh.faces()
    a, bn = xy[f["lnode"]], xy[f["rnode"]]
    bnd = f["rcell"] < 0
    left = bnd & (a[:, 0] == 0.0) & (bn[:, 0] == 0.0)
    xr = xy[:, 0].max()  # the mesh's own outlet column (see ib_case)
    right = bnd & (a[:, 0] == xr) & (bn[:, 0] == xr)
    assert right.any(), "no outlet faces"
    ubc = np.where(right, P.BC_NORMAL_GRADIENT, P.BC_FIXED).astype(np.int32)
    ub = np.zeros((mesh.n_faces, 2))
    ub[left, 0] = 1.0
    pbc = np.where(right, P.BC_FIXED, P.BC_NORMAL_GRADIENT).astype(np.int32)
    t0 = time.perf_counter()
    ib = P.HipGhostCellIB(mesh, [(10.0, 10.0, 0.5)])
    d = P.HipFractionalStep(mesh, 1.0, 0.005, ubc, ub, p_bc=pbc, pin_cell=-1)
    d.immersed_boundary(ib, [(0.0, 0.0)], P.BC_NORMAL_GRADIENT, 0.0)
    for k in ("u", "p"):
        d.setup(k + ".tolerance", args.tol)
        d.setup(k + ".maxIters", 100000)
        d.setup(k + ".mode", args.ib_flow_mode)
    d.setup("p.preconditioner", args.ib_flow_precond)
    d.setup("p.solver", args.ib_flow_psolver)
    if args.ib_flow_pmode:
        d.setup("p.mode", args.ib_flow_pmode)
    d.setup("numPressureCorrections", args.ib_flow_correctors)
    setup_s = time.perf_counter() - t0
    d.advance(2e-3, 2)
    log("ib flow: %d cells, two warm-up steps done" % mesh.n_cells)
    t0 = time.perf_counter()
    hist = d.advance(2e-3, args.ib_flow_steps)
    ms = 1e3 * (time.perf_counter() - t0) / args.ib_flow_steps
    u, p, _ = d.fields()
    force = ib.hydrodynamic_force(1.0, 0.005, P.DeviceArray(u.reshape(-1)), P.DeviceArray(p))[0]
    out = dict(workload="cylinder-ghost-cell-flow-step-30x20-quads", cells=mesh.n_cells, ms_per_step=round(ms, 2),
               steps=args.ib_flow_steps, setup_s=round(setup_s, 2), mode=args.ib_flow_mode, pressure_mode=args.ib_flow_pmode or args.ib_flow_mode,
               pressure_preconditioner=args.ib_flow_precond, pressure_solver=args.ib_flow_psolver, pressure_correctors=args.ib_flow_correctors,
               iters_u_p=[float(x) for x in hist.mean(0)[:2]], max_div=float("%.3g" % hist[:, 2].max()),
               force=[float("%.5g" % x) for x in force])
    log("ib flow: %.2f ms/step (setup %.1f s), iterations u/p %s, max div %.2e, force %s" %
        (ms, setup_s, hist.mean(0)[:2], hist[:, 2].max(), force))
    d.close()
    return out


def simple_cavity(args):
    """The C1 case on the device SIMPLE driver: lid-driven cavity on the
    128x128 unit square, Re = 100 (rho 1, mu 0.01), alpha_u 0.7, alpha_p 0.3,
    `args.simple_iters` outer iterations from rest (FP64 inner solves to 1e-8).
    Reported beside the headline; the CPU figure is the oracle's restatement
    of the same iterations (direct inner solves, 1 thread) on a 10-iteration
    sample."""
    from phase_amd import hip as P
    from phase_amd import meshgen as G

    xy, cptr, cind = G.rectilinear(128, 128)
    mesh = P.HipMesh(xy, cptr, cind)
    f = mesh.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy)
    ubc = np.full(mesh.n_faces, P.BC_FIXED, np.int32)
    ub = np.zeros((mesh.n_faces, 2))
    ub[pid == 3, 0] = 1.0
    warm = P.HipSimple(mesh, 1.0, 0.01, 0.7, 0.3, ubc, ub, 0)
    warm.iterate(3)
    warm.close()
    s = P.HipSimple(mesh, 1.0, 0.01, 0.7, 0.3, ubc, ub, 0)
    t0 = time.perf_counter()
    hist = s.iterate(args.simple_iters)
    dt = time.perf_counter() - t0
    out = dict(workload="lid-driven-cavity-128x128-re100", cells=mesh.n_cells, outer_iters=args.simple_iters,
               ms_per_outer=round(1e3 * dt / args.simple_iters, 3), total_s=round(dt, 3),
               mean_momentum_iters=round(float(hist[:, 0].mean()), 2),
               mean_pcorr_iters=round(float(hist[:, 1].mean()), 2),
               final_mass_imbalance=float("%.3g" % hist[-1, 2]), final_max_pcorr=float("%.3g" % hist[-1, 3]))
    s.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        from oracle.simple_oracle import SimpleOracle

        O.build()
        om = O.Mesh(xy, cptr, cind)
        ref = SimpleOracle(om, 1.0, 0.01, 0.7, 0.3, ubc, ub, pin=0)
        t0 = time.perf_counter()
        ref.iterate(10)
        out["cpu_port_ms_per_outer"] = round(1e3 * (time.perf_counter() - t0) / 10, 2)
    log("simple C1: %.3f ms per outer iteration (%d iterations), momentum %.1f / pCorr %.1f inner iterations"
        % (out["ms_per_outer"], args.simple_iters, out["mean_momentum_iters"], out["mean_pcorr_iters"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--edge", type=int, default=512)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--inner-tol", type=float, default=0.1)
    ap.add_argument("--block-rows", type=int, default=32, help="rows per ILU aggregate (blocks pack aggregates to 64)")
    ap.add_argument("--setup", default="",
                    help="extra solver setup of the C2, C3-pressure and C4 methods: key=value,...")
    ap.add_argument("--bubble-modes", default="fp64,fp64-cg,mixed-cg,fp64-cg-mg,mixed-cg-mg")
    ap.add_argument("--ordering", default="geometric", help="ILU aggregates: geometric (bricks) or graph")
    ap.add_argument("--modes", default="fp64,fp64-cg,mixed-cg,fp64-mg,fp64-cg-mg,mixed-cg-mg,learned-cg-mg",
                    help="precision fp64|mixed|learned, with -cg for preconditioned CG instead of BiCGStab and "
                         "-mg for the multigrid preconditioner instead of ILU(0) alone")
    ap.add_argument("--capture-solves", type=int, default=4)
    ap.add_argument("--capture-every", type=int, default=100)
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--simple-iters", type=int, default=200, help="C1 SIMPLE outer iterations (0 skips)")
    ap.add_argument("--bubble-cells", type=int, default=2_000_000, help="C3 bubble pressure solve (0 skips)")
    ap.add_argument("--multiphase-cells", type=int, default=2_000_000,
                    help="C3 whole multiphase time steps with surface tension (0 skips)")
    ap.add_argument("--multiphase-steps", type=int, default=5)
    ap.add_argument("--multiphase-dt", type=float, default=1e-4)
    ap.add_argument("--ib-flow-cells", type=int, default=1_000_000,
                    help="C5 FractionalStepGCIB time steps at this many cells (0 skips)")
    ap.add_argument("--ib-flow-steps", type=int, default=3)
    ap.add_argument("--ib-flow-mode", default="fp64")
    ap.add_argument("--ib-cells", type=int, default=1_000_000,
                    help="C5 ghost-cell pressure system at this many cells (0 skips)")
    ap.add_argument("--multiphase-mode", default="mixed", help="solver mode of the multiphase driver's solves")
    ap.add_argument("--multiphase-precond", default="multigrid", help="pressure preconditioner (ilu or multigrid)")
    ap.add_argument("--ib-flow-precond", default="multigrid", help="pressure preconditioner (ilu or multigrid)")
    ap.add_argument("--ib-flow-psolver", default="gmres", help="pressure Krylov method (bicgstab or gmres)")
    ap.add_argument("--ib-flow-pmode", default="mixed", help="pressure precision (fp64 or mixed)")
    ap.add_argument("--ib-modes", default="fp64,mixed,fp64-mg,mixed-mg,fp64-gmres-mg,mixed-gmres-mg",
                    help="methods of the ghost-cell pressure case")
    ap.add_argument("--multiphase-correctors", type=int, default=1,
                    help="pressure-corrector passes per multiphase step (two diverge at this dt and mesh: the "
                         "surface-tension time-step limit is below dt)")
    ap.add_argument("--ib-flow-correctors", type=int, default=2,
                    help="pressure-corrector passes per immersed-boundary flow step")
    ap.add_argument("--copy-probe-mb", type=int, default=0, help="run one copy kernel of this size (counter scale)")
    ap.add_argument("--no-host-path", dest="host_path", action="store_false",
                    help="skip the host-buffer (PCIe-inclusive) solve")
    ap.add_argument("--c4-cells", type=int, default=1_000_000,
                    help="learned case, sigma 2, at this many cells (0 skips)")
    ap.add_argument("--c4-modes", default="fp64,fp64-cg,mixed-cg,fp64-cg-mg,mixed-cg-mg,learned-cg-mg")
    ap.add_argument("--spmv-cells", type=int, default=16_000_000,
                    help="SpMV bandwidth point beyond the last-level cache at this many cells (0 skips)")
    ap.add_argument("--c4-pairs", type=int, default=2000, help="training pairs for the learned case")
    ap.add_argument("--fallback-ratio", type=float, default=0.5,
                    help="learned modes: the residual ratio a network step must reach to be kept")
    ap.add_argument("--learned-objective", default="residual",
                    help="the network's training loss: residual (what the solver's step keeps) or error")
    args = ap.parse_args()
    global FALLBACK_RATIO
    FALLBACK_RATIO = args.fallback_ratio
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the path would shard by "
              "cell partition with one halo exchange per SpMV)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import phase_amd as P
    from phase_amd import meshgen as G

    t0 = time.perf_counter()
    xy, cptr, cind = build_problem(args.cells, args.edge, args.sigma)
    mesh = P.HipMesh(xy, cptr, cind)
    f = mesh.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy)
    bc = np.where(pid == 2, P.BC_FIXED, P.BC_NORMAL_GRADIENT).astype(np.int32)
    g = G.lognormal_face_coeffs(mesh.n_faces, args.sigma, seed=2)
    b = G.zero_mean_normal(mesh.n_cells, seed=3)
    d_g, d_bc, d_pf, d_b = (P.DeviceArray(g), P.DeviceArray(bc), P.DeviceArray(np.zeros(mesh.n_faces)),
                            P.DeviceArray(b))
    log("setup: %d cells, %d faces, %.1fs" % (mesh.n_cells, mesh.n_faces, time.perf_counter() - t0))

    results = {}
    kernels = None
    nnz = None
    for mode in args.modes.split(","):
        kw = solver_kw(mode)
        prec = kw["mode"]
        s = P.HipSparseMatrixSolver(0, tolerance=args.tol, maxIters=50000, innerTolerance=args.inner_tol,
                                    blockRows=args.block_rows, ordering=args.ordering, **kw)
        apply_setup(s, args)

        def step():
            s.assemble_laplacian_device(mesh, 3, gamma_face=d_g, bc_type=d_bc, phi_face=d_pf)
            s.sub_rhs_device(d_b)
            s.zero_guess()
            return s.solve(raise_on_noconv=False)

        train_info = None
        if prec == "learned":
            train_info = train_network(s, mesh, d_g, d_bc, d_pf, args, kw)
        for _ in range(args.warmup):
            step()
        t0 = time.perf_counter()
        rel = []
        for _ in range(args.steps):
            rel.append(step())
        dt = (time.perf_counter() - t0) / args.steps
        st = s.stats()
        results[mode] = dict(ms_per_step=1e3 * dt, relres=max(rel), iters=st["inner_iters"],
                             passes=st["outer_passes"], solve_ms=st["solve_ms"])
        results[mode]["fallbacks"] = st["fallbacks"]
        if train_info is not None:
            results[mode].update(train_info)
        log("%s: %.2f ms/step, relres %.2e, %d iterations, %d passes" % (mode, 1e3 * dt, max(rel),
                                                                          st["inner_iters"], st["outer_passes"]))
        if kernels is None:
            kernels = s.time_kernels(args.kernel_reps)
            kbytes = s.kernel_bytes()
            ordinfo = s.ordering_info()
            host_csr = s.export()
            nnz = len(host_csr[1])
        s.close()
    if args.copy_probe_mb > 0:
        copy_probe(args.copy_probe_mb)

    n = mesh.n_cells
    gbs = {k: kbytes[k] / (kernels[k] * 1e-3) / 1e9 for k in kernels if kernels[k] > 0}
    log("kernels (ms): " + ", ".join("%s %.4f" % (k, v) for k, v in kernels.items()))
    log("kernels (GB/s): " + ", ".join("%s %.0f" % (k, v) for k, v in gbs.items()))

    # headline: the fastest method that reached the FP64 goal
    ok = {k: v for k, v in results.items() if v["relres"] <= args.tol}
    head = min(ok, key=lambda k: ok[k]["ms_per_step"]) if ok else list(results)[0]
    # dominant kernel of the headline solve: the top row (largest share of
    # kernel time) of the committed kernel trace of the headline method; its
    # launches are timed live above with HIP events
    dom_name = dominant_from_trace()
    key = TRACE_KEYS.get(dom_name)
    if key is None or key not in kernels:
        dom_name, key = "phase::kSpmv<double>", "spmv64"
    roof = dict(kernel=dom_name, bound="hbm", achieved=round(gbs[key], 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(gbs[key] / PEAK_HBM_GBS, 4), traffic=measured_traffic(dom_name),
                algorithmic_bytes=round(kbytes[key]), avg_launch_ms=round(kernels[key], 5),
                timing_key=key, selected_from=os.path.relpath(TRACE_FILE, ROOT))

    out = {
        "metric": METRIC,
        "value": round(results[head]["ms_per_step"], 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(results[head]["ms_per_step"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head.partition("-")[0] == "fp64" else "f32-inner/f64-refine",
        "data": "synthetic",
        "config": {"workload": "pcorr-poisson-delaunay-1M", "cells": n, "faces": mesh.n_faces, "nnz": nnz,
                   "sigma": args.sigma, "tol": args.tol, "method": head},
        "modes": {k: {kk: clean(vv) for kk, vv in v.items()} for k, v in results.items()},
        "speedup_vs_plain_fp64": (round(results["fp64"]["ms_per_step"] / results[head]["ms_per_step"], 3)
                                  if "fp64" in results else None),
        "spmv_gbs": {"fp64": round(gbs["spmv64"], 1), "fp32": round(gbs["spmv32"], 1),
                     "fp64_cold": round(gbs["spmv64_cold"], 1), "fp32_cold": round(gbs["spmv32_cold"], 1)},
        "kernel_ms": {k: round(v, 5) for k, v in kernels.items()},
        "kernel_gbs": {k: round(v, 1) for k, v in gbs.items()},
        "ordering": ordinfo,
        "roofline": roof,
    }
    if args.host_path:
        out["host_path"] = host_path(args, host_csr, head)
    if args.simple_iters > 0:
        out["simple_c1"] = simple_cavity(args)
    if args.bubble_cells > 0:
        out["bubble_c3"] = bubble_case(args)
    if args.multiphase_cells > 0:
        out["multiphase_c3"] = multiphase_case(args)
    if args.ib_cells > 0:
        out["ib_c5"] = ib_case(args)
    if args.ib_flow_cells > 0:
        out["ib_flow_c5"] = ib_flow_case(args)
    if args.c4_cells > 0:
        out["learned_c4"] = learned_case(args)
    if args.spmv_cells > 0:
        out["spmv_size_point"] = spmv_size_point(args)
    if not args.no_cpu_baseline:
        it = results.get("fp64", results[head])["iters"]
        out["cpu_baseline"] = cpu_baseline(xy, cptr, cind, bc, g, b, it)
        log("cpu baseline: %s" % out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
