#!/usr/bin/env python3
"""Benchmark: flux-residual throughput (Mcells/s) of the HIP path on one MI355X.

Workload (one "step"): one full residual evaluation -- conservative->primitive,
boundary ghosts, node-stencil least-squares gradients + Venkatakrishnan
limiter, ghost gradients, and the HLLC/MUSCL + viscous flux residual -- on the
laminar flat-plate mesh: 256 x 128 x 64 hexahedra (2,097,152 cells), wall-normal
growth 1.08, +-10 % seeded interior jitter, reverse Cuthill-McKee order,
Mach 0.3, Re_L = 1e5, FP64.  Inputs are resident in HBM before timing.

Prints ONE JSON line.  Extra fields: the flux kernel's roofline, the per-stage
device times, implicit steps (Jacobian + BiCGSTAB/SGS) in the three precision
forms, the CPU baseline (the reference's own compiled code where available,
else the C++ oracle port), the time to a 1e-8 residual drop on a 262,144-cell
plate in four forms (time_to_drop), the learned warm start on a 4.19M-cell
plate timed to an equal drop (config4) and the 20.25M-tet transonic SA bump:
its residual and the forms timed to an equal drop (config5).

usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "distributed-cfd-solver_amd"))

METRIC = "Mcells/s flux residual; time to 1e-8 residual drop: plain vs mixed+learned"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", default="256x128x64", help="plate mesh ni x nj x nk")
    ap.add_argument("--flux-reps", type=int, default=20)
    ap.add_argument("--implicit-steps", type=int, default=2)
    ap.add_argument("--cpu-baseline", choices=["auto", "reference", "port", "off"], default="auto")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--drop-size", default="128x64x32",
                    help="plate mesh of the time-to-1e-8-drop runs ('off' skips them)")
    ap.add_argument("--drop-previous-update", action="store_true",
                    help="also time the previous-update guess alone (the network with all parameters zero)")
    ap.add_argument("--drop-max-steps", type=int, default=12000)
    ap.add_argument("--config4", default="256x64x256",
                    help="plate mesh of the learned warm start at 4.19M cells ('off' skips it)")
    ap.add_argument("--config4-cfl", type=float, default=100.0)
    ap.add_argument("--config4-steps", type=int, default=150,
                    help="implicit steps of the plain FP64 run that fix the drop the other forms are timed to")
    ap.add_argument("--config5", default="150",
                    help="hexahedra per direction of the transonic SA tetrahedral bump (each split into 6 Kuhn "
                         "tetrahedra; 150 = the 20.25M-tet case), or off")
    ap.add_argument("--config5-steps", type=int, default=100,
                    help="implicit steps of the plain FP64 run that fix the drop the other forms are timed to")
    ap.add_argument("--drop-profile", type=int, default=0, metavar="N",
                    help="profiling mode: N implicit steps of the time-to-drop workload in each form, nothing else")
    ap.add_argument("--pmc-run", action="store_true",
                    help="profiling mode: one 1 GiB device copy (counter calibration) and flux-kernel launches only")
    return ap.parse_args()


def flux_algorithmic_bytes(nc, nbf, nf, viscous=True):
    # per evaluation: cells 272 B (q 40, grad 120, phi 40, centroid 24, R 40, lambda 8),
    # boundary ghosts g B, faces f B (owner+neigh 8, normal 24, area 8, centroid 24 [+ d 24, 1/d 8])
    g, f = (160, 96) if viscous else (40, 64)
    return 272 * nc + g * nbf + f * nf


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: one GPU only (requested --gpus {a.gpus}); the path is not sharded across GPUs",
              file=sys.stderr)
        sys.exit(2)
    if a.drop_profile > 0:
        print(json.dumps(drop_profile(a)), flush=True)
        return

    import numpy as np

    import cfdg
    import cfdg.cases as cases

    ni, nj, nk = (int(x) for x in a.size.split("x"))
    case = cases.plate(ni=ni, nj=nj, nk=nk)
    t0 = time.time()
    mesh = cfdg.Mesh.generate(**case.mesh)
    t_mesh = time.time() - t0
    t0 = time.time()
    solver = cfdg.Solver(mesh, case.numerics, case.bcs)
    t_setup = time.time() - t0
    nc, nf, ninner = mesh.n_cells, mesh.n_faces, mesh.n_inner_faces
    nbf = nf - ninner

    solver.init_uniform(case.rho, case.velocity, case.p)
    U = solver.get_state()
    U = U * (1.0 + 0.02 * np.random.default_rng(1234).standard_normal(U.shape))
    solver.set_state(U)

    if a.pmc_run:
        # counter calibration: one device-to-device copy of exactly 1 GiB (the HIP
        # runtime's copy kernel), then the residual stages and flux launches
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        nbytes = 1 << 30
        src, dst = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
        assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
        assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        solver.eval_residual()
        solver.time_flux_kernel(5)
        solver.implicit_step(cases.default_linparams())
        print(json.dumps({"pmc_run": True, "copy_bytes": nbytes, "flux_launches": 5 + 2}))
        return

    for _ in range(a.warmup):
        solver.eval_residual()
    stage_acc = np.zeros(6)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        solver.eval_residual()
        stage_acc += solver.stage_times()
    t_total = time.perf_counter() - t0
    ms_per_step = 1e3 * t_total / a.steps
    value = nc / (t_total / a.steps) / 1e6
    stage_ms = (stage_acc / a.steps).tolist()

    flux_ms = solver.time_flux_kernel(a.flux_reps)
    B = flux_algorithmic_bytes(nc, nbf, nf, viscous=True)
    achieved = B / (flux_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "flux_kernel_plate_2M_traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                tj = json.load(f)
            if tj.get("cells") == nc:
                traffic = tj.get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    # one full implicit step (residual, dt, Jacobian, BiCGSTAB + multicolour SGS, update)
    lp = cases.default_linparams()

    def implicit_steps():
        rows = []
        for _ in range(a.implicit_steps):
            t1 = time.perf_counter()
            st = solver.implicit_step(lp)
            rows.append({"ms": 1e3 * (time.perf_counter() - t1), "linear_iterations": st.linear.iterations,
                         "linear_status": st.linear.status, "jacobian_ms": float(solver.stage_times()[4]),
                         "linear_ms": float(solver.stage_times()[5])})
        return rows

    imp = implicit_steps()
    # the same steps from the same state in the two reduced-precision forms
    imp_forms = {}
    for name, prec in (("fp32_blocks_fp32_krylov", cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, 8, 0.1)),
                       ("bf16_blocks_fp32_krylov", cfdg.Precision.make(cfdg.PREC_BF16, cfdg.PREC_FP32, 16, 0.1))):
        solver.set_precision(prec)
        solver.set_state(U)
        imp_forms[name] = implicit_steps()
    solver.set_precision(cfdg.Precision.make())
    solver.set_state(U)

    cpu = None
    if a.cpu_baseline != "off":
        cpu = cpu_baseline(a, mesh, case, U, nc)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"laminar flat plate, {ni}x{nj}x{nk} hexahedra ({nc} cells, {nf} faces), y-growth 1.08, "
                        "10% jitter, RCM order, Mach 0.3, Re_L 1e5; one step = one full residual evaluation "
                        "(primitives, ghosts, LSQ-node gradients + Venkatakrishnan limiter, ghost gradients, "
                        "HLLC/MUSCL + viscous flux residual)",
            "cells": nc, "faces": nf, "boundary_faces": nbf,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_flux (flux residual)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": B,
            "ms_per_launch": round(flux_ms, 5),
        },
        "cpu_baseline": cpu,
        "stream_ceiling": stream_ceiling(),
        "stage_ms": {"primitives_ghosts": round(stage_ms[0], 4), "gradients_limiter": round(stage_ms[1], 4),
                     "ghost_gradients": round(stage_ms[2], 4), "flux": round(stage_ms[3], 4)},
        "flux_kernel_mcells_s": round(nc / (flux_ms * 1e-3) / 1e6, 1),
        "implicit_step": imp,
        "implicit_step_forms": imp_forms,
        "setup_s": {"mesh": round(t_mesh, 2), "context": round(t_setup, 2)},
    }
    solver.close()
    gnn = None
    if a.drop_size != "off":
        out["time_to_drop"], gnn = time_to_drop(a)
    if a.config4 != "off":
        out["config4"] = config4(a)
    if a.config5 != "off":
        out["config5"] = config5(a, int(a.config5), gnn)
    print(json.dumps(out))


def sa_flux_algorithmic_bytes(nc, nbf, nf):
    # the viscous count plus the eddy viscosity of cells and ghosts (8 B each)
    return flux_algorithmic_bytes(nc, nbf, nf, viscous=True) + 8 * (nc + nbf)


def equal_drop(s, U0, forms, N, out, tag, reset=None):
    """The first form takes N implicit steps from U0 and fixes the relative
    drop it reaches; every other form is timed from U0 to that drop (at most
    4 N steps).  Adds a row per form and the speedups over the first to out."""
    import cfdg.cases as cases

    lp = cases.default_linparams()
    target = None
    for name, prec, net in forms:
        s.set_precision(prec)
        s.set_state(U0)
        if reset is not None:
            reset()
        s.set_warm_start(net)
        if target is None:
            r = s.run_steady(lp, N, rel_tol=0.0, check_interval=1)
            target = r.rel
        else:
            r = s.run_steady(lp, 4 * N, rel_tol=target, check_interval=1)
        row = {"seconds": round(r.seconds, 3), "steps": int(r.steps), "linear_iterations": int(r.linear_iterations),
               "rel": r.rel, "reached": bool(r.rel <= target), "ms_per_step": round(1e3 * r.seconds / max(r.steps, 1), 2)}
        if net is not None:
            row["warm_start_kept_rejected"] = list(s.warm_start_stats())
        out[name] = row
        print(f"bench.py: {tag} {name}: {row}", file=sys.stderr, flush=True)
    s.set_warm_start(None)
    out["drop"] = target
    first = out[forms[0][0]]["seconds"]
    for name, _, _ in forms[1:]:
        out[f"speedup_{name}"] = round(first / out[name]["seconds"], 3)


def config4(a):
    """The learned warm start on a 4.19M-cell plate: 256 x 64 x 256
    hexahedra, the laminar plate's box, stretching (1.08), 10 % jitter, RCM
    order and flow (Mach 0.3, Re_L 1e5), second-order unlimited MUSCL with the
    node LSQ gradient, CFL 100, from the freestream with 2 % seeded noise.
    The network is trained first on 8 seeded 32 x 16 x 8 plates under the
    same numerics (train_network).  The plain FP64 run takes a fixed number
    of implicit steps; FP32 Krylov with a zero initial guess and with the
    learned guess (zero guess as fallback, unchanged linear tolerance) are
    timed to the drop it reached, from the same state."""
    import numpy as np

    import cfdg
    import cfdg.cases as cases

    ni, nj, nk = (int(x) for x in a.config4.split("x"))
    gnn, training = train_network(a.config4_cfl)
    print(f"bench.py: config4 training {training}", file=sys.stderr, flush=True)
    case = cases.plate(ni=ni, nj=nj, nk=nk, limiter="UNLIMITED", cfl=a.config4_cfl)
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    s.init_uniform(case.rho, case.velocity, case.p)
    U = s.get_state()
    U0 = U * (1.0 + 0.02 * np.random.default_rng(3).standard_normal(U.shape))
    print(f"bench.py: config4 mesh and context {mesh.n_cells} cells", file=sys.stderr, flush=True)
    fp32 = cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1)
    out = {"workload": f"laminar plate {ni}x{nj}x{nk} hexahedra ({mesh.n_cells} cells), Mach 0.3, Re_L 1e5, "
                       f"unlimited MUSCL + node LSQ, CFL {a.config4_cfl:g}, BiCGSTAB + multicolour SGS (rel 0.1), "
                       "2% seeded noise on the freestream", "training": training}
    forms = [("plain_fp64", cfdg.Precision.make(), None), ("fp32_krylov", fp32, None),
             ("fp32_krylov_learned", fp32, gnn)]
    equal_drop(s, U0, forms, a.config4_steps, out, "config4")
    out["learned_over_zero_guess"] = round(out["fp32_krylov"]["seconds"] / out["fp32_krylov_learned"]["seconds"], 3)
    s.close()
    return out


def config5(a, n, gnn):
    """The transonic RANS bump: n^3 hexahedra split into 6 Kuhn tetrahedra
    each (n = 150: 20.25M tets), 5 % seeded node jitter, Gaussian bump of
    height 0.1 and width 0.25, Mach 0.8, Re 1e6, viscous + Spalart-Allmaras
    with the reference's frozen module update, node LSQ + Venkatakrishnan
    MUSCL, CFL 10, from the freestream with 2 % seeded noise.

    - residual: one evaluation with the module (primitives, ghosts, LSQ
      gradients + limiter and the nu_tilde gradient, ghost gradients, flux
      and module residual), median of 10; the flux kernel's roofline;
    - forms: this case does not reach a 1e-8 drop under any setting of the
      reference (its relative residual stalls near 1e-2), so the plain FP64
      run takes a fixed number of implicit steps and every other form is
      timed to the drop that run reached, from the same state."""
    import numpy as np

    import cfdg
    import cfdg.cases as cases

    case = cases.bump_sa(ni=n, nj=n, nk=n)
    t0 = time.perf_counter()
    mesh = cfdg.Mesh.generate(**case.mesh)
    t_mesh = time.perf_counter() - t0
    print(f"bench.py: config5 mesh {mesh.n_cells} cells in {t_mesh:.1f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    t_ctx = time.perf_counter() - t0
    s.init_uniform(case.rho, case.velocity, case.p)
    U0 = s.get_state() * (1.0 + 0.02 * np.random.default_rng(1234).standard_normal((5, mesh.n_cells)))
    s.set_state(U0)
    t0 = time.perf_counter()
    s.set_turbulence(cases.turbulence(case, update=cfdg.TURB_FROZEN))
    t_turb = time.perf_counter() - t0
    print(f"bench.py: config5 context {t_ctx:.1f} s, turbulence {t_turb:.1f} s", file=sys.stderr, flush=True)
    rnu0 = s.turbulence()["rho_nu"]
    nc, nf, nbf = mesh.n_cells, mesh.n_faces, mesh.n_faces - mesh.n_inner_faces
    for _ in range(2):
        s.eval_residual()
    t = []
    for _ in range(10):
        t1 = time.perf_counter()
        s.eval_residual()
        t.append(time.perf_counter() - t1)
    ms = 1e3 * float(np.median(t))
    stages = s.stage_times()
    flux_ms = s.time_flux_kernel(10)
    B = sa_flux_algorithmic_bytes(nc, nbf, nf)
    out = {"workload": f"transonic bump, {n}x{n}x{n} hexahedra x 6 tetrahedra ({nc} cells, {nf} faces), "
                       "Mach 0.8, Re 1e6, viscous + Spalart-Allmaras (frozen module), node LSQ + "
                       "Venkatakrishnan MUSCL, CFL 10, 2% seeded noise",
           "setup_s": {"mesh": round(t_mesh, 1), "context": round(t_ctx, 1), "turbulence": round(t_turb, 1)},
           "residual": {"ms": round(ms, 3), "mcells_s": round(nc / (ms * 1e-3) / 1e6, 1),
                        "stage_ms": [round(float(x), 3) for x in stages[:4]]},
           "flux_roofline": {"kernel": "k_flux<true> (flux residual with eddy viscosity)",
                             "ms_per_launch": round(flux_ms, 4), "algorithmic_bytes": B,
                             "achieved": round(B / (flux_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                             "frac": round(B / (flux_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    print(f"bench.py: config5 residual {out['residual']}", file=sys.stderr, flush=True)
    forms = [("plain_fp64", cfdg.Precision.make(), None),
             ("fp32_krylov", cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1), None),
             ("bf16_blocks_fp32_krylov", cfdg.Precision.make(cfdg.PREC_BF16, cfdg.PREC_FP32, max_outer=8,
                                                             inner_rel_tol=0.1), None)]
    if gnn is not None:
        forms.append(("fp32_krylov_learned", cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=8,
                                                                 inner_rel_tol=0.1), gnn))
    equal_drop(s, U0, forms, a.config5_steps, out, "config5", lambda: s.set_turbulence_state(rnu0))
    s.close()
    return out


def stream_ceiling(nbytes=1 << 31, reps=10):
    """Measured HBM ceiling of this device, for context beside the 8 TB/s
    peak: the HIP runtime's device-to-device copy of nbytes (read + write
    bytes counted) and its fill of nbytes (write bytes), timed with HIP
    events over `reps` calls after warm-up."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
        vp = C.c_void_p
        src, dst = vp(), vp()
        if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) != 0 or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) != 0:
            return {"error": "hipMalloc"}
        e0, e1 = vp(), vp()
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        out = {}
        ops = (("copy", lambda: hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3, None), 2 * nbytes),
               ("fill", lambda: hip.hipMemsetAsync(dst, 0, C.c_size_t(nbytes), None), nbytes))
        for name, fn, moved in ops:
            for _ in range(3):
                fn()
            hip.hipDeviceSynchronize()
            hip.hipEventRecord(e0, None)
            for _ in range(reps):
                fn()
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = C.c_float()
            hip.hipEventElapsedTime(C.byref(ms), e0, e1)
            out[f"{name}_GBps"] = round(moved / (ms.value / reps * 1e-3) / 1e9, 1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(src)
        hip.hipFree(dst)
        out["bytes"] = nbytes
        out["note"] = "HIP runtime device-to-device copy (read + write bytes) and fill (write bytes) of 2 GiB"
        return out
    except OSError as e:  # the ceiling is context, not part of the measured path
        return {"error": str(e)}


def drop_profile(a):
    """N implicit steps of the time-to-drop workload (the same plate, numerics,
    noise and precision forms as time_to_drop) in the plain FP64 form and the
    two mixed forms, each after 20 untimed steps, and nothing else: the
    command a kernel trace of that workload runs under."""
    import numpy as np

    import cfdg
    import cfdg.cases as cases

    ni, nj, nk = (int(x) for x in a.drop_size.split("x"))
    lp = cases.default_linparams()
    forms = {"plain_fp64": None,
             "fp32_krylov": cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1),
             "bf16_blocks_fp32_krylov": cfdg.Precision.make(cfdg.PREC_BF16, cfdg.PREC_FP32, max_outer=8,
                                                            inner_rel_tol=0.1)}
    out = {"workload": f"laminar plate {ni}x{nj}x{nk} hexahedra, unlimited MUSCL, CFL 1000, 2% seeded noise",
           "steps": a.drop_profile}
    for name, prec in forms.items():
        case = cases.plate(ni=ni, nj=nj, nk=nk, limiter="UNLIMITED", cfl=1000.0)
        s = cfdg.Solver(cfdg.Mesh.generate(**case.mesh), case.numerics, case.bcs)
        s.init_uniform(case.rho, case.velocity, case.p)
        U = s.get_state()
        s.set_state(U * (1.0 + 0.02 * np.random.default_rng(1).standard_normal(U.shape)))
        if prec is not None:
            s.set_precision(prec)
        for _ in range(20):
            s.implicit_step(lp)
        its = 0
        t0 = time.perf_counter()
        for _ in range(a.drop_profile):
            its += s.implicit_step(lp).linear.iterations
        dt = time.perf_counter() - t0
        out[name] = {"ms_per_step": round(1e3 * dt / a.drop_profile, 3), "linear_iterations": its}
        print(f"bench.py: drop profile {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def train_network(cfl=1000.0):
    """The learned warm start's network, trained on the GPU: 8 seeded
    32 x 16 x 8 plates with Mach drawn from U(0.2, 0.5), unlimited MUSCL at
    the given CFL, 2 % noise on the freestream, every 100th of 2000 implicit
    steps recorded, 30 epochs.  Returns the network and a summary."""
    import numpy as np

    import cfdg
    import cfdg.cases as cases

    lp = cases.default_linparams()
    t0 = time.perf_counter()
    gnn = cfdg.Gnn(seed=11)
    rng = np.random.default_rng(2024)
    for m in range(8):
        tcase = cases.plate(ni=32, nj=16, nk=8, mach=float(rng.uniform(0.2, 0.5)), seed=100 + m,
                            limiter="UNLIMITED", cfl=cfl)
        mesh = cfdg.Mesh.generate(**tcase.mesh)
        s = cfdg.Solver(mesh, tcase.numerics, tcase.bcs)
        s.init_uniform(tcase.rho, tcase.velocity, tcase.p)
        U = s.get_state()
        s.set_state(U * (1.0 + 0.02 * np.random.default_rng(m).standard_normal(U.shape)))
        s.set_recording(True)
        for it in range(2000):
            s.implicit_step(lp)
            if it % 100 == 99:
                gnn.record(s)
        s.close()
    loss = gnn.train(epochs=30, lr=1e-3, seed=5)
    return gnn, {"seconds": round(time.perf_counter() - t0, 2), "samples": gnn.num_samples(), "final_loss": loss}


def time_to_drop(a):
    """Wall time to a 1e-8 drop of the reference's relative residual
    (max_v L2_v / L2_v of the first step, solver.hpp:740-765) on a steady
    laminar plate, for the plain FP64 path, the two mixed forms (FP32 or BF16
    blocks, FP32 Krylov inside FP64 defect correction) and the FP32 form with
    the learned warm start.  Second-order unlimited MUSCL with the node-LSQ
    gradient, CFL 1000, from the freestream with seeded 2 % noise; the
    Venkatakrishnan-limited scheme of the 2.1M-cell plate stalls this case
    near 5e-7 (limiter cycling), so it cannot be timed to 1e-8.  The network
    is trained first on 8 seeded 32x16x8 plates with Mach drawn from
    U(0.2, 0.5) (every 100th of 2000 steps recorded); its training time is
    reported separately.  After the timed run every form continues, untimed,
    by a further 1e-4 drop (to about 1e-12 overall), and its converged
    primitive fields and wall force are compared with the plain FP64 run's
    continued the same way.  With --drop-previous-update the previous-update
    guess alone (the network with all parameters zero) is timed too."""
    import numpy as np

    import cfdg
    import cfdg.cases as cases

    ni, nj, nk = (int(x) for x in a.drop_size.split("x"))
    kw = dict(limiter="UNLIMITED", cfl=1000.0)
    lp = cases.default_linparams()
    forms = {"fp32": cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1),
             "bf16": cfdg.Precision.make(cfdg.PREC_BF16, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1)}

    def start(case, seed):
        mesh = cfdg.Mesh.generate(**case.mesh)
        s = cfdg.Solver(mesh, case.numerics, case.bcs)
        s.init_uniform(case.rho, case.velocity, case.p)
        U = s.get_state()
        s.set_state(U * (1.0 + 0.02 * np.random.default_rng(seed).standard_normal(U.shape)))
        return s

    case = cases.plate(ni=ni, nj=nj, nk=nk, **kw)
    walls = [p for p, b in enumerate(case.bcs) if b.type == cfdg.BC["NO_SLIP_WALL_HEAT_FLUX"]]

    def primitives(U):
        rho = U[0]
        u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
        p = (cases.GAMMA - 1.0) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
        return np.stack([rho, u, v, w, p])

    def run(form, gnn=None):
        s = start(case, 1)
        if form != "plain":
            s.set_precision(forms[form])
        if gnn is not None:
            s.set_warm_start(gnn)
        r = s.run_steady(lp, a.drop_max_steps, rel_tol=1e-8)
        res = {"seconds": round(r.seconds, 3), "steps": int(r.steps), "linear_iterations": int(r.linear_iterations),
               "converged": r.status == cfdg.CONVERGED, "rel": r.rel}
        if gnn is not None:
            res["warm_start_kept_rejected"] = list(s.warm_start_stats())
        # untimed: a further 1e-4 drop, then the converged fields and wall force
        c = s.run_steady(lp, 8000, rel_tol=1e-4)
        res["continued"] = {"steps": int(c.steps), "reached": c.status == cfdg.CONVERGED,
                            "rel_overall_approx": r.rel * c.rel}
        s.eval_residual()
        _, _, force = s.boundary_integrals()
        q = primitives(s.get_state())
        s.close()
        return res, (q, force[walls].sum(axis=0))

    out = {"workload": f"laminar plate {ni}x{nj}x{nk} hexahedra, Mach 0.3, Re_L 1e5, unlimited MUSCL + node LSQ, "
                       "CFL 1000, BiCGSTAB + multicolour SGS (rel 0.1), 2% seeded noise on the freestream"}
    out["plain_fp64"], ref = run("plain")
    print(f"bench.py: drop plain_fp64 {out['plain_fp64']}", file=sys.stderr, flush=True)
    conv = {}
    out["fp32_krylov"], conv["fp32_krylov"] = run("fp32")
    print(f"bench.py: drop fp32_krylov {out['fp32_krylov']}", file=sys.stderr, flush=True)
    out["bf16_blocks_fp32_krylov"], conv["bf16_blocks_fp32_krylov"] = run("bf16")
    print(f"bench.py: drop bf16_blocks_fp32_krylov {out['bf16_blocks_fp32_krylov']}", file=sys.stderr, flush=True)
    if a.drop_previous_update:
        zero = cfdg.Gnn(seed=1)
        zero.set_params(np.zeros(zero.num_params, np.float32))
        out["fp32_krylov_previous_update"], _ = run("fp32", zero)
    gnn, out["training"] = train_network()
    print(f"bench.py: drop training {out['training']}", file=sys.stderr, flush=True)
    out["fp32_krylov_learned"], conv["fp32_krylov_learned"] = run("fp32", gnn)
    print(f"bench.py: drop fp32_krylov_learned {out['fp32_krylov_learned']}", file=sys.stderr, flush=True)
    # converged fields against the plain FP64 run's: rho, p against the
    # freestream values, u, v, w against the freestream speed; wall force relative
    speed = float(np.linalg.norm(case.velocity))
    scale = np.array([case.rho, speed, speed, speed, case.p])[:, None]
    q0, f0 = ref
    out["converged_error_vs_fp64"] = {
        name: {"rho": float(np.max(np.abs(q[0] - q0[0]) / scale[0])), "u": float(np.max(np.abs(q[1] - q0[1]) / scale[1])),
               "v": float(np.max(np.abs(q[2] - q0[2]) / scale[2])), "w": float(np.max(np.abs(q[3] - q0[3]) / scale[3])),
               "p": float(np.max(np.abs(q[4] - q0[4]) / scale[4])),
               "wall_force": float(np.linalg.norm(f - f0) / np.linalg.norm(f0))}
        for name, (q, f) in conv.items()}
    t_plain = out["plain_fp64"]["seconds"]
    out["speedup_mixed_fp32"] = round(t_plain / out["fp32_krylov"]["seconds"], 3)
    out["speedup_mixed_bf16"] = round(t_plain / out["bf16_blocks_fp32_krylov"]["seconds"], 3)
    out["speedup_mixed_fp32_learned"] = round(t_plain / out["fp32_krylov_learned"]["seconds"], 3)
    return out, gnn


def host_cpu():
    """Host cores and CPU model of the machine the bench runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": usable, "logical_cpus": os.cpu_count(), "model": model}


def cpu_baseline(a, mesh, case, U, nc):
    """The reference's Solver::evaluate_residual (compiled from its own sources,
    oracle/_ref) on the same mesh and state, or the C++ oracle port; one rank
    on one host core (the host's core count is recorded beside it)."""
    import numpy as np
    cores = 1
    host = host_cpu()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    if a.cpu_baseline in ("auto", "reference"):
        try:
            import ref_run
            if ref_run.available():
                times, _ = ref_run.time_residual(mesh.arrays(), mesh.n_patches, case.numerics, case.bcs, U,
                                                 reps=a.cpu_reps)
                t = float(np.median(times))
                return {"value": round(nc / t / 1e6, 4), "unit": "Mcells/s", "cores": cores, "kind": "reference",
                        "sample": f"{len(times)} residual evaluations of the full {nc}-cell mesh by the reference's "
                                  "Solver::evaluate_residual, 1 MPI rank, median",
                        "s_per_eval": round(t, 3), "host": host}
        except Exception as e:  # noqa: BLE001  (fall back to the port and say so)
            print(f"bench.py: reference CPU baseline unavailable ({e}); timing the oracle port", file=sys.stderr)
        if a.cpu_baseline == "reference":
            return None
    from oracle_py import Oracle
    o = Oracle(mesh, case.numerics, case.bcs)
    o.set_state(U)
    times = []
    for _ in range(a.cpu_reps):
        t0 = time.perf_counter()
        o.eval_residual()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": round(nc / t / 1e6, 4), "unit": "Mcells/s", "cores": cores, "kind": "port",
            "sample": f"{len(times)} residual evaluations of the full {nc}-cell mesh by the C++ oracle, median",
            "s_per_eval": round(t, 3), "host": host}


if __name__ == "__main__":
    main()
