"""Throughput benchmark of the 3D pure-qP VTI forward-modeling time loop on one MI355X.

Workload (the 512^3 single-GPU configuration): 512x512x512 interior, 32-cell PML on all
faces (padded 576^3), dx=dy=dz=12.5 m, dt=1 ms, 3000 time steps, 25 Hz Ricker at the surface
centre, the full 512x512 surface receiver plane, interior snapshots every 500 steps; model =
four VTI layers with a smooth +-3% random perturbation (seed 1). A bench "step" is one
complete shot (reset + 3000 time steps) with the model already resident in HBM; the timed
shots replay the captured 10-step hipGraph (--graph 0: direct launches).

metric: Gcell-updates/s = updated cells (nx-10)(ny-10)(nz-10) x time steps / wall time.
roofline: the dominant launch, the interior update kernel: its algorithmic bytes per launch
(24 B per interior cell updated: u_cur, u_prev, three model fields, u_next; 20 B where the
quiet map spares the eps - del load; 0 B for cells left out at rest), averaged over one
instrumented shot, over its mean launch time, from HIP events on the plan's stream around every
interior-update launch of a shot run right after the timed shots (replayed graphs carry no
per-launch events). config.roofline_all_cells counts every cell at 24 B.
roofline_step: the whole time step: algorithmic bytes of all its launches (the same per-cell
counts; 64 B per PML cell; 4 B per cell of the gradient tiles evaluated) over the timed shots'
wall time per time step.
cpu_baseline: the CPU oracle (the reference's arithmetic, bit-exact with its CPU kernels)
on the host's CPU share, on a 20-step window that starts from the GPU's state at step 300;
config.cpu_baseline_single_thread: the same on one thread (the reference's CPU path is
single-threaded) over 2 steps.

    python bench.py [--steps K] [--warmup W] [--cpu-steps S] [--no-cpu]

Other workloads (--workload): fwi (adjoint-state gradient), precision (16-bit storage against
the FP64 oracle), batch (many shot windows in one launch per stage), dropin (the same shot
through modeling3d() with host buffers, PCIe transfers included).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, os.path.join(REPO, "gpu-c_3dqmodeling_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_update_576cube_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calib_copy(nbytes):
    """A plain device-to-device copy of nbytes through the HIP runtime: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE give the counters' scale for a known traffic of nbytes each way."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    t = time.perf_counter()
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    log(f"calibration copy of {nbytes} bytes: {nbytes / (time.perf_counter() - t) / 1e9:.1f} GB/s (read) ")
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_baseline(setup, plan, start, runs):
    """Time the oracle for each (threads, nsteps) of `runs`, starting from the GPU's wavefield
    after `start` steps (threads 0 = the host's CPU share)."""
    from oracle import oracle as O

    plan.reset()
    plan.run(0, start)
    plan.sync()
    fields = [plan.field(w) for w in range(9)]
    out = []
    for threads, nsteps in runs:
        threads = threads or O.default_threads()
        o = O.OracleState(setup, "f32", nthreads=threads)
        o.import_state(*fields)
        t0 = time.perf_counter()
        for k in range(start, start + nsteps):
            o.step(setup.wavelet[k])
        dt = time.perf_counter() - t0
        o.close()
        out.append({"value": round(setup.n_upd * nsteps / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
                    "kind": "port",
                    "sample": f"CPU oracle (FP32, reference arithmetic) on the {setup.nx}^3 grid, {nsteps} time "
                              f"steps {start}..{start + nsteps - 1} from the GPU state at step {start}, "
                              f"{threads} thread(s); {dt:.1f} s"})
    del fields
    return out


def bench_fwi(args):
    """Adjoint-state FWI gradient: 256^3 interior, pml 32, 2000 steps, one shot at the surface
    centre, 256x256 surface receivers; observed data from the seed-2 model, gradient at that
    model smoothed with sigma = 15 cells; checkpoints every 50 steps (FP32 and BF16) and, with
    --fwi-full, full-wavefield storage. Then the finite-difference check on a 64^3 grid, 400
    steps, random joint directions in (vp, eps, del), against FP64 oracle misfits."""
    import qp3d
    n = 256 if args.n == 512 else args.n
    nt = 2000 if args.nt == 3000 else args.nt
    t = time.perf_counter()
    s_true = qp3d.config_setup(hx=n, hy=n, nz0=n, pml=args.pml, nt=nt, seed=2)
    s0 = qp3d.smoothed_config_setup(n, n, n, args.pml, nt, seed=2, sigma=15.0)
    log(f"fwi setup {s0.nx}^3, nt {nt}, {time.perf_counter() - t:.1f} s")
    with qp3d.Plan(s_true, gmax_mode=args.gmax_mode) as p:
        obs = p.forward()
        t = time.perf_counter()
        p.forward()
        t_fwd = time.perf_counter() - t
    log(f"plain forward shot {t_fwd:.3f} s")
    runs = {}
    ref = None
    # the adjoint sweep leaves out the planes its touched map shows at rest, except in the
    # rest_map_off run (same gradient, every plane swept)
    # rest map off: the same gradient, every plane swept; resident: the forward run keeps the
    # u_k of as many of the last segments as fit in HBM, which the gradient does not recompute
    cases = [("ckpt50_fp32", 50, "fp32", True, 1), ("ckpt50_fp32_rest_map_off", 50, "fp32", False, 1),
             ("ckpt50_bf16", 50, "bf16", True, 1), ("ckpt50_fp32_resident", 50, "fp32", True, -1)]
    if args.fwi_full:
        cases.append(("full_storage", nt, "fp32", True, 1))
    for key, interval, ck, rest, resident in cases:
        kept = 1
        try:
            with qp3d.Plan(s0, gmax_mode=args.gmax_mode) as p, qp3d.Fwi(p, interval=interval, ckpt_storage=ck) as f:
                f.set_rest_map(rest)
                if resident != 1:
                    kept = f.set_resident_segments(resident)
                f.forward()
                p.sync()
                t = time.perf_counter()
                rec = f.forward()
                g = f.gradient(rec - obs)
                wall = time.perf_counter() - t
                tm = f.times()
        except qp3d.Qp3dError as e:
            log(f"{key}: {e}")
            runs[key] = {"error": str(e)}
            continue
        # s_per_shot: wall time with the record, residual and gradients crossing PCIe (host
        # arrays); device_s_per_shot: the device time of the forward run, the recomputation and
        # the adjoint sweep
        dev = (tm["forward"] + tm["recompute"] + tm["adjoint"]) * 1e-3
        r = {"s_per_shot": round(wall, 3), "x_forward": round(wall / t_fwd, 2),
             "device_s_per_shot": round(dev, 3),
             "ms_forward": round(tm["forward"], 1), "ms_recompute": round(tm["recompute"], 1),
             "ms_adjoint": round(tm["adjoint"], 1), "misfit": float(0.5 * np.sum((rec.astype(np.float64) - obs) ** 2)),
             "resident_segments": kept}
        if ref is None:
            ref = g
        else:
            r["rel_diff_vs_ckpt50_fp32"] = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(g, ref)]
        runs[key] = r
        log(f"{key}: {r}")
        del g
    fd = fd_study(qp3d, args.fd_dirs) if args.fd_dirs > 0 else None
    out = {"metric": "FWI gradient s/shot (256^3, 2000 steps, checkpoints every 50)",
           "value": runs["ckpt50_fp32"]["s_per_shot"], "unit": "s/shot", "n_gpus": 1,
           "higher_is_better": False, "dtype": "fp32 (FP64 gradient accumulation)", "data": "synthetic",
           "config": {"workload": "adjoint-state FWI gradient w.r.t. vp, eps, del; 256^3 interior, pml 32, "
                                  "2000 steps, seed-2 true model, start model smoothed sigma=15",
                      "padded": [s0.nx, s0.ny, s0.nz], "plain_forward_s": round(t_fwd, 3), "runs": runs,
                      "fd_check": fd}}
    print(json.dumps(out), flush=True)


def fd_study(qp3d, ndirs, n=64, nt=400, pml=20, rel=1e-4):
    """Directional derivatives of the GPU adjoint against central differences of FP64 oracle
    misfits on the 64^3 grid (pml 20, 400 steps, seed-2 observed data, start model smoothed
    sigma = 15): random smooth directions cycling through vp, eps and del; the perturbed oracle
    runs replay the base run's threshold decisions (the linearisation the adjoint transposes;
    with free decisions the misfit jumps wherever a cell crosses the 0.01*G threshold)."""
    import dataclasses
    from scipy.ndimage import gaussian_filter
    from oracle import oracle as O
    s_true = qp3d.config_setup(hx=n, hy=n, nz0=n, pml=pml, nt=nt, seed=2)
    with qp3d.Plan(s_true) as p:
        obs = p.forward().astype(np.float64)
    s0 = qp3d.smoothed_config_setup(n, n, n, pml, nt, seed=2, sigma=15.0)
    o = O.OracleState(s0, "f64")
    bits = O.decision_bits(s0)
    o.set_decisions(bits, 1)
    o.run_record()
    o.close()
    with qp3d.Plan(s0) as p, qp3d.Fwi(p, interval=50) as f:
        rec = f.forward().astype(np.float64)
        g = f.gradient((rec - obs).astype(np.float32))
    out = {"grid": [s0.nx, s0.ny, s0.nz], "nt": nt, "step": rel, "decisions": "frozen at the base run"}
    t0 = time.perf_counter()
    names = ("vp", "eps", "dele")
    errs = {k: [] for k in names}
    for i in range(ndirs):
        name = names[i % 3]
        d = gaussian_filter(np.random.default_rng(100 + i).standard_normal((s0.nz, s0.nx, s0.ny)), 2.0)
        d = d / np.abs(d).max()
        h = rel * (float(np.abs(s0.vp).max()) if name == "vp" else 1.0)
        base = getattr(s0, name).astype(np.float64)
        mp, mm = (base + h * d).astype(np.float32), (base - h * d).astype(np.float32)
        jp = 0.5 * np.sum((O.record_frozen(dataclasses.replace(s0, **{name: mp}), bits) - obs) ** 2)
        jm = 0.5 * np.sum((O.record_frozen(dataclasses.replace(s0, **{name: mm}), bits) - obs) ** 2)
        dj = jp - jm
        ad = float(np.sum(g[names.index(name)] * (mp.astype(np.float64) - mm.astype(np.float64))))
        errs[name].append(abs(ad - dj) / abs(dj))
        log(f"fd {name} {i}: adjoint {ad:.9e} frozen fd {dj:.9e} rel {errs[name][-1]:.2e} "
            f"({time.perf_counter() - t0:.0f} s)")
    out["rel_err"] = {k: {"n": len(v), "max": float(max(v)), "median": float(np.median(v))} for k, v in errs.items() if v}
    return out


def bench_dropin(args):
    """The drop-in entry point with host buffers: qp3d_modeling3d (the reference's modeling3d
    signature) on the 512^3 shot, timed around the whole call (plan creation, model upload over
    PCIe, nt steps, record copy back, teardown); one untimed call first."""
    import qp3d
    s = qp3d.config_setup(hx=args.n, hy=args.n, nz0=args.n, pml=args.pml, nt=args.nt, seed=1, fm=25.0, src_depth=0)
    record = np.zeros((s.nt, s.hx, s.hy), np.float32)
    t = time.perf_counter()
    qp3d.modeling3d(s, record=record)
    log(f"untimed drop-in call {time.perf_counter() - t:.3f} s")
    t = time.perf_counter()
    qp3d.modeling3d(s, record=record)
    wall = time.perf_counter() - t
    host_bytes = 4 * (4 * s.nx * s.ny * s.nz + record.size)
    out = {"metric": f"Gcell-updates/s through modeling3d() with host buffers ({args.n}^3, {s.nt} steps)",
           "value": round(s.n_upd * s.nt / wall / 1e9, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "fp32", "data": "synthetic (seed 1)",
           "config": {"workload": "drop-in call including PCIe transfers", "padded": [s.nx, s.ny, s.nz], "nt": s.nt,
                      "s_per_call": round(wall, 3), "host_bytes_moved": host_bytes}}
    print(json.dumps(out), flush=True)


def bench_precision(args):
    """FP32, BF16 and FP16 storage against the FP64 oracle on a 256^3 crop of the 512^3 model
    (centre in x and y, top 256 cells in z), pml 32, 1500 steps: relative L2 gather error and
    peak snapshot error (steps 500 and 1000). Throughput of each storage at 512^3 comes from
    the forward workload with --storage."""
    import qp3d
    from oracle import oracle as O
    n, nt, crop = 512, 1500, 256
    vp, ep, de = qp3d.genmodel(n, n, n, 1)
    a, b = (n - crop) // 2, (n + crop) // 2
    g = [qp3d.file_to_grid(np.ascontiguousarray(f[a:b, a:b, :crop])) for f in (vp, ep, de)]
    s = qp3d.build_setup(*g, args.pml, nt, 12.5, 12.5, 12.5, 0.001, 25.0)
    log(f"precision crop {s.nx}^3, nt {nt}")
    res = {}
    gpu = {}
    for st in ("fp32", "bf16", "fp16"):
        with qp3d.Plan(s, storage=st, snapflag=1, ntsnap=500) as p:
            t = time.perf_counter()
            rec = p.forward()
            res[st] = {"s_per_shot": round(time.perf_counter() - t, 3)}
            gpu[st] = (rec.astype(np.float64), {k: v.astype(np.float64) for k, v in p.snapshots()})
    o = O.OracleState(s, "f64", nthreads=O.default_threads())
    orec = np.zeros((nt, s.hx, s.hy))
    osnap = {}
    t0 = time.perf_counter()
    pm = s.pml
    for k in range(nt):
        o.step(s.wavelet[k])
        u = o.field("u2")
        orec[k] = u[s.sz, pm:s.nx - pm, pm:s.ny - pm]
        if k % 500 == 0:
            osnap[k] = np.array(u[pm:s.nz - pm, pm:s.nx - pm, pm:s.ny - pm])
        if k % 50 == 0:
            log(f"oracle f64 step {k} ({time.perf_counter() - t0:.0f} s)")
    o.close()
    for st, (rec, snaps) in gpu.items():
        res[st]["gather_rel_l2"] = float(np.linalg.norm(rec - orec) / np.linalg.norm(orec))
        res[st]["snapshot_peak_err"] = {k: float(np.abs(snaps[k] - osnap[k]).max() / np.abs(osnap[k]).max())
                                        for k in (500, 1000) if k in snaps}
    out = {"metric": "relative L2 gather error vs FP64 oracle (256^3 crop, 1500 steps)",
           "value": res["bf16"]["gather_rel_l2"], "unit": "relative L2", "n_gpus": 1, "higher_is_better": False,
           "dtype": "fp32 arithmetic; storage fp32/bf16/fp16", "data": "synthetic",
           "config": {"workload": "256^3 crop of the seed-1 512^3 model, pml 32, 1500 steps",
                      "padded": [s.nx, s.ny, s.nz], "results": res}}
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """Batched shots: an 800x800x300 model (seed 3), shots on a random surface grid (seed 3),
    each with its own window of 160-320 cells in x and y by 300 in z (pml 32), 20 Hz Ricker at
    the window centre, the full surface receiver plane of the window, 2500 steps; all shots in
    the form the autotuner finds fastest: one persistent launch per run, or one launch per stage
    and step replayed from a captured graph (or each shot's own launches within a step), with
    its tile shapes. The other of the first two forms is timed too. The same shots run one after
    another (each plan alone, also with its graph) give the baseline."""
    import qp3d
    nt = 2500 if args.nt == 3000 else args.nt
    t = time.perf_counter()
    model = qp3d.genmodel(800, 800, args.depth, 3)
    setups = [qp3d.window_setup(model, x0, y0, wx, wy, args.pml, nt, 20.0)
              for x0, y0, wx, wy in qp3d.survey_windows(800, 800, args.shots, args.win[0], args.win[1], 3)]
    del model
    log(f"batch setup: {args.shots} shots in {time.perf_counter() - t:.1f} s")
    plans = [qp3d.Plan(s, storage=args.storage, use_graph=True) for s in setups]
    cells = sum(s.n_upd for s in setups) * nt
    out_runs = {}
    with qp3d.Batch(plans, use_graph=True) as b:
        t = time.perf_counter()
        tune = b.autotune(steps=4)
        log(f"autotune {tune} in {time.perf_counter() - t:.1f} s")
        def timed_run():
            b.reset()
            b.run(0, 20)
            b.sync()
            t0 = time.perf_counter()
            b.reset()
            b.run()
            b.sync()
            return time.perf_counter() - t0

        wall = timed_run()
        kb = b.step_bytes()
        log(f"batched ({'persistent' if tune['persistent'] >= 0 else 'launches'}): {wall:.2f} s")
        # the other form, forced: one persistent launch per run (the autotuner's fastest tile set
        # and cut), or the per-stage launches
        pset, pshare = b.persistent_choice()
        if tune["persistent"] >= 0:
            b.set_persistent(-1, False)
            b.set_per_shot(tune["per_shot"])
            wall_p, wall_l = wall, timed_run()
        else:
            b.set_persistent(pset, pshare)
            wall_p, wall_l = timed_run(), wall
        log(f"persistent launch (tile set {pset}, share {pshare}): {wall_p:.2f} s; launches: {wall_l:.2f} s")
    seq = None
    if args.sequential:
        t0 = time.perf_counter()
        for p in plans:
            p.reset()
            p.run()
            p.sync()
        seq = time.perf_counter() - t0
        log(f"one after another: {seq:.2f} s")
    for p in plans:
        p.close()
    step_bytes = kb["gradmax"] + kb["update"]
    out = {"metric": f"Gcell-updates/s ({args.shots} shots batched, {nt} steps)", "value": round(cells / wall / 1e9, 3),
           "unit": "Gcell-updates/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(wall * 1e3, 1),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32" if args.storage == "fp32" else f"fp32 arithmetic, {args.storage} storage",
           "data": "synthetic (layered VTI + smooth random perturbation, seed 3)",
           "config": {"workload": f"{args.shots} shots, windows {args.win[0]}-{args.win[1]} x {args.win[0]}-{args.win[1]} "
                                  f"x {args.depth} of an 800x800x{args.depth} model, "
                                  f"pml {args.pml}, {nt} steps, 20 Hz, storage {args.storage}",
                      "cells_updated_per_step": kb["cells_updated"], "autotuned_tiles": tune["tiles"],
                      "autotuned_per_shot_launches": tune["per_shot"],
                      "autotuned_items_per_shot_share": tune["share"],
                      "autotune_ms_per_step": [round(x, 4) for x in tune["ms"]],
                      "autotuned_persistent_tile_set": tune["persistent"],
                      "batched_s": round(wall, 3), "persistent_launch_s": round(wall_p, 3),
                      "persistent_tile_set_share": [pset, pshare], "stage_launches_s": round(wall_l, 3),
                      "sequential_s": round(seq, 3) if seq else None,
                      "speedup_vs_sequential": round(seq / wall, 3) if seq else None,
                      "hbm_gbs_algorithmic": round(step_bytes * nt / wall / 1e9, 1)}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots")
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--n", type=int, default=512, help="interior cells per axis")
    ap.add_argument("--pml", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-start", type=int, default=300)
    ap.add_argument("--cpu-steps-1", type=int, default=2, help="time steps of the single-thread CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", type=int, default=1, help="1: timed shots replay the captured graph; 0: direct")
    ap.add_argument("--tile-update", type=int, default=0)
    ap.add_argument("--tile-grad", type=int, default=0)
    ap.add_argument("--tile-pml", type=int, default=0)
    ap.add_argument("--tile-pml-y", type=int, default=0)
    ap.add_argument("--streams", type=int, default=0, help="2: interior and PML updates concurrent; 1: serial")
    ap.add_argument("--zchunk", type=int, default=0, help="z planes per update item (0: the plan's choice)")
    ap.add_argument("--gmax-mode", type=int, default=0,
                    help="1: gradient pass before the update; 2: predicted threshold with fix-up; 0: default")
    ap.add_argument("--storage", default="fp32", choices=["fp32", "bf16", "fp16"],
                    help="storage of wavefields and PML state (arithmetic is always FP32)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a 1 GiB device-to-device copy (the reference kernel for scaling HBM counters)")
    ap.add_argument("--shots", type=int, default=32, help="batch: number of shots")
    ap.add_argument("--win", type=int, nargs=2, default=[160, 320], help="batch: window size range in x and y")
    ap.add_argument("--depth", type=int, default=300, help="batch: model depth (cells)")
    ap.add_argument("--sequential", type=int, default=1, help="batch: also run the shots one after another")
    ap.add_argument("--amp-bound", type=int, default=1,
                    help="forward: gradient pass under the amplitude bound (1) or every tile evaluated (0)")
    ap.add_argument("--workload", default="forward", choices=["forward", "fwi", "precision", "batch", "dropin"],
                    help="forward: the 512^3 shot (default); fwi: the 256^3 adjoint gradient; "
                         "precision: FP32/BF16/FP16 storage against the FP64 oracle on a 256^3 crop; "
                         "dropin: the 512^3 shot through modeling3d() with host buffers")
    ap.add_argument("--fwi-full", action="store_true", help="fwi: also run with full-wavefield storage")
    ap.add_argument("--fd-dirs", type=int, default=20, help="fwi: directions of the finite-difference check")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the shot-parallel and z-slab "
              f"designs for several GPUs are described, not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "fwi":
        return bench_fwi(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "dropin":
        return bench_dropin(args)

    import qp3d

    if args.calib_copy:
        calib_copy(1 << 30)
    t = time.perf_counter()
    s = qp3d.config_setup(hx=args.n, hy=args.n, nz0=args.n, pml=args.pml, nt=args.nt, seed=1, fm=25.0, src_depth=0)
    log(f"setup {s.nx}x{s.ny}x{s.nz} in {time.perf_counter() - t:.1f} s")
    ntsnap = 500
    plan = qp3d.Plan(s, snapflag=1, ntsnap=ntsnap, use_graph=bool(args.graph), tile_update=args.tile_update,
                     tile_grad=args.tile_grad, tile_pml=args.tile_pml, tile_pml_y=args.tile_pml_y,
                     streams=args.streams, storage=args.storage, gmax_mode=args.gmax_mode, zchunk=args.zchunk)
    log("model uploaded")
    if not args.amp_bound:
        plan.set_amp_bound(False)
    for i in range(args.warmup):
        t = time.perf_counter()
        plan.forward_async()
        plan.sync()
        log(f"warmup shot {i}: {time.perf_counter() - t:.3f} s")
    # Timed shots: graph replays (no per-launch events); with --graph 0, direct launches with
    # events around the interior update only (two per step). After the timed region one shot
    # with events around the interior update only gives the roofline kernel's mean launch time,
    # and one shot with events around every launch the other launches' times.
    plan.set_instrument(0 if args.graph else 2)
    plan.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.forward_async()
        plan.sync()  # also harvests this shot's kernel events
    t1 = time.perf_counter()
    wall = t1 - t0
    kt_timed = plan.kernel_times()

    if args.graph:
        plan.set_instrument(2)
        plan.forward_async()
        plan.sync()
        kt_timed = plan.kernel_times()
    # interior-update time over the shot, in twelve stretches of steps (from rest to the end)
    st_int = plan.step_times()
    nb = max(1, s.nt // 12)
    int_bins = [round(float(st_int[i:i + nb].mean()), 4) for i in range(0, nb * 12, nb)] if s.nt >= 12 else None
    plan.set_instrument(1)
    plan.forward_async()
    plan.sync()
    kt = plan.kernel_times()
    gtiles = plan.gradient_tiles()  # evaluated, skipped over the instrumented shot
    ucells = plan.update_cells()  # per update kind: cells, at rest, without eps - del (same shot)
    kt["update_interior"] = kt_timed["update_interior"]
    kb = plan.step_bytes()
    plan.set_instrument(0)
    from qp3d.threshold import steps_in_band
    gmode = plan.gmax_mode()
    in_band = int(steps_in_band(plan.step_maxima(), plan.step_predictions()).sum()) if gmode == 2 else None
    log(f"timed {args.steps} shots in {wall:.3f} s; per step: {kt}")
    gbs = {k: kb[k] / (kt[k] * 1e-3) / 1e9 for k in ("gradmax", "update_interior", "update_pml") if kt[k] > 0}
    log(f"bytes per step {kb}; achieved GB/s {gbs}")

    cells = s.n_upd * s.nt * args.steps
    value = cells / wall / 1e9
    ms_time_step = wall / args.steps / s.nt * 1e3
    # Bytes the launches had to move: cells at rest are left out (0 B), quiet cells skip the
    # 4-byte eps - del load, the gradient pass reads only the tiles it evaluates. Without the
    # amplitude bound (or with gmax_mode 2) the counters stay 0 and every cell counts in full.
    b_int = kb["update_interior"] / kb["cells_interior"]
    b_pml = kb["update_pml"] / kb["cells_pml"] if kb["cells_pml"] else 0.0
    int_c = ucells[0]
    pml_c = ucells[1] + ucells[2]
    counted = int_c[0] > 0
    if counted:
        moved_int = ((int_c[0] - int_c[1]) * b_int - int_c[2] * 4.0) / s.nt
        moved_pml = ((pml_c[0] - pml_c[1]) * b_pml - pml_c[2] * 4.0) / s.nt
        geval = gtiles[0] / max(1, gtiles[0] + gtiles[1])
    else:
        moved_int, moved_pml, geval = kb["update_interior"], kb["update_pml"], 1.0
    moved_grad = kb["gradmax"] * geval
    achieved = moved_int / (kt["update_interior"] * 1e-3) / 1e9
    achieved_all = gbs["update_interior"]
    step_bytes = kb["gradmax"] + kb["update_interior"] + kb["update_pml"]
    step_moved = moved_grad + moved_int + moved_pml
    step_gbs = step_moved / (ms_time_step * 1e-3) / 1e9
    step_gbs_all = step_bytes / (ms_time_step * 1e-3) / 1e9
    upd_total = int(int_c[0] + pml_c[0])
    rest_frac = float(int_c[1] + pml_c[1]) / max(1, upd_total)
    quiet_frac = float(int_c[2] + pml_c[2]) / max(1, upd_total)
    # The counter passes (FETCH_SIZE, WRITE_SIZE) measured a launch that processes every cell,
    # on an earlier build: they give the kernels' re-read ratio at full load, not the traffic of
    # this build's launches (which leave out cells at rest), so they go into config, and
    # roofline.traffic stays null.
    traffic = step_traffic = None
    traffic_full = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            with open(TRAFFIC_FILE) as f:
                tr = json.load(f)
            if tr.get("nx") == s.nx and tr.get("storage", "fp32") == args.storage:
                traffic_full = {"interior_bytes_per_launch": tr.get("bytes_per_launch"),
                                "ratio_to_algorithmic": {k: v.get("ratio") for k, v in tr.get("per_kernel", {}).items()},
                                "step_bytes": tr.get("bytes_per_step"),
                                "source": "profiles/k_update_576cube_traffic.md (every cell processed; earlier build)"}
        except (OSError, ValueError):
            traffic_full = None
    cpu = cpu1 = None
    if not args.no_cpu and args.cpu_steps > 0:
        runs = [(0, args.cpu_steps)] + ([(1, args.cpu_steps_1)] if args.cpu_steps_1 > 0 else [])
        res = cpu_baseline(s, plan, min(args.cpu_start, s.nt - args.cpu_steps), runs)
        cpu = res[0]
        cpu1 = res[1] if len(res) > 1 else None
        log(f"cpu baseline {res}")
    plan.close()

    out = {
        "metric": f"Gcell-updates/s ({args.n}^3 FP32 forward, {s.nt} steps)" if args.storage == "fp32" else
                  f"Gcell-updates/s ({args.n}^3 forward, {args.storage.upper()} storage / FP32 arithmetic, {s.nt} steps)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.storage == "fp32" else f"fp32 arithmetic, {args.storage} storage",
        "data": "synthetic (layered VTI + smooth random perturbation, seed 1)",
        "config": {"workload": f"3D pure-qP VTI forward shot, {args.n}^3 interior, pml {args.pml} "
                               f"({s.nx}^3 padded), {s.nt} steps, dt 1 ms, dx 12.5 m, 25 Hz Ricker, "
                               f"{args.n}x{args.n} receivers, snapshots every 500 steps",
                   "step_unit": f"one shot: reset + {s.nt} time steps (ms_per_step is per shot)",
                   "ms_per_shot": round(wall / args.steps * 1e3, 3),
                   "ms_per_time_step": round(ms_time_step, 5),
                   "time_loop": "replayed 10-step hipGraph" if args.graph else "direct launches",
                   "padded": [s.nx, s.ny, s.nz], "nt": s.nt, "n_upd": s.n_upd, "n_int": s.n_int,
                   "gcells_int_per_s": round(s.n_int * s.nt * args.steps / wall / 1e9, 3),
                   "ms_gradmax": round(kt["gradmax"], 5), "ms_update_interior": round(kt["update_interior"], 5),
                   "ms_update_pml": round(kt["update_pml"], 5), "ms_update_stage": round(kt["update_stage"], 5),
                   "gbs_gradmax": round(gbs.get("gradmax", 0), 1), "gbs_update_pml": round(gbs.get("update_pml", 0), 1),
                   "storage": args.storage, "gmax_mode": gmode, "steps_in_predicted_band": in_band,
                   "gradient_amp_bound": bool(args.amp_bound) and gmode == 1,
                   "gradient_tiles_evaluated_skipped": list(gtiles),
                   "gradient_tiles_skipped_frac": round(gtiles[1] / max(1, gtiles[0] + gtiles[1]), 4),
                   "update_cells_at_rest_frac": round(rest_frac, 4) if counted else None,
                   "update_cells_quiet_frac": round(quiet_frac, 4) if counted else None,
                   "update_cells_at_rest_frac_interior": round(int_c[1] / max(1, int_c[0]), 4) if counted else None,
                   "ms_update_interior_by_twelfths": int_bins,
                   # the last twelfth of the shot: every interior cell active (the wave fills the
                   # grid long before), so each launch moves its full 24 B per cell
                   "roofline_interior_last_twelfth": ({"achieved": round(kb["update_interior"] / (int_bins[-1] * 1e-3) / 1e9, 1),
                                                       "frac": round(kb["update_interior"] / (int_bins[-1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                                                      if int_bins and int_bins[-1] > 0 else None),
                   "roofline_kernel": f"k_update<interior> ({b_int:.0f} B per interior cell updated, "
                                      f"{b_int - 4:.0f} B without eps - del, 0 B at rest; "
                                      f"{moved_int:.0f} B per launch over the shot, "
                                      f"{kb['cells_interior']:.0f} cells per launch)",
                   "traffic_full_load": traffic_full,
                   "roofline_all_cells": {"achieved": round(achieved_all, 1), "frac": round(achieved_all / HBM_PEAK_GBS, 4),
                                          "step_achieved": round(step_gbs_all, 1),
                                          "step_frac": round(step_gbs_all / HBM_PEAK_GBS, 4), "step_bytes": step_bytes},
                   "cpu_baseline_single_thread": cpu1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "roofline_step": {"bound": "hbm", "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4), "traffic": step_traffic,
                          "bytes": round(step_moved)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
