"""Throughput benchmark of the 3D pure-qP VTI forward-modeling time loop on one MI355X.

Workload (the 512^3 single-GPU configuration): 512x512x512 interior, 32-cell PML on all
faces (padded 576^3), dx=dy=dz=12.5 m, dt=1 ms, 3000 time steps, 25 Hz Ricker at the surface
centre, the full 512x512 surface receiver plane, interior snapshots every 500 steps; model =
four VTI layers with a smooth +-3% random perturbation (seed 1). A bench "step" is one
complete shot (reset + 3000 time steps) with the model already resident in HBM.

metric: Gcell-updates/s = updated cells (nx-10)(ny-10)(nz-10) x time steps / wall time.
roofline: the dominant launch, the interior update kernel: its algorithmic bytes per launch
(24 B per interior cell: u_cur, u_prev, three model fields, u_next) over its mean launch time,
from HIP events on the plan's stream.
cpu_baseline: the CPU oracle (the reference's arithmetic, bit-exact with its CPU kernels)
on the host's CPU share, on a 20-step window that starts from the GPU's state at step 300.

    python bench.py [--steps K] [--warmup W] [--cpu-steps S] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

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


def cpu_baseline(setup, plan, start, nsteps):
    """Time the oracle for nsteps starting from the GPU's wavefield after `start` steps."""
    import numpy as np
    from oracle import oracle as O

    plan.reset()
    plan.run(0, start)
    plan.sync()
    fields = [plan.field(w) for w in range(9)]
    threads = O.default_threads()
    o = O.OracleState(setup, "f32", nthreads=threads)
    o.import_state(*fields)
    del fields
    t0 = time.perf_counter()
    for k in range(start, start + nsteps):
        o.step(setup.wavelet[k])
    dt = time.perf_counter() - t0
    o.close()
    return {"value": round(setup.n_upd * nsteps / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"CPU oracle (FP32, reference arithmetic) on the 576^3 grid, {nsteps} time steps "
                      f"{start}..{start + nsteps - 1} from the GPU state at step {start}; {dt:.1f} s"}


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
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--tile-update", type=int, default=0)
    ap.add_argument("--tile-grad", type=int, default=0)
    ap.add_argument("--tile-pml", type=int, default=0)
    ap.add_argument("--streams", type=int, default=0, help="2: interior and PML updates concurrent; 1: serial")
    ap.add_argument("--storage", default="fp32", choices=["fp32", "bf16", "fp16"],
                    help="storage of wavefields and PML state (arithmetic is always FP32)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a 1 GiB device-to-device copy (the reference kernel for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the shot-parallel and z-slab "
              f"designs for several GPUs are described, not built)", file=sys.stderr)
        sys.exit(2)

    import qp3d

    if args.calib_copy:
        calib_copy(1 << 30)
    t = time.perf_counter()
    s = qp3d.config_setup(hx=args.n, hy=args.n, nz0=args.n, pml=args.pml, nt=args.nt, seed=1, fm=25.0, src_depth=0)
    log(f"setup {s.nx}x{s.ny}x{s.nz} in {time.perf_counter() - t:.1f} s")
    ntsnap = 500
    plan = qp3d.Plan(s, snapflag=1, ntsnap=ntsnap, use_graph=bool(args.graph), tile_update=args.tile_update,
                     tile_grad=args.tile_grad, tile_pml=args.tile_pml, streams=args.streams, storage=args.storage)
    log("model uploaded")
    for i in range(args.warmup):
        t = time.perf_counter()
        plan.forward_async()
        plan.sync()
        log(f"warmup shot {i}: {time.perf_counter() - t:.3f} s")
    # Kernel times come from HIP events around every launch of the timed shots. Replayed
    # graphs carry no events, so with --graph the events come from one extra shot of direct
    # launches after the timed region.
    plan.set_instrument(not args.graph)
    plan.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.forward_async()
        plan.sync()  # also harvests this shot's kernel events
    t1 = time.perf_counter()
    wall = t1 - t0
    if args.graph:
        plan.set_instrument(True)
        plan.forward_async()
        plan.sync()
    kt = plan.kernel_times()
    kb = plan.step_bytes()
    plan.set_instrument(False)
    log(f"timed {args.steps} shots in {wall:.3f} s; per step: {kt}")
    gbs = {k: kb[k] / (kt[k] * 1e-3) / 1e9 for k in ("gradmax", "update_interior", "update_pml") if kt[k] > 0}
    log(f"bytes per step {kb}; achieved GB/s {gbs}")

    cells = s.n_upd * s.nt * args.steps
    value = cells / wall / 1e9
    achieved = gbs["update_interior"]
    step_bytes = kb["gradmax"] + kb["update_interior"] + kb["update_pml"]
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            with open(TRAFFIC_FILE) as f:
                tr = json.load(f)
            if tr.get("nx") == s.nx and tr.get("storage", "fp32") == args.storage:
                traffic = tr.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        cpu = cpu_baseline(s, plan, min(args.cpu_start, s.nt - args.cpu_steps), args.cpu_steps)
        log(f"cpu baseline {cpu}")
    plan.close()

    out = {
        "metric": "Gcell-updates/s (512^3 FP32 forward, 3000 steps)" if args.storage == "fp32" else
                  f"Gcell-updates/s (512^3 forward, {args.storage.upper()} storage / FP32 arithmetic, 3000 steps)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (layered VTI + smooth random perturbation, seed 1)",
        "config": {"workload": "3D pure-qP VTI forward shot, 512^3 interior, pml 32 (576^3 padded), "
                               "3000 steps, dt 1 ms, dx 12.5 m, 25 Hz Ricker, 512x512 receivers, "
                               "snapshots every 500 steps",
                   "padded": [s.nx, s.ny, s.nz], "nt": s.nt, "n_upd": s.n_upd, "n_int": s.n_int,
                   "gcells_int_per_s": round(s.n_int * s.nt * args.steps / wall / 1e9, 3),
                   "ms_per_time_step": round(wall / args.steps / s.nt * 1e3, 5),
                   "ms_gradmax": round(kt["gradmax"], 5), "ms_update_interior": round(kt["update_interior"], 5),
                   "ms_update_pml": round(kt["update_pml"], 5), "ms_update_stage": round(kt["update_stage"], 5),
                   "gbs_gradmax": round(gbs.get("gradmax", 0), 1), "gbs_update_pml": round(gbs.get("update_pml", 0), 1),
                   "storage": args.storage,
                   "roofline_kernel": f"k_update<interior> ({kb['update_interior'] / kb['cells_interior']:.0f} B per interior cell)",
                   "hbm_frac_step": round(step_bytes * s.nt * args.steps / wall / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
