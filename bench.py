"""Benchmark of the se_e2_a energy/force/virial path on one MI355X.

Workload (the 12,288-atom water configuration): 4096 H2O in a 49.68 A cubic
box, se_e2_a rcut 6.0 / rcut_smth 0.5, sel [46, 92], embedding [25, 50, 100]
compressed at strides 0.01/0.1, axis 16, fitting [240, 240, 240] tanh, seeded
random weights, zero biases, davg 0 / dstd 1; synthetic seeded coordinates.
Launch shapes and kernel variants are autotuned per engine before timing.
Because that model's energies are ~1e-4 eV/atom, the mixed-vs-FP64 errors
are also reported on the same architecture with davg/dstd taken from the
environment statistics of a water frame (energies of order 1e-2 eV/atom).  One step = one full E+F+V
evaluation from raw coordinates (wrap, ghosts, cell-list neighbour list,
environment matrix, table embedding + contraction, fitting forward and
backward, force/virial, ghost fold, per-frame sums) with the coordinates
already resident in HBM.

Prints one JSON line (rank 0).  Also reports per-stage device times, the
roofline of the dominant kernel (live HIP-event timing over the timed
region) and the reference CPU path timed on a bounded sample of the same
workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "deepmd-kit-lr-pt_amd"))

METRIC = "atom-steps/s (E+F+V), 12k-atom water, mixed vs FP64 GPU; force RMSE vs FP64"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3


def stage_model(m, nloc, nall, precision):
    """Algorithmic work per launch of each engine stage: per local atom
    figures times nloc.  N = sum(sel), M = embedding width, A = axis,
    D = M*A, s = bytes per stored float (8 FP64, 4 mixed), h = bytes per
    descriptor element (8 FP64, 2 mixed), rho = nall / nloc.  Entries are
    (bound, amount, peak): "hbm" stages are bytes against HBM bandwidth,
    "flop" stages are flops against the arithmetic peak of the path (table
    stages: FP32 / FP64 vector; fitting: FP16 / FP64 matrix).  The table
    stages also carry their algorithmic bytes."""
    N, M, A = m.nnei, m.M, m.axis_neuron
    D = M * A
    mixed = precision == "mixed"
    s = 4 if mixed else 8
    h = 2 if mixed else 8
    rho = nall / nloc
    fit = [D] + list(m.fit_neuron) + [1]
    fit_fwd = 2 * sum(fit[i] * fit[i + 1] for i in range(len(fit) - 1))
    fit_bwd = fit_fwd - 2 * fit[-2]  # no product for the output layer
    mm_peak = FP16_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    vec_peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    per_atom = {
        # reads nlist (4N) and extended coordinates, writes em (4 N s); the
        # slot derivative and displacement are rebuilt by the force stage
        "env_mat": ("hbm", 4 * N + 4 * N * s + 24 * rho, None, None),
        # quintic + contraction (18 N M), descriptor (8 M A); reads em, writes xyz and D
        "embed_fwd": ("flop", 18 * N * M + 8 * M * A, vec_peak, 4 * N * s + 4 * M * s + D * h),
        # dT from dE/dD (16 M A), table adjoint (36 N M); reads em, xyz, dE/dD, writes net_deriv
        "embed_bwd": ("flop", 36 * N * M + 16 * M * A, vec_peak, 8 * N * s + 4 * M * s + D * s),
        # reads net_deriv and nlist, the extended coordinates once; scatters
        # forces (FP64) and writes the per-row virial
        "force_virial": ("hbm", 4 * N * s + 4 * N + 24 * rho + 72, None, None),
        # reads the cell lists and ghost coordinates, writes the sorted rows
        "nlist": ("hbm", 8 * N, None, None),
        # the same, storing the environment matrix of the rows as well
        "nlist_env": ("hbm", 8 * N + 4 * N * s + 24 * rho, None, None),
        "fitting": ("flop", fit_fwd + fit_bwd, mm_peak, None),
        "fitting_fwd": ("flop", fit_fwd, mm_peak, None),
        "fitting_bwd": ("flop", fit_bwd, mm_peak, None),
    }
    return {k: (b, amt * nloc, pk, None if by is None else by * nloc) for k, (b, amt, pk, by) in per_atom.items()}


# HBM bytes per launch (read MiB, write MiB) measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE on this workload, scaled by a 256 MiB copy; the
# readings and the commands are in TRAFFIC_PROFILE.
MEASURED_TRAFFIC_MIB = {
    "mixed": {"embed_bwd": (123.05, 26.19), "embed_fwd": (31.32, 54.89), "fitting": (61.83, 93.03),
              "force_virial": (40.09, 19.78), "nlist_env": (13.4, 33.0)},
    "fp64": {"embed_bwd": (261.61, 50.70), "embed_fwd": (85.80, 202.84), "force_virial": (60.19, 11.39),
             "nlist_env": (14.9, 58.9)},
}


def host_cpu():
    """lscpu model name, sockets and physical cores of the host."""
    import subprocess

    info = {"model": None, "sockets": None, "physical_cores": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except Exception:  # noqa: BLE001
        return info
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    info["model"] = kv.get("Model name")
    try:
        sockets = int(kv.get("Socket(s)", "0"))
        cps = int(kv.get("Core(s) per socket", "0"))
        info["sockets"] = sockets
        info["physical_cores"] = sockets * cps
    except ValueError:
        pass
    return info


TRAFFIC_PROFILE = "profiles/water12k_kernels_and_traffic.txt"


def cpu_baseline(m, coord, atype, box, evals=10, warmup=3, what="the same 12,288-atom water frame"):
    """Reference CPU path, as its benchmark plan states it: the compiled
    reference lib (the oracle restatement where the reference is absent) for
    normalize/copy_coord, build_nlist, prod_env_mat_a, tabulate and
    prod_force/prod_virial, the fitting GEMMs in libtorch FP64 on the CPU;
    OpenMP threads bound to cores (OMP_PROC_BIND=close, OMP_PLACES=cores,
    set before the library's OpenMP runtime starts), `warmup` untimed
    evaluations, then the median of `evals` timed ones with the per-stage
    breakdown."""
    import statistics

    import torch

    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", 0) or cpu["physical_cores"] or os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    torch.set_num_threads(threads)
    from oracle import oracle_lib, ref_available, ref_lib
    from oracle import sea
    from oracle.sea import evaluate

    sea.USE_TORCH = True
    lib = ref_lib() if ref_available() else oracle_lib()
    kind = "reference" if ref_available() else "port"
    r = None
    for _ in range(warmup):
        r = evaluate(m, coord, atype, box, lib)  # also the CPU result
    times, stages = [], []
    for _ in range(evals):
        st = {}
        t0 = time.perf_counter()
        r = evaluate(m, coord, atype, box, lib, stage_seconds=st)
        times.append(time.perf_counter() - t0)
        stages.append(st)
    med = statistics.median(times)
    k_med = times.index(sorted(times)[len(times) // 2])
    return r, {"value": len(atype) / med, "unit": "atom-steps/s", "cores": threads, "kind": kind,
               "sample": f"median of {evals} E+F+V evaluations (after {warmup} warm-up) of {what}: reference "
                         f"CPU lib + libtorch FP64 fitting, OMP_NUM_THREADS={threads}, OMP_PROC_BIND=close, "
                         f"OMP_PLACES=cores",
               "median_s": med, "min_s": min(times), "max_s": max(times),
               "stages_s_of_median_eval": {k: round(v, 5) for k, v in stages[k_med].items()},
               "host_cpu": cpu}


def time_engine(eng, x, box, steps, warmup):
    """Warm-up, then `steps` timed evaluations bracketed by device syncs;
    returns (seconds, per-stage device ms averaged over the timed calls)."""
    import torch

    for _ in range(warmup):
        eng.compute(x, box)  # capacity growth, if any, happens here
    eng.status()
    # the capacities now fit this configuration: the timed steps run without
    # the per-step status read-back (status() is checked after them)
    eng.set_growth(False)
    eng.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.compute(x, box)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    eng.status()
    stages = eng.timing()
    eng.set_timing(False)
    eng.set_growth(True)
    return dt, stages


def time_graph(eng, x, box, steps, warmup):
    """The step captured into a hipGraph on a side stream, `steps` timed
    replays bracketed by device syncs; returns seconds."""
    import torch

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eng.capture(x, box)
        for _ in range(warmup):
            eng.replay()
    torch.cuda.synchronize()
    eng.status()
    t0 = time.perf_counter()
    with torch.cuda.stream(s):
        for _ in range(steps):
            eng.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    eng.status()
    return dt


def roofline(m, nloc, nall, precision, stages, embed_mask=0):
    model = stage_model(m, nloc, nall, precision)
    dom = max((k for k in stages if k in model), key=lambda k: stages[k])
    kind, amount, peak, nbytes = model[dom]
    t_ms = stages[dom]
    if kind == "hbm":
        roof = {"kernel": dom, "bound": "hbm", "achieved": amount / (t_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s"}
    else:
        # the fitting runs on the matrix cores; the table stages are counted
        # as the reference's per-neighbour FP32 (FP64) work against the vector
        # peak -- in the mixed engine the row-run kernels move the column sums
        # onto FP16 matrix-core products (once per table-row run), so the
        # figure is an FP32-equivalent rate
        bound = "mfma" if dom.startswith("fitting") else "valu"
        roof = {"kernel": dom, "bound": bound, "achieved": amount / (t_ms * 1e-3) / 1e12, "peak": peak,
                "unit": "TFLOP/s"}
        runs = (dom == "embed_fwd" and embed_mask & 8) or (dom == "embed_bwd" and embed_mask & 4)
        if precision == "mixed" and runs:
            roof["note"] = ("achieved = reference per-neighbour FP32 flops / kernel time; the row-run kernel "
                            "forms the column sums once per table-row run on v_mfma_f32_16x16x16_f16 split "
                            "products")
    roof["frac"] = roof["achieved"] / roof["peak"]
    tr = MEASURED_TRAFFIC_MIB.get(precision, {}).get(dom)
    roof["traffic"] = None if tr is None else (tr[0] + tr[1]) * 1048576.0
    if tr is not None:
        roof["traffic_source"] = TRAFFIC_PROFILE
    roof["avg_ms"] = t_ms
    if nbytes is not None:
        roof["algorithmic_hbm_GBps"] = nbytes / (t_ms * 1e-3) / 1e9
    if roof["traffic"] is not None:
        roof["measured_hbm_GBps"] = roof["traffic"] / (t_ms * 1e-3) / 1e9
    return roof


def data_statistics_errors(nmol: int = 512):
    """Mixed-vs-FP64 errors of the same architecture with davg/dstd from the
    environment statistics of a water frame (as training sets them): the
    benchmark model's davg 0 / dstd 1 gives energies of ~1e-4 eV/atom, this
    one of order 1e-2 eV/atom with forces of order 0.4 eV/A, so the errors
    are measured at the scale of a trained model."""
    import numpy as np
    import torch

    from dpmi import model as dmodel
    from dpmi import systems
    from dpmi.engine import DeviceModel, Engine

    m = dmodel.water_model(seed=1, env_stat=True)
    dm = DeviceModel(m)
    coord, atype, box = systems.water(nmol, seed=7)
    n = len(atype)
    res = {}
    for prec in ("fp64", "mixed"):
        eng = Engine(dm, [n], atype, precision=prec)
        e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box)
        eng.status()
        res[prec] = [t.cpu().numpy().copy() for t in (e, f, v)]
        del eng
    (e_m, f_m, v_m), (e_d, f_d, v_d) = res["mixed"], res["fp64"]
    return {"workload": f"water {nmol} H2O ({n} atoms), davg/dstd from a water frame",
            "energy_eV_atom": float(e_d[0] / n), "force_rms_eV_A": float(np.sqrt(np.mean(f_d ** 2))),
            "energy_meV_atom": float(abs(e_m[0] - e_d[0]) / n * 1e3),
            "force_rmse_meV_A": float(np.sqrt(np.mean((f_m - f_d) ** 2)) * 1e3),
            "force_max_meV_A": float(np.max(np.abs(f_m - f_d)) * 1e3),
            "virial_meV_atom": float(np.max(np.abs(v_m[0] - v_d[0])) / n * 1e3)}


def batched_leg(frames: int, steps: int, warmup: int, eager: bool = True) -> dict:
    """Times the batched workload (see run_batched); returns its JSON object."""
    class A:
        pass

    a = A()
    a.frames, a.steps, a.warmup, a.eager = frames, steps, warmup, eager
    return run_batched(a, emit=False)


def run_batched(args, emit: bool = True):
    """256 water frames with atom counts uniform over {96, 99, ..., 3000}, each
    in its own 1.0 g/cm^3 box, in one mixed-precision engine: autotuned launch
    shapes, one hipGraph launch per step, per-frame errors against the FP64
    engine on the same frames."""
    import numpy as np
    import torch

    from dpmi import model as dmodel
    from dpmi import systems
    from dpmi.engine import DeviceModel, Engine

    m = dmodel.water_model(seed=1)
    frames = systems.uneven_water_frames(args.frames, seed=2024, lo=96, hi=3000)
    natoms = [len(a) for (_, a, _) in frames]
    coord = np.concatenate([c for (c, _, _) in frames])
    atype = np.concatenate([a for (_, a, _) in frames])
    box = np.stack([b for (_, _, b) in frames])
    ntot = int(sum(natoms))
    dm = DeviceModel(m)
    x = torch.as_tensor(coord, device="cuda")
    eng = Engine(dm, natoms, atype, precision="mixed")
    knobs = eng.autotune(x, box, reps=8)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eng.capture(x, box)
        for _ in range(args.warmup):
            eng.replay()
    s.synchronize()
    eng.status()
    t0 = time.perf_counter()
    with torch.cuda.stream(s):
        for _ in range(args.steps):
            eng.replay()
    s.synchronize()
    dt = time.perf_counter() - t0
    eng.status()
    e_m, f_m, v_m = [t.cpu().numpy().copy() for t in (eng.energy, eng.force, eng.virial)]
    dt_eager = None
    if getattr(args, "eager", True):  # eager (no graph) beside it
        eng.set_growth(False)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            eng.compute(x, box)
        torch.cuda.synchronize()
        dt_eager = time.perf_counter() - t1
        eng.status()
    del eng
    ref = Engine(dm, natoms, atype, precision="fp64")
    e_d, f_d, v_d = [t.cpu().numpy().copy() for t in ref.compute(x, box)]
    ref.status()
    errs = []
    off = 0
    for k, n in enumerate(natoms):
        errs.append((abs(e_m[k] - e_d[k]) / n * 1e3,
                     float(np.sqrt(np.mean((f_m[off:off + n] - f_d[off:off + n]) ** 2)) * 1e3),
                     float(np.max(np.abs(v_m[k] - v_d[k]))) / n * 1e3))
        off += n
    errs = np.array(errs)
    out = {
        "metric": "atom-steps/s (E+F+V), 256 uneven water frames (96-3000 atoms), mixed, one hipGraph launch per step",
        "value": ntot * args.steps / dt,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-mfma/fp32 (fp64 geometry, s(r), sums)",
        "data": "synthetic",
        "config": {"workload": f"{args.frames} water frames, {ntot} atoms, 1.0 g/cm3 boxes", "frames": args.frames,
                   "atoms": ntot, "parallelism": "single-gpu"},
        "eager_value": None if dt_eager is None else ntot * args.steps / dt_eager,
        "tuned_knobs": knobs,
        "per_frame_error_vs_fp64": {
            "energy_meV_atom_max": float(errs[:, 0].max()), "energy_meV_atom_mean": float(errs[:, 0].mean()),
            "force_rmse_meV_A_max": float(errs[:, 1].max()), "force_rmse_meV_A_mean": float(errs[:, 1].mean()),
            "virial_meV_atom_max": float(errs[:, 2].max())},
    }
    if not emit:
        return out
    print(json.dumps(out), flush=True)
    return 0


def run_copper(args) -> int:
    """1,048,576-atom fcc Cu (64^3 cells, a = 3.615 A, 0.05 A Gaussian
    displacements), one frame: se_e2_a rcut 8.0 / rcut_smth 2.0, sel [200],
    embedding [32, 64, 128] compressed, axis 16, fitting [240, 240, 240];
    mixed engine timed over `steps` evaluations, the FP64 engine over a few
    beside it for the error and the ratio.  Per-stage device time, algorithmic
    HBM GB/s of every stage and the fitting's MFMA utilisation."""
    import numpy as np
    import torch

    from dpmi import model as dmodel
    from dpmi import systems
    from dpmi.engine import DeviceModel, Engine

    m = dmodel.copper_model(seed=1)
    coord, atype, box = systems.fcc_copper(args.ncell, seed=7)
    nloc = len(atype)
    dm = DeviceModel(m)
    x = torch.as_tensor(coord, device="cuda")
    res, knobs = {}, {}
    for prec, steps in (("mixed", args.steps), ("fp64", max(2, args.steps // 4))):
        print(f"copper {nloc} atoms: {prec} engine, {steps} steps", file=sys.stderr, flush=True)
        eng = Engine(dm, [nloc], atype, precision=prec)
        knobs[prec] = eng.autotune(x, box, reps=3)
        print(f"  tuned {knobs[prec]}", file=sys.stderr, flush=True)
        dt, stages = time_engine(eng, x, box, steps, min(args.warmup, 3))
        e, f, v = eng.compute(x, box)
        torch.cuda.synchronize()
        nall = int(eng.fetch("nall")[0])
        free, total = torch.cuda.mem_get_info()
        res[prec] = {"dt": dt / steps, "stages": stages, "nall": nall, "e": e.cpu().numpy().copy(),
                     "f": f.cpu().numpy().copy(), "v": v.cpu().numpy().copy(), "hbm_gib": (total - free) / 2**30}
        del eng, e, f, v
        torch.cuda.empty_cache()
    rm, rd = res["mixed"], res["fp64"]
    model = stage_model(m, nloc, rm["nall"], "mixed")
    per_stage = {}
    for k, t_ms in rm["stages"].items():
        row = {"ms": round(t_ms, 4)}
        if k in model:
            kind, amount, peak, nbytes = model[k]
            if kind == "hbm":
                row["hbm_GBps"] = round(amount / (t_ms * 1e-3) / 1e9, 1)
            else:
                row["TFLOPs"] = round(amount / (t_ms * 1e-3) / 1e12, 2)
                row["frac_of_peak"] = round(amount / (t_ms * 1e-3) / 1e12 / peak, 4)
                if nbytes is not None:
                    row["hbm_GBps"] = round(nbytes / (t_ms * 1e-3) / 1e9, 1)
        per_stage[k] = row
    err = {
        "energy_meV_atom": float(abs(rm["e"][0] - rd["e"][0]) / nloc * 1e3),
        "force_rmse_meV_A": float(np.sqrt(np.mean((rm["f"] - rd["f"]) ** 2)) * 1e3),
        "force_max_meV_A": float(np.max(np.abs(rm["f"] - rd["f"])) * 1e3),
        "virial_meV_atom": float(np.max(np.abs(rm["v"][0] - rd["v"][0])) / nloc * 1e3),
    }
    out = {
        "metric": "atom-steps/s (E+F+V), 1M-atom fcc Cu, mixed engine",
        "value": nloc / rm["dt"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": min(args.warmup, 3),
        "ms_per_step": rm["dt"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-mfma/fp32 (fp64 geometry, s(r), sums)",
        "data": "synthetic",
        "config": {"workload": f"fcc Cu {args.ncell}^3 cells ({nloc} atoms, box {box[0]:.2f} A), rcut 8.0, "
                               f"sel [200], embedding [32, 64, 128] compressed, one frame, E+F+V per step",
                   "nloc": nloc, "nall": rm["nall"], "parallelism": "single-gpu"},
        "roofline": roofline(m, nloc, rm["nall"], "mixed", rm["stages"], knobs["mixed"]["embed_matrix_cores"]),
        "stages": per_stage,
        "fp64_value": nloc / rd["dt"],
        "mixed_over_fp64": rd["dt"] / rm["dt"],
        "fp64_stages_ms": {k: round(vv, 4) for k, vv in rd["stages"].items()},
        "mixed_error_vs_fp64": err,
        "hbm_in_use_GiB": {"mixed": round(rm["hbm_gib"], 2), "fp64": round(rd["hbm_gib"], 2)},
        "tuned_knobs": knobs,
    }
    out["roofline"]["traffic"] = None
    out["roofline"].pop("measured_hbm_GBps", None)
    if not args.no_cpu_baseline:
        cc, ca, cbx = systems.fcc_copper(6, seed=7)
        _, cb = cpu_baseline(m, cc, ca, cbx, what=f"a {len(ca)}-atom fcc Cu frame (6^3 cells)")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"],
                    help="path reported as value (the other one is timed beside it)")
    ap.add_argument("--nmol", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-frame batched leg of the default run")
    ap.add_argument("--eager", action="store_true", help="time the mixed engine with eager launches only")
    ap.add_argument("--workload", default="water12k", choices=["water12k", "batched", "copper1m"],
                    help="water12k: the 12,288-atom water frame (default); batched: uneven water frames; "
                         "copper1m: 1M-atom fcc copper")
    ap.add_argument("--ncell", type=int, default=64, help="fcc cells per edge of the copper workload")
    ap.add_argument("--frames", type=int, default=256, help="frames of the batched workload")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run 5 plain device-to-device copies of this many MiB (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "copper1m":
        return run_copper(args)
    if args.copy_mib > 0:
        import torch

        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.copy_mib, "copies": 5}), flush=True)
        return 0

    import numpy as np
    import torch

    from dpmi import model as dmodel
    from dpmi import systems
    from dpmi.engine import DeviceModel, Engine

    m = dmodel.water_model(seed=1)
    # the 12,288-atom configuration: 4096 molecules in a 49.68 A cubic box
    coord, atype, box = systems.water(args.nmol, seed=2024, box=49.68 if args.nmol == 4096 else None)
    nloc = len(atype)
    dm = DeviceModel(m)
    x = torch.as_tensor(coord, device="cuda")
    other = "fp64" if args.precision == "mixed" else "mixed"
    res = {}
    knobs = {}
    for prec in (args.precision, other):
        eng = Engine(dm, [nloc], atype, precision=prec)
        knobs[prec] = eng.autotune(x, box, reps=8)
        dt_eager, stages = time_engine(eng, x, box, args.steps, args.warmup)
        dt = dt_eager
        if not args.eager:
            # the same step captured once into a hipGraph and replayed (the
            # coordinates are re-read from x on every replay); both engines
            # are timed this way
            dt = time_graph(eng, x, box, args.steps, args.warmup)
        e, f, v = eng.compute(x, box)
        torch.cuda.synchronize()
        nall = int(eng.fetch("nall")[0])
        res[prec] = {"dt": dt, "dt_eager": dt_eager, "stages": stages, "nall": nall, "e": e.cpu().numpy().copy(),
                     "f": f.cpu().numpy().copy(), "v": v.cpu().numpy().copy()}
        del eng
    r = res[args.precision]
    ms = r["dt"] / args.steps * 1e3
    value = nloc * args.steps / r["dt"]
    fp64_value = nloc * args.steps / res["fp64"]["dt"]
    mixed_value = nloc * args.steps / res["mixed"]["dt"]
    rm, rd = res["mixed"], res["fp64"]
    mixed_err = {
        "energy_meV_atom": float(abs(rm["e"][0] - rd["e"][0]) / nloc * 1e3),
        "force_rmse_meV_A": float(np.sqrt(np.mean((rm["f"] - rd["f"]) ** 2)) * 1e3),
        "force_max_meV_A": float(np.max(np.abs(rm["f"] - rd["f"])) * 1e3),
        "virial_meV_atom": float(np.max(np.abs(rm["v"][0] - rd["v"][0])) / nloc * 1e3),
    }

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-mfma/fp32 (fp64 geometry, s(r), sums)" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"water {args.nmol} H2O ({nloc} atoms, box {box[0]:.2f} A), se_e2_a compressed, "
                               f"{args.precision} engine, one frame, E+F+V per step", "nloc": nloc,
                   "nall": r["nall"], "precision": args.precision, "parallelism": "single-gpu"},
        "roofline": roofline(m, nloc, r["nall"], args.precision, r["stages"],
                             knobs[args.precision]["embed_matrix_cores"]),
        "stages_ms": {k: round(vv, 5) for k, vv in r["stages"].items()},
        "fp64_value": fp64_value,
        "mixed_value": mixed_value,
        "launch": "eager" if args.eager else "hipGraph replay, both engines (stages_ms from eager runs)",
        "mixed_eager_value": nloc * args.steps / res["mixed"]["dt_eager"],
        "fp64_eager_value": nloc * args.steps / res["fp64"]["dt_eager"],
        "mixed_over_fp64": mixed_value / fp64_value,
        "mixed_error_vs_fp64": mixed_err,
        "other_stages_ms": {k: round(vv, 5) for k, vv in res[other]["stages"].items()},
        "tuned_knobs": knobs,
    }
    out["mixed_error_vs_fp64_data_statistics"] = data_statistics_errors()
    if not args.no_batched:
        b = batched_leg(256, steps=20, warmup=3, eager=False)
        out["batched_256_frames"] = {k: b[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "config",
                                                       "tuned_knobs", "per_frame_error_vs_fp64")}
    if not args.no_cpu_baseline:
        ref, cb = cpu_baseline(m, coord, atype, box)
        out["cpu_baseline"] = cb
        out["fp64_accuracy_vs_cpu_reference"] = {
            "energy_rel": abs(rd["e"][0] - ref.energy) / abs(ref.energy),
            "force_max_rel": float(np.max(np.abs(rd["f"] - ref.force)) / np.max(np.abs(ref.force))),
            "virial_max_rel": float(np.max(np.abs(rd["v"][0] - ref.virial)) / np.max(np.abs(ref.virial))),
        }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
