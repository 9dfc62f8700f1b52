"""Benchmark: Mpts*RKstage/s of the fused pseudo-spectral RK4 step, 512^3 incompressible NS.

Workload (the 512^3 decaying-turbulence configuration): FOUR basis, 2/3 dealiasing, RK4 with the
FP64 integrating factor, nu = 1e-3, seeded random-phase divergence-free IC with
E(k) ~ k^4 exp(-2 (k/4)^2), total energy 0.5, dt fixed once from CFL 0.5 at t = 0.
value = FP32-mixed throughput (grid points x RK stages per second, inputs resident in HBM);
the FP64 GPU path is timed on the same IC and reported beside it, with the E(k) error of the
FP32-mixed run against the FP64 run after the same number of steps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu] [--no-fp64]
Prints one JSON line.  Runs on one GPU; --gpus > 1 exits with a message (single-GPU build).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PASS_NAMES = ["x_inverse", "y_inverse", "z_fused_c2r_products_r2c", "y_forward", "x_forward_epilogue"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cfl_dt(plan64, V, cfl=0.5):
    """dt = cfl * dx / max|u| with |u| the Euclidean speed in real space (FP64, computed once)."""
    import torch
    speed2 = None
    for a in V:
        b = a.clone()
        plan64.inverse(b)
        N3 = plan64.N[2]
        r = torch.view_as_real(b).reshape(plan64.N[0], plan64.N[1], -1)[:, :, :N3]
        speed2 = r * r if speed2 is None else speed2 + r * r
        del b
    umax = float(torch.sqrt(speed2.max()))
    del speed2
    dx = 2 * math.pi / plan64.N[0]
    return cfl * dx / umax, umax


def time_steps(state, steps, dt, nu):
    import torch
    torch.cuda.synchronize()
    state.plan._post(0, "")
    t0 = time.perf_counter()
    state.step(dt, nu=nu, steps=steps)          # enqueues hipGraph replays, then syncs the stream
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def spectrum_err(plan64, Va, Vb):
    import numpy as np
    Ea = plan64.vector_spectrum(Va)
    Eb = plan64.vector_spectrum(Vb)
    sel = Eb >= 1e-10 * Eb.max()
    maxrel = float((np.abs(Ea - Eb)[sel] / Eb[sel]).max())
    l2 = float(np.linalg.norm(Ea - Eb) / np.linalg.norm(Eb))
    return maxrel, l2


def cpu_baseline(V_host, N, nu, dt, threads):
    """The FP64 oracle (the reference algorithm restated in C, OpenMP over slabs) timed on one RK
    stage of the same 512^3 workload on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    o = Oracle(N, threads=threads)
    V = [v.copy() for v in V_host]
    t0 = time.perf_counter()
    o.rk4_stage1_ns(V, nu, dt)
    el = time.perf_counter() - t0
    return el, o.L.orc_get_threads()


def load_traffic(workload):
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs one box per GPU on a single MI355X; multi-GPU runs "
              "(--gpus > 1) are not part of it", flush=True)
        sys.exit(2)

    import numpy as np
    import torch
    from tarang_amd import Context, Plan, State

    N = (args.n, args.n, args.n)
    nu = 1e-3
    if args.calib_copy_mib > 0:
        src = torch.ones(args.calib_copy_mib * (1 << 18), dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {args.calib_copy_mib} MiB done")
        del src, dst
    ctx = Context(0)
    p64 = Plan(ctx, N, precision="fp64")
    p32 = Plan(ctx, N, precision="fp32")
    log(f"bench: IC {N} seed 1")
    V = p64.ic_random(1, 0.5, 4.0)
    dt, umax = cfl_dt(p64, V)
    E0 = p64.total_energy(V)
    log(f"bench: E0={E0:.6f} umax={umax:.4f} dt={dt:.6e}")

    s32 = State(p32)
    s32.upload([v.to(torch.complex64) for v in V])
    # FP32-mixed timing
    s32.step(dt, nu=nu, steps=args.warmup)
    torch.cuda.synchronize()
    t32 = time_steps(s32, args.steps, dt, nu)
    log(f"bench: fp32-mixed {args.steps} steps {t32:.4f} s")
    # per-pass timing with HIP events on the library stream (stage 2 of the pipeline)
    ms, nbytes = s32.profile_passes(dt, nu=nu, stage=2, nrep=5)
    out32 = s32.download()

    pts = float(N[0]) * N[1] * N[2]
    value = pts * 4 * args.steps / t32 / 1e6
    result = {"fp64": None}
    if not args.no_fp64:
        s64 = State(p64)
        s64.upload(V)
        s64.step(dt, nu=nu, steps=args.warmup)
        torch.cuda.synchronize()
        t64 = time_steps(s64, args.steps, dt, nu)
        log(f"bench: fp64 {args.steps} steps {t64:.4f} s")
        out64 = s64.download()
        v64 = pts * 4 * args.steps / t64 / 1e6
        maxrel, l2 = spectrum_err(p64, [a.to(torch.complex128) for a in out32], out64)
        result["fp64"] = {"value": round(v64, 1), "ms_per_step": round(1e3 * t64 / args.steps, 3),
                          "speedup_fp32_mixed": round(value / v64, 3),
                          "Ek_rel_err_fp32_vs_fp64_max": maxrel, "Ek_rel_err_fp32_vs_fp64_l2": l2,
                          "steps_compared": args.warmup + args.steps}
        del s64, out64

    # roofline of the dominant kernel (the longest pass)
    k = int(np.argmax(ms))
    ach = nbytes[k] / (ms[k] * 1e-3) / 1e9
    workload = f"ns{args.n}_fp32mixed"
    traffic = load_traffic(workload)
    roof = {"bound": "hbm", "kernel": PASS_NAMES[k], "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": (traffic or {}).get(PASS_NAMES[k])}
    stage_bytes = float(sum(nbytes))
    stage_ms = float(sum(ms))
    passes = {PASS_NAMES[i]: {"ms": round(float(ms[i]), 4), "GB": round(nbytes[i] / 1e9, 4),
                              "GB_s": round(nbytes[i] / (ms[i] * 1e-3) / 1e9, 1)} for i in range(5)}

    cpu = None
    if not args.no_cpu:
        Vh = [a.cpu().numpy() for a in V]
        threads = args.cpu_threads if args.cpu_threads > 0 else min(os.cpu_count() or 1, 16)
        log(f"bench: CPU oracle, one RK stage at {N}, {threads} threads")
        el, used = cpu_baseline(Vh, N, nu, dt, threads)
        cpu = {"value": round(pts / el / 1e6, 3), "unit": "Mpts*RKstage/s", "cores": int(used),
               "kind": "port", "seconds": round(el, 2),
               "sample": f"1 RK4 stage (rhs: 9 whole-array transforms + projection + IF update) of the "
                         f"same {N[0]}^3 FP64 workload by the oracle (reference algorithm restated in C, "
                         f"own FFT, OpenMP)"}

    line = {
        "metric": "Mpts*RKstage/s at 512^3 NS: FP32-mixed vs FP64 GPU; % HBM roofline; E(k) err",
        "value": round(value, 1),
        "unit": "Mpts*RKstage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t32 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp32 storage/FFT; fp64 integrating factor, twiddles, sums)",
        "data": "synthetic (seeded random-phase IC generated on the GPU)",
        "config": {"workload": f"ns_{args.n}^3_rk4_decaying_turbulence", "N": args.n, "nu": nu,
                   "dt": dt, "cfl": 0.5, "ic": "random-phase k^4 exp(-2(k/4)^2), E=0.5, seed 1",
                   "dealias": "2/3", "precision": "fp32-mixed"},
        "fp64_gpu": result["fp64"],
        "roofline": roof,
        "stage": {"ms_per_stage_sum_of_passes": round(stage_ms, 4), "algorithmic_GB_per_stage":
                  round(stage_bytes / 1e9, 3), "GB_s": round(stage_bytes / (stage_ms * 1e-3) / 1e9, 1),
                  "passes": passes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
