"""Benchmark: Mpts*RKstage/s of the fused pseudo-spectral RK4 step, 512^3 incompressible NS.

Workload (the 512^3 decaying-turbulence configuration): FOUR basis, 2/3 dealiasing, RK4 with the
FP64 integrating factor, nu = 1e-3, seeded random-phase divergence-free IC with
E(k) ~ k^4 exp(-2 (k/4)^2), total energy 0.5, dt fixed once from CFL 0.5 at t = 0.
value = FP32-mixed throughput (grid points x RK stages per second, inputs resident in HBM);
the FP64 GPU path is timed on the same IC and reported beside it, with the E(k) error of the
FP32-mixed run against the FP64 run after the same number of steps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu] [--no-fp64]
  python bench.py --workload ensemble [--steps K] [--warmup W] [--no-cpu]
  python bench.py --workload mhd [--n 512]      python bench.py --workload rb [--rb-n 256]
  python bench.py --workload forced [--n 512]   (the N^3 step with energy-supply forcing)
  python bench.py --workload dft16              (16-point DFTs: VALU codelet vs MFMA DFT-as-GEMM)
The default run also times the MHD 512^3, RB 256^3 and 64-box ensemble workloads and reports the main fields
of their own lines under other_workloads (--no-extra skips them).
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


# Canonical (unpruned) bytes per pass in units of S = N1 N2 (N3/2+1) elements: every pass reads and
# writes whole spectral-size component arrays once (NS: 3 in / 6 products; MHD: 6 in / 9 products).
# The fused kernels move less (dealias-compacted rows, kz < N3/3); their own byte counts come from
# th_profile_passes and the measured HBM traffic from rocprofv3.
CANON_S = {"ns": [6, 6, 9, 12, 18], "mhd": [12, 12, 15, 18, 33], "rb": [8, 8, 13, 18, 25]}


def pass_times(state, dt, nrep=3, **phys):
    """Mean HIP-event duration (ms) and algorithmic bytes of each of the five passes over RK stages
    1..4 (stages 1 and 4 touch one state array less), replayed pass by pass on the plan's stream
    with the state left by the timed steps; the mean matches a kernel trace over whole steps."""
    import numpy as np
    ms, nb = np.zeros(5), np.zeros(5)
    for stage in (1, 2, 3, 4):
        m, b = state.profile_passes(dt, stage=stage, nrep=nrep, **phys)
        ms += m / 4
        nb += b / 4
    return ms, nb


def pass_names(nbytes):
    """The five pass names; a chained step (the x-inverse done inside the previous stage's x-forward, so no bytes
    of its own) names the x-forward after both."""
    names = list(PASS_NAMES)
    if nbytes[0] == 0:
        names[4] = "x_forward_epilogue_chained_x_inverse"
    return names


def roofline(ms, nbytes, N, esize, physics, traffic=None):
    """Roofline of the dominant (longest) pass: `achieved` = the bytes that kernel must move per launch (its
    dealias-pruned reads and writes of whole compacted arrays, each once) / its HIP-event duration.  The
    canonical bytes of the unpruned pipeline (every pass reading and writing whole (N1, N2, N3/2+1) arrays)
    divided by the same duration are reported beside it as an effective rate, not as a fraction of peak."""
    import numpy as np
    S = float(N[0]) * N[1] * (N[2] // 2 + 1) * esize
    k = int(np.argmax(ms))
    alg = float(nbytes[k])
    ach = alg / (ms[k] * 1e-3) / 1e9
    canon = CANON_S[physics][k] * S
    return {"bound": "hbm", "kernel": pass_names(nbytes)[k], "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": alg,
            "launch_ms": round(float(ms[k]), 4),
            "bytes_model": "compacted kept-mode arrays, each read and written once per launch",
            "canonical_bytes_per_launch": canon,
            "effective_canonical_GB_s": round(canon / (ms[k] * 1e-3) / 1e9, 1)}, sum(CANON_S[physics]) * S


def host_info(threads):
    """nproc, the OpenMP threads used, OMP_NUM_THREADS and the CPU model of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "threads": threads, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": model}


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
    """Max relative E(k) error of Va against Vb over shells with E_b(k) >= 1e-10 max E_b, the l2 relative error
    over all shells, and the shell that sets the maximum (k, E_b(k) / max E_b, its error)."""
    import numpy as np
    Ea = plan64.vector_spectrum(Va)
    Eb = plan64.vector_spectrum(Vb)
    sel = Eb >= 1e-10 * Eb.max()
    rel = np.where(sel, np.abs(Ea - Eb) / np.where(Eb > 0, Eb, 1.0), 0.0)
    k = int(np.argmax(rel))
    l2 = float(np.linalg.norm(Ea - Eb) / np.linalg.norm(Eb))
    return float(rel[k]), l2, {"k": k, "E_over_max": float(Eb[k] / Eb.max()), "rel_err": float(rel[k])}


def divergence_rel(V, N):
    """max_k |K . V(k)| / max_k |K| |V(k)| of a reference-layout vector field (device tensors)."""
    import torch
    dev = V[0].device

    def signed(n):
        l = torch.arange(n, device=dev, dtype=torch.float64)
        return torch.where(l <= n // 2, l, l - n)
    kx = signed(N[0]).view(-1, 1, 1)
    ky = signed(N[1]).view(1, -1, 1)
    kz = torch.arange(N[2] // 2 + 1, device=dev, dtype=torch.float64).view(1, 1, -1)
    div = (kx * V[0] + ky * V[1] + kz * V[2]).abs().max()
    mag = (torch.sqrt(kx * kx + ky * ky + kz * kz) *
           torch.sqrt(V[0].abs() ** 2 + V[1].abs() ** 2 + V[2].abs() ** 2)).max()
    return float(div / mag)


def cpu_baseline(V_host, N, nu, dt, threads):
    """The FP64 oracle (the reference algorithm restated in C: split transforms, 9 whole-array transforms per
    stage, separate array passes, OpenMP over x-slabs) timed on one full RK4 step (4 stages) of the same
    workload on this host's cores; returns (seconds, threads used, the oracle's result fields)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    o = Oracle(N, threads=threads)
    V = [v.copy() for v in V_host]
    t0 = time.perf_counter()
    o.rk4_ns(V, nu, dt, steps=1)
    el = time.perf_counter() - t0
    return el, o.L.orc_get_threads(), V


def config1(ctx, threads):
    """The CPU-runnable configuration: Taylor-Green 64^3, FP64, RK4, dt = 1e-3, nu = 0.01, 100 steps.  The
    fused FP64 GPU step and the oracle each run it in full; E(k) over unit shells and Pi(R) over spheres
    R = 1..21, INF of the two runs are compared (max relative error over |ref| >= 1e-10 max |ref|)."""
    import numpy as np
    import torch
    from tarang_amd import Plan, State
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    N = (64, 64, 64)
    nu, dt, steps = 0.01, 1e-3, 100
    p = Plan(ctx, N, precision="fp64")
    s = State(p)
    s.upload(p.ic_taylor_green(1, 1.0))
    s.step(dt, nu=nu, steps=1)                   # graph capture outside the timed run (then restart from TG)
    s.upload(p.ic_taylor_green(1, 1.0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(dt, nu=nu, steps=steps)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    g = s.download()
    o = Oracle(N, threads=threads)
    V = o.ic_taylor_green(1, 1.0)
    t0 = time.perf_counter()
    o.rk4_ns(V, nu, dt, steps=steps)
    tc = time.perf_counter() - t0
    Er, Eg = o.vector_spectrum(V), p.vector_spectrum(g)
    sel = Er >= 1e-10 * Er.max()
    radii = np.concatenate([np.arange(0, 22, dtype=float), [10000.0]])
    fo = o.flux_ns(V, radii)
    s2 = State(p)
    s2.upload(g)
    fg, _ = s2.flux_shell_tr(radii)
    fsel = np.abs(fo) >= 1e-10 * np.abs(fo).max()
    pts = 64.0 ** 3 * 4 * steps
    return {"workload": "taylor_green_64^3_fp64_rk4_dt1e-3_nu0.01_100_steps",
            "gpu_fp64_fused_s": round(tg, 4), "gpu_value": round(pts / tg / 1e6, 1),
            "cpu_oracle_s": round(tc, 3), "cpu_value": round(pts / tc / 1e6, 3), "cpu_threads": o.L.orc_get_threads(),
            "Ek_rel_err_gpu_vs_cpu_max": float((np.abs(Eg - Er)[sel] / Er[sel]).max()),
            "Pi_rel_err_gpu_vs_cpu_max": float((np.abs(fg - fo)[fsel] / np.abs(fo)[fsel]).max()),
            "unit": "Mpts*RKstage/s"}


def load_traffic(workload):
    """Measured HBM bytes per launch of each pass (profiles/traffic_<workload>.json: rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs scaled by a copy kernel of known size), keyed by pass and then by the launch variant the
    counters were read on: {pass: {"<variant>": {"kernel": ..., "bytes": ..., ...}}}."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def traffic_of(table, name, variant):
    """The measured bytes of pass `name` when the counters were read on the same launch variant, else None."""
    e = (table or {}).get(name)
    e = e.get(str(variant)) if isinstance(e, dict) else None
    return e.get("bytes") if isinstance(e, dict) else None


ENSEMBLE_SHAPES = [64] * 24 + [96] * 24 + [128] * 16


def bench_ensemble(args, emit=True):
    """The uneven 64-box ensemble (24 x 64^3, 24 x 96^3, 16 x 128^3), FP32-mixed, box b seeded
    100 + b, nu_b ~ U[5e-4, 5e-3] (numpy default_rng(5)), dt_b from CFL 0.5 at t = 0 (FP64).
    value = sum_b N_b^3 x 4 stages per step / wall of the batched, graph-captured ensemble step;
    the same boxes stepped one after another with the single-box fused step are timed beside it."""
    import ctypes as C
    import numpy as np
    import torch
    from tarang_amd import Context, Ensemble, Plan, State, abi

    shapes = [(n, n, n) for n in ENSEMBLE_SHAPES]
    nb = len(shapes)
    nu = np.random.default_rng(5).uniform(5e-4, 5e-3, nb)
    ctx = Context(0)
    ens = Ensemble(ctx, shapes, precision="fp32")
    p64 = {n: Plan(ctx, (n, n, n), precision="fp64") for n in sorted(set(ENSEMBLE_SHAPES))}
    dt = np.zeros(nb)
    E0 = np.zeros(nb)
    singles = []
    host_ic = []
    for b, N in enumerate(shapes):
        V = p64[N[0]].ic_random(100 + b, 0.5, 4.0)
        dt[b], _ = cfl_dt(p64[N[0]], V)
        E0[b] = p64[N[0]].total_energy(V)
        V32 = [v.to(torch.complex64) for v in V]
        ens.upload(b, V32)
        ps = Plan(ctx, N, precision="fp32")
        st = State(ps)
        st.upload(V32)
        singles.append(st)
        if not args.no_cpu and b in (0, 24, 48):
            host_ic.append((b, [v.cpu().numpy() for v in V]))
        del V, V32
    log(f"bench: ensemble of {nb} boxes ready, dt in [{dt.min():.3e}, {dt.max():.3e}]")
    if not args.no_tune:
        ens.tune(dt, nu, nrep=3)
        tuned = {}
        for b, st in enumerate(singles):
            n = shapes[b][0]
            if n not in tuned:
                tuned[n] = st.tune(float(dt[b]), nu=float(nu[b]), nrep=3)
            else:
                st.variants = tuned[n]
        log(f"bench: tuned single-box variants {tuned}")
    L = abi.lib()
    phys = [abi.Phys(float(nu[b]), 0.0, 0.0, 0.0, 0.0) for b in range(nb)]

    def seq(steps):
        # each box advanced alone by its own captured single-box step, one box after another
        for b, st in enumerate(singles):
            abi._check(L.th_step(st.plan.h, st.h, float(dt[b]), C.byref(phys[b]), steps), "th_step")
        for st in singles:
            abi._check(L.th_sync(st.plan.h), "th_sync")

    def batched(steps):
        ens.step(dt, nu, steps=steps)

    res = {}
    for name, fn in (("ensemble", batched), ("sequential", seq)):
        fn(args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(args.steps)
        torch.cuda.synchronize()
        res[name] = time.perf_counter() - t0
        log(f"bench: {name} {args.steps} steps {res[name]:.4f} s")
    # the batched and the sequential runs advanced every box by the same steps from the same IC: bit-identical
    # where a class runs the same launch variants as its single boxes (the two are tuned separately; other
    # variants round differently in FP32)
    worst, same_var_diff = 0.0, 0.0
    var_cls = {}
    for b in (0, 24, 48, nb - 1):
        a, r = ens.download(b), singles[b].download()
        d = max(float((x - y).abs().max()) for x, y in zip(a, r))
        worst = max(worst, d)
        ve, vs = ens.variants(b), singles[b].variants
        var_cls[shapes[b][0]] = {"ensemble": ve, "sequential": vs}
        if ve == vs:
            same_var_diff = max(same_var_diff, d)
    pts = float(sum(N[0] * N[1] * N[2] for N in shapes))
    value = pts * 4 * args.steps / res["ensemble"] / 1e6
    seqv = pts * 4 * args.steps / res["sequential"] / 1e6
    # bytes the pruned kernels must move per RK stage (compacted arrays, each read and written once; mean over
    # stages 1-4 of one box per size class, from th_profile_passes), summed over the boxes; the canonical model
    # (51 S_b per box, S_b = N1 N2 (N3/2+1) x 8 B) beside it
    per_class = {}
    for b, N in enumerate(shapes):
        if N[0] not in per_class:
            nbytes = 0.0
            for stage in (1, 2, 3, 4):
                _, by = singles[b].profile_passes(float(dt[b]), nu=float(nu[b]), stage=stage, nrep=1)
                nbytes += float(by.sum()) / 4
            per_class[N[0]] = nbytes
    stage_bytes = sum(per_class[N[0]] for N in shapes)
    canon_bytes = sum(51.0 * N[0] * N[1] * (N[2] // 2 + 1) * 8 for N in shapes)
    ach = stage_bytes * 4 * args.steps / res["ensemble"] / 1e9
    cpu = None
    if not args.no_cpu and host_ic:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from pyoracle import Oracle
        threads = args.cpu_threads if args.cpu_threads > 0 else min(os.cpu_count() or 1, 16)
        tot_pts, tot_s, used = 0.0, 0.0, 1
        for b, Vh in host_ic:
            N = shapes[b]
            o = Oracle(N, threads=threads)
            V = [v.copy() for v in Vh]
            t0 = time.perf_counter()
            o.rk4_ns(V, float(nu[b]), float(dt[b]), steps=1)
            tot_s += time.perf_counter() - t0
            tot_pts += 4.0 * N[0] * N[1] * N[2]
            used = o.L.orc_get_threads()
        cpu = {"value": round(tot_pts / tot_s / 1e6, 3), "unit": "Mpts*RKstage/s", "cores": int(used),
               "kind": "port", "seconds": round(tot_s, 2),
               "sample": "1 full RK4 step of one 64^3, one 96^3 and one 128^3 box of the ensemble "
                         "(boxes 0, 24, 48) by the FP64 oracle, OpenMP"}
    line = {
        "metric": "Mpts*RKstage/s, uneven 64-box NS ensemble, FP32-mixed; speedup over sequential per-box",
        "value": round(value, 1),
        "unit": "Mpts*RKstage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * res["ensemble"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp32 storage/FFT; fp64 integrating factor, twiddles, sums)",
        "data": "synthetic (seeded random-phase ICs generated on the GPU)",
        "config": {"workload": "ensemble_24x64^3_24x96^3_16x128^3_ns_rk4", "boxes": nb,
                   "points": int(pts), "nu": "U[5e-4,5e-3] rng 5", "dt": "per box, CFL 0.5 at t=0",
                   "ic": "random-phase k^4 exp(-2(k/4)^2), E=0.5, seed 100+b", "precision": "fp32-mixed"},
        "sequential": {"value": round(seqv, 1), "ms_per_step": round(1e3 * res["sequential"] / args.steps, 3),
                       "speedup_ensemble_over_sequential": round(value / seqv, 3),
                       "max_abs_diff_sampled_boxes": worst,
                       "max_abs_diff_sampled_boxes_same_variants": same_var_diff,
                       "launch_variants_by_size": var_cls},
        "roofline": {"bound": "hbm", "kernel": "whole_ensemble_step", "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_stage": stage_bytes,
                     "bytes_model": "pruned kernels' compacted arrays, each read and written once per pass, summed "
                                    "over the 5 passes and 64 boxes (the 64^3 and 96^3 working sets fit in the "
                                    "256 MB last-level cache, so HBM traffic can be lower)",
                     "canonical_bytes_per_stage": canon_bytes,
                     "effective_canonical_GB_s": round(canon_bytes * 4 * args.steps / res["ensemble"] / 1e9, 1)},
        "cpu_baseline": cpu,
    }
    return _emit(line, emit)


def _emit(line, emit):
    if emit:
        print(json.dumps(line), flush=True)
    return line


def _line(metric, value, args, ms_step, config, extra):
    line = {"metric": metric, "value": round(value, 1), "unit": "Mpts*RKstage/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None}
    line.update(extra)
    line["config"] = config
    return line


def bench_mhd(args, emit=True):
    """512^3 incompressible MHD (Elsasser products), FP32-mixed fused step, nu = eta = 1e-3, U seed 2 and
    B seed 3 with the random-phase k^4 exp(-2(k/4)^2) spectrum (E = 0.5 each; B's component along U removed
    so the cross helicity is 0), dt from CFL 0.5 on max(|u|, |b|) at t = 0.  Algorithmic bytes per stage: 90 S (S = N^2 (N/2+1) x 8 B)."""
    import numpy as np
    import torch
    from tarang_amd import Context, Plan, State, abi
    n = args.n
    N = (n, n, n)
    nu = eta = 1e-3
    ctx = Context(0)
    p64 = Plan(ctx, N, precision="fp64")
    p32 = Plan(ctx, N, precision="fp32")
    U = p64.ic_random(2, 0.5, 4.0)
    B = p64.ic_random(3, 0.5, 4.0)
    # cross helicity near 0: the part of B along U is removed (B <- B - (Hc / E_u) U, divergence-free and
    # dealiased like both) and B is rescaled to E_b = 0.5
    Eu, Eb0 = p64.total_energy(U), p64.total_energy(B)
    Hc0 = sum(p64.energy(a, b) for a, b in zip(U, B))
    B = [b - (Hc0 / Eu) * a for a, b in zip(U, B)]
    sc = math.sqrt(0.5 / p64.total_energy(B))
    B = [b * sc for b in B]
    dtu, umax = cfl_dt(p64, U)
    dtb, bmax = cfl_dt(p64, B)
    dt = min(dtu, dtb)
    Eu, Eb = p64.total_energy(U), p64.total_energy(B)
    Hc = sum(p64.energy(a, b) for a, b in zip(U, B))
    s = State(p32, abi.TH_MHD)
    s.upload([a.to(torch.complex64) for a in U + B])
    del U, B
    if args.variants:
        s.variants = [int(x) for x in args.variants.split(",")]
    var = s.tune(dt, nu=nu, eta=eta, nrep=3) if not (args.no_tune or args.variants) else s.variants
    s.step(dt, nu=nu, eta=eta, steps=args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(dt, nu=nu, eta=eta, steps=args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out = s.download()
    E1 = p32.total_energy(out[:3]) + p32.total_energy(out[3:])
    ms, nbytes = pass_times(s, dt, nu=nu, eta=eta)
    # the nine MHD fluxes (u<u, b<b, u<b, u>b and the rest of Compute_flux(IncVF& W)) over 32 unit shells from
    # the fused giver-shell pipeline, timed once (outside the step metric)
    import numpy as np
    radii = np.concatenate([np.arange(32), [10000.0]]).astype(float)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fl = s.flux_mhd(radii)
    flux_s = time.perf_counter() - t0
    pts = float(n) ** 3
    value = pts * 4 * args.steps / el / 1e6
    traffic = load_traffic(f"mhd{n}_fp32mixed")
    k = int(np.argmax(ms))
    roof, canon_stage = roofline(ms, nbytes, N, 8, "mhd", traffic_of(traffic, pass_names(nbytes)[k], var[k]))
    roof["traffic_source"] = (f"profiles/traffic_mhd{n}_fp32mixed.md (same launch variant)" if roof["traffic"] is not None
                              else "no counter reading for this launch variant")
    passes = {pass_names(nbytes)[i]: {"ms": round(float(ms[i]), 4), "GB": round(nbytes[i] / 1e9, 4),
                              "GB_s": round(nbytes[i] / (ms[i] * 1e-3) / 1e9, 1) if ms[i] > 0 else None} for i in range(5)}
    return _emit(_line(
        f"Mpts*RKstage/s at {n}^3 MHD (Elsasser), FP32-mixed fused step", value, args, 1e3 * el / args.steps,
        {"workload": f"mhd_{n}^3_rk4_elsasser", "N": n, "nu": nu, "eta": eta, "dt": dt, "cfl": 0.5,
         "ic": "U seed 2, B seed 3 minus its component along U, k^4 exp(-2(k/4)^2), E=0.5 each",
         "precision": "fp32-mixed"},
        {"dtype": "fp32-mixed (fp32 storage/FFT; fp64 integrating factor, twiddles, sums)",
         "data": "synthetic (seeded random-phase IC generated on the GPU)",
         "ic_check": {"E_u": Eu, "E_b": Eb, "Hc_over_E": abs(Hc) / (Eu + Eb), "Hc_over_E_before_projection":
                      abs(Hc0) / (Eu + Eb0), "Hc_over_E_lt_1e-2": bool(abs(Hc) / (Eu + Eb) < 1e-2), "umax": umax, "bmax": bmax},
         "energy_after": E1, "launch_variants": var,
         "flux_mhd_fused": {"seconds": round(flux_s, 3), "spheres": 32,
                            "u2u_max": float(np.abs(fl[0]).max()), "b2b_max": float(np.abs(fl[1]).max()),
                            "u2b_in_out_max": float(np.abs(fl[2]).max()), "u_out_b_out_max": float(np.abs(fl[6]).max())},
         "roofline": roof,
         "stage": {"canonical_GB_per_stage": round(canon_stage / 1e9, 3),
                   "pruned_GB_per_stage": round(float(sum(nbytes)) / 1e9, 3),
                   "GB_s_canonical": round(canon_stage / (float(sum(ms)) * 1e-3) / 1e9, 1), "passes": passes},
         "cpu_baseline": None}), emit)


def bench_forced(args):
    """N^3 forced turbulence: the decaying-turbulence workload (seed 1, nu = 1e-3, dt from CFL 0.5) with the
    energy-supply forcing of force_field_proc 1 on the band 1.5 < |K| <= 3.5 (energy_supply 0.1,
    epsh_by_k_epse 0.2), FP32-mixed fused step, recomputed at every RK stage.  The unforced step of the
    same state is timed beside it, so the line shows what the forcing costs."""
    import torch
    from tarang_amd import Context, Plan, State
    n = args.n
    N = (n, n, n)
    nu = 1e-3
    ctx = Context(0)
    p64 = Plan(ctx, N, precision="fp64")
    p32 = Plan(ctx, N, precision="fp32")
    V = p64.ic_random(1, 0.5, 4.0)
    dt, umax = cfl_dt(p64, V)
    E0 = p64.total_energy(V)
    s = State(p32)
    s.upload([v.to(torch.complex64) for v in V])
    del V
    var = s.tune(dt, nu=nu, nrep=3) if not args.no_tune else s.variants
    res = {}
    for name, force in (("unforced", None), ("forced", (1, 1.5, 3.5, 0.1, 0.2))):
        s.set_force(*force) if force else s.set_force(0)
        s.step(dt, nu=nu, steps=args.warmup)
        torch.cuda.synchronize()
        res[name] = time_steps(s, args.steps, dt, nu)
        log(f"bench: {name} {args.steps} steps {res[name]:.4f} s")
    E1 = p32.total_energy(s.download())
    pts = float(n) ** 3
    value = pts * 4 * args.steps / res["forced"] / 1e6
    print(json.dumps(_line(
        f"Mpts*RKstage/s at {n}^3 forced turbulence (energy supply), FP32-mixed fused step", value, args,
        1e3 * res["forced"] / args.steps,
        {"workload": f"ns_{n}^3_rk4_forced_supply", "N": n, "nu": nu, "dt": dt, "cfl": 0.5,
         "force": {"force_field_proc": 1, "inner_radius": 1.5, "outer_radius": 3.5, "energy_supply": 0.1,
                   "epsh_by_k_epse": 0.2},
         "ic": "seed 1, k^4 exp(-2(k/4)^2), E=0.5", "precision": "fp32-mixed"},
        {"dtype": "fp32-mixed (fp32 storage/FFT; fp64 integrating factor, twiddles, sums, force)",
         "data": "synthetic (seeded random-phase IC generated on the GPU)",
         "ic_check": {"E0": E0, "umax": umax}, "energy_after": E1, "launch_variants": var,
         "unforced": {"value": round(pts * 4 * args.steps / res["unforced"] / 1e6, 1),
                      "ms_per_step": round(1e3 * res["unforced"] / args.steps, 3)},
         "forcing_overhead_ms_per_step": round(1e3 * (res["forced"] - res["unforced"]) / args.steps, 4),
         "roofline": None, "cpu_baseline": None})), flush=True)


def bench_rb(args, emit=True):
    """Rayleigh-Benard, SCFT basis with the sin/cos axis first, N^3 (default 256^3), kfactor
    (pi, pi/sqrt2, pi/sqrt2), Pr = 1, Ra = 1e7 (PRLARGE/ULARGE scaling), V = 0, temperature
    perturbation amplitude 1e-3 (seed 4), dt = 1e-3 fixed.  The fused RB step (DCT/DST x passes, 9
    products in the z pass, force + projection + IF-RK4 epilogue) is timed in FP32-mixed and FP64; the
    plain path (separate transforms and array passes, the reference's operation sequence) is timed beside
    it in FP32-mixed.  Reports Nu after the timed steps from each path.  Algorithmic bytes per stage:
    72 S (S = N1 N2 (N3/2+1) x 8 B)."""
    import numpy as np
    import torch
    from tarang_amd import Context, Plan, State, abi
    n = args.rb_n
    N = (n, n, n)
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    Ra, Pr, dt = 1e7, 1.0, 1e-3
    ctx = Context(0)
    res = {}
    # --no-fp64: the FP32-mixed fused step alone (counter runs); --variants fixes its launch variants
    for prec in (("fp32",) if args.no_fp64 else ("fp32", "fp64")):
        p = Plan(ctx, N, kfactor=kf, basis="SCFT", precision=prec)
        s = State(p, abi.TH_RB)
        s.upload(p.empty(3) + [p.ic_random_scalar(4, 1e-3)])
        if args.variants and prec == "fp32":
            s.variants = [int(x) for x in args.variants.split(",")]
        var = s.tune(dt, Ra=Ra, Pr=Pr, nrep=3) if not (args.no_tune or (args.variants and prec == "fp32")) else s.variants
        s.step(dt, Ra=Ra, Pr=Pr, steps=args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(dt, Ra=Ra, Pr=Pr, steps=args.steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out = s.download()
        ms, nbytes = pass_times(s, dt, Ra=Ra, Pr=Pr)
        res[prec] = (el, p.nusselt(out[:3], out[3], Ra, Pr), ms, nbytes, var)
        del s, out, p
    el_plain, Nu_plain = None, None
    if not args.no_fp64:
        p = Plan(ctx, N, kfactor=kf, basis="SCFT", precision="fp32")
        f = p.empty(3) + [p.ic_random_scalar(4, 1e-3)]
        p.step_plain(abi.TH_RB, f, dt, Ra=Ra, Pr=Pr, steps=args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p.step_plain(abi.TH_RB, f, dt, Ra=Ra, Pr=Pr, steps=args.steps)
        torch.cuda.synchronize()
        el_plain = time.perf_counter() - t0
        Nu_plain = p.nusselt(f[:3], f[3], Ra, Pr)
        del f, p
    pts = float(n) ** 3
    el, Nu, ms, nbytes, var = res["fp32"]
    value = pts * 4 * args.steps / el / 1e6
    v64 = pts * 4 * args.steps / res["fp64"][0] / 1e6 if "fp64" in res else None
    traffic = load_traffic(f"rb{n}_fp32mixed")
    k = int(np.argmax(ms))
    roof, canon_stage = roofline(ms, nbytes, N, 8, "rb", traffic_of(traffic, pass_names(nbytes)[k], var[k]))
    roof["traffic_source"] = (f"profiles/traffic_rb{n}_fp32mixed.md (same launch variant)" if roof["traffic"] is not None
                              else "no counter reading for this launch variant")
    passes = {pass_names(nbytes)[i]: {"ms": round(float(ms[i]), 4), "GB": round(nbytes[i] / 1e9, 4),
                              "GB_s": round(nbytes[i] / (ms[i] * 1e-3) / 1e9, 1) if ms[i] > 0 else None} for i in range(5)}
    return _emit(_line(
        f"Mpts*RKstage/s at {n}^3 Rayleigh-Benard (SCFT), FP32-mixed fused step", value, args,
        1e3 * el / args.steps,
        {"workload": f"rb_scft_{n}^3_rk4", "N": n, "kfactor": "pi, pi/sqrt2, pi/sqrt2", "Ra": Ra, "Pr": Pr, "dt": dt,
         "ic": "V = 0, theta amplitude 1e-3 seed 4", "precision": "fp32-mixed"},
        {"dtype": "fp32-mixed (fp32 storage/FFT; fp64 integrating factor, twiddles, sums)",
         "data": "synthetic (seeded temperature perturbation generated on the GPU)",
         "path": "fused (5 passes per RK stage, hipGraph replay)", "launch_variants": var,
         "Nu_fp32": Nu, "Nu_fp64": res["fp64"][1] if "fp64" in res else None, "Nu_plain_fp32": Nu_plain,
         "fp64_gpu": ({"value": round(v64, 1), "ms_per_step": round(1e3 * res["fp64"][0] / args.steps, 3)}
                      if "fp64" in res else None),
         "plain_fp32_gpu": ({"value": round(pts * 4 * args.steps / el_plain / 1e6, 1),
                             "ms_per_step": round(1e3 * el_plain / args.steps, 3)} if el_plain else None),
         "roofline": roof,
         "stage": {"canonical_GB_per_stage": round(canon_stage / 1e9, 3),
                   "pruned_GB_per_stage": round(float(sum(nbytes)) / 1e9, 3),
                   "GB_s_canonical": round(canon_stage / (float(sum(ms)) * 1e-3) / 1e9, 1), "passes": passes},
         "cpu_baseline": None}), emit)


def bench_dft16(args):
    """The DFT-as-GEMM measurement: the 16-point DFTs of the radix-16 line-FFT stages on 2^22 FP32 complex
    vectors (512 MiB in, 512 MiB out, resident in HBM) by the vector-ALU codelet the line FFTs use and by the
    FP32 matrix cores (th_dft16_probe; the DFT as a real 32 x 32 matrix times 32 vectors, 16
    v_mfma_f32_32x32x2f32 per 32 vectors).  Timed at 1, 8 and 32 applications per launch (the data read
    and written once, the applications in registers), HIP events over 10 launches each; the cost of one
    application is the increment from 8 to 32 (memory traffic identical).  Parity of both against
    numpy's FFT on the first 65536 vectors."""
    import numpy as np
    import torch
    from tarang_amd import Context
    ctx = Context(0)
    nvec = 1 << 22
    g = torch.Generator(device="cuda:0")
    g.manual_seed(5)
    x = torch.randn((16, nvec), dtype=torch.complex64, device="cuda:0", generator=g)
    y = torch.empty_like(x)
    xs = x[:, :65536].cpu().numpy().astype(np.complex128)
    res = {}
    for method, name in ((0, "valu"), (1, "mfma")):
        r = {}
        for reps in (1, 8, 32):
            ms = ctx.dft16_probe(x, y, method, sign=-1, reps=reps, ntimed=10)
            ref = xs.copy()
            for _ in range(reps):
                ref = np.fft.fft(ref, axis=0) / 4
            got = y[:, :65536].cpu().numpy().astype(np.complex128)
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            r[f"reps{reps}"] = {"ms": round(ms, 4), "GDFT16_s": round(nvec * reps / (ms * 1e-3) / 1e9, 2),
                                "GB_s": round(2 * 16 * 8 * nvec / (ms * 1e-3) / 1e9, 1), "max_rel_err": err}
            log(f"bench: dft16 {name} reps={reps} {ms:.4f} ms err {err:.2e}")
        # one application in registers: the 8 -> 32 increment
        inc = (r["reps32"]["ms"] - r["reps8"]["ms"]) / 24
        r["per_application_us"] = round(1e3 * inc, 3)
        r["GDFT16_s_compute"] = round(nvec / (inc * 1e-3) / 1e9, 1)
        res[name] = r
    # executed MFMA work: 16 instructions x 4096 flops per 32 vectors per application
    mfma_tflops = 16 * 4096 * (nvec / 32) / (res["mfma"]["per_application_us"] * 1e-6) / 1e12
    ratio = res["mfma"]["per_application_us"] / res["valu"]["per_application_us"]
    roof = {"bound": "mfma", "kernel": "k_dft16_mfma (one application, 8 -> 32 increment)",
            "achieved": round(mfma_tflops, 1), "peak": 157.3, "unit": "TFLOP/s", "frac": round(mfma_tflops / 157.3, 4),
            "traffic": None, "flops_model": "16 v_mfma_f32_32x32x2f32 x 4096 flops per 32 vectors per application"}
    print(json.dumps(_line(
        "16-point FP32 complex DFTs per second in registers: VALU codelet vs MFMA DFT-as-GEMM (2^22 vectors)",
        res["valu"]["GDFT16_s_compute"], args, res["valu"]["reps8"]["ms"],
        {"workload": "dft16_2^22_vectors_fp32", "nvec": nvec, "layout": "point-major", "sign": -1},
        {"unit": "GDFT16/s", "dtype": "fp32", "data": "synthetic (torch.randn on the GPU, seed 5)",
         "valu": res["valu"], "mfma": res["mfma"], "mfma_over_valu_time_per_application": round(ratio, 2),
         "winner": "valu" if ratio >= 1 else "mfma", "roofline": roof, "cpu_baseline": None})), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["ns512", "ensemble", "mhd", "rb", "forced", "dft16"], default="ns512",
                    help="ns512: the N^3 decaying-turbulence step (default, N from --n); "
                         "ensemble: the uneven 64-box ensemble; mhd: N^3 MHD; rb: Rayleigh-Benard (SCFT); "
                         "forced: the N^3 step with energy-supply forcing")
    ap.add_argument("--rb-n", type=int, default=256)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--no-extra", action="store_true",
                    help="ns512: skip the MHD 512^3, RB 256^3 and 64-box ensemble runs summarised under other_workloads")
    ap.add_argument("--variants", default="",
                    help="ns512 and mhd: launch variants of the five FP32-mixed passes, comma-separated (no autotuning; "
                         "counter runs use this to read each variant)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs one box per GPU on a single MI355X; multi-GPU runs "
              "(--gpus > 1) are not part of it", flush=True)
        sys.exit(2)
    if args.calib_copy_mib > 0:
        # a device copy of known size before the workload: counter runs scale their byte counters by it
        import torch
        src = torch.ones(args.calib_copy_mib * (1 << 18), dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {args.calib_copy_mib} MiB done")
        del src, dst
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "mhd":
        bench_mhd(args)
        return
    if args.workload == "rb":
        bench_rb(args)
        return
    if args.workload == "forced":
        bench_forced(args)
        return
    if args.workload == "dft16":
        bench_dft16(args)
        return

    import numpy as np
    import torch
    from tarang_amd import Context, Plan, State

    N = (args.n, args.n, args.n)
    nu = 1e-3
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
    # launch shapes chosen by the autotuner on this workload (outside the timed region)
    if args.variants:
        s32.variants = [int(x) for x in args.variants.split(",")]
        var32 = s32.variants
    else:
        var32 = s32.tune(dt, nu=nu, nrep=3) if not args.no_tune else s32.variants
    log(f"bench: fp32-mixed launch variants {var32}")
    # FP32-mixed timing
    s32.step(dt, nu=nu, steps=args.warmup)
    torch.cuda.synchronize()
    t32 = time_steps(s32, args.steps, dt, nu)
    log(f"bench: fp32-mixed {args.steps} steps {t32:.4f} s")
    # per-pass timing with HIP events on the library stream (mean over RK stages 1-4)
    out32 = s32.download()
    ms, nbytes = pass_times(s32, dt, nu=nu)

    pts = float(N[0]) * N[1] * N[2]
    value = pts * 4 * args.steps / t32 / 1e6
    result = {"fp64": None}
    if not args.no_fp64:
        s64 = State(p64)
        s64.upload(V)
        var64 = s64.tune(dt, nu=nu, nrep=3) if not args.no_tune else s64.variants
        s64.step(dt, nu=nu, steps=args.warmup)
        torch.cuda.synchronize()
        t64 = time_steps(s64, args.steps, dt, nu)
        log(f"bench: fp64 {args.steps} steps {t64:.4f} s")
        out64 = s64.download()
        v64 = pts * 4 * args.steps / t64 / 1e6
        maxrel, l2, worst = spectrum_err(p64, [a.to(torch.complex128) for a in out32], out64)
        # flux Pi(R) and shell-to-shell transfer T(n,m) over 32 unit shells (radii 0..31, inf) from
        # the fused per-giver-shell pipeline, FP32-mixed against FP64
        radii = np.concatenate([np.arange(32), [10000.0]]).astype(float)
        Pi32, T32 = s32.flux_shell_tr(radii)
        Pi64, T64 = s64.flux_shell_tr(radii)
        result["fp64"] = {"value": round(v64, 1), "ms_per_step": round(1e3 * t64 / args.steps, 3),
                          "speedup_fp32_mixed": round(value / v64, 3),
                          "Ek_rel_err_fp32_vs_fp64_max": maxrel, "Ek_rel_err_fp32_vs_fp64_l2": l2,
                          "Ek_gate": {"max_rel_err_le": 1e-5, "pass": bool(maxrel <= 1e-5), "worst_shell": worst,
                                      "shells": "E_fp64(k) >= 1e-10 max E_fp64",
                                      "horizon_steps": args.warmup + args.steps},
                          "Pi_rel_err_fp32_vs_fp64_max": float(np.abs(Pi32 - Pi64).max() / np.abs(Pi64).max()),
                          "T_rel_err_fp32_vs_fp64_max": float(np.abs(T32 - T64).max() / np.abs(T64).max()),
                          "divergence_rel_fp64": divergence_rel(out64, N),
                          "steps_compared": args.warmup + args.steps, "launch_variants": var64}
        del s64, out64

    # roofline of the dominant kernel (the longest pass)
    traffic = load_traffic(f"ns{args.n}_fp32mixed")
    k = int(np.argmax(ms))
    roof, canon_stage = roofline(ms, nbytes, N, 8, "ns", traffic_of(traffic, pass_names(nbytes)[k], var32[k]))
    roof["traffic_source"] = ("profiles/traffic_ns512_fp32mixed.md (same launch variant)" if roof["traffic"] is not None
                              else "no counter reading for this launch variant")
    stage_bytes = float(sum(nbytes))
    stage_ms = float(sum(ms))
    passes = {pass_names(nbytes)[i]: {"ms": round(float(ms[i]), 4), "GB": round(nbytes[i] / 1e9, 4),
                              "GB_s": round(nbytes[i] / (ms[i] * 1e-3) / 1e9, 1) if ms[i] > 0 else None} for i in range(5)}

    cpu = None
    if not args.no_cpu:
        # the GPU box gives one GPU's share of the host to a job: 16 threads unless --cpu-threads says otherwise
        threads = args.cpu_threads if args.cpu_threads > 0 else min(os.cpu_count() or 1, 16)
        Vh = [a.cpu().numpy() for a in V]
        del V
        log(f"bench: CPU oracle, one RK4 step at {N}, {threads} threads")
        el, used, _ = cpu_baseline(Vh, N, nu, dt, threads)
        del Vh
        log(f"bench: CPU oracle step {el:.2f} s; config 1 (TG 64^3, 100 steps) on GPU and CPU")
        c1 = config1(ctx, threads)
        cpu = {"value": round(pts * 4 / el / 1e6, 3), "unit": "Mpts*RKstage/s", "cores": int(used),
               "kind": "port", "seconds": round(el, 2), "steps_measured": 1,
               "sample": f"1 full RK4 step (4 stages, each: 9 whole-array transforms + projection + IF update) of "
                         f"the same {N[0]}^3 FP64 workload by the oracle (reference algorithm restated in C, own "
                         f"FFT, OpenMP over x-slabs)",
               "host": host_info(int(used)), "config1": c1}
    else:
        del V

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
        "divergence_rel_fp32": divergence_rel(out32, N),
        "roofline": roof,
        "launch_variants": {"passes": PASS_NAMES, "fp32_mixed": var32},
        "stage": {"ms_per_stage_sum_of_passes": round(stage_ms, 4),
                  "canonical_GB_per_stage": round(canon_stage / 1e9, 3),
                  "GB_s_canonical": round(canon_stage / (stage_ms * 1e-3) / 1e9, 1),
                  "pruned_GB_per_stage": round(stage_bytes / 1e9, 3),
                  "GB_s_pruned": round(stage_bytes / (stage_ms * 1e-3) / 1e9, 1), "passes": passes},
        "cpu_baseline": cpu,
    }
    if not args.no_extra and not args.variants:
        line["other_workloads"] = other_workloads(args)
    print(json.dumps(line), flush=True)


def other_workloads(args):
    """The MHD 512^3, RB 256^3 and 64-box ensemble benchmarks (`--workload mhd | rb | ensemble`, same steps and
    warm-up, no CPU legs), summarised so that the default run measures every configuration."""
    import copy
    import gc
    import torch
    gc.collect()
    torch.cuda.empty_cache()
    ea = copy.copy(args)
    ea.no_cpu = True
    ea.variants = None       # the NS line's fixed launch variants do not apply to the other workloads
    ea.no_fp64 = False       # and the RB line keeps its FP64 and plain-path legs
    out = {}
    for name, fn in (("mhd_512", bench_mhd), ("rb_256", bench_rb), ("ensemble_64_boxes", bench_ensemble)):
        log(f"bench: other workload {name}")
        d = fn(ea, emit=False)
        r = d.get("roofline") or {}
        e = {"metric": d["metric"], "value": d["value"], "unit": d["unit"], "ms_per_step": d["ms_per_step"],
             "workload": d["config"]["workload"],
             "roofline": {k: r.get(k) for k in ("kernel", "achieved", "frac", "unit", "traffic")}}
        for k in ("flux_mhd_fused", "energy_after", "Nu_fp32", "Nu_fp64", "Nu_plain_fp32", "fp64_gpu", "plain_fp32_gpu",
                  "sequential", "launch_variants"):
            if k in d:
                e[k] = d[k]
        out[name] = e
        gc.collect()
        torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
