"""Benchmark: particle-steps/s of the 3D 1M-particle Oldroyd-B step on one MI355X.

Workload: 100^3 lattice, walls normal to y moving at -/+0.25, 16 rotating
magnetic colloids (R = 0.05), tau = 1, eta = 1, c = 20, np = 100, kT = 1e-3,
dt = 2e-5, seeded velocity noise sigma = 1e-3 (1,059,xxx particles incl. walls
and colloid shells).  One step = VV1 + neighbour search + density + velocity
gradient/dc-dt + forces + wall/colloid forces + VV2, all on the GPU, state
resident in HBM.  `value` is the FP32 form (pair arithmetic in float,
positions and reductions in double); the FP64 form is reported beside it.

Also measured: the 64^3 periodic sine-forced flow in both forms (the
"c2_periodic_64cubed" object), and (the "sweep" object) the batched parameter
sweep - 256
independent 2D wall + colloid simulations of uneven size (n per side 32-160,
V, tau, colloid count, field frequency and strength drawn from a seed), FP32
form, all in one batch whose step is one launch per phase, captured as a
hipGraph, with the launch shape picked by the autotuner - against the same
simulations run one after another.

Prints one JSON line.  --gpus above 1 is not supported (one-GPU build).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sph-magnetic-particles-oldroyd-b-in-cuda_amd")
ORACLE = os.path.join(ROOT, "oracle")
sys.path[:0] = [PKG, ORACLE]

import inputs  # noqa: E402
import sph_system  # noqa: E402
import surrogate  # noqa: E402
import sweep  # noqa: E402

HBM_PEAK_GBS = 8000.0
# MI355X FP32 vector peak: 157.3 TFLOP/s counts packed FP32 (v_pk_fma_f32, two
# lanes per instruction); this build compiles without packed FP32, whose
# ceiling is half of it.
FP32_VALU_PEAK_TFLOPS = 157.3
FP32_VALU_UNPACKED_TFLOPS = 78.65
# Compulsory HBM bytes per particle-step of the whole FP32-form 3D step
# (hash 32 + sort 16 + reorder 136 + pass A 92 + pass B 156).
STEP_BYTES_3D_FP32 = 432
# Flops of one fluid-fluid pair in k_pairs from the reference's formulas
# (src/kernel_vel_gradients.cu:632-640, src/kernel_forces.cu:429-440), an FMA
# counted as two: separation and r^2 8, r and 1/r 2, e 3, gradW 16, v_ij 3,
# velocity gradient 22, reversible stress force 30, e.v_ij 5, viscous
# coefficients 6, viscous force 12.
FLOPS_PER_PAIR = 107


def kernel_bytes(phase, dim, rbytes, store_gradv=False):
    """Algorithmic (compulsory) HBM bytes per particle for one launch of the
    phase's main kernel: every array the operation must read or write, once
    (neighbour data re-reads and the neighbour list are not counted)."""
    pos = 8 * dim
    v, c = rbytes * dim, rbytes * (6 if dim == 3 else 3)
    if phase == "forces":      # k_pairs: x v c type dens press mass -> F dc/dt, kicked v c
        return pos + v + c + 4 + 3 * rbytes + v + c + v + c
    if phase == "density":     # k_density: x mass -> dens press
        return pos + rbytes + 2 * rbytes
    if phase == "gradv_dcdt":  # x v type dens c -> dcdt (+ gradv)
        return pos + v + 4 + rbytes + c + c + (rbytes * dim * dim if store_gradv else 0)
    if phase == "vv1":         # type x v F mass c dcdt -> x v c
        return 4 + pos + v + v + rbytes + c + c + pos + v + c
    if phase == "vv2":         # type v F mass c dcdt -> v c
        return 4 + v + v + rbytes + c + c + v + c
    if phase == "neighbours":  # hash + scatter + rank/gather of x v c mass type id
        moving = pos + v + c + rbytes + 4 + 4
        return (pos + 8) + 20 + (8 + 4 + moving + moving + 8)
    return 0


TRAFFIC_FILE = "profiles/traffic_walls_colloids_1M_fp32.json"
PAIR_PROFILE = "profiles/pair_passes_walls_colloids_1M_fp32.md"
# VALU instructions x 4 cycles / SIMD busy time of the FP32 pair passes on the
# 1M-particle workload (SQ_INSTS_VALU, SQ_BUSY_CYCLES; see PAIR_PROFILE)
VALU_BUSY = {"k_pairs": 0.63, "k_density": 0.88}


def measured_traffic(kernel, N):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (scaled by a copy kernel of known size) on this workload, from
    the committed profile; None when no measurement of this kernel and size."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None
    if kernel not in t or f"N={N}" not in t.get("workload", ""):
        return None
    return t[kernel]["traffic_bytes"]


def make_system(precision, n, n_colloids, tmpdir, seed=3):
    cfg = inputs.walls_colloids_3d(n=n, n_colloids=n_colloids, seed=seed)
    path = os.path.join(tmpdir, f"input_{precision}")
    inputs.write_input(path, cfg)
    s = sph_system.System(precision)
    s.read_input(path)
    s.initialize()
    inputs.seeded_velocity(s, 1e-3, seed)
    return s, cfg


def time_gpu(precision, n, n_colloids, steps, warmup, tmpdir, amortize_steps=300):
    s, cfg = make_system(precision, n, n_colloids, tmpdir)
    s.prepare()
    s.step(warmup)
    s.synchronize()
    c0 = s.counters()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    t1 = time.perf_counter()
    c1 = s.counters()
    elapsed = t1 - t0
    # A window of several Verlet-list rebuild periods: the amortized cost of a
    # step, rebuilds included (the timed window above may hold none).
    t2 = time.perf_counter()
    s.step(amortize_steps)
    s.synchronize()
    t3 = time.perf_counter()
    c2 = s.counters()
    # Per-phase device time with HIP events on the library's own stream.
    s.set_timing(True)
    s.get_timing()  # reset
    s.step(max(3, min(steps, 10)))
    s.synchronize()
    phases = s.get_timing()
    s.set_timing(False)
    counts = s.counters()
    ke = s.kinetic_energy()
    N = s.N
    s.close()
    return dict(N=N, elapsed=elapsed, phases=phases, ke=ke, cfg=cfg,
                rebuilds_in_window=c1["rebuilds"] - c0["rebuilds"],
                amortized=dict(steps=amortize_steps, seconds=t3 - t2,
                               rebuilds=c2["rebuilds"] - c1["rebuilds"]),
                pairs=counts["pairs"], list_entries=counts["list_entries"])


def _time_oracle(o, warmup, steps, repeats):
    """Median wall time of `steps` oracle steps over `repeats` repeats after
    `warmup` untimed steps."""
    o.prepare()
    o.step(warmup)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        o.step(steps)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times


def cpu_baseline(n, n_colloids, tmpdir, steps=5, warmup=10, repeats=3, with_small=True):
    """The FP64 OpenMP oracle on this host's cores: the 1M-particle workload
    (the bench's), and the 2D Couette (full 2000 steps) and 3D 64^3 periodic
    flows (20 steps), each after `warmup` steps, median of `repeats`; the
    AVX-512 build on hosts that have it."""
    import oracle

    lib = oracle.LIB_V4_PATH if oracle.host_has_avx512() and os.path.exists(oracle.LIB_V4_PATH) \
        else oracle.LIB_PATH
    out = {}
    s, _ = make_system(sph_system.FP64, n, n_colloids, tmpdir)
    o = oracle.Oracle.from_system(s, lib_path=lib)
    N = s.N
    s.close()
    med, times = _time_oracle(o, warmup, steps, repeats)
    del o
    out["main"] = dict(value=N * steps / med, N=N, steps=steps, seconds=times)
    if with_small:
        for key, cfg, nsteps in (("couette_2d_64x64", inputs.couette_2d(n=64), 2000),
                                 ("periodic_3d_64cubed", inputs.sine_3d(n=64), 20)):
            path = os.path.join(tmpdir, f"input_cpu_{key}")
            inputs.write_input(path, cfg)
            q = sph_system.System(sph_system.FP64)
            q.read_input(path)
            q.initialize()
            inputs.seeded_velocity(q, 1e-3, 1)
            o = oracle.Oracle.from_system(q, lib_path=lib)
            Nq = q.N
            q.close()
            med, times = _time_oracle(o, warmup, nsteps, repeats)
            del o
            out[key] = dict(value=Nq * nsteps / med, N=Nq, steps=nsteps, seconds=times)
    out["cores"] = oracle.num_threads(lib)
    out["cpu"] = oracle.cpu_model()
    out["nproc"] = os.cpu_count()
    out["isa"] = "x86-64-v4 (AVX-512)" if lib == oracle.LIB_V4_PATH else "x86-64-v3 (AVX2)"
    out["warmup"], out["repeats"] = warmup, repeats
    return out


def time_c2(steps, warmup, tmpdir, seed=5):
    """The 64^3 periodic sine-forced Oldroyd-B flow (262,144 particles, tau 0.5,
    dt 5e-5, force amplitude 1 along x varying in y): particle-steps/s of each
    precision form, with rebuilds counted as in the main workload."""
    out = {"workload": "3D periodic sine-forced flow, 64^3 lattice, tau 0.5, dt 5e-5"}
    for prec, name in ((sph_system.FP32, "fp32"), (sph_system.FP64, "fp64")):
        path = os.path.join(tmpdir, f"input_c2_{name}")
        inputs.write_input(path, inputs.sine_3d(n=64))
        s = sph_system.System(prec)
        s.read_input(path)
        s.initialize()
        inputs.seeded_velocity(s, 1e-3, seed)
        s.prepare()
        s.step(warmup)
        s.synchronize()
        c0 = s.counters()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        t = time.perf_counter() - t0
        c1 = s.counters()
        out[name] = {"value": s.N * steps / t, "ms_per_step": 1e3 * t / steps,
                     "rebuilds_in_window": c1["rebuilds"] - c0["rebuilds"]}
        out["particles"] = s.N
        s.close()
    return out


def time_sweep(n_sims, steps, seq_steps, tmpdir, seed=4):
    """Batched sweep throughput vs the same simulations run one at a time."""
    pts = sweep.sample_points(n_sims, seed=seed)
    systems = [sweep.build_system(p, steps, sph_system.FP32, tmpdir) for p in pts]
    n_part = sum(s.N for s in systems)
    b = sph_system.Batch(sph_system.FP32)
    for s in systems:
        b.add(s)
    b.prepare()
    block, xcd, _ = b.autotune(steps=3)
    b.step(5)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    b.close()
    # one simulation at a time: same kernels, one launch set per simulation
    t_seq = 0.0
    failed = 0
    for s in systems:
        try:
            s.prepare()
            s.step(2)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(seq_steps)
            s.synchronize()
            t_seq += time.perf_counter() - t0
        except sph_system.SphError:
            failed += 1
        s.close()
    batched = n_part * steps / t_batch
    sequential = n_part * seq_steps / t_seq
    return {
        "workload": f"{n_sims} 2D wall + colloid simulations, n per side 32-160 (seed {seed})",
        "sims": n_sims, "particles": n_part, "steps": steps,
        "value": batched, "unit": "particle-steps/s",
        "ms_per_step": 1e3 * t_batch / steps,
        "launch": {"block": block, "xcd": xcd},
        "sequential": sequential, "sequential_steps": seq_steps, "sequential_failed": failed,
        "batch_speedup": batched / sequential,
    }


def time_pipeline(n_train, n_hold, n_sweep, steps, n_range=(32, 160), seed=21, fit_steps=2000,
                  lr=1e-3, log=None):
    """The learned sweep: SPH runs train the MLP ensemble; a new sweep is then
    answered by the ensemble, with the points whose spread of the wall stress
    exceeds 5 % re-run by batched SPH.  Compared with running the whole new
    sweep with batched SPH: wall time and wall-stress error.  Points are drawn
    as in the uneven-size sweep (sides n_range, 10,000 steps by default, the
    observables averaged over the second half).  Runs whose wall stress is not
    that of a steady sheared flow (its samples swing by more than the
    Newtonian wall stress: a colloid configuration gone unstable) count as
    failed, in the training data and in the plain sweep alike
    (`unstable_sph_runs`); left in the training data, a dozen such runs with
    stresses of either sign at many times their scale made the members
    disagree at about a tenth of the sweep's points.  2000 Adam steps."""
    train = sweep.resolve(sweep.sample_points(n_train + n_hold, seed=seed, n_range=n_range))
    Y, tinfo = sweep.run_points(train, steps, log=log)
    X = surrogate.feature_matrix(train)
    model = surrogate.Surrogate(members=5, seed=seed)
    t0 = time.perf_counter()
    model.fit(X[:n_train], Y[:n_train], steps=fit_steps, batch=256, lr=lr, lr_end=1e-5)
    t_fit = time.perf_counter() - t0
    mh, _ = model.predict(X[n_train:])
    ok = np.isfinite(Y[n_train:, 0])
    held_err = np.abs(mh[ok, 0] - Y[n_train:][ok, 0]) / np.abs(Y[n_train:][ok, 0])
    pts = sweep.resolve(sweep.sample_points(n_sweep, seed=seed + 1, n_range=n_range))
    t0 = time.perf_counter()
    Yref, pinfo = sweep.run_points([dict(p) for p in pts], steps, log=log)
    t_plain = time.perf_counter() - t0
    Ys, info = surrogate.sweep_with_fallback(pts, model, steps, threshold=0.05, log=log)
    ok = np.isfinite(Yref[:, 0]) & np.isfinite(Ys[:, 0])
    err = np.abs(Ys[ok, 0] - Yref[ok, 0]) / np.abs(Yref[ok, 0])
    lattice = (f"{n_range[0]}x{n_range[0]} lattice" if n_range[0] == n_range[1] else
               f"n per side {n_range[0]}-{n_range[1]}")
    settle = np.concatenate([tinfo["settle"], pinfo["settle"]])
    return {
        "workload": f"{n_sweep}-point 2D sweep ({lattice}, {steps} steps, dt {sweep.DT}); "
                    f"surrogate trained on {n_train} SPH runs, {n_hold} held out",
        "wall_stress_settle_median": float(np.nanmedian(settle)),
        "plain_batched_s": t_plain, "surrogate_s": info["seconds"],
        "speedup": t_plain / info["seconds"],
        "fallback_points": info["fallback"], "inference_s": info["inference_s"],
        "wall_stress_rel_err_median": float(np.median(err)),
        "wall_stress_rel_err_p90": float(np.percentile(err, 90)),
        "held_out_rel_err_median": float(np.median(held_err)),
        "failed_sph_runs": int(pinfo["failed"] + tinfo["failed"]),
        "unstable_sph_runs": int(pinfo["unstable"] + tinfo["unstable"]),
        "train_data_s": tinfo["seconds"], "fit_s": t_fit,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=100, help="lattice sites per axis")
    ap.add_argument("--colloids", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-sims", type=int, default=256)
    ap.add_argument("--sweep-steps", type=int, default=200)
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--pipeline-points", type=int, default=1000)
    ap.add_argument("--pipeline-steps", type=int, default=10000, help="SPH steps per sweep point")
    ap.add_argument("--fit-steps", type=int, default=2000, help="Adam steps of the surrogate")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the benchmark runs on exactly one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    with tempfile.TemporaryDirectory() as tmpdir:
        r32 = time_gpu(sph_system.FP32, a.n, a.colloids, a.steps, a.warmup, tmpdir)
        print(f"# fp32: N={r32['N']} {a.steps} steps in {r32['elapsed']:.3f}s", file=sys.stderr,
              flush=True)
        r64 = None
        if not a.no_fp64:
            r64 = time_gpu(sph_system.FP64, a.n, a.colloids, a.steps, a.warmup, tmpdir)
            print(f"# fp64: {a.steps} steps in {r64['elapsed']:.3f}s", file=sys.stderr,
                  flush=True)
        c2 = time_c2(max(a.steps, 100), a.warmup, tmpdir)
        print(f"# c2 64^3: fp32 {c2['fp32']['ms_per_step']:.3f} ms/step, fp64 "
              f"{c2['fp64']['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
        sw = None
        if not a.no_sweep:
            sw = time_sweep(a.sweep_sims, a.sweep_steps, 20, tmpdir)
            print(f"# sweep: {sw['sims']} sims, {sw['particles']} particles, "
                  f"{sw['ms_per_step']:.3f} ms/step batched, x{sw['batch_speedup']:.1f} vs one "
                  f"at a time", file=sys.stderr, flush=True)
        pipe = None
        if not a.no_pipeline:
            pipe = time_pipeline(1800, 200, a.pipeline_points, a.pipeline_steps,
                                 fit_steps=a.fit_steps)
            print(f"# pipeline: x{pipe['speedup']:.1f} vs plain batched, "
                  f"{pipe['fallback_points']} of {a.pipeline_points} re-run, median wall-stress "
                  f"error {pipe['wall_stress_rel_err_median']:.3%}", file=sys.stderr, flush=True)
        cpu = None
        if not a.no_cpu:
            cpu = cpu_baseline(a.n, a.colloids, tmpdir, steps=a.cpu_steps)
            print(f"# cpu oracle ({cpu['isa']}, {cpu['cores']} threads, {cpu['cpu']}): "
                  f"{cpu['main']['value']:.3e} particle-steps/s", file=sys.stderr, flush=True)

    N = r32["N"]
    value = N * a.steps / r32["elapsed"]
    # dominant kernel by measured device time
    ph = r32["phases"]
    dom = max((k for k in ph if ph[k][1] > 0), key=lambda k: ph[k][0])
    avg_ms = ph[dom][0] / ph[dom][1]
    bpp = kernel_bytes(dom, 3, 4)
    achieved = bpp * N / (avg_ms * 1e-3) / 1e9
    kname = {"forces": "k_pairs", "density": "k_density", "neighbours": "k_rank_gather",
             "vv1": "k_vv1", "vv2": "k_vv2", "walls_colloids": "k_walls"}.get(dom, dom)
    traffic = measured_traffic(kname, N)
    step_ms = 1e3 * r32["elapsed"] / a.steps
    am = r32["amortized"]
    amortized_value = N * am["steps"] / am["seconds"]
    # flop ceiling of the pair pass: evaluated pairs = neighbours within rcut
    pair_tflops = FLOPS_PER_PAIR * r32["pairs"] / (ph["forces"][0] / ph["forces"][1] * 1e-3) / 1e12 \
        if ph.get("forces", (0, 0))[1] else None
    line = {
        "metric": "particle-steps/s, 3D 1M-particle Oldroyd-B (FP32 vs FP64); % HBM peak; sweep speedup",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pairs, f64 positions/reductions",
        "data": "synthetic (seeded lattice + N(0,1e-3) velocity noise, seeded colloid centres)",
        "config": {
            "workload": "3D walls + 16 magnetic colloids, 100^3 lattice",
            "particles": N,
            "dim": 3, "L": 1.0, "Vwall": 0.25, "tau": 1.0, "eta": 1.0, "c": 20.0,
            "np": 100.0, "kT": 1e-3, "n_colloids": 16, "coll_R": 0.05,
            "F0_repulsion": 10.0, "tau_repulsion": 1.25, "cutoff_magnetic": 0.3,
            "F0_magnetic": 1e-3, "omega_magnetic": 6.283185307179586,
            "dt": 2e-5, "precision": "fp32-form",
        },
        "fp64": None if r64 is None else {
            "value": N * a.steps / r64["elapsed"],
            "ms_per_step": 1e3 * r64["elapsed"] / a.steps,
        },
        "phases_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in ph.items() if v[1]},
        # Verlet lists are rebuilt when a particle has moved half the skin:
        # rebuilds inside the timed window, and a longer window's amortized rate
        "rebuilds_in_window": r32["rebuilds_in_window"],
        "amortized": {"steps": am["steps"], "rebuilds": am["rebuilds"],
                      "ms_per_step": 1e3 * am["seconds"] / am["steps"],
                      "value": amortized_value},
        # the whole step's compulsory bytes (432 B per particle-step) at `value`
        "step_hbm_frac": STEP_BYTES_3D_FP32 * value / (HBM_PEAK_GBS * 1e9),
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": TRAFFIC_FILE if traffic is not None else None,
            "bytes_per_particle": bpp,
            "avg_launch_ms": avg_ms,
            # what bounds it instead (SQ counters of the same kernel and workload,
            # from the committed profile; not measured in this run)
            "valu_busy_profiled": VALU_BUSY.get(kname),
            "valu_busy_source": PAIR_PROFILE if kname in VALU_BUSY else None,
        },
        # what bounds the pair pass instead: FP32 vector arithmetic
        "valu_roofline": None if pair_tflops is None else {
            "kernel": "k_pairs", "bound": "valu",
            "achieved": pair_tflops, "unit": "TFLOP/s",
            "peak": FP32_VALU_PEAK_TFLOPS, "frac": pair_tflops / FP32_VALU_PEAK_TFLOPS,
            "peak_unpacked": FP32_VALU_UNPACKED_TFLOPS,
            "frac_unpacked": pair_tflops / FP32_VALU_UNPACKED_TFLOPS,
            "flops_per_pair": FLOPS_PER_PAIR, "pairs_per_launch": r32["pairs"],
            "list_entries_per_launch": r32["list_entries"],
            "note": "peak counts packed FP32 (v_pk_fma_f32); this build is compiled without "
                    "packed FP32, whose ceiling is peak_unpacked",
        },
        "c2_periodic_64cubed": c2,
        "sweep": sw,
        "sweep_pipeline": pipe,
        "cpu_baseline": None if cpu is None else {
            "value": cpu["main"]["value"],
            "unit": "particle-steps/s",
            "cores": cpu["cores"],
            "kind": "port",
            "sample": f"FP64 OpenMP oracle ({cpu['isa']}, {cpu['cpu']}, nproc "
                      f"{cpu['nproc']}), same workload (N={cpu['main']['N']}): "
                      f"{cpu['main']['steps']} steps after {cpu['warmup']} warm-up steps, "
                      f"median of {cpu['repeats']}",
            "other_configs": {k: {"value": v["value"], "N": v["N"], "steps": v["steps"]}
                              for k, v in cpu.items()
                              if isinstance(v, dict) and k != "main"},
        },
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
