"""Benchmark: contact-solves/s at fc3d residual <= 1e-6 on one MI355X.

Workload (default): the 1M-contact granular pile (400k spheres, mu ~ U[0.1,0.9],
v_free ~ N(0,1) scaled by 0.1 on 10% of the bodies), solved to an FP64
fc3d_compute_error <= 1e-6.  One step = one complete solve from a zero
reaction with the data resident in HBM.  The reported method ("fast") is the
coloured NSGS in the velocity form (local problems from the body velocities,
W's factors H and M^-1 resident) with FP32 sweeps using the undamped
Alart-Curnier Newton local solver, then FP64 sweeps with the reference's
hybrid local solver until the FP64 residual meets the tolerance.  Timed
beside it on the same input: the plain FP64 coloured NSGS (reference
options, rows of W, zero start), the fast method on the rows of W, with
safeguarded Anderson acceleration, and from the learned warm start.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 is refused: the library
runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "contact-solves/s at fc3d residual ≤1e-6: warm-start vs plain GPU NSGS"

WORKLOADS = {
    "pile1m": dict(gen="pile", nbodies=400_000, nc=1_000_000, mu=(0.1, 0.9), seed=5,
                   desc="granular pile, 1M contacts / 400k bodies, tol 1e-6"),
    "pairs50k": dict(gen="pairs", nbodies=20_000, nc=50_000, mu=(0.2, 0.5), seed=2,
                     desc="random pairs, 50k contacts / 20k bodies, tol 1e-6"),
}


# HBM traffic per launch of the reported method's dominant kernel
# (k_sweep_vel<float>), from two rocprofv3 --pmc passes (FETCH_SIZE,
# WRITE_SIZE) over `bench.py --steps 1 --warmup 0 --no-cpu --no-batch`,
# scaled by the readings of a plain 1 GiB device copy (FETCH x2.000,
# WRITE x1.000).  A profiler pass cannot run inside the timed bench, so the
# figure is carried here with its source.  It was measured before the FP32
# sweeps stopped loading and storing the PLI state for local solver 550,
# which saves 8 bytes a contact (about 0.7 MB of a big colour's launch).
PMC_TRAFFIC = {
    "pile1m": {"bytes_per_launch": 44.783e6, "fetch": 36.662e6, "write": 8.121e6,
               "source": "profiles/k_sweep_vel_pile1m_hbm_traffic.md"},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def generate(w):
    from siconos_amd import numerics as N
    f = N.generate_granular_pile if w["gen"] == "pile" else N.generate_random_pairs
    return f(w["nbodies"], w["nc"], w["mu"][0], w["mu"][1], w["seed"])


def time_method(H, params, steps, warmup):
    from siconos_amd import numerics as N
    zero = np.zeros(3 * H.nc)
    stats = None
    for _ in range(warmup):
        H.set_reaction(zero)
        stats = H.solve(params)
    total = 0.0
    dev_ms = 0.0
    for _ in range(steps):
        H.set_reaction(zero)  # host->device reset, outside the timed region
        t0 = time.perf_counter()
        stats = H.solve(params)  # returns after the device finished
        total += time.perf_counter() - t0
        dev_ms += stats.solve_ms
    return total, dev_ms, stats


def cpu_baseline(g, sweeps_needed, sample_sweeps):
    """The oracle (C restatement of the reference fc3d_nsgs, natural order,
    one core) timed on `sample_sweeps` sweeps of the same problem."""
    from oracle import pyoracle as O
    t0 = time.perf_counter()
    info, r, u, it, res = O.nsgs(g.W, g.q, g.mu, tol=1e-300, max_iter=sample_sweeps)
    dt = time.perf_counter() - t0
    per_sweep = dt / max(it, 1)
    value = g.nc / (per_sweep * sweeps_needed)
    return {"value": value, "unit": "contact-solves/s", "cores": 1, "kind": "port",
            "sample": (f"{it} sweeps of the oracle fc3d_nsgs on the same {g.nc}-contact problem "
                       f"({dt:.1f} s, {per_sweep * 1e3:.1f} ms/sweep), extrapolated to the "
                       f"{sweeps_needed} sweeps the same algorithm needs to reach 1e-6; "
                       f"host {platform.processor() or platform.machine()}, "
                       f"{os.cpu_count()} logical cpus")}


def train_warm_start(w, tol):
    """The warm-start network trained on 200k converged contacts of four
    problems of the workload's generator with other seeds."""
    from siconos_amd import numerics as N
    from siconos_amd import warmstart
    if w["gen"] == "pile":
        probs = (N.generate_granular_pile(20_000, 50_000, w["mu"][0], w["mu"][1], 1000 + k)
                 for k in range(4))
    else:
        probs = (N.generate_random_pairs(20_000, 50_000, w["mu"][0], w["mu"][1], 1000 + k)
                 for k in range(4))
    x, t, _ = warmstart.collect(((g.W, g.q, g.mu) for g in probs), tol=tol)
    model, loss = warmstart.train(x, t, seed=1, epochs=30, batch=1024, lr=1e-3)
    return model, loss, len(x)


def batch_leg(steps, warmup, cpu):
    """4096 independent random-pair problems, nc ~ U[8, 512], bodies = nc / 2.5,
    mu ~ U[0.1, 0.9], tol 1e-6 each, solved in one persistent launch; with the
    oracle over the host cores (one problem per thread) beside it."""
    from siconos_amd import numerics as N
    probs = N.generate_batch(4096, 8, 512, 0.1, 0.9, 3)
    B = N.Batch(probs)
    nc = int(B.ntotal)
    out = {"workload": "4096 problems, nc ~ U[8,512], mu ~ U[0.1,0.9], tol 1e-6",
           "contacts": nc}
    for name, prm in (("fp64", N.hip_params(tol=1e-6, max_iter=1000)),
                      ("fp32_fp64", N.hip_params(tol=1e-6, max_iter=1000, precision=1,
                                                 fp32_tol=1e-6, fp32_local_solver=550,
                                                 fp32_local_tol=1e-7))):
        zero = np.zeros(3 * nc)
        for _ in range(warmup):
            B.set_reaction(zero)
            B.solve(prm)
        total, conv = 0.0, 0
        for _ in range(steps):
            B.set_reaction(zero)
            t0 = time.perf_counter()
            s, info, iters, res = B.solve(prm)
            total += time.perf_counter() - t0
            conv += int(s.converged_contacts)
        out[name] = {"value": conv / total, "ms_per_step": total / steps * 1e3,
                     "converged_problems": int(s.converged), "max_residual": s.max_residual,
                     "schedule": int(s.schedule)}
        log(f"[bench] batch {name}: {total / steps * 1e3:.1f} ms, {s.converged}/4096 converged")
    if cpu:
        from oracle import pyoracle as O
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        ok, iters, res, _ = O.solve_many(probs, 1e-6, 1000, threads)
        dt = time.perf_counter() - t0
        conv = int(sum(W.nc for (W, _, _), x in zip(probs, res) if x <= 1e-6))
        out["cpu_baseline"] = {"value": conv / dt, "unit": "contact-solves/s", "cores": threads,
                               "kind": "port",
                               "sample": f"all 4096 problems, oracle fc3d_nsgs one problem per "
                                         f"thread ({dt:.1f} s, {ok} converged)"}
        log(f"[bench] batch CPU baseline: {dt:.1f} s on {threads} threads")
    return out


def fc2d_leg(steps, warmup, cpu):
    """The reference's largest fc2d fixture (fc2d_granularflowonwall_07002_00801,
    7002 contacts, CSparse storage, from data_collection_fc2d_1.c) solved by
    the fc2d HIP path at tol 1e-5 from a zero reaction, resident on the
    device; the oracle's fc2d_nsgs (natural order, one core) beside it."""
    import gzip
    import tempfile
    from siconos_amd import numerics as N
    src = os.path.join(REPO, "tests", "golden", "fixtures",
                       "fc2d_granularflowonwall_07002_00801.dat.gz")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "fc2d.dat")
        with gzip.open(src, "rb") as f, open(path, "wb") as g:
            g.write(f.read())
        p = N.FrictionContactProblem.from_file(path)
        H = N.Fc2dHipProblem(p)
        nc = p.numberOfContacts
        zero = np.zeros(2 * nc)
        for _ in range(warmup):
            H.set_reaction(zero)
            H.solve(max_iter=10000, tol=1e-5)
        total, dev = 0.0, 0.0
        for _ in range(steps):
            H.set_reaction(zero)
            t0 = time.perf_counter()
            s = H.solve(max_iter=10000, tol=1e-5)
            total += time.perf_counter() - t0
            dev += s.solve_ms
        out = {"workload": "fc2d_granularflowonwall_07002_00801 (7002 contacts, CSparse), "
                           "tol 1e-5", "contacts": nc, "colours": H.num_colours,
               "sweeps": s.iterations, "residual": s.residual, "info": s.info,
               "ms_per_step": total / steps * 1e3, "device_ms": dev / steps,
               "value": nc * steps / total if s.info == 0 else 0.0, "unit": "contact-solves/s"}
        log(f"[bench] fc2d: {out['ms_per_step']:.1f} ms, {s.iterations} sweeps, "
            f"residual {s.residual:.2e}")
        if cpu:
            from oracle import pyoracle as O
            o = O.Fc2dProblem(path)
            t0 = time.perf_counter()
            info, r, u, it, res = o.nsgs(tol=1e-5, max_iter=10000, mode=2)
            dt = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": nc / dt if info == 0 else 0.0,
                                   "unit": "contact-solves/s", "cores": 1, "kind": "port",
                                   "sample": f"the whole solve, oracle fc2d_nsgs in natural "
                                             f"order ({dt * 1e3:.0f} ms, {it} sweeps)"}
    return out


def copy_calibration(mib):
    """A plain device-to-device copy of known size: its per-dispatch
    FETCH_SIZE and WRITE_SIZE readings convert the counters to bytes."""
    import torch
    n = mib << 20
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "copies": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="pile1m", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sweeps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 4096-problem batched leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (scale for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus must be 1 "
                                   f"(got {args.gpus})"}))
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from siconos_amd import numerics as N
    w = WORKLOADS[args.workload]
    tol = 1e-6
    log(f"[bench] generating {w['desc']}")
    t0 = time.perf_counter()
    g = generate(w)
    log(f"[bench] generated in {time.perf_counter() - t0:.1f} s; nnzb={g.W.nnzb}")
    t0 = time.perf_counter()
    H = N.HipProblem(g.W, g.q, g.mu)
    setup_s = time.perf_counter() - t0
    log(f"[bench] device problem ready in {setup_s:.1f} s; colours={H.num_colours}")

    # the learned warm start: trained (outside the timed region) on converged
    # contacts of other problems of the same generator
    t0 = time.perf_counter()
    model, mlp_loss, train_nc = train_warm_start(w, tol)
    H.set_model(model)
    log(f"[bench] warm-start network trained on {train_nc} contacts in "
        f"{time.perf_counter() - t0:.1f} s (loss {mlp_loss:.4f})")

    # W's factors for the velocity form (resident, outside the timed region)
    t0 = time.perf_counter()
    H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    log(f"[bench] body factors attached in {time.perf_counter() - t0:.1f} s")

    plain = N.hip_params(tol=tol, max_iter=3000, sweeps_per_launch=16)
    # FP32 local tolerance 1e-7 instead of the default 1e-6: the FP32 iterate
    # then needs 2-4 FP64 sweeps instead of 26-36 (1e-7 measured best of
    # 1e-3 .. 1e-8 on four piles)
    fast_kw = dict(tol=tol, max_iter=3000, sweeps_per_launch=16, precision=1, fp32_tol=1e-6,
                   fp32_local_solver=550, fp32_local_tol=1e-7, warm_start=1)
    legs = {
        "fast": N.hip_params(velocity_form=1, **fast_kw),
        "fast_rows_of_W": N.hip_params(**fast_kw),
        "fast_with_anderson": N.hip_params(velocity_form=1, accel_depth=3, accel_every=2,
                                           **fast_kw),
        "fast_with_learned_start": N.hip_params(**dict(fast_kw, warm_start=2)),
    }
    log("[bench] timing plain FP64 coloured NSGS")
    tp, devp, sp = time_method(H, plain, args.steps, args.warmup)
    log(f"[bench] plain: {tp / args.steps * 1e3:.1f} ms/solve, sweeps={sp.iterations}, "
        f"residual={sp.residual:.3e}")
    res = {}
    for name, prm in legs.items():
        t, dev, st = time_method(H, prm, args.steps, args.warmup)
        res[name] = (t, st)
        log(f"[bench] {name}: {t / args.steps * 1e3:.1f} ms/solve, sweeps={st.iterations} "
            f"(fp32 {st.fp32_sweeps}, anderson {st.accel_steps}/{st.accel_rejects} rejected), "
            f"start used {st.warm_start_used}, residual={st.residual:.3e}")
    tf, sf = res["fast"]
    ok_plain = sp.residual <= tol
    ok_fast = sf.residual <= tol

    # roofline of the dominant kernel, the FP32 velocity-form colour pass
    # k_sweep_vel<float>: one instrumented fast solve; every launch carries a
    # pair of HIP events stamped with the dispatch's own start and end on the
    # library's stream; the FP32 launches are summed apart, and the tail
    # launches (k_sweep_vel_tail, the last small colours in one workgroup)
    # are taken out of the kernel's figures and reported beside them
    H.set_reaction(np.zeros(3 * H.nc))
    si = H.solve(N.hip_params(**dict(fast_kw, velocity_form=1, instrument=1)))
    kms = si.sweep_kernel_ms_fp32 - si.sweep_tail_ms_fp32
    kby = si.sweep_kernel_bytes_fp32 - si.sweep_tail_bytes_fp32
    nl = max(si.sweep_kernel_launches_fp32 - si.sweep_tail_launches_fp32, 1)
    achieved = kby / (kms * 1e-3) / 1e9
    pmc = PMC_TRAFFIC.get(args.workload)
    ntl = max(si.sweep_tail_launches_fp32, 1)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0,
                "traffic": pmc["bytes_per_launch"] if pmc else None,
                "traffic_source": pmc["source"] if pmc else None,
                "kernel": "k_sweep_vel<float> (one colour pass, velocity form)",
                "launches": int(nl),
                "launch_avg_us": kms * 1e3 / nl,
                "bytes_per_launch_avg": kby / nl,
                "sweep_fp32_us": si.sweep_kernel_ms_fp32 * 1e3 / max(si.fp32_sweeps, 1),
                "tail": {"kernel": "k_sweep_vel_tail<float>",
                         "launches": int(si.sweep_tail_launches_fp32),
                         "launch_avg_us": si.sweep_tail_ms_fp32 * 1e3 / ntl,
                         "bytes_per_launch_avg": si.sweep_tail_bytes_fp32 / ntl}}
    log(f"[bench] sweep kernel: {achieved:.1f} GB/s over {nl} FP32 launches, "
        f"{roofline['launch_avg_us']:.2f} us each; {si.sweep_tail_launches_fp32} tail launches, "
        f"{roofline['tail']['launch_avg_us']:.2f} us each")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline sample ({args.cpu_sweeps} sweeps)")
        cpu = cpu_baseline(g, sp.iterations, args.cpu_sweeps)

    batch = None if args.no_batch else batch_leg(args.steps, args.warmup, not args.no_cpu)
    fc2d = fc2d_leg(args.steps, args.warmup, not args.no_cpu)

    # W = H M^-1 H^T and q = H v_free for the same pile, assembled on the GPU
    asm_ms = []
    for _ in range(3):
        Wa, qa, ms = N.assemble(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
        asm_ms.append(ms)
    same = bool(np.array_equal(Wa.col_idx, g.W.col_idx) and
                np.allclose(Wa.values, g.W.values, rtol=0, atol=1e-13))
    t_asm = min(asm_ms)
    assembly = {"device_ms": t_asm, "blocks": int(Wa.nnzb),
                "blocks_per_s": Wa.nnzb / (t_asm * 1e-3), "matches_host_assembly": same}
    log(f"[bench] GPU assembly of W: {t_asm:.2f} ms for {Wa.nnzb} blocks (matches host: {same})")

    nc = g.nc
    value = nc * args.steps / tf if ok_fast else 0.0
    plain_value = nc * args.steps / tp if ok_plain else 0.0

    def leg(name):
        t, st = res[name]
        d = {"value": nc * args.steps / t if st.residual <= tol else 0.0,
             "ms_per_step": t / args.steps * 1e3, "sweeps": st.iterations,
             "fp32_sweeps": st.fp32_sweeps, "residual": st.residual}
        if st.accel_steps:
            d.update(anderson_steps=st.accel_steps, anderson_rejected=st.accel_rejects)
        return d

    learned = leg("fast_with_learned_start")
    tn, sn = res["fast_with_learned_start"]
    learned.update(
        warm_start_used={1: "zero (fallback)", 2: "learned"}.get(sn.warm_start_used),
        residual_after_first_sweep={"zero": sn.start_residual_fallback,
                                    "learned": sn.start_residual_learned})
    out = {
        "metric": METRIC, "value": value, "unit": "contact-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": tf / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 sweeps + f64 residual/refinement", "data": "synthetic",
        "config": {"workload": w["desc"], "contacts": nc, "bodies": w["nbodies"],
                   "nnz_blocks": int(g.W.nnzb), "colours": H.num_colours, "tol": tol,
                   "method": "coloured NSGS from a zero start in the velocity form: FP32 "
                             "sweeps with the undamped AC Newton local solver (local "
                             "tolerance 1e-7), then FP64 sweeps with the reference hybrid "
                             "local solver to the FP64 residual"},
        "plain_fp64": {"value": plain_value, "ms_per_step": tp / args.steps * 1e3,
                       "sweeps": sp.iterations, "residual": sp.residual},
        "fast": leg("fast"),
        "fast_rows_of_W": leg("fast_rows_of_W"),
        "fast_with_anderson": leg("fast_with_anderson"),
        "fast_with_learned_start": learned,
        "warm_start_training": {"contacts": train_nc, "final_loss": mlp_loss},
        "speedup_vs_plain": (value / plain_value) if plain_value else None,
        "roofline": roofline, "cpu_baseline": cpu, "batch": batch, "assembly": assembly,
        "fc2d": fc2d,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
