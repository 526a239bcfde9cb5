"""Benchmark: explicit internal-force step on the 128^3 two-block HEX8 cube.

One step = one central-difference step of the whole mesh on the device:
fused element kernel (F -> Neo-Hookean stress -> B^T.sigma) over all blocks
with the nodal force sum and update (mixed precision: tile-local assembly in
LDS, then the nodes shared between tiles; FP64: element force slots, then a
CSR node gather), and the top-face reaction sum.  Inputs are resident in HBM
before timing starts.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--precision mixed|fp64]
                    [--workload cube128_2blk] [--cpu-steps 40] [--no-cpu] [--uq NE,NA]
                    [--no-fp64] [--fp64-steps 30] [--no-ensemble]

Prints one JSON line.  `value` is element-steps/s of the timed K steps.
`roofline` covers the dominant (element) kernel: its algorithmic bytes per
launch (algorithmic_bytes below) over its average launch duration from HIP
events that every 10th launch of the timed steps records at its own start
and end (hipExtLaunchKernelGGL).  `cpu_baseline` times the FP64 CPU oracle (OpenMP, deterministic
gather) on a bounded sample of the same workload.

The same line carries an `fp64` object (the plain FP64 path on the same
workload, --fp64-steps timed steps) and an `ensemble` object: the batched
ensemble of 256 samples on the 32^3 cube plus 64 on cubes with edge counts
drawn from {16, 24, 32, 40, 48} (per-sample E, nu, rho lognormal around
200 GPa / 0.3 / 7800, top velocity uniform in [0.5, 2] m/s, 400 steps),
replayed from a captured HIP graph: samples/s of whole runs, the mean-QoI
error of the mixed ensemble against the FP64 ensemble, and plain
one-at-a-time solves of the model path (create, lumped mass and dt, 400
steps, QoI) timed on two samples per edge count.

Short timed runs: a GPU coming out of idle runs its first steps slower while
its clocks rise (20-step chunks of the headline step after a 5-step warmup
from idle: 0.276, 0.254, then 0.245 ms/step on one box), so a run of 20 timed
steps reads a few percent slower than one of 500.  The timed steps continue
from the warmed state without a host reset, which would idle the device
between them.

A `surrogate` object covers the learned control variate on the 24^3 cube:
512 plain FP64 solves (batched) train the 4-256-256-256-64 network on the
device; 8192 draws are estimated with 5% plain solves for bias correction;
samples/s of the whole estimator, its error against plain solves of all
draws, and plain one-at-a-time solves timed beside it.

--uq NE,NA adds an off-nominal ensemble (NE exact samples with the bulk and
shear moduli of every block scaled by 1 + 0.2 x_k, x_k evenly spaced in
[-1, 1], and NA approximate samples with closure coefficients interpolating
between neighbouring exact samples) and reports the closure kernel in a "uq"
object; the headline line is the default run without it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# MI355X vector peaks (TFLOP/s): FP32 (packed), FP64
VALU_PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6}
# FP operations per element-step of the reference's algorithm (FMA = 2):
# 1.23k per Gauss point (Jacobians, inverses, F, polar decomposition,
# Neo-Hookean, dN/dx, force contraction) x 8
COUNTED_FLOPS_PER_ELEMENT_STEP = 9840
# The FP64 element kernel is bound by VALU issue, not HBM: its VALU
# instructions per element (8 lanes per element; rocprofv3 SQ_INSTS_VALU on
# the 128^3 cube, 1,897 per lane) and the issue peak those are measured
# against: 1024 SIMDs, one wave64 FP64 instruction per 4 cycles, 2.4 GHz
FP64_VALU_INSTS_PER_ELEMENT = 8 * 1897
VALU_ISSUE_PEAK_PER_S = 1024 * 2.4e9 / 4
# the same for the tiled mixed element kernel (one lane per element; SQ_INSTS_VALU
# per wave of 64 elements on the 128^3 cube, 512-element tiles)
MIXED_VALU_INSTS_PER_ELEMENT = 2388


MAX_BLOCKS_PER_LAUNCH = 8  # element blocks sharing one element-kernel launch
TIMING_EVERY = 10  # kernel launches timed by HIP events: every 10th of each kernel


def algorithmic_bytes(prob, s_bytes, assembly):
    """Entry-point byte model (s = bytes per integration-point word), per step:
    E*(32 + 120*s) + N*152 -- int32 connectivity, F and sigma written once per
    Gauss point, and per node X, u, v, m read and u, v, f_int written once.

    Split over the launches that do the work:
      slot path, element kernel per group of blocks: E_g*(32 + 120*s) + N_g*48
        (X and u of the N_g distinct nodes of the group); the node kernel
        carries the remaining N*104;
      tiled path, element kernel: E*(32 + 120*s) + N_int*152 + N_bnd*48 (it
        completes the step of the N_int tile-interior nodes and reads X, u of
        the others); the boundary kernel carries N_bnd*104."""
    ranges = prob.block_ranges()
    per_launch = []
    if assembly["tiled"]:
        if len(ranges) > MAX_BLOCKS_PER_LAUNCH:
            raise SystemExit("bench.py: the tiled byte split assumes one element launch (at most 8 blocks)")
        n_int, n_bnd = assembly["interior_nodes"], assembly["boundary_nodes"]
        per_launch.append(prob.num_elem * (32 + 120 * s_bytes) + n_int * 152 + n_bnd * 48)
    else:
        for g in range(0, len(ranges), MAX_BLOCKS_PER_LAUNCH):
            e0, e1 = ranges[g][0], ranges[min(g + MAX_BLOCKS_PER_LAUNCH, len(ranges)) - 1][1]
            ng = len(np.unique(prob.conn[e0:e1]))
            per_launch.append((e1 - e0) * (32 + 120 * s_bytes) + ng * 48)
    step = prob.num_elem * (32 + 120 * s_bytes) + prob.num_nodes * 152
    return per_launch, step


def closure_bytes(num_nodes, ne, na):
    """Algorithmic bytes of one closure launch: per node the nominal and NE
    exact forces (24 B each), flags and mass (9 B), and per approximate
    sample u, v read and f, u, v written (120 B)."""
    return num_nodes * (33 + 24 * ne + 120 * na)


def uq_ensemble(prob, ne, na):
    """Moduli [ne, blocks] and closure coefficients [na, ne + 1] of --uq."""
    K0 = np.array([b.material.bulk_modulus for b in prob.blocks()])
    G0 = np.array([b.material.shear_modulus for b in prob.blocks()])
    x = np.linspace(-1.0, 1.0, ne) if ne > 1 else np.zeros(ne)
    K = K0[None, :] * (1 + 0.2 * x[:, None])
    G = G0[None, :] * (1 + 0.2 * x[:, None])
    c = np.zeros((na, ne + 1))
    for s in range(na):
        if ne == 0:
            c[s, 0] = 1.0
            continue
        p = s * (ne - 1) / max(na - 1, 1)  # position among the exact samples
        k = min(int(p), ne - 1)
        w = p - k
        c[s, 1 + k] += 1.0 - w
        if w > 0:
            c[s, 2 + k] += w
    return K, G, c


def load_traffic(workload, precision):
    """HBM bytes per element-kernel launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), as
    committed under profiles/; None when not measured for this workload."""
    path = os.path.join(ROOT, "profiles", f"element_{precision}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prob, steps, runs=5):
    """FP64 CPU oracle, OpenMP over elements + deterministic CSR node gather:
    median of `runs` runs of steps/runs steps after one warm-up step, and one
    step of the serial (single-thread) oracle."""
    import statistics

    from oracle import oracle as O  # the checker: only this leg of bench.py touches it
    from tests.helpers import oracle_model_from_problem

    threads = min(16, os.cpu_count() or 1)
    om = oracle_model_from_problem(prob)
    dt = prob.dt_factor * om.critical_dt()
    run = O.ExplicitRun(om, prob.num_steps * dt, prob.num_steps, 0, threads=threads)
    run.advance(1)  # warm-up (page-in, OpenMP pool)
    per = max(steps // runs, 1)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        run.advance(per)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    serial = O.ExplicitRun(om, prob.num_steps * dt, prob.num_steps, 0, threads=1)
    t0 = time.perf_counter()
    serial.advance(1)
    ts = time.perf_counter() - t0
    # NimbleSM's own element and material code (oracle/_ref, compiled from the
    # reference's sources where they were present), serial as the reference's
    # CPU path runs: F, stress and forces of a sample of the workload's
    # elements, gathered from the current state
    reference = None
    R = O.ref_lib()
    if R is not None:
        blk = prob.blocks()[0]
        S = min(262144, len(blk.conn))
        conn = np.asarray(blk.conn[:S], np.int64)
        Xg = np.ascontiguousarray(prob.X[conn])
        ug = np.ascontiguousarray(run.u[conn])
        F, sig, f = np.zeros((S, 8, 9)), np.zeros((S, 8, 6)), np.zeros((S, 8, 3))
        K, G = blk.material.bulk_modulus, blk.material.shear_modulus
        rt = []
        for _ in range(3):
            t0 = time.perf_counter()
            R.ref_deformation_gradients(S, Xg, ug, F)
            R.ref_stress(8 * S, K, G, F, sig)
            R.ref_nodal_forces(S, Xg, ug, sig, f)
            rt.append(time.perf_counter() - t0)
        reference = {
            "value": S / statistics.median(rt), "unit": "element-evaluations/s", "cores": 1, "kind": "reference",
            "sample": f"F, Neo-Hookean stress and nodal forces of NimbleSM's own element/material code on {S} "
                      f"pre-gathered elements of {prob.name}'s first block, serial, median of 3 runs "
                      f"({sum(rt):.1f} s in all); no gather, scatter or node update",
            "runs_s": rt,
        }
    return {
        "value": prob.num_elem * per / med,
        "unit": "element-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of {runs} runs of {per} FP64 steps of {prob.name} ({prob.num_elem} elements) after 1 "
                  f"warm-up step, {threads} OpenMP threads ({sum(times):.1f} s in all); serial: 1 step on one "
                  f"thread ({ts:.1f} s)",
        "runs_s": times,
        "serial_value": prob.num_elem / ts,
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "reference_code": reference,
    }


def fp64_leg(N, workload, steps, warmup):
    """The plain FP64 path on the same workload: a few timed steps (the
    element kernel is VALU-bound, so a short run gives its rate)."""
    import torch

    prob = N.workload(workload).build()
    model = prob.make_model("fp64")
    integ = prob.make_integrator(model, num_steps=warmup + steps)
    integ.Advance(warmup)
    model.EnableKernelTiming(True)
    model.Synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    integ.Advance(steps)
    model.Synchronize()
    elapsed = time.perf_counter() - t0
    el_ms, _ = model.KernelTime("element")
    nd_ms, _ = model.KernelTime("node")
    model.EnableKernelTiming(False)
    per_launch, step_bytes = algorithmic_bytes(prob, 8, model.AssemblyInfo())
    ms = elapsed * 1e3 / steps
    model.close()
    return {
        "steps": steps, "warmup": warmup, "ms_per_step": ms, "value": prob.num_elem * steps / elapsed,
        "unit": "element-steps/s", "element_ms_avg": el_ms, "node_ms_avg": nd_ms,
        "element_alg_bytes_per_launch": float(np.mean(per_launch)),
        "element_frac": float(np.mean(per_launch)) / (el_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if el_ms > 0 else 0.0,
        "step_frac": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        # its binding limit: VALU issue (counter-measured instructions per
        # element over the launch time, against the nominal-clock issue peak)
        "element_valu_issue_frac": (FP64_VALU_INSTS_PER_ELEMENT * prob.num_elem / 64 / (el_ms * 1e-3)
                                    / VALU_ISSUE_PEAK_PER_S) if el_ms > 0 else 0.0,
    }


def host_seam_leg(N, workload, calls=10):
    """The internal-force seam driven with host buffers, as a caller that
    keeps its own integrator loop on the host would drive it
    (ModelDataBase::ComputeInternalForce takes host views of u and f): per
    call, u goes to the device, the force evaluation runs, f comes back.
    Reported beside the same evaluation on device-resident data."""
    prob = N.workload(workload).build()
    rng = np.random.default_rng(3)
    u = 1e-4 / prob.n * rng.uniform(-1.0, 1.0, (prob.num_nodes, 3))
    out = {"calls": calls, "bytes_per_call": 2 * u.nbytes}
    for precision in ("mixed", "fp64"):
        model = prob.make_model(precision)
        model.ComputeLumpedMass()
        model.SetVectorNodeData("displacement", u)
        model.ComputeInternalForce()
        f = model.GetVectorNodeData("internal_force")
        t0 = time.perf_counter()
        for _ in range(calls):
            model.SetVectorNodeData("displacement", u)
            model.ComputeInternalForce()
            f = model.GetVectorNodeData("internal_force")
        seam = (time.perf_counter() - t0) / calls
        model.Synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            model.ComputeInternalForce()
        model.Synchronize()
        dev = (time.perf_counter() - t0) / calls
        assert np.isfinite(f).all()
        model.close()
        out[precision] = {"ms_per_call": seam * 1e3, "device_ms_per_call": dev * 1e3,
                          "element_evaluations_per_s": prob.num_elem / seam,
                          "device_element_evaluations_per_s": prob.num_elem / dev}
    return out


def ensemble_leg(N, num_steps=400, runs=3, plain_per_size=2):
    """Batched ensemble of 256 samples on the 32^3 cube and 64 on uneven
    cubes (edge counts from {16, 24, 32, 40, 48}), num_steps steps each:
    samples/s of whole runs (graph replays, mixed precision), the mean-QoI
    error against the same ensemble in FP64, and plain one-at-a-time solves of
    the model path timed on a stratified subset (plain_per_size samples per
    edge count, scaled by the counts)."""
    import torch

    from nimblesm_amd.ensemble import NOMINAL, EnsembleProblem, config2_draws

    d = config2_draws()
    out = {"workload": "256 x 32^3 + 64 x n^3, n uniform in {16,24,32,40,48}", "samples": len(d),
           "steps": num_steps, "elements_per_step": int(np.sum(np.asarray(d.n) ** 3))}
    res = {}
    for precision in ("mixed", "fp64"):
        t_setup = time.perf_counter()
        prob = EnsembleProblem(d, num_steps=num_steps, precision=precision).build()
        t_setup = time.perf_counter() - t_setup
        q = prob.run(use_graph=True)  # capture + first replay
        torch.cuda.synchronize()
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            q = prob.run(use_graph=True)  # replay + QoI readout
            times.append(time.perf_counter() - t0)
        info = prob.ens.Info()
        # per-kernel durations from one eager run with HIP events
        prob.ens.EnableKernelTiming(True)
        prob.run(use_graph=False)
        el_ms, _ = prob.ens.KernelTime("element")
        nd_ms, _ = prob.ens.KernelTime("node")
        prob.ens.EnableKernelTiming(False)
        prob.close()
        run_s = float(np.median(times))
        res[precision] = q
        out[precision] = {
            "run_s": run_s, "samples_per_s": len(d) / run_s,
            "element_steps_per_s": out["elements_per_step"] * num_steps / run_s,
            "element_ms_avg": el_ms, "node_ms_avg": nd_ms, "tile_elems": info["tile_elems"],
            "work_items_per_step": info["work_items"], "setup_s": t_setup, "graph": True,
        }
    ref = res["fp64"]
    err = np.linalg.norm(res["mixed"].mean(0) - ref.mean(0)) / np.linalg.norm(ref.mean(0))
    per = np.linalg.norm(res["mixed"] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    out["qoi_mean_rel_err_mixed_vs_fp64"] = float(err)
    out["qoi_worst_sample_rel_err_mixed_vs_fp64"] = float(per.max())

    # plain one-at-a-time solves: model creation (tiles), lumped mass and dt,
    # num_steps steps, QoI sampling -- and the steps alone
    plain = {}
    for precision in ("mixed", "fp64"):
        solve_total = steps_total = 0.0
        detail = {}
        for n in sorted(set(int(x) for x in d.n)):
            # the model's one-time tile autotuning of this shape (cached for
            # the process) is not part of a solve
            N.CubeProblem(n=n, materials=[NOMINAL], num_steps=1).make_model(precision).close()
            idx = np.nonzero(d.n == n)[0][:plain_per_size]
            ts, tt = [], []
            for i in idx:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                prob1 = N.CubeProblem(n=n, materials=[(d.E[i], d.nu[i], d.rho[i])], num_steps=num_steps,
                                      v_top=float(d.v_top[i]), name="plain")
                m = prob1.make_model(precision)
                integ = prob1.make_integrator(m)
                m.Synchronize()
                t1 = time.perf_counter()
                integ.Integrate()
                m.Synchronize()
                t2 = time.perf_counter()
                m.SampleReactionHistory(num_steps, 64)
                t3 = time.perf_counter()
                m.close()
                ts.append(t3 - t0)
                tt.append(t2 - t1)
            cnt = int(np.sum(d.n == n))
            solve_total += cnt * float(np.mean(ts))
            steps_total += cnt * float(np.mean(tt))
            detail[str(n)] = {"count": cnt, "solve_ms": 1e3 * float(np.mean(ts)), "steps_ms": 1e3 * float(np.mean(tt))}
        plain[precision] = {"samples_per_s": len(d) / solve_total, "steps_only_samples_per_s": len(d) / steps_total,
                            "per_edge_count": detail, "timed_per_edge_count": plain_per_size}
    out["plain"] = plain
    out["speedup_vs_plain"] = out["mixed"]["samples_per_s"] / plain["mixed"]["samples_per_s"]
    out["speedup_vs_plain_steps_only"] = out["mixed"]["samples_per_s"] / plain["mixed"]["steps_only_samples_per_s"]
    out["fp64_speedup_vs_plain_fp64_steps_only"] = (out["fp64"]["samples_per_s"] /
                                                    plain["fp64"]["steps_only_samples_per_s"])
    return out


def surrogate_leg(N, num_steps=400, n_train=512, n_eval=8192, fraction=0.05, iterations=3000):
    """Learned control variate on the 24^3 cube: n_train plain FP64 solves
    (batched) train the network on the device; n_eval draws are estimated
    with `fraction` of them solved plainly (batched, mixed precision) for
    bias correction and error judging.  samples/s = n_eval / wall time of the
    whole estimator (training solves, training, inference, correction
    solves); against plain one-at-a-time solves of the model path timed on 4
    draws.  The error is against the mean of plain solves of all n_eval draws
    (batched, untimed)."""
    import torch

    from nimblesm_amd.ensemble import NOMINAL, EnsembleProblem, draw_parameters
    from nimblesm_amd.surrogate import LinearSurrogate, MLPSurrogate, control_variate

    train = draw_parameters(n_train, seed=101, n=24)
    ev = draw_parameters(n_eval, seed=202, n=24)

    def solve(d, precision):
        prob = EnsembleProblem(d, num_steps=num_steps, precision=precision).build()
        q = prob.run()
        prob.close()
        return q

    # one-time autotuning of the tile size for this shape (cached for the process)
    t0 = time.perf_counter()
    solve(ev.subset(np.arange(8)), "mixed")
    autotune_s = time.perf_counter() - t0
    truth_all = solve(ev, "mixed")
    truth = truth_all.mean(axis=0)
    torch.cuda.synchronize()

    phases = {}
    t = time.perf_counter()
    y_train = solve(train, "fp64")
    phases["training_solves_s"] = time.perf_counter() - t
    t = time.perf_counter()
    mlp = MLPSurrogate(seed=0).fit(train, y_train, iterations=iterations)
    phases["training_s"] = time.perf_counter() - t
    t = time.perf_counter()
    pred = mlp.predict(ev)
    phases["inference_s"] = time.perf_counter() - t
    t = time.perf_counter()
    est = control_variate(mlp, ev, lambda d: solve(d, "mixed"), fraction=fraction, seed=7, pred=pred)
    phases["correction_solves_s"] = time.perf_counter() - t
    total = sum(phases.values())
    err = float(np.linalg.norm(est.mean - truth) / np.linalg.norm(truth))

    # equal cost comparisons: the same number of plain solves
    rng = np.random.default_rng(8)
    n_plain = n_train + est.num_plain
    mc = rng.choice(n_eval, size=n_plain, replace=False)
    err_mc = float(np.linalg.norm(truth_all[mc].mean(0) - truth) / np.linalg.norm(truth))
    idx = np.sort(np.random.default_rng(7).choice(n_eval, size=est.num_plain, replace=False))
    lin = LinearSurrogate().fit(train, y_train)
    est_lin = control_variate(lin, ev, lambda d: truth_all[idx], plain_idx=idx)
    err_lin = float(np.linalg.norm(est_lin.mean - truth) / np.linalg.norm(truth))

    # plain one-at-a-time solves of the model path (24^3), after the model's
    # one-time tile autotuning of the shape
    plain = {}
    for precision in ("mixed", "fp64"):
        N.CubeProblem(n=24, materials=[NOMINAL], num_steps=1).make_model(precision).close()
        ts, tt = [], []
        for i in range(4):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            prob1 = N.CubeProblem(n=24, materials=[(ev.E[i], ev.nu[i], ev.rho[i])], num_steps=num_steps,
                                  v_top=float(ev.v_top[i]), name="plain")
            m = prob1.make_model(precision)
            integ = prob1.make_integrator(m)
            m.Synchronize()
            t1 = time.perf_counter()
            integ.Integrate()
            m.Synchronize()
            t2 = time.perf_counter()
            m.SampleReactionHistory(num_steps, 64)
            ts.append(time.perf_counter() - t0)
            tt.append(t2 - t1)
            m.close()
        plain[precision] = {"solve_ms": 1e3 * float(np.median(ts)), "steps_ms": 1e3 * float(np.median(tt)),
                            "samples_per_s": 1.0 / float(np.median(ts)),
                            "steps_only_samples_per_s": 1.0 / float(np.median(tt))}
    sps = n_eval / total
    mlp.close()
    return {
        "workload": f"24^3 cube, {n_train} FP64 training solves, {n_eval} draws, {fraction:.0%} plain solves",
        "network": "4-256-256-256-64 GELU, BF16 weights/activations, FP32 accumulation (MFMA 16x16x32)",
        "iterations": iterations, "final_loss": float(mlp.loss[-1]), "initial_loss": float(mlp.loss[0]),
        "samples_per_s": sps, "wall_s": total, "phases": phases, "autotune_s": autotune_s,
        "qoi_mean_rel_err": err, "judged_rel_err": est.est_rel_err, "heldout_rel_err": est.heldout_rel_err,
        "heldout_mean_rel_err": est.heldout_mean_rel_err,
        "fallback": est.fallback, "plain_solves": n_plain,
        "equal_cost_plain_mc_rel_err": err_mc, "equal_cost_linear_cv_rel_err": err_lin,
        "plain_one_at_a_time": plain,
        "speedup_vs_plain": sps / plain["mixed"]["samples_per_s"],
        "speedup_vs_plain_steps_only": sps / plain["mixed"]["steps_only_samples_per_s"],
        "speedup_vs_plain_fp64_steps_only": sps / plain["fp64"]["steps_only_samples_per_s"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--workload", default="cube128_2blk")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--uq", default="", metavar="NE,NA", help="add NE exact and NA approximate UQ samples")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg")
    ap.add_argument("--fp64-steps", type=int, default=30)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-control-variate leg")
    ap.add_argument("--no-seam", action="store_true", help="skip the host-buffer internal-force seam leg")
    ap.add_argument("--only", choices=["ensemble", "surrogate"], default=None,
                    help="run only that leg and print its object (for profiling)")
    ap.add_argument("--surrogate-iterations", type=int, default=3000)
    ap.add_argument("--hbm-probe", type=int, default=0, metavar="MIB",
                    help="only time device fills and copies of MIB MiB (the practical HBM bandwidth)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs (domain decomposition with a "
              "shared-node force exchange) are not part of it. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import nimblesm_amd as N

    if not torch.cuda.is_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 5}), flush=True)
        return

    if args.hbm_probe:
        # practical ceilings under the nominal 8 TB/s: a write-only stream
        # (fill), a read-only stream (sum) and a read+write stream (copy),
        # each the median of 20 timed repetitions after 3 untimed ones
        nbytes = args.hbm_probe << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        res = {}
        for name, fn, moved in (("fill", lambda: a.fill_(2.0), nbytes), ("read", lambda: a.sum(), nbytes),
                                ("copy", lambda: b.copy_(a), 2 * nbytes)):
            ts = []
            for i in range(23):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                fn()
                e1.record()
                e1.synchronize()
                if i >= 3:
                    ts.append(e0.elapsed_time(e1) * 1e-3)
            res[name + "_GBs"] = moved / float(np.median(ts)) / 1e9
        print(json.dumps({"hbm_probe_bytes": nbytes, **res, "peak_GBs": HBM_PEAK_GBS}), flush=True)
        return

    if args.only == "ensemble":
        print(json.dumps({"ensemble": ensemble_leg(N)}), flush=True)
        return
    if args.only == "surrogate":
        print(json.dumps({"surrogate": surrogate_leg(N, iterations=args.surrogate_iterations)}), flush=True)
        return

    prob = N.workload(args.workload).build()
    s_bytes = 4 if args.precision == "mixed" else 8
    total = args.warmup + args.steps
    model = prob.make_model(args.precision)
    ne = na = 0
    if args.uq:
        ne, na = (int(x) for x in args.uq.split(","))
        uq = uq_ensemble(prob, ne, na)
        model.SetUqSamples(*uq)

    # warmup steps, then the timed steps continue from that state (no host
    # reset between them: it would leave the device idle)
    integ = prob.make_integrator(model, num_steps=max(total, 1))
    integ.Advance(args.warmup)

    # every TIMING_EVERY-th launch of each kernel inside the timed region
    # records a start and a stop event (timing every launch would cost a few
    # microseconds per launch of the measured throughput)
    model.EnableKernelTiming(True, every=TIMING_EVERY)
    torch.cuda.synchronize()
    model.Synchronize()
    t0 = time.perf_counter()
    integ.Advance(args.steps)
    model.Synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    el_ms, el_n = model.KernelTime("element")
    nd_ms, nd_n = model.KernelTime("node")
    rx_ms, rx_n = model.KernelTime("reaction")
    cl_ms, cl_n = model.KernelTime("closure")
    model.EnableKernelTiming(False)

    assembly = model.AssemblyInfo()
    per_launch, step_bytes = algorithmic_bytes(prob, s_bytes, assembly)
    launches_per_step = len(per_launch)
    # mean algorithmic bytes per launch over the launches of one step
    bytes_per_launch = float(np.mean(per_launch))
    achieved = bytes_per_launch / (el_ms * 1e-3) / 1e9 if el_ms > 0 else 0.0
    traffic = load_traffic(args.workload, args.precision)
    ms_per_step = elapsed * 1e3 / args.steps
    value = prob.num_elem * args.steps / elapsed
    R = model.GetReactionHistory(args.warmup, args.steps)
    # compulsory floor: connectivity, X, u, v, m read and u, v written once
    # (no integration-point state, no internal-force write)
    floor_bytes = prob.num_elem * 32 + prob.num_nodes * 128

    out = {
        "metric": "element-steps/s + % HBM roofline",
        "value": value,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded jittered cube, SplitMix64)",
        "config": {
            "workload": args.workload,
            "elements": prob.num_elem,
            "nodes": prob.num_nodes,
            "blocks": len(prob.materials),
            "precision": args.precision,
            "assembly": "tiles" if assembly["tiled"] else "slots",
            "steps": args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "kernels": {
            "element_ms_avg": el_ms, "element_launches_timed": el_n, "timed_every": TIMING_EVERY,
            "element_launches_per_step": launches_per_step,
            "element_alg_bytes_per_launch": bytes_per_launch,
            "node_ms_avg": nd_ms, "node_launches_timed": nd_n,
            "tile_elems": model.TileElems() if assembly["tiled"] else 0,
            "tiles": assembly["num_tiles"], "interior_nodes": assembly["interior_nodes"],
            "boundary_nodes": assembly["boundary_nodes"],
            "reaction_ms_avg": rx_ms, "reaction_launches": rx_n,
        },
        "step_roofline": {
            "alg_bytes_per_step": step_bytes,
            "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "reaction_final": float(R[-1]),
        "secondary": {
            "counted_flops_per_element_step": COUNTED_FLOPS_PER_ELEMENT_STEP,
            "valu_peak_tflops": VALU_PEAK_TFLOPS[args.precision],
            "valu_frac": COUNTED_FLOPS_PER_ELEMENT_STEP * value / (VALU_PEAK_TFLOPS[args.precision] * 1e12),
            "compulsory_floor_bytes_per_element": floor_bytes / prob.num_elem,
            "compulsory_floor_frac": floor_bytes / prob.num_elem * value / (HBM_PEAK_GBS * 1e9),
            "element_measured_hbm_GBs": traffic / (el_ms * 1e-3) / 1e9 if traffic and el_ms > 0 else None,
            # counter-measured VALU instructions over the launch time against the
            # 2.4 GHz issue peak (packed FP32 is not used, so valu_frac's packed
            # peak overstates what the element arithmetic could reach)
            "element_valu_issue_frac": (MIXED_VALU_INSTS_PER_ELEMENT * prob.num_elem / 64 / (el_ms * 1e-3)
                                        / VALU_ISSUE_PEAK_PER_S)
            if args.precision == "mixed" and assembly["tiled"] and el_ms > 0 else None,
        },
    }
    if args.uq:
        cb = closure_bytes(prob.num_nodes, ne, na)
        out["config"]["uq"] = {"exact_samples": ne, "approx_samples": na}
        out["uq"] = {
            "trajectories": 1 + ne + na,
            "trajectory_steps_per_s": (1 + ne + na) * args.steps / elapsed,
            "closure_ms_avg": cl_ms, "closure_launches": cl_n, "closure_alg_bytes_per_launch": cb,
            "closure_achieved_GBs": cb / (cl_ms * 1e-3) / 1e9 if cl_ms > 0 else 0.0,
            "closure_frac": cb / (cl_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if cl_ms > 0 else 0.0,
        }
    model.close()
    if not args.no_fp64 and args.precision == "mixed" and not args.uq:
        out["fp64"] = fp64_leg(N, args.workload, args.fp64_steps, 20)
    if not args.no_ensemble and not args.uq:
        out["ensemble"] = ensemble_leg(N)
    if not args.no_surrogate and not args.uq:
        out["surrogate"] = surrogate_leg(N, iterations=args.surrogate_iterations)
    if not args.no_seam and not args.uq:
        out["host_seam"] = host_seam_leg(N, args.workload)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
