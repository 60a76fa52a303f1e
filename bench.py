#!/usr/bin/env python3
"""Newton DOF-steps/s of the device time step on one MI355X.

Workload (default): the 3D Mitchell-Schaefer excitable model on a 96^3 Q1 hex
mesh of the unit cube (912,673 nodes, 1,825,346 DOF), D = 1e-3 on u, Alexander2
with dt = 0.1, BiCGSTAB + Jacobi, Newton tolerance 1e-12 on the squared
l2 residual, linear tolerance 1e-10.  A "step" is one time step (two implicit
Runge-Kutta stages, each a Newton solve); Newton DOF-steps = stages x DOF.

Prints one JSON line with the throughput, per-step p50/p99 latency, the
roofline of the dominant kernel (BCSR SpMV, timed live with HIP events on the
model's stream) and a CPU baseline: the oracle (a C++ restatement of the
reference step, OpenMP-coloured assembly, sequential BiCGSTAB) timed on the
host cores for one time step of the same workload.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "dune-copasi_amd"
HBM_PEAK_GBS = 8000.0


def _load(name, path, pkg=False):
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=[str(path.parent)] if pkg else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def make_workload(wl, name, precision, precond="Jacobi"):
    if name == "ms96":
        return wl.mitchell_schaefer(n=96, cell_type="cube", precond=precond, nl_rtol=1e-12, lin_rtol=1e-10,
                                    precision=precision), "mitchell_schaefer_3d_q1_96^3_alexander2_dt0.1"
    if name == "ms48":
        return wl.mitchell_schaefer(n=48, cell_type="cube", precond=precond, nl_rtol=1e-12, lin_rtol=1e-10,
                                    precision=precision), "mitchell_schaefer_3d_q1_48^3_alexander2_dt0.1"
    if name == "gs128":
        return wl.gray_scott(n=128, cell_type="cube", precond="ILU", nl_rtol=1e-10, lin_rtol=1e-10,
                             precision=precision), "gray_scott_2d_q1_128^2_implicit_euler_dt1"
    if name == "gs1024":
        return wl.gray_scott(n=1024, cell_type="cube", precond=precond, nl_rtol=1e-16, lin_rtol=1e-10,
                             precision=precision), "gray_scott_2d_q1_1024^2_implicit_euler_dt1"
    if name == "two1024":
        # Newton to 1e-20 on the squared l2 defect (1e-10 in the norm): the
        # FP64 and mixed results then agree to equal accuracy (1e-8)
        return wl.two_compartments(n=1024, cell_type="cube", precision=precision, nl_rtol=1e-20, precond=precond), \
            "two_compartments_2d_q1_1024^2_implicit_euler_dt0.05"
    raise SystemExit(f"unknown workload {name}")


TRAFFIC_FILES = {("ms96", "fp64"): "traffic_mitchell_schaefer_q1_96_fp64.txt"}
VARIANT_KERNELS = {"lds_blocks": "spmv_blk_k", "lds_entries": "spmv_k", "rows_g8": "spmv_rows_k",
                   "rows_g16": "spmv_rows_k", "rows_g32": "spmv_rows_k", "rows_g64": "spmv_rows_k"}


def committed_traffic(workload, precision, chosen_kernel):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (rocprofv3 FETCH_SIZE/WRITE_SIZE passes calibrated on a plain
    copy; one `kernel =` line, then its `traffic_bytes =`, per kernel
    variant measured), when it covers the variant this run uses; None
    otherwise."""
    name = TRAFFIC_FILES.get((workload, precision))
    if not name or not (ROOT / "profiles" / name).exists():
        return None
    kernel, found = None, {}
    for line in (ROOT / "profiles" / name).read_text().splitlines():
        if "=" in line and not line.startswith("#"):
            k, v = (x.strip() for x in line.split("=", 1))
            if k == "kernel":
                kernel = v
            elif k == "traffic_bytes" and kernel is not None:
                found[kernel] = float(v)
    return found.get(chosen_kernel)


def _host_threads():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() else n


def _oracle_steps(o, w, steps):
    t0 = time.perf_counter()
    for k in range(steps):
        rc, _, _ = o.step(w.dt)
        if rc:
            raise RuntimeError(f"oracle step failed: {rc}")
        print(f"# oracle step {k + 1}/{steps} done at {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    return steps * w.stages * o.n_dof, el


def cpu_baseline(wl_obj, steps: int, small=None, gpu_state_after=None):
    """Oracle on the host cores: `steps` time steps of the same workload with
    OpenMP-coloured assembly and a sequential BiCGSTAB (as istl's), then one
    more step with the Krylov solve threaded too (the stronger baseline), and
    one step of `small` (a smaller mesh of the same model, or the workload
    itself) on one thread.  With gpu_state_after(k) (the device's state after
    k steps of the workload from its initial state), the oracle's state after
    its first 1 + `steps` steps is the parity check of the full-size run."""
    sys.path.insert(0, str(ROOT))
    from tests import oracle_lib  # the oracle is only ever the baseline here

    L = oracle_lib.load()
    cores = _host_threads()
    L.orc_set_threads(cores)
    L.orc_set_krylov_threads(0)
    o = oracle_lib.OracleModel(wl_obj.ini)
    o.set_state(wl_obj.initial_state(o), 0.0)
    _oracle_steps(o, wl_obj, 1)  # the first step (the stimulus transient) is a warm-up
    dof_steps, el = _oracle_steps(o, wl_obj, steps)
    parity = None
    if gpu_state_after is not None:
        u_orc, _ = o.get_state()
        u_gpu, info = gpu_state_after(1 + steps)
        parity = {"steps": 1 + steps,
                  "max_rel_l2_per_species": max(oracle_lib.rel_l2_per_species(info, u_gpu, u_orc)),
                  "max_rel_mass_l2_per_species": max(oracle_lib.rel_mass_l2_per_species(o, u_gpu, u_orc))}
    L.orc_set_krylov_threads(1)
    t_dof, t_el = _oracle_steps(o, wl_obj, 1)
    L.orc_set_krylov_threads(0)
    o.close()
    w1 = small or wl_obj
    L.orc_set_threads(1)
    o1 = oracle_lib.OracleModel(w1.ini)
    o1.set_state(w1.initial_state(o1), 0.0)
    s_dof, s_el = _oracle_steps(o1, w1, 1)
    o1.close()
    L.orc_set_threads(cores)
    return {"value": dof_steps / el, "unit": "Newton DOF-steps/s", "cores": cores, "kind": "port",
            "sample": f"time step(s) 2..{steps + 1} of the same workload from the same initial state, step 1 "
                      f"untimed ({dof_steps} Newton DOF-steps, {el:.1f} s, OpenMP-coloured assembly, "
                      f"sequential BiCGSTAB)",
            "threaded_krylov": {"value": t_dof / t_el, "cores": cores,
                                "sample": f"the next time step with SpMV, dots and updates on all threads "
                                          f"({t_dof} Newton DOF-steps, {t_el:.1f} s)"},
            "device_vs_oracle": parity,
            "one_thread": {"value": s_dof / s_el, "cores": 1,
                           "sample": f"one time step of {'the ' + str(o1.n_dof) + '-DOF mesh of the same model' if small else 'the same workload'} "
                                     f"({s_dof} Newton DOF-steps, {s_el:.1f} s)"}}


def cpu_baseline_batch(ws, dof_steps_budget: float):
    """Oracle on the host cores: one time step of as many members (in batch
    order) as fit the DOF-step budget."""
    sys.path.insert(0, str(ROOT))
    from tests import oracle_lib  # the oracle is only ever the baseline here

    L = oracle_lib.load()
    L.orc_set_threads(_host_threads())
    L.orc_set_krylov_threads(0)
    done, el, used = 0, 0.0, 0
    for w in ws:
        o = oracle_lib.OracleModel(w.ini)
        o.set_state(w.initial_state(o), 0.0)
        t0 = time.perf_counter()
        rc, _, _ = o.step(w.dt)
        el += time.perf_counter() - t0
        if rc:
            raise RuntimeError(f"oracle step failed: {rc}")
        done += o.n_dof * w.stages
        used += 1
        if done >= dof_steps_budget:
            break
    return {"value": done / el, "unit": "Newton DOF-steps/s", "cores": int(L.orc_max_threads()), "kind": "port",
            "sample": f"one time step of each of the first {used} members ({done} Newton DOF-steps, {el:.1f} s, "
                      f"OpenMP-coloured assembly, sequential BiCGSTAB)"}


def plain_fp64(dca, items, steps, precond="Jacobi", stride=1):
    """The plain FP64 path on the same models, one model at a time: each model
    built alone (FP64 BiCGSTAB, same preconditioner), one untimed step, then `steps`
    timed steps.  items: (ini, dt, stages, total steps the workload runs,
    initial state of a model handle -> packed state, size key).  stride > 1
    times every stride-th model in order of size (a stratified sample: each
    timed model stands for the `stride` models around it in size).
    Returns Newton DOF-steps/s of the whole workload at the measured per-step
    times and the sampled DOF-steps and seconds."""
    import torch

    ov = ("model.time_step_operator.linear_solver.precision = fp64\n"
          f"model.time_step_operator.linear_solver.preconditioner.type = {precond}\n")
    order = sorted(range(len(items)), key=lambda j: items[j][5])
    strata = [order[i:i + stride] for i in range(0, len(order), stride)]
    work, secs, sampled_work, sampled_secs = 0.0, 0.0, 0.0, 0.0
    for stratum in strata:
        ini, dt, stages, total, init, _ = items[stratum[len(stratum) // 2]]
        weight = len(stratum)
        m = dca.Model(ini, ov, 0, runtime_parameters=True)
        m.set_state(init(m), 0.0)
        t = 0.0
        if m.step(t, dt).status:
            raise SystemExit("plain FP64 step failed")
        t += dt
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            if m.step(t, dt).status:
                raise SystemExit("plain FP64 step failed")
            t += dt
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        work += weight * m.n_dof * stages * total
        secs += weight * el * total
        sampled_work += m.n_dof * stages * steps
        sampled_secs += el * steps
        m.close()
    return {"value": work / secs, "unit": "Newton DOF-steps/s", "models": len(items), "timed_models": len(strata),
            "timed_steps_per_model": steps, "sampled_dof_steps": sampled_work, "sampled_seconds": sampled_secs}


def batch_measure(args, dca, wl, precision, warmup, steps, plain_stride, cpu=True):
    """Config: 256 independent Gray-Scott instances advanced by one Batch.
    Returns the measurement as a dict (the workload's own JSON line, or an
    object of the default line)."""
    import torch

    ws = wl.gray_scott_batch(count=args.members, seed=0, precision=precision, nl_rtol=args.batch_nl_rtol,
                             nl_atol=0.0, lin_rtol=1e-10)
    t_setup = time.perf_counter()
    b = dca.Batch(f"precision = {precision}\ngraph_iterations = {args.graph_iterations}\n"
                  f"preconditioner = {args.preconditioner}\n"
                  f"refinement.inner_relative_tolerance = {args.ir_inner:g}\n"
                  f"refinement.tolerance_floor = {args.ir_floor:g}\n"
                  f"point_tile = {args.point_tile}\n", 0)
    ids = []
    for w in ws:
        i = b.add(w.ini)
        ids.append(i)
        mv = b.member(i)
        mv.set_state(w.initial_state(mv), 0.0)
        mv.close()
    n_dof = 0
    for i in ids:
        mv = b.member(i)
        n_dof += mv.n_dof
        mv.close()
    setup_s = time.perf_counter() - t_setup
    print(f"# batch set-up {setup_s:.1f}s: {len(ids)} members, {n_dof} DOF, "
          f"{b.device_bytes() / 2**30:.2f} GiB on the device", file=sys.stderr, flush=True)

    def one_step(k):
        st = b.step()
        bad = [x.status for x in st if x.status]
        if bad:
            raise SystemExit(f"batch step {k}: {len(bad)} members failed (status {bad[:5]})")
        return st

    for k in range(warmup):
        one_step(k)
    print(f"# batch warmup done ({warmup} steps)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lat, newton, lin, refs = [], 0, 0, 0
    t0 = time.perf_counter()
    for k in range(steps):
        s0 = time.perf_counter()
        st = one_step(k)
        lat.append(time.perf_counter() - s0)
        newton += max(x.newton_iterations for x in st)
        lin += max(x.linear_iterations for x in st)
        refs += max(x.refinement_steps for x in st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"# batch: timed {steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)
    # kernel roofline: profiled steps after the timed region (plain launches,
    # HIP events around every batched SpMV on the batch stream)
    b.profile_enable(True)
    for k in range(args.profile_steps):
        one_step(k)
    kind = "spmv32" if precision == "mixed" else "spmv64"
    n, ms, by = b.profile(kind)
    b.profile_enable(False)
    roofline = None
    if n:
        gbs = (by / n) / (ms / n * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": f"batched {'FP32' if kind == 'spmv32' else 'FP64'} Krylov operator "
                                              f"(point apply dcg_tile_0_pfbatch / BCSR bspmv_k)",
                    "achieved": gbs, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None, "avg_launch_us": 1e3 * ms / n,
                    "algorithmic_bytes_per_launch": by / n, "launches": n}
    total_steps = warmup + steps + args.profile_steps
    # error of sampled members against single FP64 models converged to round-off
    sample = sorted(range(len(ws)), key=lambda j: ws[j].cells[0])
    sample = sorted(set([sample[0], sample[len(sample) // 2], sample[-1]]))[: args.error_members]
    errs = {}
    tight = ("model.time_step_operator.linear_solver.precision = fp64\n"
             "model.time_step_operator.nonlinear_solver.convergence_condition.relative_tolerance = 1e-20\n"
             "model.time_step_operator.nonlinear_solver.convergence_condition.absolute_tolerance = 1e-28\n")
    for j in sample:
        w = ws[j]
        mv = b.member(ids[j])
        u, _ = mv.get_state()
        info = mv.compartments()
        mv.close()
        ref = dca.Model(w.ini, tight + "model.time_step_operator.linear_solver.preconditioner.type = Jacobi\n", 0)
        ref.set_state(w.initial_state(ref), 0.0)
        t = 0.0
        for _ in range(total_steps):
            if ref.step(t, w.dt).status:
                raise SystemExit("FP64 reference step failed")
            t += w.dt
        r, _ = ref.get_state()
        ref.close()
        e = []
        for c in info:
            for sp in range(c["n_species"]):
                a = u[c["dof_offset"] + sp: c["dof_offset"] + c["n_species"] * c["n_nodes"]: c["n_species"]]
                bb = r[c["dof_offset"] + sp: c["dof_offset"] + c["n_species"] * c["n_nodes"]: c["n_species"]]
                e.append(float(np.linalg.norm(a - bb) / max(np.linalg.norm(bb), 1e-300)))
        errs[w.name] = max(e)
    cpu_b = cpu_baseline_batch(ws, args.cpu_dof_steps) if cpu and not args.no_cpu_baseline else None
    value = steps * n_dof / elapsed
    plain = None
    if args.plain_steps > 0:
        print("# plain FP64 path on the same models ...", file=sys.stderr, flush=True)
        plain = plain_fp64(dca, [(w.ini, w.dt, 1, steps, w.initial_state, w.cells[0]) for w in ws],
                           args.plain_steps, args.preconditioner, plain_stride)
        plain["speedup_of_this_line"] = value / plain["value"]
    out = {
        "metric": "Newton DOF-steps/s (plain FP64 vs mixed/batched), p50/p99 latency, HBM GB/s",
        "value": value,
        "unit": "Newton DOF-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32-krylov/f64-refinement",
        "data": "synthetic (seeded initial states and parameters)",
        "config": {"workload": f"gray_scott_batch_{len(ws)}_2d_q1_sides32-512_implicit_euler_dt1",
                   "members": len(ws), "n_dof": n_dof, "precision": precision, "preconditioner": args.preconditioner,
                   "refinement_tolerance_floor": args.ir_floor if precision == "mixed" else None,
                   "graph_iterations": args.graph_iterations,
                   "newton_relative_tolerance_two_norm2": args.batch_nl_rtol},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "newton_iterations_per_step_max_member": newton / steps,
        "krylov_iterations_per_step_max_member": lin / steps,
        "refinements_per_step_max_member": refs / steps,
        "error_vs_fp64": {"max_rel_l2_per_species": max(errs.values()) if errs else None, "members": errs,
                          "steps": total_steps},
        "setup_s": setup_s,
        "device_bytes": b.device_bytes(),
        "plain_fp64_same_models": plain,
        "roofline": roofline,
        "cpu_baseline": cpu_b,
    }
    b.close()
    return out


def run_batch(args, dca, wl):
    print(json.dumps(batch_measure(args, dca, wl, args.precision, args.warmup, args.steps, args.plain_stride)))


def serve_measure(args, dca, wl, precision, R, qps, min_steps, max_steps, plain_steps=None, error_members=None):
    """Config: a Poisson stream of seeded random models served by one Server
    (qps <= 0: every request arrives at once, the saturated server).
    Returns the measurement as a dict."""
    import torch

    plain_steps = args.plain_steps if plain_steps is None else plain_steps
    error_members = args.error_members if error_members is None else error_members
    reqs = [wl.random_request(args.serve_seed * 100000 + i, precision=precision,
                              steps_range=(min_steps, max_steps)) for i in range(R)]
    rng = np.random.Generator(np.random.PCG64(args.serve_seed))
    if qps > 0:
        arrivals = np.cumsum(rng.exponential(1.0 / qps, size=R))
        arrivals -= arrivals[0]
    else:
        arrivals = np.zeros(R)
    # the compiled kernels of every request's model are cached on disk (a
    # parameter sweep or fit sends the same model structure again and again):
    # one untimed pass of one step per request builds and compiles them, so
    # the timed stream measures serving, not first-time hipRTC compilation
    t_warm = time.perf_counter()
    warm = dca.Server(f"precision = {precision}\nmax_members = {args.max_members}\nworkers = {args.workers}\n"
                      f"preconditioner = {args.preconditioner}\n", 0)
    wids = [warm.submit(r.ini, 1) for r in reqs]
    while warm.tick() > 0:
        pass
    if any(warm.result(i)[1].status for i in wids):
        raise SystemExit("warm-up request failed")
    warm.close()
    warm_s = time.perf_counter() - t_warm
    srv = dca.Server(f"precision = {precision}\nmax_members = {args.max_members}\n"
                     f"workers = {args.workers}\ngraph_iterations = {args.graph_iterations}\n"
                     f"preconditioner = {args.preconditioner}\n"
                     f"refinement.tolerance_floor = {args.ir_floor:g}\n", 0)
    ids = [None] * R
    t0 = time.perf_counter()
    sub, last_print = 0, 0.0
    while True:
        now = time.perf_counter() - t0
        while sub < R and arrivals[sub] <= now:
            ids[sub] = srv.submit(reqs[sub].ini, reqs[sub].steps)
            sub += 1
        open_ = srv.tick()
        if sub == R and open_ == 0:
            break
        if now - last_print > 10.0:
            st = srv.stats()
            print(f"# t={now:.0f}s submitted {sub}/{R} resident {st['resident']} finished {st['finished']} "
                  f"used {st['used_bytes'] / 2**30:.1f} GiB", file=sys.stderr, flush=True)
            last_print = now
        if open_ == 0 and sub < R:
            time.sleep(max(0.0, min(0.05, arrivals[sub] - (time.perf_counter() - t0))))
    torch.cuda.synchronize()
    infos = [srv.result(i)[1] for i in ids]
    bad = [(reqs[k].name, x.status) for k, x in enumerate(infos) if x.status]
    if bad:
        raise SystemExit(f"failed requests: {bad}")
    first = min(x.t_arrival for x in infos)
    makespan = max(x.t_finished for x in infos) - first
    work = sum(x.n_dof * x.steps * reqs[k].stages for k, x in enumerate(infos))
    lat = np.array([x.t_finished - x.t_arrival for x in infos])
    queue = np.array([x.t_admitted - x.t_arrival for x in infos])
    print(f"# served {R} requests in {makespan:.1f}s", file=sys.stderr, flush=True)
    # error of sampled results against single FP64 models converged to round-off
    tight = ("model.time_step_operator.linear_solver.precision = fp64\n"
             "model.time_step_operator.linear_solver.preconditioner.type = Jacobi\n"
             "model.time_step_operator.nonlinear_solver.convergence_condition.relative_tolerance = 1e-20\n"
             "model.time_step_operator.nonlinear_solver.convergence_condition.absolute_tolerance = 1e-28\n")
    order = sorted(range(R), key=lambda k: infos[k].n_dof * infos[k].steps)
    errs = {}
    for k in order[: error_members]:
        u, _ = srv.result(ids[k])
        ref = dca.Model(reqs[k].ini, tight, 0)
        ref.set_state(ref.initial_state(), 0.0)
        t = 0.0
        for _ in range(reqs[k].steps):
            if ref.step(t, reqs[k].dt).status:
                raise SystemExit("FP64 reference step failed")
            t += reqs[k].dt
        r, _ = ref.get_state()
        e = []
        for c in ref.compartments():
            for sp in range(c["n_species"]):
                sl = slice(c["dof_offset"] + sp, c["dof_offset"] + c["n_species"] * c["n_nodes"], c["n_species"])
                e.append(float(np.linalg.norm(u[sl] - r[sl]) / max(np.linalg.norm(r[sl]), 1e-300)))
        ref.close()
        errs[reqs[k].name] = max(e)
    srv.close()
    plain = None
    if plain_steps > 0:
        print("# plain FP64 path on the same requests ...", file=sys.stderr, flush=True)
        plain = plain_fp64(dca, [(r.ini, r.dt, r.stages, r.steps, lambda m: m.initial_state(), k)
                                 for k, r in enumerate(reqs)], plain_steps, args.preconditioner)
        plain["speedup_of_this_line"] = (work / makespan) / plain["value"]
    out = {
        "metric": "Newton DOF-steps/s (plain FP64 vs mixed/batched), p50/p99 latency, HBM GB/s",
        "value": work / makespan,
        "unit": "Newton DOF-steps/s",
        "n_gpus": 1,
        "steps": R,
        "warmup": 0,
        "ms_per_step": 1e3 * makespan / max(1, sum(x.steps for x in infos)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32-krylov/f64-refinement",
        "data": "synthetic (seeded random models and Poisson arrivals)",
        "config": {"workload": "serving_poisson_random_mass_action_2d_64-512_3d_32-96",
                   "requests": R, "qps": qps, "steps_range": [min_steps, max_steps],
                   "refinement_tolerance_floor": args.ir_floor if precision == "mixed" else None,
                   "precision": precision, "max_members": args.max_members, "workers": args.workers},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(lat.max())},
        "queue_s": {"p50": float(np.percentile(queue, 50)), "p99": float(np.percentile(queue, 99))},
        "compile_warmup_s": warm_s,
        "makespan_s": makespan,
        "newton_dof_steps": work,
        "error_vs_fp64": {"max_rel_l2_per_species": max(errs.values()) if errs else None, "requests": errs},
        "plain_fp64_same_requests": plain,
        "roofline": None,
        "cpu_baseline": None,
    }
    return out


def run_serve(args, dca, wl):
    print(json.dumps(serve_measure(args, dca, wl, args.precision, args.requests, args.qps, args.min_steps,
                                   args.max_steps)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ms96")
    ap.add_argument("--preconditioner", default="Jacobi", choices=["Jacobi", "VertexBlockJacobi"],
                    help="Jacobi: the reference's (diagonal); VertexBlockJacobi: inverse of each vertex's species block")
    ap.add_argument("--profile-steps-single", type=int, default=2,
                    help="single models: steps after the timed ones with per-kernel HIP-event timing")
    ap.add_argument("--readback-iterations", type=int, default=0,
                    help="single-model workloads: BiCGSTAB iterations between read-backs (0: adaptive; 1 launches "
                         "none past convergence, so a profiler's per-kernel averages count only working launches)")
    ap.add_argument("--storage", default="auto", choices=["matrix", "points", "auto"],
                    help="single models: the Jacobian as an assembled BCSR matrix, or as point coefficients per "
                         "cell applied by the Krylov operator (linear_solver.storage)")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="default: fp64 for single models, mixed for gs_batch and serve")
    ap.add_argument("--point-tile", default="auto",
                    help="point storage: the apply's tile shape, auto (timed at set-up) or a candidate index")
    ap.add_argument("--ir-floor", type=float, default=None,
                    help="mixed precision: lowest relative tolerance a Newton correction is solved to (0: as requested);"
                         " Newton's FP64 defect test still decides convergence.  Default 1e-4, and 1e-5 for the"
                         " two-compartment workload, whose states then agree with the FP64 run within the 1e-8 bar"
                         " (1.6e-8 at 1e-4)")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="single-model FP64 runs: skip the mixed-precision run of the same workload")
    ap.add_argument("--copy-probe-mib", type=int, default=1024,
                    help="after the timed steps, a plain device copy of this size each way: the box's copy"
                         " bandwidth beside the roofline, and the counter calibration under rocprofv3 (0: none)")
    ap.add_argument("--ir-inner", type=float, default=1e-3,
                    help="mixed precision: relative tolerance of each FP32 Krylov solve inside the FP64 refinement")
    ap.add_argument("--members", type=int, default=256, help="batch workload: number of models")
    ap.add_argument("--graph-iterations", type=int, default=4, help="batch workload: iterations per hipGraph")
    ap.add_argument("--batch-nl-rtol", type=float, default=1e-16,
                    help="batch workload: Newton tolerance on the squared l2 residual")
    ap.add_argument("--profile-steps", type=int, default=1, help="batch workload: profiled steps after timing")
    ap.add_argument("--error-members", type=int, default=3, help="batch workload: members checked against FP64")
    ap.add_argument("--cpu-dof-steps", type=float, default=4e6, help="batch workload: CPU baseline sample size")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="gs_batch / serve: timed steps per model of the plain FP64 path on the same models (0: skip)")
    ap.add_argument("--plain-stride", type=int, default=1,
                    help="gs_batch: time the plain FP64 path on every n-th model by size (stratified sample)")
    ap.add_argument("--no-batch", action="store_true",
                    help="default line: skip the batched-models object (256 Gray-Scott models)")
    ap.add_argument("--no-serve", action="store_true", help="default line: skip the serving object (12 random requests)")
    ap.add_argument("--line-batch-steps", type=int, default=3, help="default line: timed batch steps")
    ap.add_argument("--line-plain-stride", type=int, default=4,
                    help="default line: stratified sample of the plain FP64 path beside the batch")
    ap.add_argument("--line-requests", type=int, default=12, help="default line: served requests")
    ap.add_argument("--line-qps", type=float, default=0.5, help="default line: serving arrival rate")
    ap.add_argument("--line-min-steps", type=int, default=20, help="default line: fewest steps of a served request")
    ap.add_argument("--line-max-steps", type=int, default=60, help="default line: most steps of a served request")
    ap.add_argument("--requests", type=int, default=16, help="serve workload: number of requests")
    ap.add_argument("--qps", type=float, default=1.0, help="serve workload: Poisson arrival rate")
    ap.add_argument("--serve-seed", type=int, default=1, help="serve workload: seed of models and arrivals")
    ap.add_argument("--min-steps", type=int, default=50, help="serve workload: fewest steps of a request")
    ap.add_argument("--max-steps", type=int, default=500, help="serve workload: most steps of a request")
    ap.add_argument("--max-members", type=int, default=64, help="serve workload: resident request slots")
    ap.add_argument("--workers", type=int, default=8, help="serve workload: host build/compile threads")
    args = ap.parse_args()
    if args.ir_floor is None:
        args.ir_floor = 1e-5 if args.workload == "two1024" else 1e-4
    if args.precision is None:
        args.precision = "mixed" if args.workload in ("gs_batch", "serve") else "fp64"
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        raise SystemExit(2)

    import torch

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dca = _load("dune_copasi_amd", PKG / "__init__.py", pkg=True)
    wl = _load("dca_workloads", PKG / "workloads.py")
    if args.workload == "gs_batch":
        return run_batch(args, dca, wl)
    if args.workload == "serve":
        return run_serve(args, dca, wl)
    w, wname = make_workload(wl, args.workload, args.precision, args.preconditioner)

    storage = f"model.time_step_operator.linear_solver.storage = {args.storage}\n"
    if args.point_tile != "auto":
        storage += f"model.time_step_operator.linear_solver.point_tile = {args.point_tile}\n"
    if args.readback_iterations:
        storage += f"model.time_step_operator.linear_solver.readback_iterations = {args.readback_iterations}\n"
    t_setup = time.perf_counter()
    m = dca.Model(w.ini, storage, 0)
    u0 = w.initial_state(m)
    m.set_state(u0, 0.0)
    setup_s = time.perf_counter() - t_setup
    tuning, chosen = m.spmv_tuning()
    ptuning = m.point_tuning()
    print(f"# setup {setup_s:.1f}s, {m.n_dof} DOF, {m.n_blocks} blocks of {m.block_size}x{m.block_size}; "
          f"SpMV autotune {tuning} -> {chosen}; point-apply tiles {ptuning}", file=sys.stderr, flush=True)

    t = 0.0
    for k in range(args.warmup):
        st = m.step(t, w.dt)
        if st.status:
            raise SystemExit(f"warmup step {k} failed with status {st.status}")
        t += w.dt
    print(f"# warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    lat, newton, lin = [], 0, 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        s0 = time.perf_counter()
        st = m.step(t, w.dt)
        lat.append(time.perf_counter() - s0)
        if st.status:
            raise SystemExit(f"step {k} failed with status {st.status}")
        newton += st.newton_iterations
        lin += st.linear_iterations
        t += w.dt
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"# timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)

    u_timed, _ = m.get_state()  # the state the mixed run is compared with
    # per-kernel HIP-event timing over further steps (events around every
    # launch; these steps are not part of the timed region)
    m.profile_enable(True)
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    for k in range(args.profile_steps_single):
        st = m.step(t, w.dt)
        if st.status:
            raise SystemExit(f"profiled step {k} failed with status {st.status}")
        t += w.dt
    torch.cuda.synchronize()
    prof_elapsed = time.perf_counter() - p0

    kinds = {}
    for kname in ["spmv64", "spmv32", "assembly_jacobian", "assembly_residual", "precond_setup", "point_setup",
                  "point_apply64", "point_apply32"]:
        n, ms, b = m.profile(kname)
        if n:
            kinds[kname] = {"launches": n, "avg_us": 1e3 * ms / n, "gbs": (b / n) / (ms / n * 1e-3) / 1e9 if ms else 0.0,
                            "bytes_per_launch": b / n, "share_of_step_time": ms / (1e3 * prof_elapsed)}
    # the dominant kernel: the Krylov operator (assembled SpMV or point-coefficient apply)
    apply_kinds = [k for k in ("spmv64", "spmv32", "point_apply64", "point_apply32") if k in kinds]
    spmv_kind = max(apply_kinds, key=lambda k: kinds[k]["share_of_step_time"]) if apply_kinds else None
    sp = kinds.get(spmv_kind)
    roofline = None
    if sp:
        if spmv_kind.startswith("point"):
            # the generated apply of compartment 0 with the tile shape in use
            shapes = [ch for c, _, _, ch in ptuning if c == 0]
            k = shapes.index(True) if True in shapes else 0
            kernel = f"dcg_tile_0_p{'f' if spmv_kind.endswith('32') else ''}{k}"
        else:
            kernel = VARIANT_KERNELS.get(chosen, "?")
        roofline = {"bound": "hbm", "kernel": spmv_kind, "kernel_name": kernel, "achieved": sp["gbs"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sp["gbs"] / HBM_PEAK_GBS,
                    "traffic": committed_traffic(args.workload, args.precision, kernel),
                    "avg_launch_us": sp["avg_us"], "algorithmic_bytes_per_launch": sp["bytes_per_launch"]}

    probe = None
    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib * (1 << 20)
        ms = m.probe_copy(nbytes)
        probe = {"bytes_each_way": nbytes, "ms": ms, "gbs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}

    value = args.steps * w.stages * m.n_dof / elapsed
    mixed = None
    if args.precision == "fp64" and not args.no_mixed:
        # the FP32-Krylov / FP64-refinement form on the same workload and steps,
        # with its error against this FP64 run
        u64 = u_timed
        w32, _ = make_workload(wl, args.workload, "mixed", args.preconditioner)
        m32 = dca.Model(w32.ini, storage + "model.time_step_operator.linear_solver.refinement.inner_relative_tolerance = "
                        f"{args.ir_inner:g}\nmodel.time_step_operator.linear_solver.refinement.tolerance_floor = "
                        f"{args.ir_floor:g}\n", 0)
        m32.set_state(u0, 0.0)
        t32 = 0.0
        for k in range(args.warmup):
            if m32.step(t32, w32.dt).status:
                raise SystemExit("mixed warmup step failed")
            t32 += w32.dt
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        refs = 0
        for k in range(args.steps):
            st = m32.step(t32, w32.dt)
            if st.status:
                raise SystemExit("mixed step failed")
            refs += st.refinement_steps
            t32 += w32.dt
        torch.cuda.synchronize()
        el32 = time.perf_counter() - s0
        u32, _ = m32.get_state()
        errs = []
        for c in m.compartments():
            for sp in range(c["n_species"]):
                sl = slice(c["dof_offset"] + sp, c["dof_offset"] + c["n_species"] * c["n_nodes"], c["n_species"])
                errs.append(float(np.linalg.norm(u32[sl] - u64[sl]) / max(np.linalg.norm(u64[sl]), 1e-300)))
        mixed = {"value": args.steps * w32.stages * m32.n_dof / el32, "ms_per_step": 1e3 * el32 / args.steps,
                 "refinement_tolerance_floor": args.ir_floor,
                 "speedup_vs_fp64": (args.steps * w32.stages * m32.n_dof / el32) / value,
                 "refinements_per_step": refs / args.steps,
                 "error_vs_fp64_run": {"max_rel_l2_per_species": max(errs), "steps": args.warmup + args.steps}}
        m32.close()
        print(f"# mixed {mixed['value']:.3e} DOF-steps/s, error vs FP64 {max(errs):.2e}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        # one thread on the 48^3 mesh of the same model (a 96^3 step takes minutes on one core)
        small = make_workload(wl, "ms48", "fp64", args.preconditioner)[0] if args.workload == "ms96" else None

        def gpu_state_after(k):
            # the device's state after k steps from the workload's initial state
            m.set_state(u0, 0.0)
            tk = 0.0
            for _ in range(k):
                st = m.step(tk, w.dt)
                if st.status:
                    raise SystemExit(f"parity step failed with status {st.status}")
                tk += w.dt
            return m.get_state()[0], m.compartments()

        cpu_w, same = w, True
        if args.workload == "two1024":
            # a 1024^2 step takes minutes on the host: the 256^2 mesh of the
            # same model (one thread: 128^2), no full-size parity check
            cpu_w = wl.two_compartments(n=256, cell_type="cube", precision="fp64", nl_rtol=1e-20,
                                        precond=args.preconditioner)
            small = wl.two_compartments(n=128, cell_type="cube", precision="fp64", nl_rtol=1e-20,
                                        precond=args.preconditioner)
            same = False
        cpu = cpu_baseline(cpu_w, args.cpu_steps, small, gpu_state_after if same else None)
        if not same:
            cpu["sample"] = cpu["sample"].replace("of the same workload", "of the 256^2 mesh of the same model")
    n_dof = m.n_dof
    m.close()
    batch_obj = serve_obj = None
    if args.workload == "ms96" and not args.no_batch:
        # the north-star forms on their own workloads: the 256-member batch
        # (mixed precision, beside the plain FP64 path on the same models)
        print("# batch of 256 Gray-Scott models ...", file=sys.stderr, flush=True)
        batch_obj = batch_measure(args, dca, wl, "mixed", 1, args.line_batch_steps, args.line_plain_stride, cpu=False)
        batch_obj = {k: v for k, v in batch_obj.items() if k not in ("metric", "n_gpus", "higher_is_better", "scaling",
                                                                    "vs_baseline", "data")}
    if args.workload == "ms96" and not args.no_serve:
        # the same requests twice: all at once (the saturated server's
        # throughput beside the plain FP64 path), then as a Poisson stream at
        # a rate the server sustains (latency of service rather than queueing)
        drop = ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data", "roofline", "cpu_baseline",
                "warmup")
        print("# serving sample, all requests at once ...", file=sys.stderr, flush=True)
        sat = serve_measure(args, dca, wl, "mixed", args.line_requests, 0.0, args.line_min_steps, args.line_max_steps)
        print(f"# serving sample at {args.line_qps} requests/s ...", file=sys.stderr, flush=True)
        paced = serve_measure(args, dca, wl, "mixed", args.line_requests, args.line_qps, args.line_min_steps,
                              args.line_max_steps, plain_steps=0, error_members=0)
        serve_obj = {"saturated": {k: v for k, v in sat.items() if k not in drop},
                     "paced": {k: v for k, v in paced.items() if k not in drop + ("error_vs_fp64",
                                                                                 "plain_fp64_same_requests")}}
    out = {
        "metric": "Newton DOF-steps/s (plain FP64 vs mixed/batched), p50/p99 latency, HBM GB/s",
        "value": value,
        "unit": "Newton DOF-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-krylov/f64-refinement",
        "data": "synthetic (seeded initial state)",
        "config": {"workload": wname, "n_dof": n_dof, "stages_per_step": w.stages,
                   "precision": args.precision,
                   "preconditioner": "ILU" if args.workload == "gs128" else args.preconditioner,
                   "jacobian_storage": args.storage, "readback_iterations": args.readback_iterations},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "newton_iterations_per_step": newton / args.steps,
        "krylov_iterations_per_step": lin / args.steps,
        "kernels": kinds,
        "roofline": roofline,
        # the SpMV variants exist only for an assembled matrix
        "spmv_autotune": {"candidates_us": tuning, "chosen": chosen} if "spmv64" in kinds or "spmv32" in kinds else None,
        # the point apply's tile shapes (cells per workgroup) timed at set-up
        "point_tile_autotune": [{"compartment": c, "tile": list(tl), "us": round(us, 2), "chosen": ch}
                                for c, tl, us, ch in ptuning] or None,
        "mixed_precision_same_workload": mixed,
        "copy_probe": probe,
        "cpu_baseline": cpu,
        "batch_same_models": batch_obj,
        "serving": serve_obj,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
