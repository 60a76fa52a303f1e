"""The request loop (dcg_server_*) on the device, through the C ABI.

Requests are small seeded random models of the serving stream: 2D and 3D
meshes, 2-8 species, one or two compartments with a transmission flux,
random mass-action networks, implicit Euler and Alexander2.  Checked:
  * a request's result is the same alone and in a mix where other requests
    arrive, run beside it and leave (the serving correctness condition);
  * every served (mixed-precision, batched) result matches the CPU oracle
    and a single FP64 device model, both converged to round-off, to 1e-8
    relative L2 per species;
  * admission by HBM footprint: with a budget that holds one request at a
    time everything still finishes, one after the other.
"""
import time

import numpy as np
import pytest

from tests.oracle_lib import OracleModel, rel_l2_per_species, rel_mass_l2_per_species

pytestmark = pytest.mark.gpu

SMALL = dict(side2=(12, 20), side3=(5, 8), steps_range=(3, 6), nl_rtol=1e-20, nl_atol=1e-28)
MIX = [0, 1, 4, 6, 9, 11]  # 3D/2D, 2-8 species, 1-2 compartments, both schemes


def _drain(srv, seconds=200.0):
    t0 = time.time()
    while time.time() - t0 < seconds:
        if srv.tick() == 0:
            return
    raise AssertionError("server did not finish")


def _info(dca, ini):
    m = dca.Model(ini, "", -1)
    info = m.compartments()
    m.close()
    return info


def test_result_same_alone_and_in_mix(dca, wl):
    print("case: alone vs mix", flush=True)
    reqs = [wl.random_request(s, **SMALL) for s in MIX]
    target = reqs[0]
    alone = dca.Server("max_members = 8\nworkers = 2\n", 0)
    a = alone.submit(target.ini, target.steps)
    _drain(alone)
    ua, ia = alone.result(a)
    assert ia.status == 0 and ia.steps_done == target.steps
    alone.close()
    mix = dca.Server("max_members = 8\nworkers = 2\n", 0)
    others = [mix.submit(r.ini, r.steps) for r in reqs[1:3]]
    for _ in range(2):
        mix.tick()
    t = mix.submit(target.ini, target.steps)
    others += [mix.submit(r.ini, r.steps) for r in reqs[3:]]
    _drain(mix)
    ut, it = mix.result(t)
    assert it.status == 0 and it.steps_done == target.steps
    for o in others:
        assert mix.result(o)[1].status == 0
    err = max(rel_l2_per_species(_info(dca, target.ini), ut, ua))
    print(f"alone vs mix rel L2 {err:.3e}", flush=True)
    assert err <= 1e-12
    mix.close()


def test_served_results_match_fp64(dca, wl):
    print("case: served vs FP64", flush=True)
    reqs = [wl.random_request(s, **SMALL) for s in MIX]
    srv = dca.Server("max_members = 8\nworkers = 3\n", 0)
    ids = [srv.submit(r.ini, r.steps) for r in reqs]
    _drain(srv)
    tight = ("model.time_step_operator.linear_solver.precision = fp64\n"
             "model.time_step_operator.linear_solver.preconditioner.type = Jacobi\n")
    for rid, r in zip(ids, reqs):
        u, info = srv.result(rid)
        assert info.status == 0, (r.name, info.status)
        assert info.t_arrival <= info.t_ready <= info.t_admitted <= info.t_finished
        ref = dca.Model(r.ini, tight, 0)
        ref.set_state(ref.initial_state(), 0.0)
        t = 0.0
        for _ in range(r.steps):
            assert ref.step(t, r.dt).status == 0
            t += r.dt
        uref, _ = ref.get_state()
        err = max(rel_l2_per_species(ref.compartments(), u, uref))
        ref.close()
        print(f"{r.name}: rel L2 vs FP64 {err:.3e}", flush=True)
        assert err < 1e-8, (r.name, err)
    srv.close()


def test_admission_by_footprint(dca, wl):
    print("case: admission by footprint", flush=True)
    reqs = [wl.random_request(s, **SMALL) for s in MIX[:4]]
    est = [dca.estimate_device_bytes(r.ini) for r in reqs]
    budget = int(1.05 * max(est))
    srv = dca.Server(f"hbm_budget_bytes = {budget}\nmax_members = 8\nworkers = 2\n", 0)
    ids = [srv.submit(r.ini, r.steps) for r in reqs]
    most = 0
    t0 = time.time()
    while time.time() - t0 < 200.0:
        open_ = srv.tick()
        st = srv.stats()
        most = max(most, st["resident"])
        assert st["used_bytes"] <= budget * 1.5
        if open_ == 0:
            break
    assert most >= 1
    infos = [srv.result(i)[1] for i in ids]
    assert all(i.status == 0 and i.finished for i in infos)
    # two requests were resident together only if the later one's estimate
    # fitted beside the earlier one's actual bytes
    for i in range(len(infos)):
        for j in range(len(infos)):
            a, b = infos[i], infos[j]
            if i == j or not (a.t_admitted <= b.t_admitted < a.t_finished):
                continue
            assert a.device_bytes + est[j] <= budget, (i, j)
    srv.close()


def test_served_results_match_oracle(dca, wl):
    """Every served result of a mix of requests (mixed precision, batched,
    arriving while others run; one and two compartments with a transmission
    flux, 2D and 3D, both schemes) against the CPU oracle stepping the same
    request alone, in the nodal and the mass-matrix norm."""
    print("case: served vs oracle", flush=True)
    seeds = MIX + [2, 5]
    reqs = [wl.random_request(s, **SMALL) for s in seeds]
    assert any(len(_info(dca, r.ini)) == 2 for r in reqs)
    srv = dca.Server("max_members = 8\nworkers = 3\n", 0)
    ids = [srv.submit(r.ini, r.steps) for r in reqs[:4]]
    for _ in range(2):
        srv.tick()
    ids += [srv.submit(r.ini, r.steps) for r in reqs[4:]]
    _drain(srv)
    for rid, r in zip(ids, reqs):
        u, info = srv.result(rid)
        assert info.status == 0, (r.name, info.status)
        orc = OracleModel(r.ini)
        orc.set_state(orc_initial(dca, r.ini), 0.0)
        for _ in range(r.steps):
            assert orc.step(r.dt)[0] == 0
        uo, _ = orc.get_state()
        err = max(rel_l2_per_species(orc.compartments(), u, uo))
        errm = max(rel_mass_l2_per_species(orc, u, uo))
        print(f"{r.name}: served vs oracle rel L2 {err:.3e} (mass norm {errm:.3e})", flush=True)
        assert err < 1e-8 and errm < 1e-8, (r.name, err, errm)
        orc.close()
    srv.close()


def orc_initial(dca, ini):
    """The request's initial state from its initial.expression keys (the
    server starts every request from them)."""
    m = dca.Model(ini, "", -1)
    u = m.initial_state()
    m.close()
    return u
