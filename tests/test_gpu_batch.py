"""Batches and run-time parameters on the device, through the C ABI.

Every batch member must give the result the same model gives when it is
stepped alone: members are compared with the CPU oracle and with single FP64
device models (folded constants, compile-time geometry) after the same steps,
to 1e-8 relative L2 per species (the parity bar).  Batches are exercised with uneven
mesh sizes, different parameters, block sizes 2 and 4, one- and two-stage
schemes, interface facets, and members joining and leaving between steps.
"""
import numpy as np
import pytest

from tests.oracle_lib import OracleModel, rel_l2_per_species, rel_mass_l2_per_species

pytestmark = pytest.mark.gpu

TOL = 1e-8
# Newton to round-off on both sides (the l_2 criterion is on two_norm2), so
# that what is compared is the discretisation, not where Newton stopped
NL = dict(nl_rtol=1e-20, nl_atol=1e-28)


def _solo(dca, w, steps, overrides=""):
    """u after `steps` steps of a single FP64 model with the workload's settings."""
    m = dca.Model(w.ini, overrides, 0)
    m.set_state(w.initial_state(m), 0.0)
    t = 0.0
    for _ in range(steps):
        st = m.step(t, w.dt)
        assert st.status == 0, (w.name, st.status)
        t += w.dt
    u, t1 = m.get_state()
    info = m.compartments()
    m.close()
    return u, t1, info


def _solo_fp64():
    return ("model.time_step_operator.linear_solver.precision = fp64\n"
            "model.time_step_operator.linear_solver.preconditioner.type = Jacobi\n")


def test_runtime_parameters_match_folded(dca, wl):
    print("case: runtime parameters vs folded constants, gray_scott n=24", flush=True)
    w = wl.gray_scott(n=24, F=0.05, k=0.055, Du=2.5e-5, Dv=1e-5, constants=True, precond="Jacobi",
                      nl_rtol=1e-12, lin_rtol=1e-12, steps=3)
    folded = dca.Model(w.ini, "", 0)
    rt = dca.Model(w.ini, "", 0, runtime_parameters=True)
    assert rt.param_names() == ["Du", "Dv", "F", "k"]
    assert np.allclose(rt.params(), [2.5e-5, 1e-5, 0.05, 0.055], rtol=0, atol=0)
    for m in (folded, rt):
        m.set_state(w.initial_state(m), 0.0)
        for s in range(3):
            assert m.step(s * w.dt, w.dt).status == 0
    a, _ = folded.get_state()
    b, _ = rt.get_state()
    # the unit-cell tables scaled at run time differ from the folded ones in
    # the last bits only
    assert max(rel_l2_per_species(folded.compartments(), b, a)) < 1e-10
    # new values in place == a model built with those constants
    rt.set_params([3e-5, 1e-5, 0.03, 0.06])
    w2 = wl.gray_scott(n=24, F=0.03, k=0.06, Du=3e-5, Dv=1e-5, precond="Jacobi", nl_rtol=1e-12, lin_rtol=1e-12)
    ref = dca.Model(w2.ini, "", 0)
    for m in (ref, rt):
        m.set_state(w2.initial_state(m), 0.0)
        for s in range(2):
            assert m.step(s * w.dt, w.dt).status == 0
    a, _ = ref.get_state()
    b, _ = rt.get_state()
    assert max(rel_l2_per_species(ref.compartments(), b, a)) < 1e-10


# point_tile: the batched point apply's tile shape, timed ("auto") or forced
# to each candidate of the 2D Gray-Scott module (the first is the assembly tile)
@pytest.mark.parametrize("precision,graphs,tile", [("mixed", 4, "auto"), ("fp64", 4, "auto"), ("mixed", 0, "auto"),
                                                   ("mixed", 4, "1"), ("mixed", 4, "2"), ("fp64", 4, "3"),
                                                   ("mixed", 4, "4"), ("fp64", 4, "5")])
def test_batch_gray_scott_matches_single(dca, wl, precision, graphs, tile):
    print(f"case: batch gray_scott x6 precision={precision} graphs={graphs} point_tile={tile}", flush=True)
    steps = 3
    ws = wl.gray_scott_batch(count=6, seed=3, side_min=8, side_max=40, precision=precision, steps=steps, **NL)
    b = dca.Batch(f"precision = {precision}\ngraph_iterations = {graphs}\npoint_tile = {tile}\n", 0)
    ids = [b.add(w.ini) for w in ws]
    assert b.ids() == ids
    for i, w in zip(ids, ws):
        mv = b.member(i)
        mv.set_state(w.initial_state(mv), 0.0)
        mv.close()
    for s in range(steps):
        st = b.step()
        assert [x.status for x in st] == [0] * len(ws), s
        assert all(x.newton_iterations >= 1 for x in st)
        if precision == "mixed":
            assert all(x.refinement_steps >= 1 for x in st)
    for i, w in zip(ids, ws):
        mv = b.member(i)
        u, t = mv.get_state()
        mv.close()
        ref, tref, info = _solo(dca, w, steps, _solo_fp64())
        assert abs(t - tref) < 1e-12
        err = max(rel_l2_per_species(info, u, ref))
        assert err < TOL, (w.name, err)
    assert b.device_bytes() > 0
    b.close()


@pytest.mark.parametrize("tile", [str(k) for k in range(10)])
def test_batch_3d_point_tiles_match_single(dca, wl, tile):
    """3D Mitchell-Schaefer members of uneven sides (Alexander2, block size 2)
    with the batched point apply forced to each candidate tile shape, against
    each model stepped alone with the assembled FP64 solve."""
    print(f"case: batch mitchell_schaefer 3D point_tile={tile}", flush=True)
    steps = 2
    ws = [wl.mitchell_schaefer(n=n, dt=0.5, **NL) for n in (5, 9, 13)]
    b = dca.Batch(f"precision = mixed\npoint_tile = {tile}\n", 0)
    ids = [b.add(w.ini) for w in ws]
    for i, w in zip(ids, ws):
        b.set_time_step(i, w.dt)
        mv = b.member(i)
        mv.set_state(w.initial_state(mv), 0.0)
        mv.close()
    for s in range(steps):
        st = b.step()
        assert [x.status for x in st] == [0] * len(ws), s
    for i, w in zip(ids, ws):
        mv = b.member(i)
        u, _ = mv.get_state()
        mv.close()
        ref, _, info = _solo(dca, w, steps, _solo_fp64())
        err = max(rel_l2_per_species(info, u, ref))
        assert err < TOL, (w.name, err)
    b.close()


def test_batch_heterogeneous_members(dca, wl):
    """Block sizes 2 and 4, ImplicitEuler and Alexander2, a two-compartment
    member with interface facets, a simplex member."""
    print("case: heterogeneous batch", flush=True)
    members = [
        wl.gray_scott(n=16, precond="Jacobi", **NL),
        wl.mitchell_schaefer(n=6, dt=0.5, **NL),
        wl.two_compartments(n=20, dt=0.05, **NL),
        wl.gray_scott(n=12, cell_type="simplex", precond="Jacobi", **NL),
    ]
    steps = 2
    b = dca.Batch("precision = mixed\n", 0)
    ids = [b.add(w.ini) for w in members]
    for i, w in zip(ids, members):
        b.set_time_step(i, w.dt)
        mv = b.member(i)
        mv.set_state(w.initial_state(mv), 0.0)
        mv.close()
    for s in range(steps):
        st = b.step()
        assert [x.status for x in st] == [0] * len(members), s
        assert st[1].stages == 2 and st[0].stages == 1
    for i, w in zip(ids, members):
        mv = b.member(i)
        u, _ = mv.get_state()
        mv.close()
        ref, _, info = _solo(dca, w, steps, _solo_fp64())
        err = max(rel_l2_per_species(info, u, ref))
        assert err < TOL, (w.name, err)
    b.close()


def test_batch_members_join_and_leave(dca, wl):
    """Continuous batching: a member's result does not depend on who else is
    in the batch or on members joining and leaving between its steps."""
    print("case: join/leave", flush=True)
    ws = wl.gray_scott_batch(count=3, seed=7, side_min=10, side_max=30, precision="mixed", **NL)
    b = dca.Batch("precision = mixed\n", 0)
    a = b.add(ws[0].ini)
    c = b.add(ws[1].ini)
    for i, w in ((a, ws[0]), (c, ws[1])):
        mv = b.member(i)
        mv.set_state(w.initial_state(mv), 0.0)
        mv.close()
    b.step()
    d = b.add(ws[2].ini)
    mv = b.member(d)
    mv.set_state(ws[2].initial_state(mv), 0.0)
    mv.close()
    b.step()
    b.remove(a)
    b.step()
    done = {c: (ws[1], 3), d: (ws[2], 2)}
    for i, (w, steps) in done.items():
        mv = b.member(i)
        u, t = mv.get_state()
        mv.close()
        ref, tref, info = _solo(dca, w, steps, _solo_fp64())
        assert abs(t - tref) < 1e-12
        err = max(rel_l2_per_species(info, u, ref))
        assert err < TOL, (w.name, err)
    b.close()


def _oracle_steps(w, steps, dt=None):
    orc = OracleModel(w.ini)
    orc.set_state(w.initial_state(orc), 0.0)
    for _ in range(steps):
        assert orc.step(dt or w.dt)[0] == 0
    return orc


@pytest.mark.parametrize("floor", [0.0, 1e-4])
def test_batch_members_match_oracle(dca, wl, floor):
    """Every member of a mixed-precision batch against the CPU oracle stepped
    alone (the reference's Newton per model, make_step_operator.hh:246-288):
    uneven Gray-Scott members up to side 128, a Mitchell-Schaefer member
    (Alexander2, block size 2 in 3D) and a two-compartment member with
    transmission (block size 4), in the nodal and the mass-matrix norm."""
    print("case: batch members vs oracle", flush=True)
    steps = 3
    ws = wl.gray_scott_batch(count=4, seed=5, side_min=16, side_max=64, precision="mixed", steps=steps, **NL)
    ws.append(wl.gray_scott(n=128, seed=9, precond="Jacobi", precision="mixed", steps=steps, **NL))
    ws.append(wl.mitchell_schaefer(n=8, dt=0.5, **NL))
    ws.append(wl.two_compartments(n=24, dt=0.05, **NL))
    b = dca.Batch(f"precision = mixed\nrefinement.tolerance_floor = {floor}\n", 0)
    ids = [b.add(w.ini) for w in ws]
    for i, w in zip(ids, ws):
        b.set_time_step(i, w.dt)
        mv = b.member(i)
        mv.set_state(w.initial_state(mv), 0.0)
        mv.close()
    for s in range(steps):
        st = b.step()
        assert [x.status for x in st] == [0] * len(ws), s
    for i, w in zip(ids, ws):
        mv = b.member(i)
        u, t = mv.get_state()
        info = mv.compartments()
        mv.close()
        orc = _oracle_steps(w, steps)
        uo, to = orc.get_state()
        assert abs(t - to) < 1e-12
        err = max(rel_l2_per_species(info, u, uo))
        errm = max(rel_mass_l2_per_species(orc, u, uo))
        print(f"{w.name}: vs oracle rel L2 {err:.3e} (mass norm {errm:.3e})", flush=True)
        assert err < TOL and errm < TOL, (w.name, err, errm)
        orc.close()
    b.close()
