"""CPU: the host restatement of the CFL loop with redo (tests/cfl_loop.py)
on scripted verdicts, against counts derived by hand from
examples/shock.cpp:59-79 and src/mesh/meshblock.cpp:511-530,974-1028; and
the NaN propagation of the oracle's CFL reduction."""
import math

import numpy as np
import pytest

import oracle as O
from cfl_loop import DENSITY, DT_FROM_U, NONFINITE, PRESSURE, Loop


def _scripted(verdicts):
    """Step and verdict stand-ins: the state is the number of accepted steps;
    the CFL step of every state is 1."""
    seq = iter(verdicts)
    return (lambda dt, u: (u + 1, "w")), (lambda w: 1.0), (lambda u: next(seq))


def test_redo_halves_dt_restores_and_gives_up():
    step, cfl_dt, verdict = _scripted([0, DENSITY, PRESSURE, 0, DENSITY, DENSITY | PRESSURE,
                                       DENSITY])
    L = Loop(cfl=0.5, max_redo=2)
    u, w = L.run(10, 0, "w0", step, cfl_dt, verdict)
    # attempts 0 and 3 accepted; 1-2 redo (dt / 2, / 4), then 4-6 redo until
    # the third in a row exceeds max_redo = 2; 7-9 do nothing
    assert L.dt_log == [0.5, 0.5, 0.25, 0.125, 0.5, 0.25, 0.125, 0.0, 0.0, 0.0]
    assert L.status == -1 and L.redo == 3 and L.redos == 5 and L.steps == 2
    # cycle: +1 per attempt, -1 per restored redo; the give-up keeps its +1
    assert L.cycle == 3
    assert L.time == 0.5 + 0.125
    # the failed step is kept when the loop gives up (u = input + 1)
    assert u == 3


def test_nonfinite_only_with_mask_and_stop_limits():
    step, cfl_dt, verdict = _scripted([NONFINITE, 0, 0])
    L = Loop(cfl=1.0)
    L.run(2, 0, "w", step, cfl_dt, verdict)
    assert L.redos == 0 and L.steps == 2 and L.dt_log == [1.0, 1.0]
    step, cfl_dt, verdict = _scripted([NONFINITE, 0, 0, 0])
    L = Loop(cfl=1.0, redo_mask=DENSITY | PRESSURE | NONFINITE, nlim=2)
    u, _ = L.run(5, 0, "w", step, cfl_dt, verdict)
    assert L.dt_log == [1.0, 0.5, 1.0, 0.0, 0.0]
    assert (L.status, L.cycle, L.steps, u) == (1, 2, 2, 2)
    step, cfl_dt, verdict = _scripted([0] * 5)
    L = Loop(cfl=1.0, tlim=2.5)
    L.run(5, 0, "w", step, cfl_dt, verdict)
    assert L.dt_log == [1.0, 1.0, 1.0, 0.0, 0.0] and L.time == 3.0 and L.status == 1


def test_dt_source_reads_input_primitives():
    seen = []
    L = Loop(cfl=1.0, dt_source=DT_FROM_U)
    L.run(2, 10, "stale", lambda dt, u: (u + 1, "last stage"),
          lambda w: seen.append(w) or 1.0, lambda u: 0, cons2prim=lambda u: ("prim", u))
    assert seen == [("prim", 10), ("prim", 11)]
    seen.clear()
    L = Loop(cfl=1.0)
    L.run(2, 10, "initial", lambda dt, u: (u + 1, "last stage"),
          lambda w: seen.append(w) or 1.0, lambda u: 0)
    assert seen == ["initial", "last stage"]


def test_oracle_cfl_propagates_nan():
    """torch's min() propagates NaN, so one interior cell with a negative
    pressure makes HydroImpl::max_time_step NaN; a ghost cell does not."""
    oc = O.sod_config(32)
    w, _ = O.initialize(oc, O.sod_init(oc))
    dt = O.max_time_step(oc, w, 1.0)
    # closed form: min dx / (|v| + sqrt(gamma p / rho)) over the two states
    dx = 1.0 / 32
    assert dt == pytest.approx(min(dx / math.sqrt(1.4), dx / math.sqrt(1.4 * 0.1 / 0.125)),
                               rel=1e-15)
    bad = w.copy()
    bad[4, 0, 0, 10] = -1.0
    assert math.isnan(O.max_time_step(oc, bad, 0.9))
    ghost = w.copy()
    ghost[4, 0, 0, 1] = -1.0
    assert O.max_time_step(oc, ghost, 1.0) == dt
    inf_v = w.copy()
    inf_v[1, 0, 0, 12] = np.inf  # dx / inf = 0: finite
    assert O.max_time_step(oc, inf_v, 1.0) == 0.0
