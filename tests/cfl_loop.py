"""The CFL time loop with redo, restated on the host around the oracle: the
loop of examples/shock.cpp:59-79 with MeshBlockImpl::max_time_step
(src/mesh/meshblock.cpp:511-530) and MeshBlockImpl::check_redo
(src/mesh/meshblock.cpp:974-1028).  One call of ``attempt`` is one pass of
the loop body; passes after the loop has stopped change nothing.

The step, the CFL reduction and the positivity verdict are arguments, so the
bookkeeping can be checked on scripted verdicts (tests/test_cfl_loop.py) and
then run with the oracle's FP64 step as the checker of the device loop
(tests/test_gpu_control.py)."""
import math

DENSITY, PRESSURE, NONFINITE = 1, 2, 4
DT_FROM_W, DT_FROM_U = 0, 1


class Loop:
    def __init__(self, cfl, tlim=math.inf, nlim=-1, max_redo=5, redo_mask=DENSITY | PRESSURE,
                 dt_source=DT_FROM_W):
        self.cfl, self.tlim, self.nlim = cfl, tlim, nlim
        self.max_redo, self.redo_mask, self.dt_source = max_redo, redo_mask, dt_source
        self.time, self.cycle, self.redo, self.status = 0.0, 0, 0, 0
        self.steps, self.redos = 0, 0
        self.dt_log = []

    def attempt(self, u, w, step, cfl_dt, verdict, cons2prim=None):
        """One pass: returns the new (u, w).  step(dt, u) -> (u_new, w_last
        stage); cfl_dt(w) -> HydroImpl::max_time_step(w); verdict(u) -> bits."""
        stop = (self.status != 0 or (self.nlim >= 0 and self.cycle >= self.nlim)
                or self.time >= self.tlim)
        if stop:
            if self.status == 0:
                self.status = 1
            self.dt_log.append(0.0)
            return u, w
        self.cycle += 1
        if self.dt_source == DT_FROM_U:
            w = cons2prim(u)
        # meshblock.cpp:529: pow(2, -current_redo) * cfl * dt
        dt = 2.0 ** -self.redo * self.cfl * cfl_dt(w)
        self.dt_log.append(dt)
        un, w = step(dt, u)
        if verdict(un) & self.redo_mask:
            self.redo += 1
            self.redos += 1
            if self.redo > self.max_redo:
                self.status = -1  # check_redo returns -1 before restoring
                return un, w
            self.cycle -= 1
            return u, w  # the step's input (w keeps the failed step's last stage)
        self.redo = 0
        self.time += dt
        self.steps += 1
        return un, w

    def run(self, n, u, w, step, cfl_dt, verdict, cons2prim=None):
        for _ in range(n):
            u, w = self.attempt(u, w, step, cfl_dt, verdict, cons2prim)
        return u, w
