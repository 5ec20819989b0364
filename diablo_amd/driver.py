"""The channel time loop of DIABLO's main program (for/diablo.f:55-138) on
the GPU: INITIALIZE from the decks (input.dat, input_chan.dat, grid_def), a
new flow or a restart, RK_CHAN_1 three times per step with COURANT under
VARIABLE_DT (and the LES model once past its spin-up when USE_LES is set),
SAVE_STATS and SAVE_FLOW at their intervals, TIME += DELTA_T,
and END_RUN's stop conditions (wall time beyond TIME_LIMIT, a file named
stop.now; for/diablo_io.f:412-435).

Output files follow the reference's text layout where it has one: mean.txt,
mean_th.txt (for/save_stats.f:430-447, 670-686), tke.txt
(for/save_stats.f:1222-1235) and, with LES, tke_les.txt
(for/save_stats.f:940-958).  Restart files are the .npz archives of
diablo_amd.restart (out.npz, end.npz with the pressure; start.npz is read
when CREATE_NEW_FLOW is false).

    python -m diablo_amd.driver RUN_DIR [--precision fp64|mixed|mixed_f64state]
"""
from __future__ import annotations

import argparse
import math
import os
import time as _time
from typing import Optional

import numpy as np

from . import case, restart
from .gpu import CU2, GpuChannel


def interval_due(interval: int, time_step: int, time: float, delta_t: float) -> bool:
    """The SAVE_STATS_INT / SAVE_FLOW_INT test of for/diablo.f:86-107: a
    negative interval is a period in time units, a positive one in steps."""
    if interval < 0:
        period = float(abs(interval))
        return math.fmod(time, period) + delta_t > period
    return interval != 0 and time_step % interval == 0


def end_run(start_wall: float, time_limit: float, directory: str, now: Optional[float] = None) -> Optional[str]:
    """END_RUN (for/diablo_io.f:412-435): the reason to stop, or None."""
    t = _time.time() if now is None else now
    if t - start_wall > time_limit:
        return "wall-time hit"
    if os.path.exists(os.path.join(directory, "stop.now")):
        return "stop.now file"
    return None


def _fmt(x: float) -> str:
    return f"{x:30.20f}"


class ChannelRun:
    """One run directory driven like for/diablo.f."""

    def __init__(self, directory: str, precision: str = "fp64", gy_1: Optional[np.ndarray] = None,
                 log=print):
        self.dir = directory
        nx, ny, nz, n_th = case.read_grid_def(directory)
        self.cfg = case.read_decks(directory, nx, ny, nz, n_th)
        self.gy = case.grid_gy(self.cfg) if gy_1 is None else np.ascontiguousarray(gy_1, dtype=np.float64)
        self.chan = GpuChannel(self.cfg, self.gy, precision=precision)
        self.time_step = 0
        self.log = log
        gyf = 0.5 * (self.gy[:-1] + self.gy[1:])       # GYF(1..NY)
        self.gyf = gyf

    # -- INITIALIZE (for/diablo_io.f:140-176) --
    def initialize(self):
        cfg, g = self.cfg, self.chan
        if cfg.les:
            g.set_les(True, model=False)      # NU_T = 0 until les_chan first runs
        if cfg.create_new_flow:
            g.initialize(*case.initial_physical(cfg, self.gy))
            self.log("A new flowfield has been created")
        else:
            restart.load_flow(os.path.join(self.dir, "start.npz"), g)
            self.log("Done reading flow")
        if cfg.reset_time or cfg.create_new_flow:
            self.time_step = 0
            g.time = 0.0
        # both branches write a step-0 record: CREATE_FLOW_CHAN ends with
        # SAVE_STATS_CHAN(.FALSE.) (for/set_ics.f:181-182), READ_FLOW's branch
        # calls SAVE_STATS(.FALSE.) (for/diablo_io.f:173)
        self.save_stats(final=False)

    # -- SAVE_STATS: the profiles in the reference's text files --
    def save_stats(self, final: bool):
        g, cfg = self.chan, self.cfg
        st, ub = g.stats_chan()
        eps = g.tke_chan()
        ny = cfg.ny
        cu2 = g.get_field(CU2)
        vme = case.fourier_view(cu2)[:, 0, 0].real                    # vme(0..NY+1)
        head = f" {self.time_step} {g.time!r} {g.delta_t!r}\n {ub!r}\n"
        with open(os.path.join(self.dir, "mean.txt"), "a") as f:
            f.write(head)
            for j in range(1, ny + 1):
                row = [st["ume"][j - 1], 0.5 * (vme[j + 1] + vme[j]), st["wme"][j - 1], st["urms"][j - 1],
                       st["vrms"][j - 1], st["wrms"][j - 1], st["uv"][j - 1], st["uw"][j - 1], st["wv"][j - 1],
                       st["dudy"][j - 1], st["dwdy"][j - 1], st["cp"][j - 1], st["shear"][j - 1],
                       st["omega_x"][j - 1], st["omega_y"][j - 1], st["omega_z"][j - 1]]
                f.write(f"{j:3d} " + " ".join(_fmt(v) for v in [self.gyf[j - 1]] + row) + " \n")
        if cfg.n_th > 0:
            with open(os.path.join(self.dir, "mean_th.txt"), "a") as f:
                f.write(head)
                for j in range(1, ny + 1):
                    row = [st[k][j - 1] for k in ("thme", "dthdy", "thrms", "thv", "pe_diss")]
                    f.write(f"{j:3d} " + " ".join(_fmt(v) for v in [self.gyf[j - 1]] + row) + " \n")
        with open(os.path.join(self.dir, "tke.txt"), "a") as f:
            f.write(f" {self.time_step} {g.time!r} {g.delta_t!r}\n")
            for j in range(2, ny):
                f.write(f"{j:3d} {self.gyf[j - 1]:20.9f} {eps[j - 1]:20.9f} \n")
        if cfg.les:
            # tkebudget_chan_les's text output (for/save_stats.f:940-958)
            eps_sgs = g.tke_les()
            with open(os.path.join(self.dir, "tke_les.txt"), "a") as f:
                f.write(f" {self.time_step} {g.time!r} {g.delta_t!r}\n")
                for j in range(1, ny + 1):
                    f.write(f"{j:3d} {_fmt(self.gyf[j - 1])} {_fmt(eps_sgs[j - 1])} \n")

    # -- SAVE_FLOW (for/diablo_io.f:305-341) --
    def save_flow(self, final: bool):
        name = "end.npz" if final else "out.npz"
        restart.save_flow(os.path.join(self.dir, name), self.chan, save_pressure=final)

    # -- the time loop (for/diablo.f:62-127) --
    def run(self, n_steps: Optional[int] = None) -> dict:
        cfg, g = self.cfg, self.chan
        n = cfg.n_steps if n_steps is None else n_steps
        start = _time.time()
        stop = None
        done = 0
        first = self.time_step + 1
        for ts in range(first, first + n):
            self.time_step = ts
            if cfg.les:
                # the model waits out a spin-up of 100 steps on a new flow
                # (for/channel.f:303)
                g.set_les(True, (not cfg.create_new_flow) or ts > 100)
            for k in (1, 2, 3):
                cfl_now = cfg.variable_dt and k == 3 and ts % cfg.update_dt == 0
                if cfl_now:
                    g._chk(g.L.dgpu_courant_request(g.h), "dgpu_courant_request")
                g.rk_chan_1(k)
                if cfl_now:
                    g.courant(cfg.cfl)
            if interval_due(cfg.save_stats_int, ts, g.time, g.delta_t):
                self.save_stats(final=False)
            if interval_due(cfg.save_flow_int, ts, g.time, g.delta_t):
                self.save_flow(final=False)
            g.time += g.delta_t
            done += 1
            stop = end_run(start, cfg.time_limit, self.dir)
            if stop:
                self.log(f" STOP because of {stop}!")
                break
        g.synchronize()
        elapsed = _time.time() - start
        # TIME_STEP = TIME_STEP - 1 before the final files (for/diablo.f:130):
        # after a completed DO loop that is the last step; after END_RUN's
        # EXIT the loop variable still holds the last step, so the final
        # records carry the step before it, as the reference's do
        if stop:
            self.time_step -= 1
        self.save_flow(final=True)
        self.save_stats(final=True)
        return {"steps": done, "time_step": self.time_step, "time": g.time, "delta_t": g.delta_t,
                "elapsed_s": elapsed, "stopped": stop}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("directory")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed", "mixed_f64state"))
    ap.add_argument("--grid", help="text file with GY(1..NY+1) (default: the tanh grid of the decks)")
    a = ap.parse_args(argv)
    gy = np.loadtxt(a.grid) if a.grid else None
    r = ChannelRun(a.directory, a.precision, gy)
    r.initialize()
    out = r.run()
    print(f"Elapsed Time (sec): {out['elapsed_s']}")
    print(f"Seconds per Iteration: {out['elapsed_s'] / max(out['steps'], 1)}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
