"""Time-to-permeability on the GPU engine.

Runs a single-phase lattice to the reference's convergence criterion and
returns the Darcy permeability, everything on the device except a few FP64
scalars per report:

* the stop rule is Simulator::updateConvergence (solver/src/lbm/simulator.cpp:
  248-325): at every report, the velocity residual sqrt(sum (u-u_prev)^2 /
  sum u^2) and the adjacent (t-1) residual must both be below the tolerance
  (or their RMS change <= 1e-12), together with the pressure-based K of
  SCSP::report (scsp.cpp:343-368) when a pressure drop exists, for 3
  consecutive reports;
* the permeability is the reference study scripts' Darcy estimate k = nu <u_x> / F with the
  mean over all voxels, solids counted as zero (scripts/rock_workflow.py:63-69,
  scripts/porous_media.py:20-44).
"""
import math
import time

from .engine import STEP_ADJACENT, STEP_FIELDS


def darcy_k(engine, axis, force, tau):
    """nu <u_axis> / F over all voxels from the engine's current fields."""
    n_axis = engine.size[axis]
    r = engine.report(axis, 0, n_axis - 1, convergence=False, adjacent=False)
    nu = (tau - 0.5) / 3.0
    return nu * (r.vel_range / engine.N) / force


def converge(engine, axis, first, last, tol, interval, max_steps, force, tau, kpar=None, log=None):
    """Advance an initialised engine until the reference's criterion fires.

    Returns dict(steps, converged, K, seconds, reports)."""
    t0 = time.perf_counter()
    stable = 0
    k_old = 0.0
    has_prev = False
    reports = []
    t = engine.t
    converged = False
    while t < max_steps:
        n = min(interval, max_steps - t)
        if n > 2:
            engine.advance(n - 2)
        if n > 1:
            engine.step(STEP_ADJACENT)
        # the adjacent history is double-buffered, so this step may also record
        # its velocity for a report at the very next step
        engine.step(STEP_FIELDS | STEP_ADJACENT)
        t = engine.t
        r = engine.report(axis, first, last, convergence=True, adjacent=True)
        c0, c1, c2, c3 = r.conv
        dP = (r.rho_slice[0][0] / r.count_slice[0] - r.rho_slice[1][0] / r.count_slice[1]) / 3.0
        drop = abs(dP) > 1e-14
        Q = r.vel_range / (last - first + 1)
        K = kpar * Q / dP if (drop and kpar) else 0.0
        errK = abs(K - k_old) / max(abs(K), 1e-30) if drop else 0.0
        k_old = K
        if r.nonfinite:
            raise RuntimeError('Non-finite velocity in convergence report.')
        if has_prev and r.has_previous:
            rms = math.sqrt(c0 / c2)
            res = math.sqrt(c0 / c1) if c1 > 0 else 0.0
            arms = math.sqrt(c3 / c2)
            ares = math.sqrt(c3 / c1) if c1 > 0 else 0.0
            ok = (not drop or errK < tol) and (rms <= 1e-12 or res < tol) and (arms <= 1e-12 or ares < tol)
            stable = stable + 1 if ok else 0
            reports.append(dict(step=t, residual=res, adjacent=ares))
            if log:
                log(f'step {t} residual {res:.3e} adjacent {ares:.3e} stable {stable}')
        has_prev = True
        if stable >= 3:
            converged = True
            break
    engine.synchronize()
    engine.check_status()
    k = darcy_k(engine, axis, force, tau)
    return dict(steps=t, converged=converged, K=k, seconds=time.perf_counter() - t0, reports=reports)
