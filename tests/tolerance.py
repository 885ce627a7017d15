"""Per-quantity error measures of one GPU step against the FP64 oracle.

Every error is relative to the magnitude the FP32 result's rounding error is
proportional to:
  inv_density  |w - 1/rho*| * rho*
  pressure     |p - p*| / max(|p*|, gas (rho*/rho0)^7): the state equation
               (rho/rho0)^7 - 1 cancels near rho0, so an FP32 p carries an
               absolute error relative to its uncancelled part
  acceleration |a - a*| / max(|a*|, S) with S the oracle's sum of term
               magnitudes (pressure terms taken with the uncancelled p)
  position     |x - x*| / |x*|
  velocity     |v - v*| / max(|v*|, dt (w* S + |g|))
"""
import numpy as np

REL_TOL = 1e-5


def step_errors(pos1, vel1, ev1, acc1, ref, op):
    rho = ref["rho"]
    pr = ref["pressure"]
    gas = float(op.gas_constant)
    rho0 = float(op.rest_density)
    dt = float(op.time_step)
    g = np.linalg.norm(np.array(op.gravity[:], np.float64))
    out = {}
    out["inv_density"] = float(np.max(np.abs(pos1[:, 3].astype(np.float64) - 1.0 / rho) * rho))
    pscale = np.maximum(np.abs(pr), gas * (rho / rho0) ** 7)
    out["pressure"] = float(np.max(np.abs(ev1[:, 3] - pr) / pscale))
    a_ref = ref["acc"]
    ascale = np.maximum(np.linalg.norm(a_ref, axis=1), ref["acc_scale"])
    ascale = np.maximum(ascale, 1e-30)
    out["acceleration"] = float(np.max(np.linalg.norm(acc1 - a_ref, axis=1) / ascale))
    x_ref = ref["pos"][:, :3]
    out["position"] = float(np.max(np.linalg.norm(pos1[:, :3] - x_ref, axis=1) /
                                   np.linalg.norm(x_ref, axis=1)))
    v_ref = ref["vel"]
    vscale = np.maximum(np.linalg.norm(v_ref, axis=1), dt * (ref["acc_scale"] / rho + g))
    out["velocity"] = float(np.max(np.linalg.norm(vel1 - v_ref, axis=1) / vscale))
    return out
