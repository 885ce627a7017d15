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

The surface force is discontinuous: it switches on at |grad c| =
surface_normal.  A particle whose FP64 |grad c| lies within the bar of that
threshold (the oracle's `branch` mask) may take either side in FP32, so its
acceleration, velocity and position are measured against the nearer of the
two outcomes the oracle returns.
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
    branch = ref.get("branch")

    def nearer(got, key):
        """Per-particle distance to the oracle's outcome, or to the nearer of
        its two outcomes on the surface-force threshold."""
        d = np.linalg.norm(got - ref[key][:, :got.shape[1]], axis=1)
        if branch is not None and branch.any():
            alt = np.linalg.norm(got - ref[key + "_alt"][:, :got.shape[1]], axis=1)
            d = np.where(branch, np.minimum(d, alt), d)
        return d

    out["acceleration"] = float(np.max(nearer(acc1, "acc") / ascale))
    x_ref = ref["pos"][:, :3]
    out["position"] = float(np.max(nearer(pos1[:, :3], "pos") / np.linalg.norm(x_ref, axis=1)))
    v_ref = ref["vel"]
    vscale = np.maximum(np.linalg.norm(v_ref, axis=1), dt * (ref["acc_scale"] / rho + g))
    out["velocity"] = float(np.max(nearer(vel1, "vel") / vscale))
    return out
