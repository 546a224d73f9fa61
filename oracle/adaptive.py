"""CPU restatement of libetd's adaptive ETD driver (etd_run_adaptive with the
plain controller) on the FP64 oracle -- test infrastructure only: the GPU
driver's accept / reject decisions and accepted states are checked against it.

Attempt from u_n with step dt (ETD1): one shared N(u_n) gives u_full = u_n +
dt and the first half step; a second half step gives u_half;
err = max |u_half - u_full| over all fields; with extrapolation the candidate
is 2 u_half - u_full.  Accept when err <= tol and the check_soln_new energy
(AC + CH) has not increased (slack 1e-12 relative); the next proposal is
dt * clamp(safety sqrt(tol / err), shrink, grow); a rejection for err
shrinks by clamp(safety sqrt(tol / err), shrink, 1), for energy by 1/2.
The last step is clipped to t_end; a step at dt_min is accepted when its
error estimate and energy are finite.  A non-finite candidate (NaN or
infinite err or energy) is rejected with the factor shrink; at dt_min the run
raises NonFinite with the fields restored to the last accepted state, and a
non-finite initial energy raises NonFinite before any step."""
import math

import numpy as np


class NonFinite(ArithmeticError):
    pass


def nanmax(values):
    """max that returns NaN when any value is NaN (numpy's max does too)."""
    return float(np.max(values))


def energy(o):
    st, _ = o.diag()
    e = float(st[0, 3]) if o.nac > 0 else 0.0
    for q in range(o.nch):
        e += float(st[o.nac + q, 3])
    return e


def fields(o):
    return np.stack([o.get_field(v) for v in range(o.V)]).reshape(o.V, -1)


def set_fields(o, arr):
    for v in range(o.V):
        o.set_field(v, arr[v].reshape(o.n, o.n, o.n))


def run(o, t_end, dt0, dt_min, dt_max, tol, safety=0.9, grow=2.0, shrink=0.2, extrapolate=True, energy_test=True,
        max_attempts=100000):
    """Returns (t, log rows [(t, dt, err, E)], counts dict); the oracle's fields
    end at the accepted state."""
    clamp = lambda x, lo, hi: min(max(x, lo), hi)
    t = 0.0
    E = energy(o)
    if not math.isfinite(E):
        raise NonFinite(f"the energy of the initial state is not finite ({E})")
    dt = dt0
    plain_next = dt0
    rej_since = 0
    accepted = rej_err = rej_en = attempts = 0
    rows = []
    while t < t_end * (1.0 - 1e-12) and attempts < max_attempts:
        if rej_since == 0 and accepted + attempts > 0:
            dt = plain_next
        dt = clamp(dt, dt_min, dt_max)
        dt = min(dt, t_end - t)
        attempts += 1
        u0 = fields(o)
        full = o.step_pair(dt)
        o.set_dt(0.5 * dt)
        o.step()
        half = fields(o)
        err = nanmax(np.abs(half - full))
        cand = 2.0 * half - full if extrapolate else half
        set_fields(o, cand)
        Enew = energy(o)
        finite = math.isfinite(err) and math.isfinite(Enew)
        fac = shrink if not finite else (safety * math.sqrt(tol / err) if err > 0.0 else grow)
        err_ok = err <= tol and finite
        en_ok = finite and ((not energy_test) or Enew <= E + 1e-12 * abs(E))
        at_min = dt <= dt_min * (1.0 + 1e-12)
        if at_min and not finite:
            set_fields(o, u0)
            raise NonFinite(f"the candidate at dt_min = {dt} is not finite (err {err}, energy {Enew}) at t = {t}")
        forced = at_min and finite
        if (err_ok and en_ok) or forced:
            t += dt
            accepted += 1
            rows.append((t, dt, err, Enew))
            E = Enew
            plain_next = dt * clamp(fac, shrink, grow)
            rej_since = 0
        else:
            set_fields(o, u0)
            if not err_ok:
                rej_err += 1
                dt = dt * clamp(fac, shrink, 1.0)
            else:
                rej_en += 1
                dt = 0.5 * dt
            rej_since += 1
    return t, rows, dict(accepted=accepted, rejected_err=rej_err, rejected_energy=rej_en, attempts=attempts)
