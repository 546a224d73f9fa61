"""Side-by-side runs of two step configurations from identical initial fields
(the mixed-precision accuracy report): every `every` steps, the largest
pointwise difference of the AC order parameters (max|d eta|), of the CH
concentrations (max|d c|), and the relative difference of the AC energy of
check_soln_new (computed in FP64 on both)."""
import numpy as np

from . import Etd
from . import workloads as W


def ac_energy(stats, nac):
    """check_soln_new's eng column of the AC fields (the same for every AC field)."""
    return float(stats[4 * 0 + 3]) if nac > 0 else 0.0


def compare(prm_ref, prm_test, steps, every=100, fields=None, log=print, keep_ref=False):
    """Runs both configurations for `steps` steps of their own dt (the two dt
    must be equal) and returns a list of dicts {step, t, max_deta, max_dc,
    rel_energy, energy_ref, energy_test}; with keep_ref, (rows, the reference's
    final fields, its final check_soln_new statistics)."""
    assert abs(prm_ref["dt"] - prm_test["dt"]) == 0.0
    a = Etd.from_dict(prm_ref)
    b = Etd.from_dict(prm_test)
    V = prm_ref["nac"] + prm_ref["nch"]
    init = fields if fields is not None else W.initial_fields(prm_ref)
    for v, f in enumerate(init):
        a.set_field(v, f)
        b.set_field(v, f)
    rows = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        a.step(k)
        b.step(k)
        done += k
        sa, _ = a.diag()
        sb, _ = b.diag()
        deta, dc = 0.0, 0.0
        for v in range(V):
            d = float(np.abs(a.get_field(v) - b.get_field(v)).max())
            if v < prm_ref["nac"]:
                deta = max(deta, d)
            else:
                dc = max(dc, d)
        ea, eb = ac_energy(sa.reshape(-1), prm_ref["nac"]), ac_energy(sb.reshape(-1), prm_ref["nac"])
        row = dict(step=done, t=done * prm_ref["dt"], max_deta=deta, max_dc=dc,
                   rel_energy=abs(eb - ea) / max(abs(ea), 1e-300),
                   energy_ref=ea, energy_test=eb)
        rows.append(row)
        if log:
            log(row)
    ref = [a.get_field(v) for v in range(V)] if keep_ref else None
    st_ref = a.diag()[0] if keep_ref else None
    a.close()
    b.close()
    return (rows, ref, st_ref) if keep_ref else rows
