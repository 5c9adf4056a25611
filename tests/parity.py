"""Shared parity measures of the GPU tests (imported by them; no tests here)."""
import numpy as np


def max_rel(a, ref, floor=1e-20):
    """max |a - ref| / ref over entries of ref above floor (abundances per H)"""
    m = ref > floor
    if not np.any(m):
        return 0.0
    return float(np.max(np.abs(a[m] - ref[m]) / ref[m]))


def charge_sum(ab, charges):
    """x_e implied by the returned abundances: sum of charge * x over the
    equations (the slots before the electron)"""
    q = np.asarray(charges, dtype=np.float64)
    return ab[:, : q.size] @ q


def electron_bound(got, ref, charges, bar=1e-4):
    """The electron slot under calculate_abundances.c:317 semantics (x_e of
    the last RHS evaluation, at a Newton iterate of the step that passed
    t_end).  Each side's value differs from the charge sum of its own
    returned abundances by its own overshoot; the species agree to `bar`.
    So |x_e,gpu - x_e,oracle| <= bar * q.y_oracle + |x_e,gpu - q.y_gpu| +
    |x_e,oracle - q.y_oracle|, every term measured, none set by hand.
    Returns (ratio of the difference to that bound per point, |diff|/x_e)."""
    qg, qo = charge_sum(got, charges), charge_sum(ref, charges)
    xg, xo = got[:, -1], ref[:, -1]
    bound = bar * np.abs(qo) + np.abs(xg - qg) + np.abs(xo - qo) + 1e-300
    return np.abs(xg - xo) / bound, np.abs(xg - xo) / np.abs(xo)


def composition_matrix(net):
    """E[element, species] of the equations' species (element counts parsed
    from the species names), and which species could be parsed"""
    import test_netgen
    els = sorted({e for s in net.species[:-1] for e in (test_netgen.composition(s.name) or {})})
    E = np.zeros((len(els), net.neq))
    ok = np.ones(net.neq, bool)
    for i, s in enumerate(net.species[:-1]):
        c = test_netgen.composition(s.name) if not s.name.startswith("PAH") else None
        if c is None:
            ok[i] = False
            continue
        for e, k in c.items():
            E[els.index(e), i] = k
    return E, ok
