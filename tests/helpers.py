"""Shared set-up for the parity tests: descriptors for the golden cases."""
import numpy as np

from pysic_amd.interactions import TERSOFF_SI, build_descriptors, tersoff_si


def tersoff_descriptors(ters=None):
    """Descriptor rows of Tersoff Si (species 0), optionally with the
    7 bond parameters [beta, eta, a, c, d, h, mu] replaced."""
    p = None
    if ters is not None:
        t = [float(x) for x in ters]
        p = dict(beta=t[0], eta=t[1], a=t[2], c=t[3], d=t[4], h=t[5], mu=t[6])
    pd, bd, _ = build_descriptors(tersoff_si(p), {"Si": 0})
    return pd, bd


def ewald_descriptors(g):
    """(species, potdesc, ewald8) of an Ewald golden case."""
    symbols = [str(s) for s in g["symbols"]]
    names = sorted(set(symbols))
    sp = np.array([names.index(s) for s in symbols], dtype=np.int32)
    rows = []
    for sym, (chi0, J) in zip(g["charge_self_symbols"], g["charge_self"]):
        s = names.index(str(sym))
        rows.append([2, 1, s, 0, chi0, 1.0, 0.0, 0.0, -1, 0])
        rows.append([2, 1, s, 0, 0.5 * J, 2.0, 0.0, 0.0, -1, 0])
    pd = np.array(rows, dtype=np.float64).reshape(-1, 10)
    rc, kc, sigma, eps = g["ewald"]
    K = g["kcells"]
    ew = np.array([1, rc, kc, K[0], K[1], K[2], sigma, eps], dtype=np.float64)
    return sp, pd, ew


def rel_maxnorm(a, b, floor=1e-300):
    """max|a-b| / max(max|b|, floor)"""
    return float(np.abs(a - b).max() / max(np.abs(b).max(), floor))
