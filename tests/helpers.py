"""Shared set-up for the parity tests: descriptors for the golden cases."""
import math

import numpy as np

from pysic_amd.interactions import (COMB, POTENTIAL_CODES, POTENTIAL_TYPES,  # noqa: F401
                                    TERSOFF_SI, BondOrderParameters, Coordinator, Potential,
                                    ProductPotential, build_descriptors, comb_bend_terms,
                                    comb_charge_model, tersoff_si)


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


def engine_pair_set(eng):
    """The engine's neighbor lists as sorted (i, j, o1, o2, o3) rows."""
    counts, idx, off = eng.neighbor_lists()
    i = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    rows = np.column_stack([i, idx.astype(np.int64), off.astype(np.int64).reshape(-1, 3)])
    return rows[np.lexsort(rows.T[::-1])]


def assert_same_pairs(eng, nl):
    """Neighbor lists bit-exact as sorted (i, j, offset) sets against the
    oracle's lists nl (Core.f90:8148-8280): the whole list radius, the
    shell between the cutoff and cutoff + skin included."""
    got, want = engine_pair_set(eng), nl.pair_set()
    print("  pairs: %d (oracle %d)" % (len(got), len(want)), flush=True)
    assert got.shape == want.shape and np.array_equal(got, want)


def rel_maxnorm(a, b, floor=1e-300):
    """max|a-b| / max(max|b|, floor)"""
    return float(np.abs(a - b).max() / max(np.abs(b).max(), floor))


# A two-species model with every pair type of the pair kernel, a product and
# the one-body 'constant' and 'force' terms.  Table 7 is an exponential
# tabulated on 60 points over 7.5 A (column 2: derivative times the range).
MIXTURE_TABLE_ID = 7


def mixture_table():
    x = np.linspace(0.0, 7.5, 60)
    return np.column_stack([0.3 * np.exp(-x / 2.0), -0.3 / 2.0 * np.exp(-x / 2.0) * 7.5])


def mixture_model():
    P = Potential
    return [
        P("LJ", symbols=[["Ar", "Ar"]], parameters=[0.04, 3.4], cutoff=8.0, cutoff_margin=1.0),
        P("Buckingham", symbols=[["Ar", "Kr"]], parameters=[500.0, 5e3, 0.35], cutoff=7.0),
        P("spring", symbols=[["Kr", "Kr"]], parameters=[0.5, 4.0], cutoff=6.0,
          cutoff_margin=0.5),
        P("power", symbols=[["Ar", "Kr"]], parameters=[0.01, 3.5, 7.5], cutoff=7.0,
          cutoff_margin=1.0),
        P("shift_power", symbols=[["Ar", "Ar"]], parameters=[0.02, 7.0, 3.0, 2.0], cutoff=6.5),
        P("step", symbols=[["Kr", "Kr"]], parameters=[0.1, 0.0, 3.5, 5.0], cutoff=6.0),
        P("tabulated", symbols=[["Ar", "Kr"]], parameters=[MIXTURE_TABLE_ID, 7.5, 1.3],
          cutoff=7.4, cutoff_margin=0.5),
        ProductPotential([
            P("LJ", symbols=[["Kr", "Kr"]], parameters=[0.05, 3.6], cutoff=7.0,
              cutoff_margin=1.0),
            P("power", symbols=[["Kr", "Kr"]], parameters=[1.0, 4.0, 1.0]),
            P("exponential", symbols=[["Kr", "Kr"]], parameters=[2.0, 0.3]),
        ]),
        P("constant", symbols=[["Ar"]], parameters=[-0.3]),
        P("force", symbols=[["Kr"]], parameters=[0.01, -0.02, 0.03]),
    ]


def comb_asymmetric_terms():
    """Charge terms whose value depends on which atom of a pair comes
    first: a Si-Si charge_exp with unequal halves, a product written for
    (Si, O) and (O, O) whose charge_abs multiplier keeps the first tuple's
    original targets, so on O-O pairs its halves always swap, charge_pair
    q1^2 q2 on Si-O and O-O, and a power x charge_pair q1 q2^3 product on
    Si-Si."""
    C = COMB
    P = Potential
    si_si = P("charge_exp", symbols=[["Si", "Si"]], cutoff=3.0, cutoff_margin=0.3,
              parameters=[5.0, C["Si"]["rmax"], C["Si"]["rmin"], C["Si"]["qmax"], C["Si"]["qmin"],
                          C["O"]["rmax"], C["O"]["rmin"], C["O"]["qmax"], C["O"]["qmin"],
                          0.9, 1.6])
    mixed = ProductPotential([
        P("exponential", symbols=[["Si", "O"], ["O", "O"]], parameters=[3.0, 1.1], cutoff=3.05,
          cutoff_margin=0.5),
        P("charge_abs", symbols=[["Si", "O"], ["O", "O"]],
          parameters=[1.2, -0.05, 0.3, 2.0, 0.8, -0.02, -0.7, 3.0])])
    qpair = P("charge_pair", symbols=[["Si", "O"], ["O", "O"]], parameters=[0.7, 2.0, 1.0],
              cutoff=3.05, cutoff_margin=0.5)
    qprod = ProductPotential([
        P("power", symbols=[["Si", "Si"]], parameters=[1.0, 2.5, 1.0], cutoff=3.0,
          cutoff_margin=0.4),
        P("charge_pair", symbols=[["Si", "Si"]], parameters=[-0.3, 1.0, 3.0])])
    return [si_si, mixed, qpair, qprod]


def comb_charges(symbols, seed):
    """Random charges in the ranges COMB SiO2 visits: Si in [0.5, 2.2],
    O in [-1.3, -0.2] (not neutral: no Ewald term here)."""
    rng = np.random.default_rng(seed)
    return np.array([rng.uniform(0.5, 2.2) if s == "Si" else rng.uniform(-1.3, -0.2)
                     for s in symbols])


def ref_entries(pd, names):
    """RefCore potential entries (kind, params, cutoff, soft, symbols, group,
    multipliers, original symbols) from engine descriptor rows (12 doubles
    each; a link-2 row continues the row before it with the parameters 5-13
    and the species of the original first target)."""
    inv = {v: k for k, v in POTENTIAL_CODES.items()}
    out = []
    last = None
    for r in np.asarray(pd, dtype=np.float64).reshape(-1, 12):
        kind = int(r[0])
        npar = len(POTENTIAL_TYPES[inv[kind]][1])
        if r[11] == 2 and kind == POTENTIAL_CODES["bond_bend"]:
            # the third target and the original targets
            syms = last[2] + (names[int(r[1])],)
            orig = tuple("Xx" if o < 0 else names[int(o)] for o in r[2:5])
            e = out[-1]
            out[-1] = e[:4] + (syms,) + e[5:7] + (orig,)
            continue
        if r[11] == 2:
            params = last[1] + list(r[2:11])
            o1 = "Xx" if r[1] < 0 else names[int(r[1])]
            syms = last[2]
            orig = (o1, syms[1] if syms[0] == o1 else syms[0])
            if last[0] == 0:
                e = out[-1]
                out[-1] = e[:1] + (params[:npar],) + e[2:7] + (orig,)
            else:
                out[-1][6][-1] = (kind, params[:npar], orig)
            continue
        params = [r[4], r[5], r[9], r[10]]
        nt = int(r[1])
        syms = tuple(names[int(e)] for e in r[2:2 + min(nt, 2)])
        last = (int(r[11]), params, syms)
        if r[11] != 0:
            out[-1][6].append((kind, params[:npar]))
            continue
        out.append((kind, params[:npar], float(r[6]), float(r[7]), syms, int(r[8]), [], syms))
    return out


def ref_bonds(bd, names):
    """RefCore 'tersoff' entries (params7, cutoff, soft, symbols, group) from
    engine bond-order descriptor rows (14 doubles each)."""
    out = []
    for r in np.asarray(bd, dtype=np.float64).reshape(-1, 14):
        out.append((list(r[3:10]), float(r[10]), float(r[11]),
                    tuple(names[int(e)] for e in r[0:3]), int(r[12])))
    return out


# Two-species Tersoff (Si, C; J. Tersoff, PRB 39, 5566 (1989) element
# parameters).  Pair terms mix as A = sqrt(A_i A_j), lambda = (l_i + l_j)/2,
# R = sqrt(R_i R_j), S = sqrt(S_i S_j); the bond-order factor of triplet
# (i, j, k) takes centre i's beta, eta, c, d, h and the (i, k) cutoffs.
SIC = {"Si": dict(A=1830.8, l1=2.4799, B=471.18, l2=1.7322, beta=1.1e-6, n=0.78734,
                  c=100390.0, d=16.217, h=-0.59825, R=2.7, S=3.0),
       "C": dict(A=1393.6, l1=3.4879, B=346.74, l2=2.2119, beta=1.5724e-7, n=0.72751,
                 c=38049.0, d=4.3484, h=-0.57058, R=1.8, S=2.1)}


def sic_model():
    els = ["C", "Si"]

    def rs(i, j):
        return (math.sqrt(SIC[i]["R"] * SIC[j]["R"]), math.sqrt(SIC[i]["S"] * SIC[j]["S"]))

    bonds = []
    for i in els:
        p = SIC[i]
        for j in els:
            for k in els:
                R, S = rs(i, k)
                bonds.append(BondOrderParameters(
                    "tersoff", cutoff=S, cutoff_margin=S - R,
                    parameters=[[p["beta"], p["n"]], [0.0, p["c"], p["d"], p["h"], 3.0]],
                    symbols=[[i, j, k]]))
    coord = Coordinator(bonds)
    pots = []
    for a, i in enumerate(els):
        for j in els[a:]:
            R, S = rs(i, j)
            A = math.sqrt(SIC[i]["A"] * SIC[j]["A"])
            B = math.sqrt(SIC[i]["B"] * SIC[j]["B"])
            l1 = 0.5 * (SIC[i]["l1"] + SIC[j]["l1"])
            l2 = 0.5 * (SIC[i]["l2"] + SIC[j]["l2"])
            pots.append(Potential("exponential", symbols=[[i, j]], parameters=[A, l1],
                                  cutoff=S, cutoff_margin=S - R))
            pots.append(Potential("exponential", symbols=[[i, j]], parameters=[-B, l2],
                                  cutoff=S, cutoff_margin=S - R, coordinator=coord))
    return pots
