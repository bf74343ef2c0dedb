"""Writes the golden vectors in tests/golden/ from the reference Fortran core.

Runs only in a container that has the reference sources: build the checker
first with oracle/build_ref.sh, then

    python tests/golden/make_golden.py          # every fixture
    python tests/golden/make_golden.py pairs    # the pair-potential fixtures
    python tests/golden/make_golden.py comb     # COMB with its coordinators

Every fixture is data only (inputs and the reference's outputs).  The
system set-up replays pysic/calculator.py:1595-1637 through the reference's
own f2py interface routines (see oracle/ref_driver.f90).
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from pysic_amd import structures as st  # noqa: E402
from pysic_amd.interactions import build_descriptors  # noqa: E402

T = dict(A=1830.8, l1=2.4799, B=471.18, l2=1.7322)
TERS = [1.1e-6, 0.78734, 0.0, 100390.0, 16.217, -0.59825, 3.0]


def tersoff_setup(struct, ters=TERS, cut=3.0, margin=0.3, skin=0.5):
    pots = [(1, [T["A"], T["l1"]], cut, cut - margin, ("Si", "Si"), -1),
            (1, [-T["B"], T["l2"]], cut, cut - margin, ("Si", "Si"), 0)]
    bonds = [(ters, cut, cut - margin, ("Si", "Si", "Si"), 0)]
    return pots, bonds, np.full(len(struct), cut + skin)


def kcells(cell, kc):
    rec = 2 * math.pi * np.linalg.inv(cell).T
    vol = rec[0] @ np.cross(rec[1], rec[2])
    return [int(kc * np.linalg.norm(np.cross(rec[(a + 1) % 3], rec[(a + 2) % 3])) / vol + 0.5)
            for a in range(3)]


def run_tersoff(name, struct, ters=TERS, with_lists=True):
    R = O.RefCore()
    pots, bonds, cuts = tersoff_setup(struct, ters)
    R.setup(struct.symbols, struct.positions, struct.cell, struct.pbc, potentials=pots,
            tersoff=bonds, ngroups=1, cutoffs=cuts)
    e = R.energy()
    f, s = R.forces()
    out = dict(positions=struct.positions, cell=struct.cell, pbc=struct.pbc.astype(np.int32),
               cutoffs=cuts, tersoff=np.array(ters), energy=e, forces=f, stress=s)
    if with_lists:
        nl = R.neighbor_lists()
        out.update(nl_counts=nl.counts, nl_idx=nl.idx, nl_off=nl.off)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, len(struct), "E = %.12f" % e)


def run_ewald(name, struct, rc, accuracy_scale=3.0, charge_self=None, scaler=None):
    sigma = rc / accuracy_scale
    kc = accuracy_scale / sigma
    eps = 0.00552635
    K = kcells(struct.cell, kc)
    pots = []
    if charge_self:
        for sym, (chi0, J) in charge_self.items():
            pots.append((2, [chi0, 1.0], 0.0, 0.0, (sym,), -1))
            pots.append((2, [0.5 * J, 2.0], 0.0, 0.0, (sym,), -1))
    R = O.RefCore()
    R.setup(struct.symbols, struct.positions, struct.cell, struct.pbc, charges=struct.charges,
            potentials=pots, tersoff=[], ngroups=0, cutoffs=np.full(len(struct), rc + 0.5),
            ewald=(rc, kc, K, sigma, eps), scaler=scaler)
    e = R.energy()
    f, s = R.forces()
    x = R.electronegativities()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), positions=struct.positions,
                        cell=struct.cell, charges=struct.charges,
                        symbols=np.array(struct.symbols), ewald=np.array([rc, kc, sigma, eps]),
                        kcells=np.array(K), energy=e, forces=f, stress=s, enegs=x,
                        scaler=np.ones(len(struct)) if scaler is None else np.asarray(scaler),
                        charge_self=np.array([[chi0, J] for chi0, J in (charge_self or {}).values()])
                        if charge_self else np.zeros((0, 2)),
                        charge_self_symbols=np.array(list((charge_self or {}).keys())))
    print(name, len(struct), "E = %.12f" % e, "K", K)


def sheared_ewald():
    """Ewald in a triclinic cell: rocksalt 2x2x2 with displacements, cell
    and positions sheared by the same matrix (same fractional coordinates)."""
    s = st.rocksalt((2, 2, 2), sigma=0.08, seed=6)
    shear = np.array([[1.0, 0.0, 0.0], [0.2, 1.0, 0.0], [-0.1, 0.15, 1.0]])
    s.cell = s.cell @ shear
    s.positions = s.positions @ shear
    run_ewald("nacl_2x2x2_sheared_rc5", s, rc=5.0)


def scaled_ewald():
    """Ewald in an anisotropic cell (K = 6, 3, 2) with per-atom charge
    scalers uniform in [0.5, 1.5]: the scaled charges are not neutral, so the
    background term enters too."""
    s = st.rocksalt((4, 2, 1), sigma=0.08, seed=9)
    sc = np.random.default_rng(10).uniform(0.5, 1.5, len(s))
    run_ewald("nacl_4x2x1_scaled_rc5", s, rc=5.0, scaler=sc)


def sic_mixed():
    """Mixed Si/C Tersoff: 192 atoms (2:1 Si:C) by random addition with a
    1.7 A minimum separation at 0.05 atoms/A^3, so Si-Si, Si-C and C-C pairs
    all fall inside their cutoffs."""
    from helpers import ref_entries, sic_model
    s = st.random_mixture(192, (192 / 0.05) ** (1.0 / 3.0), 1.7, ["Si", "Si", "C"], seed=21)
    names = sorted(set(s.symbols))
    pd, bd, ng = build_descriptors(sic_model(), {x: k for k, x in enumerate(names)})
    tersoff = [(list(r[3:10]), float(r[10]), float(r[11]),
                tuple(names[int(e)] for e in r[0:3]), int(r[12])) for r in bd]
    cuts = np.full(len(s), 3.5)
    R = O.RefCore()
    R.setup(s.symbols, s.positions, s.cell, s.pbc, potentials=ref_entries(pd, names),
            tersoff=tersoff, ngroups=ng, cutoffs=cuts)
    e = R.energy()
    f, st6 = R.forces()
    np.savez_compressed(os.path.join(HERE, "sic_mixed_192.npz"), positions=s.positions,
                        cell=s.cell, pbc=s.pbc.astype(np.int32), symbols=np.array(s.symbols),
                        cutoffs=cuts, potdesc=pd, bonddesc=bd, energy=e, forces=f, stress=st6)
    print("sic_mixed_192", len(s), "E = %.12f" % e, "groups", ng, "bond rows", len(bd))


def run_pairs(name, struct, model, tables=None, skin=0.5):
    """A pair-potential case: descriptor rows (engine format) in, the
    reference's energy, forces and stress out."""
    from helpers import ref_entries
    names = sorted(set(struct.symbols))
    pd, bd, _ = build_descriptors(model, {s: k for k, s in enumerate(names)})
    sp = np.array([names.index(x) for x in struct.symbols], dtype=np.int32)
    rc = max(p.get_cutoff() for p in model)
    cuts = np.full(len(struct), rc + skin)
    R = O.RefCore()
    R.setup(struct.symbols, struct.positions, struct.cell, struct.pbc,
            potentials=ref_entries(pd, names), cutoffs=cuts, tables=tables)
    e = R.energy()
    f, s = R.forces()
    tid, tab = (next(iter(tables.items())) if tables else (-1, np.zeros((0, 2))))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), positions=struct.positions,
                        cell=struct.cell, pbc=struct.pbc.astype(np.int32), species=sp,
                        symbols=np.array(struct.symbols), potdesc=pd, cutoffs=cuts,
                        table_id=np.array(tid), table=np.asarray(tab), energy=e, forces=f,
                        stress=s)
    print(name, len(struct), "E = %.12f" % e)


def comb_charge_pairs(name="sio2_48_comb_charge_pairs", asymmetric=False, seed=21):
    """COMB's charge-dependent pair terms on amorphous SiO2 (48 atoms, no
    coordinators, random charges): energy, forces, stress and the
    electronegativities -dE/dq; asymmetric adds terms whose value depends
    on which atom of a pair the reference takes first."""
    from helpers import comb_asymmetric_terms, comb_charge_model, comb_charges, ref_entries
    s = st.amorphous_sio2(16, 2.2, seed=seed)
    q = comb_charges(s.symbols, seed=seed + 1)
    names = sorted(set(s.symbols))
    model = comb_charge_model() + (comb_asymmetric_terms() if asymmetric else [])
    pd, _, _ = build_descriptors(model, {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], dtype=np.int32)
    cuts = np.full(len(s), 3.05 + 0.5)
    R = O.RefCore()
    R.setup(s.symbols, s.positions, s.cell, s.pbc, charges=q,
            potentials=ref_entries(pd, names), cutoffs=cuts)
    e = R.energy()
    f, w = R.forces()
    x = R.electronegativities()
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        positions=s.positions, cell=s.cell, pbc=s.pbc.astype(np.int32),
                        species=sp, symbols=np.array(s.symbols), charges=q, potdesc=pd,
                        cutoffs=cuts, energy=e, forces=f, stress=w, enegs=x)
    print(name, len(s), "E = %.12f" % e)


def comb_bend():
    """COMB's bond bending (Si-O-Si, O-Si-O) alone and with the charge
    terms on amorphous SiO2 (48 atoms, random charges)."""
    from helpers import comb_bend_terms, comb_charge_model, comb_charges, ref_entries
    for name, extra in (("sio2_48_comb_bend", []), ("sio2_48_comb_bend_charge", None)):
        s = st.amorphous_sio2(16, 2.2, seed=25)
        q = comb_charges(s.symbols, seed=26)
        names = sorted(set(s.symbols))
        model = comb_bend_terms() + (comb_charge_model() if extra is None else extra)
        pd, _, _ = build_descriptors(model, {x: k for k, x in enumerate(names)})
        sp = np.array([names.index(x) for x in s.symbols], dtype=np.int32)
        cuts = np.full(len(s), 3.05 + 0.5)
        R = O.RefCore()
        R.setup(s.symbols, s.positions, s.cell, s.pbc, charges=q,
                potentials=ref_entries(pd, names), cutoffs=cuts)
        e = R.energy()
        f, w = R.forces()
        x = R.electronegativities()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), positions=s.positions,
                            cell=s.cell, pbc=s.pbc.astype(np.int32), species=sp,
                            symbols=np.array(s.symbols), charges=q, potdesc=pd, cutoffs=cuts,
                            energy=e, forces=f, stress=w, enegs=x)
        print(name, len(s), "E = %.12f" % e)


def comb_full():
    """COMB as the reference's builder makes it (comb.py:630-990): bending,
    charge-dependent repulsion and attraction, each attractive product
    scaled by its coordinator's 'tersoff' bond-order factors in three
    groups, on amorphous SiO2 with random charges: energy, forces, stress
    and electronegativities."""
    from helpers import comb_charges, ref_bonds, ref_entries
    from pysic_amd.interactions import comb_sio2
    for name, n_si, seed in (("sio2_48_comb_full", 16, 27), ("sio2_96_comb_full", 32, 29)):
        s = st.amorphous_sio2(n_si, 2.2, seed=seed)
        q = comb_charges(s.symbols, seed=seed + 1)
        names = sorted(set(s.symbols))
        pd, bd, ng = build_descriptors(comb_sio2(), {x: k for k, x in enumerate(names)})
        sp = np.array([names.index(x) for x in s.symbols], dtype=np.int32)
        cuts = np.full(len(s), 3.05 + 0.5)
        R = O.RefCore()
        R.setup(s.symbols, s.positions, s.cell, s.pbc, charges=q,
                potentials=ref_entries(pd, names), tersoff=ref_bonds(bd, names), ngroups=ng,
                cutoffs=cuts)
        e = R.energy()
        f, w = R.forces()
        x = R.electronegativities()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), positions=s.positions,
                            cell=s.cell, pbc=s.pbc.astype(np.int32), species=sp,
                            symbols=np.array(s.symbols), charges=q, potdesc=pd, bonddesc=bd,
                            cutoffs=cuts, energy=e, forces=f, stress=w, enegs=x)
        print(name, len(s), "E = %.12f" % e)


def pair_cases():
    from helpers import MIXTURE_TABLE_ID, mixture_model, mixture_table
    from pysic_amd.interactions import lj_argon
    # every pair type, a product, one-body terms: Ar/Kr random mixture
    mix = st.random_mixture(120, 18.0, 2.9, ["Kr", "Ar", "Ar"], seed=0)
    run_pairs("pairs_arkr_mixture_120", mix, mixture_model(),
              tables={MIXTURE_TABLE_ID: mixture_table()})
    # argon fcc, LJ at 2.5 sigma (sharp and smoothed cutoff)
    ar = st.fcc((3, 3, 3), a=5.26, sigma=0.08, seed=2)
    run_pairs("lj_argon_fcc_108", ar, lj_argon())
    run_pairs("lj_argon_fcc_108_smooth", ar, lj_argon(margin=1.0))


def boundary_cases():
    """The rest of the f2py boundary pysic calls on this path, as the
    reference core answers them: its descriptor registry
    (PyInterface.f90:706-1011), the bond-order factor queries of a Tersoff
    coordinator (579-634) and the stand-alone Ewald energy (1050)."""
    R = O.RefCore()
    reg = R.registry()
    pots, bonds = reg["potentials"], reg["bond_order_factors"]

    def pad(names):
        return list(names) + [""] * (12 - len(names))
    np.savez_compressed(
        os.path.join(HERE, "registry_reference.npz"),
        pot_names=np.array([p[0] for p in pots]), pot_targets=np.array([p[1] for p in pots]),
        pot_nparams=np.array([len(p[2]) for p in pots]),
        pot_param_names=np.array([pad(p[2]) for p in pots]),
        bond_names=np.array([b[0] for b in bonds]), bond_targets=np.array([b[1] for b in bonds]),
        bond_levels=np.array([b[2] for b in bonds]),
        bond_nparams=np.array([[len(b[3][0]), len(b[3][1])] for b in bonds]),
        bond_param_names=np.array([[pad(b[3][0]), pad(b[3][1])] for b in bonds]))
    print("registry_reference", len(pots), "potentials", len(bonds), "bond-order factors")
    # bond-order queries of the Tersoff group on a small displaced diamond
    s = st.diamond_si((2, 2, 2), sigma=0.05, seed=1)
    p_t, b_t, cuts = tersoff_setup(s)
    R.setup(s.symbols, s.positions, s.cell, s.pbc, potentials=p_t, tersoff=b_t, ngroups=1,
            cutoffs=cuts)
    atom = 3
    np.savez_compressed(os.path.join(HERE, "tersoff_bond_order_queries.npz"),
                        positions=s.positions, cell=s.cell, cutoffs=cuts, group=0, atom=atom,
                        factors=R.bond_order_factors(0),
                        gradients=R.bond_order_gradients(0, atom),
                        gradients_of_factor=R.bond_order_gradients_of_factor(0, atom))
    print("tersoff_bond_order_queries", len(s))
    # get_ewald_energy on the sheared rocksalt cell: two parameter sets,
    # neither the one set with set_ewald_parameters (scalers there are not 1)
    r = st.rocksalt((2, 2, 2), sigma=0.08, seed=6)
    shear = np.array([[1.0, 0.0, 0.0], [0.2, 1.0, 0.0], [-0.1, 0.15, 1.0]])
    r.cell = r.cell @ shear
    r.positions = r.positions @ shear
    eps = 0.00552635
    rc = 5.0
    sc = np.random.default_rng(12).uniform(0.5, 1.5, len(r))
    R.setup(r.symbols, r.positions, r.cell, r.pbc, charges=r.charges, potentials=[],
            tersoff=[], ngroups=0, cutoffs=np.full(len(r), rc + 0.5),
            ewald=(rc, 1.8, kcells(r.cell, 1.8), rc / 3.0, eps), scaler=sc)
    params, energies = [], []
    for sigma in (rc / 3.0, 1.4):
        kc = 3.0 / sigma
        K = kcells(r.cell, kc)
        params.append([rc, kc, K[0], K[1], K[2], sigma, eps])
        energies.append(R.ewald_energy(rc, kc, K, sigma, eps))
    np.savez_compressed(os.path.join(HERE, "ewald_energy_reference.npz"), positions=r.positions,
                        cell=r.cell, charges=r.charges, symbols=np.array(r.symbols), scaler=sc,
                        params=np.array(params), energies=np.array(energies))
    print("ewald_energy_reference", len(r), energies)


def main():
    if not O.ref_available():
        raise SystemExit("build the reference checker first: oracle/build_ref.sh")
    if sys.argv[1:] == ["pairs"]:
        pair_cases()
        return
    if sys.argv[1:] == ["comb"]:
        comb_full()
        return
    # Tersoff Si, the parity configuration at reduced and full size
    run_tersoff("si_diamond_2x2x2_s005", st.diamond_si((2, 2, 2), sigma=0.05, seed=1))
    run_tersoff("si_diamond_4x4x4_s005", st.diamond_si((4, 4, 4), sigma=0.05, seed=0))
    run_tersoff("si_diamond_8x8x8_s005_seed0", st.diamond_si((8, 8, 8), sigma=0.05, seed=0),
                with_lists=False)
    run_tersoff("si_diamond_3x3x3_perfect", st.diamond_si((3, 3, 3)))
    # amorphous Si: neighbors inside the smooth-cutoff shell
    run_tersoff("si_amorphous_216", st.amorphous_si(216, 2.3, seed=7))
    # a != 0 exercises the exp((a (r_ij - r_ik))^mu) branch
    ters_a = list(TERS)
    ters_a[2] = 1.3
    run_tersoff("si_amorphous_128_a13", st.amorphous_si(128, 2.3, seed=11), ters=ters_a)
    # open boundaries: padding subcells
    clus = st.diamond_si((3, 3, 3), sigma=0.05, seed=5)
    clus.pbc[:] = [True, False, False]
    clus.cell = np.diag([16.293, 20.0, 21.0])
    run_tersoff("si_slab_open_yz", clus)
    # Ewald: rocksalt with displacements
    run_ewald("nacl_2x2x2_s008_rc5", st.rocksalt((2, 2, 2), sigma=0.08, seed=3), rc=5.0)
    run_ewald("nacl_3x3x3_s008_rc7", st.rocksalt((3, 3, 3), sigma=0.08, seed=4), rc=7.0)
    sheared_ewald()
    scaled_ewald()
    sic_mixed()
    comb_charge_pairs()
    comb_charge_pairs("sio2_48_charge_terms_asymmetric", asymmetric=True, seed=23)
    comb_bend()
    comb_full()
    # EEM: small random SiO2 with charge_self + Ewald and random charges
    sio2 = st.amorphous_sio2(16, 2.2, seed=2)
    rng = np.random.default_rng(8)
    q = rng.normal(0.0, 0.5, len(sio2))
    q -= q.mean()
    sio2.set_initial_charges(q)
    run_ewald("sio2_48_eem_rc5", sio2, rc=5.0,
              charge_self={"Si": (4.77, 8.0), "O": (8.74, 13.4)})
    pair_cases()
    boundary_cases()


if __name__ == "__main__":
    import sys
    if sys.argv[1:] == ["sheared_ewald"]:
        sheared_ewald()
    elif sys.argv[1:] == ["sic_mixed"]:
        sic_mixed()
    elif sys.argv[1:] == ["comb_charge_pairs"]:
        comb_charge_pairs()
        comb_charge_pairs("sio2_48_charge_terms_asymmetric", asymmetric=True, seed=23)
    elif sys.argv[1:] == ["comb_bend"]:
        comb_bend()
    elif sys.argv[1:] == ["scaled_ewald"]:
        scaled_ewald()
    elif sys.argv[1:] == ["boundary"]:
        boundary_cases()
    else:
        main()
