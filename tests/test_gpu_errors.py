"""Argument and state errors of the C-ABI (include/pysic_hip.h).

pysic reports misuse of its core from Python before the Fortran runs:
calculator.py:1057 and 1574 (MissingAtomsError, lists or potentials before
the atoms), 1174-1192 and 1285 (InvalidPotentialError), 1395 and 1436
(InvalidParametersError), 1494 (LockedCoreError, a count mismatch).  The
C-ABI checks the same conditions on the host before it launches anything:
each refused call returns its PYSIC_E* code with a message naming the
problem, and the handle then evaluates a sound system as before.
"""
import ctypes

import numpy as np
import pytest

from helpers import tersoff_descriptors
from pysic_amd import structures as st

pytestmark = pytest.mark.gpu

EINVAL, ESTATE = -1, -3


@pytest.fixture(scope="module")
def sound():
    """A small Si cell, its descriptors and the energy a fresh handle gives."""
    from pysic_amd.engine import Engine
    s = st.diamond_si((3, 3, 3), sigma=0.05, seed=11)
    pd, bd = tersoff_descriptors()
    e = Engine(0)
    load(e, s, pd, bd)
    e_ref = e.energy()
    e.close()
    return s, pd, bd, e_ref


@pytest.fixture
def eng():
    from pysic_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def load(e, s, pd, bd):
    n = len(s)
    e.create_atoms(np.zeros(n, np.int32), s.positions)
    e.create_cell(s.cell, (True, True, True))
    e.set_potentials(pd)
    e.set_bond_order_factors(bd)
    e.ewald_off()
    e.generate_neighbor_lists(np.full(n, 3.5))


def refused(e, rc, code, text):
    """rc is the code expected and the handle's message names the problem."""
    msg = e.L.pysic_last_error(e.h).decode()
    print("  %d: %s" % (rc, msg), flush=True)
    assert rc == code
    assert text in msg


def still_sound(e, sound):
    s, pd, bd, e_ref = sound
    load(e, s, pd, bd)
    assert abs(e.energy() - e_ref) <= 1e-12 * abs(e_ref)


def test_no_such_device():
    from pysic_amd.engine import Engine, PysicEngineError
    print("case: a handle on device 64", flush=True)
    with pytest.raises(PysicEngineError, match="no usable HIP device 64"):
        Engine(64)


def test_calls_out_of_order(eng, sound):
    s, pd, bd, _ = sound
    n = len(s)
    L, h = eng.L, eng.h
    out = np.zeros(1)
    dp = out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    cut = np.full(n, 3.5)
    cp = cut.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    print("case: evaluation and list build on an empty handle", flush=True)
    refused(eng, L.pysic_calculate_energy(h, dp), ESTATE, "atoms not created")
    refused(eng, L.pysic_generate_neighbor_lists(h, n, cp), ESTATE,
            "atoms and cell must be created first")
    print("case: atoms without a cell", flush=True)
    eng.create_atoms(np.zeros(n, np.int32), s.positions)
    refused(eng, L.pysic_calculate_energy(h, dp), ESTATE, "cell not created")
    print("case: no neighbor lists", flush=True)
    eng.create_cell(s.cell, (True, True, True))
    eng.set_potentials(pd)
    eng.set_bond_order_factors(bd)
    eng.ewald_off()
    refused(eng, L.pysic_calculate_energy(h, dp), ESTATE, "neighbor lists not generated")
    still_sound(eng, sound)


def test_bad_arguments(eng, sound):
    """Refused calls change nothing: the loaded system still evaluates to
    the same energy after each."""
    s, pd, bd, e_ref = sound
    n = len(s)
    L, h = eng.L, eng.h
    load(eng, s, pd, bd)
    dbl = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)

    def unchanged():
        assert abs(eng.energy() - e_ref) <= 1e-12 * abs(e_ref)

    f = np.zeros((2 * n, 3))
    w = np.zeros(6)
    print("case: null and mismatched outputs", flush=True)
    refused(eng, L.pysic_calculate_energy(h, None), EINVAL, "calculate_energy: null output")
    refused(eng, L.pysic_calculate_forces(h, n + 1, f.ctypes.data_as(dbl),
                                          w.ctypes.data_as(dbl)), EINVAL,
            "calculate_forces: bad arguments")
    refused(eng, L.pysic_calculate_forces(h, n, f.ctypes.data_as(dbl), None), EINVAL,
            "calculate_forces: bad arguments")
    refused(eng, L.pysic_get_state(h, n - 1, f.ctypes.data_as(dbl), f.ctypes.data_as(dbl)),
            EINVAL, "get_state: bad arguments")
    unchanged()
    print("case: create_atoms with no atoms, and with twice the atoms and a species 4",
          flush=True)
    pos2 = np.ascontiguousarray(np.vstack([s.positions, s.positions + 0.7]))
    sp2 = np.zeros(2 * n, np.int32)
    refused(eng, L.pysic_create_atoms(h, 0, None, None, pos2.ctypes.data_as(dbl), None,
                                      sp2.ctypes.data_as(ip)), EINVAL,
            "create_atoms: bad arguments")
    sp2[n + 7] = 4
    refused(eng, L.pysic_create_atoms(h, 2 * n, None, None, pos2.ctypes.data_as(dbl), None,
                                      sp2.ctypes.data_as(ip)), EINVAL, "species must be 0..3")
    unchanged()
    print("case: cutoffs all zero, and a count that does not match", flush=True)
    zero = np.zeros(n)
    refused(eng, L.pysic_generate_neighbor_lists(h, n, zero.ctypes.data_as(dbl)), EINVAL,
            "neighbor cutoff must be positive")
    refused(eng, L.pysic_generate_neighbor_lists(h, n - 1, zero.ctypes.data_as(dbl)), EINVAL,
            "generate_neighbor_lists: bad arguments")
    unchanged()
    print("case: a NaN position, an infinite momentum, a zero mass, a NaN charge, a NaN "
          "time step", flush=True)
    sp = np.zeros(n, np.int32)
    bad = np.ascontiguousarray(s.positions.copy())
    bad[5, 1] = np.nan
    refused(eng, L.pysic_create_atoms(h, n, None, None, bad.ctypes.data_as(dbl), None,
                                      sp.ctypes.data_as(ip)), EINVAL, "non-finite position")
    refused(eng, L.pysic_update_atom_coordinates(h, n, bad.ctypes.data_as(dbl), None), EINVAL,
            "non-finite position")
    good = np.ascontiguousarray(s.positions)
    mom = np.zeros((n, 3))
    mom[3, 0] = np.inf
    refused(eng, L.pysic_update_atom_coordinates(h, n, good.ctypes.data_as(dbl),
                                                 mom.ctypes.data_as(dbl)), EINVAL,
            "non-finite momentum")
    mass = np.ones(n)
    mass[2] = 0.0
    refused(eng, L.pysic_create_atoms(h, n, mass.ctypes.data_as(dbl), None,
                                      good.ctypes.data_as(dbl), None, sp.ctypes.data_as(ip)),
            EINVAL, "masses must be positive")
    q = np.zeros(n)
    q[1] = np.nan
    refused(eng, L.pysic_update_atom_charges(h, n, q.ctypes.data_as(dbl)), EINVAL,
            "non-finite charge")
    refused(eng, L.pysic_md_run(h, 1, float("nan"), 1, None), EINVAL, "md_run: bad arguments")
    unchanged()
    print("case: cell vectors that span no volume, or are not finite", flush=True)
    flat = np.array([[5.0, 0.0, 0.0], [0.0, 5.0, 0.0], [5.0, 5.0, 0.0]])
    inv = np.zeros((3, 3))
    pb = np.ones(3, np.int32)
    refused(eng, L.pysic_create_cell(h, flat.ctypes.data_as(dbl), inv.ctypes.data_as(dbl),
                                     pb.ctypes.data_as(ip)), EINVAL, "span no volume")
    flat[2] = [0.0, 0.0, np.inf]
    refused(eng, L.pysic_create_cell(h, flat.ctypes.data_as(dbl), inv.ctypes.data_as(dbl),
                                     pb.ctypes.data_as(ip)), EINVAL, "non-finite cell vector")
    unchanged()
    print("case: run length, kernel shape, grid size, batch evaluation", flush=True)
    refused(eng, L.pysic_md_run(h, -1, 0.1, 1, None), EINVAL, "md_run: bad arguments")
    refused(eng, L.pysic_set_kernel_variant(h, 99), EINVAL, "no such kernel variant")
    refused(eng, L.pysic_set_site_grid(h, -1), EINVAL, "set_site_grid: blocks must be >= 0")
    refused(eng, L.pysic_evaluate_cells(h, 1, 1), ESTATE, "evaluate_cells: no batch of cells set")
    unchanged()
    still_sound(eng, sound)


@pytest.mark.parametrize("case", ["type", "species", "exponentials", "factors"])
def test_bad_descriptors(eng, sound, case):
    from pysic_amd.engine import PysicEngineError
    s, pd, bd, e_ref = sound
    bad = pd.copy()
    if case == "type":
        bad[0, 0], text = 99, "unsupported potential type"
    elif case == "species":
        bad[1, 3], text = 7, "potential species out of range"
    elif case == "exponentials":   # 6 exponential terms for the Si-Si pair (at most 4)
        bad, text = np.vstack([pd] * 3), "too many exponential terms for one pair"
    else:                          # an exponential times 4 more (at most 4 factors)
        mult = pd[0].copy()
        mult[11] = 1.0
        bad, text = np.vstack([pd[0]] + [mult] * 4), "at most 4 factors in a product"
    print("case: descriptor with %s out of range" % case, flush=True)
    load(eng, s, pd, bd)
    with pytest.raises(PysicEngineError, match=text):
        eng.set_potentials(bad)
    # the refused rows are not kept: the model is the one loaded before
    assert abs(eng.energy() - e_ref) <= 1e-12 * abs(e_ref)
    with pytest.raises(PysicEngineError, match="bond-order species out of range"):
        b = bd.copy()
        b[0, 0] = 9
        eng.set_bond_order_factors(b)
    assert abs(eng.energy() - e_ref) <= 1e-12 * abs(e_ref)
    still_sound(eng, sound)


def test_learned_model_refusals(eng, sound):
    """The warm-start network (include/pysic_learned.h) refuses use before
    it is configured, sample indices past its training set, species it was
    not configured for, systems whose lists are shorter than its cutoff,
    and more corrector stages than it holds."""
    from pysic_amd.engine import Mlp, PysicEngineError
    s, pd, bd, _ = sound
    load(eng, s, pd, bd)
    m = Mlp(0)
    try:
        print("case: network not configured", flush=True)
        with pytest.raises(PysicEngineError, match="network not configured"):
            m.predict_charges(eng)
        mu, eta = np.linspace(0.5, 3.0, 16), np.full(16, 4.0)
        m.configure(1, 3.0, mu, eta, np.zeros(32), np.ones(32))
        with pytest.raises(PysicEngineError, match="network parameters not set"):
            m.warm_start(eng, 0)
        m.set_params(np.zeros(m.nparams))
        print("case: training samples and species", flush=True)
        with pytest.raises(PysicEngineError, match="training species out of range"):
            m.set_training_data(np.zeros((4, 32)), [0, 0, 0, 9], np.zeros(4))
        m.set_training_data(np.zeros((4, 32)), [0, 0, 0, 0], np.zeros(4))
        with pytest.raises(PysicEngineError, match="sample index out of range: 4"):
            m.evaluate([0, 4])
        with pytest.raises(PysicEngineError, match="at most 262144 samples"):
            m.gradients(np.zeros((1 << 18) + 1, np.int32))
        assert np.isfinite(m.evaluate([0, 3])).all()
        print("case: corrector stages", flush=True)
        with pytest.raises(PysicEngineError, match="set_corrector: bad arguments"):
            m.set_corrector(np.zeros((17, m.CORRECTOR_NCOEF)))
        bad = np.zeros((2, m.CORRECTOR_NCOEF))
        bad[1, 5] = np.nan
        with pytest.raises(PysicEngineError, match="non-finite coefficient"):
            m.set_corrector(bad)
        assert m.corrector_stages() == 0
        with pytest.raises(PysicEngineError, match="more stages asked for"):
            m.warm_start(eng, 2)
        print("case: lists shorter than the cutoff, and a second species", flush=True)
        m.configure(1, 4.0, mu, eta, np.zeros(32), np.ones(32))
        with pytest.raises(PysicEngineError, match="shorter than the descriptor cutoff"):
            m.predict_charges(eng)
        m.configure(1, 3.0, mu, eta, np.zeros(32), np.ones(32))
        sp = np.zeros(len(s), np.int32)
        sp[::2] = 1
        eng.create_atoms(sp, s.positions)
        eng.generate_neighbor_lists(np.full(len(s), 3.5))
        with pytest.raises(PysicEngineError, match="more species than the network"):
            m.predict_charges(eng)
        print("case: configured for the system: a prediction", flush=True)
        load(eng, s, pd, bd)
        q = m.predict_charges(eng)
        assert q.shape == (len(s),) and np.isfinite(q).all()
    finally:
        m.close()
    still_sound(eng, sound)


def test_coordinate_update_without_momenta_keeps_them(eng, sound):
    """update_atom_coordinates (PyInterface.f90:277) with no momenta moves
    the atoms and keeps their momenta; with momenta it sets both.  Checked
    after a list build has reordered the atoms on the device."""
    s, pd, bd, _ = sound
    n = len(s)
    rng = np.random.default_rng(13)
    mom0 = rng.normal(size=(n, 3))
    eng.create_atoms(np.zeros(n, np.int32), s.positions, momenta=mom0)
    eng.create_cell(s.cell, (True, True, True))
    eng.set_potentials(pd)
    eng.set_bond_order_factors(bd)
    eng.ewald_off()
    eng.generate_neighbor_lists(np.full(n, 3.5))
    print("case: positions moved, momenta not given", flush=True)
    moved = s.positions + rng.normal(scale=0.01, size=(n, 3))
    eng.update_coordinates(moved)
    pos, mom = eng.state()
    assert np.array_equal(pos, moved) and np.array_equal(mom, mom0)
    print("case: positions and momenta given", flush=True)
    mom1 = rng.normal(size=(n, 3))
    eng.update_coordinates(s.positions, mom1)
    pos, mom = eng.state()
    assert np.array_equal(pos, s.positions) and np.array_equal(mom, mom1)
