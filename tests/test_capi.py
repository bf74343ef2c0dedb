"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
entry point declared in include/pysic_hip.h, and fails cleanly (an error
code, no crash) where no GPU is present."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h)
           for h in ("pysic_hip.h", "pysic_learned.h", "pysic_batch.h", "pysic_registry.h")]
LIB = os.path.join(ROOT, "pysic_amd", "libpysic_hip.so")


def declared_symbols():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pysic_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("pysic_create", "pysic_create_atoms", "pysic_create_cell", "pysic_set_potentials",
              "pysic_set_bond_order_factors", "pysic_set_ewald",
              "pysic_generate_neighbor_lists", "pysic_calculate_energy",
              "pysic_calculate_forces", "pysic_calculate_electronegativities",
              "pysic_relax_charges", "pysic_md_run", "pysic_mlp_predict_charges",
              "pysic_mlp_train", "pysic_mlp_gradients", "pysic_set_cells",
              "pysic_evaluate_cells", "pysic_get_cell_results"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libpysic_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from pysic_amd.engine import SIGNATURES
    assert sorted(SIGNATURES) == declared_symbols()


def test_engine_without_gpu_raises_cleanly():
    import torch  # noqa: F401  (only to know whether a GPU is visible)
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible here")
    from pysic_amd.engine import Engine, PysicEngineError
    with pytest.raises(PysicEngineError):
        Engine(0)


def test_kernel_variants_are_listed():
    from pysic_amd.engine import load
    L = load()
    n = L.pysic_kernel_variant_count()
    assert n >= 2
    names = [L.pysic_kernel_variant_name(v).decode() for v in range(n)]
    assert all(names)
