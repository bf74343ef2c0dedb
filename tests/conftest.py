import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run on the GPU box)")


@pytest.fixture(scope="session")
def oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    import oracle as O
    O.lib()
    return O


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


TERSOFF_CASES = ["si_diamond_2x2x2_s005", "si_diamond_4x4x4_s005", "si_diamond_3x3x3_perfect",
                 "si_amorphous_216", "si_amorphous_128_a13", "si_slab_open_yz"]
EWALD_CASES = ["nacl_2x2x2_s008_rc5", "nacl_3x3x3_s008_rc7", "sio2_48_eem_rc5",
               "nacl_2x2x2_sheared_rc5", "nacl_4x2x1_scaled_rc5"]
