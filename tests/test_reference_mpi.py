"""The MPI build of the reference core (the multi-core CPU baseline) gives
the serial build's forces, energy and stress: mpiexec -n 2 of
oracle/ref_md.py on a 6^3-cell Tersoff Si sample against the serial core in
this process.  Skipped where either build or mpiexec is missing."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

MPIEXEC = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"


@pytest.mark.skipif(not (O.ref_available() and os.path.exists(O.REF_MPI_PATH)
                         and os.path.exists(MPIEXEC)),
                    reason="reference builds or mpiexec not present")
def test_mpi_reference_matches_serial(tmp_path):
    import ref_md
    out = str(tmp_path / "mpi.npz")
    rank = "ulimit -s $(ulimit -H -s) && exec %s %s --reps 6 --steps 1 --forces-out %s" % (
        sys.executable, os.path.join(ROOT, "oracle", "ref_md.py"), out)
    r = subprocess.run([MPIEXEC, "-n", "2", "-launcher", "fork", "bash", "-c", rank],
                       capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS="1"), cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-500:]
    assert '"ranks": 2' in r.stdout
    m = np.load(out)
    R, s, cut = ref_md.setup(6)
    f, stress = R.forces()
    e = R.energy()
    print("max |dF| %.2e  |dE| %.2e" % (np.abs(m["forces"] - f).max(), abs(m["energy"] - e)))
    assert np.abs(m["forces"] - f).max() <= 1e-12 * np.abs(f).max()
    assert abs(float(m["energy"]) - e) <= 1e-12 * abs(e)
    assert np.abs(m["stress"] - stress).max() <= 1e-12 * np.abs(stress).max()
