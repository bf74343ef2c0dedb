"""pysic_amd -- an MI355X-native force engine with the pysic calculator API.

The Tersoff bond-order, pair-potential, Ewald and EEM charge-relaxation
path of pysic, evaluated by hand-written HIP kernels for gfx950 behind a C-ABI
(include/pysic_hip.h, libpysic_hip.so) that replaces the reference's f2py
interface.  The Python objects mirror the reference's: Pysic,
FastNeighborList, Potential, ProductPotential, BondOrderParameters,
Coordinator, CoulombSummation, ChargeRelaxation.
"""
__version__ = "0.1.0"

from .interactions import (BondOrderParameters, Coordinator, InvalidParametersError,  # noqa: F401
                           InvalidPotentialError, Potential, ProductPotential, comb_sio2,
                           tersoff_si)
from .coulomb import ChargeRelaxation, CoulombSummation, estimate_ewald_parameters  # noqa: F401
from .structures import Structure  # noqa: F401


def __getattr__(name):
    # the calculator loads the native library; import it lazily
    if name in ("Pysic", "FastNeighborList"):
        from . import calculator
        return getattr(calculator, name)
    raise AttributeError(name)
