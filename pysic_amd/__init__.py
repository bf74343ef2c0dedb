"""pysic_amd -- an MI355X-native force engine with the pysic calculator API.

The Tersoff bond-order, Ewald and EEM charge-relaxation path of pysic,
evaluated by hand-written HIP kernels for gfx950 behind a C-ABI
(include/pysic_hip.h, libpysic_hip.so) that replaces the reference's f2py
interface.
"""
__version__ = "0.1.0"
