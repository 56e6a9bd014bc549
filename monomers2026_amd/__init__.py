"""MI355X-native SELM fluctuating-hydrodynamics step (HIP/CDNA4 behind a C ABI).

The product is ``libselm_hip.so`` (kernels + C ABI, include/selm_hip.h) and
the C++ driver that reads the reference's Model.SELM / LAMMPS decks.  This
Python package is the ctypes binding used by tests and benchmarks.
"""
from . import selm  # noqa: F401
from .selm import System, Pool, make_params, load, PREC_FP64, PREC_FP32_BULK  # noqa: F401
