"""MI355X-native (gfx950) exact diagonalisation of the t-J model on two-leg ladders.

A drop-in for the data-parallel core of the reference ``exactdiag`` package:
symmetry-adapted basis, matrix-free H·v, Lanczos and continued-fraction
spectra run as hand-written HIP kernels in ``libexactdiag_hip.so``; the
Python modules mirror the reference's configuration and API surface.
"""

__version__ = "0.1.0"
