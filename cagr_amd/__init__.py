"""cagr_amd -- MI355X-native treecode Gauss-formula operator of CAGR.

The HIP library (cagr_amd/csrc, built into cagr_amd/libcagr_amd.so) holds
the octree build, the upward pass and every treecode operator; this package
is the thin Python host layer above its C ABI (include/cagr_amd.h).
"""
from .wn_treecode import WindingNumberTreecode  # noqa: F401
from .compat import _cpu, _cuda  # noqa: F401  (the reference's native-module functions)

__all__ = ["WindingNumberTreecode", "_cpu", "_cuda"]
