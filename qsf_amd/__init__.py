"""qsf_amd: MI355X-native split-operator time step of QSF (HIP kernels behind a C ABI).

The compute path is libqsf_hip.so (qsf_amd/csrc, hand-written HIP for gfx950); this Python
package only binds it (qsf_amd.hip) and holds the descriptor mirror (qsf_amd.qsf_types).
"""
from . import qsf_types  # noqa: F401

__all__ = ["qsf_types"]
