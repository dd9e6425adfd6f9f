"""nwgrad_amd: the batched alignment hot path of nwgrad, native to AMD Instinct MI355X (gfx950).

Drop-in surface (the reference's ``nwgrad`` names and argument meanings):

* ``SeqPairBatch`` (float32 Viterbi / soft) and ``SeqPairBatchDouble`` (float64):
  ``score_and_grad``, ``compute_grad``, ``scores``, ``grads``, ``weighted_grad``,
  ``set_params``, ``fill``, soft settings.
* ``AlignParams`` / ``SubstMatrix``: parameters and gradients.

Every alignment runs on the GPU through ``libnwgrad_hip.so`` (HIP kernels behind the C ABI of
``include/nwgrad_hip.h``).  There is no CPU fallback: importing works anywhere the native module
was built, but constructing a batch without a usable gfx950 device raises ``RuntimeError``.
"""
from __future__ import annotations

import os as _os

_here = _os.path.dirname(_os.path.abspath(__file__))


def _share_torch_hip_runtime():
    """PyTorch-ROCm ships its own HIP runtime under the same soname as the system one.  Two
    copies in one process cannot both open the GPU, and the library would otherwise bind the
    system copy first; so when torch is installed, its copy is loaded (only that library, torch
    itself is not imported) before the native module, which then binds to it."""
    import ctypes
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    lib = _os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if _os.path.exists(lib):
        try:
            ctypes.CDLL(lib, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            pass


_share_torch_hip_runtime()

try:
    from ._nwgrad_amd import AlignParams, SeqPairBatch, SeqPairBatchDouble, SubstMatrix, device_available
except ImportError as _e:  # the native module is the product: fail loudly
    raise ImportError(
        "nwgrad_amd: the native module is not built (run `python nwgrad_amd/_build.py` or "
        "`python -c 'import __graft_entry__ as g; g.build()'`): " + str(_e)) from _e

from . import tuning as _tuning

_tuning.install()   # the measured lane-width table (nwgrad_amd/lane_table.json), when present

AA_ORDER = "ACDEFGHIKLMNPQRSTVWY"
DNA = "ACGT"

__all__ = ["AlignParams", "SeqPairBatch", "SeqPairBatchDouble", "SubstMatrix", "device_available",
           "AA_ORDER", "DNA", "lib_path"]


def lib_path() -> str:
    """Path of the HIP library this package runs on."""
    return _os.path.join(_here, "libnwgrad_hip.so")
