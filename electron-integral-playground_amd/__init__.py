"""MI355X-native McMurchie-Davidson integral path (overlap, nuclear attraction, (P|Q), (ij|P)).

Host layer mirrors the reference package's interface (readers, basis_utility, pair_utility,
molecular_integral.*); the integrals run in the HIP library ``lib/libmolecular_integral.so``.
"""

import ctypes
import pathlib

PACKAGE_DIR = pathlib.Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "lib" / "libmolecular_integral.so"

_lib = None


def _bind_torch_hip_runtime():
    """Load PyTorch (when installed) before the library, so the process holds ONE HIP runtime.

    PyTorch ships its own libamdhip64 (SONAME libamdhip64.so.7) and its libraries ask for it by the
    unversioned name.  Loaded first, it satisfies this library's libamdhip64.so.7 dependency too; loaded
    after the system runtime, it would come in as a second copy whose initialisation finds no device,
    and torch streams passed to this library would belong to the other runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def library() -> ctypes.CDLL:
    """Load the HIP library.  There is no CPU fallback: a missing library is an error."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(f"HIP library {LIBRARY_PATH} is not built; run build() first")
        _bind_torch_hip_runtime()
        _lib = ctypes.CDLL(str(LIBRARY_PATH))
    return _lib
