"""les_amd: MI355X (gfx950) HIP implementation of the KhaledGhannam/LES
per-timestep hot path behind a C ABI (include/les_amd.h)."""
from ._lib import LIB_PATH, declared_symbols, load  # noqa: F401


def __getattr__(name):
    # torch-dependent API is imported lazily so that the C-ABI checks need no GPU
    if name in ("Les", "LesSim", "LesError"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
