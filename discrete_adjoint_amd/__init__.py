"""MI355X hot path of the TITAN2D-derived granular-flow code: the order-1
finite-volume step and its discrete adjoint as HIP kernels behind a C ABI
(include/tg.h).  Python here is host plumbing over that ABI."""
from . import _lib  # noqa: F401
from ._lib import TG_FP64, TG_MIXED  # noqa: F401


def load():
    return _lib.load()
