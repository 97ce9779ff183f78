"""Python binding of the MI355X DEM step library (libbdem.so).

types      ctypes mirrors of the reference's boundary structs
lib        the C ABI and a KDevice-style Device wrapper
scenarios  seeded synthetic decks (box fill, silo, dense random packs)
deck       the reference project decks and state files
"""
from . import types, lib, scenarios, deck, adjoint  # noqa: F401
from .lib import Batch, BdemError, Device, load  # noqa: F401
