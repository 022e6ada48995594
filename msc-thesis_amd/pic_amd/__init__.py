"""MI355X-native 2D electrostatic PIC engine (Python host side of libpicamd.so)."""
from . import _lib
from ._lib import PicError, check, default_config, lib, pic_config, pic_grid
from .engine import (Batch, Simulation, checkpoint_info, lattice_count, load_landau,
                     load_reference, load_two_stream, run_compat)

__all__ = ["Batch", "Simulation", "PicError", "check", "default_config", "lib", "pic_config",
           "pic_grid", "checkpoint_info", "lattice_count", "load_landau", "load_reference", "load_two_stream",
           "run_compat", "_lib"]
