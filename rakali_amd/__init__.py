"""MI355X (gfx950) structured 2-D shallow-water step of Rakali.

The compute path is the HIP library ``librakali_amd.so`` behind the
``rkg_*`` C ABI (``include/rakali_amd.h``); this package is its host side.
"""
from .grid import GRAVITY, Grid, make_config  # noqa: F401

__all__ = ["Grid", "make_config", "GRAVITY"]
