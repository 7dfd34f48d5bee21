"""MI355X (gfx950) structured 2-D shallow-water step of Rakali.

The compute path is the HIP library ``librakali_amd.so`` behind the
``rkg_*`` C ABI (``include/rakali_amd.h``); ``librakali.so`` layers the
reference's handle API (``rki_*``, ``include/rakali.h``) on top.  This
package is their host side: ``Grid`` binds ``rkg_*`` directly, ``Rakali``
is the reference's simulation class over ``rki_*``.
"""
from .grid import GRAVITY, Grid, make_config  # noqa: F401
from .rakali import Rakali  # noqa: F401

__all__ = ["Grid", "Rakali", "make_config", "GRAVITY"]
