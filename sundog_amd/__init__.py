"""sundog_amd: sundog's render path for one AMD Instinct MI355X (gfx950).

The product is libsundog.so (sundog_amd/lib), a C-ABI library exporting the
reference's scene-construction and render entry points plus resident render
sessions. This package holds its build script, its ctypes bindings and the
scene-program driver.
"""

from ._lib import load, last_error, SundogError  # noqa: F401
from .scenelib import (Scene, Session, SceneHandle, SceneError, scale, rotate_x, rotate_y,  # noqa: F401
                       rotate_z, translate, rigid_body, static_body, device_pcg32, probe, rel_rmse)
