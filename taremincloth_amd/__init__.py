"""MI355X-native XPBD cloth substep engine (HIP for gfx950) behind the
reference's ``taremin_cloth_core`` API.

``taremincloth_amd.taremin_cloth_core`` is the drop-in module; register it
under the reference's name with ``taremincloth_amd.install_as_taremin_cloth_core()``.
"""
from __future__ import annotations

import sys

from . import taremin_cloth_core  # noqa: F401

__version__ = "0.1.0"


def install_as_taremin_cloth_core() -> None:
    """Make ``import taremin_cloth_core`` resolve to this package's module."""
    sys.modules["taremin_cloth_core"] = taremin_cloth_core
