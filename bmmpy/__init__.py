"""The reference's import name, ``bmmpy``, over this package.

``import bmmpy as bmm`` gives the same public names as the reference's
``bmmpy/__init__.py:34-105`` (every search, replay and GA runs on the GPU
through ``bmmpy_amd``), and ``bmmpy._bmmpy``, ``bmmpy.search``,
``bmmpy.solver``, ``bmmpy.ga``, ``bmmpy.apply``, ``bmmpy.matrix``,
``bmmpy.workflows`` and ``bmmpy.candidate`` are the modules a reference user
imports.  ``bmmpy.algebra`` (Mastrovito generators, row transforms) is not
part of this build.
"""
from __future__ import annotations

import sys

import bmmpy_amd as _impl
from bmmpy_amd import *  # noqa: F401,F403
from bmmpy_amd import __all__, __version__  # noqa: F401
from bmmpy_amd import _bmmpy, apply, ga, matrix, search, solver, workflows  # noqa: F401

from . import candidate  # noqa: E402,F401

for _name in ("_bmmpy", "apply", "ga", "matrix", "search", "solver", "workflows"):
    sys.modules[f"{__name__}.{_name}"] = getattr(_impl, _name)
del _name
