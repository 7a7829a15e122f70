"""mom25d on MI355X: the per-frequency solve (Green's tables, RWG/MPIE fill,
dense LU and port de-embedding) of the mom25d 2.5D method of moments, on
hand-written gfx950 kernels behind the reference's Python API.

The names that need torch and the GPU library (Solver and friends) load on
first use, so that the host-only modules (mesh, layered, deembed) can be
imported without them, e.g. by the worker processes of the wave fits."""

import importlib

from . import analytic
from .layered import GND, Dielectric, Stackup
from .mesh import CoplanarGround, Layout, Mesh, Port, Via

mm = 1e-3
GHz = 1e9

_LAZY = {"Solver": "solver", "Substrate": "solver", "SweepResult": "solver", "fit_waves": "deembed",
         "via_impedance": "solver", "via_inductance": "solver", "sweep_batch": "solver"}

__all__ = ["analytic", "GND", "Dielectric", "Stackup", "CoplanarGround", "Layout", "Mesh", "Port", "Via", "Solver", "Substrate",
           "SweepResult", "fit_waves", "via_impedance", "via_inductance", "sweep_batch", "mm", "GHz"]


def __getattr__(name):
    mod = _LAZY.get(name)
    if mod is None:
        raise AttributeError(f"module 'mom25d_amd' has no attribute {name!r}")
    value = getattr(importlib.import_module(f".{mod}", __name__), name)
    globals()[name] = value
    return value
