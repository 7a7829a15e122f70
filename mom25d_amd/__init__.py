"""mom25d on MI355X: the per-frequency solve (Green's tables, RWG/MPIE fill,
dense LU and port de-embedding) of the mom25d 2.5D method of moments, on
hand-written gfx950 kernels behind the reference's Python API."""

from . import analytic
from .layered import GND, Dielectric, Stackup
from .mesh import Layout, Mesh, Port, Via
from .solver import Solver, Substrate, SweepResult, fit_waves, via_impedance, via_inductance

mm = 1e-3
GHz = 1e9

__all__ = ["analytic", "GND", "Dielectric", "Stackup", "Layout", "Mesh", "Port", "Via", "Solver", "Substrate",
           "SweepResult", "fit_waves", "via_impedance", "via_inductance", "mm", "GHz"]
