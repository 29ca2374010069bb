"""periodicflow on MI355X: the Fourier-Galerkin RHS and time stepping as HIP kernels.

Same import name and Python API as the reference package
(src/periodicflow/__init__.py, simulation.py, space_solver.py,
time_integrator.py, basis.py, math.py); the array work runs in libpfhip.so.
Submodules that need torch and the HIP library are imported on first use.
"""
import logging

try:  # rich formatting as in the reference's logger, when available
    from rich.console import Console
    from rich.logging import RichHandler

    _handler = RichHandler(console=Console(width=120), show_time=False, show_level=False, show_path=False)
except ImportError:  # pragma: no cover
    _handler = logging.StreamHandler()

logger = logging.getLogger("periodicflow")
if not logger.handlers:
    logger.addHandler(_handler)
    logger.setLevel(logging.INFO)

__version__ = "0.1.0"
