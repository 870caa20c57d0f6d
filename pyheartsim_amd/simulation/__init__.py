"""Drop-in namespace for the reference's ``simulation`` package
(simulation/__init__.py:15-19): equation_parts, pacing, simulation_gpu,
unipolar_electrogram_equation_parts, unipolar_electrogram_gpu.
"""
from . import equation_parts, pacing, unipolar_electrogram_equation_parts  # noqa: F401
from . import simulation_gpu, unipolar_electrogram_gpu  # noqa: F401
