"""MI355X (gfx950) DEM time step and Darcy pressure solve behind the Python
``sim`` API of the sphere granular simulator.

``sim`` mirrors python/sphere.py's setup, binary I/O and ``run()``;
``Engine`` is the in-process C ABI context of libsphere_hip.so; ``Ensemble``
steps many independent Engines together; ``Net`` is the Darcy correction
network on its kernels; ``NavierStokes`` is the reference's cfd_solver = 0
fluid step without grains.
"""
from .sim import sim, status, run, cleanup, V_sphere, VERSION  # noqa: F401
from ._abi import Engine, Ensemble, Net, NavierStokes, SphereError, load_library, ns_params  # noqa: F401
