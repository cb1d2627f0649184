"""CPU FP64 restatement of the reference's fast spectral inelastic collision path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports, calls or
links anything under ``oracle/``.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Modules
-------
quadrature  : radial Gauss-Legendre, circle, Lebedev and tabulated sphere rules
              (reference ``src/collision/spectral_mesh.py``,
              ``src/collision/spherical_design/__init__.py``)
collision   : kernel precomputation and ``collide``/``laplacian``
              (reference ``src/collision/inelastic_collisions.py``,
              ``src/collision/base.py``), plus the moments and the FP64
              conservative projection that the reference lacks
transport   : ghost cells, ``advection_1D``, ``tvd.limit``/``minmod`` and
              ``BoltzmannSolver1D.df_advection``
              (reference ``src/solver.py``, ``src/riemann``, ``src/limiters``,
              ``src/boltzmann.py``)
integrators : Euler / PFE / TPI2 / TPI3 steps and ``evolve_to_time``
              (reference ``src/boltzmann.py``, ``src/solver.py``)
brute       : direct O(n^2) evaluation of the same discrete gain sum, used to
              pin ``collision.collide`` independently of any FFT library.

Pinning status of each function is listed in each module's docstring.
"""
