"""CPU oracle for the triply periodic Navier-Stokes RK4 path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker (or the timed CPU
baseline), never as the thing measured or shipped.

``oracle.ns`` is a NumPy/SciPy restatement of the reference solver
(spectralDNS/solvers/NS.py, spectralDNS/maths/{cross,integrators}.py and the
``solve`` loop of spectralDNS/__init__.py).  Each function cites the
reference file:line it follows.  How each part is pinned (reference golden
numbers, closed forms, brute-force sums) is stated in ``oracle/ns.py``.
The other modules restate, on oracle.ns's transforms, the reference's other
integrators (``oracle.integrators``), the velocity-vorticity solver
(``oracle.vv``), the MHD solver (``oracle.mhd``), the NS Standard /
Divergence / Skewed convection forms (``oracle.convection``) and the learned
closure (``oracle.closure``); each module's header states its pins.
"""
