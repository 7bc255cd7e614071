"""Import path of the reference package (``synchrad``), served by ``synchrad_amd``.

``from synchrad.calc import SynchRad`` and ``from synchrad.utils import J_in_um``
resolve to the MI355X implementation, so scripts written for the reference
(e.g. its tests/test_undulator_analytic.py) run unchanged.
"""
__version__ = "0.0.3"
