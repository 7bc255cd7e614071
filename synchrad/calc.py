"""``synchrad.calc`` import path; see ``synchrad_amd.calc``."""
from synchrad_amd.calc import SynchRad, mpi_installed, h5py_installed  # noqa: F401
