"""``synchrad.utils`` import path; see ``synchrad_amd.utils``."""
from synchrad_amd.utils import (  # noqa: F401
    J_in_um, Utilities, alpha_fs, energy_1m_eV, omega_1m, r_e,
)
