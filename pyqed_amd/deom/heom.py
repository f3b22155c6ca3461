"""pyqed.deom.heom: the reference's second copy of pyqed/HEOM/heom.py (identical files); re-exports
pyqed_amd.HEOM.heom."""
from ..HEOM.heom import *  # noqa: F401,F403
from ..HEOM.heom import (HEOMSolver, _calc_matsubara_params, _heom, _heom_propagator,  # noqa: F401
                         enr_state_dictionaries, rk_run_params, state_number_enumerate)
