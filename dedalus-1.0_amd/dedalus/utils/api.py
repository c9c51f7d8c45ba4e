"""Public utilities API (dedalus/utils/api.py)."""

from .parallelism import com_sys, load_all, swap_indices
from .timer import Timer
from .misc_numeric import integrate_quad, integrate_simp, interp_linear, find_zero
from .restart import restart, restart_into


def identify_version():
    return "dedalus-1.0 interface on MI355X"
