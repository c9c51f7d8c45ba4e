"""Public time-stepping API (dedalus/time_stepping/api.py)."""

from .time_step import RK2mid, RK2trap, RK4, CrankNicholsonVisc
