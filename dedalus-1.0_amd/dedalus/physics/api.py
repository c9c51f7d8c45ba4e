"""Public physics API (dedalus/physics/api.py)."""

from .physics import IncompressibleHydro, BoussinesqHydro, IncompressibleMHD
