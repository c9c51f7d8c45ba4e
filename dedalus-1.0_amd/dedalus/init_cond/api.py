"""Public initial-condition API (dedalus/init_cond/api.py)."""

from .init_cond import (taylor_green, sin_k, cos_k, turb, turb_new, MIT_vortices,
                        vorticity_wave, alfven, kida_vortex, add_gaussian_white_noise,
                        constant, mcwilliams_spec)
