"""Learned sub-grid closure for coarse 2D runs.

A small periodic CNN (3x3 convolutions, GELU) maps the coarse velocity to the
sub-grid tendency that coarse-graining leaves out; the prediction enters the
coarse solver through the velocity forcing hook
(dedalus/physics/physics.py:524-525, applied at :569-572).  The convolutions
run on hand-written BF16 MFMA kernels of libddx (include/ddx_closure.h)
behind a torch.autograd.Function; pair generation, coarse-graining and the
coarse runs use the same library's pseudospectral step.
"""

from .net import ClosureNet, conv3x3, pad16  # noqa: F401
