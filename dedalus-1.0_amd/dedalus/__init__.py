"""dedalus on MI355X: drop-in Python 3 package for dedalus-1.0's pseudospectral step.

The numerical work (transforms, dealiasing, right-hand sides, ETD updates) runs
in hand-written HIP kernels of libddx.so (``dedalus._hip``); the modules keep
the reference's package layout (``dedalus.mods``, ``dedalus.physics.api``,
``dedalus.data_objects.api``, ``dedalus.time_stepping.api`` ...).
"""
