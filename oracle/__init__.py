"""CPU oracle for the pseudospectral step -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch numpy restatement of dedalus-1.0's numpy
path (``FFT.method = 'numpy'``, ``FFT.dealiasing = '2/3'``).  It exists to
check the HIP implementation: only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product
package (``dedalus`` under ``dedalus-1.0_amd/``) never imports it and has no
CPU fallback.

Modules
-------
numpy_path
    Field component / field / state containers with the reference's lazy
    space-switching semantics, the physics right-hand sides (incompressible
    hydro, Boussinesq, MHD), the exponential-time-differencing kernels and
    the RK2mid / RK2trap integrators.  Every function cites the reference
    file:line it restates.
ics
    Seeded initial conditions for the workloads (Taylor-Green 64^2, random
    solenoidal 3D fields with prescribed shell spectra, Boussinesq noise),
    built directly in the reference k layout.

Pinning (what checks each oracle function; see tests/test_oracle_*.py)
-----------------------------------------------------------------------
The reference ships no tests, golden vectors or fixtures for this path, so
the oracle is pinned by mathematics:

* k grids -- closed form (integer multiples of 2*pi/L, +Nyquist sign).
* forward/backward transforms -- brute-force O(N^2) DFT on small grids,
  single-mode known answers, Parseval, round trip, Hermitian-part rule of
  c2r on a non-Hermitian kx=0 plane.
* 2/3 dealias -- closed-form kept-mode count and per-axis cutoffs.
* deriv / k2 / projection -- single-mode derivatives, P^2 = P,
  k.P(F) = 0, P(grad phi) = 0.
* nonlinear products -- direct convolution sums on 8^2 / 8^3 grids.
* hydro + RK2mid -- Taylor-Green closed form u(t) = u(0) exp(-2 nu t).
* Boussinesq -- linear internal-gravity-wave frequency N kx/|k|.
* MHD -- linear Alfven-wave frequency k.B0/sqrt(4 pi rho0), div B = 0.
* ETD kernels -- phi functions against mpmath at 50 digits, branch points.
* ICs -- Hermitian symmetry, solenoidality, prescribed shell spectrum, rms.

No function of the oracle is left "parity unpinned".
"""
