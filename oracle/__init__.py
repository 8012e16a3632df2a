"""FP64 CPU restatement of SchwarMAX's model-evaluation path (the oracle).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or, for the CPU baseline, as the thing timed on the host).  The product
path in ``schwarmax_amd`` never imports it and has no CPU fallback.

Every function cites the reference file:line it restates.  The reference (JAX)
cannot be imported in this environment (jax, jaxlib and astropy are absent), so
each part is pinned by closed forms, invariants, library routines and brute
force in ``tests/test_oracle_*.py``:

=========================  ===================================================
part                       pin
=========================  ===================================================
constants                  closed form from IAU/Julian definitions
NFW / Miyamoto-Nagai       closed forms, Poisson (FD Laplacian = 4 pi G rho)
T3 / V4 needle bars        brute-force needle convolution of the axisymmetric
                           kernel, Poisson pairing potential <-> density,
                           far field -GM/r
gradients (dual numbers)   central finite differences
BS3 integrator             exactness on quadratic-in-t RHS, order-3 global
                           convergence, harmonic closed form, Jacobi integral
grid binning               brute-force pure-Python loop on tiny inputs
Gauss-Hermite basis        closed-form H_k, orthonormality by quadrature
h_to_V_sigma               Gaussian-LOSVD moments by quadrature
ADMM NNLS                  scipy.optimize.nnls on the stacked problem (KKT)
theta maps                 round trip with the forward NFW map
rotation matrix            orthogonality, det = 1, beta = 0 special case
Jeans moments              scipy.integrate.quad on the same integrand;
                           v_phi on a separable model in closed form
rotation curve, T_orb      Kepler closed forms; the R_safe floor and clamps
clockwise cyl -> Cartesian speed, radial component, L_z = -R v_phi
density target             tensor Gauss-Legendre cell integrals of rho R
data normalisation,        hand-derived two-aperture example (mass per
gamma_kin                  orbit, M/L, pixel area, GH normalisation)
bootstrap chi^2            hand-derived rows: broadcast density term,
                           log-mean-exp, sigma_amplify, NaN dropping
realization average        hand values with the (sum valid + 0.1) divisor
trajectory path            t_traj = running sum of accepted steps by the
                           single-step function; km/s factor; dt / T
                           normalisation sums to 1 inside the grid
ADMM at full shape         torch FP64 restatement (tests/fp64_admm.py)
                           equals this one on a small problem
=========================  ===================================================

Pins live in ``tests/test_oracle_*.py``.

Parts with no independent pin are listed in ``UNPINNED`` below.
"""

# Functions whose exact arithmetic no check here can pin.  The reference's
# JAX PRNG streams (keys 42/109/2026/911) cannot be reproduced without JAX,
# so random normals / uniforms are explicit inputs everywhere instead.
UNPINNED = (
    "JAX threefry PRNG streams (replaced by explicit noise inputs)",
    "XLA summation order of segment_sum / trapezoid / cholesky",
)
