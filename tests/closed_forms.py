"""Closed forms of restricted diffusion used as pins by the tests."""
import numpy as np
from scipy.special import jnp_zeros


def burcaw_dperp(t, a, D0, nterms=100):
    """Radial D(t) = <dx^2 + dy^2> / (4 t) of diffusion inside an impermeable
    cylinder of radius a from a uniform start: a^2/(4t) - (2a^2/t) sum_k
    exp(-b_k^2 D0 t/a^2) / (b_k^2 (b_k^2 - 1)), b_k the zeros of J1'
    (Burcaw et al. 2015, from Stepisnik 1993 and Callaghan 1995; the form of
    calculateDanalytical_longtime, tests/calibration/
    num_molecules_and_time_step_calibration/
    calibration_for_num_molecules_and_time_step_search.py:35-61)."""
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    b = jnp_zeros(1, nterms)[None, :]
    s = (np.exp(-(b ** 2) * D0 * t[:, None] / a ** 2) / (b ** 2 * (b ** 2 - 1))).sum(axis=1)
    return a ** 2 / (4 * t) - (2 * a ** 2 / t) * s


def cylinder_chain(SimGeometry3D, Structure3D, a, Lc, spacing, nimg, D0, T2=100.0):
    """A cylinder of radius a along z as a chain of spheres spanning the box
    (linspace from -Lc/2 to Lc/2 at about `spacing`) plus nimg image spheres
    beyond each end, as the calibration study builds it
    (calibration_for_num_molecules_and_time_step_search.py:115-130)."""
    g = SimGeometry3D(Lc, Lc, Lc, D0, T2, 1)
    n = int(round(Lc / spacing)) + 1
    sz = np.linspace(-Lc / 2, Lc / 2, n)
    sz = np.concatenate((sz[-nimg:] - Lc, sz, sz[:nimg] + Lc))
    g.add_structure(Structure3D(np.zeros_like(sz), np.zeros_like(sz), sz, np.full(sz.shape, a),
                                D0, T2, 1))
    return g
