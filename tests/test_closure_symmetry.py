"""The square-symmetry augmentation of closure training (CPU, with the oracle).

Training maps each batch by an element of the square's symmetry group
(dedalus.closure.pipeline.square_symmetry).  That is only right if the
sub-grid target tau = B(NL(u)) - NL(B(u)) -- NL the projected nonlinear term
of the reference's IncompressibleHydro RHS, B the block-mean coarse-graining
-- and the input features (u, v, omega) of B(u) transform as the map says.
Here the oracle computes both for a random fine field and for its image
under each of the 8 elements, and they must agree to rounding.
"""

import numpy as np
import pytest
import torch

from helpers import random_k
from oracle import numpy_path as npth

TWO_PI = 2 * np.pi
FINE, FACTOR = (32, 32), 4
COARSE = (FINE[0] // FACTOR, FINE[1] // FACTOR)


def _x(kc, shape):
    c = npth.Component(shape, (TWO_PI, TWO_PI))
    c['kspace'] = kc
    return c['xspace'].copy()


def _k(x, shape):
    c = npth.Component(shape, (TWO_PI, TWO_PI))
    c['xspace'] = x
    return c['kspace'].copy()


def _nl(uk, shape):
    """-P[(u.grad)u] in x space (the RHS with nu = 0; viscosity sits in the
    integrating factor, physics.py:504-586)."""
    ph = npth.IncompressibleHydro(shape)
    ph.parameters['nu'] = 0.
    data = npth.load_state(ph, {'u': uk})
    deriv = ph.create_fields()
    ph.RHS(data, deriv)
    return [deriv['u'][i]['xspace'].copy() for i in range(2)]


def _block(f):
    h, w = f.shape
    return f.reshape(h // FACTOR, FACTOR, w // FACTOR, FACTOR).mean(axis=(1, 3))


def _target(ux):
    """(features (H, W, 3), tau (H, W, 2)) of a fine x-space velocity."""
    uk = [_k(a, FINE) for a in ux]
    nlf = _nl(uk, FINE)
    ubk = [_k(_block(a), COARSE) for a in ux]       # B: block mean, coarse transform
    nlc = _nl(ubk, COARSE)
    tau = np.stack([_block(nlf[i]) - nlc[i] for i in range(2)], axis=-1)
    c = npth.Component(COARSE, (TWO_PI, TWO_PI))
    kx, ky = c.k['x'], c.k['y']
    wk = 1j * kx * ubk[1] - 1j * ky * ubk[0]
    feats = np.stack([_x(ubk[0], COARSE), _x(ubk[1], COARSE), _x(wk, COARSE)], axis=-1)
    return feats, tau


@pytest.mark.parametrize("g", range(8))
def test_subgrid_target_is_equivariant(g):
    from dedalus.closure.pipeline import square_symmetry
    ux = [_x(random_k(FINE, 60 + i, slope=1.0), FINE) for i in range(2)]
    feats, tau = _target(ux)
    img = torch.tensor(np.stack(ux, axis=-1)[None])
    ux2 = square_symmetry(img, g)[0].numpy()
    feats2, tau2 = _target([ux2[..., 0], ux2[..., 1]])
    want_f = square_symmetry(torch.tensor(feats[None]), g, pseudo=(2,))[0].numpy()
    want_t = square_symmetry(torch.tensor(tau[None]), g)[0].numpy()
    scale = np.abs(tau).max()
    assert scale > 1e-3
    assert np.abs(tau2 - want_t).max() < 1e-12 * max(1.0, scale)
    assert np.abs(feats2 - want_f).max() < 1e-12 * max(1.0, np.abs(feats).max())
    if g:
        # the map moves the data: an identity would pass trivially
        assert np.abs(want_t - tau).max() > 1e-3 * scale


def test_square_symmetry_is_a_group_action():
    """g = 0 is the identity and each element has an inverse within the
    eight (the maps are bijections on the images)."""
    from dedalus.closure.pipeline import square_symmetry
    a = torch.randn(2, 6, 6, 3, dtype=torch.float64)
    assert torch.equal(square_symmetry(a, 0, pseudo=(2,)), a)
    for g in range(8):
        b = square_symmetry(a, g, pseudo=(2,))
        back = [h for h in range(8) if torch.equal(square_symmetry(b, h, pseudo=(2,)), a)]
        assert len(back) == 1, g
