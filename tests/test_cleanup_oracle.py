"""The CPU restatement of the divergence cleanup and the Fourier filtering (oracle/cleanup_oracle.py)
against the compiled reference's golden vectors (tests/golden/cleanup_ref.npz, filter_ref.npz: the
mini case with cleanupDivergence, resp. filterFields(2), after every push) and against what the
mathematics fixes:
- the cleanup operator is the divergence of the node gradient (div of the correction = -L phi), so a
  cleaned field satisfies div E = 4 pi rho exactly;
- the filter multiplies the modes with unfolded index k = m 2 pi / L above the cut by -1: a mode m = 1
  then keeps its part at m = 1 and loses the sign of its conjugate at m = n - 1, so cos(2 pi x / L)
  becomes zero (the real part of i sin); a constant stays; and the y component reads k_y from the z
  index, so its cos(2 pi y / L_y) stays while the x component's is removed.
A sign slip, a transposed stencil weight, a wrong mode index or an off-by-one ghost fails one of them.
"""
import os

import numpy as np

from conftest import relerr
from oracle import cleanup_oracle as co

HERE = os.path.dirname(os.path.abspath(__file__))
D = (1.0, 1.0, 1.0)


def test_cleanup_matches_reference_to_its_solver_tolerance():
    d = np.load(os.path.join(HERE, "golden", "cleanup_ref.npz"))
    for step in (1, 2):
        E, rho = d["s%d/Enew" % step], d["s%d/rho" % step]
        got, _ = co.cleanup_divergence(E, rho, D)
        want = d["s%d/Enew_post" % step]
        # the reference's GMRES stops at a Givens residual of 1e-8 / (n0 n1 n2)
        assert relerr(got, want) <= 1e-9, (step, relerr(got, want))
        # and the cleanup changed E by far more than that
        assert relerr(want, E) > 1e-3


def test_magnetic_cleanup_matches_reference():
    d = np.load(os.path.join(HERE, "golden", "cleanup_eb_ref.npz"))
    # the initial B (cosine perturbations) is not solenoidal; the B update keeps div B, the cleanup
    # removes it in the first step
    B = d["s1/Bnew"]
    o = slice(1, -2)
    assert np.linalg.norm(co.div_b(B, D)[o, o, o]) > 0.1
    for step in (1, 2):
        got, _ = co.cleanup_divergence_magnetic(d["s%d/Bnew" % step], D)
        assert relerr(got, d["s%d/Bnew_post" % step]) <= 1e-12, step
        assert np.linalg.norm(co.div_b(got, D)[o, o, o]) <= 1e-13 * np.linalg.norm(got)


def test_cleaned_field_satisfies_gauss_law():
    # odd sizes: with an even size the averaging in the gradient and the divergence removes the
    # modes with k = pi across an axis, which the charge density may carry
    rng = np.random.default_rng(4)
    n = (9, 5, 7)
    E = co._periodic_nodes(rng.normal(size=n + (3,)))
    rho_o = rng.normal(size=n)
    rho_o -= rho_o.mean()  # a neutral box: the periodic system has a solution
    rho = co._periodic_cells(rho_o)
    Ec, phi = co.cleanup_divergence(E, rho, D)
    o = slice(2, -2)
    assert relerr(co.div_e(Ec, D)[o, o, o], 4 * np.pi * rho_o) <= 1e-12
    assert relerr(co.div_e(E, D)[o, o, o], 4 * np.pi * rho_o) > 0.1


def test_cleanup_operator_is_div_of_the_gradient():
    rng = np.random.default_rng(5)
    n = (7, 8, 6)
    phi = rng.normal(size=n)
    c = co._periodic_nodes(co.grad_correction(phi, D))
    o = slice(2, -2)
    assert relerr(co.div_e(c, D)[o, o, o], -co.apply_cleanup_operator(phi, D)) <= 1e-13
    # anisotropic cells: the weights carry 1/dx^2, 1/dy^2, 1/dz^2 on their own axes
    d2 = (0.7, 1.3, 0.9)
    c = co._periodic_nodes(co.grad_correction(phi, d2))
    assert relerr(co.div_e(c, d2)[o, o, o], -co.apply_cleanup_operator(phi, d2)) <= 1e-13


def test_filter_matches_reference_and_is_an_involution():
    d = np.load(os.path.join(HERE, "golden", "filter_ref.npz"))
    for step in (1, 2):
        E, B = d["s%d/Enew" % step], d["s%d/Bnew" % step]
        Ef, Bf = co.filter_fields(E, B, 2, D)
        assert relerr(Ef, d["s%d/Enew_post" % step]) <= 1e-13
        assert relerr(Bf, d["s%d/Bnew_post" % step]) <= 1e-13
        assert relerr(Ef, E) > 0.1


def test_filter_closed_forms():
    n = (6, 5, 7)
    N = tuple(v + 4 for v in n)
    g = np.meshgrid(*[np.arange(v) for v in n], indexing="ij")
    cy = np.cos(2 * np.pi * g[1] / n[1])
    E = np.zeros(n + (3,))
    E[..., 0] = cy + 0.5   # x component: cos(2 pi y / Ly) is removed, the constant stays
    E[..., 1] = cy         # y component: k_y comes from the z index, so it stays
    E[..., 2] = np.cos(2 * np.pi * g[0] / n[0])  # removed
    B = np.zeros(n + (3,))
    B[..., 1] = 2.0
    Ef, Bf = co.filter_fields(co._periodic_nodes(E), co._periodic_cells(B), 2, D)
    o = slice(2, -3)
    assert np.allclose(Ef[o, o, o, 0], 0.5, atol=1e-14)
    assert np.allclose(Ef[o, o, o, 1], cy, atol=1e-14)
    assert np.allclose(Ef[o, o, o, 2], 0.0, atol=1e-14)
    assert np.allclose(Bf[2:-2, 2:-2, 2:-2], B, atol=1e-14)
