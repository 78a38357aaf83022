"""CPU restatement of the divergence cleanup of E and the Fourier short-wave filtering of the new
fields (numpy, FP64), on the reference's ghosted arrays of one periodic rank.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; the product path never imports it.

- cleanup_divergence follows Simulation::cleanupDivergence (PIC++/divergencecleanup.cpp:20-392):
  right part -(4 pi rho - div E) per owned cell (cleanUpRightPart :832-837, evaluateDivEgeneral
  PIC++/fields.cpp:2967-2978), the 27-point operator of createDivergenceCleanupInternalEquation
  (:620-813), the node gradient of evaluateDivergenceCleaningField (:553-618) subtracted from E
  (updateFieldByCleaning :531-551). The reference solves for the potential with GMRES to a
  relative tolerance; this restatement solves the periodic system exactly in Fourier space (the
  operator is circulant: its eigenvalues are its stencil's symbol), with the constant mode of the
  potential set to zero. The two agree to the reference's solver tolerance.
- cleanup_divergence_magnetic follows Simulation::cleanupDivergenceMagnetic (:939-1306): the same
  operator for a node potential with div B at the nodes as right part, and the cell gradient
  subtracted from B; solved exactly here, with BiCGStab in the reference.
- filter_fields follows Simulation::filterFields / filterFieldGeneral, periodic branch
  (PIC++/fields_filtering.cpp:21-295) with fourierTranslation (PIC++/fourier.cpp:13-192): forward
  DFT of each component over the owned points, modes with |k| > 2 pi / (cut dx) multiplied by -1
  (k from the unfolded mode indices; the y component's k_y from the z index), inverse DFT, real
  part back; then the ghost exchanges.
"""
from __future__ import annotations

import numpy as np


def _periodic_nodes(owned):
    n = owned.shape[:3]
    idx = [np.mod(np.arange(v + 5) - 2, v) for v in n]
    return owned[np.ix_(idx[0], idx[1], idx[2])]


def _periodic_cells(owned):
    n = owned.shape[:3]
    idx = [np.mod(np.arange(v + 4) - 2, v) for v in n]
    return owned[np.ix_(idx[0], idx[1], idx[2])]


def div_e(E, d):
    """evaluateDivEgeneral at every cell of a ghosted node field E (N+1, N+1, N+1, 3) -> (N, N, N)."""
    Ex, Ey, Ez = E[..., 0], E[..., 1], E[..., 2]
    rx = (Ex[1:, :-1, :-1] + Ex[1:, 1:, :-1] + Ex[1:, :-1, 1:] + Ex[1:, 1:, 1:]) / 4.0
    lx = (Ex[:-1, :-1, :-1] + Ex[:-1, 1:, :-1] + Ex[:-1, :-1, 1:] + Ex[:-1, 1:, 1:]) / 4.0
    ry = (Ey[:-1, 1:, :-1] + Ey[1:, 1:, :-1] + Ey[:-1, 1:, 1:] + Ey[1:, 1:, 1:]) / 4.0
    ly = (Ey[:-1, :-1, :-1] + Ey[1:, :-1, :-1] + Ey[:-1, :-1, 1:] + Ey[1:, :-1, 1:]) / 4.0
    rz = (Ez[:-1, :-1, 1:] + Ez[1:, :-1, 1:] + Ez[:-1, 1:, 1:] + Ez[1:, 1:, 1:]) / 4.0
    lz = (Ez[:-1, :-1, :-1] + Ez[1:, :-1, :-1] + Ez[:-1, 1:, :-1] + Ez[1:, 1:, :-1]) / 4.0
    return (rx - lx) / d[0] + (ry - ly) / d[1] + (rz - lz) / d[2]


def cleanup_stencil(d):
    """The 27 weights of createDivergenceCleanupInternalEquation (3-D, _27points), [a+1][b+1][c+1]."""
    X, Y, Z = (v * v for v in d)
    w = np.zeros((3, 3, 3))
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                nz = (a != 0) + (b != 0) + (c != 0)
                if nz == 0:
                    v = -1 / (2 * X) - 1 / (2 * Y) - 1 / (2 * Z)
                elif nz == 3:
                    v = 1 / (16 * X) + 1 / (16 * Y) + 1 / (16 * Z)
                else:
                    f = 8 if nz == 2 else 4
                    v = sum((1 if o else -1) / (f * s) for o, s in ((a, X), (b, Y), (c, Z)))
                w[a + 1, b + 1, c + 1] = v
    return w


def apply_cleanup_operator(phi, d):
    """L phi on the owned cells, periodic."""
    w = cleanup_stencil(d)
    out = np.zeros_like(phi)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                out += w[a + 1, b + 1, c + 1] * np.roll(phi, (-a, -b, -c), axis=(0, 1, 2))
    return out


def grad_correction(phi, d):
    """evaluateDivergenceCleaningField at the owned nodes from owned-cell phi (periodic images):
    (n0, n1, n2, 3)."""
    P = {}
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                P[di, dj, dk] = np.roll(phi, (1 - di, 1 - dj, 1 - dk), axis=(0, 1, 2))
    cx = -(P[1, 1, 1] + P[1, 0, 1] + P[1, 1, 0] + P[1, 0, 0] - P[0, 1, 1] - P[0, 0, 1] - P[0, 1, 0] - P[0, 0, 0]) / (4 * d[0])
    cy = -(P[1, 1, 1] + P[0, 1, 1] + P[1, 1, 0] + P[0, 1, 0] - P[1, 0, 1] - P[0, 0, 1] - P[1, 0, 0] - P[0, 0, 0]) / (4 * d[1])
    cz = -(P[1, 1, 1] + P[1, 0, 1] + P[0, 1, 1] + P[0, 0, 1] - P[1, 1, 0] - P[1, 0, 0] - P[0, 1, 0] - P[0, 0, 0]) / (4 * d[2])
    return np.stack([cx, cy, cz], -1)


def cleanup_divergence(E, rho, d):
    """E (ghosted nodes, N+1 per axis, 3), rho (ghosted cells) -> cleaned E (ghosts exchanged) and
    the potential on the owned cells."""
    o = slice(2, -2)
    rhs = -(4 * np.pi * rho[o, o, o] - div_e(E, d)[o, o, o])
    phi = _solve_periodic(rhs, d)
    owned = E[2:-3, 2:-3, 2:-3] + grad_correction(phi, d)
    return _periodic_nodes(owned), phi


def div_b(B, d):
    """evaluateDivBgeneral at every node i = 1 .. N-1 of a ghosted cell field B (N, N, N, 3):
    (N-1, N-1, N-1), index i - 1 for node i."""
    Bx, By, Bz = B[..., 0], B[..., 1], B[..., 2]
    rx = (Bx[1:, :-1, :-1] + Bx[1:, 1:, :-1] + Bx[1:, :-1, 1:] + Bx[1:, 1:, 1:]) / 4.0
    lx = (Bx[:-1, :-1, :-1] + Bx[:-1, 1:, :-1] + Bx[:-1, :-1, 1:] + Bx[:-1, 1:, 1:]) / 4.0
    ry = (By[:-1, 1:, :-1] + By[1:, 1:, :-1] + By[:-1, 1:, 1:] + By[1:, 1:, 1:]) / 4.0
    ly = (By[:-1, :-1, :-1] + By[1:, :-1, :-1] + By[:-1, :-1, 1:] + By[1:, :-1, 1:]) / 4.0
    rz = (Bz[:-1, :-1, 1:] + Bz[1:, :-1, 1:] + Bz[:-1, 1:, 1:] + Bz[1:, 1:, 1:]) / 4.0
    lz = (Bz[:-1, :-1, :-1] + Bz[1:, :-1, :-1] + Bz[:-1, 1:, :-1] + Bz[1:, 1:, :-1]) / 4.0
    return (rx - lx) / d[0] + (ry - ly) / d[1] + (rz - lz) / d[2]


def _solve_periodic(rhs, d):
    n = rhs.shape
    w = cleanup_stencil(d)
    k = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(v) for v in n], indexing="ij")
    lam = np.zeros(n, complex)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                lam += w[a + 1, b + 1, c + 1] * np.exp(1j * (a * k[0] + b * k[1] + c * k[2]))
    R = np.fft.fftn(rhs)
    small = np.abs(lam) < 1e-12 * np.max(np.abs(lam))
    lam[small] = 1.0
    R[small] = 0.0
    return np.fft.ifftn(R / lam).real


def cleanup_divergence_magnetic(B, d):
    """Simulation::cleanupDivergenceMagnetic (PIC++/divergencecleanup.cpp:939-1306) of a ghosted cell
    field B: right part div B at the owned nodes (-cleanUpRightPartMagnetic), the same 27-point
    operator for a node potential, B -= grad phi at the owned cells
    (evaluateDivergenceCleaningFieldMagnetic); exact periodic solve as in cleanup_divergence (the
    reference uses BiCGStab to 1e-8). Returns the cleaned B (ghosts exchanged) and phi."""
    o = slice(1, -2)  # div_b index i - 1: owned nodes 2 .. n+1
    phi = _solve_periodic(div_b(B, d)[o, o, o], d)
    P = {}
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                P[di, dj, dk] = np.roll(phi, (-di, -dj, -dk), axis=(0, 1, 2))
    cx = -(P[1, 0, 0] + P[1, 1, 0] + P[1, 0, 1] + P[1, 1, 1] - P[0, 0, 0] - P[0, 1, 0] - P[0, 0, 1] - P[0, 1, 1]) / (4 * d[0])
    cy = -(P[0, 1, 0] + P[1, 1, 0] + P[0, 1, 1] + P[1, 1, 1] - P[0, 0, 0] - P[1, 0, 0] - P[0, 0, 1] - P[1, 0, 1]) / (4 * d[1])
    cz = -(P[0, 0, 1] + P[0, 1, 1] + P[1, 0, 1] + P[1, 1, 1] - P[0, 0, 0] - P[0, 1, 0] - P[1, 0, 0] - P[1, 1, 0]) / (4 * d[2])
    owned = B[2:-2, 2:-2, 2:-2] + np.stack([cx, cy, cz], -1)
    return _periodic_cells(owned), phi


def filter_fields(E, B, cut, d):
    """filterFields(cut) of ghosted newE (nodes) and newB (cells): the filtered fields."""
    kcut = 2 * np.pi / (cut * d[0])

    def one(F, nodes):
        out = F.copy()
        hi = -3 if nodes else -2
        for comp in range(3):
            a = F[2:hi, 2:hi, 2:hi, comp].astype(complex)
            n = a.shape
            S = np.fft.fftn(a)
            m = np.meshgrid(*[np.arange(v) for v in n], indexing="ij")
            L = np.array(n, float) * np.array(d)
            kx = m[0] * 2 * np.pi / L[0]
            ky = (m[2] if comp == 1 else m[1]) * 2 * np.pi / L[1]
            kz = m[2] * 2 * np.pi / L[2]
            S[np.sqrt(kx * kx + ky * ky + kz * kz) > kcut] *= -1
            out[2:hi, 2:hi, 2:hi, comp] = np.fft.ifftn(S).real
        owned = out[2:hi, 2:hi, 2:hi]
        return _periodic_nodes(owned) if nodes else _periodic_cells(owned)

    return one(E, True), one(B, False)
