"""The mathematics the capacitance solve rests on (f2d_capacitance.hip),
checked on the CPU with the oracle's restatement of the reference's matrix
(elliptic.py:114-195, "v") and scipy's SuperLU:

* the masked operator is the principal submatrix of the rectangle operator
  on the wet unknowns (same diagonal everywhere, couplings between unknown
  neighbours only);
* the capacitance matrix C = P_B A_R^-1 P_B^T of the dry unknowns B next to
  the wet ones, built as the device builds it -- the Green's function of the
  periodic doubled grid summed over the odd images of the closed
  directions -- equals SuperLU's columns of A_R^-1 at B;
* x = A_R^-1 (b_W + P_B^T s), s = (-C)^-1 (A_R^-1 b_W)_B solves the masked
  problem, with x = 0 on B.

A planted mistake (an even image instead of the odd one) must fail."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import oracle as O
from helpers import P, random_mask


def _geometry(ny, nx, xper, yper, maindiag, seed):
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper,
          Lx=nx / 64.0, Ly=ny / 80.0)
    mesh = O.Mesh(p, msk=random_mask(ny, nx, seed, xper, yper, nisl=5))
    unk = O.unknown_mask(p, mesh.mskv)
    rect = O.unknown_mask(p, O.Mesh(p).mskv)
    GR, NR = O.gindex(rect)
    AR = O.laplacian(GR, NR, mesh.dx, mesh.dy, maindiag, False, xper, yper, 3)
    GW, NW = O.gindex(unk)
    AW = O.laplacian(GW, NW, mesh.dx, mesh.dy, maindiag, False, xper, yper, 3)
    NYP, NXP = unk.shape

    def wi(i):
        return (i + nx if i < 3 else (i - nx if i >= nx + 3 else i)) if xper else i

    def wj(j):
        return (j + ny if j < 3 else (j - ny if j >= ny + 3 else j)) if yper else j
    B = []
    for j in range(NYP):
        for i in range(NXP):
            if rect[j, i] and not unk[j, i] and any(
                    0 <= wj(a) < NYP and 0 <= wi(b) < NXP and unk[wj(a), wi(b)]
                    for a, b in ((j, i - 1), (j, i + 1), (j - 1, i), (j + 1, i))):
                B.append((j, i))
    return p, mesh, unk, GR, AR, GW, AW, B


def _green_capacitance(mesh, B, nx, ny, xper, yper, maindiag, odd_sign=-1.0):
    """C from the periodic doubled grid's Green's function (numpy FFT, the
    zero-mean pseudo-inverse when singular) and the images of closed
    directions, local coordinates X = i - 3, Y = j - 3 (k_cap_build)."""
    nx2, ny2 = (nx if xper else 2 * nx), (ny if yper else 2 * ny)
    cx, cy = mesh.dy / mesh.dx, mesh.dx / mesh.dy
    kx = 2 * np.pi * np.fft.fftfreq(nx2)
    ky = 2 * np.pi * np.fft.fftfreq(ny2)
    lam = 2 * cx * (np.cos(kx)[None, :] - 1) + 2 * cy * (np.cos(ky)[:, None] - 1) - maindiag
    d = np.zeros((ny2, nx2))
    d[0, 0] = 1.0
    dh = np.fft.fft2(d)
    if maindiag == 0:
        lam[0, 0], dh[0, 0] = 1.0, 0.0
    G = np.real(np.fft.ifft2(dh / lam))
    bx = np.array([b[1] - 3 for b in B])
    by = np.array([b[0] - 3 for b in B])
    C = np.zeros((len(B), len(B)))
    for sy in ([0] if yper else [0, 1]):
        dy = (by[:, None] + by[None, :]) if sy else (by[:, None] - by[None, :])
        for sx in ([0] if xper else [0, 1]):
            dx = (bx[:, None] + bx[None, :]) if sx else (bx[:, None] - bx[None, :])
            C += (odd_sign ** (sx + sy)) * G[dy % ny2, dx % nx2]
    return C


CASES = [(48, 64, False, False, 0.0, 1), (32, 64, True, False, 0.0, 2),
         (64, 32, False, True, 0.4, 3), (32, 32, True, True, 0.9, 4)]


@pytest.mark.parametrize("ny,nx,xper,yper,maindiag,seed", CASES)
def test_masked_operator_is_principal_submatrix(ny, nx, xper, yper, maindiag, seed):
    p, mesh, unk, GR, AR, GW, AW, B = _geometry(ny, nx, xper, yper, maindiag, seed)
    iW = GR[unk == 1]
    assert (GW[unk == 1] == np.arange(len(iW))).all()  # same row-major order
    assert abs(AR[iW][:, iW] - AW).max() == 0.0


@pytest.mark.parametrize("ny,nx,xper,yper,maindiag,seed", CASES)
def test_capacitance_matrix_and_solve(ny, nx, xper, yper, maindiag, seed):
    p, mesh, unk, GR, AR, GW, AW, B = _geometry(ny, nx, xper, yper, maindiag, seed)
    iB = np.array([GR[j, i] for j, i in B])
    iW = GR[unk == 1]
    LU = spl.splu(AR.tocsc())
    E = np.zeros((AR.shape[0], len(B)))
    E[iB, np.arange(len(B))] = 1.0
    Cref = LU.solve(E)[iB]
    C = _green_capacitance(mesh, B, nx, ny, xper, yper, maindiag)
    assert np.abs(C - Cref).max() <= 1e-12 * np.abs(Cref).max()
    # the solve, and x = 0 at B
    rng = np.random.default_rng(seed)
    bW = rng.standard_normal(len(iW))
    bt = np.zeros(AR.shape[0])
    bt[iW] = bW
    s = np.linalg.solve(-C, LU.solve(bt)[iB])
    bt[iB] += s
    x = LU.solve(bt)
    xref = spl.splu(AW.tocsc()).solve(bW)
    assert np.abs(x[iW] - xref).max() <= 1e-12 * np.abs(xref).max()
    assert np.abs(x[iB]).max() <= 1e-12 * np.abs(xref).max()


def test_even_images_are_wrong():
    """The closed direction's images must carry the odd (Dirichlet) sign:
    the even images of a Neumann wall give another matrix."""
    ny, nx, xper, yper, maindiag, seed = CASES[0]
    p, mesh, unk, GR, AR, GW, AW, B = _geometry(ny, nx, xper, yper, maindiag, seed)
    iB = np.array([GR[j, i] for j, i in B])
    E = np.zeros((AR.shape[0], len(B)))
    E[iB, np.arange(len(B))] = 1.0
    Cref = spl.splu(AR.tocsc()).solve(E)[iB]
    C_even = _green_capacitance(mesh, B, nx, ny, xper, yper, maindiag, odd_sign=1.0)
    assert np.abs(C_even - Cref).max() > 1e-3 * np.abs(Cref).max()
