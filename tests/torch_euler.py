"""FP64 torch restatement of the eulerpsi RK3 step on a fully wet, doubly
periodic grid (interior arrays, periodic rolls), differentiable by
torch.autograd.  Test infrastructure: the reference for the hand-written
adjoint of the library (same discretisation: WENO-Z5 fluxes weno.py:167-197,
303-364, div operators.py:104-107, c2v :126-133, Poisson elliptic.py:114-195
through its DFT closed form, perpgrad :144-149, rk3 integrators.py:92-107).
"""
import math

import torch


def weno5z(qmm, qm, q0, qp, qpp):
    eps = 1e-16
    qi1 = (1. / 3. * qmm - 7. / 6. * qm) + 11. / 6. * q0
    qi2 = (-1. / 6. * qm + 5. / 6. * q0) + 1. / 3. * qp
    qi3 = (1. / 3. * q0 + 5. / 6. * qp) - 1. / 6. * qpp
    k1, k2 = 13. / 12., 0.25
    b1 = k1 * ((qmm - 2 * qm) + q0) ** 2 + k2 * ((qmm - 4 * qm) + 3 * q0) ** 2
    b2 = k1 * ((qm - 2 * q0) + qp) ** 2 + k2 * (qm - qp) ** 2
    b3 = k1 * ((q0 - 2 * qp) + qpp) ** 2 + k2 * ((3 * q0 - 4 * qp) + qpp) ** 2
    tau = torch.abs(b1 - b3)
    w1 = 0.1 * (1 + tau / (b1 + eps))
    w2 = 0.6 * (1 + tau / (b2 + eps))
    w3 = 0.3 * (1 + tau / (b3 + eps))
    return ((w1 * qi1 + w2 * qi2) + w3 * qi3) / ((w1 + w2) + w3)


def face_flux(q, U, dim):
    """U r on the faces between cell i-1 and i along `dim` (U > 0: the
    left-biased stencil q[i-3..i+1], else the mirrored one)."""
    r = lambda s: torch.roll(q, s, dims=dim)  # noqa: E731  r(s)[i] = q[i - s]
    rpos = weno5z(r(3), r(2), r(1), q, r(-1))
    rneg = weno5z(r(-2), r(-1), q, r(1), r(2))
    return U * torch.where(U > 0, rpos, rneg)


def velocity(w, dx, dy):
    """U = perpgrad(A^-1 c2v(w)), contravariant (eulerpsi diag)."""
    ny, nx = w.shape
    b = 0.25 * (((torch.roll(w, (1, 1), (0, 1)) + torch.roll(w, 1, 1)) +
                 torch.roll(w, 1, 0)) + w)
    kx = 2 * math.pi * torch.fft.fftfreq(nx, dtype=torch.float64)
    ky = 2 * math.pi * torch.fft.fftfreq(ny, dtype=torch.float64)
    lam = (2 * (dy / dx) * (torch.cos(kx)[None, :] - 1) +
           2 * (dx / dy) * (torch.cos(ky)[:, None] - 1))
    inv = torch.where(lam != 0, 1 / torch.where(lam != 0, lam, torch.ones_like(lam)),
                      torch.zeros_like(lam))
    psi = torch.real(torch.fft.ifft2(torch.fft.fft2(b) * inv))
    Ux = -(torch.roll(psi, -1, 0) - psi) * (1 / dy ** 2)
    Uy = (torch.roll(psi, -1, 1) - psi) * (1 / dx ** 2)
    return Ux, Uy


def tendency(w, dx, dy):
    Ux, Uy = velocity(w, dx, dy)
    Fx = face_flux(w, Ux, 1)
    Fy = face_flux(w, Uy, 0)
    return -(torch.roll(Fx, -1, 1) - Fx) - (torch.roll(Fy, -1, 0) - Fy)


def rk3_step(w, dt, dx, dy):
    d1 = tendency(w, dx, dy)
    w1 = w + dt * d1
    d2 = tendency(w1, dx, dy)
    w2 = w1 + ((-3 * dt) / 4 * d1 + dt / 4 * d2)
    d3 = tendency(w2, dx, dy)
    return w2 + (((-dt) / 12 * d1 + (-dt) / 12 * d2) + (2 * dt) / 3 * d3)


def run(w0, nsteps, dt, dx, dy):
    w = w0
    for _ in range(nsteps):
        w = rk3_step(w, dt, dx, dy)
    return w
