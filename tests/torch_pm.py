"""FP64 CPU restatement of the differentiable PM forward map in torch.

TEST INFRASTRUCTURE ONLY: the reference gradient that the GPU adjoint
(pm_diff_backward) is checked against.  It is never called by the library.

    noise w -> scale-free field            (src/ic/ic.cc:37-83)
            -> potential                   (src/ic/ic.cc:100-117)
            -> Zel'dovich start            (src/run/nbody.hh:26-51, 88-95; numeric.hh:48-70)
            -> leapfrog drift / kick       (src/run/leapfrog.hh:33-46, nbody.hh:97-133)
            -> final CIC density           (src/run/numeric.hh:106-139)

The gather interpolates at folded coordinates (x mod L), as the FP32 form of
the library does.  Its forward output is pinned against the C++ oracle in
tests/test_torch_reference.py; torch autograd of it is the FP64 gradient.
Arrays are flat with axis 0 fastest; reshaped to (N,)*R, axis k is tensor
dimension R-1-k.
"""
from __future__ import annotations

import math

import torch

DT = torch.float64


def _wave(N: int, Lk: float) -> torch.Tensor:
    # src/base/fourier.hh:42-55: k = (n > N/2 ? n - N : n) * 2 pi / L
    n = torch.arange(N)
    return torch.where(n > N // 2, n - N, n).to(DT) * (2 * math.pi / Lk)


def _kaxes(R: int, N: int, Lk: float):
    """K[k]: wave number along axis k, broadcastable over (N,)*R."""
    out = []
    for k in range(R):
        shape = [1] * R
        shape[R - 1 - k] = N
        out.append(_wave(N, Lk).reshape(shape))
    return out


def _ksqr(K) -> torch.Tensor:
    s = K[0] ** 2
    for k in K[1:]:
        s = s + k ** 2
    return s


def ic_fields(w, R, mbits, L, ns, smooth=False, scale=8.0, sigma=0.83):
    """Scale-free density and its potential on the mass grid (flat)."""
    N = 1 << mbits
    W = w.reshape((N,) * R)
    K = _kaxes(R, N, float(N))  # kspace<R>(N, N): grid units
    k2 = _ksqr(K)
    kk = torch.sqrt(k2)
    P = torch.where(k2 > 0, torch.sqrt(torch.where(k2 > 0, kk, torch.ones_like(kk)) ** ns), torch.zeros_like(kk))
    t = scale * N / L
    S = torch.ones_like(kk)
    for k in range(R):
        S = S * torch.exp(t * (torch.cos(K[k]) - 1))
    fB = P * S
    fA = fB if smooth else P
    Wh = torch.fft.fftn(W)
    uB = torch.fft.ifftn(Wh * fB).real
    var = (uB ** 2).mean()
    delta = torch.fft.ifftn(Wh * fA).real * (sigma / torch.sqrt(var))
    Kp = _kaxes(R, N, L)
    k2p = _ksqr(Kp)
    g = torch.where(k2p > 0, -1.0 / torch.where(k2p > 0, k2p, torch.ones_like(k2p)), torch.zeros_like(k2p))
    phi = torch.fft.ifftn(torch.fft.fftn(delta) * g).real
    return delta.reshape(-1), phi.reshape(-1)


def _fd_ref(f, k: int, R: int, res: float):
    # gradient_2nd_order (numeric.hh:96-101): ((f-2/12 - 2 f-1/3) + 2 f+1/3) - f+2/12, / res
    d = R - 1 - k
    fm2, fm1 = torch.roll(f, 2, d), torch.roll(f, 1, d)
    fp1, fp2 = torch.roll(f, -1, d), torch.roll(f, -2, d)
    return (fm2 / 12 - 2 * fm1 / 3 + 2 * fp1 / 3 - fp2 / 12) / res


def zeldovich(phi, R, Nm, L, a_pos, a_mom):
    """X = q res - a_pos u, P = -a_mom u with u = D phi / res (nbody.hh:40-50); (R, Np)."""
    res = L / Nm
    F = phi.reshape((Nm,) * R)
    X, P = [], []
    for k in range(R):
        u = _fd_ref(F, k, R, res).reshape(-1)
        idx = torch.arange(Nm ** R)
        q = ((idx >> (k * int(math.log2(Nm)))) & (Nm - 1)).to(DT)
        X.append(q * res - u * a_pos)
        P.append(-u * a_mom)
    return torch.stack(X), torch.stack(P)


def _cic(Y, R, Nf, L):
    res = L / Nf
    p = torch.remainder(Y, L) / res
    o = torch.floor(p.detach())
    f = p - o
    o = o.long()
    corners = []
    for j in range(1 << R):
        w = torch.ones_like(f[0])
        idx = torch.zeros_like(o[0])
        for k in range(R):
            b = (j >> k) & 1
            w = w * (f[k] if b else 1 - f[k])
            idx = idx + ((o[k] + b) % Nf) * (Nf ** k)
        corners.append((w, idx))
    return corners


def deposit(Y, R, Nf, L, mass):
    rho = torch.zeros(Nf ** R, dtype=DT)
    for w, idx in _cic(Y, R, Nf, L):
        rho = rho.index_add(0, idx, w * mass)
    return rho


def poisson(rho, R, Nf, L, a, G):
    Kp = _kaxes(R, Nf, L)
    k2 = _ksqr(Kp)
    g = torch.where(k2 > 0, -1.0 / torch.where(k2 > 0, k2, torch.ones_like(k2)), torch.zeros_like(k2))
    phi = torch.fft.ifftn(torch.fft.fftn((rho - 1).reshape((Nf,) * R)) * g).real * (G / a)
    return phi


def gather(phi, Y, R, Nf, L):
    res = L / Nf
    grads = [_fd_ref(phi, k, R, res).reshape(-1) for k in range(R)]
    F = [torch.zeros_like(Y[0]) for _ in range(R)]
    for w, idx in _cic(Y, R, Nf, L):
        for k in range(R):
            F[k] = F[k] + w * grads[k][idx]
    return torch.stack(F)


def adot(a, H0, Om, OL):
    Ok = 1 - Om - OL
    return H0 * a * math.sqrt(OL + Om * a ** -3 + Ok * a ** -2)


def forward(w, R, mbits, f, L, ns, a0, da, nsteps, H0=68.5, Om=1.0, OL=0.0, smooth=False, scale=8.0,
            sigma=0.83):
    """Final CIC density (flat, (f N)^R) of the run from noise w."""
    Nm = 1 << mbits
    Nf = f * Nm
    G = 1.5 * Om * H0 * H0
    mass = (Nf / Nm) ** R
    _, phi0 = ic_fields(w, R, mbits, L, ns, smooth, scale, sigma)
    Y, P = zeldovich(phi0, R, Nm, L, a0, a0 + da / 2)
    a_seq = [a0]
    for _ in range(nsteps):
        a_seq.append(a_seq[-1] + da)
    cd = lambda a: da / (a * a * adot(a, H0, Om, OL))  # noqa: E731
    Y = Y + cd(a_seq[0]) * P
    for i in range(nsteps):
        ak = a_seq[i] + da / 2
        phi = poisson(deposit(Y, R, Nf, L, mass), R, Nf, L, ak, G)
        P = P - gather(phi, Y, R, Nf, L) * (da / adot(ak, H0, Om, OL))
        if i + 1 < nsteps:
            Y = Y + cd(a_seq[i + 1]) * P
    return deposit(Y, R, Nf, L, mass), Y, P
