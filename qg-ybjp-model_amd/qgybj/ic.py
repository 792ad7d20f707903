"""Seeded synthetic initial conditions.

Eddies: random-phase streamfunction with an isotropic Gaussian-shell
spectrum, the shape of the reference's generic initialiser
(init.f90:472-507, amp_width = 2):
    psi = sum_{k in [-K,K]^2, k != 0} exp(-(|k|-k0)^2/(2*2)) cos(kx x + ky y + kz z_s + phi_k),
kz = |k_h|/sqrt(Bu), phases phi_k from numpy's PCG64 with the given seed,
normalised to unit RMS horizontal velocity.
Waves: horizontally uniform B = L+A, BR = exp(-((z_s - L3)/(0.1 L3))^2),
BI = 0 (a surface-trapped layer).
Fields are rounded through float32, the reference's input file format
(IO_ncf.f90:230, 305).
"""
from __future__ import annotations

import math

import numpy as np

TWOPI = 4.0 * math.asin(1.0)


def _grid(n1: int, n3: int):
    x = TWOPI * np.arange(n1) / n1
    dz = TWOPI / n3
    zs = TWOPI * np.arange(1, n3 + 1) / n3 - dz / 2
    return x, zs


def eddy_psi(n1: int, n3: int, k0: int, seed: int, Bu: float = 1.0, amp_width: float = 2.0) -> np.ndarray:
    """(n3, n1, n1) float64 physical streamfunction (rounded through float32)."""
    _, zs = _grid(n1, n3)
    K = min(4 * k0, n1 // 2 - 1)
    rng = np.random.default_rng(seed)
    ks = np.arange(-K, K + 1)
    kx, ky = np.meshgrid(ks, ks, indexing="xy")
    kx = kx.ravel()
    ky = ky.ravel()
    keep = (kx != 0) | (ky != 0)
    kx = kx[keep]
    ky = ky[keep]
    kh = np.sqrt(kx * kx + ky * ky)
    amp = np.exp(-((kh - k0) ** 2) / (2.0 * amp_width))
    phi = rng.uniform(0.0, TWOPI, size=kx.shape[0])
    kz = kh / math.sqrt(Bu)
    # spectrum on the full FFT grid: psi = Re sum c_k e^{i k.x}
    G = np.zeros((n3, n1, n1), dtype=np.complex128)
    c = amp[None, :] * np.exp(1j * (kz[None, :] * zs[:, None] + phi[None, :]))  # (n3, nmodes)
    iy = np.mod(ky, n1)
    ix = np.mod(kx, n1)
    iym = np.mod(-ky, n1)
    ixm = np.mod(-kx, n1)
    for m in range(kx.shape[0]):
        G[:, iy[m], ix[m]] += 0.5 * c[:, m]
        G[:, iym[m], ixm[m]] += 0.5 * np.conj(c[:, m])
    psi = np.real(np.fft.ifft2(G, axes=(-2, -1))) * (n1 * n1)
    # unit RMS horizontal velocity
    kk = np.fft.fftfreq(n1, 1.0 / n1)
    KX, KY = np.meshgrid(kk, kk, indexing="xy")
    u = np.real(np.fft.ifft2(-1j * KY * G, axes=(-2, -1))) * (n1 * n1)
    v = np.real(np.fft.ifft2(1j * KX * G, axes=(-2, -1))) * (n1 * n1)
    rms = math.sqrt(float(np.mean(u * u + v * v)))
    if rms > 0:
        psi = psi / rms
    return psi.astype(np.float32).astype(np.float64)


def wave_b(n1: int, n3: int, depth: float = 0.1) -> tuple[np.ndarray, np.ndarray]:
    """(BR, BI) physical fields (n3, n1, n1), surface-trapped Gaussian."""
    _, zs = _grid(n1, n3)
    prof = np.exp(-(((zs - TWOPI) / (depth * TWOPI)) ** 2))
    BR = np.broadcast_to(prof[:, None, None], (n3, n1, n1)).astype(np.float32).astype(np.float64)
    return np.ascontiguousarray(BR), np.zeros((n3, n1, n1))


def perturbed_waves(n1: int, n3: int, seed: int, amp: float = 0.2, depth: float = 0.1):
    """Surface-trapped waves with a seeded horizontal modulation (used by
    tests that need horizontal structure in B from the first step)."""
    BR, BI = wave_b(n1, n3, depth)
    rng = np.random.default_rng(seed + 1000)
    x = TWOPI * np.arange(n1) / n1
    X, Y = np.meshgrid(x, x, indexing="xy")
    mod = np.zeros((n1, n1))
    for _ in range(4):
        kx, ky = rng.integers(1, 4, size=2)
        mod += np.cos(kx * X + ky * Y + rng.uniform(0, TWOPI))
    BR2 = BR * (1.0 + amp * mod[None, :, :])
    BI2 = BR * (amp * np.sin(X + 2 * Y)[None, :, :])
    return (BR2.astype(np.float32).astype(np.float64), BI2.astype(np.float32).astype(np.float64))
