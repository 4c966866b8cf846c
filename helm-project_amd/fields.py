"""Seeded synthetic fields for the regrid workloads.

RNG: counter-based SplitMix64 (value k of stream `seed` is the SplitMix64
finaliser of seed + (k + 1) * 0x9E3779B97F4A7C15), uniforms from the top 53
bits, normals by Box-Muller.  Identical on every machine, so CPU and GPU runs
see the same inputs.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """U[0, 1) doubles."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """N(0, 1) by Box-Muller on consecutive uniform pairs."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, 2 * offset)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * math.pi * u2)
    z[1::2] = r * np.sin(2.0 * math.pi * u2)
    return z[:n]


def cfg1_field(lon_deg: np.ndarray, lat_deg: np.ndarray, seed: int = 1001) -> np.ndarray:
    """f = 2 + cos^2(lat) cos(2 lon) + N(0, 0.01^2) at source-cell centres."""
    lo, la = np.radians(lon_deg), np.radians(lat_deg)
    return 2.0 + np.cos(la) ** 2 * np.cos(2.0 * lo) + 0.01 * normal(seed, lon_deg.shape[0])


def harmonic_coeffs(seed: int, lmax: int = 32) -> np.ndarray:
    """Real SH coefficients a[l, m'] (m' = m + lmax, |m| <= l) ~ N(0,1)/(l+1)."""
    z = normal(seed, (lmax + 1) * (2 * lmax + 1)).reshape(lmax + 1, 2 * lmax + 1)
    l = np.arange(lmax + 1)[:, None]
    m = np.arange(-lmax, lmax + 1)[None, :]
    return np.where(np.abs(m) <= l, z / (l + 1.0), 0.0)


def harmonic_sum(xyz: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """Σ_{l<=L, |m|<=l} a_lm Y_lm(p) for unit vectors p [n, 3] (float64).

    Y_lm are orthonormal real spherical harmonics, evaluated with the
    standard fully normalised associated-Legendre recurrence in z = cos θ.
    """
    lmax = coeffs.shape[0] - 1
    z = np.clip(xyz[:, 2], -1.0, 1.0)
    phi = np.arctan2(xyz[:, 1], xyz[:, 0])
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    out = np.zeros(xyz.shape[0])
    pmm = np.full(xyz.shape[0], math.sqrt(1.0 / (4.0 * math.pi)))  # P̄_00
    for m in range(lmax + 1):
        if m > 0:
            pmm = -math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * s * pmm
        cm = np.cos(m * phi)
        sm = np.sin(m * phi)
        fac = math.sqrt(2.0) if m > 0 else 1.0
        # l = m
        p_lm2 = None
        p_lm1 = pmm
        out += coeffs[m, lmax + m] * fac * p_lm1 * cm
        if m > 0:
            out += coeffs[m, lmax - m] * fac * p_lm1 * sm
        for l in range(m + 1, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            if p_lm2 is None:
                p_l = a * z * p_lm1
            else:
                b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                p_l = a * (z * p_lm1 - b * p_lm2)
            p_lm2, p_lm1 = p_lm1, p_l
            out += coeffs[l, lmax + m] * fac * p_l * cm
            if m > 0:
                out += coeffs[l, lmax - m] * fac * p_l * sm
    return out


def harmonic_sum_torch(xyz, coeffs: np.ndarray):
    """harmonic_sum for a float64 torch tensor xyz [n, 3] (same recurrence,
    evaluated on the tensor's device)."""
    import torch

    lmax = coeffs.shape[0] - 1
    z = torch.clamp(xyz[:, 2], -1.0, 1.0)
    phi = torch.atan2(xyz[:, 1], xyz[:, 0])
    s = torch.sqrt(torch.clamp(1.0 - z * z, min=0.0))
    out = torch.zeros_like(z)
    pmm = torch.full_like(z, math.sqrt(1.0 / (4.0 * math.pi)))
    for m in range(lmax + 1):
        if m > 0:
            pmm = -math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * s * pmm
        cm, sm = torch.cos(m * phi), torch.sin(m * phi)
        fac = math.sqrt(2.0) if m > 0 else 1.0
        p_lm2, p_lm1 = None, pmm
        out += float(coeffs[m, lmax + m]) * fac * p_lm1 * cm
        if m > 0:
            out += float(coeffs[m, lmax - m]) * fac * p_lm1 * sm
        for l in range(m + 1, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            if p_lm2 is None:
                p_l = a * z * p_lm1
            else:
                b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                p_l = a * (z * p_lm1 - b * p_lm2)
            p_lm2, p_lm1 = p_lm1, p_l
            out += float(coeffs[l, lmax + m]) * fac * p_l * cm
            if m > 0:
                out += float(coeffs[l, lmax - m]) * fac * p_l * sm
    return out


def harmonic_sums_torch(xyz, coeff_list):
    """[F, n] float64: harmonic_sum_torch for F coefficient sets at once (one
    Legendre recurrence, a rank-1 update per (l, m) term)."""
    import torch

    C = np.stack(coeff_list)  # [F, L+1, 2L+1]
    lmax = C.shape[1] - 1
    Ct = torch.tensor(C, dtype=torch.float64, device=xyz.device)
    z = torch.clamp(xyz[:, 2], -1.0, 1.0)
    phi = torch.atan2(xyz[:, 1], xyz[:, 0])
    s = torch.sqrt(torch.clamp(1.0 - z * z, min=0.0))
    out = torch.zeros((C.shape[0], z.shape[0]), dtype=torch.float64, device=xyz.device)
    pmm = torch.full_like(z, math.sqrt(1.0 / (4.0 * math.pi)))
    for m in range(lmax + 1):
        if m > 0:
            pmm = -math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * s * pmm
        fac = math.sqrt(2.0) if m > 0 else 1.0
        cm, sm = fac * torch.cos(m * phi), fac * torch.sin(m * phi)
        p_lm2, p_lm1 = None, pmm
        for l in range(m, lmax + 1):
            if l > m:
                a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
                if p_lm2 is None:
                    p_l = a * z * p_lm1
                else:
                    b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                    p_l = a * (z * p_lm1 - b * p_lm2)
                p_lm2, p_lm1 = p_lm1, p_l
            out.addr_(Ct[:, l, lmax + m], p_lm1 * cm)
            if m > 0:
                out.addr_(Ct[:, l, lmax - m], p_lm1 * sm)
    return out


def level_scales(seed: int, n: int) -> np.ndarray:
    """Per-level scale 10^U(-3, 3)."""
    return 10.0 ** (6.0 * uniform(seed, n) - 3.0)
