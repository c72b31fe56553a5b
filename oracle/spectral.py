"""Oracle of the spectral resolve: TEST INFRASTRUCTURE ONLY (imported by tests/,
never by the product library).

A numpy FP32 restatement of the CIE 2015 10-degree colour-matching fit
(pk.h:960-1012), xyz2srgblinear (pk.h:1014-1022), srgb_oetf (pk.h:1023-1031),
packRGBA (main_gpu.cu:9-14) and the per-chain weighting of solveSpecularPath
(main_gpu.cu:473-483) applied per (receiver point, wavelength) to the
per-wavelength irradiance images.  Exponentials and powers come from numpy, so
results agree with the GPU's to a few ulp, not bit for bit.
"""
from __future__ import annotations

import numpy as np

INTEGRAL_OF_CMF_Y_IN_NM = np.float32(118.51810464018001)  # pk.h:1033
F = np.float32


def asymmetric_gaussian(x, mean, sigma, a):
    """CIE_2015_10deg::asymmetric_gaussian (pk.h:962-968)."""
    x = np.asarray(x, dtype=np.float32)
    denom = F(sigma) + F(a) * (x - F(mean))
    denom = np.where(denom < F(1.0e-15), F(1.0e-15), denom)
    denom = np.where(F(sigma) * F(2.0) < denom, F(sigma) * F(2.0), denom)
    k = (x - F(mean)) / denom
    return np.exp(-k * k).astype(np.float32)


def cmf_x(x):
    """pk.h:969-973."""
    a = F(0.42) * asymmetric_gaussian(x, 445.5849609375, 31.146467208862305, 0.06435633450746536)
    b = F(1.16) * asymmetric_gaussian(x, 594.5570068359375, 48.602108001708984, -0.04772702232003212)
    return (a + b - a * b * F(42.559776306152344)) * F(1.0008530432426956)


def cmf_y(x):
    """pk.h:974-976."""
    return F(1.0) * asymmetric_gaussian(x, 556.8383178710938, 66.54190826416016, -0.026492968201637268) * \
        F(1.004315086937574)


def cmf_z(x):
    """pk.h:977-979."""
    return F(2.146832) * asymmetric_gaussian(x, 445.9251708984375, 30.91781997680664, 0.08379141241312027) * \
        F(0.9975112815948937)


M_XYZ_TO_SRGB = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.204, 1.057]],
                         dtype=np.float32)


def xyz2srgblinear(xyz):
    """pk.h:1014-1022: three dot products, summed left to right."""
    xyz = np.asarray(xyz, dtype=np.float32)
    m = M_XYZ_TO_SRGB
    return np.stack([m[i, 0] * xyz[..., 0] + m[i, 1] * xyz[..., 1] + m[i, 2] * xyz[..., 2] for i in range(3)],
                    axis=-1)


def srgb_oetf(r):
    """pk.h:1023-1031."""
    r = np.maximum(np.asarray(r, dtype=np.float32), F(0.0))
    return np.where(r <= F(0.003130), F(12.92) * r,
                    F(1.055) * np.power(r, F(1.0) / F(2.4)).astype(np.float32) - F(0.055)).astype(np.float32)


def pack_rgba(rgb):
    """packRGBA (main_gpu.cu:9-14): int(c * 255 + 0.5) clamped to [0, 255], ABGR bytes."""
    c = np.asarray(rgb, dtype=np.float32) * F(255.0) + F(0.5)
    i = np.clip(np.trunc(c), 0, 255).astype(np.uint32)
    return np.uint32(0xFF000000) | (i[..., 2] << np.uint32(16)) | (i[..., 1] << np.uint32(8)) | i[..., 0]


def resolve(images, lambdas, weights, R, n_frames):
    """accum (n, 4) and pixels (n,) of palz_spectral_resolve."""
    images = np.asarray(images, dtype=np.float32)
    L, n = images.shape
    lam = np.asarray(lambdas, dtype=np.float32)
    xyz = np.stack([cmf_x(lam) / INTEGRAL_OF_CMF_Y_IN_NM, cmf_y(lam) / INTEGRAL_OF_CMF_Y_IN_NM,
                    cmf_z(lam) / INTEGRAL_OF_CMF_Y_IN_NM], axis=-1)
    w = xyz2srgblinear(xyz) * np.asarray(weights, dtype=np.float32)[:, None]
    acc = np.zeros((n, 3), dtype=np.float32)
    for l in range(L):
        acc = acc + w[l][None, :] * images[l][:, None]
    acc = acc * (np.asarray(R, dtype=np.float32) * F(1.0 / np.pi))
    out = np.concatenate([acc, np.full((n, 1), n_frames, dtype=np.float32)], axis=1)
    pixels = pack_rgba(srgb_oetf(acc / F(n_frames)))
    return out, pixels
