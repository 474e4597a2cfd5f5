"""Host-side draws from the backend's Philox4x32-10 stream (vectorised).

Used where the reference's host layer draws from R's generator: initial
values (R/api.R:306-307, ``rnorm``) and parallel-tempering swap uniforms
(R/adaptive.R:530-546, ``runif``).  Same generator and uniform map as the
kernel (csrc/philox.h); the counter's fourth word is the purpose tag.
"""
from __future__ import annotations

import numpy as np

TAG_STEP, TAG_SWAP, TAG_INIT, TAG_DA, TAG_DATA, TAG_MLP = 0, 1, 2, 3, 4, 5

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox(ctr: np.ndarray, key0, key1) -> np.ndarray:
    """Philox4x32-10 of an (n, 4) uint32 counter array under one key."""
    c = np.asarray(ctr, dtype=np.uint32).reshape(-1, 4)
    x0, x1, x2, x3 = (c[:, i].astype(np.uint64) for i in range(4))
    k0 = np.uint32(key0 & 0xFFFFFFFF)
    k1 = np.uint32(key1 & 0xFFFFFFFF)
    for _ in range(10):
        p0 = _M0 * x0
        p1 = _M1 * x2
        y0 = (p1 >> np.uint64(32)) ^ x1 ^ np.uint64(k0)
        y1 = p1 & _MASK
        y2 = (p0 >> np.uint64(32)) ^ x3 ^ np.uint64(k1)
        y3 = p0 & _MASK
        x0, x1, x2, x3 = y0, y1, y2, y3
        k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
    return np.stack([x0, x1, x2, x3], axis=1).astype(np.uint32)


def to_unit(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.float64) + 0.5) / 4294967296.0


def normals(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n standard normals: Box-Muller cos branch of words (0, 1) of block k."""
    k = np.arange(n, dtype=np.uint32)
    ctr = np.stack([k, np.zeros_like(k), np.zeros_like(k),
                    np.full_like(k, TAG_INIT)], axis=1)
    w = philox(ctr, seed, stream)
    u1, u2 = to_unit(w[:, 0]), to_unit(w[:, 1])
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def uniforms(seed: int, batch: int, n: int, stream: int = 0,
             group: int = 0) -> np.ndarray:
    """n uniforms for swap step `batch` of ladder `group`: word 0 of Philox
    block (batch, k, group, TAG_SWAP), k = position in the sweep."""
    k = np.arange(n, dtype=np.uint32)
    ctr = np.stack([np.full_like(k, batch & 0xFFFFFFFF), k,
                    np.full_like(k, group & 0xFFFFFFFF),
                    np.full_like(k, TAG_SWAP)], axis=1)
    return to_unit(philox(ctr, seed, stream)[:, 0])
