"""Counter-based random draws shared by the host functions and the MD kernels.

The reference threads JAX threefry keys through its thermostats and
Maxwell-Boltzmann draws (dynamics/thermal.py:73-97, integrators.py:99-176);
those streams cannot be reproduced without JAX.  This module uses
Philox-4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC 2011) with the same functional key threading: a key is
two uint32 words, ``split`` derives new keys from it, and draws read a key
and a counter.  The device kernels restate exactly these functions, so a
thermostat step on the host and on the GPU draws the same numbers; the
distributions are the reference's, the streams are not.

Definitions (uint32 arithmetic):

* ``philox(ctr, key)``: 10 rounds of (c0, c1, c2, c3) ->
  (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0)) with
  M0 = 0xD2511F53, M1 = 0xCD9E8D57, the key bumped by (0x9E3779B9,
  0xBB67AE85) after each round.
* ``split(key, num)``: words of philox((i, 0, 0, TAG_SPLIT), key) for
  i = 0, 1, ..., key j = (words[2j], words[2j + 1]).
* ``uniform53(hi, lo)`` = ((hi >> 5) 2^26 + (lo >> 6)) 2^-53 in [0, 1).
* ``normals(key, count)``: pair i from philox((i, 0, 0, TAG_NORMAL), key):
  u1 = uniform53(w0, w1), u2 = uniform53(w2, w3), rho = sqrt(-2 log1p(-u1)),
  (rho cos 2 pi u2, rho sin 2 pi u2) (Box-Muller).
* ``uniform(key)`` = uniform53(w0, w1) and ``randint(key, n)`` =
  min(floor(uniform53(w2, w3) n), n - 1) of philox((0, 0, 0, TAG_UNIF), key).
"""

from __future__ import annotations

import math
from typing import List

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
TAG_SPLIT = 0x53504C54
TAG_NORMAL = 0x4E524D4C
TAG_UNIF = 0x554E4946
MASK = 0xFFFFFFFF


def prng_key(seed: int) -> np.ndarray:
    """A raw key (2,) uint32 from an integer seed (high word, low word)."""
    seed = int(seed)
    if seed < 0:
        raise ValueError(f"seed must be >= 0, got {seed}")
    return np.array([(seed >> 32) & MASK, seed & MASK], dtype=np.uint32)


def require_key(key: np.ndarray) -> None:
    """Raise TypeError unless ``key`` is a raw uint32 key of shape (2,)."""
    k = np.asarray(key)
    if k.dtype != np.uint32 or k.shape != (2,):
        raise TypeError(f"key must be a uint32 array of shape (2,), got {k.dtype} {k.shape}")


def philox(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    """Philox-4x32-10 of counters (m, 4) under one key (2,): (m, 4) uint32."""
    c = np.asarray(ctr, dtype=np.uint64).reshape(-1, 4) & MASK
    c0, c1, c2, c3 = (c[:, i].copy() for i in range(4))
    k0, k1 = int(key[0]) & MASK, int(key[1]) & MASK
    for _ in range(10):
        p0 = np.uint64(M0) * c0
        p1 = np.uint64(M1) * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c0, c1, c2, c3 = hi1 ^ c1 ^ np.uint64(k0), lo1, hi0 ^ c3 ^ np.uint64(k1), lo0
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return np.stack([c0, c1, c2, c3], axis=1).astype(np.uint32)


def split(key: np.ndarray, num: int = 2) -> List[np.ndarray]:
    """``num`` keys derived from ``key`` (the first usually replaces it)."""
    require_key(np.asarray(key, dtype=np.uint32))
    blocks = (2 * num + 3) // 4
    ctr = np.zeros((blocks, 4), dtype=np.uint64)
    ctr[:, 0] = np.arange(blocks)
    ctr[:, 3] = TAG_SPLIT
    w = philox(ctr, key).reshape(-1)
    return [np.array([w[2 * j], w[2 * j + 1]], dtype=np.uint32) for j in range(num)]


def _uniform53(hi: np.ndarray, lo: np.ndarray) -> np.ndarray:
    return ((hi.astype(np.uint64) >> np.uint64(5)).astype(np.float64) * 67108864.0
            + (lo.astype(np.uint64) >> np.uint64(6)).astype(np.float64)) * (1.0 / 9007199254740992.0)


def normals(key: np.ndarray, count: int) -> np.ndarray:
    """``count`` standard normal draws (float64)."""
    pairs = (int(count) + 1) // 2
    ctr = np.zeros((pairs, 4), dtype=np.uint64)
    ctr[:, 0] = np.arange(pairs)
    ctr[:, 3] = TAG_NORMAL
    w = philox(ctr, key)
    u1 = _uniform53(w[:, 0], w[:, 1])
    u2 = _uniform53(w[:, 2], w[:, 3])
    rho = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.empty(2 * pairs)
    z[0::2] = rho * np.cos(2.0 * math.pi * u2)
    z[1::2] = rho * np.sin(2.0 * math.pi * u2)
    return z[: int(count)]


def uniform(key: np.ndarray) -> float:
    """One uniform draw in [0, 1)."""
    w = philox(np.array([[0, 0, 0, TAG_UNIF]], dtype=np.uint64), key)[0]
    return float(_uniform53(w[0:1], w[1:2])[0])


def randint(key: np.ndarray, n: int) -> int:
    """One uniform integer in [0, n)."""
    w = philox(np.array([[0, 0, 0, TAG_UNIF]], dtype=np.uint64), key)[0]
    return int(min(math.floor(float(_uniform53(w[2:3], w[3:4])[0]) * n), n - 1))
