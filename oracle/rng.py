"""Counter-based Langevin noise: Philox4x32-10 and Box-Muller, in NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The reference draws
its Langevin noise with ``torch.randn`` (integrators/nvt.py:74-83), a stream
that a graph-captured step cannot replay, so the fused engine draws it from a
counter-based generator instead: three standard normals per (seed, step,
atom), from two Philox4x32-10 blocks (Salmon, Moraes, Dror and Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11) and Box-Muller in float64.  This
module restates that definition so that tests can check the engine's noise
against it; the generator is pinned by the published Random123 known-answer
vectors (tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_LO = np.uint64(0xFFFFFFFF)
_32 = np.uint64(32)


def philox4x32_10(ctr, key):
    """Ten Philox rounds on counter words ``ctr`` (four arrays or ints) with
    key ``key`` (two ints): each round multiplies words 0 and 2 by M0 / M1,
    mixes the high halves with words 1 / 3 and the key, and bumps the key by
    the Weyl constants.  Returns four uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & _LO for c in ctr)
    k0, k1 = np.uint64(key[0]) & _LO, np.uint64(key[1]) & _LO
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        c0, c1, c2, c3 = (((p1 >> _32) ^ c1 ^ k0) & _LO, p1 & _LO,
                          ((p0 >> _32) ^ c3 ^ k1) & _LO, p0 & _LO)
        k0 = (k0 + _W0) & _LO
        k1 = (k1 + _W1) & _LO
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def u01_53(a, b) -> np.ndarray:
    """Uniform float64 in (0, 1] from two 32-bit words: 53 bits of
    (a << 21) ^ (b >> 11), plus one, times 2^-53."""
    v = (np.asarray(a, np.uint64) << np.uint64(21)) ^ (np.asarray(b, np.uint64) >> np.uint64(11))
    v = v & np.uint64((1 << 53) - 1)
    return (v.astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def langevin_normals(seed: int, step: int, n_atoms: int) -> np.ndarray:
    """[n_atoms, 3] standard normals of one step: block b of atom i is
    Philox(counter (i, step lo, step hi, b), key (seed lo, seed hi));
    z0, z1 = Box-Muller of block 0's two uniforms, z2 = the cosine normal of
    block 1's."""
    i = np.arange(n_atoms, dtype=np.uint64)
    s_lo, s_hi = step & 0xFFFFFFFF, (step >> 32) & 0xFFFFFFFF
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    z = np.empty((n_atoms, 3))
    r0 = philox4x32_10((i, s_lo, s_hi, 0), key)
    r1 = philox4x32_10((i, s_lo, s_hi, 1), key)
    a = np.sqrt(-2.0 * np.log(u01_53(r0[0], r0[1])))
    b = np.sqrt(-2.0 * np.log(u01_53(r1[0], r1[1])))
    t2 = np.pi * (2.0 * u01_53(r0[2], r0[3]))
    t4 = np.pi * (2.0 * u01_53(r1[2], r1[3]))
    z[:, 0] = a * np.cos(t2)
    z[:, 1] = a * np.sin(t2)
    z[:, 2] = b * np.cos(t4)
    return z
