"""Random streams used by the hot path (oracle side).

Two generators:

1. The reference's torch key chain.  ``PRNGKey``/``next_key`` hash the integer key
   with SHA-256 and reduce it mod 2**32-1 (cola/backends/torch_fns.py:154-155,
   222-230); ``randn`` seeds torch's global generator with that key and draws
   ``torch.randn`` on the target device (torch_fns.py:233-241).  On the CPU this is
   what the reference's CPU path draws.

2. The engine's counter-based probe stream, shared bit-for-bit by the HIP kernels
   (cola_amd/csrc/rng.hip) and this oracle: SplitMix64's output function
   (Steele, Lea & Flood, "Fast splittable pseudorandom number generators",
   OOPSLA 2014) applied twice to (seed, stream, row).  Rademacher probes take one
   64-bit word per (row, group of 64 probes) and read probe p from bit p % 64.
"""
import hashlib

import numpy as np
import torch

GOLDEN = 0x9E3779B97F4A7C15
MASK64 = (1 << 64) - 1


def sha_hash(n: int) -> int:
    """cola/backends/torch_fns.py:226-230."""
    n_bytes = n.to_bytes((n.bit_length() + 7) // 8, 'big')
    digest = hashlib.sha256(n_bytes).digest()
    return int(int.from_bytes(digest, 'big') % (2**32 - 1))


def PRNGKey(x: int) -> int:
    """cola/backends/torch_fns.py:154-155."""
    return sha_hash(x)


def next_key(key: int) -> int:
    """cola/backends/torch_fns.py:222-223."""
    return sha_hash(key)


def torch_randn(*shape, dtype=torch.float32, key=None):
    """cola/backends/torch_fns.py:233-241 on the CPU generator (state restored)."""
    if key is None:
        key = PRNGKey(0)
    old = torch.random.get_rng_state()
    torch.random.manual_seed(key)
    z = torch.randn(*shape, dtype=dtype)
    torch.random.set_rng_state(old)
    return z


# ---------------------------------------------------------------- counter stream
def mix64(z):
    """SplitMix64 output function on uint64 numpy arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over='ignore'):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def stream_word(seed: int, stream, row):
    """64 random bits for (seed, stream, row): mix64(mix64(seed + G*(stream+1)) + G*(row+1))."""
    stream = np.asarray(stream, dtype=np.uint64)
    row = np.asarray(row, dtype=np.uint64)
    with np.errstate(over='ignore'):
        h = mix64(np.uint64(seed & MASK64) + np.uint64(GOLDEN) * (stream + np.uint64(1)))
        return mix64(h + np.uint64(GOLDEN) * (row + np.uint64(1)))


def rademacher_bits(seed: int, n: int, num_probes: int):
    """uint64 words (ceil(B/64), n): bit p%64 of word [p//64, i] is probe p at row i (1 -> -1)."""
    groups = (num_probes + 63) // 64
    g = np.arange(groups, dtype=np.uint64)[:, None]
    i = np.arange(n, dtype=np.uint64)[None, :]
    return stream_word(seed, g, i)


def rademacher(seed: int, n: int, num_probes: int, dtype=torch.float32):
    """(B, n) matrix of +-1 probes, column-block layout (probe-major)."""
    words = rademacher_bits(seed, n, num_probes)
    p = np.arange(num_probes)
    bits = (words[p // 64, :] >> (p % 64).astype(np.uint64)[:, None]) & np.uint64(1)
    return torch.from_numpy(1.0 - 2.0 * bits.astype(np.float64)).to(dtype)


GAUSS_SALT = 0x5851F42D4C957F2D


def gaussian(seed: int, n: int, num_probes: int, dtype=torch.float32):
    """(B, n) standard normals by Box-Muller on two 24-bit uniforms of one word."""
    p = np.arange(num_probes, dtype=np.uint64)[:, None]
    i = np.arange(n, dtype=np.uint64)[None, :]
    w = stream_word(seed ^ GAUSS_SALT, p, i)
    u1 = ((w >> np.uint64(40)).astype(np.float64) + 0.5) / 16777216.0
    u2 = ((w >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.float64) / 16777216.0
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return torch.from_numpy(z).to(dtype)
