"""Integer keys of the reference's torch backend (cola/backends/torch_fns.py:154-155,
222-230): PRNGKey(x) = next_key(x) = SHA-256(x big-endian) mod (2**32 - 1).

The engine uses these integers as seeds of its counter-based probe streams."""
import hashlib


def sha_hash(n):
    n_bytes = n.to_bytes((n.bit_length() + 7) // 8, "big")
    return int(int.from_bytes(hashlib.sha256(n_bytes).digest(), "big") % (2**32 - 1))


def PRNGKey(x):
    return sha_hash(int(x))


def next_key(key):
    return sha_hash(int(key))
