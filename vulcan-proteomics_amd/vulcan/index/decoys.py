"""Peptide-level decoys for the L1 table (restates vulcan/index/decoys.py:28-132 of the reference).

* decoy = pseudo-reverse with the C-terminal residue fixed (decoys.py:28-29); if it is a palindrome or collides
  with a forward peptide or an earlier decoy, C-term-fixed shuffles seeded by blake2b(peptide|seed|attempt)
  (decoys.py:32-43, 109-132). Sorted processing order, so the result does not depend on input order.
"""
from __future__ import annotations

import hashlib

import numpy as np

DECOY_SEED = 1001


def pseudo_reverse(pep: str) -> str:
    """'PEPTIDEK' -> 'EDITPEPK' (reverse all but the C-terminal residue)."""
    return pep[:-1][::-1] + pep[-1]


def _rng(*keys) -> np.random.Generator:
    digest = hashlib.blake2b("|".join(str(k) for k in keys).encode(), digest_size=8).digest()
    return np.random.default_rng(int.from_bytes(digest, "little"))


def hash_shuffle_perm(pep: str, seed: int, attempt: int = 0) -> np.ndarray:
    """Permutation q with dec[j] = pep[q[j]]: a shuffle of positions 0..n-2 (C-term fixed)."""
    n = len(pep)
    head = np.arange(n - 1)
    _rng(pep, seed, attempt).shuffle(head)
    return np.append(head, n - 1).astype(np.uint8)


def apply_perm(pep: str, perm) -> str:
    return "".join(pep[int(i)] for i in perm)


def hash_shuffle(pep: str, seed: int, attempt: int = 0) -> str:
    return apply_perm(pep, hash_shuffle_perm(pep, seed, attempt))


def make_decoys(forward, seed: int = DECOY_SEED, max_tries: int = 50):
    """{forward: (decoy, perm or None)}, n_shuffled. perm None = pseudo-reverse. A forward peptide whose
    shuffles all collide gets no decoy (and is dropped by the builder, as in build.py:111)."""
    fwd = set(forward)
    used: set[str] = set()
    out: dict[str, tuple[str, np.ndarray | None]] = {}
    n_shuffled = 0
    for pep in sorted(fwd):
        dec, perm, k = pseudo_reverse(pep), None, 0
        while dec == pep or dec in fwd or dec in used:
            if k >= max_tries:
                dec = None
                break
            perm = hash_shuffle_perm(pep, seed, k)
            dec = apply_perm(pep, perm)
            k += 1
        if k:
            n_shuffled += 1
        if dec is not None:
            out[pep] = (dec, perm)
            used.add(dec)
    return out, n_shuffled
