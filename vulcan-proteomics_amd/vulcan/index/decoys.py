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


def make_entrapment(targets, seed: int, taken: set, max_tries: int = 200):
    """Paired shuffled entrapment E_s of every target (reference decoys.py:72-106): the C-term-fixed shuffle
    seeded by (target, s, attempt), drawn again while it equals the target, a target, a pseudo-reverse of a
    target, a sequence in `taken` (T and earlier entrapments), a palindrome (its pseudo-reverse is itself) or
    has a pseudo-reverse that is a target or taken. `taken` gains the new entrapments. Returns
    ({target: (entrapment, perm)}, n_redrawn, n_dropped)."""
    tgt = sorted(set(targets))
    tset = set(tgt)
    rev_t = {pseudo_reverse(t) for t in tgt}

    def clash(e: str, t: str) -> bool:
        pe = pseudo_reverse(e)
        return e == t or e in taken or e in tset or e in rev_t or pe == e or pe in taken or pe in tset

    out, n_redraw, n_drop = {}, 0, 0
    for t in tgt:
        k = 0
        perm = hash_shuffle_perm(t, seed, 0)
        e = apply_perm(t, perm)
        while clash(e, t):
            k += 1
            if k > max_tries:
                e = None
                break
            perm = hash_shuffle_perm(t, seed, k)
            e = apply_perm(t, perm)
        if e is None:
            n_drop += 1
            continue
        n_redraw += k > 0
        out[t] = (e, perm)
        taken.add(e)
    return out, n_redraw, n_drop


def make_shuffled_decoys(forward, seed: int, used: set, max_tries: int = 50):
    """One shuffled decoy set (reference decoys.py:143-168): the C-term-fixed shuffle of every forward peptide
    (T and E) seeded by (peptide, seed, attempt), drawn again while it equals the peptide, any forward peptide,
    or a decoy already drawn in this or an earlier set (`used`, updated). Returns ({fwd: (decoy, perm)},
    n_redrawn, n_dropped)."""
    fwd = sorted(set(forward))
    fset = set(fwd)
    out, n_redraw, n_drop = {}, 0, 0
    for p in fwd:
        k = 0
        perm = hash_shuffle_perm(p, seed, 0)
        d = apply_perm(p, perm)
        while d == p or d in fset or d in used:
            k += 1
            if k > max_tries:
                d = None
                break
            perm = hash_shuffle_perm(p, seed, k)
            d = apply_perm(p, perm)
        if d is None:
            n_drop += 1
            continue
        n_redraw += k > 0
        out[p] = (d, perm)
        used.add(d)
    return out, n_redraw, n_drop
