"""L1 peptidoform-table builder and loader (the input of the search path).

Restates vulcan/index/build.py:78-248 and :424-450 of the reference:
  1. digestion (Sage enzyme semantics: cleave after `cleave_at` unless followed by `restrict`, up to
     `missed_cleavages` internal sites, length filter), unique I/L-collapsed peptides (lib.rs:89-150);
  2. decoys (vulcan.index.decoys, paired with their forward peptide);
  3. variable-mod expansion, ids and Sage f32 masses (libvulcan_l1.so, lib.rs:283-399);
  4. content sort by (mass32, pf_id) (build.py:168) and the symmetric `partner` row (build.py:171-172).
The index side is label-free; labels are kept apart (labels.arrow) and only used for reporting.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np
import pyarrow as pa

from . import _core
from . import decoys as D

INDEX_VERSION = "vulcan-mi355x-l1-0.1"
DECOY_TAG = "rev_"   # FASTA entries whose accession contains it are skipped (decoys are made per peptide)


def _standard_mask() -> np.ndarray:
    """Residue letters with a Sage monoisotopic mass (lib.rs:63-65 is_standard): the 20 amino acids plus U, O."""
    from ..gpu.host import SAGE_MONO
    m = np.zeros(256, bool)
    for i in range(26):
        m[65 + i] = SAGE_MONO[i] > 0
    return m


_STANDARD = _standard_mask()


def read_fasta(paths) -> list:
    """(accession, sequence) of every FASTA entry of `paths` in file order: a '>' line starts an entry (the
    accession is its first whitespace-separated token), the following lines are its sequence."""
    out = []
    for path in ([paths] if isinstance(paths, str) else paths):
        acc, seq = None, []
        with open(path) as fh:
            for line in fh:
                line = line.strip()
                if line.startswith(">"):
                    if acc is not None:
                        out.append((acc, "".join(seq)))
                    head = line[1:].split()
                    acc, seq = (head[0] if head else ""), []
                elif line and acc is not None:
                    seq.append(line)
        if acc is not None:
            out.append((acc, "".join(seq)))
    return out


def content_hash(fasta_paths, params) -> str:
    """Hash of the FASTA bytes, the build parameters and the index version (reference build.py:66-75)."""
    h = hashlib.sha256()
    for path in fasta_paths:
        with open(path, "rb") as fh:
            for b in iter(lambda: fh.read(1 << 22), b""):
                h.update(b)
    h.update(json.dumps(params, sort_keys=True).encode())
    h.update(INDEX_VERSION.encode())
    return h.hexdigest()[:16]


def default_params() -> dict:
    """Reference defaults (build.py:52-63): trypsin KR/P, 2 missed, len 7-50, 500-5000 Da, CAM static,
    Met-ox variable (<= 2)."""
    return {
        "enzyme": {"cleave_at": "KR", "restrict": "P", "c_terminal": True, "missed_cleavages": 2,
                   "min_len": 7, "max_len": 50},
        "static_mods": {"C": 57.021464},
        "variable_mods": {"M": 15.9949},
        "max_variable_mods": 2,
        "variable_mod_caps": None,
        "peptide_min_mass": 500.0, "peptide_max_mass": 5000.0,
        "decoy_seed": D.DECOY_SEED, "id_seed": 0,
        # multi-replicate tables (build_multi_decoy_tables)
        "decoy_seeds": [D.DECOY_SEED], "entrap_seeds": [],
    }


def digest(proteins, enzyme: dict):
    """Unique I/L-collapsed tryptic peptides of `proteins` (sorted) and their minimal missed-cleavage count;
    peptides with a residue outside the Sage mass table (X, B, Z, J, ...) are dropped (lib.rs:117-121)."""
    cut = np.zeros(256, bool)
    for c in enzyme["cleave_at"]:
        cut[ord(c)] = True
    restrict = ord(enzyme["restrict"]) if enzyme.get("restrict") else -1
    mc, lo, hi = enzyme["missed_cleavages"], enzyme["min_len"], enzyme["max_len"]
    best: dict[bytes, int] = {}
    for prot in proteins:
        a = np.frombuffer(prot.encode() if isinstance(prot, str) else prot, np.uint8)
        n = len(a)
        if n == 0:
            continue
        site = cut[a[:-1]] & (a[1:] != restrict)
        b = np.concatenate([[0], np.nonzero(site)[0] + 1, [n]])
        raw = a.tobytes().replace(b"I", b"L")
        nb = len(b) - 1
        for m in range(mc + 1):
            if nb - m <= 0:
                break
            s0 = b[:nb - m]
            s1 = b[m + 1:]
            ln = s1 - s0
            ok = np.nonzero((ln >= lo) & (ln <= hi))[0]
            for i in ok:
                if not _STANDARD[a[s0[i]:s1[i]]].all():
                    continue
                k = raw[s0[i]:s1[i]]
                v = best.get(k)
                if v is None or m < v:
                    best[k] = m
    keys = sorted(best)
    return [k.decode() for k in keys], np.array([best[k] for k in keys], np.uint8)


class L1Index:
    """Index side of an L1 table (label-free), reference build.py:424-450."""

    def __init__(self, path: str | None = None, arrays: dict | None = None):
        self.path = path
        if path is not None:
            with pa.OSFile(os.path.join(path, "peptidoforms.arrow"), "rb") as src:
                t = pa.ipc.open_file(src).read_all()
            arrays = {k: t.column(k).to_numpy() for k in t.column_names}
            arrays["residues"] = np.load(os.path.join(path, "residues.npy"))
            with open(os.path.join(path, "meta.json")) as fh:
                self.meta = json.load(fh)
        else:
            self.meta = arrays.pop("meta", {})
        self.pf_id = np.ascontiguousarray(arrays["pf_id"], np.uint64)
        self.mass32 = np.ascontiguousarray(arrays["mass32"], np.float32)
        self.seq_off = np.ascontiguousarray(arrays["seq_off"], np.int64)
        self.length = np.ascontiguousarray(arrays["length"], np.uint8)
        self.mask = np.ascontiguousarray(arrays["mask"], np.uint64)
        self.n_var = np.ascontiguousarray(arrays["n_var"], np.uint8)
        self.missed = np.ascontiguousarray(arrays["missed"], np.uint8)
        self.partner = np.ascontiguousarray(arrays["partner"], np.int64)
        # multi-decoy tables: the pair partner in decoy set j >= 1 as partner_<j> (reference build.py:366-372)
        self.partners = [self.partner] + [np.ascontiguousarray(arrays[f"partner_{j}"], np.int64)
                                          for j in range(1, 64) if f"partner_{j}" in arrays]
        self.residues = np.ascontiguousarray(arrays["residues"], np.uint8)

    def __len__(self):
        return len(self.pf_id)

    def sequence(self, i: int) -> str:
        o = int(self.seq_off[i])
        return self.residues[o:o + int(self.length[i])].tobytes().decode()

    def save(self, path: str, labels: dict | None = None):
        os.makedirs(path, exist_ok=True)
        tbl = pa.table({"pf_id": pa.array(self.pf_id, pa.uint64()), "mass32": pa.array(self.mass32, pa.float32()),
                        "seq_off": pa.array(self.seq_off, pa.int64()), "length": pa.array(self.length, pa.uint8()),
                        "mask": pa.array(self.mask, pa.uint64()), "n_var": pa.array(self.n_var, pa.uint8()),
                        "missed": pa.array(self.missed, pa.uint8()), "partner": pa.array(self.partner, pa.int64())})
        for j, pj in enumerate(self.partners[1:], start=1):
            tbl = tbl.append_column(f"partner_{j}", pa.array(pj, pa.int64()))
        _write(tbl, os.path.join(path, "peptidoforms.arrow"))
        np.save(os.path.join(path, "residues.npy"), self.residues)
        if labels is not None:
            _write(pa.table({k: pa.array(v) for k, v in labels.items()}), os.path.join(path, "labels.arrow"))
        with open(os.path.join(path, "meta.json"), "w") as fh:
            json.dump(self.meta, fh, indent=1, default=str)


def _write(table: pa.Table, path: str) -> None:
    with pa.OSFile(path + ".partial", "wb") as sink, pa.ipc.new_file(sink, table.schema) as w:
        w.write_table(table)
    os.replace(path + ".partial", path)


def read_labels(path: str) -> dict:
    with pa.OSFile(os.path.join(path, "labels.arrow"), "rb") as src:
        t = pa.ipc.open_file(src).read_all()
    return {k: t.column(k).to_numpy() for k in t.column_names}


def build_tables(proteins, params: dict | None = None, shuffle_rows_seed: int | None = None,
                 swap_labels: bool = False, progress=None):
    """Build the L1 table in memory. Returns (L1Index, labels dict). shuffle_rows_seed / swap_labels exist only
    for the label-blindness test (reference build.py:80-83): the index side must come out bit-identical.
    progress: optional callable(str), told when each stage ends (digest, decoys, expansion, sort, assembly)."""
    p = default_params()
    p.update(params or {})
    t0 = time.perf_counter()
    say = progress or (lambda msg: None)
    e = p["enzyme"]
    targets, t_missed = digest(proteins, e)
    say(f"digest: {len(targets)} target peptides ({time.perf_counter() - t0:.1f} s)")
    t_missed_of = dict(zip(targets, t_missed.tolist()))
    dec, n_shuf = D.make_decoys(targets, p["decoy_seed"])
    say(f"decoys: {len(dec)} ({time.perf_counter() - t0:.1f} s)")
    fw = [t for t in targets if t in dec]
    dc = [dec[t][0] for t in fw]
    perms = [dec[t][1] for t in fw]
    order = np.arange(len(fw))
    if shuffle_rows_seed is not None:
        order = np.random.default_rng(shuffle_rows_seed).permutation(len(fw))
    a_seq = [fw[i] for i in order]
    b_seq = [dc[i] for i in order]
    pr = []
    for i in order:
        q = perms[i]
        if q is None:
            n = len(fw[i])
            q = np.array([n - 2 - j for j in range(n - 1)] + [n - 1], np.uint8)
        pr.append(q)
    if swap_labels:
        a_seq, b_seq = b_seq, a_seq
        pr = [np.argsort(q).astype(np.uint8) for q in pr]
    r = _core.expand_pairs(a_seq, pr, p["static_mods"], p["variable_mods"], p["max_variable_mods"],
                           p["peptide_min_mass"], p["peptide_max_mass"], p["id_seed"], p.get("variable_mod_caps"))
    n_pairs = len(r["seq"])
    say(f"expansion: {2 * n_pairs} peptidoforms ({time.perf_counter() - t0:.1f} s)")
    seq_i = r["seq"].astype(np.int64)
    lens = np.array([len(s) for s in a_seq], np.int64)
    a_off = np.zeros(len(a_seq) + 1, np.int64)
    np.cumsum(lens, out=a_off[1:])
    residues = np.concatenate([np.frombuffer("".join(a_seq).encode(), np.uint8),
                               np.frombuffer("".join(b_seq).encode(), np.uint8)])
    pf_id = np.concatenate([r["id_f"], r["id_d"]])
    mass32 = np.concatenate([r["mass32_f"], r["mass32_d"]])
    seq_off = np.concatenate([a_off[seq_i], a_off[-1] + a_off[seq_i]])
    length = np.concatenate([lens[seq_i], lens[seq_i]]).astype(np.uint8)
    mask = np.concatenate([r["mask_f"], r["mask_d"]])
    n_var = np.concatenate([r["n_var"], r["n_var"]])
    fwd_of = [y if swap_labels else x for x, y in zip(a_seq, b_seq)]
    missed_seq = np.array([t_missed_of[f] for f in fwd_of], np.uint8)
    missed = np.concatenate([missed_seq[seq_i], missed_seq[seq_i]])
    if len(np.unique(pf_id)) != len(pf_id):
        raise RuntimeError("peptidoform id collision")
    say(f"columns ({time.perf_counter() - t0:.1f} s)")
    srt = np.lexsort((pf_id, mass32))
    say(f"sort ({time.perf_counter() - t0:.1f} s)")
    inv = np.empty_like(srt)
    inv[srt] = np.arange(len(srt))
    partner_unsorted = np.concatenate([np.arange(n_pairs) + n_pairs, np.arange(n_pairs)])
    partner = inv[partner_unsorted][srt]
    lens_s = length[srt].astype(np.int64)
    new_off = np.zeros(len(srt) + 1, np.int64)
    np.cumsum(lens_s, out=new_off[1:])
    gather = np.repeat(seq_off[srt] - new_off[:-1], lens_s) + np.arange(new_off[-1])
    is_dec_side = np.concatenate([np.zeros(n_pairs, bool), np.ones(n_pairs, bool)])
    is_decoy = (~is_dec_side if swap_labels else is_dec_side)[srt]
    meta = {"index_version": INDEX_VERSION, "params": p, "n_targets": len(targets), "n_forward": len(fw),
            "n_decoy_shuffled": n_shuf, "n_peptidoforms": int(len(pf_id)), "build_s": time.perf_counter() - t0}
    l1 = L1Index(arrays={"pf_id": pf_id[srt], "mass32": mass32[srt], "seq_off": new_off[:-1],
                         "length": length[srt], "mask": mask[srt], "n_var": n_var[srt], "missed": missed[srt],
                         "partner": partner, "residues": residues[gather], "meta": meta})
    labels = {"pf_id": l1.pf_id, "is_decoy": is_decoy}
    return l1, labels


FAMILY = {"T": 0, "E": 1, "decoy_T": 2, "decoy_E": 3}


def build_multi_decoy_tables(proteins, params: dict | None = None):
    """Multi-replicate table for family top-2 retention (reference build.py:257-418): rows T ∪ E_s ∪ dec_j(T) ∪
    dec_j(E_s) for every entrapment seed s of params['entrap_seeds'] and decoy seed j of params['decoy_seeds']
    (shuffled decoys only). Forward peptides lacking a decoy in some set are dropped. partners[j] pairs each
    forward row with its set-j decoy (and back); rows of other decoy sets are their own set-j partner.
    Labels: family (0 T, 1 E, 2 dec(T), 3 dec(E)), entrap_seed, decoy_set (-1 forward), decoy_seed, is_decoy,
    is_entrapment. Returns (L1Index, labels dict)."""
    p = default_params()
    p.update(params or {})
    dseeds = list(p["decoy_seeds"])
    J = len(dseeds)
    if J < 1:
        raise ValueError("decoy_seeds must name at least one decoy set")
    t0 = time.perf_counter()
    e = p["enzyme"]
    targets, t_missed = digest(proteins, e)
    t_missed_of = dict(zip(targets, t_missed.tolist()))
    taken = set(targets)
    fam_of = {t: (0, 0, t) for t in targets}             # forward peptide -> (family, entrapment seed, target)
    ent_stats = {}
    for s_ in p["entrap_seeds"]:
        m, nr, nd = D.make_entrapment(targets, s_, taken)
        ent_stats[str(s_)] = {"n": len(m), "redrawn": nr, "dropped": nd}
        for t, (ent, _) in m.items():
            fam_of[ent] = (1, s_, t)
    forward = sorted(fam_of)
    used: set = set()
    decs, dec_stats = [], {}
    for ds in dseeds:
        dj, nr, nd = D.make_shuffled_decoys(forward, ds, used)
        decs.append(dj)
        dec_stats[str(ds)] = {"n": len(dj), "redrawn": nr, "dropped": nd}
    fw = [f for f in forward if all(f in dj for dj in decs)]
    res = []
    for dj in decs:
        r = _core.expand_pairs(fw, [dj[f][1] for f in fw], p["static_mods"], p["variable_mods"],
                               p["max_variable_mods"], p["peptide_min_mass"], p["peptide_max_mass"], p["id_seed"],
                               p.get("variable_mod_caps"))
        res.append(([dj[f][0] for f in fw], r))
    r0 = res[0][1]
    for _, r in res[1:]:
        if not (np.array_equal(r["seq"], r0["seq"]) and np.array_equal(r["id_f"], r0["id_f"])):
            raise RuntimeError("the forward side differs between decoy sets")
    nf = len(r0["seq"])
    seq_i = r0["seq"].astype(np.int64)
    lens = np.array([len(x) for x in fw], np.int64)
    a_off = np.zeros(len(fw) + 1, np.int64)
    np.cumsum(lens, out=a_off[1:])
    L = int(a_off[-1])
    residues = np.concatenate([np.frombuffer("".join(fw).encode(), np.uint8)] +
                              [np.frombuffer("".join(b).encode(), np.uint8) for b, _ in res])
    pf_id = np.concatenate([r0["id_f"]] + [r["id_d"] for _, r in res])
    mass32 = np.concatenate([r0["mass32_f"]] + [r["mass32_d"] for _, r in res])
    seq_off = np.concatenate([a_off[seq_i] + j * L for j in range(J + 1)])
    length = np.tile(lens[seq_i], J + 1).astype(np.uint8)
    mask = np.concatenate([r0["mask_f"]] + [r["mask_d"] for _, r in res])
    n_var = np.tile(r0["n_var"], J + 1)
    tgt_of = [fam_of[f][2] for f in fw]
    missed = np.tile(np.array([t_missed_of[t] for t in tgt_of], np.uint8)[seq_i], J + 1)
    if len(np.unique(pf_id)) != len(pf_id):
        raise RuntimeError("peptidoform id collision")
    N = len(pf_id)
    srt = np.lexsort((pf_id, mass32))
    inv = np.empty_like(srt)
    inv[srt] = np.arange(N)
    base = np.arange(nf)
    partners = []
    for j in range(J):
        pu = np.arange(N)
        pu[:nf] = (j + 1) * nf + base
        pu[(j + 1) * nf:(j + 2) * nf] = base
        partners.append(inv[pu][srt])
    lens_s = length[srt].astype(np.int64)
    new_off = np.zeros(N + 1, np.int64)
    np.cumsum(lens_s, out=new_off[1:])
    gather = np.repeat(seq_off[srt] - new_off[:-1], lens_s) + np.arange(new_off[-1])
    side = np.repeat(np.arange(-1, J), nf).astype(np.int8)
    fam_fwd = np.array([fam_of[f][0] for f in fw], np.uint8)[seq_i]
    seed_fwd = np.array([fam_of[f][1] for f in fw], np.int32)[seq_i]
    fam = np.tile(fam_fwd, J + 1) + np.where(side >= 0, 2, 0).astype(np.uint8)
    meta = {"index_version": INDEX_VERSION + "+multidecoy", "params": p, "n_targets": len(targets),
            "n_forward_peptides": len(fw), "n_forward_without_decoy": len(forward) - len(fw),
            "decoy_sets": dec_stats, "entrapment": ent_stats, "n_decoy_sets": J, "n_peptidoforms": int(N),
            "n_peptidoforms_by_family": {k: int((fam == v).sum()) for k, v in FAMILY.items()},
            "build_s": time.perf_counter() - t0}
    arrays = {"pf_id": pf_id[srt], "mass32": mass32[srt], "seq_off": new_off[:-1], "length": length[srt],
              "mask": mask[srt], "n_var": n_var[srt], "missed": missed[srt], "partner": partners[0],
              "residues": residues[gather], "meta": meta}
    for j in range(1, J):
        arrays[f"partner_{j}"] = partners[j]
    l1 = L1Index(arrays=arrays)
    side_s = side[srt]
    labels = {"pf_id": l1.pf_id, "family": fam[srt], "entrap_seed": np.tile(seed_fwd, J + 1)[srt],
              "decoy_set": side_s, "decoy_seed": np.array([0] + dseeds, np.int32)[side_s + 1],
              "is_decoy": side_s >= 0, "is_entrapment": (fam[srt] % 2) == 1}
    return l1, labels


def build_multi_decoy(proteins, out_dir: str, params: dict | None = None) -> dict:
    """Build and write a multi-replicate table (same file layout as build(), plus partner_<j> columns and the
    family / decoy-set label columns)."""
    l1, labels = build_multi_decoy_tables(proteins, params)
    l1.path = out_dir
    l1.save(out_dir, labels)
    return l1.meta


def build(proteins, out_dir: str, params: dict | None = None, shuffle_rows_seed: int | None = None,
          swap_labels: bool = False) -> dict:
    """Build and write peptidoforms.arrow, residues.npy, labels.arrow, meta.json (reference file layout)."""
    l1, labels = build_tables(proteins, params, shuffle_rows_seed, swap_labels)
    l1.save(out_dir, labels)
    return l1.meta


def build_fasta(fasta_paths, out_dir: str, params: dict | None = None) -> dict:
    """The index of FASTA files (reference build.py:78-93): entries whose accession contains DECOY_TAG are
    skipped, the rest digested and expanded as build(); meta.json records the FASTA paths and content hash."""
    fasta_paths = [fasta_paths] if isinstance(fasta_paths, str) else list(fasta_paths)
    p = default_params()
    p.update(params or {})
    entries = read_fasta(fasta_paths)
    proteins = [seq for acc, seq in entries if DECOY_TAG not in acc]
    l1, labels = build_tables(proteins, p)
    l1.meta.update(fasta=[os.path.abspath(f) for f in fasta_paths], content_hash=content_hash(fasta_paths, p),
                   n_proteins=len(proteins), n_skipped_decoy_entries=len(entries) - len(proteins))
    l1.save(out_dir, labels)
    return l1.meta


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Build an L1 peptidoform index from FASTA files.")
    ap.add_argument("--fasta", nargs="+", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--params", default="{}", help="JSON overrides of the build parameters")
    a = ap.parse_args(argv)
    meta = build_fasta(a.fasta, a.out, json.loads(a.params))
    print(json.dumps({k: meta[k] for k in ("n_proteins", "n_targets", "n_peptidoforms", "content_hash", "build_s")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
