"""FASTA input of the L1 index builder (reference build.py:66-93, lib.rs:91-127): parsing, decoy-tag and
non-standard-residue filtering, content hash, CLI (CPU only)."""
import json

import numpy as np

from vulcan.index import build as B

P1 = "MKWVTFISLLFLFSSAYSRGVFRRDTHKSEIAHRFKDLGEEHFKGLVLIAFSQYLQQCPFDEHVKLVNELTEFAKTCVADESHAGCEK"
P2 = "MSDKPDMAEIEKFDKSKLKKTETQEKNPLPSKETIEQEKQAGESGSAAVPTNPIEQEFKHHKAEHVLK"
PX = "MTTRPEGKXGTAVLLRGGVEKPLIDNGSDQRGTVLK"


def _write_fasta(path, entries):
    with open(path, "w") as fh:
        for acc, seq in entries:
            fh.write(f">{acc} some description\n")
            for i in range(0, len(seq), 30):   # multi-line sequences
                fh.write(seq[i:i + 30] + "\n")


def _small_params():
    return {"enzyme": {"cleave_at": "KR", "restrict": "P", "c_terminal": True, "missed_cleavages": 1,
                       "min_len": 5, "max_len": 30}, "peptide_min_mass": 300.0}


def test_read_fasta_and_digest_filters(tmp_path):
    f = tmp_path / "a.fasta"
    _write_fasta(f, [("sp|P1|A", P1), ("rev_sp|P1|A", P1[::-1]), ("sp|P2|B", P2), ("tr|PX|C", PX)])
    ent = B.read_fasta([str(f)])
    assert [a for a, _ in ent] == ["sp|P1|A", "rev_sp|P1|A", "sp|P2|B", "tr|PX|C"]
    assert ent[0][1] == P1 and ent[3][1] == PX
    peps, missed = B.digest([PX], _small_params()["enzyme"])
    assert all("X" not in p for p in peps)
    # the peptides around the X survive (I -> L), the ones holding it (XGTAVLLR, its missed-cleavage joins) do not
    assert {"MTTRPEGK", "GGVEKPLLDNGSDQR", "GTVLK"} <= set(peps)
    assert not any("GTAVLL" in p for p in peps)


def test_build_fasta_equals_build_tables(tmp_path):
    f1, f2 = tmp_path / "a.fasta", tmp_path / "b.fasta"
    _write_fasta(f1, [("sp|P1|A", P1), ("rev_sp|P1|A", P1[::-1])])
    _write_fasta(f2, [("sp|P2|B", P2), ("tr|PX|C", PX)])
    out = tmp_path / "idx"
    meta = B.build_fasta([str(f1), str(f2)], str(out), _small_params())
    assert meta["n_proteins"] == 3 and meta["n_skipped_decoy_entries"] == 1
    l1 = B.L1Index(str(out))
    ref, _ = B.build_tables([P1, P2, PX], {**B.default_params(), **_small_params()})
    assert len(l1) == len(ref) > 0
    for k in ("pf_id", "mass32", "length", "mask", "partner", "residues"):
        assert np.array_equal(getattr(l1, k), getattr(ref, k)), k
    assert b"X" not in l1.residues.tobytes()
    assert l1.meta["content_hash"] == meta["content_hash"]


def test_content_hash_tracks_bytes_and_params(tmp_path):
    f = tmp_path / "a.fasta"
    _write_fasta(f, [("sp|P1|A", P1)])
    p = {**B.default_params(), **_small_params()}
    h0 = B.content_hash([str(f)], p)
    assert B.content_hash([str(f)], p) == h0 and len(h0) == 16
    assert B.content_hash([str(f)], {**p, "max_variable_mods": 1}) != h0
    _write_fasta(f, [("sp|P1|A", P1 + "K")])
    assert B.content_hash([str(f)], p) != h0


def test_cli(tmp_path, capsys):
    f = tmp_path / "a.fasta"
    _write_fasta(f, [("sp|P2|B", P2)])
    rc = B.main(["--fasta", str(f), "--out", str(tmp_path / "idx"), "--params", json.dumps(_small_params())])
    assert rc == 0
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert out["n_proteins"] == 1 and out["n_peptidoforms"] == len(B.L1Index(str(tmp_path / "idx")))
