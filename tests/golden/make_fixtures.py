"""Writes the small genotype fixtures under tests/golden/ from the reference tree.

Run once where the reference tree is mounted (default /root/reference):
    python tests/golden/make_fixtures.py [/path/to/reference]

Inputs (data files the reference's own tests use):
  test.zip:test/data/uk10k_chr1_1mb.bed (+ .fam/.bim), N = 3642, M = 6717
  test.gauss/data.bed (+ .fam/.bim),                    N = 1000, M = 11054
Outputs: the 3 magic bytes plus the first COLS SNP columns of each .bed, and a
JSON with the full files' sizes, SHA-256 digests and N/M read from .fam/.bim.
Only bytes of the data files are copied; no reference source is read.
"""
from __future__ import annotations

import hashlib
import io
import json
import sys
import zipfile
from pathlib import Path

COLS = 48
HERE = Path(__file__).resolve().parent


def _slice(bed: bytes, n: int, cols: int) -> bytes:
    stride = (n + 3) // 4
    return bed[: 3 + stride * cols]


def main(ref: Path) -> None:
    meta = {}
    with zipfile.ZipFile(ref / "test.zip") as z:
        bed = z.read("test/data/uk10k_chr1_1mb.bed")
        fam = z.read("test/data/uk10k_chr1_1mb.fam").decode()
        bim = z.read("test/data/uk10k_chr1_1mb.bim").decode()
    n = len([l for l in fam.splitlines() if l.strip()])
    m = len([l for l in bim.splitlines() if l.strip()])
    (HERE / "uk10k_chr1_1mb.slice.bed").write_bytes(_slice(bed, n, COLS))
    meta["uk10k_chr1_1mb"] = dict(num_inds=n, num_snps=m, bytes=len(bed),
                                  sha256=hashlib.sha256(bed).hexdigest(), slice_cols=COLS)

    bed = (ref / "test.gauss" / "data.bed").read_bytes()
    n = len([l for l in (ref / "test.gauss" / "data.fam").read_text().splitlines() if l.strip()])
    m = len([l for l in (ref / "test.gauss" / "data.bim").read_text().splitlines() if l.strip()])
    (HERE / "gauss_data.slice.bed").write_bytes(_slice(bed, n, COLS))
    meta["gauss_data"] = dict(num_inds=n, num_snps=m, bytes=len(bed),
                              sha256=hashlib.sha256(bed).hexdigest(), slice_cols=COLS)
    (HERE / "bed_fixtures.json").write_text(json.dumps(meta, indent=2) + "\n")
    print(json.dumps(meta, indent=2))


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
