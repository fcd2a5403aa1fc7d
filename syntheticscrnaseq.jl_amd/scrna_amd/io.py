"""On-disk count matrices for downstream comparison with real data.

``write_10x(directory, counts)`` writes the 10x Genomics v2 layout
(matrix.mtx, genes.tsv, barcodes.tsv) that the reference's real-data
comparison reads (experiments/compare_real_scrnaseq.jl:62-96); the writing is
done by libscrna's native code.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib


def write_10x(directory: str, counts, gene_names=None, barcode_prefix: str | None = None) -> int:
    """Write a cells x G integer count matrix (e.g. SimResult.counts).  Returns the nonzero count."""
    X = np.asarray(counts)
    if X.ndim != 2:
        raise ValueError("counts must be cells x G")
    if not np.issubdtype(X.dtype, np.integer):
        if not np.all(X == np.round(X)):
            raise ValueError("counts must be integers")
    Xf = np.asfortranarray(X, dtype=np.int32)
    cells, G = Xf.shape
    names = None
    keep = []
    if gene_names is not None:
        if len(gene_names) != G:
            raise ValueError(f"need {G} gene names")
        enc = [str(n).encode() for n in gene_names]
        keep.append(enc)
        names = (C.c_char_p * G)(*enc)
    nnz = C.c_int64(0)
    lib = _lib.load()
    rc = lib.scrna_write_10x(os.fsencode(directory), Xf.ctypes.data_as(_lib.c_int32_p), cells, G, cells, names,
                             barcode_prefix.encode() if barcode_prefix else None, C.byref(nnz))
    _lib.check(rc, "write_10x")
    return int(nnz.value)
