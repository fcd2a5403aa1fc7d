"""10x v2 writer against scipy's MatrixMarket reader and a restatement of the
reference's reader (experiments/compare_real_scrnaseq.jl:69-96)."""
import os

import numpy as np
import pytest
import scipy.io

import scrna_amd as sc


def read_10x_like_reference(path):
    """experiments/compare_real_scrnaseq.jl:69-96: skip '%' lines, header
    'n_genes n_cells nnz', then 'gene cell count' triples (1-based)."""
    lines = open(os.path.join(path, "matrix.mtx")).read().splitlines()
    k = 0
    while lines[k].startswith("%"):
        k += 1
    n_genes, n_cells, nnz = (int(x) for x in lines[k].split())
    X = np.zeros((n_genes, n_cells), dtype=np.int64)
    for line in lines[k + 1:k + 1 + nnz]:
        i, j, v = (int(x) for x in line.split())
        X[i - 1, j - 1] += v
    names = [ln.split("\t")[1] for ln in open(os.path.join(path, "genes.tsv")).read().splitlines()]
    return X, names


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    Y = rng.poisson(0.7, size=(333, 17)).astype(np.int32)
    Y[:, 3] = 0                                       # an all-zero gene
    d = str(tmp_path / "out")
    nnz = sc.io.write_10x(d, Y)
    assert nnz == np.count_nonzero(Y)
    M = scipy.io.mmread(os.path.join(d, "matrix.mtx")).toarray()
    assert M.shape == (17, 333) and np.array_equal(M, Y.T)
    X, names = read_10x_like_reference(d)
    assert np.array_equal(X, Y.T)
    assert names == [f"gene{g + 1}" for g in range(17)]
    bc = open(os.path.join(d, "barcodes.tsv")).read().splitlines()
    assert len(bc) == 333 and len(set(bc)) == 333
    assert all(len(b) == 18 and b.endswith("-1") and set(b[:16]) <= set("ACGT") for b in bc)


def test_gene_names_and_prefix(tmp_path):
    Y = np.array([[0, 5], [2, 0], [1, 1]], dtype=np.int32)
    d = str(tmp_path / "named")
    sc.io.write_10x(d, Y, gene_names=["Sox2", "Nanog"], barcode_prefix="run7_")
    genes = open(os.path.join(d, "genes.tsv")).read().splitlines()
    assert genes == ["SYNG000001\tSox2", "SYNG000002\tNanog"]
    assert open(os.path.join(d, "barcodes.tsv")).read().splitlines()[0] == "run7_AAAAAAAAAAAAAAAA-1"
    txt = open(os.path.join(d, "matrix.mtx")).read().splitlines()
    assert txt[0] == "%%MatrixMarket matrix coordinate integer general"
    assert txt[2] == "2 3 4"
    assert txt[3:] == ["1 2 2", "1 3 1", "2 1 5", "2 3 1"]


def test_rejects_bad_counts(tmp_path):
    with pytest.raises(ValueError):
        sc.io.write_10x(str(tmp_path / "a"), np.array([[1, -1]]))
    with pytest.raises(ValueError):
        sc.io.write_10x(str(tmp_path / "b"), np.array([[1.5, 2.0]]))
