"""CPU tests of the gradient operator's row classes (fsxh_g_classes, host
code of libfsx.so; no device call), the form in which step3 and step2's
right-hand side read G and G^T at matrix precision 32. G comes from the
oracle's assembly (pinned in test_oracle_pins.py).

The classes are checked against an independent numpy partition of the rows
by their (column offsets, FP32 values) lists, the values the form represents
against G's own FP32 rounding, and the products G x and G^T y of the form
against the FP64 CSR products."""
import numpy as np
import pytest

import meshgen

PRUNE = 2.0 ** -24


def oracle_g(oracle_lib, mesh):
    o = oracle_lib.case_from_mesh(mesh, NGP=27)
    rp, ci = o.i("Grow"), o.i("Gcol")
    g3 = np.stack([o.d("G1"), o.d("G2"), o.d("G3")], 1)
    nnp = o.sizes()["NNp"]
    o.close()
    return rp, ci, g3, nnp


def flushed_f32(rp, g3):
    """G's FP32 rounding with components below 2^-24 of their row's largest
    magnitude set to 0, as the class form keeps them."""
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    m = np.zeros(rp.size - 1)
    np.maximum.at(m, rows, np.abs(g3).max(axis=1))
    v = g3.copy()
    v[np.abs(v) < PRUNE * m[rows, None]] = 0.0
    return v.astype(np.float32)


def partition(keys):
    """First-occurrence ids of a list of hashable keys."""
    ids, out = {}, np.empty(len(keys), np.int64)
    for i, k in enumerate(keys):
        out[i] = ids.setdefault(k, len(ids))
    return out


@pytest.mark.parametrize("n", [6, 12])
def test_g_classes_on_the_uniform_cavity(fsx_lib, oracle_lib, n):
    """On the uniform cavity the rows of G fall into 64 classes and those of
    G^T into 125 (node kinds and positions against the walls), whatever the
    grid size; the classes are exactly the partition of the rows by their
    lists, and the form's values are G's FP32 rounding (flushed residues
    aside)."""
    rp, ci, g3, nnp = oracle_g(oracle_lib, meshgen.cavity(n))
    f = fsx_lib.g_classes(rp, ci, g3, nnp)
    assert f is not None
    ncls, nent, ntcls, ntent, toff = f["counts"]
    assert (ncls, ntcls) == (64, 125), f["counts"]
    assert toff == 0   # the reference's velocity numbering: G^T's rows keep their columns
    v32 = flushed_f32(rp, g3)
    assert np.array_equal(f["g3_f32"].view(np.uint32), v32.view(np.uint32))   # no NaN: G^T agrees
    keys = [(tuple(ci[rp[r]:rp[r + 1]] - ci[rp[r]]), v32[rp[r]:rp[r + 1]].tobytes()) for r in range(rp.size - 1)]
    assert np.array_equal(f["row_cls"], partition(keys))
    # class entries: the rows' lists, once per class
    assert nent == sum(rp[r + 1] - rp[r] for r in np.unique(f["row_cls"], return_index=True)[1])
    # products against the FP64 CSR operator
    rng = np.random.default_rng(n)
    x = rng.standard_normal(nnp)
    y = rng.standard_normal(rp.size - 1)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    for d in range(3):
        gx = np.bincount(rows, g3[:, d] * x[ci], minlength=rp.size - 1)
        cx = np.bincount(rows, f["g3_f32"][:, d].astype(np.float64) * x[ci], minlength=rp.size - 1)
        assert np.abs(cx - gx).max() <= 2e-7 * np.abs(g3[:, d]).max() * np.abs(x).max() * 27
        gty = np.bincount(ci, g3[:, d] * y[rows], minlength=nnp)
        cty = np.bincount(ci, f["g3_f32"][:, d].astype(np.float64) * y[rows], minlength=nnp)
        assert np.abs(cty - gty).max() <= 2e-7 * np.abs(g3[:, d]).max() * np.abs(y).max() * 125


def test_g_transpose_classes_carry_offsets_in_grid_order(fsx_lib, oracle_lib):
    """With the velocity nodes numbered in grid order (z, y, x), as the
    device numbers them, G^T's rows repeat their column offsets too: keyed
    by (offset, values) in column order they form 27 classes (below, inside
    and above in each axis), which carry the offsets, and the form still
    represents G exactly."""
    o = oracle_lib.case_from_mesh(meshgen.cavity(6), NGP=27)
    rp, ci = o.i("Grow"), o.i("Gcol")
    g3 = np.stack([o.d("G1"), o.d("G2"), o.d("G3")], 1)
    xyz = o.d("coord").reshape(-1, 3)
    nnp = o.sizes()["NNp"]
    o.close()
    q = np.round(xyz * 1e9).astype(np.int64)   # mid-node coordinates carry rounding
    order = np.lexsort((q[:, 0], q[:, 1], q[:, 2]))   # new row -> old row
    lens = np.diff(rp)[order]
    rp2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in order])
    f = fsx_lib.g_classes(rp2, ci[idx], g3[idx], nnp)
    assert f["counts"][2] == 27 and f["counts"][4] == 1, f["counts"]
    assert np.array_equal(f["g3_f32"].view(np.uint32), flushed_f32(rp2, g3[idx]).view(np.uint32))


def test_g_classes_grow_on_a_stretched_mesh(fsx_lib, oracle_lib):
    """The sinh-clustered cavity's elements differ in size along each axis
    (symmetric about the centre): G's classes grow with the element count
    (1000 on the 8^3 grid's 4913 rows), so beyond the device's 65535 at
    64^3 the kernels read the CSR copies there. The form stays exact."""
    rp, ci, g3, nnp = oracle_g(oracle_lib, meshgen.cavity(8, cluster=2.0))
    f = fsx_lib.g_classes(rp, ci, g3, nnp)
    assert f is not None
    assert f["counts"][0] >= 8 ** 3, f["counts"]
    assert np.array_equal(f["g3_f32"].view(np.uint32), flushed_f32(rp, g3).view(np.uint32))


def test_g_classes_flush_only_residues(fsx_lib):
    """A planted row pair: equal but for a component below 2^-24 of the row's
    largest magnitude share a class (the component reads 0); a difference
    above it keeps them apart."""
    rp = np.array([0, 2, 4, 6], np.int32)
    ci = np.array([0, 1, 1, 2, 2, 3], np.int32)   # rows 0, 1, 2: same offsets from their first column
    g3 = np.array([[1.0, 0.5, 0.0], [2.0, -1.0, 1e-9],
                   [1.0, 0.5, 0.0], [2.0, -1.0, -3e-10],
                   [1.0, 0.5, 0.0], [2.0, -1.0, 1e-6]])
    f = fsx_lib.g_classes(rp, ci, g3, 4)
    assert f["row_cls"].tolist() == [0, 0, 1]
    assert f["g3_f32"][1, 2] == 0.0 and f["g3_f32"][3, 2] == 0.0 and f["g3_f32"][5, 2] == np.float32(1e-6)
    assert f["counts"][:2] == (2, 4)


def test_g_classes_reject_bad_input(fsx_lib):
    g3 = np.zeros((3, 3))
    with pytest.raises(fsx_lib.FsxError):   # column out of range
        fsx_lib.g_classes(np.array([0, 2, 3], np.int32), np.array([0, 5, 1], np.int32), g3, 2)
    with pytest.raises(fsx_lib.FsxError):   # columns not ascending
        fsx_lib.g_classes(np.array([0, 2, 3], np.int32), np.array([1, 0, 1], np.int32), g3, 2)
