"""CPU tests of the direct solve's nested-dissection order (fsxh_nd_order,
host code of libfsx.so; no device call): a permutation in postorder of a
separator tree whose dense fronts hold every entry of the true Cholesky
factor, checked against the oracle's symbolic Cholesky (itself pinned
against numeric factors in test_oracle_pins.py)."""
import numpy as np
import pytest


def z_pattern(fsx, mesh):
    t = fsx.Topology(mesh)
    rp, ci = t.get("Zrow"), t.get("Zcol")
    n = rp.size - 1
    xyz = np.asarray(t.get("coord")).reshape(-1, 3)[:n]   # corner (pressure) nodes come first
    t.close()
    return rp, ci, xyz


@pytest.mark.parametrize("case", ["cavity16", "cavity12_sinh", "duct"])
def test_nd_fronts_hold_the_factor(fsx_lib, oracle_lib, case):
    """The order is a permutation; the dense fronts' entry count and flops
    bound the true factor's (a separator that left an entry between its two
    parts would put fill outside the fronts) and exceed them by little; on
    the 16^3 cavity and the permuted duct the order fills less than the grid
    (band) order and than the permuted numbering."""
    if case == "cavity16":
        mesh = fsx_lib.Mesh.cavity(16)
    elif case == "cavity12_sinh":
        mesh = fsx_lib.Mesh.cavity(12, cluster=2.0)
    else:
        mesh = fsx_lib.Mesh.duct(n=4, n_in=6, n_arc=8, n_out=10, seed=5)
    rp, ci, xyz = z_pattern(fsx_lib, mesh)
    n = rp.size - 1
    perm, st = fsx_lib.nd_order(rp, ci, xyz)
    assert np.array_equal(np.sort(perm), np.arange(n))
    nnz, flops, h = oracle_lib.symbolic_cholesky(rp, ci, perm)
    assert nnz <= st["nnz_l"] <= 1.15 * nnz, (nnz, st)
    assert flops <= st["flops"] <= 1.3 * flops, (flops, st)
    assert st["fronts"] > 1 and st["depth"] >= 1 and h < n
    ident = oracle_lib.symbolic_cholesky(rp, ci, np.arange(n, dtype=np.int32))
    if case != "cavity12_sinh":
        assert st["nnz_l"] < ident[0] and st["flops"] < ident[1], (st, ident)


def test_nd_order_at_64_cubed(fsx_lib):
    """The 64^3 cavity's pressure grid (65^3 nodes): L's dense fronts hold
    about 6.6e8 entries against the band's 2.30e9 (grid order,
    2 nx^2 + 2 nx + 2 = 8582 half-bandwidth), factor flops about 3.7e12."""
    rp, ci, xyz = z_pattern(fsx_lib, fsx_lib.Mesh.cavity(64))
    n = rp.size - 1
    perm, st = fsx_lib.nd_order(rp, ci, xyz)
    assert np.array_equal(np.sort(perm), np.arange(n))
    w = 2 * 65 * 65 + 2 * 65 + 2
    band = n * (w + 1) - w * (w + 1) / 2
    assert st["nnz_l"] < 0.3 * band and st["flops"] < 4e12, st


def test_nd_order_rejects_bad_input(fsx_lib):
    rp = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 5, 1], np.int32)   # column 5 out of range
    with pytest.raises(fsx_lib.FsxError):
        fsx_lib.nd_order(rp, ci, np.zeros((2, 3)))
    with pytest.raises(fsx_lib.FsxError):
        fsx_lib.nd_order(np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.full((2, 3), np.nan))
