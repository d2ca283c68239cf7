"""CPU tests of Z's tensor-product form (fsxh_z_tensor, host code of
libfsx.so; no device call): on box grids with separable coordinates the
oracle's assembled Z (pinned in test_oracle_pins.py) equals
Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz from 1D tables, row for row, the
pinned node's penalty aside; a planted change to one entry, a non-separable
mesh or another density breaks it."""
import numpy as np
import pytest

import meshgen


def oracle_z(oracle_lib, mesh, ngp=27, rho=1.0):
    o = oracle_lib.case_from_mesh(mesh, NGP=ngp, density=rho)
    s = o.sizes()
    rp, ci, v = o.i("Zrow"), o.i("Zcol"), o.d("Z")
    xc = o.d("coord").reshape(-1, 3)[:s["NNp"]]   # corner (pressure) nodes first
    zp = s["zeroP"]
    o.close()
    return rp, ci, v, xc, zp


def kron_z(coef, shape):
    """Z from the tables by explicit Kronecker products (x fastest)."""
    import scipy.sparse as sp
    nx, ny, nz = shape
    mats = []
    at = 0
    for n in shape:
        K = sp.lil_matrix((n, n)); M = sp.lil_matrix((n, n))
        for i in range(n):
            for d in range(5):
                j = i + d - 2
                if 0 <= j < n:
                    K[i, j] = coef[at + i, d]
                    M[i, j] = coef[at + i, 5 + d]
        mats.append((K.tocsr(), M.tocsr()))
        at += n
    (Kx, Mx), (Ky, My), (Kz, Mz) = mats
    return (sp.kron(sp.kron(Mz, My), Kx) + sp.kron(sp.kron(Mz, Ky), Mx) + sp.kron(sp.kron(Kz, My), Mx)).tocsr()


@pytest.mark.parametrize("case", [("uniform", 6, 0.0, 27, 1.0), ("sinh", 6, 2.0, 27, 1.0),
                                  ("sinh_ngp8_rho", 6, 1.5, 8, 1.7)])
def test_z_is_a_sum_of_kronecker_products(fsx_lib, oracle_lib, case):
    _, n, cl, ngp, rho = case
    rp, ci, v, xc, zp = oracle_z(oracle_lib, meshgen.cavity(n, cluster=cl), ngp, rho)
    shape = (n + 1,) * 3
    t = fsx_lib.z_tensor(shape, xc, ngp, rho, rp, ci, v, zp)
    assert t["holds"] and t["err"] < 1e-12, t["err"]
    # independent of the library's own row check: the explicit Kronecker
    # sum against Z, entry by entry, the pin's diagonal aside
    import scipy.sparse as sp
    A = sp.csr_matrix((v, ci, rp), shape=(rp.size - 1,) * 2)
    D = (A - kron_z(t["coef"], shape)).tolil()
    pen = D[zp, zp]
    D[zp, zp] = 0.0
    scale = abs(A).max(axis=1).toarray().ravel()
    Dc = D.tocsr().tocoo()
    assert np.all(np.abs(Dc.data) <= 1e-11 * scale[Dc.row])
    assert pen == pytest.approx(t["pen"], rel=1e-12)
    # the reference pins the node by scaling its diagonal (a penalty)
    assert pen > 0.0
    # closed forms of the 1D tables: B 1 = m (the linear basis sums to one),
    # so M's rows sum to the pressure basis integrals (h_left + h_right) / 2;
    # D 1 vanishes at velocity nodes inside the axis, so K's rows sum to 0
    # two or more points from its ends
    at = 0
    for d, n1 in enumerate(shape):
        ax = np.unique(np.round(xc[:, d], 12))
        h = np.diff(ax)
        lumped = np.concatenate([[h[0] / 2], (h[:-1] + h[1:]) / 2, [h[-1] / 2]])
        K, M = t["coef"][at:at + n1, :5], t["coef"][at:at + n1, 5:]
        assert np.allclose(M.sum(axis=1), lumped, rtol=1e-12)
        assert np.all(np.abs(K[2:-2].sum(axis=1)) <= 1e-12 * np.abs(K).max())
        at += n1


def test_z_tensor_rejects(fsx_lib, oracle_lib):
    rp, ci, v, xc, zp = oracle_z(oracle_lib, meshgen.cavity(4))
    shape = (5, 5, 5)
    # a planted change to one off-diagonal entry (in the oracle's output)
    v2 = v.copy()
    v2[rp[40] + 1] *= 1.0 + 1e-6
    t = fsx_lib.z_tensor(shape, xc, 27, 1.0, rp, ci, v2, zp)
    assert not t["holds"] and 1e-10 < t["err"] < 1e-6
    # another density than the one Z was assembled with
    assert not fsx_lib.z_tensor(shape, xc, 27, 2.0, rp, ci, v, zp)["holds"]
    # the 8-point rule integrates the 1D products (degree <= 3) exactly, as
    # the 27-point rule Z was assembled with does: the same tables; the
    # 1-point rule gives the quadratic basis' vertex nodes no lumped mass
    assert fsx_lib.z_tensor(shape, xc, 8, 1.0, rp, ci, v, zp)["holds"]
    assert fsx_lib.z_tensor(shape, xc, 1, 1.0, rp, ci, v, zp)["coef"] is None
    # coordinates that do not separate: one interior node moved in x
    xc2 = xc.copy()
    xc2[1 + 5 * (1 + 5 * 1), 0] += 1e-3
    t = fsx_lib.z_tensor(shape, xc2, 27, 1.0, rp, ci, v, zp)
    assert not t["holds"] and t["coef"] is None
    with pytest.raises(fsx_lib.FsxError):
        fsx_lib.z_tensor((5, 5, 4), xc, 27, 1.0, rp, ci, v, zp)
