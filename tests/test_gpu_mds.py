"""GPU MDS against the FP64 classical-MDS oracle (sign-aligned relative
Frobenius error and largest principal angle; the eigengap is asserted so the
comparison is well posed)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.proximity import Accumulator, DeviceForest, mds_dense, unpack_upper

from helpers import centred, eig_residuals, small_forest

pytestmark = pytest.mark.gpu


def _counts(kind, n, p, ncl, T, seed, variant, mtry=0):
    x, y, f = small_forest(kind, n, p, ncl, T, seed, mtry=mtry)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    U = acc.accumulate(keys, variant, _lib.LAYOUT_UPPER, _lib.ALGO_DENSE)
    return f, U


@pytest.mark.parametrize("precision,tol_rel", [(_lib.PREC_FP64, 1e-6), (_lib.PREC_FP32, 2e-3)])
def test_mds_counts_over_T(gpu, precision, tol_rel):
    print(f"mds counts/T precision {precision}", flush=True)
    f, U = _counts(_lib.DATA_COVTYPE_LIKE, 1800, 54, 7, 60, 11, _lib.PROX_ALL, mtry=7)
    n, T = f.n, f.ntree
    X, eig, it = mds_dense(U, _lib.LAYOUT_UPPER, n, _lib.MDS_COUNTS, 1.0 / T, precision=precision)
    C = unpack_upper(U.cpu().numpy(), n).astype(np.float64) / T
    Xref, w = O.classical_mds(C, 3)
    assert w[2] / w[3] > 1.02, f"eigengap too small for a stable comparison: {w[:5]}"
    e = O.embedding_errors(X.cpu().numpy(), Xref)
    print(it, eig[:4], w[:4], e, flush=True)
    assert e["rel_frobenius"] < tol_rel
    assert np.allclose(eig[:3], w[:3], rtol=tol_rel)


# FP64: the VALU operator; FP32: the matrix-core operator (E formed in FP64,
# rounded to FP32, products on v_mfma_f32_16x16x4f32)
@pytest.mark.parametrize("precision,eig_rtol,res_tol,x_tol", [(_lib.PREC_FP64, 1e-9, 1e-5, 1e-5),
                                                              (_lib.PREC_FP32, 1e-5, 1e-3, 2e-3)])
@pytest.mark.parametrize("kind,n,p,ncl,T,seed", [(_lib.DATA_COVTYPE_LIKE, 1500, 54, 7, 100, 12),
                                                 (_lib.DATA_OVERLAP_MIXTURE, 1200, 8, 3, 80, 12),
                                                 (_lib.DATA_GAUSS_MIXTURE, 333, 9, 3, 40, 5)])
def test_mds_reference_chain(gpu, kind, n, p, ncl, T, seed, precision, eig_rtol, res_tol, x_tol):
    print(f"mds reference chain kind {kind} n {n} precision {precision}", flush=True)
    f, U = _counts(kind, n, p, ncl, T, seed, _lib.PROX_OOB_UPPER)
    n = f.n
    nout = f.nout(2)
    X, eig, it = mds_dense(U, _lib.LAYOUT_UPPER, n, _lib.MDS_REFERENCE, nout=torch.from_numpy(nout).to(gpu),
                           precision=precision)
    Uh = unpack_upper(U.cpu().numpy(), n).astype(np.float64)
    # finishprox from the symmetric pair sums: P_ij + P_ji = U_ij
    S = O.finishprox(n, nout, (Uh / 2.0).reshape(-1, order="F"))  # P = U/2 gives the same pair sums
    M = O.mds_input_reference(S.reshape(n, n, order="F"), nout)
    Xref, w = O.classical_mds(M, 3)
    Xg = X.cpu().numpy()
    # eigenvalues and eigen-residuals pin the result whatever the eigengap
    assert np.allclose(eig[:3], w[:3], rtol=eig_rtol)
    res = eig_residuals(centred(M), Xg, eig[:3])
    print(it, eig[:5], w[:5], res, flush=True)
    assert res.max() < res_tol
    if w[2] / w[3] > 1.02:  # a well-separated third eigenvalue: compare the embedding itself
        e = O.embedding_errors(Xg, Xref)
        print(e, flush=True)
        assert e["rel_frobenius"] < x_tol


@pytest.mark.parametrize("precision,tol", [(_lib.PREC_FP64, 1e-8), (_lib.PREC_FP32, 1e-5)])
def test_mds_full_layout_matches_upper(gpu, precision, tol):
    f, U = _counts(_lib.DATA_OVERLAP_MIXTURE, 700, 6, 4, 40, 13, _lib.PROX_ALL)
    n = f.n
    full = torch.from_numpy(unpack_upper(U.cpu().numpy(), n).reshape(-1)).to(gpu)
    Xu, eu, _ = mds_dense(U, _lib.LAYOUT_UPPER, n, _lib.MDS_COUNTS, 1.0 / f.ntree, precision=precision)
    Xf, ef, _ = mds_dense(full, _lib.LAYOUT_FULL, n, _lib.MDS_COUNTS, 1.0 / f.ntree, precision=precision)
    assert O.embedding_errors(Xf.cpu().numpy(), Xu.cpu().numpy())["rel_frobenius"] < tol


def test_mds_wide_block_at_n_100000(gpu):
    # n = 100,000 takes the 32-column block on the matrix cores.  Residuals of
    # the top three eigenpairs of the centred operator B = -1/2 J E J,
    # E = (max M - M)^2 off the diagonal, M = C / T, formed here in float64
    # from the FULL count matrix in row chunks (torch on the device)
    n, T = 100000, 60
    print(f"mds wide block: grow n={n} T={T}", flush=True)
    x, y, f = small_forest(_lib.DATA_COVTYPE_LIKE, n, 54, 7, T, 3, mtry=7)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    C = acc.accumulate(keys, _lib.PROX_ALL, _lib.LAYOUT_FULL, _lib.ALGO_GATHER)
    del acc, df, keys
    torch.cuda.empty_cache()
    print("mds wide block: FP64 and FP32 MDS", flush=True)
    X64, e64, it64 = mds_dense(C, _lib.LAYOUT_FULL, n, _lib.MDS_COUNTS, 1.0 / T, precision=_lib.PREC_FP64)
    X32, e32, it32 = mds_dense(C, _lib.LAYOUT_FULL, n, _lib.MDS_COUNTS, 1.0 / T, precision=_lib.PREC_FP32)
    print("mds wide block: residuals", flush=True)
    Cm = C.view(n, n)
    mx = float(Cm.max()) / T
    V = torch.cat([X64, X32], dim=1)  # n x 6, float64
    EV = torch.zeros_like(V)
    r = torch.zeros(n, dtype=torch.float64, device=C.device)
    for i0 in range(0, n, 2048):
        i1 = min(n, i0 + 2048)
        E = (mx - Cm[i0:i1].to(torch.float64) / T) ** 2
        E[torch.arange(i1 - i0, device=C.device), torch.arange(i0, i1, device=C.device)] = 0.0
        r[i0:i1] = E.sum(dim=1) / n
        EV[i0:i1] = E @ V
        del E
    g = float(r.mean())
    s = V.sum(dim=0)
    BV = -0.5 * (EV - r[:, None] * s[None, :] - (r @ V)[None, :] + g * s[None, :])
    lam = torch.tensor(np.concatenate([e64[:3], e32[:3]]), dtype=torch.float64, device=C.device)
    res = (BV - V * lam[None, :]).norm(dim=0) / (lam.abs() * V.norm(dim=0))
    res = res.cpu().numpy()
    print("mds wide block: iterations", it64, it32, "residuals", res, flush=True)
    assert np.all(res[:3] < 1e-6), res
    assert np.all(res[3:] < 1e-3), res
    assert np.allclose(e32[:3], e64[:3], rtol=1e-5, atol=0)
