"""GPU low-rank factors and factor MDS against FP64 references."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.proximity import DeviceForest, lowrank_factors, mds_factors, mds_factors_mixed, quantize_factors

from helpers import small_forest

pytestmark = pytest.mark.gpu


def _setup(n=1500, T=60, seed=21):
    x, y, f = small_forest(_lib.DATA_COVTYPE_LIKE, n, 54, 7, T, seed, mtry=7)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    nd = O.leaf_assign(f, x)
    C = O.counts(f, nd, O.ALL, 0).reshape(n, n, order="F").astype(np.float64)
    return f, keys, C / T


def e4m3_decode(c: np.ndarray) -> np.ndarray:
    c = c.astype(np.int64)
    e = (c >> 3) & 15
    m = c & 7
    v = np.where(e == 0, m * 2.0 ** -9, (1 + m / 8.0) * 2.0 ** (e - 7.0))
    return np.where(c & 0x80, -v, v)


def test_factor_storage_formats(gpu):
    rng = np.random.default_rng(3)
    U = rng.normal(size=(4000, 16)).astype(np.float32) * np.linspace(0.01, 3, 16).astype(np.float32)
    U[7, 3] = 0.0
    Ut = torch.from_numpy(U).to(gpu)
    amax = np.abs(U).max(axis=0)
    # INT8: q = round(u / s), s = max|u| / 127
    q, s = quantize_factors(Ut, _lib.PREC_INT8)
    s = s.cpu().numpy()
    assert np.allclose(s, amax / 127, rtol=1e-7)
    assert np.array_equal(q.cpu().numpy(), np.clip(np.rint(U / s.astype(np.float32)), -127, 127).astype(np.int8))
    # BF16: round to nearest even of the float32 bits
    q, s = quantize_factors(Ut, _lib.PREC_BF16)
    b = U.view(np.uint32).astype(np.uint64)
    want = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
    assert np.array_equal(q.cpu().numpy().view(np.uint16), want)
    # FP8 e4m3: the code decodes to the nearest representable value of u / s
    q, s = quantize_factors(Ut, _lib.PREC_FP8)
    s = s.cpu().numpy()
    assert np.allclose(s, amax / 448, rtol=1e-7)
    codes = q.cpu().numpy().view(np.uint8)
    dec = e4m3_decode(codes)
    grid = np.unique(np.abs(e4m3_decode(np.arange(0, 0x7F))))
    v = (U / s.astype(np.float32)).astype(np.float64)
    nearest = grid[np.abs(np.abs(v)[..., None] - grid[None, None, :]).argmin(-1)]
    assert np.allclose(np.abs(dec), nearest, rtol=0, atol=0) or np.all(
        np.abs(np.abs(dec) - np.abs(v)) <= np.abs(nearest - np.abs(v)) + 1e-12)
    assert e4m3_decode(np.array([0x7E]))[0] == 448.0 and e4m3_decode(np.array([0x01]))[0] == 2.0 ** -9


def test_lowrank_factors_against_fp64_eigh(gpu):
    print("lowrank factors", flush=True)
    f, keys, P = _setup()
    r = 24
    U, eig = lowrank_factors(keys, f, rank=r, oversample=16, power_iters=4)
    w = np.linalg.eigvalsh(P)[::-1]
    print("top eig gpu", eig[:6], "fp64", w[:6], flush=True)
    assert np.allclose(eig[:8], w[:8], rtol=2e-3)
    Uh = U.cpu().numpy().astype(np.float64)
    lam, V = np.linalg.eigh(P)
    Pr = (V[:, -r:] * lam[-r:]) @ V[:, -r:].T
    err = np.linalg.norm(P - Uh @ Uh.T)
    best = np.linalg.norm(P - Pr)  # Eckart-Young: the optimal rank-r error
    print("rank", r, "approximation error", err, "optimal", best, flush=True)
    assert err <= 1.01 * best


def _fp64_factor_mds(Uh: np.ndarray):
    Pt = Uh @ Uh.T
    np.fill_diagonal(Pt, 1.0)
    D = 1.0 - Pt
    np.fill_diagonal(D, 0.0)
    D2 = D * D
    rr = D2.mean(axis=1)
    B = -0.5 * (D2 - rr[:, None] - rr[None, :] + rr.mean())
    from scipy.linalg import eigh
    n = B.shape[0]
    wv, V = eigh(B, subset_by_index=[n - 8, n - 1])
    o = np.argsort(wv)[::-1]
    return V[:, o[:3]] * np.sqrt(wv[o[:3]]), wv[o]


@pytest.mark.parametrize("rank", [32, 128])
def test_factor_mds_fp64_matches_closed_form(gpu, rank):
    f, keys, P = _setup(n=1200, T=50, seed=22)
    U, _ = lowrank_factors(keys, f, rank=rank, oversample=16, power_iters=3)
    Uq, sc = quantize_factors(U, _lib.PREC_FP32)
    X, eig, it = mds_factors(Uq, sc, _lib.PREC_FP32, precision=_lib.PREC_FP64)
    Xref, w = _fp64_factor_mds(U.cpu().numpy().astype(np.float64))
    print(it, eig[:4], w[:4], flush=True)
    assert np.allclose(eig[:3], w[:3], rtol=1e-9)
    if w[2] / w[3] > 1.02:
        assert O.embedding_errors(X.cpu().numpy(), Xref)["rel_frobenius"] < 1e-6


@pytest.mark.parametrize("rank", [32, 128])
@pytest.mark.parametrize("fmt", [_lib.PREC_FP32, _lib.PREC_BF16, _lib.PREC_INT8, _lib.PREC_FP8])
def test_factor_mds_storage_precisions(gpu, fmt, rank):
    f, keys, P = _setup(n=1200, T=50, seed=22)
    U, _ = lowrank_factors(keys, f, rank=rank, oversample=16, power_iters=3)
    Xref, w = _fp64_factor_mds(U.cpu().numpy().astype(np.float64))
    Uq, sc = quantize_factors(U, fmt)
    X, eig, it = mds_factors(Uq, sc, fmt, precision=_lib.PREC_FP32)
    e = O.embedding_errors(X.cpu().numpy(), Xref)
    print("fmt", fmt, "iters", it, "eig", eig[:4], "ref", w[:4], e, flush=True)
    assert np.allclose(eig[:3], w[:3], rtol={_lib.PREC_FP32: 1e-4, _lib.PREC_BF16: 2e-2, _lib.PREC_INT8: 2e-2,
                                             _lib.PREC_FP8: 1e-1}[fmt])


def _fp64_factor_B(Uh: np.ndarray) -> np.ndarray:
    Pt = Uh @ Uh.T
    np.fill_diagonal(Pt, 1.0)
    D = 1.0 - Pt
    np.fill_diagonal(D, 0.0)
    D2 = D * D
    rr = D2.mean(axis=1)
    return -0.5 * (D2 - rr[:, None] - rr[None, :] + rr.mean())


@pytest.mark.parametrize("rank", [32, 64, 96, 128])
@pytest.mark.parametrize("fmt", [_lib.PREC_BF16, _lib.PREC_INT8, _lib.PREC_FP8, _lib.PREC_FP32])
def test_mixed_factor_mds_mfma_then_fp32_refinement(gpu, rank, fmt):
    # The bf16 matrix-core operator (MFMA) on the stored factors gives
    # eigenvalues at bf16/storage accuracy; FP32 refinement sweeps on the FP32
    # factors then reach the FP64 closed form of those factors.
    print(f"mixed factor mds rank {rank} fmt {fmt}", flush=True)
    f, keys, P = _setup(n=1300, T=50, seed=24)
    U, _ = lowrank_factors(keys, f, rank=rank, oversample=16, power_iters=3)
    Uh = U.cpu().numpy().astype(np.float64)
    B = _fp64_factor_B(Uh)
    from scipy.linalg import eigh
    n = B.shape[0]
    w, Vr = eigh(B, subset_by_index=[n - 8, n - 1])
    w = w[::-1]
    Uq, sc = quantize_factors(U, fmt)
    X0, e0, it0 = mds_factors_mixed(Uq, sc, fmt, None, max_iter=300, tol=2e-3)
    storage_tol = {_lib.PREC_FP32: 1e-2, _lib.PREC_BF16: 2e-2, _lib.PREC_INT8: 5e-2, _lib.PREC_FP8: 1e-1}[fmt]
    print("mfma only", it0, e0[:4], w[:4], flush=True)
    assert np.allclose(e0[:3], w[:3], rtol=storage_tol)
    X, e, it = mds_factors_mixed(Uq, sc, fmt, U, max_iter=300, tol=2e-3, refine_max=30, refine_tol=1e-6)
    Xh = X.cpu().numpy()
    lam = np.sum(Xh * Xh, axis=0)
    V = Xh / np.sqrt(lam)
    res = np.linalg.norm(B @ V - V * lam, axis=0) / lam
    print("refined", it, e[:4], res, flush=True)
    # the refinement operator carries u, u*v and Q as bf16 pairs: about 2^-17 relative
    assert np.allclose(e[:3], w[:3], rtol=1e-5)
    assert res.max() < 5e-5
