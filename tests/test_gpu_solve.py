"""Minimum-norm solve (MLD_SOLVE_SVD, the reference's serial_lsystem_by_svd semantics) on the
GPU against NumPy's SVD least squares, including rank-deficient systems and the fold of the
reference's full D^2 quadratic layout onto the unique columns."""
import numpy as np
import pytest

import oracle as O
from mld_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def ctx64():
    import mld_amd as M

    c = M.Context(j_max=4.0, r_cut=4.7, precision=M.PREC_FP64)
    yield c
    c.close()


def test_svd_equals_cholesky_on_a_regular_system(ctx64):
    from mld_amd.fit import GramFit, SOLVE_CHOL_IR, SOLVE_SVD

    rng = np.random.default_rng(5)
    D = 40
    A = rng.normal(size=(600, D)) @ np.diag(np.linspace(1.0, 30.0, D))
    y = rng.normal(size=600)
    wt = rng.uniform(0.5, 2.0, size=600)
    gf = GramFit(ctx64, ncol=D)
    gf.add_rows(A, y, wt)
    w1, _ = gf.solve(1e-6, SOLVE_CHOL_IR)
    w2, d2 = gf.solve(1e-6, SOLVE_SVD)
    assert rel(w2, w1) < 1e-10, rel(w2, w1)
    assert d2["rel_residual"] < 1e-12


def test_svd_minimum_norm_on_duplicate_columns(ctx64):
    """Two identical columns and lambda = 0: the Gram is singular; the minimum-norm
    solution splits the weight equally (numpy pinv), where Cholesky reports failure."""
    from mld_amd.fit import GramFit, SOLVE_SVD

    rng = np.random.default_rng(6)
    X = rng.normal(size=(300, 9))
    A = np.concatenate([X, X[:, 2:3], X[:, 5:6] * 1.0], axis=1)   # columns 9 = 2, 10 = 5
    y = rng.normal(size=300)
    wt = np.ones(300)
    gf = GramFit(ctx64, ncol=A.shape[1])
    gf.add_rows(A, y, wt)
    G, h = gf.export()
    # the null directions of the computed Gram carry rounding noise near eps * max: a
    # threshold above it (the svd_rcond key) makes the minimum-norm solution well defined
    w, d = gf.solve(0.0, SOLVE_SVD, rcond=1e-12)
    ref = np.linalg.lstsq(A, y, rcond=1e-10)[0]
    assert rel(w, ref) < 1e-9, rel(w, ref)
    assert abs(w[2] - w[9]) < 1e-10 * abs(w).max()
    assert abs(w[5] - w[10]) < 1e-10 * abs(w).max()
    assert rel(G @ w, h) < 1e-10
    # an all-zero column: exactly singular, its coefficient is zero in the minimum-norm solution
    Az = np.concatenate([X, np.zeros((300, 1))], axis=1)
    gz = GramFit(ctx64, ncol=Az.shape[1])
    gz.add_rows(Az, y, wt)
    wz, _ = gz.solve(0.0, SOLVE_SVD)
    assert wz[-1] == 0.0
    assert rel(wz[:-1], np.linalg.lstsq(X, y, rcond=None)[0]) < 1e-10


@pytest.mark.parametrize("lam", [0.0, 1e-3])
def test_svd_quadratic_fold_matches_full_layout_solution(ctx64, lam):
    """The fold of the reference's full 1 + D + D^2 quadratic layout (every off-diagonal
    product column twice) onto the unique columns.  The Gram is a quadratic one (order 2,
    from the batch path, its rows weighted 1e-30) plus 800 random unique-layout rows, so
    the 1,596 unique columns are rank deficient with a clean gap: with lambda = 0 the
    minimum-norm solution depends on the fold's column weights, with lambda > 0 on the
    lambda / 2 ridge of the off-diagonal columns.  The library's solution, expanded with
    w_pq = w_qp = u / 2, must equal NumPy's solution of the full-layout system."""
    import torch
    import mld_amd as M
    from mld_amd.fit import GramFit, SOLVE_SVD, expand_quadratic, unique_pairs

    rng = np.random.default_rng(7)
    pos, cells = [], []
    for _ in range(2):
        p, c = synth.bcc_supercell(rng, 2.855, 2, 0.1)
        pos.append(p)
        cells.append(c)
    types = [np.zeros(len(p), np.int32) for p in pos]
    b = M.Batch(ctx64, pos, cells, types)
    b.neighbours()
    e, g = b.alloc_outputs(True)
    b.bso4(e, g)
    E = np.array([synth.morse(p, c)[0] for p, c in zip(pos, cells)])
    F = np.concatenate([synth.morse(p, c)[1] for p, c in zip(pos, cells)])
    gf = GramFit(ctx64, order=2)
    gf.add_batch(b, E, torch.from_numpy(F.reshape(-1)).to("cuda:0"), w_e=1e-30, w_f=1e-30)
    D, nu = 55, gf.ncol
    Ru = rng.normal(size=(800, nu))
    yu = rng.normal(size=800)
    gf.add_rows(Ru, yu, np.ones(800))
    u, d = gf.solve(lam, SOLVE_SVD, rcond=1e-10)
    Rf = np.zeros((800, 1 + D + D * D))
    Rf[:, :1 + D] = Ru[:, :1 + D]
    for c, (p, q) in enumerate(unique_pairs(D)):
        Rf[:, 1 + D + D * p + q] = Ru[:, 1 + D + c]
        Rf[:, 1 + D + D * q + p] = Ru[:, 1 + D + c]
    if lam == 0.0:
        wf = np.linalg.lstsq(Rf, yu, rcond=1e-10)[0]
    else:
        wf = np.linalg.solve(Rf.T @ Rf + lam * np.eye(Rf.shape[1]), Rf.T @ yu)
    full = expand_quadratic(u, D)
    # lambda = 1e-3 against eigenvalues up to ~5e3: condition ~5e6 bounds the agreement
    assert rel(full, wf) < (1e-9 if lam == 0.0 else 1e-7), rel(full, wf)
    # the unweighted minimum-norm solution of the unique system is a different vector
    if lam == 0.0:
        plain = np.linalg.lstsq(Ru, yu, rcond=1e-10)[0]
        assert rel(expand_quadratic(plain, D), wf) > 1e-3


def test_dense_fit_phi_formulation(ctx64):
    """train_snap_get_parameters_linear (parameters.F90:1064-1146): phi = Amat W Amat^T + lambda I,
    rhs = Amat W ymat, solved by dgelsd (fit_lapack_svd) -- restated with scipy's gelsd on a
    rank-deficient Amat (a repeated column), against add_rows + the minimum-norm solve; and the
    full-rank fit types against the Cholesky solve."""
    import scipy.linalg as sl
    from mld_amd.fit import GramFit, SOLVE_CHOL_IR, SOLVE_SVD

    rng = np.random.default_rng(17)
    D, M = 23, 400
    Amat = rng.normal(size=(D, M)) * np.linspace(0.5, 4.0, D)[:, None]   # (dim, observations)
    ymat = rng.normal(size=M)
    wts = rng.uniform(0.2, 3.0, size=M)
    lam = 1e-3
    phi = (Amat * wts) @ Amat.T + lam * np.eye(D)
    rhs = (Amat * wts) @ ymat
    gf = GramFit(ctx64, ncol=D)
    gf.add_rows(Amat.T, ymat, wts)
    w_chol, _ = gf.solve(lam, SOLVE_CHOL_IR)
    assert rel(w_chol, np.linalg.solve(phi, rhs)) < 1e-10
    # rank-deficient: column 7 repeated as column 11, lambda = 0, svd_rcond above the rounding floor
    Ad = Amat.copy()
    Ad[11] = Ad[7]
    phi0 = (Ad * wts) @ Ad.T
    rhs0 = (Ad * wts) @ ymat
    ref = sl.lstsq(phi0, rhs0, cond=1e-12, lapack_driver="gelsd")[0]
    gd = GramFit(ctx64, ncol=D)
    gd.add_rows(Ad.T, ymat, wts)
    w_svd, _ = gd.solve(0.0, SOLVE_SVD, rcond=1e-12)
    assert rel(w_svd, ref) < 1e-8, rel(w_svd, ref)
    assert abs(w_svd[7] - w_svd[11]) < 1e-9 * abs(w_svd).max()


@pytest.mark.parametrize("method", ["chol_ir", "svd"])
def test_two_stage_qnml_solve_against_numpy(method):
    """mld_gram_solve_qnml on a 2j_max = 4 context (D = 14, 1 + 14 + 105 columns): the linear block
    (G_LL + lambda) w_L = h_L, then the quadratic block (G_QQ + Lambda_Q) w_Q = h_Q - G_LQ^T w_L with
    the unique-column ridge (lambda / 2 off the diagonal pairs), against NumPy on the same G, h."""
    import mld_amd as M
    from mld_amd.fit import GramFit, SOLVE_CHOL_IR, SOLVE_SVD

    c = M.Context(j_max=2.0, r_cut=4.7)
    try:
        rng = np.random.default_rng(11)
        gf = GramFit(c, order=2)
        n, nL = gf.ncol, 1 + c.D
        assert (n, nL) == (120, 15)
        A = rng.normal(size=(900, n)) * np.exp(rng.uniform(-2.0, 2.0, size=n))
        y = rng.normal(size=900)
        wt = rng.uniform(0.5, 2.0, size=900)
        gf.add_rows(A, y, wt)
        lam = 1e-3
        w, diag = gf.solve_qnml(lam, SOLVE_CHOL_IR if method == "chol_ir" else SOLVE_SVD)
        G, h = (A * wt[:, None]).T @ A, (A * wt[:, None]).T @ y
        r = gf.ridge_vector(lam)
        wL = np.linalg.solve(G[:nL, :nL] + np.diag(r[:nL]), h[:nL])
        wQ = np.linalg.solve(G[nL:, nL:] + np.diag(r[nL:]), h[nL:] - G[:nL, nL:].T @ wL)
        assert rel(w[:nL], wL) < 1e-10, rel(w[:nL], wL)
        assert rel(w[nL:], wQ) < 1e-10, rel(w[nL:], wQ)
        # not the one-stage solution of the whole system
        wfull = np.linalg.solve(G + np.diag(r), h)
        assert rel(w, wfull) > 1e-3
    finally:
        c.close()


def test_recorded_solve_loops_equal_launch_by_launch(ctx64, monkeypatch):
    """The Jacobi sweeps of the minimum-norm solve and the preconditioner applications of the
    refinement run as recorded graphs (one launch per sweep / application, replayed with the host's
    convergence checks between replays); MLD_SOLVE_GRAPHS=0 runs the same launches one by one.
    The results are the same bits, sweep counts and iteration counts."""
    from mld_amd.fit import GramFit, SOLVE_CHOL_IR, SOLVE_SVD

    rng = np.random.default_rng(23)
    D = 97  # odd: the Jacobi matrix carries a padding row
    A = rng.normal(size=(700, D)) @ np.diag(np.exp(rng.uniform(-6.0, 6.0, D)))
    y = rng.normal(size=700)
    wt = rng.uniform(0.5, 2.0, size=700)
    out = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("MLD_SOLVE_GRAPHS", graphs)
        gf = GramFit(ctx64, ncol=D)
        gf.add_rows(A, y, wt)
        w_svd, d_svd = gf.solve(1e-9, SOLVE_SVD)
        w_ir, d_ir = gf.solve(1e-9, SOLVE_CHOL_IR)
        out[graphs] = (w_svd, d_svd, w_ir, d_ir)
    (a1, b1, c1, e1), (a0, b0, c0, e0) = out["1"], out["0"]
    assert b1["sweeps"] >= 3 and e1["iterations"] >= 1
    assert np.array_equal(a1, a0) and b1 == b0
    assert np.array_equal(c1, c0) and e1 == e0
    G = (A * wt[:, None]).T @ A
    h = (A * wt[:, None]).T @ y
    ref = np.linalg.solve(G + 1e-9 * np.eye(D), h)
    # column scales e^-6..e^6 give cond(G) near 1e10: the Jacobi SVD's fitted values agree with
    # LAPACK's to about eps * cond(A) (measured 1.4e-8)
    assert rel(G @ c1 + 1e-9 * c1, h) < 1e-10 and rel(A @ a1, A @ ref) < 1e-7
