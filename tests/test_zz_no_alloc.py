"""mld_gram_add_rows_noalloc (the entry point that allocates nothing) against mld_gram_add_rows_dev
on the same rows: two calls in a row, the second smaller, each into caller-owned G and h."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _reference(ctx, D, A, y, w):
    """mld_gram_begin + mld_gram_add_rows_dev, exported as the upper triangle (row-major) and h."""
    ctx.check(ctx.L.mld_gram_begin(ctx.h, D))
    ctx.check(ctx.L.mld_gram_add_rows_dev(ctx.h, D, A.shape[0], _vp(A), _vp(y), _vp(w)))
    G = np.zeros((D, D))
    h = np.zeros(D)
    ctx.check(ctx.L.mld_gram_export(ctx.h, G.ctypes.data_as(C.c_void_p), h.ctypes.data_as(C.c_void_p)))
    return np.triu(G), h


def _noalloc(ctx, D, A, y, w):
    import torch

    G = torch.zeros((D, D), dtype=torch.float64, device=A.device)
    h = torch.zeros(D, dtype=torch.float64, device=A.device)
    torch.cuda.synchronize()
    ctx.check(ctx.L.mld_gram_add_rows_noalloc(ctx.h, D, A.shape[0], _vp(A), _vp(y), _vp(w), _vp(G), _vp(h)))
    ctx.synchronize()
    return G.cpu().numpy(), h.cpu().numpy()


def test_noalloc_equals_add_rows_dev_on_two_calls():
    import torch
    import mld_amd as M

    ctx = M.Context(j_max=4.0, r_cut=4.7)
    rng = np.random.default_rng(31)
    for mc, D in ((5000, 300), (700, 121)):
        # Ac (D, mc) column-major = (mc, D) row-major
        A = torch.from_numpy(rng.standard_normal((mc, D))).cuda()
        y = torch.from_numpy(rng.standard_normal(mc)).cuda()
        w = torch.from_numpy(rng.uniform(0.1, 2.0, mc)).cuda()
        torch.cuda.synchronize()
        Gr, hr = _reference(ctx, D, A, y, w)
        G, h = _noalloc(ctx, D, A, y, w)
        assert np.array_equal(np.triu(G), Gr) and np.array_equal(h, hr)
        assert not np.tril(G, -1).any()   # entries below the diagonal are left as they were
        An = A.cpu().numpy() * np.sqrt(w.cpu().numpy())[:, None]
        Gn = An.T @ An
        assert np.abs(np.triu(G) - np.triu(Gn)).max() <= 1e-12 * np.abs(Gn).max()
    # beyond the scratch: rejected on the host before any launch
    big = 20000
    with pytest.raises(M.MldError) as ex:
        ctx.check(ctx.L.mld_gram_add_rows_noalloc(ctx.h, 1596, big, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16),
                                                   C.c_void_p(16), C.c_void_p(16)))
    assert ex.value.code == M.MLD_E_ARG
    ctx.close()
