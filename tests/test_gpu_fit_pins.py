"""Whole fits of the reference's own tests on the GPU against the predictions of their REF files:
fit13 step 1 (BSO4, REF/*_deviation_linear.milady), fit21 (ACE, W, D = 207) and fit24 (ACE, Ta Ti V W,
D = 3,576), REF/train_{energy,force,stress}.out.  Bounds are relative to the largest |REF| value of
each kind.

Two checks, for two different things:
  - the REF predictions lie in the span of the library's energy, force and stress rows (least squares
    on the rows themselves, columns scaled to unit norm): this pins the descriptors, the rows and the
    model evaluation without any solve, to the 11 significant digits REF prints.  For ACE it says the
    library's basis functions span exactly the reference's (Ralf basis, RPI coupling, spline radials);
  - driver.fit_configs reproduces the REF predictions.  Every one of these normal matrices is
    ill-conditioned in FP64 (fit13: eigenvalues 1.7 .. 1.5e16; fit21: the reference keeps 37 of 208
    directions, REF/ndm.out "dgelsd ... RANK 37", with its cut inside a continuum of eigenvalues at
    the rounding floor), so the coefficients along the weakest directions are set by rounding and by
    the solver, and the bounds of this check are the conditioning's, not the rows'.  Two solves are
    checked: the library's own (MLD_SOLVE_SVD, Jacobi on the device), and LAPACK's gelsd -- the
    reference's own solver -- with its rcond (dlamch('E') = 2^-53) on the Gram the library
    accumulated (fit21 at the reference's rank 37).  Where the solve keeps every direction (fit13,
    fit24), the library's solve is also checked against the exact solution of its own normal
    equations: for fit13 REF is itself 1.75e-4 (energies) from that exact solution -- the rounding
    of the reference's FP64 gelsd along the weakest direction, which gelsd on the same Gram repeats.
    gelsd on fit24's Gram reproduces REF as closely as on fit13's (energies 1.9e-6, forces 1.4e-4),
    which it does only when the Gram is the reference's column for column: the library's RPI
    vectors are the reference's (svd_enforce_signs_by_ref's rule), not another basis of their span."""
import numpy as np
import pytest
import scipy.linalg as sl

import fit_pins
from mld_amd import driver, inputs

pytestmark = pytest.mark.gpu

# name, rank of the host solve, bounds of the library's solve, bounds of gelsd on the library's Gram,
# bound of the library's solve against the exact solution of its normal equations (None: not formed)
CASES = [
    ("fit13_lml", None, dict(E=5e-4, F=2.5e-2, S=2.5e-2), dict(E=1e-5, F=5e-4, S=5e-4), 1e-6),
    ("fit21_ace", 37, dict(E=0.2, F=0.6, S=1.0), dict(E=0.12, F=0.5, S=1.0), None),
    ("fit24_ace", None, dict(E=1e-3, F=2.5e-2, S=0.25), dict(E=1e-5, F=5e-4, S=2.5e-3), 3e-4),
]


def host_solve(G, h, lam, rank=None):
    Gs = np.triu(G) + np.triu(G, 1).T + lam * np.eye(G.shape[0])
    if rank is None:
        return sl.lstsq(Gs, h, cond=2.0 ** -53, lapack_driver="gelsd")[0]
    ev, V = np.linalg.eigh(Gs)
    top = np.argsort(-np.abs(ev))[:rank]
    return V[:, top] @ ((V[:, top].T @ h) / ev[top])


@pytest.mark.parametrize("name", ["fit13_lml", "fit21_ace", "fit24_ace"])
def test_reference_predictions_lie_in_the_span_of_the_rows(name):
    nml, configs, z = fit_pins.load(name)
    print(name, len(configs), "configurations", flush=True)
    A, p, kinds = fit_pins.library_rows(nml, z)
    d = np.linalg.norm(A, axis=0)
    d[d == 0] = 1.0
    x, _, rank, _ = sl.lstsq(A / d, p, cond=1e-15, lapack_driver="gelsd")
    r = (A / d) @ x - p
    err = {k: float(np.abs(r[s]).max() / np.abs(p[s]).max()) for k, s in kinds.items()}
    print(name, "rows", A.shape, "numerical rank", rank, "relative max |rows x - REF|:", err, flush=True)
    for k, v in err.items():
        assert v < 1e-9, (name, k, v)


def exact_solve(G, h, lam):
    """(G + lam I) w = h to well below FP64 solver rounding.  Up to 300 columns: LU in 50-digit
    arithmetic, rounded to FP64.  Larger: iterative refinement of the FP64 Cholesky solution with
    the residual formed in x87 extended precision (64-bit significand), which converges while
    cond * 2^-53 < 1 and reaches a forward error of about cond * 2^-64."""
    Gs = np.triu(G) + np.triu(G, 1).T + lam * np.eye(G.shape[0])
    if G.shape[0] <= 300:
        import mpmath as mp

        with mp.workdps(50):
            x = mp.lu_solve(mp.matrix(Gs.tolist()), mp.matrix(h.tolist()))
            return np.array([float(v) for v in x])
    assert np.finfo(np.longdouble).nmant >= 63, "extended precision residuals need x87 long double"
    ev = np.linalg.eigvalsh(Gs)
    cond = ev.max() / ev.min()
    assert ev.min() > 0 and cond * 2.0 ** -53 < 0.25, ("refinement needs cond * 2^-53 well below 1", cond)
    cf = sl.cho_factor(Gs)
    G_ld, h_ld = Gs.astype(np.longdouble), h.astype(np.longdouble)
    x = sl.cho_solve(cf, h).astype(np.longdouble)
    steps = []
    for _ in range(40):
        dx = sl.cho_solve(cf, (h_ld - G_ld @ x).astype(np.float64))
        x += dx
        steps.append(float(np.abs(dx).max() / np.abs(x).max()))
        if len(steps) > 3 and steps[-1] >= 0.5 * steps[-4]:
            break  # at the floor set by the extended-precision residual
    assert steps[-1] < 10.0 * cond * 2.0 ** -64, ("refinement did not converge", cond, steps)
    return x.astype(np.float64)


def relative_gap(p, q):
    """Largest |p - q| of each kind over the largest |q|, predictions of two models."""
    out = {}
    e1, e2 = np.array([a[0] for a in p]), np.array([b[0] for b in q])
    out["E"] = np.abs(e1 - e2).max() / np.abs(e2).max()
    f1, f2 = np.concatenate([a[1] for a in p]), np.concatenate([b[1] for b in q])
    out["F"] = np.abs(f1 - f2).max() / np.abs(f2).max()
    s = [(a[2], b[2]) for a, b in zip(p, q) if a[2] is not None]
    if s:
        s1, s2 = np.array([a for a, _ in s]), np.array([b for _, b in s])
        out["S"] = np.abs(s1 - s2).max() / np.abs(s2).max()
    return out


@pytest.mark.parametrize("name,rank,tol_lib,tol_host,tol_exact", CASES, ids=[c[0] for c in CASES])
def test_driver_fit_matches_reference(name, rank, tol_lib, tol_host, tol_exact):
    nml, configs, z = fit_pins.load(name)
    print(name, len(configs), "configurations", flush=True)
    w, diag = driver.fit_configs(nml, configs, predict=True, export=True)
    err = fit_pins.compare(diag["predictions"], configs, z)
    G, h = diag["G"], diag["h"]
    ev = np.linalg.eigvalsh(np.triu(G) + np.triu(G, 1).T + inputs.ridge(nml) * np.eye(G.shape[0]))
    kept = int((np.abs(ev) > 2.0 ** -53 * np.abs(ev).max()).sum())
    print(name, "library solve, relative max |ours - REF|:", err, "kept", kept, "of", len(ev),
          "eigenvalues %.3e .. %.3e" % (np.abs(ev).min(), np.abs(ev).max()), flush=True)
    wh = host_solve(G, h, inputs.ridge(nml), rank)
    err_h = fit_pins.compare(driver.predict_configs(nml, configs, wh), configs, z)
    print(name, "gelsd on the library's Gram" + (f" at rank {rank}" if rank else ""), err_h, flush=True)
    if tol_exact is not None:
        px = driver.predict_configs(nml, configs, exact_solve(G, h, inputs.ridge(nml)))
        gap = relative_gap(diag["predictions"], px)
        print(name, "library solve against the exact solution:", gap, "; REF against it:",
              fit_pins.compare(px, configs, z), flush=True)
        for k, v in gap.items():
            assert v < tol_exact, (name, "exact", k, v)
    for k, v in err.items():
        assert v < tol_lib[k], (name, "library", k, v)
    for k, v in err_h.items():
        assert v < tol_host[k], (name, "host", k, v)

