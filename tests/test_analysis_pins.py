"""The descriptor analysis of the reference's fit13 test (vacancy.ml_02, ml_type = -2: all 2,249
configurations of its classes, 19,521 atoms) restated on the oracle's descriptors, against the
files that run wrote (tests/golden/fit13_sigma.npz, every fourth atom):
  - the 55 eigenvalues of the train covariance (classes_for_mcd "07 08 13", 5,142 atoms) and of the
    full covariance (every class);
  - per atom |B|^2, |B - mean|^2 and the Mahalanobis distance of both passages;
  - the three PCA projections of both passages (each eigenvector up to its sign).
The covariance divides by D - 1 = 54, not by the atom count (compute_Sigma_class_by_Phi,
module_Sigma_matrix.F90:283): with the atom count the eigenvalues miss by the factor 5,142 / 54
(test_covariance_divisor_is_the_dimension)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import analysis as A
import oracle as O
from mld_amd import inputs

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRAIN = ("07", "08", "13")
_CACHE = {}


def descriptors():
    if not _CACHE:
        z = dict(np.load(os.path.join(G, "fit13_sigma.npz")))
        nml = inputs.read_input_ml(os.path.join(G, "fit13_sigma_input.ml"))
        off = np.concatenate([[0], np.cumsum(z["nat"])])
        jj, rc = int(round(2 * float(nml["j_max"]))), float(nml["r_cut"])

        def one(k):
            return O.bso4(z["pos"][off[k]:off[k + 1]], z["cell"][k], z["itype"][off[k]:off[k + 1]], jj, rc,
                          want_grad=False)[0]

        with ThreadPoolExecutor(8) as ex:
            B = np.concatenate(list(ex.map(one, range(len(z["nat"])))))
        cls = np.repeat(z["cls"], z["nat"])
        _CACHE.update(z=z, nml=nml, B=B, cls=cls)
    return _CACHE["z"], _CACHE["nml"], _CACHE["B"], _CACHE["cls"]


def passage(B, cls, classes):
    m = np.isin(cls, classes)
    mean = A.class_mean(B[m])
    S, Si = A.covariance(B[m], mean)
    return mean, S, Si


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("which", ["train", "full"])
def test_eigenvalues_norms_and_distances(which):
    z, nml, B, cls = descriptors()
    classes = TRAIN if which == "train" else tuple(sorted(set(cls)))
    mean, S, Si = passage(B, cls, classes)
    ev, _ = A.eigen(S)
    ref = z[f"{which}_eigenvalues"]
    # the eigenvalues run from 0.27 to 1.8e6: each to 1e-9 of the largest, and the small ones to
    # 2e-9 relative (the 10 printed digits)
    assert np.abs(ev - ref).max() < 1e-9 * ref.max()
    assert np.abs((ev - ref) / ref).max() < 2e-9
    nrm, nrm_mean, d2 = A.atom_stats(B[z["atom"]], mean, Si)
    # each value to the 10 significant digits the files print (half a unit of the last: up to 5e-10)
    assert np.abs((nrm - z[f"{which}_norm"]) / z[f"{which}_norm"]).max() < 6e-10
    assert np.abs((nrm_mean - z[f"{which}_norm_mean"]) / z[f"{which}_norm_mean"]).max() < 6e-10
    assert np.abs((d2 - z[f"{which}_d2"]) / z[f"{which}_d2"]).max() < 1e-6


@pytest.mark.parametrize("which", ["train", "full"])
def test_pca_projections(which):
    z, nml, B, cls = descriptors()
    classes = TRAIN if which == "train" else tuple(sorted(set(cls)))
    mean, S, Si = passage(B, cls, classes)
    _, V = A.eigen(S)
    P = A.pca(B[z["atom"]], mean, V, int(nml["n_pca"]))
    ref = z[f"{which}_pca"]
    for k in range(P.shape[1]):
        s = np.sign(P[:, k] @ ref[:, k])
        assert rel(s * P[:, k], ref[:, k]) < 1e-8, k


def test_covariance_divisor_is_the_dimension():
    z, nml, B, cls = descriptors()
    m = np.isin(cls, TRAIN)
    Phi = B[m] - A.class_mean(B[m])
    wrong = np.linalg.eigvalsh(Phi.T @ Phi / (m.sum() - 1) + 1e-9 * np.eye(B.shape[1]))
    assert np.abs(wrong - z["train_eigenvalues"]).max() > 0.9 * z["train_eigenvalues"].max()


def test_grid_selection_invariants():
    """One atom per non-empty open bin of d2^power, the first such atom in atom order."""
    rng = np.random.default_rng(3)
    d2 = rng.lognormal(0.0, 2.0, 5000)
    sel = A.grid_selection(d2, 0.05, 100)
    v = d2 ** 0.05
    lo, step = v.min(), (v.max() - v.min()) / 100
    bins = np.floor((v[sel] - lo) / step).astype(int)
    assert len(set(bins)) == len(sel) and np.all(np.diff(bins) > 0)
    for i, b in zip(sel, bins):
        inside = np.where((v > b * step + lo) & (v < (b + 1) * step + lo))[0]
        assert inside[0] == i
    nonempty = {int(np.floor((x - lo) / step)) for x in v if lo < x < v.max() and (x - lo) / step % 1 != 0}
    assert set(bins) == nonempty


def test_cur_leverage_closed_forms():
    rng = np.random.default_rng(4)
    Phi = rng.normal(size=(6, 40))
    Phi[:, 7] = 0.0  # a zero column has no leverage
    p = A.cur_leverage(Phi, 4)
    assert abs(p.sum() - 1.0) < 1e-12 and p[7] == 0.0 and (p >= 0).all()
    # full rank k = 6: every column's score is its leverage in the row space, |Phi^+ Phi e_j|^2 / 6
    p6 = A.cur_leverage(Phi, 6)
    H = np.linalg.pinv(Phi) @ Phi
    assert np.abs(p6 - np.diag(H) / 6).max() < 1e-12


def test_analysis_files_in_the_reference_format(tmp_path):
    """train_sigma.mcd, full_sigma.pca and train_sigma.eigenvalues written from the values of the
    reference's own lines reproduce those lines byte for byte (tests/golden/fit13_sigma_excerpt)."""
    import mld_amd.analysis as MA
    from mld_amd import io

    ex = os.path.join(G, "fit13_sigma_excerpt")
    z, nml, B, cls = descriptors()
    configs = [(io.Config(pos=np.zeros((n, 3)), cell=np.eye(3), itype=np.zeros(n, np.int32), species=["W"], name=str(nm)),
                io.DbClass(str(c), "111", 0, 0, True, False, False, (1.0, 1.0, 1.0)))
               for nm, c, n in zip(z["names"], z["cls"], z["nat"])]
    atom_cfg = np.repeat(np.arange(len(configs)), z["nat"])
    atom_ia = np.concatenate([np.arange(n) for n in z["nat"]])
    lines = open(os.path.join(ex, "train_sigma.mcd")).read().splitlines(True)
    v = np.array([[float(x) for x in ln.split()[:3]] for ln in lines])
    k = len(lines)
    MA.write_mcd(str(tmp_path / "m"), np.arange(k), configs, atom_cfg, atom_ia, v[:, 0], v[:, 1], v[:, 2],
                 prefix=nml["db_path"])
    assert open(tmp_path / "m").read().splitlines(True) == lines
    lines = open(os.path.join(ex, "full_sigma.pca")).read().splitlines(True)
    v = np.array([[float(x) for x in ln.split()[:5]] for ln in lines])
    k = len(lines)
    MA.write_pca(str(tmp_path / "p"), configs, atom_cfg[:k], atom_ia[:k], v[:, :3], v[:, 4], nml["db_path"])
    assert open(tmp_path / "p").read().splitlines(True) == lines
    ev = np.loadtxt(os.path.join(ex, "train_sigma.eigenvalues"))
    MA.write_eigenvalues(str(tmp_path / "e"), ev)
    assert open(tmp_path / "e").read() == open(os.path.join(ex, "train_sigma.eigenvalues")).read()
