"""The descriptor analysis (ml_type = -2) through the library on the reference's fit13 analysis run
(2,249 configurations, 19,521 atoms; tests/golden/fit13_sigma.npz holds every fourth atom of the
files that run wrote): covariance sums on the FP64 MFMA SYRK, per-atom statistics and projections
on the device, eigen-decomposition by the device's Jacobi solver.
  - eigenvalues, |B|^2, |B - mean|^2, Mahalanobis distances and PCA projections of both passages
    against the reference's files (bounds as the oracle's, tests/test_analysis_pins.py);
  - the files analyse() writes parse back to the same values;
  - kernel_dump 2 selects the atoms the oracle's restatement selects from the oracle's distances,
    and kernel_matrix.dat holds their descriptors;
  - the CUR scores equal the oracle's leverage scores and the draw keeps about sum min(1, c p)."""
import os

import numpy as np
import pytest

import analysis as A
import test_analysis_pins as T
from mld_amd import inputs, io

pytestmark = pytest.mark.gpu


def configs_of(z):
    return [(io.Config(pos=z["pos"][a:b], cell=z["cell"][k], itype=z["itype"][a:b], species=["W"], name=str(nm)),
             io.DbClass(str(c), "111", 0, 0, True, False, False, (1.0, 1.0, 1.0)))
            for k, (nm, c, a, b) in enumerate(zip(z["names"], z["cls"], np.cumsum(np.r_[0, z["nat"][:-1]]),
                                                  np.cumsum(z["nat"])))]


_RES = {}


def run(tmp_path_factory):
    if not _RES:
        import mld_amd.analysis as MA

        z, nml, B, cls = T.descriptors()
        configs = configs_of(z)
        print("analysis of", len(configs), "configurations", flush=True)
        out = str(tmp_path_factory.mktemp("analysis"))
        nml = {**nml, "kernel_dump": 2, "write_kernel_matrix": True}
        res = MA.analyse_configs(nml, configs, out_dir=out)
        _RES.update(res=res, out=out, nml=nml, configs=configs)
    return _RES["res"], _RES["out"], _RES["nml"], _RES["configs"]


@pytest.mark.parametrize("which", ["train", "full"])
def test_passages_against_reference_files(which, tmp_path_factory):
    res, out, nml, configs = run(tmp_path_factory)
    z, _, Bo, cls = T.descriptors()
    p = res[which]
    ref = z[f"{which}_eigenvalues"]
    assert np.abs(p.evals - ref).max() < 1e-9 * ref.max()
    assert np.abs((p.evals - ref) / ref).max() < 5e-9
    a = z["atom"]
    assert np.abs((p.norm[a] - z[f"{which}_norm"]) / z[f"{which}_norm"]).max() < 6e-10
    assert np.abs((p.norm_mean[a] - z[f"{which}_norm_mean"]) / z[f"{which}_norm_mean"]).max() < 6e-10
    assert np.abs((p.d2[a] - z[f"{which}_d2"]) / z[f"{which}_d2"]).max() < 1e-6
    for k in range(p.proj.shape[1]):
        s = np.sign(p.proj[a, k] @ z[f"{which}_pca"][:, k])
        assert T.rel(s * p.proj[a, k], z[f"{which}_pca"][:, k]) < 1e-8, k
    assert T.rel(res["B"], Bo) < 1e-12


def test_written_files_parse_back(tmp_path_factory):
    res, out, nml, configs = run(tmp_path_factory)
    for name, arr in (("train_sigma.mcd", res["train"].d2), ("full_sigma.mcd", res["full"].d2)):
        v = np.array([float(ln.split()[2]) for ln in open(os.path.join(out, name))])
        assert np.abs((v - arr) / arr).max() < 6e-10, name
    ev = np.loadtxt(os.path.join(out, "full_sigma.eigenvalues"))
    assert np.abs((ev - res["full"].evals) / res["full"].evals).max() < 6e-10
    sel = [ln.split()[5] for ln in open(os.path.join(out, "train_sigma_selected.mcd"))]
    assert set(sel) == {"07", "08", "13"} and len(sel) == 5142


def test_grid_selection_and_kernel_matrix(tmp_path_factory):
    import mld_amd.analysis as MA

    res, out, nml, configs = run(tmp_path_factory)
    z, _, Bo, cls = T.descriptors()
    # the oracle's distances of the full passage and its selections
    mean, S, Si = T.passage(Bo, cls, tuple(dict.fromkeys(cls)))
    _, _, d2o = A.atom_stats(Bo, mean, Si)
    tr = np.concatenate([np.where(cls == c)[0] for c in ("07", "08", "13")])
    power = float(nml["power_mcd"])
    want = np.concatenate([tr[A.grid_selection(d2o[tr], power, int(nml["np_kernel_ref"]))],
                           A.grid_selection(d2o, power, int(nml["np_kernel_full"]))])
    assert np.array_equal(res["kernel_points"], want)
    draft, zstd = MA.read_kernel_matrix(os.path.join(out, "kernel_matrix.dat"))
    assert draft.shape == (len(want), 55)
    assert np.abs(draft - Bo[want]).max() <= 6e-10 * np.abs(Bo[want]).max()
    assert np.abs(zstd.mean(0)).max() < 1e-10 and np.abs(zstd.std(0) - 1.0).max() < 1e-10


def test_cur_scores_and_draw(tmp_path_factory):
    import mld_amd as M
    import mld_amd.analysis as MA
    from mld_amd.driver import _batch

    res, out, nml, configs = run(tmp_path_factory)
    z, _, Bo, cls = T.descriptors()
    ctx = M.Context(**inputs.context_kwargs(nml, 0))
    try:
        b = _batch(ctx, configs, ["W"])
        e, _ = b.alloc_outputs(False)
        b.bso4(e, None)
        atoms, p, k = MA.cur_scores(ctx, b.natoms, res["cls"], ["07", "08", "13"], -1, -1.0)
        rng = np.random.default_rng(37)
        sel = MA.cur_selection(ctx, b.natoms, res["cls"], ["07", "08", "13"], 1000, -1, -1.0, rng)
    finally:
        ctx.close()
    Phi = (Bo[atoms] - Bo[atoms].mean(0)).T
    assert k == 55
    po = A.cur_leverage(Phi, k)
    assert np.abs(p - po).max() < 1e-8 * po.max() and abs(p.sum() - 1.0) < 1e-10
    expect = np.minimum(1.0, 1000 * p).sum()
    sd = np.sqrt((np.minimum(1.0, 1000 * p) * (1 - np.minimum(1.0, 1000 * p))).sum())
    assert abs(len(sel) - expect) < 6 * sd + 1
    assert np.isin(sel, atoms).all() and len(np.unique(sel)) == len(sel)


def test_kernel_fit_from_the_selected_points(tmp_path_factory, tmp_path):
    """fit13's kernel step (vacancy.ml_03: ml_type 1, kernel_po sigma 0, length 0.05, power 4) on
    fit13's training set with the first 400 points the analysis selected, read back from
    kernel_matrix.dat: the two-stage fit's linear stage is the linear fit of the same rows (fit13
    step 1: the same normal equations to 1e-13), the kernel stage lowers the training error, and the
    kernl coefficient files carry the model."""
    import fit_pins
    import mld_amd.analysis as MA
    from mld_amd import driver

    res, out, nml_a, configs_a = run(tmp_path_factory)
    z_a, _, Bo, _ = T.descriptors()
    pts = res["kernel_points"][:400]
    km = str(tmp_path / "kernel_matrix.dat")
    MA.write_kernel_matrix(km, res["B"], pts, configs_a, res["atom_cfg"], res["atom_ia"])
    draft, zk = MA.read_kernel_matrix(km)
    nml1, configs, zf = fit_pins.load("fit13_lml")
    nml3 = {**nml1, "ml_type": 1, "kernel_type": 4, "sigma_kernel": 0.0, "length_kernel": 0.05, "kernel_power": 4.0}
    assert inputs.fit_order(nml3) == 3
    w1, d1 = driver.fit_configs(nml1, configs, predict=True, export=True)
    w3, d3 = driver.fit_configs(nml3, configs, out_dir=str(tmp_path), predict=True, kernel_points=zk, export=True)
    assert w3.shape == (1 + 55 + 400,)
    # the linear block of the kernel model's normal equations is the linear fit's
    assert T.rel(d3["G"][:56, :56], d1["G"]) < 1e-13 and T.rel(d3["h"][:56], d1["h"]) < 1e-13
    # fit13's linear normal matrix has eigenvalues from 1.7 to 1.5e16: the two solves of it (other
    # summation orders in the Gram) agree to the rounding of its weakest directions
    e1 = np.array([p[0] for p in d1["predictions"]])
    el = np.array([p[0] for p in d3["predictions_linear"]])
    assert T.rel(el, e1) < 1e-4
    els = inputs.elements(nml1)
    eref = inputs.reference_energies(nml1)
    target = np.array([c.header_energy(int(nml1["iread_energy"])) - eref[driver._species_index(c, els)].sum()
                       for c, _ in configs])

    def rmse(preds):
        e = np.array([p[0] if p is not None else np.nan for p in preds])
        m = ~np.isnan(e) & np.array([cl.use_e for _, cl in configs])
        return float(np.sqrt(np.mean((e[m] - target[m]) ** 2)))

    print("energy rmse linear", rmse(d1["predictions"]), "kernel", rmse(d3["predictions"]), flush=True)
    assert rmse(d3["predictions"]) < rmse(d1["predictions"])
    pot = io.read_params_pot(str(tmp_path / "bso4_kernl_params.pot"))
    assert np.array_equal(pot.shape, w3.shape) and np.abs(pot - w3).max() <= 1e-15 * np.abs(w3).max()
    lines = open(tmp_path / "lammps_bso4_kernl_params.pot").read().splitlines()
    assert lines[7].split() == ["9", str(456 + 400 * 55 + 3), "7"]
    assert lines[9].split() == ["7", "4.0", "0", "0", "4", "55", "400"]
    tail = [float(x) for x in lines[-3:]]
    assert tail == [0.0, 0.05, 4.0]
    assert len(lines) == 11 + 456 + 400 * 55 + 3
    assert (tmp_path / "train_energy.out").exists() and (tmp_path / "energy_deviation_kernel.milady").exists()
