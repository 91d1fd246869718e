"""CPU restatement of the reference's descriptor analysis (ml_type = -2) and kernel-point
selection, for tests only (test infrastructure: nothing in the product imports it).

  class_mean            get_dimension_design_Phi_by_class (module_Sigma_matrix.F90:162-202)
  covariance            compute_Sigma_class_by_Phi (module_Sigma_matrix.F90:256-308):
                        Sigma = Phi Phi^T / (D - 1) + 1e-9 I, Phi = B - mean over the class atoms
                        (the divisor is the dimension less one, as the reference has it)
  atom_stats            compute_mcd_dist_for_one_config (main_descriptors_analysis.F90:1315-1405):
                        |B|^2, |B - mean|^2, (B - mean)^T Sigma^-1 (B - mean)
  pca                   get_pca_projection (main_descriptors_analysis.F90:1006-1043): projections
                        on the eigenvectors of the n_pca largest eigenvalues, largest first
  grid_selection        get_index_kernel_from_dist (main_descriptors_analysis.F90:477-662): d2^power
                        over np equal bins between its min and max, the first atom strictly
                        inside each non-empty bin
  cur_leverage          compute_leverage_column_score (module_Sigma_matrix.F90:322-442): column
                        scores |V_k(j, :)|^2 / k of the top-k right singular vectors of Phi

Parity: pinned against the files the reference's fit13 analysis run wrote (eigenvalues, per-atom
norms, distances and projections; tests/test_analysis_pins.py).  grid_selection is restated from
the source and pinned by its invariants (one atom per non-empty open bin, in atom order); the
reference ships no file of its selections.  cur_leverage is pinned by the closed forms
(scores sum to 1; a column orthogonal to the top-k space scores 0); the reference's CUR sampling
draws from Fortran's random_number and is parity unpinned."""
from __future__ import annotations

import numpy as np

SIGMA_RIDGE = 1e-9   # lambda_krr forced to 1e-9 before the inverse (module_Sigma_matrix.F90:293-300)


def class_mean(B: np.ndarray) -> np.ndarray:
    """Mean descriptor over the atoms B (n, D)."""
    return B.sum(axis=0) / B.shape[0]


def covariance(B: np.ndarray, mean: np.ndarray):
    """(Sigma, Sigma^-1) of the atoms B (n, D) about mean."""
    Phi = B - mean
    D = B.shape[1]
    S = Phi.T @ Phi / float(D - 1)
    S = S + SIGMA_RIDGE * np.eye(D)
    return S, np.linalg.inv(S)


def atom_stats(B: np.ndarray, mean: np.ndarray, sigma_inv: np.ndarray):
    """(|B|^2, |B - mean|^2, d2) per atom."""
    V = B - mean
    return (B * B).sum(1), (V * V).sum(1), np.einsum("ij,jk,ik->i", V, sigma_inv, V)


def eigen(S: np.ndarray):
    """Ascending eigenvalues and eigenvectors (diago_serial_real_double)."""
    return np.linalg.eigh(S)


def pca(B: np.ndarray, mean: np.ndarray, evecs: np.ndarray, n_pca: int) -> np.ndarray:
    """(n, n_pca) projections on the eigenvectors of the largest eigenvalues, largest first."""
    return (B - mean) @ evecs[:, ::-1][:, :n_pca]


def grid_selection(d2: np.ndarray, power: float, npoints: int) -> np.ndarray:
    """Indices of the atoms get_index_kernel_from_dist selects from the distances d2."""
    v = d2 ** power
    lo, hi = v.min(), v.max()
    step = (hi - lo) / float(npoints)
    out = []
    for i in range(npoints):
        inside = (v > i * step + lo) & (v < (i + 1) * step + lo)
        if inside.any():
            out.append(int(np.argmax(inside)))
    return np.array(out, np.int64)


def cur_leverage(Phi: np.ndarray, k: int) -> np.ndarray:
    """Column scores of Phi (D, n): |V_k(j, :)|^2 / k, V_k the top-k right singular vectors."""
    _, _, Vt = np.linalg.svd(Phi, full_matrices=False)
    Vk = Vt[:k]
    return (Vk * Vk).sum(0) / float(k)
