"""GPU parity of the SO(4) bispectrum descriptors and gradients (C ABI) against the oracle
and the reference's golden vectors.  Tolerances: FP64 form 1e-10 relative, FP32 form 1e-5."""
import numpy as np
import pytest

import oracle as O
from conftest import fit19_weights, structures
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


@pytest.fixture(scope="module")
def ctx32():
    import mld_amd as M

    c = M.Context(j_max=4.0, r_cut=4.7, precision=M.PREC_FP32)
    yield c
    c.close()


def compact(force, nn):
    """(D, nat, nn_max+1, 3) Fortran layout -> compact slots (P+nat, 3, D)."""
    return np.stack([force[:, a, k, :].T for a in range(force.shape[1]) for k in range(nn[a] + 1)])


def test_fit17_golden_fp64(ctx64, fit17):
    for name, pos, cell, it, B in structures(fit17):
        e, _, nn, kind = ctx64.bso4(pos, cell, it, want_grad=False)
        assert rel(e.T, B) < 1e-10, name


def test_fit17_golden_fp32(ctx32, fit17):
    for name, pos, cell, it, B in structures(fit17):
        e, _, _, _ = ctx32.bso4(pos, cell, it, want_grad=False)
        assert rel(e.T, B) < 1e-5, name


def test_fit19_golden_weighted_3ch_fp64(fit19):
    import mld_amd as M

    w1, w3 = fit19_weights()
    c = M.Context(j_max=4.0, r_cut=4.7, nspecies=4, weight=w1, weight_3ch=w3)
    assert c.D == 165
    for name, pos, cell, it, B in structures(fit19):
        e, f, nn, kind = c.bso4(pos, cell, it, want_grad=True)
        assert rel(e.T, B) < 1e-10, name
        eo, go, nl = O.bso4(pos, cell, it, 8, 4.7, nch=3, w1=w1, w3=w3)
        assert rel(compact(f, nn), go) < 1e-10, name
    c.close()
    # the FP32 form of the same three-channel descriptors and gradients
    c = M.Context(j_max=4.0, r_cut=4.7, nspecies=4, weight=w1, weight_3ch=w3, precision=1)
    for name, pos, cell, it, B in structures(fit19):
        e, f, nn, kind = c.bso4(pos, cell, it, want_grad=True)
        assert rel(e.T, B) < 1e-5, name
        eo, go, nl = O.bso4(pos, cell, it, 8, 4.7, nch=3, w1=w1, w3=w3)
        assert rel(compact(f, nn), go) < 1e-5, name
    c.close()


@pytest.mark.parametrize("case", ["bcc128", "small16", "liquid"])
def test_gradients_against_oracle(ctx64, ctx32, case):
    rng = np.random.default_rng({"bcc128": 1, "small16": 2, "liquid": 3}[case])
    if case == "bcc128":
        pos, cell = synth.bcc_supercell(rng, 2.855, 4, 0.05)
    elif case == "small16":
        pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.15)
    else:
        pos, cell = synth.random_packing(rng, 60, 0.085)
        pos = pos + rng.normal(0, 0.1, pos.shape)
    it = np.zeros(len(pos), np.int32)
    eo, go, nl = O.bso4(pos, cell, it, 8, 4.7)
    e, f, nn, kind = ctx64.bso4(pos, cell, it)
    off, nbr = nl[0], nl[1]
    assert np.array_equal(nn, np.diff(off))
    for a in range(len(pos)):
        assert np.array_equal(kind[a, :nn[a]] - 1, nbr[off[a]:off[a + 1]])
        assert np.all(kind[a, nn[a]:] == 0)
    assert rel(e.T, eo) < 1e-10
    assert rel(compact(f, nn), go) < 1e-10
    e32, f32, _, _ = ctx32.bso4(pos, cell, it)
    assert rel(e32.T, eo) < 1e-5
    assert rel(compact(f32, nn), go) < 1e-5


def test_batch_equals_single_bitwise(ctx64):
    import torch
    import mld_amd as M

    rng = np.random.default_rng(21)
    structs = synth.mixed_structures(rng, 6, nmin=16, nmax=80)
    pos = [s[0] for s in structs]
    cells = [s[1] for s in structs]
    types = [np.zeros(len(p), np.int32) for p in pos]
    b = M.Batch(ctx64, pos, cells, types)
    b.neighbours()
    energy, grad = b.alloc_outputs(True)
    b.bso4(energy, grad)
    ctx64.synchronize()
    E = energy.cpu().numpy()
    Gr = grad.cpu().numpy()
    off, idx = b.pair_tables()
    a0 = 0
    s0 = 0
    for p, c, t in zip(pos, cells, types):
        e, f, nn, kind = ctx64.bso4(p, c, t)
        n = len(p)
        assert np.array_equal(E[a0:a0 + n], e.T)
        ns = int(nn.sum()) + n
        assert np.array_equal(Gr[s0:s0 + ns], compact(f, nn))
        a0 += n
        s0 += ns
    assert s0 == Gr.shape[0]


def test_translation_invariance_at_full_size(ctx32):
    """Size-independent property on a 200-structure slice of the batched workload:
    every atom's slots sum to zero (self slot = -sum of neighbour slots)."""
    import mld_amd as M

    rng = np.random.default_rng(20261203)
    structs = synth.mixed_structures(rng, 200)
    pos = [s[0] for s in structs]
    b = M.Batch(ctx32, pos, [s[1] for s in structs], [np.zeros(len(p), np.int32) for p in pos])
    b.neighbours()
    energy, grad = b.alloc_outputs(True)
    b.bso4(energy, grad)
    ctx32.synchronize()
    off, _ = b.pair_tables()
    import torch

    slot_atom = torch.repeat_interleave(torch.arange(b.natoms, device=grad.device),
                                        torch.from_numpy(np.diff(off) + 1).to(grad.device))
    sums = torch.zeros((b.natoms, 3, ctx32.D), dtype=torch.float64, device=grad.device)
    sums.index_add_(0, slot_atom, grad.double())
    scale = grad.double().abs().max()
    assert float(sums.abs().max() / scale) < 1e-5
    assert torch.isfinite(energy).all()


def test_edge_cases(ctx64):
    import mld_amd as M

    # isolated atom (no neighbour): zero descriptor, like the reference's skip
    pos = np.array([[0.0, 0.0, 0.0], [6.0, 6.0, 6.0]])
    cell = np.eye(3) * 20.0
    e, f, nn, kind = ctx64.bso4(pos, cell, np.zeros(2, np.int32))
    assert np.all(nn == 0) and np.all(e == 0) and np.all(f == 0)
    # neighbour overflow -> MLD_E_NEIGH (neighbours.F90:352-358)
    rng = np.random.default_rng(4)
    p2, c2 = synth.bcc_supercell(rng, 2.855, 4, 0.05)
    with pytest.raises(M.MldError) as ex:
        ctx64.bso4(p2, c2, np.zeros(len(p2), np.int32), nn_max=5)
    assert ex.value.code == M.MLD_E_NEIGH
    # species index outside the context
    with pytest.raises(M.MldError) as ex:
        ctx64.bso4(p2, c2, np.ones(len(p2), np.int32))
    assert ex.value.code == M.MLD_E_ARG
    # an empty structure inside a batch
    b = M.Batch(ctx64, [p2, np.zeros((0, 3)), p2[:10]], [c2, c2, c2],
                [np.zeros(len(p2), np.int32), np.zeros(0, np.int32), np.zeros(10, np.int32)])
    b.neighbours()
    energy, grad = b.alloc_outputs(True)
    b.bso4(energy, grad)
    ctx64.synchronize()
    assert energy.shape[0] == len(p2) + 10


@pytest.mark.parametrize("prec", [0, 1])
def test_inner_cut_type3_against_oracle(prec):
    """Pair inner cutoff (r_cut_pair_in, fcut type 3 of width r_cut_width_in) with neighbours in
    its taper: B and dB/dr equal the oracle's (FP64 1e-10, FP32 1e-5)."""
    import mld_amd as M

    rng = np.random.default_rng(12)
    pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.1)
    it = (np.arange(len(pos)) % 2).astype(np.int32)
    rin = np.array([2.3, 2.3, 2.3, 0.0])
    c = M.Context(j_max=4.0, r_cut=4.7, nspecies=2, r_cut_pair_in=rin, r_cut_width_in=0.4, precision=prec)
    e, f, nn, _ = c.bso4(pos, cell, it)
    c.close()
    eo, go, nl = O.bso4(pos, cell, it, 8, 4.7, rin=rin, win=0.4)
    tol = 1e-10 if prec == 0 else 1e-5
    assert rel(e.T, eo) < tol and rel(compact(f, nn), go) < tol


@pytest.mark.parametrize("prec", [0, 1])
def test_lbso4_diag_against_oracle(prec):
    import mld_amd as M

    rng = np.random.default_rng(4)
    pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.1)
    it = np.zeros(len(pos), np.int32)
    c = M.Context(j_max=4.0, r_cut=4.7, lbso4_diag=True, precision=prec)
    assert c.D == 35
    e, f, nn, _ = c.bso4(pos, cell, it)
    c.close()
    eo, go, _ = O.bso4(pos, cell, it, 8, 4.7, diag=True)
    tol = 1e-10 if prec == 0 else 1e-5
    assert rel(e.T, eo) < tol and rel(compact(f, nn), go) < tol
