#!/usr/bin/env python3
"""Benchmark of the MI355X fitting hot path (one JSON line on stdout).

Default workload (batched descriptors with gradients): 2,000 seeded structures,
N ~ U{16..432}, bcc / fcc / liquid-like random packing at 0.085 A^-3 with
Gaussian displacements sigma ~ U[0.02, 0.2] A; SO(4) bispectrum 2j_max = 8
(55 components), r_cut = 4.7 A; per-atom B plus dB/dr for every neighbour pair.
One step = neighbour lists + descriptors + gradients for the whole batch, inputs
(positions, cells, species) already in HBM.  ``value`` is the FP32 (mixed) form;
the plain FP64 form is measured in the same run and reported beside it together
with the FP32-vs-FP64 error.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nstruct S] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

SEED = 20261203
# Reference-counted work per atom (complex multiply-adds x 8 flops): U/dU recursion
# 4*285 per neighbour, three CG packings of 34,946 terms, and the zaxpy force
# contraction 3 * 5,907 per neighbour (2j_max = 8, 55 components).
NU, NCG, NZAXPY = 285, 34946, 5907


def algorithmic_flops(natoms: int, npairs: int) -> float:
    return 8.0 * (4 * NU * npairs + 3 * NCG * natoms + 3 * NZAXPY * npairs)


def algorithmic_bytes(natoms: int, npairs: int, D: int, esz: int) -> float:
    # positions + pair geometry in, B and dB/dr out
    return 24.0 * natoms + 32.0 * npairs + esz * D * (natoms + 3 * (npairs + natoms))


def make_workload(nstruct: int):
    from mld_amd import synth

    rng = np.random.default_rng(SEED)
    structs = synth.mixed_structures(rng, nstruct)
    pos = [s[0] for s in structs]
    cells = [s[1] for s in structs]
    types = [np.zeros(p.shape[0], np.int32) for p in pos]
    return pos, cells, types


def time_precision(prec: int, pos, cells, types, steps: int, warmup: int):
    import torch
    import mld_amd as M

    ctx = M.Context(j_max=4.0, r_cut=4.7, precision=prec)
    b = M.Batch(ctx, pos, cells, types)
    b.neighbours()
    energy, grad = b.alloc_outputs(True)
    stream = torch.cuda.ExternalStream(ctx.stream())
    for _ in range(warmup):
        b.neighbours()
        b.bso4(energy, grad, sync=False)
    ctx.synchronize()
    torch.cuda.synchronize()
    kern_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.neighbours()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        b.bso4(energy, grad, sync=False)
        e1.record(stream)
        e1.synchronize()
        kern_ms.append(e0.elapsed_time(e1))
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return dict(ms_per_step=1000.0 * (t1 - t0) / steps, kernel_ms=float(np.mean(kern_ms)), natoms=b.natoms,
                npairs=b.npairs, D=ctx.D, energy=energy, grad=grad, ctx=ctx, batch=b)


def cpu_baseline(pos, cells, types, budget_s: float = 15.0):
    """Oracle (C FP64 restatement of the reference loops), one core, bounded sample."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O

    t0 = time.perf_counter()
    atoms = 0
    n = 0
    for p, c, t in zip(pos, cells, types):
        O.bso4(p, c, t, 8, 4.7, want_grad=True)
        atoms += p.shape[0]
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return dict(value=atoms / dt, unit="atom-envs/s", cores=1, kind="port",
                sample=f"first {n} structures ({atoms} atom-envs) of the same workload, B + dB/dr, FP64, {dt:.1f} s")


# ----------------------------------------------------------------------------- fits
def fit_workload(kind: str, scale: float):
    """config 1: 100 x 128-atom bcc Fe (a=2.855, sigma=0.05), linear E+F fit, w_e=w_f=1.
    config 3: 8,000 bcc Fe structures with N~U{128..256}, quadratic E+F fit, w_e=1, w_f=0.1, lambda=1e-8."""
    from mld_amd import synth

    if kind == "fit-linear":
        rng = np.random.default_rng(SEED + 1)
        n = max(1, int(round(100 * scale)))
        structs = [synth.bcc_supercell(rng, 2.855, 4, 0.05) for _ in range(n)]
        return structs, dict(order=1, w_e=1.0, w_f=1.0, lam=1e-8, chunk=100)
    rng = np.random.default_rng(SEED + 3)
    n = max(1, int(round(8000 * scale)))
    structs = synth.bcc_vacancy_structures(rng, n)
    return structs, dict(order=2, w_e=1.0, w_f=0.1, lam=1e-8, chunk=500)


def run_fit(structs, cfg, prec: int, mixed: bool, steps: int, warmup: int):
    import torch
    import mld_amd as M
    from mld_amd import synth
    from mld_amd.fit import GramFit

    ctx = M.Context(j_max=4.0, r_cut=4.7, precision=prec)
    chunk = cfg["chunk"]
    batches = []
    for s0 in range(0, len(structs), chunk):
        part = structs[s0:s0 + chunk]
        b = M.Batch(ctx, [p for p, _ in part], [c for _, c in part], [np.zeros(len(p), np.int32) for p, _ in part])
        b.neighbours()
        po, pi, pd = b.pair_tensors()
        E, F = synth.morse_from_pairs(po, pi, pd, b.atom_off)
        batches.append((b, E.cpu().numpy(), F.reshape(-1).contiguous()))
    natoms = sum(b.natoms for b, _, _ in batches)
    out = None
    times = []
    for it in range(warmup + steps):
        ctx.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gf = GramFit(ctx, order=cfg["order"], mixed=mixed, keep_rows=mixed)
        for b, E, F in batches:
            b.neighbours()
            e, g = b.alloc_outputs(True)
            b.bso4(e, g, sync=False)
            gf.add_batch(b, E, F, cfg["w_e"], cfg["w_f"])
        w, diag = gf.solve(cfg["lam"])
        ctx.synchronize()
        t1 = time.perf_counter()
        if it >= warmup:
            times.append(t1 - t0)
        out = (w, diag, gf)
    w, diag, gf = out
    G, h = gf.export()
    return dict(ctx=ctx, w=w, diag=diag, G=G, h=h, natoms=natoms, rows=sum(b.nstruct + 3 * b.natoms for b, _, _ in batches),
                s_per_fit=float(np.mean(times)), batches=batches, gf=gf)


def prediction_rmse(ctx, batches, order, w, nval: int = 2):
    """Energy (per atom) and force RMSE of coefficients w against the Morse targets of the
    first batches, predicted on the device with the context's descriptor precision."""
    import torch
    from mld_amd import fit

    fit.model_set(ctx, order, w)
    e2 = f2 = 0.0
    ne = nf = 0
    for b, E, F in batches[:nval]:
        b.neighbours()
        e, g = b.alloc_outputs(True)
        b.bso4(e, g)
        Ep, Fp, _ = fit.predict(ctx, b)
        nat = torch.from_numpy(np.diff(b.atom_off).astype(np.float64)).to(Ep.device)
        Et = torch.from_numpy(E).to(Ep.device)
        e2 += float((((Ep - Et) / nat) ** 2).sum())
        f2 += float(((Fp.reshape(-1) - F) ** 2).sum())
        ne += b.nstruct
        nf += 3 * b.natoms
    return dict(energy_rmse_meV_per_atom=1e3 * math.sqrt(e2 / ne), force_rmse_meV_per_A=1e3 * math.sqrt(f2 / nf))


def main_fit(args):
    import torch

    structs, cfg = fit_workload(args.workload, args.scale)
    r64 = run_fit(structs, cfg, 0, False, args.steps, args.warmup)
    rmx = run_fit(structs, cfg, 1, True, args.steps, args.warmup)
    w64, wmx = r64["w"], rmx["w"]
    p64 = prediction_rmse(r64["ctx"], r64["batches"], cfg["order"], w64)
    pmx_on64 = prediction_rmse(r64["ctx"], r64["batches"], cfg["order"], wmx)
    pmx = prediction_rmse(rmx["ctx"], rmx["batches"], cfg["order"], wmx)
    natoms = r64["natoms"]
    value = natoms / rmx["s_per_fit"]
    D = r64["G"].shape[0]
    # the SYRK work of the Gram (upper tiles) and descriptor work
    out = {
        "metric": "atom-envs/s (desc+grads+Gram) mixed vs FP64 at stated err; serve p50/p99 ms",
        "value": value, "unit": "atom-envs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * rmx["s_per_fit"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+bf16x3", "data": f"synthetic (seed {SEED}), Morse targets",
        "config": {"workload": f"{args.workload}: {len(structs)} bcc Fe structures, {natoms} atoms, "
                               f"{r64['rows']} rows, {D} columns, order {cfg['order']}, lambda {cfg['lam']}",
                   "parallelism": "single GPU"},
        "fp64": {"value": natoms / r64["s_per_fit"], "ms_per_step": 1e3 * r64["s_per_fit"], "solve": r64["diag"]},
        "mixed_speedup_vs_fp64": r64["s_per_fit"] / rmx["s_per_fit"],
        "mixed": {"solve": rmx["diag"]},
        "err_mixed_vs_fp64": {"coef_max_rel": float(np.abs(wmx - w64).max() / np.abs(w64).max()),
                              "coef_norm_rel": float(np.linalg.norm(wmx - w64) / np.linalg.norm(w64)),
                              "rmse_fp64": p64, "rmse_mixed_coef_fp64_desc": pmx_on64, "rmse_mixed": pmx,
                              "force_rmse_diff_meV_per_A": abs(pmx["force_rmse_meV_per_A"] -
                                                               p64["force_rmse_meV_per_A"])},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nstruct", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="desc", choices=["desc", "fit-linear", "fit-quadratic"])
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the fit workload's structures")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU sharding is a design note "
                                   "(per-GPU Gram + one allreduce), run with --gpus 1"}))
        sys.exit(2)
    if args.workload != "desc":
        return main_fit(args)
    import torch

    pos, cells, types = make_workload(args.nstruct)
    r64 = time_precision(0, pos, cells, types, args.steps, args.warmup)
    r32 = time_precision(1, pos, cells, types, args.steps, args.warmup)
    # FP32 vs FP64 error, normalised by the largest magnitude of each output
    e64, e32 = r64["energy"], r32["energy"].double()
    g64, g32 = r64["grad"], r32["grad"].double()
    err_b = float((e32 - e64).abs().max() / e64.abs().max())
    err_g = float((g32 - g64).abs().max() / g64.abs().max())
    natoms, npairs, D = r32["natoms"], r32["npairs"], r32["D"]
    flops = algorithmic_flops(natoms, npairs)
    ach = flops / (r32["kernel_ms"] * 1e-3) / 1e12
    ach64 = flops / (r64["kernel_ms"] * 1e-3) / 1e12
    value = natoms / (r32["ms_per_step"] * 1e-3)
    out = {
        "metric": "atom-envs/s (desc+grads+Gram) mixed vs FP64 at stated err; serve p50/p99 ms",
        "value": value,
        "unit": "atom-envs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seed {SEED})",
        "config": {"workload": "batched SO(4) bispectrum B + dB/dr, 2j_max=8, r_cut=4.7 A, "
                               f"{args.nstruct} structures N~U{{16..432}} bcc/fcc/liquid rho=0.085",
                   "natoms": natoms, "npairs": npairs, "D": D, "parallelism": "single GPU"},
        "fp64": {"value": natoms / (r64["ms_per_step"] * 1e-3), "ms_per_step": r64["ms_per_step"],
                 "kernel_ms": r64["kernel_ms"], "achieved_tflops": ach64, "frac": ach64 / 78.6},
        "mixed_speedup_vs_fp64": r64["ms_per_step"] / r32["ms_per_step"],
        "err_fp32_vs_fp64": {"B_max_rel": err_b, "dBdr_max_rel": err_g},
        "roofline": {"bound": "mfma", "achieved": ach, "peak": 157.3, "unit": "TFLOP/s", "frac": ach / 157.3,
                     "traffic": None, "kernel": "bso4_kernel<float,8>", "kernel_ms": r32["kernel_ms"],
                     "algorithmic_flops_per_launch": flops,
                     "algorithmic_bytes_per_launch": algorithmic_bytes(natoms, npairs, D, 4)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pos, cells, types)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
