#!/usr/bin/env python3
"""Benchmark of the walker-batched VMC engine on one MI355X.

Workload (the configuration the headline metric is quoted on): C2-like dimer,
two Z_eff=4 centres at +-1.17 bohr, 4+4 valence electrons, 58 contracted
spherical GTOs, Gaussian pseudopotential with l=0,1 nonlocal channels
(12-point quadrature), 24 CSFs / 48 determinants; 16,384 walkers.
A step is one VMC step of every walker: a Metropolis sweep over the 8
electrons (one-electron ratio + accept) and the full evaluation (local energy
including the pseudopotential quadrature).  Inputs are resident in HBM before
timing; nothing is copied to the host inside the timed region.

Prints one JSON line (see README/DESIGN for the fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3    # MI355X FP32


def eval_full_flops(s, x):
    """Algorithmic FP operations of one full evaluation of every walker in x
    (counted per term of the formulas the kernel evaluates):
      e-e pair Pade+polynomial 12 + 6 nordb; e-e-n per (pair, centre)
      12 nterms + 60; e-n per (electron, centre) 12 + 6 norda;
      basis per (point, shell) 30 + 12 nprim + 10 (2l+1) (value, gradient,
      laplacian) or 8 + 8 nprim + 3 (2l+1) (value only);
      orbital contraction 2 NB per component and orbital;
      pseudopotential: per quadrature point of an electron inside a nonlocal
      channel's range, value-only basis + 2 NB nocc + the Jastrow value change
      of one electron ((N-1) pair terms + one e-n term)."""
    from champ_amd.system import nterms4
    N, M = int(s.nelec), int(s.ncent)
    nt = nterms4(s.nordc)
    npair = N * (N - 1) // 2
    pair = 12 + 6 * s.nordb + M * (12 * nt + 60)
    en = M * (12 + 6 * s.norda)
    # shells per centre type (grouped like the library does: one per radial fn)
    shells = []
    for ic in range(M):
        t = s.iwctype[ic]
        seen = {}
        for q in range(s.ctype_bas_off[t], s.ctype_bas_off[t + 1]):
            r = s.bas_rad[q]
            key = (r, s.bas_l[q])
            seen[key] = seen.get(key, 0) + 1
        for (r, l), nm in seen.items():
            npr = int(s.rad_prim_off[r + 1] - s.rad_prim_off[r])
            shells.append((npr, nm))
    basis_full = sum(30 + 12 * npr + 10 * nm for npr, nm in shells)
    basis_val = sum(8 + 8 * npr + 3 * nm for npr, nm in shells)
    nocc = len(set(np.asarray(s.det_up).ravel().tolist()))
    NB = int(s.nbasis)
    per_walker = npair * pair + N * en + N * (basis_full + 5 * 2 * NB * nocc)
    flops = float(per_walker) * x.shape[0]
    if s.nloc:
        # electrons within range of a nonlocal channel (|v_l| > 1e-4)
        cent = np.asarray(s.cent)
        r = np.linalg.norm(x[:, :, None, :] - cent[None, None, :, :], axis=-1)
        inr = np.zeros_like(r, dtype=bool)
        for ic in range(M):
            t = s.iwctype[ic]
            if s.npotd[t] <= 1:
                continue
            for l in range(s.npotd[t]):
                if l + 1 == s.lpotp1[t]:
                    continue
                b, e = s.ecp_off[t * 5 + l], s.ecp_off[t * 5 + l + 1]
                v = np.zeros_like(r[:, :, ic])
                rr = np.maximum(r[:, :, ic], 1e-10)
                for q in range(b, e):
                    p = rr ** (s.ecp_power[q] - 2) if s.ecp_power[q] != 2 else 1.0
                    v += p * s.ecp_coef[q] * np.exp(-s.ecp_expo[q] * rr * rr)
                inr[:, :, ic] |= np.abs(v) > 1e-4
        npts = int(inr.sum()) * s.nquad
        flops += float(npts) * (basis_val + 2 * NB * nocc + (N - 1) * pair + en)
    return float(flops)


def cpu_baseline(s, seconds=12.0, nthreads=None):
    """The FP64 oracle's VMC (OpenMP over walkers) on the host cores, run on
    a bounded sample: batches of 128 walkers x 2 steps until `seconds`."""
    import oracle
    from champ_amd import rng, system
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    O = oracle.Oracle(s)
    nw, nst = 128, 2
    done, t0 = 0, time.time()
    k = 0
    while time.time() - t0 < seconds:
        x = system.random_walkers(s, nw, seed=100 + k)
        seeds = rng.walker_seeds("1234567890123457", nw, stride=(1 << 30) + k)
        O.vmc_run(x, seeds, nst, 0.5, trace=False, nthreads=nthreads)
        done += nw * nst
        k += 1
    dt = time.time() - t0
    return dict(value=done / dt, unit="walker-moves/s", cores=nthreads, kind="port",
                sample=f"{k} batches of {nw} walkers x {nst} VMC steps ({done} walker-steps, {dt:.1f} s)")


def run_engine(s, W, steps, warmup, precision, delta, x, seeds):
    from champ_amd.engine import Engine
    e = Engine(s, W, precision)
    e.set_walkers(x)
    e.set_rng(seeds)
    g = e.autotune()
    print(f"[bench] {precision}: autotuned lane-group width G={g}", file=sys.stderr, flush=True)
    e.vmc_device(max(1, warmup), delta)
    e.sync()
    e.profile(True)
    t0 = time.perf_counter()
    e.vmc_device(steps, delta)
    e.sync()
    t1 = time.perf_counter()
    prof = e.profile_get()
    e.profile(False)
    return e, t1 - t0, prof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--walkers", type=int, default=16384)
    ap.add_argument("--delta", type=float, default=0.6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this run has one GPU, "
              "multi-GPU runs are not part of this build (see DESIGN.md, sharding).",
              file=sys.stderr)
        sys.exit(2)
    from champ_amd import rng, system
    s = system.c2_dimer(seed=2)
    W = args.walkers
    x0 = system.random_walkers(s, W, seed=7)
    seeds = rng.walker_seeds("1837465927472523", W)
    res = {}
    for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        e, dt, prof = run_engine(s, W, args.steps, args.warmup, prec, args.delta, x0, seeds)
        res[prec] = dict(engine=e, dt=dt, prof=prof, rate=W * args.steps / dt)
        print(f"[bench] {prec}: {args.steps} steps x {W} walkers in {dt:.3f} s "
              f"({W * args.steps / dt:.4g} walker-moves/s); phases {prof}", file=sys.stderr, flush=True)
    # exact-configuration energy bias of the mixed engine against FP64
    bias = None
    if "fp64" in res and "mixed" in res:
        xs = res["fp64"]["engine"].get_walkers()
        ef = res["fp64"]["engine"]
        em = res["mixed"]["engine"]
        ef.set_walkers(xs)
        em.set_walkers(xs)
        e64 = ef.psi_at()["eloc"]
        e32 = em.psi_at()["eloc"]
        d = e32 - e64
        bias = dict(mean_mHa=1e3 * float(d.mean()), sem_mHa=1e3 * float(d.std() / np.sqrt(len(d))),
                    max_abs_mHa=1e3 * float(np.abs(d).max()))
    main_prec = "mixed" if ("mixed" in res and bias is not None and abs(bias["mean_mHa"]) <= 0.1) \
        else ("fp64" if "fp64" in res else "mixed")
    r = res[main_prec]
    # roofline of the dominant kernel (full evaluation), HIP-event timings
    xs = r["engine"].get_walkers()
    fl = eval_full_flops(s, xs)
    ev_avg_ms = r["prof"]["eval_ms"] / max(1, r["prof"]["eval_n"])
    achieved = fl / (ev_avg_ms * 1e-3) / 1e12
    peak = PEAK_FP64_TFLOPS if main_prec == "fp64" else PEAK_FP64_TFLOPS
    out = {
        "metric": "walker-moves/s + dlnPsi/dp samples/s vs FP64 GPU path at <=0.1 mHa bias",
        "value": r["rate"],
        "unit": "walker-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["dt"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "C2-like dimer VMC: 8 valence e, 58 GTO, BFD-like ECP l=0,1 (Q=12), "
                               "24 CSF / 48 det", "walkers": W, "delta": args.delta,
                   "precision": main_prec},
        "fp64_walker_moves_per_s": res.get("fp64", {}).get("rate"),
        "mixed_walker_moves_per_s": res.get("mixed", {}).get("rate"),
        "mixed_vs_fp64": (res["mixed"]["rate"] / res["fp64"]["rate"]) if len(res) == 2 else None,
        "energy_bias_mixed_vs_fp64": bias,
        "electron_moves_per_s": r["rate"] * s.nelec,
        "phase_ms": {k: v for k, v in r["prof"].items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "k_eval_full", "avg_ms": ev_avg_ms, "flop_per_launch": fl},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.cpu_seconds)
    for v in res.values():
        v["engine"].close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
