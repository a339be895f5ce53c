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

The timed run records no events.  A second, shorter run of the same engine
records HIP events around every kernel launch on the engine's stream
(qmcw_profile) and gives the per-kernel averages the roofline entry uses.

Prints one JSON line.
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
PEAK_HBM_GBS = 8000.0       # MI355X HBM3E
TRAFFIC_FILE = os.path.join(HERE, "profiles", "hbm_traffic.json")


# ---------------------------------------------------------------- FLOP models
def _det_sizes(s, sp):
    """(N_sigma, N_occ, [excitation order of each unique determinant])."""
    dets = np.asarray(s.det_up if sp == 0 else s.det_dn)
    if dets.shape[1] == 0:
        return 0, 0, []
    ref = set(dets[0].tolist())
    occ = set(dets.ravel().tolist())
    uniq = {tuple(sorted(d.tolist())) for d in dets}
    return dets.shape[1], len(occ), [len(set(u) - ref) for u in uniq]


def _unit_costs(s):
    """Algorithmic FP operations of the building blocks (per evaluation):
      e-e pair (Pade + polynomial): value+grad+lap 12 + 6 nordb, value 6 + 2 nordb
      e-e-n per (pair, centre): value+grad+lap 12 nterms + 60, value 8 nterms + 20
      e-n per (electron, centre): value+grad+lap 12 + 6 norda, value 6 + 2 norda
      basis per point and shell: value+grad+lap 30 + 12 nprim + 10 (2l+1),
                                 value 8 + 8 nprim + 3 (2l+1)."""
    from champ_amd.system import nterms4
    nt, M = nterms4(s.nordc), int(s.ncent)
    shells = []
    for ic in range(M):
        t = s.iwctype[ic]
        seen = {}
        for q in range(s.ctype_bas_off[t], s.ctype_bas_off[t + 1]):
            key = (s.bas_rad[q], s.bas_l[q])
            seen[key] = seen.get(key, 0) + 1
        for (r, l), nm in seen.items():
            shells.append((int(s.rad_prim_off[r + 1] - s.rad_prim_off[r]), nm))
    return dict(
        pair_full=12 + 6 * s.nordb + M * (12 * nt + 60),
        pair_val=6 + 2 * s.nordb + M * (8 * nt + 20),
        en_full=M * (12 + 6 * s.norda),
        en_val=M * (6 + 2 * s.norda),
        basis_full=sum(30 + 12 * npr + 10 * nm for npr, nm in shells),
        basis_val=sum(8 + 8 * npr + 3 * nm for npr, nm in shells),
    )


def _nlpp_points(s, x):
    """Quadrature points evaluated by the pseudopotential: electrons within
    range of a nonlocal channel (|v_l| > 1e-4) times nquad."""
    M = int(s.ncent)
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
    return inr.sum(axis=2), int(s.nquad)     # [W, N] centres in range, Q


def kernel_flops(s, name, x):
    """Algorithmic FP operations of one launch of kernel `name` over the W
    walkers in x (the per-unit figures of the scope table, times the units of
    one launch).  k_ratio_1e is averaged over the electrons of a sweep."""
    N, M, NB = int(s.nelec), int(s.ncent), int(s.nbasis)
    W = x.shape[0]
    u = _unit_costs(s)
    sp = [_det_sizes(s, 0), _det_sizes(s, 1)]
    npair = N * (N - 1) // 2
    if name == "k_full_jas":
        return float(W) * (npair * u["pair_full"] + N * u["en_full"])
    if name == "k_full_det":
        f = N * (u["basis_full"] + 2 * NB * 5 * max(sp[0][1], sp[1][1]))
        for ns, nocc, ks in sp:
            f += 2 * ns ** 3 + 2 * ns * ns * nocc + sum(k ** 3 for k in ks) + 4 * ns * nocc * ns
        return float(W) * f
    if name == "k_full_nlpp":
        cnt, Q = _nlpp_points(s, x)
        f = 0.0
        for i in range(N):
            nocc = sp[0][1] if i < s.nup else sp[1][1]
            pts = float(cnt[:, i].sum()) * Q
            f += pts * (u["basis_val"] + 2 * NB * nocc + (N - 1) * u["pair_val"] + u["en_val"])
        return f
    if name == "k_ratio_1e":
        f = 0.0
        for i in range(N):
            ns, nocc, ks = sp[0] if i < s.nup else sp[1]
            f += (u["basis_full"] + 2 * NB * 5 * nocc
                  + (N - 1) * u["pair_full"] + u["en_full"]
                  + 2 * ns * ns + 2 * ns * nocc + sum(k ** 3 for k in ks))
        return float(W) * f / N
    return None


def kernel_bytes(s, name, x):
    """Algorithmic HBM bytes of one launch of the one-electron move over the W
    walkers in x, with the per-unit figures of the scope table's measurement
    plan (FP64 state): Sherman-Morrison W (2 N_s^2 8 + N_occ 5 8) -- read and
    write A^-1, read the new orbital row -- and one-electron Jastrow
    W (N 3 8 + N 3 8 2) -- positions, the old and new gradient rows --,
    averaged over the electrons of a sweep.  None for the other kernels."""
    if name != "k_ratio_1e":
        return None
    N = int(s.nelec)
    W = x.shape[0]
    sp = [_det_sizes(s, 0), _det_sizes(s, 1)]
    b = 0.0
    for i in range(N):
        ns, nocc, _ = sp[0] if i < s.nup else sp[1]
        b += 2 * ns * ns * 8 + nocc * 5 * 8 + N * 3 * 8 + N * 3 * 8 * 2
    return float(W) * b / N


# ---------------------------------------------------------------- CPU baseline
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


# ---------------------------------------------------------------- GPU runs
def run_engine(s, W, steps, warmup, precision, delta, x, seeds, prof_steps):
    from champ_amd.engine import Engine
    e = Engine(s, W, precision)
    e.set_walkers(x)
    e.set_rng(seeds)
    e.autotune()
    g = e.group_widths()
    print(f"[bench] {precision}: autotuned lane-group widths (jastrow, det, nlpp, moves) = {g}",
          file=sys.stderr, flush=True)
    e.vmc_device(max(1, warmup), delta)
    e.sync()
    t0 = time.perf_counter()
    e.vmc_device(steps, delta)
    e.sync()
    t1 = time.perf_counter()
    # per-kernel HIP-event timing in a separate run (events add host work)
    e.profile(True)
    e.vmc_device(prof_steps, delta)
    e.sync()
    prof = e.profile_get()
    e.profile(False)
    return e, t1 - t0, prof, g


def run_derivs(W, steps, warmup, precision, delta, prof_steps):
    """Derivative samples: VMC steps whose full evaluation also forms
    d ln Psi/dp and dE_L/dp (the optimisation sample) and adds them to the
    linear method's product sums, C2-like system with its parameter set
    (CSF, Jastrow, orbital rotations)."""
    from champ_amd import rng, system
    from champ_amd.engine import Engine
    s = system.c2_dimer(seed=2, derivs=True)
    e = Engine(s, W, precision)
    e.set_walkers(system.random_walkers(s, W, seed=8))
    e.set_rng(rng.walker_seeds("1837465927472523", W, stride=1 << 20))
    g = e.autotune()
    e.accum_reset()
    e.sample_derivs(2)          # derivatives and the linear method's product sums every step
    e.vmc_device(max(1, warmup), delta)
    e.sync()
    t0 = time.perf_counter()
    e.vmc_device(steps, delta)
    e.sync()
    dt = time.perf_counter() - t0
    e.profile(True)
    e.vmc_device(prof_steps, delta)
    e.sync()
    prof = e.profile_get()
    e.profile(False)
    return dict(engine=e, rate=W * steps / dt, dt=dt, prof=prof, G=g, P=e.P, system=s)


def roofline_entry(dom, avg_ms, nl, fl, achieved_fl, by, traffic):
    """The dominant kernel against its bound: the one-electron move is HBM-bound
    in the scope table's measurement plan for small systems (algorithmic bytes
    per launch over the launch time, against 8 TB/s, with the measured HBM
    traffic per launch beside it); a kernel without a byte model is measured
    in FP64 FLOPs against the FP64 peak (the MI355X's vector and matrix FP64
    peaks are equal)."""
    common = {"kernel": dom, "avg_us": 1e3 * avg_ms, "launches_timed": nl, "flop_per_launch": fl,
              "achieved_fp64_tflops": achieved_fl, "frac_fp64_peak": achieved_fl / PEAK_FP64_TFLOPS}
    if by is not None:
        ach = by / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": traffic, "bytes_per_launch": by, **common}
    return {"bound": "mfma", "achieved": achieved_fl, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved_fl / PEAK_FP64_TFLOPS, "traffic": traffic, **common}


def accum_roofline(dres, W):
    """FP64 MFMA rate of the product sums (k_accum_*: three [W x P]^T [W x P]
    products, 2 W P^2 flops each, plus O(W P) vector sums) in the derivative
    run's profile."""
    if not dres:
        return None
    d = next(iter(dres.values()))
    ms, n = d["prof"].get("k_accum", (0.0, 0))
    if not n:
        return None
    P = d["P"]
    fl = 3 * 2.0 * W * P * P
    ach = fl / (ms / n * 1e-3) / 1e12
    return {"kernel": "k_accum (gemm, vector sums, chunk reduction)", "avg_us": 1e3 * ms / n,
            "flop_per_launch": fl, "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_FP64_TFLOPS}


def run_batch(nsys, steps, warmup, precision, delta, seed0=1000, derivs_half=True,
              deriv_bytes=2 << 30):
    """Multi-system batch: nsys synthetic molecules (4-128 electrons, 1-24
    nuclei, 10-800 GTOs), 256-4096 walkers each, one handle (and HIP stream)
    per system.  With derivs_half half of the systems also form the linear
    method's derivatives and product sums each step (its Jastrow set and
    orbital rotations); such a system's walker count is capped so that its
    derivative scratch (the per-pair basis rows of the nonlocal orbital
    derivative, N M NB doubles per walker, and the orbital values and
    derivatives, 5 N NB) stays within deriv_bytes, never below 256.  Each
    system is timed alone, then all together: every handle's steps are
    queued before any is awaited, so the systems run concurrently on their
    own streams.  The one-launch batch runs plain VMC steps, so it is timed
    with the derivatives switched off."""
    from champ_amd import rng, system
    from champ_amd.engine import Engine
    rs = np.random.default_rng(seed0)
    engines = []
    for k in range(nsys):
        s = system.random_molecule(seed0 + k)
        W = int(rs.integers(256, 4097))
        # half of the systems; consecutive handles' streams share the GPU's
        # few hardware queues in turn, so the derivative systems alternate
        # with the others along every queue
        der = derivs_half and (k + k // 4) % 2 == 1
        if der:
            system.molecule_derivs(s)
            per = 8 * (s.nelec * s.ncent * s.nbasis + 5 * s.nelec * s.nbasis)
            W = min(W, max(256, deriv_bytes // per))
        e = Engine(s, W, precision)
        e.set_walkers(system.random_walkers(s, W, seed=seed0 + k))
        e.set_rng(rng.walker_seeds("1837465927472523", W, stride=(1 << 20) + k))
        g = e.autotune()
        if der:
            e.accum_reset()
            e.sample_derivs(2)
        engines.append((s, W, e, g, der))
    per = []
    for s, W, e, g, der in engines:
        e.vmc_device(max(1, warmup), delta)
        e.sync()
        t0 = time.perf_counter()
        e.vmc_device(steps, delta)
        e.sync()
        dt = time.perf_counter() - t0
        per.append(dict(nelec=int(s.nelec), ncent=int(s.ncent), nbasis=int(s.nbasis), walkers=W,
                        G=g, derivs=der, nparm=int(s.nparm) if der else 0,
                        walker_moves_per_s=W * steps / dt))
    t0 = time.perf_counter()
    for s, W, e, g, der in engines:
        e.vmc_device(steps, delta)
    for s, W, e, g, der in engines:
        e.sync()
    dt = time.perf_counter() - t0
    agg = sum(W for _, W, _, _, _ in engines) * steps / dt
    # the same systems as one batch: every kernel of the step once over the
    # walkers of all systems (each system keeps its own lane-group widths)
    from champ_amd.engine import Batch
    for _, _, e, _, der in engines:
        if der:
            e.sample_derivs(0)
    b = Batch([e for _, _, e, _, _ in engines])
    b.vmc_device(max(1, warmup), delta)
    for _, _, e, _, _ in engines:
        e.sync()
    t0 = time.perf_counter()
    b.vmc_device(steps, delta)
    for _, _, e, _, _ in engines:
        e.sync()
    dt1 = time.perf_counter() - t0
    agg1 = sum(W for _, W, _, _, _ in engines) * steps / dt1
    b.close()
    for _, _, e, _, _ in engines:
        e.close()
    return dict(aggregate=agg, per_system=per, seconds=dt, aggregate_one_launch=agg1,
                seconds_one_launch=dt1, derivs_half=derivs_half)


def cpu_baseline_dmc(s, seconds=10.0, nthreads=None, tau=0.01):
    """The FP64 oracle's DMC generations (OpenMP over walkers) on the host
    cores for the chain: batches of 16 walkers, one generation each, until
    `seconds`."""
    import oracle
    from champ_amd import rng, system
    from champ_amd.abi import DmcParams
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    nw, done, k = 16, 0, 0
    t0 = time.time()
    while time.time() - t0 < seconds:
        x = system.random_walkers(s, nw, seed=200 + k)
        d = oracle.OracleDMC(s, 2 * nw)
        d.set(x, np.asarray(rng.walker_seeds("1234567890123457", nw, stride=(1 << 30) + k), np.int32))
        d.generation(DmcParams(tau, -250.0, eest=-250.0, esigma=1.0), nthreads=nthreads)
        d.close()
        done += nw
        k += 1
    dt = time.time() - t0
    return dict(value=done / dt, unit="walker-moves/s", cores=nthreads, kind="port",
                sample=f"{k} batches of {nw} walkers x 1 DMC generation ({done} walker-moves, {dt:.1f} s)")


def run_dmc_chain(W0, steps, warmup, precision, tau=0.01, equil=20, tmoves=0, reinv=64,
                  keep_engine=False):
    """DMC of the 98-electron chain (16 Z_eff=4 centres with a
    pseudopotential, 34 H, 424 GTOs, one determinant): W0 walkers, capacity
    2 W0, time step tau, under the rundmc driver (population control),
    started, as a DMC run starts from VMC configurations, from `equil` VMC
    sweeps of the random initial walkers (step 0.3 bohr).  The `warmup`
    generations are an equilibration run: after them etrial is reset to
    eest and the averages restart (dmc_mod.f90:386-403 with
    l_reset_etrial).  The synthetic wavefunction (random orthonormal MOs)
    is far from the ground state, so its DMC energy keeps relaxing and the
    population drifts over hundreds of generations; the population range
    of the timed generations is reported."""
    from champ_amd import rng, system
    from champ_amd.dmc import DmcDriver
    from champ_amd.engine import Engine
    s = system.hchain(4)
    cap = 2 * W0
    e = Engine(s, cap, precision)
    e.set_walkers(system.random_walkers(s, cap, seed=9))
    e.set_rng(rng.walker_seeds("1837465927472523", cap, stride=1 << 20))
    e.set_branch_rng(np.array([1, 2, 3, 5], np.int32))
    e.set_reinversion(reinv)
    e.autotune()
    g = [int(v) for v in e.group_widths()]
    if equil:
        e.vmc_device(equil, 0.3)
    e.dmc_begin(W0)
    e0 = float(e.get_eloc()[:W0].mean())
    drv = DmcDriver(e, s.nelec, W0, tau=tau, etrial=e0, tmoves=tmoves)
    drv.run(max(1, warmup))
    drv.end_equilibration(reset_etrial=True)
    e.sync()
    t0 = time.perf_counter()
    h = drv.run(steps)
    e.sync()
    dt = time.perf_counter() - t0
    moves = sum(x["nwalk"] for x in h)
    # kernel times of one more generation (outside the timed region)
    e.profile(True)
    drv.run(1)
    e.sync()
    prof = e.profile_get()
    e.profile(False)
    xs = e.get_walkers()[:h[-1]["nwalk"]]
    if not keep_engine:
        e.close()
        e = None
    nw = [x["nwalk"] for x in h]
    return dict(rate=moves / dt, G=g, nwalk_final=h[-1]["nwalk"], eave=h[-1]["eave"],
                nwalk_range=[min(nw), round(float(np.mean(nw)), 1), max(nw)],
                acc=float(np.mean([x["acc"] for x in h])), seconds=dt, prof=prof, x=xs, system=s,
                engine=e)


def copy_calibration(nbytes=256 << 20):
    """Three device-to-device copies of nbytes (HIP's blit copy kernel), the
    known-size reference for scaling FETCH_SIZE / WRITE_SIZE counters
    (launched only with --copy-calibration)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(nbytes)
    assert hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    assert hip.hipMemset(a, 0, n) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, n, 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy calibration: 3 copies of {nbytes} bytes", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 200 VMC steps; 20 DMC generations with --dmc)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--walkers", type=int, default=16384)
    ap.add_argument("--delta", type=float, default=0.6)
    ap.add_argument("--prof-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--deriv-walkers", type=int, default=8192)
    ap.add_argument("--deriv-steps", type=int, default=40)
    ap.add_argument("--no-derivs", action="store_true", help="skip the derivative samples/s measurement")
    ap.add_argument("--dmc", action="store_true",
                    help="only run DMC of the 98-electron chain (4096 walkers, tau 0.01)")
    ap.add_argument("--tmoves", type=int, default=0, choices=[0, 1, 2],
                    help="with --dmc: T-moves (1 det_balance1, 2 casula1)")
    ap.add_argument("--reinversion", type=int, default=64,
                    help="with --dmc: rebuild the Slater inverses every this many moves (0: per sweep)")
    ap.add_argument("--batch", type=int, default=0,
                    help="only run the multi-system batch with this many systems")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs and reports a single GPU; "
              "multi-GPU runs (walker sharding) are not part of it.", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 20 if args.dmc else 200
    if args.dmc:
        prec = "mixed" if args.precision == "both" else args.precision
        precs = [prec] + ([q for q in ("fp64", "mixed_bf16") if q != prec] if args.precision == "both" else [])
        runs = {}
        for q in precs:
            runs[q] = run_dmc_chain(4096, args.steps, args.warmup, q, tmoves=args.tmoves,
                                    reinv=args.reinversion, keep_engine=(q == prec))
            print(f"[bench] DMC chain {q}: {runs[q]['rate']:.4g} walker-moves/s "
                  f"(population {runs[q]['nwalk_range']})", file=sys.stderr, flush=True)
        d = runs[prec]
        for k, (ms, n) in d["prof"].items():
            if n:
                print(f"[bench]   {k:12s} {n:5d} launches  avg {ms / n * 1e3:9.1f} us  "
                      f"total {ms:9.3f} ms", file=sys.stderr, flush=True)
        s_ch, xs = d["system"], d["x"]
        # energies of the final population at identical configurations
        bias = {}
        if len(precs) > 1:
            from champ_amd.engine import Engine
            el = {}
            for q in precs:
                eq = Engine(s_ch, len(xs), q)
                eq.set_walkers(xs)
                el[q] = eq.psi_at()["eloc"]
                eq.close()
            for q in precs:
                if q != "fp64":
                    dd = el[q] - el["fp64"]
                    bias[q] = dict(mean_mHa=1e3 * float(dd.mean()),
                                   sem_mHa=1e3 * float(dd.std() / np.sqrt(len(dd))),
                                   max_abs_mHa=1e3 * float(np.abs(dd).max()))
        if d["engine"] is not None:
            d["engine"].close()
        dom = max(d["prof"], key=lambda k: d["prof"][k][0])
        ms_tot, nl = d["prof"][dom]
        avg_ms = ms_tot / max(1, nl)
        npop = d["nwalk_final"]
        fl = kernel_flops(s_ch, dom, xs)
        ach = fl / (avg_ms * 1e-3) / 1e12 if fl else None
        roof = ({"bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": ach / PEAK_FP64_TFLOPS, "traffic": None, "kernel": dom,
                 "avg_us": 1e3 * avg_ms, "launches_timed": nl, "flop_per_launch": fl,
                 "walkers_per_launch": npop} if fl else None)
        out = {"metric": "DMC walker-moves/s (98-electron chain)", "value": d["rate"],
               "unit": "walker-moves/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f64" if prec == "fp64" else "f32+f64", "data": "synthetic",
               "config": {"workload": f"DMC, 98-electron chain (16 Z_eff=4 + 34 H, 424 GTOs), "
                                      f"4096 walkers (target), tau 0.01, {args.steps} timed generations "
                                      f"from 20 VMC sweeps and {args.warmup} equilibration generations (etrial reset), "
                                      f"re-inversion every {args.reinversion} moves"
                                      + (", T-moves" if args.tmoves else ""),
                          "tmoves": args.tmoves, "reinversion_every": args.reinversion,
                          "precision": prec, "group_widths": d["G"]},
               "walker_moves_per_s": {q: r["rate"] for q, r in runs.items()},
               "energy_bias_vs_fp64": bias or None,
               "nwalk_final": d["nwalk_final"], "nwalk_min_mean_max": d["nwalk_range"],
               "acceptance": d["acc"], "seconds": d["seconds"],
               "kernel_avg_us": {k: (ms / n * 1e3 if n else None) for k, (ms, n) in d["prof"].items()},
               "roofline": roof}
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline_dmc(s_ch, args.cpu_seconds)
        print(json.dumps(out), flush=True)
        return
    if args.batch:
        prec = "mixed" if args.precision == "both" else args.precision
        b = run_batch(args.batch, args.steps, args.warmup, prec, args.delta,
                      derivs_half=not args.no_derivs)
        print(json.dumps({"metric": "walker-moves/s (multi-system batch)",
                          "value": b["aggregate"],
                          "unit": "walker-moves/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "higher_is_better": True, "dtype":
                          "f64" if prec == "fp64" else "f32+f64", "data": "synthetic",
                          "config": {"workload": f"{args.batch} synthetic molecules",
                                     "precision": prec,
                                     "derivatives": "half of the systems" if b["derivs_half"]
                                     else "none"},
                          "streams_walker_moves_per_s": b["aggregate"],
                          "one_launch_walker_moves_per_s": b["aggregate_one_launch"],
                          "one_launch_derivatives": "none",
                          "per_system": b["per_system"]}), flush=True)
        return
    from champ_amd import rng, system
    s = system.c2_dimer(seed=2)
    W = args.walkers
    x0 = system.random_walkers(s, W, seed=7)
    seeds = rng.walker_seeds("1837465927472523", W)
    res = {}
    for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        e, dt, prof, g = run_engine(s, W, args.steps, args.warmup, prec, args.delta, x0, seeds,
                                    args.prof_steps)
        res[prec] = dict(engine=e, dt=dt, prof=prof, rate=W * args.steps / dt, G=g)
        print(f"[bench] {prec}: {args.steps} steps x {W} walkers in {dt:.3f} s "
              f"({W * args.steps / dt:.4g} walker-moves/s)", file=sys.stderr, flush=True)
        for k, (ms, n) in prof.items():
            if n:
                print(f"[bench]   {k:12s} {n:5d} launches  avg {ms / n * 1e3:9.1f} us  "
                      f"total {ms:9.3f} ms", file=sys.stderr, flush=True)
    # derivative samples/s (the optimisation sample)
    dres = {}
    if not args.no_derivs:
        for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
            dres[prec] = run_derivs(args.deriv_walkers, args.deriv_steps, args.warmup, prec,
                                    args.delta, max(2, args.prof_steps // 2))
            print(f"[bench] {prec} derivatives: {dres[prec]['rate']:.4g} samples/s "
                  f"(P={dres[prec]['P']}, G={dres[prec]['G']})", file=sys.stderr, flush=True)
    # exact-configuration energy bias of the mixed engine against FP64
    bias = bias_bf16 = None
    if "fp64" in res and "mixed" in res:
        xs = res["fp64"]["engine"].get_walkers()
        ef, em = res["fp64"]["engine"], res["mixed"]["engine"]
        ef.set_walkers(xs)
        em.set_walkers(xs)
        d = em.psi_at()["eloc"] - ef.psi_at()["eloc"]
        bias = dict(mean_mHa=1e3 * float(d.mean()), sem_mHa=1e3 * float(d.std() / np.sqrt(len(d))),
                    max_abs_mHa=1e3 * float(np.abs(d).max()))
        # bfloat16 inputs to the orbital contraction, same configurations
        from champ_amd.engine import Engine
        eb = Engine(s, W, "mixed_bf16")
        eb.set_walkers(xs)
        db = eb.psi_at()["eloc"] - ef.psi_at()["eloc"]
        eb.close()
        bias_bf16 = dict(mean_mHa=1e3 * float(db.mean()),
                         sem_mHa=1e3 * float(db.std() / np.sqrt(len(db))),
                         max_abs_mHa=1e3 * float(np.abs(db).max()))
    main_prec = "mixed" if ("mixed" in res and bias is not None and abs(bias["mean_mHa"]) <= 0.1) \
        else ("fp64" if "fp64" in res else "mixed")
    r = res[main_prec]
    # roofline of the dominant kernel (largest total time in the profiled run)
    xs = r["engine"].get_walkers()
    dom = max(r["prof"], key=lambda k: r["prof"][k][0])
    ms_tot, nl = r["prof"][dom]
    avg_ms = ms_tot / max(1, nl)
    fl = kernel_flops(s, dom, xs)
    by = kernel_bytes(s, dom, xs)
    achieved_fl = fl / (avg_ms * 1e-3) / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        tf = json.load(open(TRAFFIC_FILE))
        ent = tf.get("kernels", {}).get(dom)
        # the measurement holds for the same workload, precision and lane-group
        # width of the kernel's own class (jastrow, det, nlpp, moves)
        cls = {"k_full_jas": 0, "k_full_det": 1, "k_full_nlpp": 2, "k_ratio_1e": 3}.get(dom)
        if ent and cls is not None and ent.get("walkers") == W and ent.get("precision") == main_prec \
                and ent.get("group_widths", [None] * 4)[cls] == list(r["G"])[cls]:
            traffic = ent["hbm_bytes_per_launch"]
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
        "config": {"workload": "C2-like dimer VMC: 8 valence e, 58 GTO, ECP l=0,1 nonlocal (Q=12), "
                               f"24 CSF / 48 det, {W} walkers, {args.steps} timed VMC steps",
                   "walkers": W, "delta": args.delta, "precision": main_prec,
                   "group_widths": r["G"]},
        "fp64_walker_moves_per_s": res.get("fp64", {}).get("rate"),
        "mixed_walker_moves_per_s": res.get("mixed", {}).get("rate"),
        "mixed_vs_fp64": (res["mixed"]["rate"] / res["fp64"]["rate"]) if len(res) == 2 else None,
        "energy_bias_mixed_vs_fp64": bias,
        "energy_bias_bf16_inputs_vs_fp64": bias_bf16,
        "electron_moves_per_s": r["rate"] * s.nelec,
        "kernel_avg_us": {k: (1e3 * v[0] / v[1] if v[1] else None) for k, v in r["prof"].items()},
        "deriv_samples_per_s": {k: v["rate"] for k, v in dres.items()} or None,
        "deriv_config": ({"workload": "C2-like dimer with CSF, Jastrow and orbital-rotation "
                                      "parameters, VMC steps with derivatives and the linear "
                                      "method's product sums",
                          "walkers": args.deriv_walkers, "nparm": next(iter(dres.values()))["P"]}
                         if dres else None),
        "deriv_mixed_vs_fp64": (dres["mixed"]["rate"] / dres["fp64"]["rate"]) if len(dres) == 2 else None,
        "deriv_kernel_avg_us": ({k: (1e3 * v[0] / v[1] if v[1] else None)
                                 for k, v in next(iter(dres.values()))["prof"].items()} if dres else None),
        "accum_mfma": accum_roofline(dres, args.deriv_walkers),
        "roofline": roofline_entry(dom, avg_ms, nl, fl, achieved_fl, by, traffic),
    }
    if args.copy_calibration:
        copy_calibration()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.cpu_seconds)
    for v in list(res.values()) + list(dres.values()):
        v["engine"].close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
