"""Benchmark of the force-and-integration path on one MI355X.

A step is one kick-drift-kick step of PP gravity at N = 262144 (Plummer
sphere a = 1, M = 1, seeded; eps = 0.002, dt = 1e-3) in the fast FP32 policy
with the state resident in HBM: sim_step, i.e. fused kick + FP64 drift,
one O(N^2) force evaluation (20 flops per interaction), kick, per-particle
energies and FP64 reductions.  value = N^2 * steps / elapsed, in
Ginteractions/s.

  python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu-baseline]

Prints one JSON line.  Only one GPU is used; --gpus above 1 exits.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

from starflood_amd import _lib as L  # noqa: E402
from starflood_amd.sim import Simulation  # noqa: E402

METRIC = "PP gravity Ginteractions/s @N=262144; BH step ms @N=4M; adjoint/forward cost"
N_PP = 262144
EPS = 2.0e-3
DT = 1.0e-3
SEED = 11
PEAK_FP32_TFLOPS = 157.3
FLOPS_PER_INTERACTION = 20


def hip_sync():
    hip = C.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


def kernel_ms(lib, name):
    ms = C.c_double()
    cnt = C.c_ulong()
    lib.sf_prof_query(name.encode(), C.byref(ms), C.byref(cnt))
    return ms.value, cnt.value


def cpu_baseline(budget_s: float = 12.0) -> dict:
    """The reference's own OpenMP PP solver (a build of its sources in
    oracle/_ref) timed on this host, on a bounded sample: the same Plummer
    workload at a smaller N, chosen so the timed work takes ~budget_s."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    kind = "reference" if O.ref_available("pp_e2e-3_dt1e-3") else "port"
    # probe at a small N to size the sample
    def run(N, reps):
        s = Simulation(N, enable_tree=0, logging=0, epsilon=EPS, dt=DT)
        s.ic_plummer(seed=SEED, first=0, count=N)
        pos, mas = s.pos.copy(), s.mas.copy()
        s.close()
        if kind == "reference":
            r = O.RefSim("pp_e2e-3_dt1e-3", N)
            r.block(0, 3)[:] = pos.ravel()
            r.block(9)[:] = mas
            r.solve_pp()
            t0 = time.perf_counter()
            for _ in range(reps):
                r.solve_pp()
            t = (time.perf_counter() - t0) / reps
            r.close()
        else:
            O.pp_ref(pos, mas, EPS)
            t0 = time.perf_counter()
            for _ in range(reps):
                O.pp_ref(pos, mas, EPS)
            t = (time.perf_counter() - t0) / reps
        return t
    t_small = run(8192, 1)
    rate = 8192.0 ** 2 / t_small
    N = int(np.sqrt(rate * budget_s / 2.0))
    N = max(8192, min(N_PP, (N // 4096) * 4096))
    t = run(N, 2)
    return {"value": N * float(N) / t / 1e9, "unit": "Ginteractions/s", "cores": threads, "kind": kind,
            "sample": f"solve_gravity_part_part on the same Plummer workload at N={N} (2 timed evaluations, "
                      f"-fopenmp -Og -std=c99 build of the reference sources, OMP_NUM_THREADS={threads})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design "
                                   "(see DESIGN.md), not a run", "n_gpus": args.gpus}))
        sys.exit(2)

    sim = Simulation(N_PP, enable_tree=0, logging=0, verbose=0, epsilon=EPS, dt=DT,
                     precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL, mem_mode=L.SF_MEM_RESIDENT)
    lib = sim.lib
    sim.ic_plummer(seed=SEED, first=0, count=N_PP)
    sim.solve()
    for _ in range(args.warmup):
        sim.step()
    hip_sync()
    lib.sf_prof_enable(1)
    lib.sf_prof_reset()
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
    hip_sync()
    elapsed = time.perf_counter() - t0
    pp_ms, pp_n = kernel_ms(lib, "pp_fast")
    lib.sf_prof_enable(0)
    st = sim.stats()
    sim.close()

    inter = float(N_PP) * N_PP
    value = inter * args.steps / elapsed / 1e9
    pp_avg_s = pp_ms / max(pp_n, 1) / 1e3
    achieved = FLOPS_PER_INTERACTION * inter / pp_avg_s / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Plummer sphere)",
        "config": {"workload": "PP gravity KDK step, Plummer sphere N=262144 (a=1, M=1, seed 11), eps=0.002, "
                               "dt=1e-3, fast FP32 policy, state resident in HBM",
                   "N": N_PP, "epsilon": EPS, "dt": DT, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_pp_fast", "kernel_ms": pp_avg_s * 1e3,
                     "flops_per_launch": FLOPS_PER_INTERACTION * inter},
        "energy_check": {"E": st["E"], "T": st["T"], "W": st["W"]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
