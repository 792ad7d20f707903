"""Benchmark: member-steps/s of the QG-YBJ+ time step at 256^2 x 128, FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (the plain drop-in at production size): n1 = n2 = 256, n3 = 128,
FP64, exponential N^2(z) with surface/bottom ratio 100, seeded random-phase
eddies peaked at |k| = 5, surface-trapped waves, wave feedback on.  One step
is one pass of the loop body (nonlinear terms, fused update + Robert filter,
q^w feedback, psi and A inversions) on the resident state.  The Euler first
step and W leapfrog warm-up steps run untimed; then exactly K leapfrog steps
are timed between two stream synchronisations.  Prints one JSON line.

The roofline entry is for the kernel with the largest share of the timed
region, from HIP events recorded on the member's stream around every launch;
its algorithmic bytes per launch are the box-restricted reads and writes the
kernel must do (see KERNEL_BYTES).  The CPU baseline is the FP64 NumPy
restatement (oracle) timed on the host on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "qg-ybjp-model_amd"))

METRIC = "member-steps/s at 256²×128 FP64, batched vs single; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
N1, N3, K0, SEED = 256, 128, 5, 1


def kernel_bytes(n1: int, n3: int):
    """Algorithmic HBM bytes per launch of each kernel of a leapfrog step.

    G = one half-transformed intermediate restricted to the kept columns
        (KX * n2 * n3 complex), C = one spectral field restricted to the
        dealiasing box (KX * nbox * n3 complex); 16 bytes per complex."""
    import qgybj
    kmax, KX, nlo, nhi = qgybj.box(n1)
    nbox = nlo + nhi
    G = KX * n1 * n3 * 16
    C = KX * nbox * n3 * 16
    return {
        "col_inv": G + C,          # psi box in, one intermediate out
        "row_nl": 12 * G,          # 6 intermediates in, 6 out
        "col_upd": 11 * G + 14 * C,  # 6 in + 5 out; q,qo,BR,BRo,BI,BIo,AR,AI in, 6 out
        "row_qw": 6 * G,           # 4 in, 2 out
        "col_qw": 2 * G + 2 * C,   # 2 in, q in, q* out
        "solve": 2 * C,            # q* in, psi out
        "solve_a": 4 * C,          # BR, BI in; AR, AI out (side stream)
    }, {"G": G, "C": C, "KX": KX, "nbox": nbox}


def cpu_baseline(P, psi, BR, BI, nsteps: int = 2):
    """Time the oracle (NumPy/SciPy FP64 restatement) on nsteps leapfrog steps
    after an untimed Euler step."""
    sys.path.insert(0, ROOT)
    from oracle import qg_oracle as O
    workers = min(16, os.cpu_count() or 1)
    O.WORKERS = workers
    g = O.Grid(P.n1, P.n1, P.n3)
    OP = O.Params(n1=P.n1, n3=P.n3, dt=P.dt, gamma=P.gamma, Ro=P.Ro, Bu=P.Bu, W2F=P.W2F, nuz=P.nuz,
                  flags=P.flags, ifac_q=P.ifac_q, ifac_w=P.ifac_w, a_ell_ut=P.Bu / P.N2u, N2u=P.N2u, N2s=P.N2s)
    S = O.initial_state(g, OP, psi, BR, BI)
    S = O.step(g, OP, S)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        S = O.step(g, OP, S)
    dt = time.perf_counter() - t0
    return {"value": nsteps / dt, "unit": "member-steps/s", "cores": workers, "kind": "port",
            "sample": f"{nsteps} leapfrog steps of the same 256^2x128 member (NumPy/SciPy FP64 oracle, "
                      f"FFTs on {workers} threads, elementwise work single-threaded)"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=2)
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs one member per GPU on a single MI355X; "
              "multi-GPU runs (--gpus > 1) are not supported", file=sys.stderr)
        return 2

    import numpy as np
    import qgybj
    from qgybj import ic
    from qgybj.params import config_plain_production

    P = config_plain_production()
    psi = ic.eddy_psi(N1, N3, K0, SEED, Bu=P.Bu)
    BR, BI = ic.wave_b(N1, N3)
    m = qgybj.Member(P)
    m.set_state_physical(psi, BR, BI)
    m.step(1)                 # Euler first step (untimed)
    m.step(a.warmup)          # leapfrog warm-up
    m.sync()
    m.kernel_times()          # reset
    m.timing(True)
    t0 = time.perf_counter()
    m.step(a.steps)           # ends with a stream synchronisation
    t1 = time.perf_counter()
    m.timing(False)
    kt = m.kernel_times()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / a.steps
    value = a.steps / wall
    diag = m.diag()
    if not np.all(np.isfinite(diag)):
        print("bench.py: non-finite diagnostics", file=sys.stderr)
        return 1

    kb, geo = kernel_bytes(N1, N3)
    dom = max((k for k in kb), key=lambda k: kt[k][0])
    tot_ms, cnt = kt[dom]
    avg_s = tot_ms * 1e-3 / max(cnt, 1)
    achieved = kb[dom] / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": kb[dom]}
    traffic_file = os.path.join(ROOT, "profiles", f"{dom}_256x128_traffic.json")
    if os.path.exists(traffic_file):
        try:
            roofline["traffic"] = json.load(open(traffic_file)).get("bytes_per_launch")
        except Exception:
            pass
    kernels = {k: {"ms_total": round(v[0], 3), "launches": v[1]} for k, v in kt.items() if v[1]}
    step_alg = sum(kb[k] * kt[k][1] for k in kb) / a.steps
    m.close()

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(P, psi, BR, BI, a.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain drop-in step, 256^2x128 FP64, exponential N^2 (ratio 100), "
                               "eddies peaked at |k|=5 (seed 1), surface-trapped waves, feedback on",
                   "n1": N1, "n2": N1, "n3": N3, "members": 1},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "step_alg_bytes": int(step_alg),
        "step_effective_GBs": round(step_alg / (ms_per_step * 1e-3) / 1e9, 1),
        "diag": [float(x) for x in diag],
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
