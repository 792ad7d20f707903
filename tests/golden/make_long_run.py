"""Write long_run_64x32.npz: the oracle's trajectory summary for the
1000-step energy / enstrophy parity test.

    python tests/golden/make_long_run.py

Workload: 64^2 x 32, constant N^2, dt = 1e-3, gamma = 1e-3 (REAL(4)),
W2F = 0.01, del^8 hyperdiffusion, eddies peaked at |k| = 3 (seed 2) and
structured waves (seed 2), wave feedback on.  Stored: the six diagnostics
[KE, PE, WKE, WPE, WCE, ENS] after 0, 100, ..., 1000 steps, and psi-hat at
the bottom and top levels after 1000 steps.  The oracle takes about a minute
for the 1000 steps; the fixture lets the GPU test compare without rerunning
it.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "qg-ybjp-model_amd"), os.path.dirname(HERE)]

from oracle import qg_oracle as O  # noqa: E402
from helpers import oracle_params, physical_ics  # noqa: E402
from qgybj.params import make_params  # noqa: E402

OUT = os.path.join(HERE, "long_run_64x32.npz")
NSTEPS, EVERY = 1000, 100


def params():
    return make_params(64, 32, dt=1e-3, W2F=0.01)


def run(nsteps=NSTEPS, every=EVERY):
    P = params()
    g = O.Grid(P.n1, P.n1, P.n3)
    OP = oracle_params(P)
    psi, BR, BI = physical_ics(P, k0=3, seed=2)
    S = O.initial_state(g, OP, psi, BR, BI)
    diags = [O.diagnostics(g, OP, S)]
    for it in range(1, nsteps + 1):
        S = O.step(g, OP, S)
        if it % every == 0:
            diags.append(O.diagnostics(g, OP, S))
    return np.array(diags), S


if __name__ == "__main__":
    d, S = run()
    np.savez_compressed(OUT, steps=np.arange(0, NSTEPS + 1, EVERY), diag=d,
                        psik_bottom=S.psik[0], psik_top=S.psik[-1])
    print("wrote", OUT)
