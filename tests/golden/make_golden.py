"""Writes the golden fixtures in this folder from the CPU oracle.

    python tests/golden/make_golden.py

weno_faces.npz   512 seeded stencils (f+ at i-2..i+2, f- at i-1..i+3) and the
                 face flux of each for WENO5-Z p=1 eps=1e-6 (reference
                 defaults), WENO5-Z p=2 eps=1e-40 and WENO5-JS eps=1e-6.
rhs_periodic16.npz  16^3 periodic Fourier field (seed 2) and its global-LF
                 WENO5-Z p=2 right-hand side, plus the state after 3 SSP-RK3
                 steps (CFL 0.4).
heat64.npz       64^3 7-point forward-Euler state after 100 steps (r = 0.15,
                 Gaussian sigma 0.1 + 1e-3 noise, seed 1): SHA-256 of the
                 bytes and the centre plane.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from multigpu_advectiondiffusion_amd import solvers  # noqa: E402

SCHEMES = {"z1": dict(weno=O.WENO_Z, p=1, eps=1e-6), "z2": dict(weno=O.WENO_Z, p=2, eps=1e-40),
           "js": dict(weno=O.WENO_JS, p=2, eps=1e-6)}


def main():
    rng = np.random.default_rng(20261202)
    scale = rng.choice([1e-3, 1.0, 10.0], size=(512, 1))
    v = rng.uniform(-1, 1, (512, 5)) * scale
    w = rng.uniform(-1, 1, (512, 5)) * scale
    out = {"v": v, "w": w}
    for k, kw in SCHEMES.items():
        out[k] = np.array([O.face(v[i], w[i], **kw) for i in range(512)])
    np.savez(os.path.join(HERE, "weno_faces.npz"), **out)

    u = solvers.ic_fourier(16, seed=2)
    Lu = O.rhs_periodic(u, 2.0 / 16, **O.Z2_GLOBAL, alpha=O.maxabs(u))
    u3 = u.copy()
    O.burgers_run(u3, 2.0 / 16, 0.4, 0.0, 10.0, nsteps=3, **O.Z2_GLOBAL)
    np.savez(os.path.join(HERE, "rhs_periodic16.npz"), u=u, Lu=Lu, u3=u3)

    h = solvers.heat7_params(64)
    v0 = solvers.ic_gauss_noise(64, sigma=0.1, noise=1e-3, seed=1)
    O.heat7_fe(v0, h["kx"], h["ky"], h["kz"], h["dt"], 100)
    np.savez(os.path.join(HERE, "heat64.npz"), sha256=np.frombuffer(hashlib.sha256(v0.tobytes()).digest(), np.uint8),
             plane=v0[32])


if __name__ == "__main__":
    main()
