#!/usr/bin/env python3
"""Writes tests/golden/sgs.npz by running the reference's own Smagorinsky
operators (src/turbulence/base/eddy_viscosity.cpp, built by `make -C oracle
ref` into oracle/_ref/ref_sgs) on seeded FP32 states: for each case the
inputs (u, v, w, rho, theta, qv, rho0), the settings, and the reference's
nu, K_theta, K_q and the u, v, w, theta, qv tendencies. Run in the development
container:

    make -C oracle ref && python tests/golden/make_sgs_golden.py
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_sgs")
NX, NY, NZ, DX, DY, DZ = 10, 9, 8, 1000.0, 900.0, 500.0
OUT = ("nu", "kth", "kq", "du", "dv", "dw", "dth", "dqv")
# name: mode (0 3d, 1 horizontal_only, 2 vertical_only), filter (0 cubic_root, 1 dx),
#       Cs, Pr_t, Sc_t, nu_max, K_max, periodic x, periodic y, open (w, e, s, n)
CASES = {
    "3d_closed": (0, 0, 0.18, 1 / 3, 1 / 3, 1000.0, 1000.0, 0, 0, (0, 0, 0, 0)),
    "horizontal_open": (1, 0, 0.18, 1 / 3, 1 / 3, 1000.0, 1000.0, 0, 0, (1, 1, 1, 1)),
    "vertical_dx_closed": (2, 1, 0.2, 0.7, 0.5, 1000.0, 1000.0, 0, 0, (0, 0, 0, 0)),
    "3d_periodic": (0, 0, 0.18, 1 / 3, 1 / 3, 1000.0, 1000.0, 1, 1, (0, 0, 0, 0)),
    "horizontal_dx_periodic_x_open_y": (1, 1, 0.25, 1 / 3, 1 / 3, 1000.0, 1000.0, 1, 0, (0, 0, 1, 1)),
    "3d_capped": (0, 0, 0.25, 1 / 3, 0.4, 40.0, 60.0, 0, 0, (1, 0, 0, 1)),
}


def state(seed):
    rng = np.random.default_rng(seed)
    z = (np.arange(NZ) + 0.5) * DZ
    rho0 = 1.2 * np.exp(-z / 9000.0)
    x = np.arange(NX)[:, None, None] * DX
    y = np.arange(NY)[None, :, None] * DY
    zz = z[None, None, :]
    shape = (NX, NY, NZ)
    u = 8.0 * np.sin(2e-4 * x + 3e-4 * y) + 0.004 * zz + 3.0 * rng.standard_normal(shape)
    v = -4.0 * np.cos(3e-4 * x - 1e-4 * y) + 2.0 * rng.standard_normal(shape)
    w = 2.0 * rng.standard_normal(shape) * np.sin(np.pi * zz / (NZ * DZ))
    rho = rho0[None, None, :] * (1.0 + 2e-3 * rng.standard_normal(shape))
    theta = 300.0 + 0.004 * zz + 0.8 * rng.standard_normal(shape)
    qv = np.clip(0.012 * np.exp(-zz / 3000.0) * (1.0 + 0.2 * rng.standard_normal(shape)), 0.0, None)
    f = {k: np.ascontiguousarray(a, dtype=np.float32) for k, a in
         (("u", u), ("v", v), ("w", w), ("rho", rho), ("theta", theta), ("qv", qv))}
    f["rho0"] = rho0
    return f


def main():
    out = {}
    for q, (name, c) in enumerate(CASES.items()):
        st = state(100 + q)
        mode, filt, cs, prt, sct, nu_max, k_max, px, py, op = c
        with tempfile.TemporaryDirectory() as d:
            for k in ("u", "v", "w", "rho", "theta", "qv"):
                st[k].tofile(os.path.join(d, k + ".f32"))
            st["rho0"].astype(np.float64).tofile(os.path.join(d, "rho0.f64"))
            with open(os.path.join(d, "params.txt"), "w") as f:
                f.write(" ".join(repr(x) for x in (NX, NY, NZ, DX, DY, DZ, mode, filt, cs, prt, sct, nu_max, k_max,
                                                   px, py) + tuple(op)))
            subprocess.run([HARNESS, d], check=True)
            res = {k: np.fromfile(os.path.join(d, k + ".f32"), dtype=np.float32).reshape(NX, NY, NZ) for k in OUT}
        for k, a in st.items():
            out[f"{name}/in/{k}"] = a
        for k, a in res.items():
            out[f"{name}/out/{k}"] = a
        out[f"{name}/settings"] = np.array([mode, filt, cs, prt, sct, nu_max, k_max, px, py] + list(op), np.float64)
    out["grid"] = np.array([NX, NY, NZ, DX, DY, DZ])
    np.savez_compressed(os.path.join(HERE, "sgs.npz"), **out)
    print("wrote", os.path.join(HERE, "sgs.npz"))


if __name__ == "__main__":
    main()
