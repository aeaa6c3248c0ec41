"""Write the golden fixtures of the reference's CPU program.

Run in a container that holds the reference tree (it is not needed
afterwards): `make -C oracle` builds oracle/_ref/ from the reference's own
sources, then this script runs them and stores inputs and outputs only.

  cpu3d_node_vectors.npz  bgk_collide / mrt_collide (CPU_3D/CPU_3D.cpp:116,
                          :164) and DirichletWest_Regularized (:974) on seeded
                          inputs, as the reference computes them
  cpu3d_t100.npz          the reference program CPU_3D/CPU_3D.cpp run as is
                          (251x43x43, 100 steps): its output.dat reduced to
                          three z-slices of u, v, rho and per-plane sums
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def node_vectors(n=256, seed=1234):
    r = O.ref()
    rng = np.random.default_rng(seed)
    dims = np.zeros(4, np.int32)
    consts = np.zeros(4, np.float32)
    r.ref_case(dims, consts)
    fin = np.empty((n, 19), np.float32)
    omega = rng.uniform(0.55, 1.95, n).astype(np.float32)
    for k in range(n):
        u = rng.uniform(-0.08, 0.08, 3)
        fin[k] = O.feq_bgk(float(rng.uniform(0.97, 1.03)), *u) + rng.uniform(-2e-3, 2e-3, 19).astype(np.float32)
    out = {"fin": fin, "omega": omega, "umax": consts[0:1], "omega_case": consts[3:4]}
    for name, fn in (("bgk", r.ref_bgk_collide), ("mrt", r.ref_mrt_collide)):
        res = fin.copy()
        for k in range(n):
            fn(res[k], omega[k])
        out[name] = res
    lid = fin.copy()
    for k in range(n):
        r.ref_boundaries(lid[k], 5, 7, 53)
    out["lid"] = lid
    np.savez_compressed(os.path.join(HERE, "cpu3d_node_vectors.npz"), **out)


def full_run():
    dims = np.zeros(4, np.int32)
    consts = np.zeros(4, np.float32)
    O.ref().ref_case(dims, consts)
    nx, ny, nz, tmax = (int(d) for d in dims)
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([O.REF_EXE], cwd=d, check=True, stdout=subprocess.DEVNULL)
        txt = np.loadtxt(os.path.join(d, "output.dat"), skiprows=2, delimiter=",", dtype=np.float64)
    # columns: X, Y, Z, u, v, w(always 0 as written), rho; order z, y, x
    u = txt[:, 3].reshape(nz, ny, nx)
    v = txt[:, 4].reshape(nz, ny, nx)
    rho = txt[:, 6].reshape(nz, ny, nx)
    zs = np.array([1, nz // 2, nz - 2])
    np.savez_compressed(
        os.path.join(HERE, "cpu3d_t100.npz"), dims=dims, consts=consts, zs=zs,
        u=u[zs].astype(np.float32), v=v[zs].astype(np.float32), rho=rho[zs].astype(np.float32),
        plane_u=u.sum(axis=(1, 2)), plane_v=v.sum(axis=(1, 2)), plane_rho=rho.sum(axis=(1, 2)))


if __name__ == "__main__":
    node_vectors()
    full_run()
    print("golden fixtures written to", HERE)
