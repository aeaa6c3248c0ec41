"""lbm_mass_momentum_noalloc -- the entry point that allocates nothing --
against lbm_get_mass_momentum through the C ABI.

Both run the same kernels (the grid-stride FP64 partial sums and the
fixed-order final sum), so the four sums must agree bit for bit.  Two calls
in a row, the second on a smaller lattice: the first fills more partial slots
of the shared scratch array than the second reduces, so stale slots would
show.  Also checked against the populations summed on the host in FP64
(relative 1e-12 for FP64 storage, 1e-6 where each node's sum runs in FP32)
and the refusals of bad arguments.
"""
import ctypes as C

import numpy as np
import pytest

import lbm_multigpu_amd as L
from test_gpu_parity import random_state, say

pytestmark = pytest.mark.gpu

HIP_D2H = 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _case(nx, ny, nz, storage, seed):
    c = L.default_config()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.bc[0], c.bc[1], c.bc[2] = L.BC_PERIODIC, L.BC_WALL, L.BC_WALL
    c.model, c.omega, c.cs = L.MODEL_MRT_LES, 1.7, 0.01
    c.halfway, c.storage = 1, storage
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[:, ny // 2 - 2: ny // 2 + 2, nx // 4: nx // 4 + 4] = L.NODE_SOLID_FORCE
    f = random_state(np.random.default_rng(seed), (nz, ny, nx), dtype=np.float64)
    return c, mask, f


def _noalloc(s, hip, dbuf, level=0):
    rc = s.lib.lbm_mass_momentum_noalloc(s.h, level, dbuf, None)
    assert rc == 0, s.lib.lbm_last_error(s.h).decode()
    assert hip.hipDeviceSynchronize() == 0
    out = np.zeros(4, np.float64)
    assert hip.hipMemcpy(out.ctypes.data, dbuf, 32, HIP_D2H) == 0
    return out


@pytest.mark.parametrize("storage", [L.STORE_F32, L.STORE_F16, L.STORE_F64])
def test_noalloc_mass_momentum_matches_allocating_entry_point(storage):
    hip = _hip()
    dbuf = C.c_void_p()
    assert hip.hipMalloc(C.byref(dbuf), 64) == 0
    try:
        # 200x64x64: 4 segments per row, 4096 workgroups' worth of rows, so
        # the 2048-workgroup launch strides; then a 40x16x12 lattice (48)
        for k, (nx, ny, nz) in enumerate(((200, 64, 64), (40, 16, 12))):
            say(f"noalloc: storage {storage}, {nx}x{ny}x{nz}")
            c, mask, f = _case(nx, ny, nz, storage, seed=11 + k)
            with L.Solver(c) as s:
                s.set_geometry(mask)
                s.set_populations(f)
                s.step(3)
                ref = s.mass_momentum()
                assert hip.hipMemset(dbuf, 0xFF, 64) == 0
                got = _noalloc(s, hip, dbuf)
                assert np.array_equal(got, ref), (got, ref)
                # an independent FP64 sum of the same populations over the fluid nodes
                g = s.populations(f64=True)
                fluid = ~np.isin(mask, (L.NODE_SOLID, L.NODE_SOLID_FORCE))
                cvec = np.array([[0, 1, 0, -1, 0, 1, -1, -1, 1, 0, 1, 0, -1, 0, 0, 1, 0, -1, 0],
                                 [0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 0, 1, 0, -1, 0, 0, 1, 0, -1],
                                 [0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1]], np.float64)
                host = np.array([g[:, fluid].sum()] + [(cvec[a][:, None] * g[:, fluid]).sum() for a in range(3)])
                # each node's 19 values are summed in the storage's arithmetic
                # (FP32 for 32- and 16-bit storage), the nodes in FP64
                tol = 1e-12 if storage == L.STORE_F64 else 1e-6
                assert abs(got[0] - host[0]) <= tol * abs(host[0])
                assert np.abs(got[1:] - host[1:]).max() <= tol * abs(host[0])
                # a repeated call gives the same bits
                assert np.array_equal(_noalloc(s, hip, dbuf), got)
    finally:
        hip.hipFree(dbuf)


def test_noalloc_refinement_level_and_refusals():
    hip = _hip()
    dbuf = C.c_void_p()
    assert hip.hipMalloc(C.byref(dbuf), 64) == 0
    try:
        from lbm_multigpu_amd import config as cfgmod
        cc = cfgmod.channel(nx=64, ny=32, nz=32, side=6, x_center=22, tau=0.55)
        cc.refinement = True
        cc.lr_size = (32, 24, 24)
        cc.lr_loc = (12.25, 4.25, 4.25)
        cfg = cc.to_c()
        say("noalloc: refined channel, both levels")
        with L.Solver(cfg) as s:
            s.set_geometry(None)
            s.init_equilibrium(L.INIT_BGK)
            s.step(4)
            for level in (0, 1):
                ref = s.mass_momentum(level)
                assert np.array_equal(_noalloc(s, hip, dbuf, level), ref)
            assert s.lib.lbm_mass_momentum_noalloc(s.h, 2, dbuf, None) == -22
            assert s.lib.lbm_mass_momentum_noalloc(s.h, 0, None, None) == -22
    finally:
        hip.hipFree(dbuf)
