"""Pins for the oracle's y-streamwise inflow family and ghost copies.

The checks follow what the conditions prescribe, on one oracle step from
random populations (dpdy = 0, so the collide conserves rho and j):
  * im 100 North_Extrap: rho = 1 exactly, j of the unstreamed node below;
  * im 200 South_Extrap: j = (0, v_in, 0) exactly, rho of the node above;
  * im 80: rho of the unstreamed node above, j of the node below;
  * im 70: j of the node below, rho of the node above;
  * im 50-54: the output is the unstreamed source node, bit for bit;
  * the precursor wrap of bgk_channel1 (rows 0 .. DYNY1 periodic among
    themselves) by brute-force streaming;
  * South_Extrap's moments: every non-conserved moment from the neighbour.
A node reading the wrong row, or rho and j swapped between the two rows,
fails these (the random rows differ in every moment).
"""
import numpy as np

import oracle_lib as O

C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
              (1, -1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, -1, 1), (0, 0, -1), (1, 0, -1),
              (0, 1, -1), (-1, 0, -1), (0, -1, -1)])
W = np.array([1 / 3] + [1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4)


def random_field(rng, shape):
    nz, ny, nx = shape
    f = np.empty((19, nz, ny, nx))
    for idx in np.ndindex(nz, ny, nx):
        f[(slice(None),) + idx] = O.feq_bgk(float(rng.uniform(0.97, 1.03)), *rng.uniform(-0.05, 0.05, 3),
                                            dtype=np.float64)
    return f + rng.uniform(-1e-3, 1e-3, f.shape) * W[:, None, None, None]


def rho_j(f):
    return f.sum(0), np.tensordot(C.T.astype(float), f, axes=(1, 0))


def test_y_streamwise_conditions_prescribe_and_copy_the_right_moments():
    rng = np.random.default_rng(41)
    nx, ny, nz = 6, 9, 4
    f = random_field(rng, (nz, ny, nx))
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[:, ny - 1, :] = O.OUTLET_NORTH      # y = ny-1 reads y = ny-2
    mask[:, 0, :] = O.INLET_SOUTH            # y = 0 reads y = 1
    mask[:, 4, :] = O.RECYCLE                # rho of y = 5, moments of y = 3
    mask[:, 6, :] = O.VEL_FROM_BELOW         # j of y = 5, rest of y = 7
    v_in = rng.uniform(0.02, 0.06, nz * nx).astype(np.float32)
    for coll in (O.MRT_LES, O.BGK):
        p = O.params(nx, ny, nz, collision=coll, omega=1.6, cs=0.01, v_in=v_in)
        g, _ = O.step(p, f, mask=mask)
        r0, j0 = rho_j(f)
        r1, j1 = rho_j(g)
        assert np.abs(r1[:, ny - 1] - 1).max() < 1e-14
        assert np.abs(j1[:, :, ny - 1] - j0[:, :, ny - 2]).max() < 1e-14
        assert np.abs(j1[1, :, 0] - v_in.reshape(nz, nx)).max() < 1e-8  # the table is FP32
        assert np.abs(j1[[0, 2], :, 0]).max() < 1e-14
        assert np.abs(r1[:, 0] - r0[:, 1]).max() < 1e-14
        assert np.abs(r1[:, 4] - r0[:, 5]).max() < 1e-14
        assert np.abs(j1[:, :, 4] - j0[:, :, 3]).max() < 1e-14
        assert np.abs(j1[:, :, 6] - j0[:, :, 5]).max() < 1e-14
        assert np.abs(r1[:, 6] - r0[:, 7]).max() < 1e-14


def test_south_extrap_keeps_every_nonconserved_moment():
    """South_Extrap (MRTtest/DynInlet3.cu:881-924) and its four-argument
    form: the output's moments are the input's except j."""
    M = O.moment_matrix()
    rng = np.random.default_rng(42)
    for _ in range(20):
        f = random_field(rng, (1, 1, 1))[:, 0, 0, 0].copy()
        j = rng.uniform(-0.05, 0.05, 3)
        g = f.copy()
        O.lib().orc_vel_extrap_f64(g, *j)
        exp = M @ f
        exp[3], exp[5], exp[7] = j
        assert np.abs(M @ g - exp).max() < 1e-14


def test_ghost_copies_are_the_unstreamed_source_nodes():
    """im 50..54 (MRTtest/DynInlet3.cu:1747-1771): the node stores the source
    node's populations as they were at the step's start."""
    rng = np.random.default_rng(43)
    nx, ny, nz = 7, 10, 3
    f = random_field(rng, (nz, ny, nx))
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[:, 2:8, 0] = O.GHOST_W        # copy of (nx-2, y)
    mask[:, 2:8, nx - 1] = O.GHOST_E   # copy of (1, y)
    mask[:, 0, 1:-1] = O.GHOST_S       # copy of (x, DYNY1-1)
    mask[:, 9, 1:-1] = O.GHOST_N       # copy of (x, DYNY2)
    mask[:, 8, 1:-1] = O.GHOST_DYN     # copy of (x, 1)
    dyn_y1, dyn_y2 = 6, 3
    p = O.params(nx, ny, nz, collision=O.MRT_LES, omega=1.6, ghost_y=(dyn_y1 - 1, dyn_y2, 1))
    g, _ = O.step(p, f, mask=mask)
    assert np.array_equal(g[:, :, 2:8, 0], f[:, :, 2:8, nx - 2])
    assert np.array_equal(g[:, :, 2:8, nx - 1], f[:, :, 2:8, 1])
    assert np.array_equal(g[:, :, 0, 1:-1], f[:, :, dyn_y1 - 1, 1:-1])
    assert np.array_equal(g[:, :, 9, 1:-1], f[:, :, dyn_y2, 1:-1])
    assert np.array_equal(g[:, :, 8, 1:-1], f[:, :, 1, 1:-1])


def test_precursor_wrap_streams_rows_zero_to_dyny1_periodically():
    """bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1131-1132: row DYNY1 pulls its
    +y neighbours' populations from row 0 and row 0 its -y ones from row
    DYNY1; every other pull is the periodic rule."""
    rng = np.random.default_rng(44)
    nx, ny, nz, y1 = 4, 9, 3, 5
    f = rng.uniform(0, 1, (19, nz, ny, nx))
    g, _ = O.step(O.params(nx, ny, nz, collision=O.BGK, omega=0.0, y_wrap=y1), f)
    for i in range(19):
        cx, cy, cz = C[i]
        for y in range(ny):
            sy = (y - cy) % ny
            if y == y1 and cy < 0:
                sy = 0
            if y == 0 and cy > 0:
                sy = y1
            exp = np.roll(np.roll(f[i, :, sy, :], cz, axis=0), cx, axis=1)
            assert np.array_equal(g[i, :, y, :], exp), (i, y)
