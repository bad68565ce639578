"""Stage-level entry (zmpm_stage, zmpm_set_grid_velocity): the virtual steps of
MpmSimulationBase::advanceOneTimeStep (MpmSimulationBase.h:152-227) one at a time.

Bars: the five stages in order equal zmpm_substep (bit for bit in the oracle
and in the FP32 form, to the FP64 atomic-order bar in the FP64 form), on plain,
damage and implicit-damage scenes; a host override of the grid update through
zmpm_set_grid_velocity (new_v = a + B x_i written from the host) equals the
device's own explicit velocity field; stages out of order, with another dt or
t, and state changes between stages are refused.
"""
import numpy as np
import pytest

from ziran2020_amd import _abi as A, scenes
from helpers import normwise

STAGES = (A.STAGE_SORT, A.STAGE_P2G, A.STAGE_GRID_UPDATE, A.STAGE_DAMAGE_SOLVE, A.STAGE_G2P)
FIELDS = ("x", "v", "C", "F", "d", "g")


def slab(implicit=False):
    sc = scenes.damage_slab(n=6000, dx=1.0 / 64, size=(0.5, 0.15, 0.15), domain_nodes=128, sigma_crit=20.0,
                            pull_speed=0.5, seed=2, eta=1e-3 if implicit else 0.1)
    if implicit:
        sc.options = dict(implicit_damage=True)
    return sc


CASES = {
    "cube": lambda: scenes.elastic_cube(n_side_cells=6),
    "damage": slab,
    "implicit_damage": lambda: slab(True),
}


def staged(lib, sc, precision, steps):
    sim = sc.make(lib, precision)
    for k in range(steps):
        for st in STAGES:
            sim.stage(st, sc.dt, k * sc.dt)
    return sim.download()


def whole(lib, sc, precision, steps):
    sim = sc.make(lib, precision)
    for k in range(steps):
        sim.substep(sc.dt, k * sc.dt)
    return sim.download()


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_stages_equal_the_substep(oracle, case):
    sc = CASES[case]()
    a, b = staged(oracle, sc, A.PREC_FP64, 4), whole(oracle, sc, A.PREC_FP64, 4)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


def affine():
    return np.array([0.1, -0.2, 0.05]), np.array([[0.3, -0.1, 0.2], [0.05, -0.4, 0.1], [0.2, 0.0, 0.1]])


def host_override(lib, sc, precision, steps):
    """A host gridVelocityExplicitUpdate: new_v = a + B x_i from the node coordinates."""
    a, B = affine()
    sim = sc.make(lib, precision)
    for k in range(steps):
        t = k * sc.dt
        for st in (A.STAGE_SORT, A.STAGE_P2G, A.STAGE_GRID_UPDATE):
            sim.stage(st, sc.dt, t)
        g = sim.grid()
        xi = g["coords"] * sc.dx
        sim.set_grid_velocity(a[None, :] + (B[:, 0][None, :] * xi[:, :1] + B[:, 1][None, :] * xi[:, 1:2]
                                            + B[:, 2][None, :] * xi[:, 2:3]))
        sim.stage(A.STAGE_DAMAGE_SOLVE, sc.dt, t)
        sim.stage(A.STAGE_G2P, sc.dt, t)
    return sim.download()


def test_oracle_host_grid_update(oracle):
    sc = scenes.elastic_cube(n_side_cells=6)
    got = host_override(oracle, sc, A.PREC_FP64, 3)
    sc.velocity_field = affine()
    want = whole(oracle, sc, A.PREC_FP64, 3)
    for k in ("x", "v", "C", "F"):
        assert normwise(got[k], want[k]) < 1e-14, (k, normwise(got[k], want[k]))


@pytest.mark.parametrize("lib_name", ["oracle"])
def test_stage_order_is_enforced(lib_name, request):
    lib = request.getfixturevalue(lib_name)
    sc = scenes.elastic_cube(n_side_cells=4)
    sim = sc.make(lib)
    with pytest.raises(A.ZmpmError):
        sim.stage(A.STAGE_P2G, sc.dt, 0.0)
    sim.stage(A.STAGE_SORT, sc.dt, 0.0)
    with pytest.raises(A.ZmpmError):
        sim.stage(A.STAGE_P2G, 2 * sc.dt, 0.0)
    with pytest.raises(A.ZmpmError):
        sim.stage(A.STAGE_P2G, sc.dt, 1.0)
    with pytest.raises(A.ZmpmError):
        sim.substep(sc.dt, 0.0)
    with pytest.raises(A.ZmpmError):
        sim.set_materials(sc.materials)
    with pytest.raises(A.ZmpmError):
        sim.set_grid_velocity(np.zeros((1, 3)))  # before GRID_UPDATE
    for st in STAGES[1:]:
        sim.stage(st, sc.dt, 0.0)
    sim.substep(sc.dt, sc.dt)  # the substep is complete again


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_stages_equal_the_substep(gpu, case, precision):
    sc = CASES[case]()
    a, b = staged(gpu, sc, precision, 4), whole(gpu, sc, precision, 4)
    exact = precision == A.PREC_FP32 and case != "implicit_damage"  # the solve sums with double atomics
    for k in FIELDS:
        if exact:
            assert np.array_equal(a[k], b[k]), k
        else:
            assert normwise(a[k], b[k]) < (1e-9 if precision == A.PREC_FP64 else 1e-5), (k, normwise(a[k], b[k]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_gpu_host_grid_update(gpu, oracle, precision):
    sc = scenes.elastic_cube(n_side_cells=8)
    got = host_override(gpu, sc, precision, 3)
    sc.velocity_field = affine()
    want = whole(gpu, sc, precision, 3)
    ref = whole(oracle, sc, A.PREC_FP64, 3)
    tol = 1e-12 if precision == A.PREC_FP64 else 1e-6
    for k in ("x", "v", "C", "F"):
        assert normwise(got[k], want[k]) < tol, (k, normwise(got[k], want[k]))
        assert normwise(got[k], ref[k]) < (1e-9 if precision == A.PREC_FP64 else 1e-4), k


@pytest.mark.gpu
def test_gpu_stage_order_is_enforced(gpu):
    sc = scenes.elastic_cube(n_side_cells=4)
    sim = sc.make(gpu)
    with pytest.raises(A.ZmpmError):
        sim.stage(A.STAGE_G2P, sc.dt, 0.0)
    sim.stage(A.STAGE_SORT, sc.dt, 0.0)
    for bad in (lambda: sim.substep(sc.dt, 0.0), lambda: sim.run(2, sc.dt),
                lambda: sim.set_materials(sc.materials), lambda: sim.stage(A.STAGE_P2G, sc.dt, 0.5)):
        with pytest.raises(A.ZmpmError):
            bad()
    for st in STAGES[1:]:
        sim.stage(st, sc.dt, 0.0)
    sim.run(2, sc.dt, sc.dt)
