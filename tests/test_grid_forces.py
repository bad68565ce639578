"""Grid-update extras of gridVelocityExplicitUpdate (MpmSimulationBase.cpp:1179-1228):
external body forces (fext as MpmInitializationHelper::addExternalBodyForce
builds them, :1436-1444), the explicit velocity field (:1187-1195) and
ignoreCollisionObject (:1217-1221).

Oracle pins (CPU), on a stress-free cube whose particles all see the same grid
velocity: one substep from rest adds exactly g dt + dt sum_k scale_k (sin(w_k t)
+ 1) d_k at the substep's own time t; with ignoreCollisionObject a sticky floor
leaves the run equal to one without it; the velocity field a + B x is
reproduced on the particles (v = a + B x_p, C = B) whatever gravity and the
objects would do.  GPU parity: both forms against the oracle over substeps
with a time offset, and a batch whose scenes carry different forces equals
each scene alone.
"""
import copy

import numpy as np
import pytest

from ziran2020_amd import _abi as A, scenes
from helpers import normwise, run


def free_cube(cells=6, seed=0):
    sc = scenes.elastic_cube(n_side_cells=cells, seed=seed)
    sc.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 0.0, 0.0)]
    sc.colliders = []
    return sc


FORCES = [(2.0, 30.0, (1.0, 0.0, 0.0)), (0.5, -7.0, (0.0, 0.6, 0.8))]


def test_body_forces_closed_form(oracle):
    sc = free_cube()
    sc.body_forces = FORCES
    v0 = np.array([0.1, -0.2, 0.3])
    sc.v[:] = v0
    t = 0.37
    sim = sc.make(oracle)
    sim.substep(sc.dt, t)
    out = sim.download()
    want = v0 + np.array(sc.gravity) * sc.dt
    for scale, w, d in FORCES:
        want = want + sc.dt * scale * (np.sin(w * t) + 1) * np.array(d)
    assert np.abs(out["v"] - want).max() < 1e-13
    # at t + dt the forces differ: the substep's own time is used
    sim.substep(sc.dt, t + sc.dt)
    out2 = sim.download()
    want2 = want + np.array(sc.gravity) * sc.dt
    for scale, w, d in FORCES:
        want2 = want2 + sc.dt * scale * (np.sin(w * (t + sc.dt)) + 1) * np.array(d)
    assert np.abs(out2["v"] - want2).max() < 1e-13


def test_ignore_collision_objects(oracle):
    sc = scenes.elastic_cube(n_side_cells=6)
    sc.v[:, 1] = -2.0
    free = copy.copy(sc)
    free.colliders = []
    a = run(oracle, sc, A.PREC_FP64, 5, ignore_collision_objects=True).download()
    b = run(oracle, free, A.PREC_FP64, 5).download()
    c = run(oracle, sc, A.PREC_FP64, 5).download()
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(a[k], b[k]), k
    assert normwise(c["v"], b["v"]) > 1e-3  # the floor does act without the flag


def test_velocity_field_reproduced(oracle):
    sc = scenes.elastic_cube(n_side_cells=6)  # gravity, sticky floor and stress all ignored
    a = np.array([0.1, -0.2, 0.05])
    B = np.array([[0.3, -0.1, 0.2], [0.05, -0.4, 0.1], [0.2, 0.0, 0.1]])
    sc.velocity_field = (a, B)
    sim = sc.make(oracle)
    sim.substep(sc.dt, 0.0)
    out = sim.download()
    assert np.abs(out["v"] - (a + sc.x @ B.T)).max() < 1e-12
    assert np.abs(out["C"] - B.ravel()).max() < 1e-10
    g = sim.grid()
    assert np.abs(g["new_v"] - (a + (g["coords"] * sc.dx) @ B.T)).max() < 1e-13
    # removing the field restores the force update
    sim.set_velocity_field(None)
    sim.substep(sc.dt, sc.dt)
    assert np.abs(sim.download()["v"][:, 1] - out["v"][:, 1]).max() > 0


def test_grid_force_validation(oracle):
    sim = free_cube(cells=2).make(oracle)
    for _ in range(4):
        sim.add_body_force(1.0, 1.0, (0, 1, 0))
    with pytest.raises(A.ZmpmError):
        sim.add_body_force(1.0, 1.0, (0, 1, 0))
    with pytest.raises(A.ZmpmError):
        sim.add_body_force(1.0, 1.0, (0, 1, 0), scene=1)
    with pytest.raises(A.ZmpmError):
        sim.set_velocity_field((0, 0, 0), None)


def forced_scene():
    sc = scenes.elastic_cube(n_side_cells=8, seed=3)
    sc.body_forces = FORCES
    sc.v = np.random.default_rng(4).normal(scale=0.2, size=sc.v.shape)
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_gpu_body_forces_match_oracle(gpu, oracle, precision):
    sc = forced_scene()
    sc.t0 = 0.25
    g = run(gpu, sc, precision, 10).download()
    o = run(oracle, sc, A.PREC_FP64, 10).download()
    tol = 1e-9 if precision == A.PREC_FP64 else 1e-4
    for k in ("x", "v", "C", "F"):
        assert normwise(g[k], o[k]) < tol, (k, normwise(g[k], o[k]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_gpu_velocity_field_and_ignored_objects(gpu, oracle, precision):
    sc = scenes.elastic_cube(n_side_cells=8, seed=5)
    sc.v = np.random.default_rng(6).normal(scale=0.3, size=sc.v.shape)
    sc.v[:, 1] -= 1.0
    tol = 1e-9 if precision == A.PREC_FP64 else 1e-4
    g = run(gpu, sc, precision, 6, ignore_collision_objects=True).download()
    o = run(oracle, sc, A.PREC_FP64, 6, ignore_collision_objects=True).download()
    for k in ("x", "v", "C", "F"):
        assert normwise(g[k], o[k]) < tol, (k, normwise(g[k], o[k]))
    vf = copy.copy(sc)
    vf.velocity_field = (np.array([0.1, -0.2, 0.05]), np.array([[0.3, -0.1, 0.2], [0.05, -0.4, 0.1], [0.2, 0.0, 0.1]]))
    g = run(gpu, vf, precision, 6).download()
    o = run(oracle, vf, A.PREC_FP64, 6).download()
    for k in ("x", "v", "C", "F"):
        assert normwise(g[k], o[k]) < tol, (k, normwise(g[k], o[k]))


@pytest.mark.gpu
def test_gpu_batch_with_per_scene_forces(gpu):
    b = scenes.batched_scenes(n_scenes=4, n_range=(2000, 8000))
    b.scenes[1].body_forces = FORCES
    b.scenes[2].velocity_field = (np.array([0.0, 0.1, 0.0]), np.zeros((3, 3)))
    out = run(gpu, b, A.PREC_FP32, 8).download()
    off = 0
    for s in b.scenes:
        alone = run(gpu, s, A.PREC_FP32, 8).download()
        for k in ("x", "v", "C", "F"):
            assert np.array_equal(out[k][off:off + s.n], alone[k]), (s.name, k)
        off += s.n
