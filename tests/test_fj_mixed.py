"""FJMixedMpmForceHelper<F model, QuadraticVolumePenalty> on the explicit path
(Lib/MPM/Force/FJMixedMpmForceHelper.h:16-278, added by addFJMixedMpmForce,
MpmParticleHandleBase.cpp:811-834).

Oracle pins (CPU):
  * evolveStrain (:168-187): over one substep of an affine field v = a + B x,
    J <- (1 + dt tr B) J and F <- (I + dt B) F, J carried apart from det F;
  * SOLVE_F (:98-105): a solid's stress is its F model's, so an FJ solid moves
    bit for bit as the plain F model;
  * SOLVE_J (:106-121): vtau = V0 lambda_J (J - 1) J I.  With mu = 0 the
    fixed-corotated stress is lambda (det F - 1) det F I, so an FJ solid with
    J0 = s moves as a mu = 0 corotated particle with F0 = diag(s, 1, 1); the
    first substep uses the J model's own lambda, later ones the F model's (the
    copy at :186);
  * a fluid phase: F, J and lambda_J never change, and its F model has mu =
    lambda = 0 (MpmParticleHandleBase.cpp:818-821).
GPU parity: FJ solids and fluids in one scene, SOLVE_F then SOLVE_J, both
forms against the oracle.
"""
import copy

import numpy as np
import pytest

from ziran2020_amd import _abi as A, scenes
from helpers import normwise, run

FIELDS = ("x", "v", "C", "F", "J")


def small_cube(cells=6, seed=0):
    return scenes.elastic_cube(n_side_cells=cells, seed=seed)


def affine(sc):
    a = np.array([0.1, -0.2, 0.05])
    B = np.array([[0.3, -0.1, 0.2], [0.05, -0.4, 0.1], [0.2, 0.0, 0.1]])
    sc.v = a + sc.x @ B.T
    sc.C = np.tile(B.ravel(), (sc.n, 1))
    sc.colliders = []
    sc.gravity = (0.0, 0.0, 0.0)
    return a, B


def test_evolve_strain_of_an_affine_field(oracle):
    sc = small_cube()
    sc.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 0.0, 0.0, fj_mixed=A.FJ_SOLID, j_E=0.0)]
    a, B = affine(sc)
    F0 = np.diag([1.05, 0.97, 1.0])
    sc.F = np.tile(F0.ravel(), (sc.n, 1))
    sc.J = np.full(sc.n, 1.1)  # independent of det F0
    sim = sc.make(oracle)
    sim.substep(sc.dt, 0.0)
    out = sim.download()
    assert np.abs(out["J"] - (1 + sc.dt * np.trace(B)) * 1.1).max() < 1e-14
    assert np.abs(out["F"] - ((np.eye(3) + sc.dt * B) @ F0).ravel()).max() < 1e-12
    assert np.abs(out["v"] - (a + sc.x @ B.T)).max() < 1e-12


def test_solve_f_is_the_f_model(oracle):
    sc = small_cube(seed=4)
    sc.v = np.random.default_rng(1).normal(scale=0.2, size=sc.v.shape)
    plain = sc.make(oracle)
    plain.run(8, sc.dt)
    fj = copy.copy(sc)
    fj.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 1e4, 0.3, fj_mixed=A.FJ_SOLID, j_E=3e4)]
    sim = fj.make(oracle)
    sim.run(8, sc.dt)
    a, b = plain.download(), sim.download()
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(a[k], b[k]), k
    assert not np.allclose(b["J"], 1.0)


def _stretched(sc, s, lam_plain):
    """The cube with F0 = diag(s, 1, 1) and a mu = 0 corotated model of lambda lam_plain."""
    out = copy.copy(sc)
    m = A.make_material(A.MODEL_FIXED_COROTATED, 0.0, 0.0)
    m.lam = lam_plain
    out.materials = [m]
    out.F = np.tile(np.diag([s, 1.0, 1.0]).ravel(), (sc.n, 1))
    return out


def test_solve_j_stress_and_the_lambda_copy(oracle):
    sc = small_cube(seed=2)
    sc.gravity = (0.0, 0.0, 0.0)
    lam_j, lam_f, s = 2000.0, 5000.0, 1.02
    m = A.make_material(A.MODEL_FIXED_COROTATED, 0.0, 0.0, fj_mixed=A.FJ_SOLID)
    m.lam, m.j_lambda = lam_f, lam_j
    fj = copy.copy(sc)
    fj.materials = [m]
    fj.J = np.full(sc.n, s)
    sim = fj.make(oracle)
    sim.set_fj_state(A.SOLVE_J)
    sim.substep(sc.dt, 0.0)
    got = sim.download()
    # first substep: the J model's own lambda
    ref = _stretched(sc, s, lam_j).make(oracle)
    ref.substep(sc.dt, 0.0)
    want = ref.download()
    for k in ("x", "v", "C"):
        assert normwise(got[k], want[k]) < 1e-12, (k, normwise(got[k], want[k]))
    # second substep: lambda_J = the F model's lambda; a mu = 0 corotated particle
    # with det F = the J the first substep left, from the same state
    J1 = got["J"]
    nxt = _stretched(sc, 1.0, lam_f)
    nxt.x, nxt.v, nxt.C = got["x"], got["v"], got["C"]
    nxt.F = np.zeros((sc.n, 9))
    nxt.F[:, 0], nxt.F[:, 4], nxt.F[:, 8] = J1, 1.0, 1.0
    ref2 = nxt.make(oracle)
    ref2.substep(sc.dt, sc.dt)
    sim.substep(sc.dt, sc.dt)
    got2, want2 = sim.download(), ref2.download()
    for k in ("x", "v", "C"):
        assert normwise(got2[k], want2[k]) < 1e-12, (k, normwise(got2[k], want2[k]))
    # and not the J model's lambda
    nxt.materials[0].lam = lam_j
    ref3 = nxt.make(oracle)
    ref3.substep(sc.dt, sc.dt)
    assert normwise(got2["v"], ref3.download()["v"]) > 1e-6


@pytest.mark.parametrize("state", [A.SOLVE_F, A.SOLVE_J])
def test_fluid_phase_keeps_f_and_j(oracle, state):
    sc = small_cube(seed=5)
    sc.v = np.random.default_rng(2).normal(scale=0.3, size=sc.v.shape)
    m = A.make_material(A.MODEL_NEO_HOOKEAN, 1e4, 0.3, fj_mixed=A.FJ_FLUID, j_E=1e3)
    fl = copy.copy(sc)
    fl.materials = [m]
    fl.J = np.full(sc.n, 0.98)
    F0 = sc.F.copy()
    sim = fl.make(oracle)
    sim.set_fj_state(state)
    sim.run(5, sc.dt)
    out = sim.download()
    assert np.array_equal(out["F"], F0) and np.all(out["J"] == 0.98)
    if state == A.SOLVE_F:
        # the F model's mu and lambda are zeroed: no stress at all
        free = copy.copy(sc)
        free.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 0.0, 0.0)]
        ref = free.make(oracle)
        ref.run(5, sc.dt)
        want = ref.download()
        for k in ("x", "v", "C"):
            assert np.array_equal(out[k], want[k]), k


def test_fj_validation(oracle):
    sim = small_cube(cells=2).make(oracle)
    bad = [A.make_material(A.MODEL_QR_STABLE_NEO_HOOKEAN, fj_mixed=A.FJ_SOLID),
           A.make_material(A.MODEL_FIXED_COROTATED, fj_mixed=A.FJ_SOLID, has_damage=True),
           A.make_material(A.MODEL_FIXED_COROTATED, fj_mixed=A.FJ_SOLID, plasticity=A.PLASTICITY_SNOW),
           A.make_material(A.MODEL_EQUATION_OF_STATE, fj_mixed=A.FJ_FLUID),
           A.make_material(A.MODEL_FIXED_COROTATED, fj_mixed=3)]
    for m in bad:
        with pytest.raises(A.ZmpmError):
            sim.set_materials([m])
    with pytest.raises(A.ZmpmError):
        sim.set_fj_state(2)


def mixed_scene(cells=8):
    """Solid FJ corotated and neo-Hookean halves above a fluid FJ layer, moving."""
    sc = small_cube(cells=cells, seed=9)
    rng = np.random.default_rng(3)
    sc.v = rng.normal(scale=0.3, size=sc.v.shape)
    sc.F = (np.eye(3) + rng.normal(scale=0.01, size=(sc.n, 3, 3))).reshape(-1, 9)
    sc.J = 1 + rng.normal(scale=0.01, size=sc.n)
    sc.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 1e4, 0.3, fj_mixed=A.FJ_SOLID, j_E=2e4),
                    A.make_material(A.MODEL_NEO_HOOKEAN, 2e4, 0.35, fj_mixed=A.FJ_SOLID, j_E=5e3, j_nu=0.2),
                    A.make_material(A.MODEL_FIXED_COROTATED, 1e4, 0.3, fj_mixed=A.FJ_FLUID, j_E=1e4)]
    y = sc.x[:, 1]
    sc.material = np.where(y < np.quantile(y, 0.3), 2, np.where(sc.x[:, 0] < np.median(sc.x[:, 0]), 0, 1))
    sc.material = sc.material.astype(np.int32)
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_gpu_fj_mixed_matches_oracle(gpu, oracle, precision):
    sc = mixed_scene()
    outs = []
    for lib, prec in ((gpu, precision), (oracle, A.PREC_FP64)):
        sim = sc.make(lib, prec)
        sim.run(5, sc.dt)
        sim.set_fj_state(A.SOLVE_J)
        sim.run(5, sc.dt, 5 * sc.dt)
        outs.append(sim.download())
    g, o = outs
    tol = 1e-9 if precision == A.PREC_FP64 else 1e-4
    for k in FIELDS:
        assert normwise(g[k], o[k]) < tol, (k, normwise(g[k], o[k]))
    fluid = sc.material == 2
    assert np.all(o["J"][fluid] == sc.J[fluid])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_gpu_fj_solve_j_from_upload(gpu, oracle, precision):
    # SOLVE_J from the first substep: the uploaded particles' stress takes the J
    # model's own lambda, later substeps the F model's
    sc = mixed_scene(cells=6)
    outs = []
    for lib, prec in ((gpu, precision), (oracle, A.PREC_FP64)):
        sim = sc.make(lib, prec)
        sim.set_fj_state(A.SOLVE_J)
        sim.run(6, sc.dt)
        outs.append(sim.download())
    g, o = outs
    tol = 1e-9 if precision == A.PREC_FP64 else 1e-4
    for k in FIELDS:
        assert normwise(g[k], o[k]) < tol, (k, normwise(g[k], o[k]))
