"""GPU parity: the HIP library against the FP64 CPU oracle on identical seeded inputs.

Bars: binning (keys, particle order, groups, block offsets) bit-exact; FP64 plain
form within 1e-9 normwise of the oracle; FP32 form within the stated tolerance
of the FP64 oracle; FP32 form bit-reproducible run to run.
"""
import numpy as np
import pytest

from ziran2020_amd import _abi as A, scenes
from helpers import normwise, run

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9     # FP64 plain form vs FP64 oracle (atomic order only)
FP32_TOL = 1e-4     # FP32 form vs FP64 oracle, normwise per field


def small_cube(seed=1, cells=8):
    return scenes.elastic_cube(n_side_cells=cells, seed=seed)


def small_slab(n=20000, sigma_crit=20.0, seed=2):
    # low critical stress so the phase field actually evolves within a few steps
    return scenes.damage_slab(n=n, dx=1.0 / 64, size=(0.5, 0.15, 0.15), domain_nodes=128,
                              sigma_crit=sigma_crit, pull_speed=0.5, seed=seed)


def implicit_slab(n=20000, eta=1e-3, seed=2):
    # the small slab with a short damage time scale: the implicit solve runs a
    # few CG iterations per substep and the phase field grows within a few substeps
    return scenes.damage_slab(n=n, dx=1.0 / 64, size=(0.5, 0.15, 0.15), domain_nodes=128,
                              sigma_crit=20.0, pull_speed=0.5, seed=seed, eta=eta)


def compare_binning(a, b):
    assert np.array_equal(a["keys"], b["keys"])
    assert np.array_equal(a["order"], b["order"])
    assert np.array_equal(a["groups"], b["groups"])
    assert np.array_equal(a["block_offset"], b["block_offset"])


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_binning_bit_exact(gpu, oracle, precision):
    sc = small_cube(cells=12)
    g = run(gpu, sc, precision, 1)
    o = run(oracle, sc, precision, 1)
    compare_binning(g.binning(), o.binning())


def test_binning_bit_exact_after_motion(gpu, oracle):
    # the sort of a later substep runs on positions both sides computed (FP64)
    sc = small_cube(cells=10, seed=7)
    sc.v[:] = np.random.default_rng(3).normal(scale=0.3, size=sc.v.shape)
    g = run(gpu, sc, A.PREC_FP64, 5)
    o = run(oracle, sc, A.PREC_FP64, 5)
    xg, xo = g.download()["x"], o.download()["x"]
    assert normwise(xg, xo) < 1e-12
    g.substep(sc.dt, 5 * sc.dt)
    o.substep(sc.dt, 5 * sc.dt)
    compare_binning(g.binning(), o.binning())


def test_binning_scattered_blocks(gpu, oracle):
    # sparse random particles: many single-particle blocks and crowded cells
    rng = np.random.default_rng(11)
    sc = small_cube(cells=4)
    n = 3000
    x = rng.uniform(0.2, 0.8, size=(n, 3))
    x[:500] = x[0] + rng.uniform(0, 1e-3, size=(500, 3))  # 500 particles in one cell
    sc.x, sc.v, sc.C = x, np.zeros((n, 3)), np.zeros((n, 9))
    sc.F, sc.V0 = scenes.identity_F(n), np.full(n, 1e-7)
    sc.material, sc.rho = np.zeros(n, np.int32), np.full(n, 1000.0)
    sc.m = sc.rho * sc.V0
    for prec in (A.PREC_FP64, A.PREC_FP32):
        compare_binning(run(gpu, sc, prec, 1).binning(), run(oracle, sc, prec, 1).binning())


def test_grid_fp64_one_step(gpu, oracle):
    sc = small_cube(cells=12)
    sc.v[:] = np.random.default_rng(5).normal(scale=0.1, size=sc.v.shape)
    g = run(gpu, sc, A.PREC_FP64, 1).grid()
    o = run(oracle, sc, A.PREC_FP64, 1).grid()
    assert np.array_equal(g["coords"], o["coords"])
    for k in ("m", "v", "new_v"):
        assert normwise(g[k], o[k]) < 1e-12, k


def test_particles_fp64_elastic_cube(gpu, oracle):
    sc = small_cube(cells=12)
    g = run(gpu, sc, A.PREC_FP64, 30).download()
    o = run(oracle, sc, A.PREC_FP64, 30).download()
    for k in ("x", "v", "C", "F"):
        assert normwise(g[k], o[k]) < FP64_TOL, k


def test_particles_fp32_elastic_cube(gpu, oracle):
    sc = small_cube(cells=12)
    g = run(gpu, sc, A.PREC_FP32, 30).download()
    o = run(oracle, sc, A.PREC_FP64, 30).download()
    for k in ("x", "v", "C", "F"):
        assert normwise(g[k], o[k]) < FP32_TOL, k


def test_fp32_deterministic(gpu):
    sc = small_cube(cells=10)
    sc.v[:] = np.random.default_rng(9).normal(scale=0.2, size=sc.v.shape)
    a = run(gpu, sc, A.PREC_FP32, 10).download()
    b = run(gpu, sc, A.PREC_FP32, 10).download()
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_damage_slab(gpu, oracle, precision):
    sc = small_slab()
    steps = 20
    g = run(gpu, sc, precision, steps).download()
    o = run(oracle, sc, A.PREC_FP64, steps).download()
    assert o["d"].max() > 0.01, "scene must exercise the damage update"
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F", "d", "g"):
        assert normwise(g[k], o[k]) < (tol if k != "d" else 10 * tol), k


@pytest.mark.parametrize("variant", ["fibre", "isotropic", "fibre_implicit"])
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_qr_anisotropic_damage_slab(gpu, oracle, precision, variant):
    # the slab with QRAnisotropic (examples/orange.h:46-63): a stiffer undegraded
    # first fibre, or every fibre degraded (isotropic); explicit and implicit damage
    iso = variant == "isotropic"
    implicit = variant.endswith("implicit")
    sc = scenes.damage_slab(n=20000, dx=1.0 / 64, size=(0.5, 0.15, 0.15), domain_nodes=128, sigma_crit=20.0,
                            pull_speed=0.5, seed=5, model=A.MODEL_QR_ANISOTROPIC, isotropic=iso,
                            fiber_scale=(2.0, 1.0, 1.0), eta=1e-3 if implicit else 0.1)
    steps = 20
    g = run(gpu, sc, precision, steps, implicit_damage=implicit).download()
    o = run(oracle, sc, A.PREC_FP64, steps, implicit_damage=implicit).download()
    assert o["d"].max() > 0.01, "scene must exercise the damage update"
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F", "d", "g"):
        assert normwise(g[k], o[k]) < (tol if k != "d" else 10 * tol), k


def degraded_prestretched_cube(lib, precision, model, isotropic, g=0.3):
    # a pre-stretched cube at a fixed degradation g (no damage update): QRAnisotropic's
    # first fibre (stiffness 3 mu) is undegraded unless the model is isotropic
    sc = small_cube(cells=8)
    sc.materials = [A.make_material(model, 1e4, 0.3, isotropic=isotropic, fiber_scale=(3.0, 1.0, 1.0))]
    n = sc.x.shape[0]
    F = np.tile(np.diag([1.04, 0.98, 1.01]).ravel(), (n, 1))
    sim = sc.make(lib, precision)
    sim.upload(sc.x, sc.v, sc.C, F, sc.m, sc.V0, sc.material, g=np.full(n, g))
    sim.run(10, sc.dt)
    return sim.download()


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_qr_anisotropic_degraded_fibre(gpu, oracle, precision):
    # g < 1 on every particle: the isotropic flag decides whether the first fibre's
    # term is degraded, a difference far above the FP32 bar
    outs = {}
    for iso in (False, True):
        g = degraded_prestretched_cube(gpu, precision, A.MODEL_QR_ANISOTROPIC, iso)
        o = degraded_prestretched_cube(oracle, A.PREC_FP64, A.MODEL_QR_ANISOTROPIC, iso)
        tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
        for k in ("x", "v", "F"):
            assert normwise(g[k], o[k]) < tol, (iso, k)
        outs[iso] = o
    assert normwise(outs[True]["v"], outs[False]["v"]) > 100 * FP32_TOL


@pytest.mark.parametrize("implicit", [False, True])
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_neo_hookean_borden_damage_slab(gpu, oracle, precision, implicit):
    # the slab with NeoHookeanBorden, the third model updateDamage dispatches to
    sc = scenes.damage_slab(n=20000, dx=1.0 / 64, size=(0.5, 0.15, 0.15), domain_nodes=128, sigma_crit=20.0,
                            pull_speed=0.5, seed=6, model=A.MODEL_NEO_HOOKEAN_BORDEN, eta=1e-3 if implicit else 0.1)
    steps = 20
    g = run(gpu, sc, precision, steps, implicit_damage=implicit).download()
    o = run(oracle, sc, A.PREC_FP64, steps, implicit_damage=implicit).download()
    assert o["d"].max() > 0.01, "scene must exercise the damage update"
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F", "d", "g"):
        assert normwise(g[k], o[k]) < (tol if k != "d" else 10 * tol), k


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_damage_rest_state_stays_undamaged(gpu, oracle, precision):
    # Stress-free rotated particles (F0 = per-particle fibre rotation, no load):
    # the Cauchy stress is rounding noise of ~1e-10..1e-4 Pa, far below sigma_c,
    # so the phase field must stay exactly zero.  Regression for the float
    # eigenvector of a tiny stress underflowing to a NaN that clamped d to 1.
    sc = scenes.damage_slab(n=20000, dx=1.0 / 64, size=(0.5, 0.15, 0.15), domain_nodes=128,
                            sigma_crit=1e-2, pull_speed=0.0, seed=5)
    o = run(oracle, sc, A.PREC_FP64, 3).download()
    assert o["d"].max() == 0.0
    g = run(gpu, sc, precision, 3).download()
    assert np.all(np.isfinite(g["d"])) and g["d"].max() == 0.0
    assert np.all(g["g"] == 1.0)


def test_damage_grid_phase_field(gpu, oracle):
    sc = small_slab()
    sc_d = np.random.default_rng(1).uniform(0, 0.5, sc.n)
    gs = sc.make(gpu, A.PREC_FP64)
    os_ = sc.make(oracle, A.PREC_FP64)
    for s in (gs, os_):
        s.upload(sc.x, sc.v, sc.C, sc.F, sc.m, sc.V0, sc.material, d=sc_d)
        s.substep(sc.dt, 0.0)
    g, o = gs.grid(), os_.grid()
    assert np.array_equal(g["coords"], o["coords"])
    assert normwise(g["pf"], o["pf"]) < 1e-12
    gd, od = gs.download(), os_.download()
    assert normwise(gd["laplacian"], od["laplacian"]) < 1e-9
    assert normwise(gd["d"], od["d"]) < 1e-9


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_two_material(gpu, oracle, precision):
    sc = scenes.two_material(n=20000, dx=1.0 / 64, domain_nodes=128)
    g = run(gpu, sc, precision, 10).download()
    o = run(oracle, sc, A.PREC_FP64, 10).download()
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F"):
        assert normwise(g[k], o[k]) < tol, k


def test_shared_grid_scenes_small(gpu, oracle):
    sc = scenes.shared_grid_scenes(n_scenes=8, n_range=(500, 4000), domain_nodes=256)
    g = run(gpu, sc, A.PREC_FP64, 5).download()
    o = run(oracle, sc, A.PREC_FP64, 5).download()
    for k in ("x", "v", "F", "d"):
        assert normwise(g[k], o[k]) < FP64_TOL, k


FP32_TOL_100 = 2e-5  # FP32 form vs FP64 form after 100 substeps, every field


@pytest.mark.parametrize("name", ["elastic_cube", "two_material", "shared_grid_scenes", "batched_scenes"])
def test_fp32_error_after_100_substeps(gpu, name):
    # The FP64 form equals the FP64 oracle to 1e-9 (tests above); the FP32
    # form, with fixed-point positions and H = F - I, stays within 2e-5 of it
    # per field after 100 substeps on each scene type (fluid + solid, damage,
    # fixed corotated on a floor).
    if name == "elastic_cube":
        sc = scenes.elastic_cube()
    elif name == "two_material":
        sc = scenes.two_material(n=1_000_000, dx=1.0 / 128)
    elif name == "shared_grid_scenes":
        sc = scenes.shared_grid_scenes(n_scenes=8)
    else:
        sc = scenes.batched_scenes(n_scenes=8)
    a = run(gpu, sc, A.PREC_FP32, 100).download()
    b = run(gpu, sc, A.PREC_FP64, 100).download()
    for k in ("x", "v", "C", "F", "d"):
        assert normwise(a[k], b[k]) < FP32_TOL_100, (k, normwise(a[k], b[k]))


def test_sparse_spheres_small(gpu, oracle):
    sc = scenes.sparse_spheres(n=30000, n_spheres=5, dx=1.0 / 64, domain_nodes=128)
    g = run(gpu, sc, A.PREC_FP32, 10).download()
    o = run(oracle, sc, A.PREC_FP64, 10).download()
    for k in ("x", "v", "F"):
        assert normwise(g[k], o[k]) < FP32_TOL, k


def test_slip_and_separate_colliders(gpu, oracle):
    sc = small_cube(cells=8)
    sc.v[:, 0] = 1.0
    sc.colliders = [A.half_space((0, 8 * sc.dx, 0), (0, 1, 0), A.SLIP, friction=0.3),
                    A.half_space((30 * sc.dx, 0, 0), (-1, 0, 0), A.SEPARATE),
                    A.box((0, 0, 0), (20 * sc.dx, 30 * sc.dx, 20 * sc.dx), A.SLIP, b0=(0.01, 0, 0),
                          dbdt=(0.2, 0, 0))]
    g = run(gpu, sc, A.PREC_FP64, 10).download()
    o = run(oracle, sc, A.PREC_FP64, 10).download()
    for k in ("x", "v"):
        assert normwise(g[k], o[k]) < FP64_TOL, k


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_corner_slip_walls_stop_after_three(gpu, oracle, precision):
    """Nodes in a corner inside three SLIP / SEPARATE walls, then a moving STICKY box
    over everything: multiObjectCollision (CollisionObject.cpp:66-105) skips the box
    once three independent slip normals are in, so corner nodes keep the projected
    velocity while the rest take the box's."""
    sc = small_cube(cells=8)
    dx = sc.dx
    rng = np.random.default_rng(21)
    sc.v = rng.normal(scale=0.2, size=sc.v.shape)
    sc.colliders = [A.half_space((18 * dx, 0, 0), (1, 0, 0), A.SLIP),
                    A.half_space((0, 10 * dx, 0), (0, 1, 0), A.SEPARATE),
                    A.half_space((0, 0, 18 * dx), (0, 0, 1), A.SLIP, friction=0.2),
                    A.box((0, 0, 0), (40 * dx, 40 * dx, 40 * dx), A.STICKY, dbdt=(0.1, 0.2, 0.3))]
    o = run(oracle, sc, A.PREC_FP64, 1)
    go = o.grid()
    xi = go["coords"] * dx
    corner = (xi[:, 0] <= 18 * dx) & (xi[:, 1] <= 10 * dx) & (xi[:, 2] <= 18 * dx)
    assert corner.sum() > 10 and not np.any(np.all(go["new_v"][corner] == [0.1, 0.2, 0.3], axis=1))
    assert np.all(go["new_v"][~corner] == [0.1, 0.2, 0.3])
    g = run(gpu, sc, precision, 1)
    gg = g.grid()
    assert np.array_equal(gg["coords"], go["coords"]) or precision == A.PREC_FP32
    if precision == A.PREC_FP64:
        assert normwise(gg["new_v"], go["new_v"]) < 1e-12
    g.run(7, sc.dt, sc.dt)
    o.run(7, sc.dt, sc.dt)
    a, b = g.download(), o.download()
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "C", "F"):
        assert normwise(a[k], b[k]) < tol, (k, normwise(a[k], b[k]))


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_rotating_scaling_and_curved_colliders(gpu, oracle, precision):
    """Moving objects as examples/boneTwist.h sets them up: a STICKY sphere rotating
    about y while it translates, a SLIP capped cylinder with its own rotation (axis
    along x) and friction, a SEPARATE box that grows (scaling), a GHOST that is
    ignored -- the GPU against the oracle after 10 substeps."""
    sc = small_cube(cells=8)
    dx = sc.dx
    rng = np.random.default_rng(22)
    sc.v = rng.normal(scale=0.2, size=sc.v.shape)
    s2 = 2 ** -0.5
    top = 16 * dx
    sc.colliders = [
        A.sphere((0, 0, 0), 3 * dx, A.STICKY, b0=(20 * dx, top, 20 * dx), dbdt=(0, 0.02, 0), omega=(0, -0.7, 0)),
        A.capped_cylinder(2 * dx, 12 * dx, shape_rotation=(s2, 0, 0, s2), center=(20 * dx, 9 * dx, 17 * dx),
                          type_=A.SLIP, friction=0.3, omega=(0.4, 0, 0), b0=(0, 0.001, 0)),
        A.box((-2 * dx, -2 * dx, -2 * dx), (2 * dx, 2 * dx, 2 * dx), A.SEPARATE, b0=(23 * dx, 12 * dx, 23 * dx),
              scale0=1.0, dsdt=3.0),
        A.sphere((20 * dx, 12 * dx, 20 * dx), 10 * dx, A.GHOST),
        A.half_space((0, 8 * dx, 0), (0, 1, 0), A.STICKY),
    ]
    o = run(oracle, sc, A.PREC_FP64, 1)
    n_hit = o.grid()["new_v"]
    assert n_hit.shape[0] > 0
    g = run(gpu, sc, precision, 1)
    g.run(9, sc.dt, sc.dt)
    o.run(9, sc.dt, sc.dt)
    a, b = g.download(), o.download()
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "C", "F"):
        assert normwise(a[k], b[k]) < tol, (k, normwise(a[k], b[k]))


def test_empty_scene(gpu):
    sc = small_cube(cells=4)
    sim = sc.make(gpu, A.PREC_FP32)
    z3 = np.zeros((0, 3))
    sim.upload(z3, z3, np.zeros((0, 9)), np.zeros((0, 9)), np.zeros(0), np.zeros(0), np.zeros(0, np.int32))
    sim.run(3, 1e-4)
    assert sim.synchronize() == A.OK
    g = sim.binning()
    assert g["keys"].size == 0


def step_status(sim, dt):
    """Status of one substep: reported by the call itself or latched until the next sync."""
    st = sim.L.zmpm_substep(sim.h, dt, 0.0)
    return st if st != A.OK else sim.synchronize(check=False)


def test_cfl_violation_reported(gpu):
    sc = small_cube(cells=4)
    sc.v[:] = (200.0, 0.0, 0.0)  # 200 m/s * 1e-4 s > dx = 1/64 m
    sim = sc.make(gpu, A.PREC_FP64)
    assert step_status(sim, sc.dt) == A.E_CFL


def test_domain_and_nan_rejected(gpu):
    sc = small_cube(cells=4)
    sc.x[3] = (-0.5, 0.3, 0.3)
    sim = sc.make(gpu, A.PREC_FP32)
    assert step_status(sim, sc.dt) == A.E_DOMAIN
    for prec in (A.PREC_FP64, A.PREC_FP32):
        sc = small_cube(cells=4)
        sc.x[5, 1] = np.nan
        sim = sc.make(gpu, prec)
        assert step_status(sim, sc.dt) == A.E_NAN


def test_full_size_damage_slab_conservation(gpu):
    """1M-particle FP32 workload: P2G conserves mass and the fixed-point sums are exact."""
    sc = scenes.damage_slab()
    sim = sc.make(gpu, A.PREC_FP32)
    sim.run(3, sc.dt)
    grid = sim.grid()
    mp = sim.download()["m"].sum()
    assert abs(grid["m"].sum() - mp) / mp < 1e-6
    b = sim.binning()
    assert np.all(np.diff(b["keys"].astype(np.uint64)) > 0)  # strictly increasing keys
    assert np.array_equal(np.sort(b["order"]), np.arange(sc.n))


# transfer schemes (MpmSimulationBase::transfer_scheme, mls_mpm, damping passes)
TRANSFER_CASES = {
    "pic": dict(transfer_scheme=A.FLIP_BLEND_PIC, flip_pic_ratio=0.0),
    "flip95": dict(transfer_scheme=A.FLIP_BLEND_PIC, flip_pic_ratio=0.95),
    "mls": dict(mls_mpm=True),
    "apic_rpic_half": dict(apic_rpic_ratio=0.5),
    "rpic_damping": dict(rpic_damping_iteration=3),
    "pic_damping": dict(transfer_scheme=A.FLIP_BLEND_PIC, flip_pic_ratio=0.0, pic_damping_iteration=2),
}


@pytest.mark.parametrize("case", sorted(TRANSFER_CASES))
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_transfer_schemes(gpu, oracle, case, precision):
    # moving particles cross cells, so the damping passes' grid addressing through
    # the substep's binning (weights from the moved positions) is exercised
    sc = small_cube(cells=10, seed=4)
    sc.v[:] = np.random.default_rng(8).normal(scale=0.3, size=sc.v.shape)
    kw = TRANSFER_CASES[case]
    g = run(gpu, sc, precision, 12, **kw).download()
    o = run(oracle, sc, A.PREC_FP64, 12, **kw).download()
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "C", "F"):
        assert normwise(g[k], o[k]) < tol, k


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_damage_slab_pic(gpu, oracle, precision):
    # AnisoMPM with FLIP_blend_PIC at flip_pic_ratio = 0, as in the reference's
    # 3DTubeSmash / cheesePull / meatproxy examples
    sc = small_slab()
    kw = dict(transfer_scheme=A.FLIP_BLEND_PIC, flip_pic_ratio=0.0)
    g = run(gpu, sc, precision, 20, **kw).download()
    o = run(oracle, sc, A.PREC_FP64, 20, **kw).download()
    assert o["d"].max() > 0.01
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F", "d", "g"):
        assert normwise(g[k], o[k]) < (tol if k != "d" else 10 * tol), k


@pytest.mark.parametrize("precision", [A.PREC_FP32, A.PREC_FP64])
@pytest.mark.parametrize("transfer", ["apic", "flip_damped", "implicit_damage"])
def test_graph_replays_match_eager_substeps(gpu, precision, transfer):
    """zmpm_run replays captured HIP graphs of eight and of two substeps; every
    replay, with host downloads in between, matches eager one-at-a-time substeps (bitwise for the
    deterministic FP32 form, to FP64 rounding otherwise; the implicit damage
    solve sums in double with HBM atomics, so there the FP32 form agrees to float
    rounding)."""
    kw = {"apic": {}, "flip_damped": dict(transfer_scheme=A.FLIP_BLEND_PIC, flip_pic_ratio=0.5,
                                          pic_damping_iteration=2),
          "implicit_damage": dict(implicit_damage=True)}[transfer]
    sc = small_slab(n=8000) if transfer != "implicit_damage" else implicit_slab(n=8000)
    exact = precision == A.PREC_FP32 and transfer != "implicit_damage"
    tol = 1e-12 if precision == A.PREC_FP64 else 1e-5
    g = sc.make(gpu, precision, **kw)
    e = sc.make(gpu, precision, **kw)
    t = 0.0
    for rep in range(5):
        g.run(2, sc.dt, t)
        for k in range(2):
            e.substep(sc.dt, t + k * sc.dt)
        t += 2 * sc.dt
        a, b = g.download(), e.download()
        for k in ("x", "v", "C", "F", "d"):
            if exact:
                assert np.array_equal(a[k], b[k]), (rep, k)
            else:
                assert normwise(a[k], b[k]) < tol, (rep, k)
    # an odd step count ends on the other buffer; the next run captures its own
    # graphs; 11 = one eight-substep graph, a pair and one eager substep
    for n in (3, 11, 11):
        g.run(n, sc.dt, t)
        for k in range(n):
            e.substep(sc.dt, t + k * sc.dt)
        t += n * sc.dt
        a, b = g.download(), e.download()
        for k in ("x", "v", "F"):
            if exact:
                assert np.array_equal(a[k], b[k]), (n, k)
            else:
                assert normwise(a[k], b[k]) < tol, (n, k)


@pytest.mark.parametrize("case", ["corotated", "corotated_hardening", "damage_slab"])
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_von_mises_plasticity(gpu, oracle, case, precision):
    # VonMisesStvkHencky after the F update (MpmSimulationBase.cpp:1517-1529);
    # the damage slab case is the reference's vonMisesTear setup (QR-stable
    # Neo-Hookean with AnisoMPM damage and von Mises plasticity)
    if case == "damage_slab":
        sc = small_slab()
        m = sc.materials[0]
        m.plasticity, m.yield_stress, m.hardening = A.PLASTICITY_VON_MISES_STVK_HENCKY, 20.0, 0.0
        steps = 20
    else:
        sc = small_cube(cells=10, seed=6)
        sc.v[:, 1] = -1.0  # onto the sticky floor
        sc.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 1e4, 0.3, plasticity=A.PLASTICITY_VON_MISES_STVK_HENCKY,
                                        yield_stress=30.0, hardening=-800.0 if case.endswith("hardening") else 0.0)]
        steps = 25
    g = run(gpu, sc, precision, steps).download()
    o = run(oracle, sc, A.PREC_FP64, steps).download()
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    keys = ("x", "v", "F", "plastic_state") + (("d",) if case == "damage_slab" else ())
    for k in keys:
        assert normwise(g[k], o[k]) < (tol if k != "d" else 10 * tol), k
    # the scene must actually yield: elastic runs differ
    sc.materials[0].plasticity = A.PLASTICITY_NONE
    e = run(oracle, sc, A.PREC_FP64, steps).download()
    assert normwise(e["F"], o["F"]) > 1e-6
    if case == "corotated_hardening":
        # yield -= xi delta_gamma: xi < 0 hardens (PlasticityApplier.cpp:1503, "supposed
        # to only increase"); xi > 0 softens, which localises and amplifies rounding
        assert o["plastic_state"].max() > 30.0


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_drucker_prager_plasticity(gpu, oracle, precision):
    # a sand-like cube (DruckerPragerStvkHencky, 30 degrees, volume correction)
    # spreading on the sticky floor: tip, cone and inside cases all occur
    sc = small_cube(cells=10, seed=7)
    sc.v[:, 1] = -1.5
    sc.v[:, 0] = np.random.default_rng(2).normal(scale=0.3, size=sc.n)
    sc.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 1e4, 0.3, plasticity=A.PLASTICITY_DRUCKER_PRAGER_STVK_HENCKY,
                                    friction_angle=30.0, cohesion=0.0, hardening=1.0, volume_correction=True)]
    steps = 25
    g = run(gpu, sc, precision, steps).download()
    o = run(oracle, sc, A.PREC_FP64, steps).download()
    assert np.abs(o["plastic_state"]).max() > 0  # tip projections happened
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F"):
        assert normwise(g[k], o[k]) < tol, k
    assert np.abs(g["plastic_state"] - o["plastic_state"]).max() < (1e-9 if precision == A.PREC_FP64 else 1e-4) * max(
        1e-3, np.abs(o["plastic_state"]).max())


@pytest.mark.parametrize("model", [A.MODEL_FIXED_COROTATED, A.MODEL_NEO_HOOKEAN])
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_snow_plasticity(gpu, oracle, model, precision):
    # SnowPlasticity (PlasticityApplier.cpp:16-50, MpmInitializationHelper.h:883-965):
    # a pre-compressed snow cube landing on the sticky floor; singular values
    # are clamped from the first substep and the hardened moduli follow Jp
    sc = small_cube(cells=10, seed=8)
    sc.v[:, 1] = -1.0
    rng = np.random.default_rng(3)
    sc.F = (0.97 * np.eye(3)).reshape(1, 9) + 0.01 * rng.normal(size=(sc.n, 9))
    sc.materials = [A.make_material(model, 1e4, 0.3, plasticity=A.PLASTICITY_SNOW, hardening=10.0)]
    steps = 25
    g = run(gpu, sc, precision, steps).download()
    o = run(oracle, sc, A.PREC_FP64, steps).download()
    assert np.abs(o["plastic_state"] - 1.0).max() > 1e-3  # Jp moved
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F", "plastic_state"):
        assert normwise(g[k], o[k]) < tol, k
    sc.materials[0].plasticity = A.PLASTICITY_NONE
    e = run(oracle, sc, A.PREC_FP64, steps).download()
    assert normwise(e["F"], o["F"]) > 1e-4


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_calculate_dt(gpu, oracle, precision):
    sc = small_cube(cells=8)
    sc.v[:] = np.random.default_rng(12).normal(scale=0.4, size=sc.v.shape)
    g = run(gpu, sc, precision, 3)
    o = run(oracle, sc, A.PREC_FP64, 3)
    rg, ro = g.calculate_dt(1e-2), o.calculate_dt(1e-2)
    tol = 1e-12 if precision == A.PREC_FP64 else 1e-5
    assert abs(rg["max_speed"] - ro["max_speed"]) <= tol * ro["max_speed"]
    assert abs(rg["dt"] - ro["dt"]) <= tol * ro["dt"]
    for k in ("min_corner", "max_corner"):
        assert np.abs(rg[k] - ro[k]).max() <= tol * np.abs(ro[k]).max(), k
    empty = sc.make(gpu, precision)
    z3 = np.zeros((0, 3))
    empty.upload(z3, z3, np.zeros((0, 9)), np.zeros((0, 9)), np.zeros(0), np.zeros(0), np.zeros(0, np.int32))
    assert empty.calculate_dt(1e-2)["dt"] == 1e-2


def _fields(out):
    return {k: out[k] for k in ("x", "v", "C", "F", "d", "g")}


def test_reupload_after_run_matches_fresh_handle(gpu):
    # The substep that writes positions also bins them for the next substep;
    # a re-upload must discard that binning (FP32 form: bit-exact).
    sc = small_slab(n=8000)
    sim = run(gpu, sc, A.PREC_FP32, 4)
    sim.upload(sc.x, sc.v, sc.C, sc.F, sc.m, sc.V0, sc.material)
    sim.run(3, sc.dt, 0.0)
    a = _fields(sim.download())
    b = _fields(run(gpu, sc, A.PREC_FP32, 3).download())
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_material_change_after_upload_restresses(gpu):
    # New materials after the upload: the stored stress and the fixed-point
    # bounds follow them (here damage switches on, which adds the phase-field
    # quantities to P2G).
    sc = small_slab(n=8000)
    plain = [A.make_material(A.MODEL_QR_STABLE_NEO_HOOKEAN, 2e4, 0.3)]
    sim = A.Simulation(gpu, dx=sc.dx, precision=A.PREC_FP32, domain_nodes=sc.domain_nodes, gravity=sc.gravity)
    sim.set_materials(plain)
    for c in sc.colliders:
        sim.add_collider(c)
    sim.upload(sc.x, sc.v, sc.C, sc.F, sc.m, sc.V0, sc.material)
    sim.set_materials(sc.materials)
    sim.run(3, sc.dt, 0.0)
    a = _fields(sim.download())
    b = _fields(run(gpu, sc, A.PREC_FP32, 3).download())
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_material_table_must_cover_uploaded_particles(gpu):
    # a smaller table would leave particles reading an empty material slot
    sc = small_cube(cells=4)
    sc.materials = [A.make_material(A.MODEL_FIXED_COROTATED, 1e4, 0.3), A.make_material(A.MODEL_NEO_HOOKEAN, 1e4, 0.3)]
    sc.material = (np.arange(sc.n) % 2).astype(np.int32)
    sim = sc.make(gpu, A.PREC_FP32)
    with pytest.raises(A.ZmpmError) as e:
        sim.set_materials(sc.materials[:1])
    assert e.value.status == A.E_ARG
    sim.set_materials(sc.materials[::-1])


@pytest.mark.parametrize("variant", range(16))
def test_launch_config_does_not_change_results(gpu, variant):
    # everything the autotuner may change is free: the workgroup counts of the
    # grid-stride kernels and every compiled variant (P2G occupancy and block
    # order, G2P segment order, collision objects in the grid update or in a
    # launch of their own) leave the FP32 form bit-exact
    sc = small_slab(n=8000)
    a = _fields(run(gpu, sc, A.PREC_FP32, 4).download())
    sim = sc.make(gpu, A.PREC_FP32)
    sim.set_launch_config(p2g_wgs=37, grid_wgs=5, prep_wgs=3, variant=variant)
    assert sim.launch_config()["p2g_wgs"] == 37 and sim.launch_config()["variant"] == variant
    sim.run(4, sc.dt, 0.0)
    b = _fields(sim.download())
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [A.PREC_FP32, A.PREC_FP64])
@pytest.mark.parametrize("transfer", ["apic", "mls", "rpic_damped"])
def test_p2g_many_blocks_per_workgroup(gpu, precision, transfer):
    # P2G workgroups walk their blocks with the next block's cell ranges,
    # neighbour row and first particle chunk loaded ahead: one or seven
    # workgroups for ~100 blocks must give what one block per workgroup gives
    # (FP32 bit for bit; FP64 to its atomic-order bar)
    kw = {"apic": {}, "mls": dict(mls_mpm=True), "rpic_damped": dict(rpic_damping_iteration=2)}[transfer]
    sc = small_slab(n=40000)
    ref = sc.make(gpu, precision, **kw)
    ref.set_launch_config(p2g_wgs=4096)
    ref.run(4, sc.dt, 0.0)
    a = _fields(ref.download())
    for wgs in (1, 7):
        sim = sc.make(gpu, precision, **kw)
        sim.set_launch_config(p2g_wgs=wgs)
        sim.run(4, sc.dt, 0.0)
        b = _fields(sim.download())
        for k in a:
            if precision == A.PREC_FP32:
                assert np.array_equal(a[k], b[k]), (wgs, k)
            else:
                assert normwise(a[k], b[k]) < 1e-12, (wgs, k, normwise(a[k], b[k]))


def test_launch_config_rejects_unknown_variants(gpu):
    sim = small_slab(n=2000).make(gpu, A.PREC_FP32)
    with pytest.raises(A.ZmpmError):
        sim.set_launch_config(variant=16)


def test_autotune_cache_file(gpu, tmp_path):
    # the first tuning of a workload writes its line; a second handle of the same
    # workload reads it back without timing; another precision is another key
    path = str(tmp_path / "tune.txt")
    sc = small_slab(n=8000)
    s1 = sc.make(gpu, A.PREC_FP32)
    c1, hit1 = s1.autotune_cached(path, sc.dt, steps=1)
    s2 = sc.make(gpu, A.PREC_FP32)
    c2, hit2 = s2.autotune_cached(path, sc.dt, steps=1)
    assert not hit1 and hit2 and c1 == c2 and s2.launch_config() == c1
    s3 = sc.make(gpu, A.PREC_FP64)
    _, hit3 = s3.autotune_cached(path, sc.dt, steps=1)
    assert not hit3
    lines = open(path).read().splitlines()
    assert len(lines) == 2 and "gfx950" in lines[0] and "/fp32/" in lines[0] and "/fp64/" in lines[1]
    # a cached configuration leaves the results as they are
    s2.run(3, sc.dt, 0.0)
    b = _fields(s2.download())
    a = _fields(run(gpu, sc, A.PREC_FP32, 3).download())
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_autotune_restores_the_state(gpu):
    # tuning runs substeps on the live scene, then puts the state back
    sc = small_slab(n=8000)
    sim = sc.make(gpu, A.PREC_FP32)
    cfg = sim.autotune(sc.dt, 0.0, steps=2)
    assert all(v >= 1 for k, v in cfg.items() if k != "variant") and 0 <= cfg["variant"] <= 15
    sim.run(4, sc.dt, 0.0)
    a = _fields(sim.download())
    b = _fields(run(gpu, sc, A.PREC_FP32, 4).download())
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# ------------------------------------------------------- implicit damage solve
@pytest.mark.parametrize("transfer", ["apic", "pic"])
@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_implicit_damage_slab(gpu, oracle, precision, transfer):
    # AnisoFractureSimulation with implicit = true: the phase field by the
    # matrix-free CG solve (AnisoFractureSimulation.h:435-680); also under
    # FLIP_blend_PIC at ratio 0, the transfer of the AnisoMPM demos
    sc = implicit_slab()
    steps = 20
    kw = dict(implicit_damage=True)
    if transfer == "pic":
        kw.update(transfer_scheme=A.FLIP_BLEND_PIC, flip_pic_ratio=0.0)
    g = run(gpu, sc, precision, steps, **kw).download()
    os_ = run(oracle, sc, A.PREC_FP64, steps, **kw)
    o = os_.download()
    assert o["d"].max() > 0.5 and oracle.lib.zmpm_oracle_cg_iterations(os_.h) >= 1
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F", "d", "g", "max_dtilde"):
        assert normwise(g[k], o[k]) < (tol if k not in ("d", "max_dtilde") else 10 * tol), k
    assert g["laplacian"].max() == 0.0


def test_implicit_damage_solution_many_iterations(gpu, oracle):
    # V0 = 1 and eta = dt make the solve's absolute tolerance tight relative to
    # its right-hand side, so the CG runs tens of iterations; the grid's x - pf
    # and the particles' d follow the oracle
    rng = np.random.default_rng(12)
    dx = 1.0 / 32
    lo = np.full(3, 16 * dx)
    x = scenes.stratified_box(rng, lo, lo + 6 * dx, dx, 2)
    n = x.shape[0]
    F = (np.eye(3).reshape(1, 9) + 0.01 * rng.normal(size=(n, 9)))
    mt = A.make_material(A.MODEL_QR_STABLE_NEO_HOOKEAN, 1e4, 0.3, has_damage=True, alphas=(-0.9,),
                         sigma_crit=20.0, eta=1e-4, zeta=1.0, residual_phase=1e-3, l0=2 * dx)
    sc = scenes.Scene("implicit_cube", dx, 64, 1e-4, 1, (0.0, 0.0, 0.0), [mt], [], x, np.zeros(n, np.int32),
                      np.full(n, 1000.0), np.ones(n), F)
    d0 = rng.uniform(0.0, 0.6, n)
    sims = []
    for lib in (gpu, oracle):
        sim = sc.make(lib, A.PREC_FP64, implicit_damage=True)
        sim.upload(sc.x, sc.v, sc.C, sc.F, sc.m, sc.V0, sc.material, d=d0)
        sim.substep(sc.dt, 0.0)
        sims.append(sim)
    assert oracle.lib.zmpm_oracle_cg_iterations(sims[1].h) >= 10
    g, o = sims[0].grid(), sims[1].grid()
    assert np.array_equal(g["coords"], o["coords"])
    assert np.abs(o["pf"]).max() > 1e-3
    assert normwise(g["pf"], o["pf"]) < 1e-8
    gd, od = sims[0].download(), sims[1].download()
    assert (od["d"] > d0).sum() > 10
    for k in ("d", "g", "max_dtilde", "x", "v", "F"):
        assert normwise(gd[k], od[k]) < 1e-8, k


def test_implicit_damage_many_blocks(gpu, oracle):
    # more grid blocks than the solve has workgroups: its block loop strides
    sc = scenes.damage_slab(n=200_000, dx=1.0 / 128, size=(1.0, 0.2, 0.2), sigma_crit=20.0, pull_speed=0.5,
                            eta=1e-3, seed=7)
    g = run(gpu, sc, A.PREC_FP64, 3, implicit_damage=True)
    o = run(oracle, sc, A.PREC_FP64, 3, implicit_damage=True)
    assert g.binning()["groups"].shape[0] > 1024
    gd, od = g.download(), o.download()
    assert od["d"].max() > 0.0
    for k in ("x", "v", "F", "d", "g", "max_dtilde"):
        assert normwise(gd[k], od[k]) < FP64_TOL * 10, k
    gg, og = g.grid(), o.grid()
    assert normwise(gg["pf"], og["pf"]) < 1e-8


def test_invalid_options_rejected(gpu):
    # the library validates as the oracle does (tests/test_oracle.py): bad scene
    # flags at create, plasticity on a model it cannot act on at set_materials
    sc = small_cube(cells=2)
    with pytest.raises(A.ZmpmError):
        sc.make(gpu, A.PREC_FP32, implicit_damage=2)
    for model, plast in ((A.MODEL_QR_STABLE_NEO_HOOKEAN, A.PLASTICITY_SNOW),
                         (A.MODEL_EQUATION_OF_STATE, A.PLASTICITY_VON_MISES_STVK_HENCKY),
                         (A.MODEL_NEO_HOOKEAN_BORDEN, A.PLASTICITY_SNOW),
                         (A.MODEL_FIXED_COROTATED, 4), (6, A.PLASTICITY_NONE)):
        sc.materials = [A.make_material(model, plasticity=plast)]
        with pytest.raises(A.ZmpmError):
            sc.make(gpu, A.PREC_FP32)


def test_implicit_damage_without_damaged_materials(gpu):
    # no material carries the phase field: the solve is skipped and the
    # substep is the explicit one (bitwise in the deterministic FP32 form)
    sc = small_cube(cells=6, seed=4)
    a = run(gpu, sc, A.PREC_FP32, 6, implicit_damage=True).download()
    b = run(gpu, sc, A.PREC_FP32, 6).download()
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_FP32])
def test_implicit_damage_mixed_materials(gpu, oracle, precision):
    # scenes with and without the phase field in one grid: only the damaged
    # materials' particles take part in the solve; nodes reached only by the
    # others keep a zero diagonal
    sc = scenes.shared_grid_scenes(n_scenes=6, n_range=(3000, 8000), domain_nodes=256, sigma_crit=20.0)
    sc.materials[1].eta = 1e-3
    damaged = sc.material == 1
    F = sc.F.reshape(-1, 3, 3)
    F[damaged] = np.diag([1.02, 1.0, 1.0]) @ F[damaged]  # pre-stretched: the phase field grows at once
    sc.F = F.reshape(-1, 9)
    steps = 10
    g = run(gpu, sc, precision, steps, implicit_damage=True).download()
    o = run(oracle, sc, A.PREC_FP64, steps, implicit_damage=True).download()
    assert o["d"].max() > 0.01
    assert np.all(o["d"][~damaged] == 0.0)
    tol = FP64_TOL if precision == A.PREC_FP64 else FP32_TOL
    for k in ("x", "v", "F", "d", "g", "max_dtilde"):
        assert normwise(g[k], o[k]) < (tol if k not in ("d", "max_dtilde") else 10 * tol), k
