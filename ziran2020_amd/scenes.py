"""Seeded synthetic scenes for the five benchmark workloads and the tests.

Scene-building semantics follow the reference's initialisation helpers:
particle mass = rho * V / N and element measure = V / N
(MpmInitializationHelper.h:637-647, MpmParticleHandleBase.cpp:1211-1213); the
fibre of an AnisoMPM particle lives in its initial F = [a, a_perp, a x a_perp]
with a_0 = e1 (MpmParticleHandleBase.cpp:196-268); domain walls are six sticky
half-spaces at 5 dx from the faces of the addressable domain
(MpmInitializationHelper.h:1498-1513, examples/paramEffects.h:85).
Sampling uses numpy's PCG64 generator with the stated seed (the reference's
PoissonDisk sampler is not restated).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


@dataclass
class Scene:
    name: str
    dx: float
    domain_nodes: int
    dt: float
    steps: int
    gravity: tuple
    materials: list
    colliders: list
    x: np.ndarray
    material: np.ndarray
    rho: np.ndarray  # per-particle density used for mass
    V0: np.ndarray
    F: np.ndarray
    v: np.ndarray = None
    C: np.ndarray = None
    info: dict = field(default_factory=dict)
    options: dict = field(default_factory=dict)  # zmpm_scene_desc options of the workload (e.g. implicit_damage)
    t0: float = 0.0  # simulation time of the first substep
    J: np.ndarray = None  # FJ-mixed particles' J (None: 1)
    body_forces: list = field(default_factory=list)  # (scale, frequency, direction): addExternalBodyForce
    velocity_field: tuple = None  # (a, B): explicit_velocity_field v = a + B x

    def __post_init__(self):
        n = self.x.shape[0]
        if self.v is None:
            self.v = np.zeros((n, 3))
        if self.C is None:
            self.C = np.zeros((n, 9))
        self.m = self.rho * self.V0

    @property
    def n(self):
        return int(self.x.shape[0])

    def desc(self, precision=A.PREC_FP64, device=0, **transfer):
        """zmpm_scene_desc fields of the scene.  transfer: apic_rpic_ratio,
        transfer_scheme, flip_pic_ratio, mls_mpm, rpic_damping_iteration,
        pic_damping_iteration, implicit_damage; they extend and override the
        workload's own options."""
        return dict(dx=self.dx, precision=precision, domain_nodes=self.domain_nodes, gravity=self.gravity,
                    device=device, **{**self.options, **transfer})

    def make(self, lib, precision=A.PREC_FP64, device=0, **transfer):
        sim = A.Simulation(lib, **self.desc(precision, device, **transfer))
        sim.set_materials(self.materials)
        self.add_scene_setup(sim, 0, batch=False)
        sim.upload(self.x, self.v, self.C, self.F, self.m, self.V0, self.material, J=self.J)
        return sim

    def add_scene_setup(self, sim, scene, batch):
        for c in self.colliders:
            if batch:
                sim.add_scene_collider(scene, c)
            else:
                sim.add_collider(c)
        for scale, freq, direction in self.body_forces:
            sim.add_body_force(scale, freq, direction, scene=scene)
        if self.velocity_field is not None:
            sim.set_velocity_field(*self.velocity_field, scene=scene)

    def run(self, sim, steps, start=0):
        """Substeps start .. start + steps - 1 of the scene (time t0 + start dt)."""
        return sim.run(steps, self.dt, self.t0 + start * self.dt)


@dataclass
class SceneBatch:
    """Independent scenes run in one handle (zmpm_create_batch): each keeps its
    own dx, dt, start time, gravity, transfer ratio, domain and colliders; the
    material table is shared (every scene's `materials` is the same list)."""
    name: str
    scenes: list
    steps: int
    info: dict = field(default_factory=dict)

    @property
    def n(self):
        return sum(s.n for s in self.scenes)

    @property
    def materials(self):
        return self.scenes[0].materials

    @property
    def dts(self):
        return np.array([s.dt for s in self.scenes])

    @property
    def t0s(self):
        return np.array([s.t0 for s in self.scenes])

    @property
    def dt(self):
        """The smallest scene step (for reports; each scene runs its own)."""
        return float(self.dts.min())

    @property
    def scene_of(self):
        return np.concatenate([np.full(s.n, i, np.int32) for i, s in enumerate(self.scenes)])

    def cat(self, field_):
        return np.concatenate([getattr(s, field_) for s in self.scenes])

    def subset(self, keep):
        return SceneBatch(self.name, [self.scenes[i] for i in keep], self.steps,
                          {**self.info, "scenes": [self.info.get("scenes", list(range(len(self.scenes))))[i]
                                                   for i in keep]})

    def make(self, lib, precision=A.PREC_FP64, device=0, **transfer):
        assert all(s.materials is self.materials for s in self.scenes), "scenes of a batch share one material table"
        sim = A.Batch(lib, [s.desc(precision, device, **transfer) for s in self.scenes])
        sim.set_materials(self.materials)
        for i, s in enumerate(self.scenes):
            s.add_scene_setup(sim, i, batch=True)
        J = None
        if any(s.J is not None for s in self.scenes):
            J = np.concatenate([s.J if s.J is not None else np.ones(s.n) for s in self.scenes])
        sim.upload(self.cat("x"), self.cat("v"), self.cat("C"), self.cat("F"), self.cat("m"), self.cat("V0"),
                   self.cat("material"), scene=self.scene_of, J=J)
        return sim

    def run(self, sim, steps, start=0):
        return sim.run_batch(steps, self.dts, self.t0s + start * self.dts)


def domain_walls(domain_nodes, dx, type_=A.STICKY):
    """addAllWallsInDomain(domain*dx, 5*dx, type) (MpmInitializationHelper.h:1498-1513)."""
    L, off = domain_nodes * dx, 5 * dx
    out = []
    for d in range(3):
        for s in (-1, 1):
            o = [0.0, 0.0, 0.0]
            nrm = [0.0, 0.0, 0.0]
            o[d] = off if s == 1 else L - off
            nrm[d] = float(s)
            out.append(A.half_space(o, nrm, type_))
    return out


def stratified_box(rng, lo, hi, dx, ppc_axis):
    """ppc_axis^3 particles per cell, one per sub-cell, uniformly jittered inside it."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    ncell = np.round((hi - lo) / dx).astype(int)
    h = dx / ppc_axis
    nsub = ncell * ppc_axis
    g = np.stack(np.meshgrid(*[np.arange(k) for k in nsub], indexing="ij"), -1).reshape(-1, 3)
    x = lo + (g + rng.random(g.shape)) * h
    return x


def uniform_box(rng, lo, hi, n):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    return lo + rng.random((n, 3)) * (hi - lo)


def fibre_F(rng, n, max_deg=15.0):
    """Per-particle fibre a = x-hat rotated by U(-max,max) degrees about a random axis
    perpendicular to x-hat, stored as F0 = [a, a_perp, a x a_perp]
    (MpmParticleHandleBase::initializeRotatedFHelper, MpmParticleHandleBase.cpp:224-268)."""
    th = np.deg2rad(rng.uniform(-max_deg, max_deg, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    a = np.stack([np.cos(th), np.sin(th) * np.sin(ph), -np.sin(th) * np.cos(ph)], 1)
    second = np.stack([-a[:, 1], a[:, 0], np.zeros(n)], 1)
    second /= np.linalg.norm(second, axis=1, keepdims=True)
    third = np.cross(a, second)
    F = np.stack([a, second, third], 2)  # columns
    return F.reshape(n, 9)


def identity_F(n):
    return np.tile(np.eye(3).reshape(1, 9), (n, 1))


# ------------------------------------------------------------------ workloads
def elastic_cube(n_side_cells=32, dx=1.0 / 64, ppc_axis=2, seed=1, steps=100, dt=1e-4, domain_nodes=64,
                 E=1e4, nu=0.3, rho=1000.0):
    """Plain explicit APIC, FP64: elastic cube (0.5 m side at the defaults) on a
    sticky floor at y = 8 dx, fixed-corotated E=1e4 nu=0.3 rho=1000, g=-9.8,
    8 ppc stratified jitter -> 262,144 particles."""
    rng = np.random.default_rng(seed)
    lo = np.array([16 * dx, 8 * dx, 16 * dx])
    hi = lo + n_side_cells * dx
    x = stratified_box(rng, lo, hi, dx, ppc_axis)
    n = x.shape[0]
    V0 = np.full(n, (n_side_cells * dx) ** 3 / n)
    cols = domain_walls(domain_nodes, dx) + [A.half_space((0, 8 * dx, 0), (0, 1, 0), A.STICKY)]
    return Scene("elastic_cube", dx, domain_nodes, dt, steps, (0.0, -9.8, 0.0),
                 [A.make_material(A.MODEL_FIXED_COROTATED, E, nu)], cols, x,
                 np.zeros(n, np.int32), np.full(n, rho), V0, identity_F(n))


def damage_slab(n=1_000_000, dx=1.0 / 128, size=(1.0, 0.2, 0.2), seed=2, steps=200, dt=1e-4,
                domain_nodes=256, E=1e4, nu=0.3, rho=1000.0, alpha=-0.9, sigma_crit=5e3,
                eta=0.1, zeta=1.0, residual=0.001, pull_speed=0.05, model=A.MODEL_QR_STABLE_NEO_HOOKEAN,
                isotropic=False, fiber_scale=(1.0, 1.0, 1.0)):
    """AnisoMPM phase-field damage: n particles uniformly in a slab, per-particle fibre
    +-15 deg about x in F0, A = I + alpha (a x a), l0 = 2 dx, sigma_c = 5e3 Pa,
    QRStableNeoHookean; bottom face held by a fixed sticky box, top face pulled by
    a sticky box moving +y (examples/paramEffects.h:129-154).  `model` may be the
    other phase-field model, QRAnisotropic (as examples/orange.h:46-63 uses it, with
    its `isotropic` flag and fibre stiffness scales)."""
    rng = np.random.default_rng(seed)
    size = np.asarray(size, float)
    lo = np.array([16 * dx, 16 * dx, 16 * dx])
    hi = lo + size
    x = uniform_box(rng, lo, hi, n)
    V0 = np.full(n, float(np.prod(size)) / n)
    mat = A.make_material(model, E, nu, has_damage=True, directors=((1, 0, 0),),
                          alphas=(alpha,), residual_phase=residual, l0=2 * dx, sigma_crit=sigma_crit,
                          eta=eta, zeta=zeta, isotropic=isotropic, fiber_scale=fiber_scale)
    m = 4 * dx
    top = A.box((lo[0] - m, hi[1] - 2 * dx, lo[2] - m), (hi[0] + m, hi[1] + 8 * dx, hi[2] + m),
                A.STICKY, dbdt=(0.0, pull_speed, 0.0))
    bottom = A.box((lo[0] - m, lo[1] - 8 * dx, lo[2] - m), (hi[0] + m, lo[1] + 2 * dx, hi[2] + m), A.STICKY)
    cols = domain_walls(domain_nodes, dx) + [top, bottom]
    return Scene("damage_slab", dx, domain_nodes, dt, steps, (0.0, 0.0, 0.0), [mat], cols, x,
                 np.zeros(n, np.int32), np.full(n, rho), V0, fibre_F(rng, n))


def sparse_spheres(n=8_000_000, n_spheres=20, dx=1.0 / 256, seed=3, steps=100, dt=1e-4, domain_nodes=512,
                   E_range=(5e3, 5e4), nu=0.3, rho=1000.0, r_range=(0.05, 0.15)):
    """Large sparse scene: n particles uniformly inside non-overlapping spheres
    (radii U[0.05,0.15] m, centres in a 1 m^3 box), one fixed-corotated material
    per sphere with E ~ U[5e3, 5e4], gravity, sticky floor."""
    rng = np.random.default_rng(seed)
    off = 16 * dx
    cs, rs = [], []
    while len(cs) < n_spheres:
        r = rng.uniform(*r_range)
        c = off + r + rng.random(3) * (1.0 - 2 * r)
        if all(np.linalg.norm(c - c2) > r + r2 + 2 * dx for c2, r2 in zip(cs, rs)):
            cs.append(c)
            rs.append(r)
    vol = np.array([4.0 / 3.0 * np.pi * r ** 3 for r in rs])
    counts = np.floor(n * vol / vol.sum()).astype(int)
    counts[: n - counts.sum()] += 1
    xs, mats_idx, V0 = [], [], []
    for s, (c, r, k) in enumerate(zip(cs, rs, counts)):
        u = rng.normal(size=(k, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        rad = r * rng.random(k) ** (1.0 / 3.0)
        xs.append(c + u * rad[:, None])
        mats_idx.append(np.full(k, s, np.int32))
        V0.append(np.full(k, vol[s] / k))
    x = np.concatenate(xs)
    Es = rng.uniform(*E_range, n_spheres)
    mats = [A.make_material(A.MODEL_FIXED_COROTATED, float(E), nu) for E in Es]
    cols = domain_walls(domain_nodes, dx) + [A.half_space((0, off, 0), (0, 1, 0), A.STICKY)]
    nn = x.shape[0]
    return Scene("sparse_spheres", dx, domain_nodes, dt, steps, (0.0, -9.8, 0.0), mats, cols, x,
                 np.concatenate(mats_idx), np.full(nn, rho), np.concatenate(V0), identity_F(nn),
                 info=dict(radii=rs, E=list(Es)))


def batched_scenes(n_scenes=64, n_range=(10_000, 500_000), seed=4, steps=100, domain_nodes=64, E=1e4, nu=0.3,
                   rho=1000.0, sigma_crit=5e3, keep=None):
    """Batched uneven scenes in one launch per stage: n_scenes independent scenes
    (a SceneBatch), particle counts log-uniform over n_range; scene s has its own
    dx (1/96, 1/128 or 1/160 by s mod 3), dt (8e-5 .. 1.25e-4), start time,
    gravity (-9.8 m/s^2 tilted up to 15 degrees), APIC/RPIC ratio, initial
    velocity and colliders: the six sticky domain walls at 5 dx, a floor at
    10 dx that is sticky, slipping with friction, separating or frictionless
    slip (s mod 4), and on some scenes a moving sphere, a rotating box or a
    growing capped cylinder.  Even scenes are QRStableNeoHookean with AnisoMPM
    damage (l0 = 2 dx of the scene, fibres +-15 degrees about x), odd scenes
    fixed-corotated or Neo-Hookean; ~8 particles per cell in a cube of at most
    40 cells.  keep: indices of the scenes to place (the others are drawn and
    dropped, so every kept scene is exactly its copy in the full batch)."""
    rng = np.random.default_rng(seed)
    dxs = (1.0 / 96, 1.0 / 128, 1.0 / 160)
    mats = []
    for dx in dxs:  # material 3k + {0, 1, 2} for the scenes of dx option k
        mats += [A.make_material(A.MODEL_FIXED_COROTATED, E, nu),
                 A.make_material(A.MODEL_QR_STABLE_NEO_HOOKEAN, E, nu, has_damage=True, alphas=(-0.9,),
                                 l0=2 * dx, sigma_crit=sigma_crit),
                 A.make_material(A.MODEL_NEO_HOOKEAN, 2 * E, 0.35)]
    counts = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_scenes)).astype(int)
    out, kept = [], []
    for s, k in enumerate(counts):
        opt = s % 3
        dx = dxs[opt]
        dt = float(rng.uniform(0.8e-4, 1.25e-4))
        t0 = float(rng.uniform(0.0, 0.05))
        tilt, az = np.deg2rad(rng.uniform(0, 15)), rng.uniform(0, 2 * np.pi)
        gravity = (9.8 * np.sin(tilt) * np.cos(az), -9.8 * np.cos(tilt), 9.8 * np.sin(tilt) * np.sin(az))
        rpic = float(rng.choice([1.0, 0.5, 0.0]))
        cells = min((k / 8.0) ** (1.0 / 3.0), 40.0)
        side = cells * dx
        lo = np.array([12 * dx, 10 * dx, 12 * dx]) + rng.uniform(0, 1, 3) * np.array([4, 2, 4]) * dx
        xk = uniform_box(rng, lo, lo + side, k)
        v0 = rng.uniform(-0.3, 0.3, 3)
        damaged = (s % 2 == 0)
        Fk = fibre_F(rng, k) if damaged else identity_F(k)
        mat = 3 * opt + (1 if damaged else int(rng.integers(0, 2)) * 2)
        floor_type, friction = [(A.STICKY, 0.0), (A.SLIP, 0.3), (A.SEPARATE, 0.1), (A.SLIP, 0.0)][s % 4]
        cols = domain_walls(domain_nodes, dx) + [A.half_space((0, 10 * dx, 0), (0, 1, 0), floor_type,
                                                              friction=friction)]
        top = lo + side
        ctr = lo + side / 2
        if s % 5 == 0:  # a sphere falling onto the top face
            r = 0.3 * side
            cols.append(A.sphere((0, 0, 0), r, A.SLIP, b0=(ctr[0], top[1] + r + 2 * dx, ctr[2]),
                                 dbdt=(0.0, -2.0, 0.0)))
        elif s % 5 == 1:  # a box spinning about y, pressing on one edge
            h = 0.2 * side
            cols.append(A.box((-h, -h, -h), (h, h, h), A.STICKY, b0=(lo[0], top[1], ctr[2]), omega=(0.0, 4.0, 0.0)))
        elif s % 5 == 2:  # a growing capped cylinder beside the cube
            cols.append(A.capped_cylinder(2 * dx, side, center=(0, 0, 0), type_=A.SEPARATE,
                                          b0=(top[0] + 4 * dx, ctr[1], ctr[2]), scale0=1.0, dsdt=5.0))
        if keep is not None and s not in keep:
            continue
        kept.append(s)
        n = xk.shape[0]
        V0 = np.full(n, side ** 3 / n)
        out.append(Scene(f"scene{s}", dx, domain_nodes, dt, steps, gravity, mats, cols, xk,
                         np.full(n, mat, np.int32), np.full(n, rho), V0, Fk, v=np.tile(v0, (n, 1)),
                         options=dict(apic_rpic_ratio=rpic), t0=t0))
    return SceneBatch("batched_scenes", out, steps, info=dict(counts=counts.tolist(), scenes=kept))


def shared_grid_scenes(n_scenes=64, n_range=(10_000, 500_000), dx=1.0 / 128, seed=4, steps=100, dt=1e-4,
                   domain_nodes=512, tile_cells=48, E=1e4, nu=0.3, rho=1000.0, sigma_crit=5e3, keep=None):
    """Uneven cubes in one world grid (one scene, one dx and dt): n_scenes cubes
    with particle counts log-uniform over n_range, each in its own tile on a
    shared sticky floor; odd cubes fixed-corotated, even cubes QRStableNeoHookean
    with AnisoMPM damage.  keep: indices of the cubes to place (the others are
    drawn and dropped, so every kept cube is exactly its copy in the full set)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_scenes)).astype(int)
    per_row = int(np.ceil(np.sqrt(n_scenes)))
    off = 16 * dx
    xs, mid, V0, Fs = [], [], [], []
    mats = [A.make_material(A.MODEL_FIXED_COROTATED, E, nu),
            A.make_material(A.MODEL_QR_STABLE_NEO_HOOKEAN, E, nu, has_damage=True, alphas=(-0.9,),
                            l0=2 * dx, sigma_crit=sigma_crit)]
    kept = []
    for s, k in enumerate(counts):
        side = (k / 8.0) ** (1.0 / 3.0) * dx  # ~8 particles per cell
        side = min(side, (tile_cells - 8) * dx)
        ix, iz = s % per_row, s // per_row
        lo = np.array([off + ix * tile_cells * dx, off, off + iz * tile_cells * dx])
        xk = uniform_box(rng, lo, lo + side, k)
        damaged = (s % 2 == 0)
        Fk = fibre_F(rng, k) if damaged else identity_F(k)
        if keep is not None and s not in keep:
            continue
        kept.append(s)
        xs.append(xk)
        mid.append(np.full(k, 1 if damaged else 0, np.int32))
        V0.append(np.full(k, side ** 3 / k))
        Fs.append(Fk)
    x = np.concatenate(xs)
    n = x.shape[0]
    cols = domain_walls(domain_nodes, dx) + [A.half_space((0, off, 0), (0, 1, 0), A.STICKY)]
    return Scene("shared_grid_scenes", dx, domain_nodes, dt, steps, (0.0, -9.8, 0.0), mats, cols, x,
                 np.concatenate(mid), np.full(n, rho), np.concatenate(V0), np.concatenate(Fs),
                 info=dict(counts=counts.tolist(), scenes=kept))


def two_material(n=16_000_000, dx=1.0 / 256, seed=5, steps=100, dt=1e-4, domain_nodes=512,
                 E=1e5, nu=0.4, bulk=1e4, gamma=7.0, rho_solid=1000.0, rho_fluid=1000.0):
    """Neo-Hookean block (8 ppc) inside a weakly compressible J-based EOS fluid column
    (2 ppc), n particles in total."""
    rng = np.random.default_rng(seed)
    off = 16 * dx
    col_lo = np.array([off, off, off])
    col_size = np.array([0.9, 0.8, 0.9])
    blk_size = np.array([0.25, 0.25, 0.25])
    blk_lo = col_lo + np.array([0.325, 0.3, 0.325])
    v_col = float(np.prod(col_size)) - float(np.prod(blk_size))
    v_blk = float(np.prod(blk_size))
    # 2 ppc fluid, 8 ppc solid, then scaled to hit n exactly
    n_f0 = 2 * v_col / dx ** 3
    n_s0 = 8 * v_blk / dx ** 3
    n_s = int(round(n * n_s0 / (n_f0 + n_s0)))
    n_f = n - n_s
    xs = uniform_box(rng, blk_lo, blk_lo + blk_size, n_s)
    xf = np.empty((0, 3))
    while xf.shape[0] < n_f:
        cand = uniform_box(rng, col_lo, col_lo + col_size, int(1.3 * (n_f - xf.shape[0])) + 16)
        inside = np.all((cand >= blk_lo) & (cand <= blk_lo + blk_size), axis=1)
        xf = np.concatenate([xf, cand[~inside]])
    xf = xf[:n_f]
    x = np.concatenate([xs, xf])
    material = np.concatenate([np.zeros(n_s, np.int32), np.ones(n_f, np.int32)])
    V0 = np.concatenate([np.full(n_s, v_blk / n_s), np.full(n_f, v_col / n_f)])
    rho = np.concatenate([np.full(n_s, rho_solid), np.full(n_f, rho_fluid)])
    F = identity_F(n)
    F[n_s:, 0] = 1.0  # J = 1 for EOS particles
    mats = [A.make_material(A.MODEL_NEO_HOOKEAN, E, nu),
            A.make_material(A.MODEL_EQUATION_OF_STATE, 1.0, 0.0, bulk=bulk, gamma=gamma)]
    cols = domain_walls(domain_nodes, dx) + [A.half_space((0, off, 0), (0, 1, 0), A.STICKY)]
    return Scene("two_material", dx, domain_nodes, dt, steps, (0.0, -9.8, 0.0), mats, cols, x, material,
                 rho, V0, F, info=dict(n_solid=n_s, n_fluid=n_f))


def damage_slab_active(**kw):
    """The damage slab with sigma_c = 20 Pa: the phase field grows from the first
    substeps, so every substep runs the full damage update (sigma^+ and the
    structural tensor) for the loaded particles."""
    kw.setdefault("sigma_crit", 20.0)
    return damage_slab(**kw)


def damage_slab_implicit(**kw):
    """The active damage slab (sigma_c = 20 Pa) with the implicit phase-field solve
    (AnisoFractureSimulation(true), examples/3DtestImplicit.h)."""
    sc = damage_slab_active(**kw)
    sc.name = "damage_slab_implicit"
    sc.options = dict(implicit_damage=True)
    return sc


WORKLOADS = {
    "elastic_cube": elastic_cube,
    "damage_slab": damage_slab,
    "damage_slab_active": damage_slab_active,
    "damage_slab_implicit": damage_slab_implicit,
    "sparse_spheres": sparse_spheres,
    "batched_scenes": batched_scenes,
    "shared_grid_scenes": shared_grid_scenes,
    "two_material": two_material,
}
