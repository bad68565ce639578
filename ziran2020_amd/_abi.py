"""ctypes mirror of include/zmpm.h.

The same declarations bind any shared library that exports the zmpm C ABI.
The package itself only ever loads the HIP library (``libzmpm.so``); tests may
bind a second library through :class:`ZmpmLib` to compare against it.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

ZMPM_ABI_VERSION = 8

# zmpm_status
OK, E_ARG, E_HIP, E_DOMAIN, E_CFL, E_NAN, E_OVERFLOW, E_NOMEM = range(8)
STATUS_NAMES = {
    OK: "ZMPM_OK", E_ARG: "ZMPM_E_ARG", E_HIP: "ZMPM_E_HIP", E_DOMAIN: "ZMPM_E_DOMAIN",
    E_CFL: "ZMPM_E_CFL", E_NAN: "ZMPM_E_NAN", E_OVERFLOW: "ZMPM_E_OVERFLOW",
    E_NOMEM: "ZMPM_E_NOMEM",
}

PREC_FP64, PREC_FP32 = 0, 1
MODEL_FIXED_COROTATED, MODEL_QR_STABLE_NEO_HOOKEAN, MODEL_NEO_HOOKEAN, MODEL_EQUATION_OF_STATE, \
    MODEL_QR_ANISOTROPIC, MODEL_NEO_HOOKEAN_BORDEN = range(6)
STICKY, SLIP, SEPARATE, GHOST = 1, 2, 3, 4
APIC_BLEND_RPIC, FLIP_BLEND_PIC = 0, 1
PLASTICITY_NONE, PLASTICITY_VON_MISES_STVK_HENCKY, PLASTICITY_DRUCKER_PRAGER_STVK_HENCKY, PLASTICITY_SNOW = 0, 1, 2, 3
SHAPE_HALF_SPACE, SHAPE_BOX, SHAPE_SPHERE, SHAPE_CAPPED_CYLINDER = 0, 1, 2, 3
FJ_NONE, FJ_SOLID, FJ_FLUID = 0, 1, 2
STAGE_SORT, STAGE_P2G, STAGE_GRID_UPDATE, STAGE_DAMAGE_SOLVE, STAGE_G2P = range(5)
SOLVE_F, SOLVE_J = 0, 1

D3 = C.c_double * 3


class SceneDesc(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("domain_nodes", C.c_int32),
        ("dx", C.c_double),
        ("gravity", D3),
        ("apic_rpic_ratio", C.c_double),
        ("cfl", C.c_double),
        ("device", C.c_int32),
        ("transfer_scheme", C.c_int32),
        ("flip_pic_ratio", C.c_double),
        ("mls_mpm", C.c_int32),
        ("rpic_damping_iteration", C.c_int32),
        ("pic_damping_iteration", C.c_int32),
        ("implicit_damage", C.c_int32),
        ("ignore_collision_objects", C.c_int32),
    ]


class Material(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("has_damage", C.c_int32),
        ("mu", C.c_double), ("lam", C.c_double),
        ("kx", C.c_double), ("ky", C.c_double),
        ("bulk", C.c_double), ("gamma", C.c_double),
        ("num_directors", C.c_int32),
        ("plasticity", C.c_int32),
        ("a0", D3 * 2),
        ("alphas", C.c_double * 2),
        ("residual_phase", C.c_double),
        ("l0", C.c_double), ("sigma_crit", C.c_double), ("eta", C.c_double), ("zeta", C.c_double),
        ("yield_stress", C.c_double), ("hardening", C.c_double),
        ("friction_angle", C.c_double), ("cohesion", C.c_double),
        ("volume_correction", C.c_int32), ("isotropic", C.c_int32),
        ("critical_compression", C.c_double), ("critical_stretch", C.c_double),
        ("min_jp", C.c_double), ("max_jp", C.c_double),
        ("fiber_stiffness", D3),
        ("fj_mixed", C.c_int32), ("j_lambda", C.c_double),
    ]


class Collider(C.Structure):
    _fields_ = [
        ("shape", C.c_int32),
        ("type", C.c_int32),
        ("origin", D3), ("normal", D3),
        ("box_min", D3), ("box_max", D3),
        ("b0", D3), ("dbdt", D3),
        ("friction", C.c_double),
        ("radius", C.c_double), ("height", C.c_double),
        ("shape_rotation", C.c_double * 4), ("rotation", C.c_double * 4),
        ("omega", D3), ("scale0", C.c_double), ("dsdt", C.c_double),
    ]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class Particles(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("x", _DP), ("v", _DP), ("C", _DP), ("F", _DP),
        ("m", _DP), ("V0", _DP), ("material", _IP),
        ("d", _DP), ("g", _DP), ("laplacian", _DP), ("plastic_state", _DP), ("max_dtilde", _DP),
        ("J", _DP),
    ]


def _ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class ZmpmError(RuntimeError):
    def __init__(self, status: int, where: str, message: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{where}: {name}" + (f" ({message})" if message else ""))


def make_material(model=MODEL_FIXED_COROTATED, E=1e4, nu=0.3, *, has_damage=False, kx=0.0, ky=0.0,
                  bulk=0.0, gamma=7.0, directors=((1.0, 0.0, 0.0),), alphas=(0.0,),
                  residual_phase=0.001, l0=0.0, sigma_crit=1.0, eta=0.1, zeta=1.0,
                  plasticity=PLASTICITY_NONE, yield_stress=0.0, hardening=0.0, friction_angle=30.0, cohesion=0.0,
                  volume_correction=True, critical_compression=2e-2, critical_stretch=7.5e-3, min_jp=0.6,
                  max_jp=20.0, isotropic=False, fiber_scale=(1.0, 1.0, 1.0), fj_mixed=FJ_NONE,
                  j_E=None, j_nu=None) -> Material:
    """Material from Young's modulus and Poisson ratio with the reference's Lame
    conversion (CorotatedElasticity.cpp:15-20, QRStableNeoHookean.h:61-65).  The snow
    plasticity defaults are SnowPlasticity's constructor defaults (PlasticityApplier.h:103;
    its psi is `hardening`).  QRAnisotropic's fibre stiffnesses are mu times
    `fiber_scale` (setParameters then scaleFiberStiffness, QRAnisotropic.h:57-64,
    96-104, which leaves an isotropic model's fibres at mu).  fj_mixed (FJ_SOLID /
    FJ_FLUID) makes it FJMixedMpmForceHelper<model, QuadraticVolumePenalty(j_E, j_nu)>
    (addFJMixedMpmForce, MpmParticleHandleBase.cpp:811-834); j_E, j_nu default to E, nu."""
    m = Material()
    m.model = model
    m.has_damage = 1 if has_damage else 0
    m.lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    m.mu = E / (2.0 * (1.0 + nu))
    m.kx, m.ky, m.bulk, m.gamma = kx, ky, bulk, gamma
    m.num_directors = len(directors)
    for k, a in enumerate(directors):
        for c in range(3):
            m.a0[k][c] = a[c]
    for k, a in enumerate(alphas):
        m.alphas[k] = a
    m.residual_phase, m.l0, m.sigma_crit, m.eta, m.zeta = residual_phase, l0, sigma_crit, eta, zeta
    m.plasticity, m.yield_stress, m.hardening = plasticity, yield_stress, hardening
    m.friction_angle, m.cohesion, m.volume_correction = friction_angle, cohesion, int(bool(volume_correction))
    m.critical_compression, m.critical_stretch, m.min_jp, m.max_jp = critical_compression, critical_stretch, min_jp, max_jp
    m.isotropic = 1 if isotropic else 0
    for c in range(3):
        m.fiber_stiffness[c] = m.mu * (1.0 if isotropic else fiber_scale[c])
    m.fj_mixed = fj_mixed
    jE = E if j_E is None else j_E
    jnu = nu if j_nu is None else j_nu
    m.j_lambda = jE * jnu / ((1.0 + jnu) * (1.0 - 2.0 * jnu))  # QuadraticVolumePenalty::setMaterialParameters
    return m


def sigma_crit_from_percentage(m: Material, percentage: float) -> float:
    """The critical stress the reference derives from a stretch percentage when it
    attaches the phase field: the strain energy density of the undegraded,
    fibre-free (QR-NH) or fibre-isotropic (QR-aniso) model at F = (1 + p) I
    (MpmParticleHandleBase.cpp:175-222; QRStableNeoHookean::getSigmaCrit,
    QRStableNeoHookean.h:66-78; QRAnisotropic::getSigmaCrit, QRAnisotropic.h:65-77),
    or NeoHookeanBorden's energy there at g = 1 (NeoHookeanBorden.cpp:36-43, whose
    deviatoric part vanishes at F = s I).  At F = s I the QR factors are Q = I,
    R = s I and J = s^3."""
    s = 1.0 + percentage
    J = s ** 3
    vol = 0.5 * m.lam * (J - 1.0) ** 2
    if m.model == MODEL_QR_STABLE_NEO_HOOKEAN:
        return vol + m.mu / 2.0 * (3.0 * s * s - 3.0) - m.mu * (J - 1.0)
    if m.model == MODEL_QR_ANISOTROPIC:
        return vol + 3.0 * 0.5 * m.mu * (s - 1.0) ** 2
    if m.model == MODEL_NEO_HOOKEAN_BORDEN:
        kappa = 0.666666 * m.mu + m.lam
        return kappa / 2.0 * ((J * J - 1.0) / 2.0 - math.log(J))
    raise ValueError("sigma_crit_from_percentage: only the phase-field models define getSigmaCrit")


def _motion(c, b0, dbdt, rotation, omega, scale0, dsdt):
    """Rigid motion of a collision object: translation b0 + dbdt t, rotation
    q(t) = (cos(|w| t/2), sin(|w| t/2) w/|w|) * rotation, scaling scale0 + dsdt t."""
    for a in range(3):
        c.b0[a], c.dbdt[a], c.omega[a] = b0[a], dbdt[a], omega[a]
    for a in range(4):
        c.rotation[a] = rotation[a]
    c.scale0, c.dsdt = scale0, dsdt
    return c


def half_space(origin, normal, type_=STICKY, friction=0.0, b0=(0, 0, 0), dbdt=(0, 0, 0),
               rotation=(1, 0, 0, 0), omega=(0, 0, 0), scale0=1.0, dsdt=0.0) -> Collider:
    c = Collider()
    c.shape, c.type, c.friction = SHAPE_HALF_SPACE, type_, friction
    for a in range(3):
        c.origin[a], c.normal[a] = origin[a], normal[a]
    return _motion(c, b0, dbdt, rotation, omega, scale0, dsdt)


def box(box_min, box_max, type_=STICKY, b0=(0, 0, 0), dbdt=(0, 0, 0), friction=0.0,
        rotation=(1, 0, 0, 0), omega=(0, 0, 0), scale0=1.0, dsdt=0.0) -> Collider:
    c = Collider()
    c.shape, c.type, c.friction = SHAPE_BOX, type_, friction
    for a in range(3):
        c.box_min[a], c.box_max[a] = box_min[a], box_max[a]
    return _motion(c, b0, dbdt, rotation, omega, scale0, dsdt)


def sphere(center, radius, type_=STICKY, b0=(0, 0, 0), dbdt=(0, 0, 0), friction=0.0,
           rotation=(1, 0, 0, 0), omega=(0, 0, 0), scale0=1.0, dsdt=0.0) -> Collider:
    c = Collider()
    c.shape, c.type, c.friction, c.radius = SHAPE_SPHERE, type_, friction, radius
    for a in range(3):
        c.origin[a] = center[a]
    return _motion(c, b0, dbdt, rotation, omega, scale0, dsdt)


def capped_cylinder(radius, height, shape_rotation=(1, 0, 0, 0), center=(0, 0, 0), type_=STICKY, b0=(0, 0, 0),
                    dbdt=(0, 0, 0), friction=0.0, rotation=(1, 0, 0, 0), omega=(0, 0, 0), scale0=1.0,
                    dsdt=0.0) -> Collider:
    """CappedCylinder(radius, height, q, b) (AnalyticLevelSet.h:252-340): axis along the
    primitive's y, its own rotation quaternion q = (w, x, y, z) and centre b."""
    c = Collider()
    c.shape, c.type, c.friction, c.radius, c.height = SHAPE_CAPPED_CYLINDER, type_, friction, radius, height
    for a in range(3):
        c.origin[a] = center[a]
    for a in range(4):
        c.shape_rotation[a] = shape_rotation[a]
    return _motion(c, b0, dbdt, rotation, omega, scale0, dsdt)


class LaunchConfig(C.Structure):
    """zmpm_launch_config (include/zmpm.h)."""
    _fields_ = [("p2g_wgs", C.c_int32), ("grid_wgs", C.c_int32),
                ("prep_wgs", C.c_int32), ("p2g_split", C.c_int32), ("variant", C.c_int32),
                ("reserved", C.c_int32 * 3)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in ("p2g_wgs", "grid_wgs", "prep_wgs", "p2g_split", "variant")}


VARIANT_P2G_LOW_OCC, VARIANT_G2P_RR, VARIANT_P2G_XCD, VARIANT_GRID_SPLIT = 1, 2, 4, 8


class ZmpmLib:
    """Bindings of one shared library exporting the zmpm C ABI."""

    EXPORTS = (
        "zmpm_abi_version", "zmpm_build_info", "zmpm_create", "zmpm_destroy", "zmpm_last_error",
        "zmpm_set_materials", "zmpm_add_collider", "zmpm_clear_colliders", "zmpm_upload",
        "zmpm_download", "zmpm_substep", "zmpm_run", "zmpm_synchronize", "zmpm_get_binning",
        "zmpm_get_grid", "zmpm_set_timing", "zmpm_get_timings", "zmpm_calculate_dt", "zmpm_get_launch_config",
        "zmpm_set_launch_config", "zmpm_autotune", "zmpm_create_batch", "zmpm_batch_origin",
        "zmpm_add_scene_collider", "zmpm_upload_batch", "zmpm_run_batch", "zmpm_calculate_dt_noalloc",
        "zmpm_set_fj_state", "zmpm_add_body_force", "zmpm_set_velocity_field", "zmpm_stage",
        "zmpm_set_grid_velocity", "zmpm_autotune_cached",
    )

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = C.CDLL(path)
        L = self.lib
        L.zmpm_abi_version.restype = C.c_int32
        L.zmpm_build_info.restype = C.c_char_p
        L.zmpm_last_error.restype = C.c_char_p
        L.zmpm_last_error.argtypes = [C.c_void_p]
        for name in self.EXPORTS:
            fn = getattr(L, name)
            if name not in ("zmpm_abi_version", "zmpm_build_info", "zmpm_last_error"):
                fn.restype = C.c_int32
        L.zmpm_create.argtypes = [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]
        L.zmpm_destroy.argtypes = [C.c_void_p]
        L.zmpm_set_materials.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Material)]
        L.zmpm_add_collider.argtypes = [C.c_void_p, C.POINTER(Collider), C.POINTER(C.c_int32)]
        L.zmpm_clear_colliders.argtypes = [C.c_void_p]
        L.zmpm_upload.argtypes = [C.c_void_p, C.POINTER(Particles)]
        L.zmpm_download.argtypes = [C.c_void_p, C.POINTER(Particles)]
        L.zmpm_substep.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.zmpm_run.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double]
        L.zmpm_synchronize.argtypes = [C.c_void_p]
        L.zmpm_get_binning.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
        L.zmpm_get_grid.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), _DP, _DP, _DP, _DP]
        L.zmpm_set_timing.argtypes = [C.c_void_p, C.c_int32]
        L.zmpm_get_timings.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_char_p), _DP]
        L.zmpm_calculate_dt.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.zmpm_get_launch_config.argtypes = [C.c_void_p, C.POINTER(LaunchConfig)]
        L.zmpm_set_launch_config.argtypes = [C.c_void_p, C.POINTER(LaunchConfig)]
        L.zmpm_autotune.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                    C.POINTER(LaunchConfig)]
        L.zmpm_create_batch.argtypes = [C.POINTER(SceneDesc), C.c_int32, C.POINTER(C.c_void_p)]
        L.zmpm_batch_origin.argtypes = [C.c_void_p, C.c_int32, _IP]
        L.zmpm_add_scene_collider.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Collider), C.POINTER(C.c_int32)]
        L.zmpm_upload_batch.argtypes = [C.c_void_p, C.POINTER(Particles), _IP]
        L.zmpm_run_batch.argtypes = [C.c_void_p, C.c_int32, _DP, _DP]
        L.zmpm_calculate_dt_noalloc.argtypes = [C.c_void_p, C.c_double, C.c_int32, _DP]
        L.zmpm_set_fj_state.argtypes = [C.c_void_p, C.c_int32]
        L.zmpm_add_body_force.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, _DP]
        L.zmpm_set_velocity_field.argtypes = [C.c_void_p, C.c_int32, _DP, _DP]
        L.zmpm_stage.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double]
        L.zmpm_set_grid_velocity.argtypes = [C.c_void_p, C.c_int64, _DP]
        L.zmpm_autotune_cached.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_char_p,
                                           C.POINTER(LaunchConfig), C.POINTER(C.c_int32)]
        if L.zmpm_abi_version() != ZMPM_ABI_VERSION:
            raise RuntimeError(f"{path}: ABI version {L.zmpm_abi_version()} != {ZMPM_ABI_VERSION}")

    @property
    def build_info(self) -> str:
        return self.lib.zmpm_build_info().decode()


def scene_desc(*, dx, precision=PREC_FP64, domain_nodes=256, gravity=(0, 0, 0), apic_rpic_ratio=1.0, cfl=0.6,
               device=0, transfer_scheme=APIC_BLEND_RPIC, flip_pic_ratio=0.0, mls_mpm=False, rpic_damping_iteration=0,
               pic_damping_iteration=0, implicit_damage=False, ignore_collision_objects=False) -> SceneDesc:
    d = SceneDesc()
    d.precision, d.domain_nodes, d.dx = precision, domain_nodes, dx
    for a in range(3):
        d.gravity[a] = gravity[a]
    d.apic_rpic_ratio, d.cfl, d.device = apic_rpic_ratio, cfl, device
    d.transfer_scheme, d.flip_pic_ratio, d.mls_mpm = transfer_scheme, flip_pic_ratio, int(bool(mls_mpm))
    d.rpic_damping_iteration, d.pic_damping_iteration = rpic_damping_iteration, pic_damping_iteration
    d.implicit_damage = int(implicit_damage)
    d.ignore_collision_objects = int(bool(ignore_collision_objects))
    return d


class Simulation:
    """One zmpm handle: scene, materials, colliders and particle state."""

    def __init__(self, lib: ZmpmLib, **desc):
        self.lib = lib
        self.L = lib.lib
        d = scene_desc(**desc)
        self.desc = d
        h = C.c_void_p()
        self._check(self.L.zmpm_create(C.byref(d), C.byref(h)), "zmpm_create", handle=False)
        self.h = h
        self.n = 0

    def _check(self, st, where, handle=True):
        if st != OK:
            msg = self.L.zmpm_last_error(self.h).decode() if handle else ""
            raise ZmpmError(st, where, msg)
        return st

    def close(self):
        if getattr(self, "h", None):
            self.L.zmpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_materials(self, mats):
        arr = (Material * len(mats))(*mats)
        self._check(self.L.zmpm_set_materials(self.h, len(mats), arr), "zmpm_set_materials")

    def add_body_force(self, scale, frequency, direction, scene=0):
        """addExternalBodyForce: f(t) = scale (sin(frequency t) + 1) direction."""
        d = np.ascontiguousarray(direction, np.float64)
        self._check(self.L.zmpm_add_body_force(self.h, scene, scale, frequency, _ptr(d)), "zmpm_add_body_force")

    def set_velocity_field(self, a=None, B=None, scene=0):
        """explicit_velocity_field as v(x) = a + B x; a=None removes it."""
        av = None if a is None else np.ascontiguousarray(a, np.float64)
        Bv = None if B is None else np.ascontiguousarray(B, np.float64).reshape(9)
        self._check(self.L.zmpm_set_velocity_field(self.h, scene, _ptr(av), _ptr(Bv)), "zmpm_set_velocity_field")

    def set_fj_state(self, state):
        """FJMixedMpmForceHelper::state: SOLVE_F (0) or SOLVE_J (1)."""
        self._check(self.L.zmpm_set_fj_state(self.h, state), "zmpm_set_fj_state")

    def add_collider(self, c: Collider) -> int:
        i = C.c_int32()
        self._check(self.L.zmpm_add_collider(self.h, C.byref(c), C.byref(i)), "zmpm_add_collider")
        return i.value

    def upload(self, x, v, Cm, F, m, V0, material, d=None, g=None, plastic_state=None, max_dtilde=None,
               scene=None, J=None):
        n = int(x.shape[0])
        arrs = dict(
            x=np.ascontiguousarray(x, np.float64).reshape(n, 3),
            v=np.ascontiguousarray(v, np.float64).reshape(n, 3),
            C=np.ascontiguousarray(Cm, np.float64).reshape(n, 9),
            F=np.ascontiguousarray(F, np.float64).reshape(n, 9),
            m=np.ascontiguousarray(m, np.float64),
            V0=np.ascontiguousarray(V0, np.float64),
            material=np.ascontiguousarray(material, np.int32),
            d=None if d is None else np.ascontiguousarray(d, np.float64),
            g=None if g is None else np.ascontiguousarray(g, np.float64),
            plastic_state=None if plastic_state is None else np.ascontiguousarray(plastic_state, np.float64),
            max_dtilde=None if max_dtilde is None else np.ascontiguousarray(max_dtilde, np.float64),
            J=None if J is None else np.ascontiguousarray(J, np.float64),
        )
        p = Particles()
        p.n = n
        for k in ("x", "v", "C", "F", "m", "V0", "d", "g", "plastic_state", "max_dtilde", "J"):
            setattr(p, k, _ptr(arrs[k]))
        p.material = _ptr(arrs["material"], C.c_int32)
        p.laplacian = None
        self._keep = arrs
        if scene is None:
            self._check(self.L.zmpm_upload(self.h, C.byref(p)), "zmpm_upload")
        else:
            arrs["scene"] = np.ascontiguousarray(scene, np.int32)
            self._check(self.L.zmpm_upload_batch(self.h, C.byref(p), _ptr(arrs["scene"], C.c_int32)),
                        "zmpm_upload_batch")
        self.n = n

    def download(self, check=True):
        n = self.n
        out = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), C=np.zeros((n, 9)), F=np.zeros((n, 9)),
                   m=np.zeros(n), V0=np.zeros(n), material=np.zeros(n, np.int32), d=np.zeros(n),
                   g=np.zeros(n), laplacian=np.zeros(n), plastic_state=np.zeros(n), max_dtilde=np.zeros(n),
                   J=np.zeros(n))
        p = Particles()
        p.n = n
        for k in ("x", "v", "C", "F", "m", "V0", "d", "g", "laplacian", "plastic_state", "max_dtilde", "J"):
            setattr(p, k, _ptr(out[k]))
        p.material = _ptr(out["material"], C.c_int32)
        st = self.L.zmpm_download(self.h, C.byref(p))
        out["status"] = st
        if check and st != OK:
            self._check(st, "zmpm_download")
        return out

    def substep(self, dt, t):
        return self._check(self.L.zmpm_substep(self.h, dt, t), "zmpm_substep")

    def stage(self, stage, dt, t):
        """One stage of the substep at t (zmpm_stage): STAGE_SORT .. STAGE_G2P in order."""
        return self._check(self.L.zmpm_stage(self.h, stage, dt, t), "zmpm_stage")

    def set_grid_velocity(self, new_v):
        """Replace new_v of the grid nodes zmpm_get_grid reports (between GRID_UPDATE and G2P)."""
        nv = np.ascontiguousarray(new_v, np.float64).reshape(-1, 3)
        self._check(self.L.zmpm_set_grid_velocity(self.h, nv.shape[0], _ptr(nv)), "zmpm_set_grid_velocity")

    def run(self, steps, dt, t0=0.0):
        return self._check(self.L.zmpm_run(self.h, steps, dt, t0), "zmpm_run")

    def synchronize(self, check=True):
        st = self.L.zmpm_synchronize(self.h)
        if check:
            self._check(st, "zmpm_synchronize")
        return st

    def binning(self):
        G = C.c_int64()
        self._check(self.L.zmpm_get_binning(self.h, C.byref(G), None, None, None, None), "zmpm_get_binning")
        n, g = self.n, G.value
        keys = np.zeros(n, np.uint64)
        order = np.zeros(n, np.int32)
        groups = np.zeros((g, 2), np.int32)
        blocks = np.zeros(g, np.uint64)
        self._check(self.L.zmpm_get_binning(self.h, C.byref(G), _ptr(keys, C.c_uint64), _ptr(order, C.c_int32),
                                            _ptr(groups, C.c_int32), _ptr(blocks, C.c_uint64)), "zmpm_get_binning")
        return dict(keys=keys, order=order, groups=groups, block_offset=blocks)

    def grid(self):
        N = C.c_int64()
        self._check(self.L.zmpm_get_grid(self.h, C.byref(N), None, None, None, None, None), "zmpm_get_grid")
        k = N.value
        out = dict(coords=np.zeros((k, 3), np.int32), m=np.zeros(k), v=np.zeros((k, 3)),
                   new_v=np.zeros((k, 3)), pf=np.zeros(k))
        self._check(self.L.zmpm_get_grid(self.h, C.byref(N), _ptr(out["coords"], C.c_int32), _ptr(out["m"]),
                                         _ptr(out["v"]), _ptr(out["new_v"]), _ptr(out["pf"])), "zmpm_get_grid")
        return out

    def calculate_dt(self, max_dt):
        """MpmSimulationBase::calculateDt: CFL step from the largest particle speed,
        with the speed and the particles' bounding box."""
        dt, sp = C.c_double(), C.c_double()
        lo, hi = np.zeros(3), np.zeros(3)
        self._check(self.L.zmpm_calculate_dt(self.h, max_dt, C.byref(dt), C.byref(sp), _ptr(lo), _ptr(hi)),
                    "zmpm_calculate_dt")
        return dict(dt=dt.value, max_speed=sp.value, min_corner=lo, max_corner=hi)

    def launch_config(self):
        cfg = LaunchConfig()
        self._check(self.L.zmpm_get_launch_config(self.h, C.byref(cfg)), "zmpm_get_launch_config")
        return cfg.as_dict()

    def set_launch_config(self, **kw):
        cfg = LaunchConfig()
        self._check(self.L.zmpm_get_launch_config(self.h, C.byref(cfg)), "zmpm_get_launch_config")
        for k, v in kw.items():
            setattr(cfg, k, int(v))
        self._check(self.L.zmpm_set_launch_config(self.h, C.byref(cfg)), "zmpm_set_launch_config")

    def autotune(self, dt, t0=0.0, steps=4, variants=True):
        """Tune the launch configuration on this scene (the state is restored)."""
        cfg = LaunchConfig()
        self._check(self.L.zmpm_autotune(self.h, dt, t0, steps, int(bool(variants)), C.byref(cfg)), "zmpm_autotune")
        return cfg.as_dict()

    def autotune_cached(self, cache_path, dt, t0=0.0, steps=4, variants=True):
        """autotune through a cache file keyed by device and workload; returns
        (configuration, whether it came from the file)."""
        cfg = LaunchConfig()
        hit = C.c_int32(0)
        self._check(self.L.zmpm_autotune_cached(self.h, dt, t0, steps, int(bool(variants)),
                                                os.fsencode(cache_path), C.byref(cfg), C.byref(hit)),
                    "zmpm_autotune_cached")
        return cfg.as_dict(), bool(hit.value)

    def timings(self):
        cnt = C.c_int32(64)
        names = (C.c_char_p * 64)()
        ms = np.zeros(64)
        self._check(self.L.zmpm_get_timings(self.h, C.byref(cnt), names, _ptr(ms)), "zmpm_get_timings")
        return {names[i].decode(): float(ms[i]) for i in range(cnt.value)}


class Batch(Simulation):
    """One handle running several independent scenes (zmpm_create_batch): each
    scene its own dx, dt, gravity, transfer ratio, CFL number and colliders."""

    def __init__(self, lib: ZmpmLib, descs):
        self.lib = lib
        self.L = lib.lib
        arr = (SceneDesc * len(descs))(*[scene_desc(**d) for d in descs])
        self.descs = arr
        self.desc = arr[0]
        self.n_scenes = len(descs)
        h = C.c_void_p()
        self._check(self.L.zmpm_create_batch(arr, len(descs), C.byref(h)), "zmpm_create_batch", handle=False)
        self.h = h
        self.n = 0

    def add_scene_collider(self, scene: int, c: Collider) -> int:
        i = C.c_int32()
        self._check(self.L.zmpm_add_scene_collider(self.h, scene, C.byref(c), C.byref(i)), "zmpm_add_scene_collider")
        return i.value

    def origin(self, scene: int):
        o = np.zeros(3, np.int32)
        self._check(self.L.zmpm_batch_origin(self.h, scene, _ptr(o, C.c_int32)), "zmpm_batch_origin")
        return o

    def run_batch(self, steps, dts, t0s):
        dts = np.ascontiguousarray(dts, np.float64)
        t0s = np.ascontiguousarray(t0s, np.float64)
        return self._check(self.L.zmpm_run_batch(self.h, steps, _ptr(dts), _ptr(t0s)), "zmpm_run_batch")
