/*
 * zmpm.h -- C ABI of the MI355X explicit APIC MPM substep (+ AnisoMPM damage).
 *
 * This is the drop-in boundary that replaces the explicit-substep body of
 * Ziran's MpmSimulationBase<T,3>.  Every declaration names the reference
 * interface (file:line under the Ziran tree) whose behaviour it carries.
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 *
 * Conventions
 *   - dim = 3 only.  Vectors are 3 doubles, matrices 9 doubles ROW-major
 *     (element (r,c) at [3*r+c]).  Host arrays are caller-owned and copied.
 *   - Every call returns a zmpm_status; zmpm_last_error() gives the message.
 *   - zmpm_substep()/zmpm_run() are asynchronous on the handle's HIP stream.
 *     Device-side conditions (CFL violation, NaN, domain exit, fixed-point
 *     overflow) are latched on the device and reported by the next
 *     zmpm_synchronize(), zmpm_download(), zmpm_get_binning() or
 *     zmpm_get_grid().
 *   - Calls on one handle are not thread-safe; separate handles are
 *     independent (one HIP stream each).
 */
#ifndef ZMPM_H
#define ZMPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZMPM_ABI_VERSION 8

/* Error model.  Replaces ZIRAN_ASSERT -> std::runtime_error
 * (Lib/Ziran/CS/Util/Debug.h:19-37) and the CFL RestartException
 * (Lib/MPM/MpmSimulationBase.cpp:1480-1499). */
typedef enum {
    ZMPM_OK = 0,
    ZMPM_E_ARG = 1,      /* bad argument / size / state                          */
    ZMPM_E_HIP = 2,      /* HIP runtime error                                     */
    ZMPM_E_DOMAIN = 3,   /* a particle's base node left [0, domain_nodes-5]      */
    ZMPM_E_CFL = 4,      /* a particle moved more than dx in one substep          */
    ZMPM_E_NAN = 5,      /* non-finite particle position                          */
    ZMPM_E_OVERFLOW = 6, /* 64-bit fixed-point P2G accumulator out of range       */
    ZMPM_E_NOMEM = 7     /* device allocation failed                              */
} zmpm_status;

/* Precision form chosen at create time. */
typedef enum {
    ZMPM_PREC_FP64 = 0, /* plain form: everything in double, f64 atomics          */
    ZMPM_PREC_FP32 = 1  /* particle state + stress in float, P2G in int64 fixed
                           point (deterministic), FP64 polar only near inversion */
} zmpm_precision;

/* Constitutive models on the path.
 *  FIXED_COROTATED          Physics/ConstitutiveModel/CorotatedElasticity.cpp:55-76
 *  QR_STABLE_NEO_HOOKEAN    Physics/ConstitutiveModel/QRStableNeoHookean.h:54-213
 *  NEO_HOOKEAN              Physics/ConstitutiveModel/NeoHookean.cpp:15-48
 *  EQUATION_OF_STATE (J-based helper)  Physics/ConstitutiveModel/EquationOfState.h:47-120,
 *                           MPM/Force/JBasedMpmForceHelper.cpp:28-50,79-94
 *  QR_ANISOTROPIC           Physics/ConstitutiveModel/QRAnisotropic.h:47-169 (k_shear =
 *                           mu, k_volume = lambda, k_fiber_isotropic = mu as setParameters
 *                           :57-64 sets them; k_fiber_0..2 from fiber_stiffness)
 *  NEO_HOOKEAN_BORDEN       Physics/ConstitutiveModel/NeoHookeanBorden.cpp:15-108 (kappa =
 *                           0.666666 mu + lambda as setLameParameters :15-21 sets it)   */
typedef enum {
    ZMPM_MODEL_FIXED_COROTATED = 0,
    ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN = 1,
    ZMPM_MODEL_NEO_HOOKEAN = 2,
    ZMPM_MODEL_EQUATION_OF_STATE = 3,
    ZMPM_MODEL_QR_ANISOTROPIC = 4,
    ZMPM_MODEL_NEO_HOOKEAN_BORDEN = 5
} zmpm_model;

/* Plasticity applied after the F update (MpmSimulationBase::applyPlasticity,
 * MpmSimulationBase.cpp:1517-1529, PlasticityApplier.h:27-57):
 *  VON_MISES_STVK_HENCKY     VonMisesStvkHencky::projectStrain (PlasticityApplier.cpp:1480-1516):
 *                            radial return of the Hencky strain's deviator onto
 *                            |eps_hat| = yield_stress / (2 mu), yield_stress -=
 *                            hardening * delta_gamma; per-particle state: yield stress
 *  DRUCKER_PRAGER_STVK_HENCKY DruckerPragerStvkHencky::projectStrain (PlasticityApplier.cpp:811-876):
 *                            friction cone alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)),
 *                            cohesion, tip projection with volume correction
 *                            log Jp += beta tr(eps); per-particle state: log Jp
 *  SNOW                      SnowPlasticity::projectStrain (PlasticityApplier.cpp:16-50):
 *                            singular values clamped to [1 - critical_compression,
 *                            1 + critical_stretch], Jp <- Jp det F / det Fe clamped to
 *                            [min_jp, max_jp], mu and lambda scaled by
 *                            exp(hardening (Jp_old - Jp)); per-particle state: Jp (the
 *                            material's mu, lambda are the values at Jp = 1).  Models
 *                            whose stress is linear in (mu, lambda): fixed corotated,
 *                            neo-Hookean. */
typedef enum {
    ZMPM_PLASTICITY_NONE = 0,
    ZMPM_PLASTICITY_VON_MISES_STVK_HENCKY = 1,
    ZMPM_PLASTICITY_DRUCKER_PRAGER_STVK_HENCKY = 2,
    ZMPM_PLASTICITY_SNOW = 3
} zmpm_plasticity;

/* Collision object types; values equal AnalyticCollisionObject::COLLISION_OBJECT_TYPE
 * (Lib/Ziran/Math/Geometry/CollisionObject.h:47-52). */
typedef enum {
    ZMPM_STICKY = 1,
    ZMPM_SLIP = 2,
    ZMPM_SEPARATE = 3,
    ZMPM_GHOST = 4
} zmpm_collider_type;

/* Level-set shapes (in the collision object's material space):
 *  HALF_SPACE       HalfSpace (Math/Geometry/AnalyticLevelSet.cpp:318-363)
 *  BOX              AxisAlignedAnalyticBox (AnalyticLevelSet.cpp:412-428 over
 *                   AnalyticBox :545-600), normal by the reference's AutoDiff rules
 *  SPHERE           Sphere (AnalyticLevelSet.cpp:446-522): inside when |X - c|^2 <
 *                   r^2 (strict), normal (X - c)/|X - c|, e_x within 1e-7 of c
 *  CAPPED_CYLINDER  CappedCylinder (AnalyticLevelSet.h:252-340): radius, height along
 *                   its local y, its own rotation quaternion and centre            */
typedef enum {
    ZMPM_SHAPE_HALF_SPACE = 0,
    ZMPM_SHAPE_BOX = 1,
    ZMPM_SHAPE_SPHERE = 2,
    ZMPM_SHAPE_CAPPED_CYLINDER = 3
} zmpm_shape;

/* Particle-grid transfer scheme; values equal MpmSimulationBase::TransferScheme
 * (Lib/MPM/MpmSimulationBase.h:47-51).  Quadratic B-splines (ZIRAN_MPM_DEGREE 2). */
typedef enum {
    ZMPM_APIC_BLEND_RPIC = 0, /* APIC, C blended with its skew part (apic_rpic_ratio)   */
    ZMPM_FLIP_BLEND_PIC = 1   /* FLIP blended with PIC (flip_pic_ratio; 0 = PIC), no C   */
} zmpm_transfer;

/* Scene-wide parameters.  Mirrors the public members of MpmSimulationBase
 * (Lib/MPM/MpmSimulationBase.h:47-76) that the explicit step reads. */
typedef struct {
    int32_t precision;       /* zmpm_precision                                            */
    int32_t domain_nodes;    /* grid nodes per axis that may be addressed: power of two in
                                [64, 4096], the reference SPGrid's range (MpmGrid.h:111-112).
                                Sizes the page flags and table (4 B per page each; 12 GB
                                at 4096 in the FP32 form, 24 GB in the FP64 form), which
                                each substep visits only where particles are.  The FP32
                                form's fixed-point positions resolve dx / 2^21 up to 1024
                                nodes, dx / 2^19 above.                                   */
    double dx;               /* MpmSimulationBase::dx                                     */
    double gravity[3];       /* MpmSimulationBase::gravity                                */
    double apic_rpic_ratio;  /* MpmSimulationBase::apic_rpic_ratio (default 1)            */
    double cfl;              /* MpmSimulationBase::cfl (half-cell flag only)              */
    int32_t device;          /* HIP device ordinal                                        */
    int32_t transfer_scheme; /* zmpm_transfer, MpmSimulationBase::transfer_scheme
                                (MpmSimulationBase.h:52; default APIC_blend_RPIC = 0)      */
    double flip_pic_ratio;   /* MpmSimulationBase::flip_pic_ratio (FLIP_blend_PIC)        */
    int32_t mls_mpm;         /* MpmSimulationBase::mls_mpm: MLS-MPM force and grad v = C
                                (APIC only, MpmSimulationBase.cpp:108-112)                */
    int32_t rpic_damping_iteration; /* MpmSimulationBase::rpic_damping_iteration (APIC)   */
    int32_t pic_damping_iteration;  /* MpmSimulationBase::pic_damping_iteration (FLIP)    */
    int32_t implicit_damage; /* AnisoFractureSimulation::implicit (AnisoFractureSimulation.h:61,
                                129-141): 1 = the phase field of the damaged particles by the
                                matrix-free CG solve of solveDamageSystem (:435-632), 0 = the
                                explicit Laplacian update                                 */
    int32_t ignore_collision_objects; /* MpmSimulationBase::ignoreCollisionObject
                                (MpmSimulationBase.cpp:1217-1221): the grid update leaves
                                the velocities the collision objects would change         */
} zmpm_scene_desc;

/* One material = one constitutive-model parameter set shared by its particles,
 * plus the optional AnisotropicPhaseField (Projects/anisofracture/
 * AnisotropicPhaseField.h:10-33) attached by
 * MpmParticleHandleBase::addFBasedMpmForceWithAnisotropicPhaseField
 * (Lib/MPM/MpmParticleHandleBase.cpp:175-222). */
typedef struct {
    int32_t model;           /* zmpm_model                                                */
    int32_t has_damage;      /* 1: phase-field damage updated each substep (QR-NH,
                                QR-aniso and NH-Borden; AnisoFractureSimulation.h:342-363
                                dispatches no other model on the path)                    */
    double mu, lambda;       /* Lame parameters (every model but EOS)                     */
    double kx, ky;           /* QRStableNeoHookean extra fibre stiffness                  */
    double bulk, gamma;      /* EquationOfState                                           */
    int32_t num_directors;   /* 1 or 2 structural directors a_0                           */
    int32_t plasticity;      /* zmpm_plasticity (models with mu: all but EOS)             */
    double a0[2][3];         /* structural directors (reference uses unit e1, e2)         */
    double alphas[2];
    double residual_phase;   /* k in g = (1-d)^2 (1-k) + k                                */
    double l0, sigma_crit, eta, zeta;
    double yield_stress;     /* VonMisesStvkHencky::yield_stress (initial, per particle)  */
    double hardening;        /* VonMisesStvkHencky::xi / DruckerPragerStvkHencky::beta /
                                SnowPlasticity::psi                                       */
    double friction_angle;   /* DruckerPragerStvkHencky friction angle (degrees)          */
    double cohesion;         /* DruckerPragerStvkHencky::cohesion                          */
    int32_t volume_correction; /* DruckerPragerStvkHencky::volume_correction              */
    int32_t isotropic;       /* QRAnisotropic::isotropic: every fibre term degraded and
                                scaleFiberStiffness a no-op (QRAnisotropic.h:96-104,139-147) */
    double critical_compression; /* SnowPlasticity::theta_c (reference default 2e-2)      */
    double critical_stretch;     /* SnowPlasticity::theta_s (7.5e-3)                       */
    double min_jp, max_jp;       /* SnowPlasticity::min_Jp, max_Jp (0.6, 20)               */
    double fiber_stiffness[3];   /* QRAnisotropic::k_fiber_0..2 (mu each after setParameters,
                                    times scaleFiberStiffness's factor)                   */
    int32_t fj_mixed;        /* zmpm_fj_phase: 0 = the model's own force helper; SOLID /
                                FLUID = FJMixedMpmForceHelper<model, QuadraticVolumePenalty>
                                (Lib/MPM/Force/FJMixedMpmForceHelper.h:16-278) as
                                addFJMixedMpmForce adds it (MpmParticleHandleBase.cpp:
                                811-834): fixed-corotated or neo-Hookean F model, a
                                per-particle J, no phase field; a FLUID phase zeroes
                                the F model's mu and lambda                              */
    double j_lambda;         /* QuadraticVolumePenalty::lambda of the J model
                                (psi = lambda/2 (J - 1)^2)                               */
} zmpm_material;

/* Material phase of an FJ-mixed material (MATERIAL_PHASE_ENUM). */
typedef enum {
    ZMPM_FJ_NONE = 0,
    ZMPM_FJ_SOLID = 1,
    ZMPM_FJ_FLUID = 2
} zmpm_fj_phase;

/* Analytic collision object (AnalyticCollisionObject, CollisionObject.h:20-165):
 * a level set in material space X moved rigidly to world space as
 * x = s(t) R(t) X + b(t), with
 *   b(t) = b0 + dbdt t                     (setTranslation, CollisionObject.h:144-148)
 *   R(t) = rotation of the quaternion q(t) = (cos(|w| t/2), sin(|w| t/2) w/|w|) * q0
 *                                          (setRotation / setAngularVelocity :150-160,
 *                                           a constant angular velocity w as the
 *                                           examples/boneTwist.h:144-170 callbacks set)
 *   s(t) = s0 + dsdt t                     (setScaling :162-166)
 * and object velocity v = w x (x - b) + (dsdt / s)(x - b) + dbdt
 * (detectAndResolveCollision, CollisionObject.cpp:341-364).  A zero rotation
 * quaternion means identity and a zero scale0 means 1, so a zero-initialised
 * record is a static, unscaled object. */
typedef struct {
    int32_t shape;           /* zmpm_shape                                                */
    int32_t type;            /* zmpm_collider_type                                        */
    double origin[3];        /* half-space origin; sphere / capped-cylinder centre        */
    double normal[3];        /* half-space outward normal (normalised by the library)     */
    double box_min[3];       /* box corners in material space                             */
    double box_max[3];
    double b0[3];            /* translation at t = 0                                      */
    double dbdt[3];          /* translation velocity                                      */
    double friction;         /* SLIP / SEPARATE kinematic friction                        */
    double radius;           /* sphere, capped cylinder                                   */
    double height;           /* capped cylinder, along its local y axis                   */
    double shape_rotation[4];/* capped cylinder's own rotation (w, x, y, z); 0 = identity */
    double rotation[4];      /* object rotation q0 at t = 0 (w, x, y, z); 0 = identity    */
    double omega[3];         /* object angular velocity                                   */
    double scale0;           /* object scaling at t = 0 (0 = 1)                           */
    double dsdt;             /* object scaling rate                                       */
} zmpm_collider;

/* Host particle arrays (caller-owned).  Stands for the Particles/DataManager
 * attributes the explicit step touches (Math/Geometry/Particles.h:7-45;
 * "F" FBasedMpmForceHelper.h:219-222; "element measure" :215-218; C added at
 * MpmSimulationBase.cpp:118; phase field AnisotropicPhaseField.h:13-25). */
typedef struct {
    int64_t n;
    double* x;          /* [n][3]                                                         */
    double* v;          /* [n][3]                                                         */
    double* C;          /* [n][9] APIC affine matrix                                      */
    double* F;          /* [n][9] deformation gradient; EOS particles keep J in F[0]      */
    double* m;          /* [n] mass                                                       */
    double* V0;         /* [n] element measure (initial volume)                           */
    int32_t* material;  /* [n] index into the material table                              */
    double* d;          /* [n] damage, may be NULL (0 on upload)                          */
    double* g;          /* [n] degradation g of the QR model, may be NULL (1 on upload)   */
    double* laplacian;  /* [n] damage Laplacian (download only), may be NULL              */
    double* plastic_state; /* [n] plasticity state (von Mises: yield stress, Drucker-Prager:
                              log Jp, snow: Jp), may be NULL (upload: the material's yield
                              stress / 0 / 1)                                              */
    double* max_dtilde; /* [n] AnisotropicPhaseField::maxDTilde, the history maximum of the
                           crack driving force the implicit damage solve reads
                           (AnisoFractureSimulation.h:284-337), may be NULL (0 on upload)  */
    double* J;          /* [n] the "J" attribute of FJ-mixed particles
                           (FJMixedMpmForceHelper.h:36), may be NULL (1 on upload)        */
} zmpm_particles;

typedef struct zmpm_context* zmpm_handle;

/* Library identity: ABI version and the gfx target the device code was built for. */
int32_t zmpm_abi_version(void);
const char* zmpm_build_info(void);

/* Construct a simulation.  Replaces MpmSimulationBase::MpmSimulationBase +
 * initialize() (MpmSimulationBase.cpp:37-133): D^-1 = 4/dx^2, C attribute. */
zmpm_status zmpm_create(const zmpm_scene_desc* desc, zmpm_handle* out);
zmpm_status zmpm_destroy(zmpm_handle h);
const char* zmpm_last_error(zmpm_handle h);

/* Material table (<= 256 entries).  Replaces the per-particle constitutive
 * model attribute added by addFBasedMpmForce* / addJBasedMpmForce
 * (MpmParticleHandleBase.cpp:175-222,796-809). */
zmpm_status zmpm_set_materials(zmpm_handle h, int32_t count, const zmpm_material* mats);

/* FJMixedMpmForceHelper::state (FJMixedMpmForceHelper.h:38-40), the helper's
 * static SOLVE_F / SOLVE_J switch, for the FJ-mixed particles of this handle:
 * 0 = SOLVE_F (default; updateState adds the F model's V0 tau(F)), 1 = SOLVE_J
 * (adds V0 lambda_J (J - 1) J I, :98-121).  Takes effect at the next substep. */
zmpm_status zmpm_set_fj_state(zmpm_handle h, int32_t state);

/* Collision objects (<= 16).  Replaces MpmSimulationBase::addCollisionObject
 * (MpmSimulationBase.cpp:87-91); evaluation order = insertion order, as in
 * AnalyticCollisionObject::multiObjectCollision (CollisionObject.cpp:66-105). */
zmpm_status zmpm_add_collider(zmpm_handle h, const zmpm_collider* c, int32_t* id_out);
zmpm_status zmpm_clear_colliders(zmpm_handle h);

/* Particle state in / out.  Upload replaces the scene-building writes of
 * MpmInitializationHelper::sampleInAnalyticLevelSet (MpmInitializationHelper.h:637-647);
 * download returns particles in their upload order (the reference never
 * reorders the DataManager arrays). */
zmpm_status zmpm_upload(zmpm_handle h, const zmpm_particles* p);
zmpm_status zmpm_download(zmpm_handle h, zmpm_particles* p);

/* One explicit substep at time t with step dt.  Replaces
 * MpmSimulationBase::advanceOneTimeStep (MpmSimulationBase.cpp:262-357) with
 * symplectic = true, APIC, quadratic B-splines, including the AnisoMPM
 * particlesToGrid override (Projects/anisofracture/AnisoFractureSimulation.h:129-141)
 * and the collision-object updateState at t and t+dt (MpmSimulationBase.cpp:291-293). */
zmpm_status zmpm_substep(zmpm_handle h, double dt, double t);

/* `steps` substeps t0, t0+dt, ...; the GPU library replays a captured HIP graph. */
zmpm_status zmpm_run(zmpm_handle h, int32_t steps, double dt, double t0);

/* Stages of one substep, for a host that replaces one of the virtual steps of
 * MpmSimulationBase::advanceOneTimeStep (MpmSimulationBase.h:152-227) and keeps
 * the device for the others.  zmpm_stage(h, s, dt, t) runs stage s of the
 * substep at time t with step dt; the stages of a substep run in this order,
 * each once, with the same dt and t, and together equal zmpm_substep(h, dt, t):
 *   SORT          beginTimeStep + sortParticlesAndPolluteGrid (:238-245, 1531-1573)
 *   P2G           forcesUpdateParticleState + particlesToGrid (:953-979, 618-667)
 *                 with the AnisoMPM phase-field transfer (AnisoFractureSimulation.h:
 *                 129-141); zmpm_get_grid then reports m and v = mv / m
 *   GRID_UPDATE   gridVelocityExplicitUpdate (:1179-1228); zmpm_get_grid then
 *                 reports new_v, which zmpm_set_grid_velocity may replace
 *   DAMAGE_SOLVE  solveDamageSystem (AnisoFractureSimulation.h:435-612) when
 *                 implicit_damage is set (else nothing)
 *   G2P           gridToParticles + evolveStrain + the damage and plasticity
 *                 updates + the damping passes (:1412-1529, 318-348)
 * The particle state changes at G2P only (the explicit damage update, which the
 * reference runs inside particlesToGrid, reads only F_n and the phase-field
 * Laplacian, so it is applied there).  Between the stages of a substep only the
 * queries (zmpm_get_grid, zmpm_get_binning, zmpm_download) and
 * zmpm_set_grid_velocity are allowed; the other entry points return ZMPM_E_ARG
 * until its G2P stage has run.  A single-scene handle only. */
typedef enum {
    ZMPM_STAGE_SORT = 0,
    ZMPM_STAGE_P2G = 1,
    ZMPM_STAGE_GRID_UPDATE = 2,
    ZMPM_STAGE_DAMAGE_SOLVE = 3,
    ZMPM_STAGE_G2P = 4
} zmpm_stage_id;
zmpm_status zmpm_stage(zmpm_handle h, int32_t stage, double dt, double t);

/* Replace the updated grid velocities (new_v, 3 doubles per node) of the
 * substep in progress, after its GRID_UPDATE stage and before its G2P stage:
 * n_nodes and the node order are those zmpm_get_grid reports (nodes with mass,
 * ordered by (x, y, z) node coordinates).  For a host override of
 * gridVelocityExplicitUpdate. */
zmpm_status zmpm_set_grid_velocity(zmpm_handle h, int64_t n_nodes, const double* new_v);

/* Wait for queued work; returns the first latched device condition. */
zmpm_status zmpm_synchronize(zmpm_handle h);

/* Binning of the LAST substep, as built by
 * MpmSimulationBase::sortParticlesAndPolluteGrid (MpmSimulationBase.cpp:1531-1573):
 *   keys[n]         sorted particle_sorter values ((offset>>data_bits)<<(32-block_bits)) + i
 *   order[n]        particle_order
 *   groups[2*G]     particle_group (first, last) pairs
 *   block_offset[G] page index of each group
 * Pass NULL arrays to query n_groups only. */
zmpm_status zmpm_get_binning(zmpm_handle h, int64_t* n_groups, uint64_t* keys,
                             int32_t* order, int32_t* groups, uint64_t* block_offset);

/* Grid nodes with m != 0 (the reference's idx >= 0 nodes) after the LAST
 * substep, sorted by (i, j, k):
 *   coords[3N] int32, m[N], v[3N] = mv/m (MpmSimulationBase.cpp:655-666),
 *   new_v[3N] after gridVelocityExplicitUpdate (:1179-1228),
 *   pf[N] normalised phase field (AnisoFractureSimulation.h:177-179; 0 if no damage).
 * Pass NULL arrays to query the count only. */
zmpm_status zmpm_get_grid(zmpm_handle h, int64_t* n_nodes, int32_t* coords,
                          double* m, double* v, double* new_v, double* pf);

/* External body force of a scene, MpmInitializationHelper::addExternalBodyForce
 * (Lib/MPM/MpmInitializationHelper.h:1436-1444) appended to MpmSimulationBase::fext:
 * f(x, t) = scale (sin(frequency t) + 1) direction, added as v_i += dt f to every
 * grid node with mass after gravity and the internal forces, before the collision
 * objects, at the scene's time t of the substep (MpmSimulationBase.cpp:1208-1212).
 * At most 4 per scene, applied in the order added. */
zmpm_status zmpm_add_body_force(zmpm_handle h, int32_t scene, double scale, double frequency,
                                const double* direction);

/* MpmSimulationBase::explicit_velocity_field (MpmSimulationBase.h:124, set by
 * MpmInitializationHelper::addExplicitVelocityField :1416-1419) as the affine map
 * v(x) = a + B x (B row-major): when set, the grid update writes new_v = a + B x_i
 * at every node instead of the force update and the collision objects
 * (MpmSimulationBase.cpp:1187-1195).  The reference takes any function of (m, x,
 * v); the device takes the affine family, which holds the rigid motions and
 * uniform strain rates.  a == NULL removes the field. */
zmpm_status zmpm_set_velocity_field(zmpm_handle h, int32_t scene, const double* a, const double* B);

/* Batched independent scenes in one handle: every launch of a substep covers
 * all of them.  No counterpart in the reference, whose simulations run one at a
 * time (one MpmSimulationBase each); each scene keeps the reference's
 * semantics.  descs[s] gives scene s its own dx, gravity, apic_rpic_ratio, cfl
 * and domain_nodes; precision, device, transfer scheme, flip_pic_ratio, mls_mpm
 * and the damping iterations must be equal in every scene, and implicit_damage
 * needs a single scene.  Scene s lives in its own cubic region of the handle's
 * grid (side = the largest domain_nodes, regions in page order; at most 4096
 * nodes per axis in all): positions in and out are metres in the scene's own
 * frame, as if it ran alone, and zmpm_get_grid / zmpm_get_binning report the
 * handle's grid (a scene's node (i, j, k) is node origin + (i, j, k)).  The
 * material table is shared (particles name their material); colliders belong
 * to one scene; statuses are latched for the whole batch.  Each scene's result
 * equals its run alone: FP32 bit for bit (per-scene fixed-point scales), FP64
 * to the rounding of its double atomics. */
zmpm_status zmpm_create_batch(const zmpm_scene_desc* descs, int32_t n_scenes, zmpm_handle* out);
/* First node of scene s's region in the handle's grid (3 int32). */
zmpm_status zmpm_batch_origin(zmpm_handle h, int32_t scene, int32_t* origin);
/* Collision object of one scene (zmpm_add_collider adds to scene 0). */
zmpm_status zmpm_add_scene_collider(zmpm_handle h, int32_t scene, const zmpm_collider* c, int32_t* id_out);
/* Upload with scene[i] the scene of particle i (zmpm_upload puts all in scene 0). */
zmpm_status zmpm_upload_batch(zmpm_handle h, const zmpm_particles* p, const int32_t* scene);
/* `steps` substeps of every scene s with its own step dt[s] from its own time
 * t0[s] (zmpm_run gives every scene the same dt and t0). */
zmpm_status zmpm_run_batch(zmpm_handle h, int32_t steps, const double* dt, const double* t0);

/* Per-stage device timing (HIP events on the handle's stream) of the substeps
 * run while enabled; off by default.  No counterpart in the reference, whose
 * ZIRAN_TIMER scopes (Lib/Ziran/CS/Util/Timer.h:35-60) play this role on the CPU. */
zmpm_status zmpm_set_timing(zmpm_handle h, int32_t enable);

/* Mean per-substep stage times (ms) of the last timed zmpm_run/zmpm_substep.
 * *count: capacity of names/ms on entry, entries written on return.
 * names[] receives static strings.  The oracle reports none. */
zmpm_status zmpm_get_timings(zmpm_handle h, int32_t* count, const char** names, double* ms);

/* MpmSimulationBase::calculateDt (Lib/MPM/MpmSimulationBase.cpp:1127-1151) with
 * evalMaxParticleSpeed (:1613-1638): the largest particle speed |v| and the
 * particles' bounding box, reduced on the device (synchronous).  *dt = cfl dx /
 * max speed, or max_dt when every particle is at rest (collision-object speeds
 * are not included, as in the reference); for a batch the smallest of the
 * scenes' steps and the box over the scenes' own frames.  Output pointers may
 * be NULL; min_corner / max_corner take 3 doubles. */
zmpm_status zmpm_calculate_dt(zmpm_handle h, double max_dt, double* dt, double* max_speed, double* min_corner,
                              double* max_corner);

/* zmpm_calculate_dt without allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The same device
 * reduction as zmpm_calculate_dt, asynchronous on the handle's stream: its
 * per-workgroup partial maxima go to a fixed static device array (1025 rows of
 * 8 doubles, 64.1 KiB per GPU, held until the process ends) instead of the
 * handle's buffer, and the result to `out`, caller-owned DEVICE memory of 8
 * doubles: dt, the largest particle speed, min corner[3], max corner[3].
 * reduction_wgs (1..1024) sets the workgroups of the first pass; larger values
 * are rejected with ZMPM_E_ARG.  Calls nothing that allocates.  Every call on
 * one GPU shares the array (each call overwrites the rows it uses): calls that
 * overlap on one GPU -- other handles' streams, other threads -- race. */
zmpm_status zmpm_calculate_dt_noalloc(zmpm_handle h, double max_dt, int32_t reduction_wgs, double* out);

/* Launch configuration of the grid-stride kernels (workgroups of 256 threads),
 * the compiled kernel variants and the P2G lane split.  The workgroup counts
 * and the variants do not change any result; the P2G split cap (most lanes
 * sharing one stencil column of a cell) changes how a cell's particles are
 * grouped into float partial sums before the 64-bit fixed-point accumulation
 * (results deterministic for a given value; the autotuner never changes it).
 * No counterpart in the reference (launch configuration of this library). */
typedef struct {
    int32_t p2g_wgs;     /* in-block sort + P2G: workgroups over grid blocks      */
    int32_t grid_wgs;    /* grid update: workgroups over nodes                    */
    int32_t prep_wgs;    /* accumulator zeroing / neighbour table                 */
    int32_t p2g_split;   /* P2G lanes per stencil column, cap (>= 1)              */
    int32_t variant;     /* compiled variants, bits: 1 = P2G built for one wave per
                            SIMD fewer, 2 = G2P segments round-robin over the XCDs,
                            4 = P2G blocks XCD by XCD, 8 = collision objects
                            resolved in a launch after the grid update (default
                            for batches) instead of inside it                     */
    int32_t reserved[3];
} zmpm_launch_config;

zmpm_status zmpm_get_launch_config(zmpm_handle h, zmpm_launch_config* cfg);
zmpm_status zmpm_set_launch_config(zmpm_handle h, const zmpm_launch_config* cfg);

/* Autotuner: times candidate workgroup counts for each grid-stride kernel on
 * the uploaded scene (coordinate descent, each candidate over two rounds of
 * `steps` event-timed substeps, its faster round kept, colliders at times from
 * t0) and keeps the fastest; with tune_variants != 0 also the compiled kernel
 * variants.  Nothing it changes changes a result.  The particle state is
 * saved before and restored after, so the simulation continues from where it
 * was (the grid and binning queries then describe the tuner's last substep).
 * The chosen configuration is returned in *cfg (may be NULL).  In a batch
 * every scene steps with dt while tuning.
 * No counterpart in the reference (launch configuration of this library). */
zmpm_status zmpm_autotune(zmpm_handle h, double dt, double t0, int32_t steps, int32_t tune_variants,
                          zmpm_launch_config* cfg);

/* zmpm_autotune with a cache file of tuned configurations, one text line per
 * (device, workload) key: the device's gfx target and compute-unit count, the
 * precision, the number of scenes, the particle count rounded to a quarter
 * power of two, and the options that select kernels (damage, implicit solve,
 * transfer scheme, MLS, damping, FJ-mixed, plasticity).  A key found in the
 * file applies its configuration without timing (*from_cache = 1); otherwise
 * the tuner runs and appends its result.  The file is created when missing;
 * processes tuning the same file at once may both append (the first line of
 * a key is the one read). */
zmpm_status zmpm_autotune_cached(zmpm_handle h, double dt, double t0, int32_t steps, int32_t tune_variants,
                                 const char* cache_path, zmpm_launch_config* cfg, int32_t* from_cache);

#ifdef __cplusplus
}
#endif

#endif /* ZMPM_H */
