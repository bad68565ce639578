// zmpm_kernels.hpp -- gfx950 kernels of the explicit MPM substep.
//
// Data layout in HBM (per handle, one GPU):
//   particles  SoA, field-major: P[f * cap + s] for the NPF scalar fields below,
//              double-buffered; after each substep the live buffer holds the
//              particles in the sorted order of that substep (G2P writes them
//              there), so the next substep's gathers are nearly sequential.
//              The stress V0*tau(F, g) of the current state is a particle field:
//              G2P computes it right after F_{n+1} and g_{n+1} exist, so P2G reads
//              6 stress components instead of F and every P2G contribution is
//              bounded before P2G runs (exact fixed-point scales).
//   page table int32[domain^3 / NB]: SPGrid page index -> dense block id or -1,
//              rebuilt every substep by one scan over the pages' activation
//              flags and particle counts.
//   grid       per dense block, NB nodes (4x4x4 for float, 2x4x4 for double, the
//              reference's SPGrid block shapes), field-major accumulators:
//              double (FP64 form) or int64 fixed point (FP32 form).
#pragma once

#include "zmpm_math.hpp"
#include "zmpm.h"

namespace zmpm {

// particle scalar fields (field-major SoA)
//   FP64 form: x in metres and F, as the reference stores them.
//   FP32 form: x as 32-bit fixed point in units of dx / 2^pos_bits (the bits of
//   an int32 in the float slot), and H = F - I instead of F (EOS: J - 1 in
//   PF+0).  A float x carries ~|x| 2^-24 of rounding, which near 1 m is 1e-5 of
//   a cell at dx = 1/256, and a float F near I loses the strain F - I to
//   cancellation; both fed O(1e-4) per-field differences from the FP64 form
//   after 100 substeps.  The fixed-point position resolves dx / 2^21 everywhere
//   in the domain, advects by exact integer adds, and floors to its base node
//   without rounding; H keeps the strain to float precision.
enum : int {
    PX = 0,   // x[3] (FP32 form: fixed point, see above)
    PV = 3,   // v[3]
    PC = 6,   // C[9] row-major
    PF = 15,  // F[9] row-major (EOS: J in PF+0); FP32 form: F - I (EOS: J - 1)
    PM = 24,  // mass
    PVOL = 25, // element measure V0
    PD = 26,  // damage d
    PG = 27,  // degradation g
    PLAP = 28, // damage Laplacian (output)
    PS = 29,   // V0 * Kirchhoff stress of the current F, symmetric: xx xy xz yy yz zz
    PYS = 35,  // plasticity state: von Mises yield stress / Drucker-Prager log Jp / snow Jp
    PDT = 36,  // implicit damage: history maximum of the crack driving force D~
    PJ = 37,   // FJ-mixed particles: J (FP32 form: J - 1)
    NPF = 38
};

// FP32-form position: X = round(x / dx * 2^b) as int32 in the scene's own
// frame, b = posBits(the scene's domain): 21 for domains of up to 1024 nodes per
// axis, 19 up to 4096, so every in-domain X is below 2^31.  POS_NAN marks a
// non-finite position (an upload of NaN / Inf or a non-finite advection);
// out-of-range values are clamped to [POS_MIN, POS_MAX], where the binning's
// domain check rejects them.
__host__ __device__ constexpr int posBits(int domain_nodes) { return domain_nodes <= 1024 ? 21 : 19; }
constexpr int POS_NAN = (int)0x80000000;
constexpr int POS_MIN = -(1 << 30);
constexpr int POS_MAX = 0x7ffffffe;

// Stored position of one axis: fixed point (FP32 form) or metres (FP64 form).
template <class T> struct PosRep;
template <> struct PosRep<float> { using type = int; };
template <> struct PosRep<double> { using type = double; };
template <class T> using PosT = typename PosRep<T>::type;

// grid accumulator fields
enum : int {
    GM = 0,    // mass
    GMV = 1,   // momentum [3]
    GFO = 4,   // force [3]
    GPF = 7,   // sum w d
    GPW = 8,   // sum w
    NGF = 9
};

// fixed-point scale groups
enum : int { SC_M = 0, SC_MV = 1, SC_F = 2, SC_PF = 3, SC_PW = 4, NSC = 5 };

// latched device status bits
enum : int {
    ST_DOMAIN = 1,
    ST_NAN = 2,
    ST_CFL = 4,
    ST_OVERFLOW = 8,
    ST_CAPACITY = 16,
    ST_HALF_CFL = 32,
    ST_SCAN = 64, // a page-scan tile's look-back hit its poll limit
    ST_BARRIER = 128 // a damage-solve grid barrier hit its poll limit
};

// SPGrid block shape of GridState<T,3> (64 B float -> 4x4x4, 128 B double -> 2x4x4)
// and the axis that fills page bit 0, 1, 2 of the interleave (SPGrid_Mask.h:31-47:
// float z/y/x, double x/z/y).
template <class T> struct Blk;
template <> struct Blk<float> {
    static constexpr int xbits = 2, ybits = 2, zbits = 2, data_bits = 6;
    static constexpr int xpos = 2, ypos = 1, zpos = 0;
};
template <> struct Blk<double> {
    static constexpr int xbits = 1, ybits = 2, zbits = 2, data_bits = 7;
    static constexpr int xpos = 0, ypos = 2, zpos = 1;
};

template <class T>
struct MatDev {
    int model, has_damage, nd;
    int iso;           // QRAnisotropic::isotropic
    T mu, lambda, kx, ky, bulk, gamma;
    T kf[3];           // QRAnisotropic::k_fiber_0..2
    T kappa;           // NeoHookeanBorden::kappa
    T a0[2][3];
    T alphas[2];
    T k, l0, sigma_crit, eta, zeta;
    T inv_sigma_crit2; // 1 / sigma_c^2 (AnisoFractureSimulation.h:262)
    T phi_bound_scale; // (1 + sum |alpha_k| |a0_k|^2)^2 / sigma_c^2: phi <= this * |sigma|_F^2
    int plasticity;    // zmpm_plasticity
    int volume_correction; // DruckerPragerStvkHencky::volume_correction
    T hardening;       // VonMisesStvkHencky::xi / DruckerPragerStvkHencky::beta
    T dp_alpha;        // DruckerPragerStvkHencky::alpha (from the friction angle)
    T cohesion;        // DruckerPragerStvkHencky::cohesion
    // SnowPlasticity: singular-value bounds, Jp bounds, and the bounds on the
    // eigenvalues of F^T F - I that the float G2P gate tests (hardening is psi)
    double snow_lo, snow_hi, min_jp, max_jp;
    float snow_lam_lo, snow_lam_hi;
    // FJMixedMpmForceHelper: ZMPM_FJ_SOLID / ZMPM_FJ_FLUID (a fluid's mu and
    // lambda above are 0) and the J model's own lambda
    int fj;
    T jl;
};

// Models that carry the phase field: updateDamage dispatches to these three only
// (AnisoFractureSimulation.h:342-363).
__host__ __device__ constexpr bool phaseFieldModel(int model)
{
    return model == ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN || model == ZMPM_MODEL_QR_ANISOTROPIC
        || model == ZMPM_MODEL_NEO_HOOKEAN_BORDEN;
}

constexpr int MAX_COLLIDERS = 16; // collision objects per handle
constexpr int MAX_BODY_FORCES = 4; // external body forces per scene

// A collision object (AnalyticCollisionObject, CollisionObject.h:20-165) as the
// kernels read it: its level set, its rigid motion and, for the capped cylinder,
// the primitive's own rotation and its inverse (formed on the host).
struct ColDev {
    int shape, type;
    int moving; // rotation or scaling: R(t), s(t) evaluated; otherwise x - b(t) is material space
    int scene;  // the scene it belongs to
    double origin[3], n[3], bmin[3], bmax[3], b0[3], dbdt[3], friction;
    double radius, height;
    double Rc[9], Rci[9]; // capped cylinder primitive rotation and its inverse (row-major)
    double q0[4], omega[3], s0, dsdt;
};

// Per-substep state of a collision object: rotation R(t), translation b(t),
// scaling s(t) and 1 / s at its scene's time (the reference's updateState at the
// substep's time, MpmSimulationBase.cpp:209-213).
struct ColState {
    double R[9];
    double b[3];
    double s, os;
};

struct Masks {
    uint64_t x, y, z;
};

// One scene of a handle: a batch holds several independent scenes, each in its
// own cubic region of the handle's grid (side `region` nodes, a power of two,
// at `origin`), with its own dx, dt, gravity, transfer ratio, CFL number,
// collision objects and clock.  Particles carry their scene in the high bits
// of their material word (PMAT_SCENE_SHIFT); a grid block's scene follows from
// its coordinates (region_scene).
template <class T>
struct SceneDev {
    T dx, odx;
    T pos_scale;   // FP32 form: metres -> position units, odx * 2^pos_bits
    int pos_bits;  // FP32 form: fraction bits of the fixed-point position (posBits)
    T Dinv;        // 4 / dx^2 (MpmSimulationBase.cpp:119-122)
    T gravity[3];
    T rpic;        // apic_rpic_ratio
    T cfl;
    T dt;          // this substep's dt
    double ddt;    // the same in double (the clock)
    double ddx;    // dx in double: node positions for the collision objects
    double inv_cfl_dx; // 1 / (cfl dx): a particle speed times this is 1 / its CFL step
    int origin[3]; // first node of the scene's region
    int domain;    // the scene's own addressable nodes per axis (base nodes in [0, domain - 4))
    int col0, ncol; // its collision objects
    int ignore_cols; // MpmSimulationBase::ignoreCollisionObject
    // external body forces (MpmSimulationBase::fext as addExternalBodyForce makes
    // them): f(t) = scale (sin(frequency t) + 1) direction, evaluated per substep
    int nbf;
    double bf_scale[MAX_BODY_FORCES], bf_freq[MAX_BODY_FORCES], bf_dir[MAX_BODY_FORCES][3];
    // explicit velocity field v(x) = a + B x (row-major B), when vfield != 0
    int vfield;
    double va[3], vB[9];
};

constexpr int PMAT_SCENE_SHIFT = 8; // material index in bits 0-7, scene above
constexpr int PMAT_MAT_MASK = 0xff;

// Kernel-visible state of one handle.
template <class T, class AccT>
struct Dev {
    // particles
    T* P;          // live buffer (sorted order of the previous substep)
    T* Q;          // G2P output buffer
    int* pid;      // original index, live order
    int* pmat;     // material index, live order
    int* qid;
    int* qmat;
    int64_t cap;   // particle capacity (field stride)
    int n;
    // binning
    uint64_t* key;     // (page << block_bits | elem) per live particle
    int* rank;         // slot inside the particle's page (k_bin)
    int* rare;         // sorted positions G2P deferred to k_rare (count in counters[4])
    int* tmp;          // sorted position -> live index, before the in-block sort
    uint32_t* tmpk;    // the same positions: cell << (32 - block bits) | original index
    int* perm;         // sorted position -> live index
    int* sorted_cell;  // sorted position -> dense cell id
    int* spid;         // sorted position -> original index
    int* touched;      // per page: 1 if a particle block or its -1 neighbour has particles
    int* page_count;   // per page: particles whose base node is in the page
    int* page_table;   // per page: dense block id or -1
    int64_t num_pages;
    uint64_t* scan_state; // per page-scan tile: published prefix (see k_page_scan)
    // super pages (runs of PSCAN_TILE consecutive pages, one page-scan tile each):
    // flags set with the page flags, compacted in order into super_list, so the
    // page scan visits only the touched ones
    int64_t num_super;
    int* super_touched;
    int* super_list;
    uint64_t* super_state; // per super-page-scan tile: published prefix
    int* block_page;   // dense block id -> page index
    int* block_start;  // dense block id -> first sorted position
    int* block_count;  // dense block id -> particles
    int* nbr;          // [cap_blocks][8] forward neighbour block ids
    int* cell_count;   // [cap_blocks*NB]
    int* cell_start;   // [cap_blocks*NB]
    int* counters;     // [0] num_blocks, [1] status bits, [2] scan epoch, [3] scan ticket, [4] rare count
    double* clock;     // per scene: [2 s] time of the next substep, [2 s + 1] time of the running substep
    int cap_blocks;
    int domain;
    Masks masks;
    // grid
    AccT* acc;         // [NGF][cap_nodes]
    T* newv;           // [cap_nodes][4]: new velocity (0-2) and phase field (3), one record per node
    T* oldv;           // [cap_nodes][4]: v = mv/m before the update (0-2; FLIP_blend_PIC)
    int64_t cap_nodes;
    // fixed-point scales (float form only), per scene
    float* scales;     // [scenes][NSC] current scale 2^e
    unsigned int* maxbits; // [scenes][NSC][BOUND_SLOTS] max |contribution| bound (float bits), atomic-max slots
    double* extent;        // [EXTENT_WGS + 1][8] particle speed / bounding-box partial maxima
    int max_slots;         // workgroups of the particle-order launches (G2P / upload)
    // tables
    const MatDev<T>* mats;
    const ColDev* cols;  // grouped by scene (SceneDev::col0, ncol)
    ColState* colstate;  // [num_cols] states at the running substep's time (written by k_scatter)
    double* bfval;       // [nscenes][MAX_BODY_FORCES][3] body forces at the running substep's time
    int num_cols;
    int any_damage;
    int any_plastic;
    int any_fj;        // some material is FJ-mixed: G2P carries J
    int fj_solve_j;    // FJMixedMpmForceHelper::state == SOLVE_J
    // launch parameters
    int p2g_split; // most lanes sharing one stencil column of a cell in P2G (>= 1)
    // transfer scheme (MpmSimulationBase::transfer_scheme, mls_mpm, flip_pic_ratio)
    int flip;      // FLIP_blend_PIC: P2G without C, G2P v <- r v + sum w (v_new - r v_old)
    int mls;       // MLS-MPM: force S (x_i - x_p) w D^-1, grad v = C
    T flip_ratio;
    // implicit damage (AnisoFractureSimulation::implicit, solveDamageSystem)
    int implicit;      // phase field by the CG solve instead of the explicit update
    int cg_wgs;        // workgroups of k_cg (all resident at once)
    double cg_rel;     // CG stop below this fraction of the first residual norm
    double* cgv;       // [5][cap_nodes] x, r, s, q, diag (x and r first hold sum vol w and the D~ part of rhs)
    T* cgp;            // [4][cap] per sorted particle: d0 x, y, z; vol dt / eta l0^2 (0: not in the solve)
    double* cg_part;   // [3][CG_MAX_WGS] per-workgroup partial sums
    unsigned* cg_bar;  // [0] barrier arrivals, [1] iterations of the last solve, [2] block tickets
    // scenes
    const SceneDev<T>* sc;     // [nscenes]
    int nscenes;
    const int* region_scene;   // [region_dim^3] scene of each region, -1 for none
    int region_shift;          // log2 of the region side in nodes
    int region_dim;            // regions per axis
};

} // namespace zmpm
