// zmpm_api.hip -- host side of the C ABI declared in include/zmpm.h: device
// memory, the substep launch sequence on the handle's HIP stream, uploads,
// downloads and the binning / grid exports used for parity checks.
#include "zmpm_device.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

using namespace zmpm;

// Scratch of zmpm_calculate_dt_noalloc: the per-workgroup partial maxima of its
// first reduction pass (one row of 8 doubles per workgroup) and the folded row,
// at most NOALLOC_MAX_WGS + 1 rows (64.1 KiB), one copy per GPU held until the
// process ends.  Host code never passes this name to a kernel or a copy: it
// stands for a host-side shadow; the device address comes from
// hipGetSymbolAddress once per device (noallocScratch below).
constexpr int NOALLOC_MAX_WGS = 1024;
static __device__ __align__(256) double scratch[(NOALLOC_MAX_WGS + 1) * 8];

namespace {

constexpr int kMaxMaterials = 256;
constexpr int kMaxColliders = MAX_COLLIDERS;
constexpr int kTimedGroups = 5;
constexpr int kPageScanWgs = 1024; // persistent page-scan workgroups (at most one per super page)
constexpr int kGraphSteps = 8;     // substeps per captured graph of zmpm_run (even)
// compiled kernel variants the launch configuration selects
constexpr int VARIANT_P2G_LOW_OCC = 1; // P2G compiled for one wave per SIMD fewer
constexpr int VARIANT_G2P_RR = 2;      // G2P segments round-robin over the XCDs
constexpr int VARIANT_P2G_XCD = 4;     // P2G blocks XCD by XCD
constexpr int VARIANT_GRID_SPLIT = 8;  // collision objects in a launch after the grid update
constexpr int VARIANT_ALL = 15;
const char* const kTimedNames[kTimedGroups + 1] = { "bin_sort", "p2g", "grid_update", "damage_solve", "g2p", "substep" };

uint64_t spreadMask(int log2_struct, int axis)
{
    // SPGrid_Mask (Lib/SPGrid/Core/SPGrid_Mask.h:22-123) aggregate masks for log2_page = 12
    const int page_bits = 12, data_bits = log2_struct, block_bits = page_bits - data_bits;
    const int zb = block_bits / 3 + (block_bits % 3 > 0), yb = block_bits / 3 + (block_bits % 3 > 1), xb = block_bits / 3;
    const int sh = 3 - block_bits % 3;
    const uint64_t hi = ~0ull << page_bits;
    if (axis == 0) return ((0x4924924924924924ull << sh) & hi) | ((uint64_t)((1 << xb) - 1) << (data_bits + zb + yb));
    if (axis == 1) return ((0x2492492492492492ull << sh) & hi) | ((uint64_t)((1 << yb) - 1) << (data_bits + zb));
    return ((0x9249249249249249ull << sh) & hi) | ((uint64_t)((1 << zb) - 1) << data_bits);
}

uint32_t hostBitPack(uint64_t data, uint64_t mask)
{
    uint32_t r = 0, bit = 0;
    for (int b = 0; b < 64 && bit < 32; b++)
        if (mask & (1ull << b)) {
            if (data & (1ull << b)) r |= 1u << bit;
            bit++;
        }
    return r;
}

} // namespace

struct zmpm_context {
    zmpm_scene_desc desc {}; // scene 0's description: the options every scene shares
    std::string err;
    const char* err_static = nullptr; // message of the allocation-free entry point
    hipStream_t stream = nullptr;
    bool fp32 = false;
    bool any_damage = false;
    std::vector<zmpm_material> mats;
    // scenes: one (zmpm_create) or a batch (zmpm_create_batch), each in its own
    // cubic region of the grid, regions laid out in page (Morton) order
    std::vector<zmpm_scene_desc> scenes;
    std::vector<std::vector<zmpm_collider>> scene_cols; // per scene, insertion order
    int region = 0;                  // region side in nodes (the largest scene domain)
    int region_dim = 1;              // regions per axis
    int grid_domain = 0;             // nodes per axis of the whole grid
    std::vector<int> region_scene;   // [region_dim^3] scene or -1
    std::vector<std::array<int, 3>> origin; // per scene: first node of its region
    std::vector<int32_t> pscene;     // scene of each uploaded particle (upload order)
    std::vector<double> step_dt;     // per scene: dt the device table holds (0: none yet)
    std::vector<unsigned char> scene_host; // host mirror of the device scene table
    void* scenes_dev = nullptr;
    void* bfval_dev = nullptr; // body forces at the running substep's time, per scene
    // per scene: external body forces (scale, frequency, direction) and the
    // explicit velocity field (enabled, a, B)
    struct BodyForce {
        double scale, freq, dir[3];
    };
    struct VelocityField {
        bool on = false;
        double a[3] {}, B[9] {};
    };
    std::vector<std::vector<BodyForce>> body_forces;
    std::vector<VelocityField> vfields;
    void* regions_dev = nullptr;
    void* colstate_dev = nullptr;
    Dev<float, long long> Df {};
    Dev<double, double> Dd {};
    std::vector<void*> allocs; // particle + grid buffers (freed on reallocation)
    void* mats_dev = nullptr;
    void* cols_dev = nullptr;
    int n = 0;
    int max_mat = -1; // largest material index of the uploaded particles
    int64_t cap = 0;
    bool timing = false;
    std::vector<hipEvent_t> events;
    int timed_steps = 0;
    double timed_ms[kTimedGroups + 1] = {};
    bool stepped = false;
    std::vector<double> clock_host; // per scene: next and running substep time (upload source)
    // captured graphs of kGraphSteps or two substeps, keyed by live buffer and
    // length (dt and t are device state)
    struct Graph {
        const void* live;
        int steps;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    bool use_graphs = true;
    int fj_state = 0;       // FJMixedMpmForceHelper::state: 0 SOLVE_F, 1 SOLVE_J
    hipError_t launch_err = hipSuccess; // a failed launch that returns its error (the cooperative k_cg)
    // zmpm_stage: the next stage of the substep in progress (0 = none), its dt and t
    int stage_next = 0;
    double stage_dt = 0, stage_t = 0;
    int p2g_split = 64;
    // workgroups of the grid-stride launches (results do not depend on them)
    int p2g_wgs = 2048, grid_wgs = 2048, prep_wgs = 8192;
    bool p2g_wgs_env = false; // P2G workgroups set by ZMPM_P2G_WGS
    int rare_wgs = 1024; // k_rare's grid: its workgroups resident at once on the device
    int variant = 0; // compiled kernel variants (VARIANT_* bits), none changes a result
    bool launch_set = false; // configured by zmpm_set_launch_config or the autotuner
};

namespace {

zmpm_status fail(zmpm_context* h, zmpm_status s, const std::string& msg)
{
    if (h) {
        h->err = msg;
        h->err_static = nullptr;
    }
    return s;
}

// Entry points that change the state: refused between the stages of a substep.
bool midSubstep(zmpm_context* h)
{
    if (h->stage_next == 0) return false;
    (void)fail(h, ZMPM_E_ARG, "a substep is in progress (zmpm_stage): run its stages to G2P first");
    return true;
}

// Error of the allocation-free entry point: a static message, no string copy.
zmpm_status failStatic(zmpm_context* h, zmpm_status s, const char* msg)
{
    if (h) h->err_static = msg;
    return s;
}

// Device address of `scratch` on each device, taken once per device (with that
// device current) by zmpm_create.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

zmpm_status setUpNoallocScratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return ZMPM_E_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return ZMPM_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return ZMPM_E_HIP;
    }
    g_scratch[device] = static_cast<double*>(p);
    return ZMPM_OK;
}

#define HIPCHK(h, x)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) return fail((h), ZMPM_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class X>
zmpm_status dalloc(zmpm_context* h, X** out, size_t count)
{
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(X));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(h, ZMPM_E_NOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(X)) + " bytes failed");
    }
    h->allocs.push_back(p);
    *out = (X*)p;
    return ZMPM_OK;
}

void dropGraphs(zmpm_context* h)
{
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}

void freeAll(zmpm_context* h)
{
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    dropGraphs(h);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    h->Df = Dev<float, long long> {};
    h->Dd = Dev<double, double> {};
    h->cap = 0;
    h->n = 0;
}

template <class T, class AccT>
Dev<T, AccT>& devOf(zmpm_context* h);
template <>
Dev<float, long long>& devOf<float, long long>(zmpm_context* h) { return h->Df; }
template <>
Dev<double, double>& devOf<double, double>(zmpm_context* h) { return h->Dd; }

template <class T, class AccT>
zmpm_status allocate(zmpm_context* h, int n)
{
    freeAll(h);
    Dev<T, AccT> D {};
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    const int64_t dom = h->grid_domain;
    const int64_t ns = (int64_t)h->scenes.size();
    const int64_t cap = std::max(n, 1);
    D.cap = cap;
    D.domain = (int)dom;
    D.num_pages = dom * dom * dom / NB;
    int64_t cap_blocks = (int64_t)n / 16 + 16384;
    if (const char* e = std::getenv("ZMPM_BLOCK_CAPACITY")) cap_blocks = std::max<int64_t>(64, std::atoll(e));
    cap_blocks = std::min<int64_t>(cap_blocks, D.num_pages);
    cap_blocks = std::min<int64_t>(cap_blocks, (int64_t)(0x7fffffff / NB) - 1);
    D.cap_blocks = (int)cap_blocks;
    D.cap_nodes = cap_blocks * NB;
    zmpm_status s;
#define ALLOC(ptr, cnt) \
    if ((s = dalloc(h, &(ptr), (size_t)(cnt))) != ZMPM_OK) return s
    ALLOC(D.P, NPF * cap);
    ALLOC(D.Q, NPF * cap);
    ALLOC(D.pid, cap);
    ALLOC(D.pmat, cap);
    ALLOC(D.qid, cap);
    ALLOC(D.qmat, cap);
    ALLOC(D.key, cap);
    ALLOC(D.rank, cap);
    ALLOC(D.tmp, cap);
    ALLOC(D.rare, cap);
    ALLOC(D.tmpk, cap);
    ALLOC(D.perm, cap);
    ALLOC(D.sorted_cell, cap);
    ALLOC(D.spid, cap);
    ALLOC(D.touched, D.num_pages);
    ALLOC(D.page_count, D.num_pages);
    ALLOC(D.page_table, D.num_pages);
    const int64_t scan_tiles = (D.num_pages + PSCAN_TILE - 1) / PSCAN_TILE;
    ALLOC(D.scan_state, scan_tiles);
    D.num_super = scan_tiles; // one page-scan tile per super page
    const int64_t super_tiles = (D.num_super + PSCAN_TILE - 1) / PSCAN_TILE;
    ALLOC(D.super_touched, D.num_super);
    ALLOC(D.super_list, D.num_super);
    ALLOC(D.super_state, super_tiles);
    ALLOC(D.block_page, cap_blocks);
    ALLOC(D.block_start, cap_blocks);
    ALLOC(D.block_count, cap_blocks);
    ALLOC(D.nbr, cap_blocks * 8);
    ALLOC(D.cell_count, D.cap_nodes);
    ALLOC(D.cell_start, D.cap_nodes);
    ALLOC(D.counters, 8);
    ALLOC(D.clock, 2 * ns);
    ALLOC(D.acc, (int64_t)NGF * D.cap_nodes);
    ALLOC(D.newv, 4 * D.cap_nodes);
    ALLOC(D.oldv, 4 * D.cap_nodes);
    ALLOC(D.scales, NSC * ns);
    ALLOC(D.maxbits, (int64_t)NSC * BOUND_SLOTS * ns);
    ALLOC(D.extent, (int64_t)(EXTENT_WGS + 1) * 8);
    if (h->desc.implicit_damage) {
        ALLOC(D.cgv, 5 * D.cap_nodes);
        ALLOC(D.cgp, 4 * cap);
        ALLOC(D.cg_part, 3 * CG_MAX_WGS);
        ALLOC(D.cg_bar, 4);
    }
#undef ALLOC
    if (h->desc.implicit_damage) {
        // k_cg's grid barriers need every workgroup resident: one per CU (the
        // occupancy calculator must allow at least that; fewer workgroups make
        // each barrier cheaper, ~9 us at 256)
        int occ = 0, cus = 0;
        HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_cg<T, AccT>, CGW, 0));
        HIPCHK(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->desc.device));
        if (occ < 1) return fail(h, ZMPM_E_HIP, "damage solve kernel cannot be resident");
        int coop = 0;
        HIPCHK(h, hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, h->desc.device));
        if (!coop) return fail(h, ZMPM_E_HIP, "the implicit damage solve needs cooperative launches");
        D.cg_wgs = std::max(1, std::min(CG_MAX_WGS, cus));
        HIPCHK(h, hipMemsetAsync(D.cg_bar, 0, 4 * sizeof(unsigned), h->stream));
    }
    {
        // k_rare walks its queue grid-stride: one resident round of workgroups
        // serves any queue, and an empty one costs a single round of starts
        int occ = 0, cus = 0;
        HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rare<T, AccT>, WG, 0));
        HIPCHK(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->desc.device));
        h->rare_wgs = std::max(1, occ * cus);
    }
    HIPCHK(h, hipMemsetAsync(D.touched, 0, D.num_pages * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.page_count, 0, D.num_pages * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.scan_state, 0, scan_tiles * sizeof(uint64_t), h->stream));
    HIPCHK(h, hipMemsetAsync(D.super_touched, 0, D.num_super * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.super_state, 0, super_tiles * sizeof(uint64_t), h->stream));
    HIPCHK(h, hipMemsetAsync(D.tmp, 0, cap * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.tmpk, 0, cap * sizeof(uint32_t), h->stream));
    HIPCHK(h, hipMemsetAsync(D.perm, 0, cap * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.sorted_cell, 0, cap * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.spid, 0, cap * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.counters, 0, 8 * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.maxbits, 0, ns * NSC * BOUND_SLOTS * sizeof(unsigned int), h->stream));
    const std::vector<float> ones((size_t)(NSC * ns), 1.f);
    HIPCHK(h, hipMemcpyAsync(D.scales, ones.data(), ones.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemsetAsync(D.clock, 0, 2 * ns * sizeof(double), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    D.masks.x = spreadMask(B::data_bits, 0);
    D.masks.y = spreadMask(B::data_bits, 1);
    D.masks.z = spreadMask(B::data_bits, 2);
    devOf<T, AccT>(h) = D;
    h->cap = cap;
    return ZMPM_OK;
}

// Collision objects of every scene, grouped by scene in insertion order.
size_t colliderCount(const zmpm_context* h)
{
    size_t c = 0;
    for (const auto& v : h->scene_cols) c += v.size();
    return c;
}

// The device scene table from the scene descriptions and the current per-scene
// dt (the host mirror stays alive for the asynchronous copy).
template <class T>
zmpm_status uploadScenes(zmpm_context* h)
{
    const size_t ns = h->scenes.size();
    h->scene_host.assign(ns * sizeof(SceneDev<T>), 0);
    SceneDev<T>* sd = reinterpret_cast<SceneDev<T>*>(h->scene_host.data());
    int col0 = 0;
    for (size_t s = 0; s < ns; s++) {
        const zmpm_scene_desc& d = h->scenes[s];
        SceneDev<T>& e = sd[s];
        e.dx = (T)d.dx;
        e.ddx = d.dx;
        e.odx = T(1) / (T)d.dx;
        e.pos_bits = posBits(d.domain_nodes);
        e.pos_scale = (T)((double)(1 << e.pos_bits) / d.dx);
        e.Dinv = T(4) / ((T)d.dx * (T)d.dx);
        for (int a = 0; a < 3; a++) e.gravity[a] = (T)d.gravity[a];
        e.rpic = (T)d.apic_rpic_ratio;
        e.cfl = (T)d.cfl;
        e.dt = (T)h->step_dt[s];
        e.ddt = h->step_dt[s];
        e.inv_cfl_dx = 1.0 / (d.cfl * d.dx);
        for (int a = 0; a < 3; a++) e.origin[a] = h->origin[s][a];
        e.domain = d.domain_nodes;
        e.col0 = col0;
        e.ncol = (int)h->scene_cols[s].size();
        col0 += e.ncol;
        e.ignore_cols = d.ignore_collision_objects ? 1 : 0;
        e.nbf = (int)h->body_forces[s].size();
        for (int k = 0; k < e.nbf; k++) {
            const auto& bf = h->body_forces[s][(size_t)k];
            e.bf_scale[k] = bf.scale;
            e.bf_freq[k] = bf.freq;
            for (int a = 0; a < 3; a++) e.bf_dir[k][a] = bf.dir[a];
        }
        const auto& vf = h->vfields[s];
        e.vfield = vf.on ? 1 : 0;
        for (int a = 0; a < 3; a++) e.va[a] = vf.a[a];
        for (int k = 0; k < 9; k++) e.vB[k] = vf.B[k];
    }
    HIPCHK(h, hipMemcpyAsync(h->scenes_dev, sd, ns * sizeof(SceneDev<T>), hipMemcpyHostToDevice, h->stream));
    return ZMPM_OK;
}

template <class T, class AccT>
void refreshTables(zmpm_context* h)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    D.mats = (const MatDev<T>*)h->mats_dev;
    D.cols = (const ColDev*)h->cols_dev;
    D.colstate = (ColState*)h->colstate_dev;
    D.num_cols = (int)colliderCount(h);
    D.sc = (const SceneDev<T>*)h->scenes_dev;
    D.bfval = (double*)h->bfval_dev;
    D.nscenes = (int)h->scenes.size();
    D.region_scene = (const int*)h->regions_dev;
    D.region_dim = h->region_dim;
    int rs = 0;
    while ((1 << rs) < h->region) rs++;
    D.region_shift = rs;
    D.any_damage = h->any_damage ? 1 : 0;
    D.any_plastic = 0;
    D.any_fj = 0;
    for (const auto& m : h->mats) {
        D.any_plastic |= m.plasticity != ZMPM_PLASTICITY_NONE ? 1 : 0;
        D.any_fj |= m.fj_mixed != ZMPM_FJ_NONE ? 1 : 0;
    }
    D.fj_solve_j = h->fj_state;
    D.p2g_split = h->p2g_split;
    D.flip = h->desc.transfer_scheme == ZMPM_FLIP_BLEND_PIC ? 1 : 0;
    D.mls = h->desc.mls_mpm ? 1 : 0;
    D.flip_ratio = (T)h->desc.flip_pic_ratio;
    D.implicit = h->desc.implicit_damage ? 1 : 0;
    // ConjugateGradient's relative tolerance for its scalar type (ConjugateGradient.h:70)
    D.cg_rel = std::is_same<T, float>::value ? 1e-6 : 1e-12;
    (void)uploadScenes<T>(h);
}

template <class T>
zmpm_status uploadMaterials(zmpm_context* h)
{
    std::vector<MatDev<T>> md(kMaxMaterials);
    for (size_t i = 0; i < h->mats.size(); i++) {
        const zmpm_material& m = h->mats[i];
        MatDev<T>& d = md[i];
        d.model = m.model;
        d.has_damage = m.has_damage;
        d.nd = m.num_directors == 2 ? 2 : 1;
        d.mu = (T)m.mu;
        d.lambda = (T)m.lambda;
        d.kx = (T)m.kx;
        d.ky = (T)m.ky;
        d.iso = m.isotropic ? 1 : 0;
        d.kappa = (T)(0.666666 * m.mu + m.lambda); // NeoHookeanBorden.cpp:15-21
        for (int c = 0; c < 3; c++) d.kf[c] = (T)m.fiber_stiffness[c];
        d.bulk = (T)m.bulk;
        d.gamma = (T)m.gamma;
        for (int k = 0; k < 2; k++) {
            for (int c = 0; c < 3; c++) d.a0[k][c] = (T)m.a0[k][c];
            d.alphas[k] = (T)m.alphas[k];
        }
        d.k = (T)m.residual_phase;
        d.l0 = (T)m.l0;
        d.sigma_crit = (T)m.sigma_crit;
        d.eta = (T)m.eta;
        d.zeta = (T)m.zeta;
        d.inv_sigma_crit2 = (T)(1.0 / (m.sigma_crit * m.sigma_crit));
        double na = 1.0;
        for (int k = 0; k < d.nd; k++)
            na += std::fabs(m.alphas[k]) * (m.a0[k][0] * m.a0[k][0] + m.a0[k][1] * m.a0[k][1] + m.a0[k][2] * m.a0[k][2]);
        // rounded up so the float form never under-estimates the bound
        d.phi_bound_scale = (T)(na * na / (m.sigma_crit * m.sigma_crit) * (1.0 + 1e-6));
        d.plasticity = m.plasticity;
        d.hardening = (T)m.hardening;
        d.volume_correction = m.volume_correction ? 1 : 0;
        d.cohesion = (T)m.cohesion;
        // DruckerPragerStvkHencky::DruckerPragerStvkHencky (PlasticityApplier.cpp:811-819)
        const double sin_phi = std::sin(m.friction_angle / 180.0 * 3.141592653);
        d.dp_alpha = (T)(std::sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi));
        // SnowPlasticity (PlasticityApplier.cpp:16-50)
        d.snow_lo = 1.0 - m.critical_compression;
        d.snow_hi = 1.0 + m.critical_stretch;
        d.min_jp = m.min_jp;
        d.max_jp = m.max_jp;
        d.snow_lam_lo = (float)(d.snow_lo * d.snow_lo - 1.0);
        d.snow_lam_hi = (float)(d.snow_hi * d.snow_hi - 1.0);
        // addFJMixedMpmForce (MpmParticleHandleBase.cpp:811-834): a fluid phase
        // zeroes the F model's mu and lambda
        d.fj = m.fj_mixed;
        d.jl = (T)m.j_lambda;
        if (m.fj_mixed == ZMPM_FJ_FLUID) d.mu = d.lambda = T(0);
    }
    HIPCHK(h, hipMemcpyAsync(h->mats_dev, md.data(), sizeof(MatDev<T>) * kMaxMaterials, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ZMPM_OK;
}

zmpm_status uploadColliders(zmpm_context* h)
{
    std::vector<ColDev> cd(std::max<size_t>(colliderCount(h), 1));
    std::vector<std::pair<int, const zmpm_collider*>> flat;
    for (size_t sc = 0; sc < h->scene_cols.size(); sc++)
        for (const zmpm_collider& c : h->scene_cols[sc]) flat.push_back({ (int)sc, &c });
    for (size_t i = 0; i < flat.size(); i++) {
        const zmpm_collider& c = *flat[i].second;
        ColDev& d = cd[i];
        d.scene = flat[i].first;
        d.shape = c.shape;
        d.type = c.type;
        double nn = std::sqrt(c.normal[0] * c.normal[0] + c.normal[1] * c.normal[1] + c.normal[2] * c.normal[2]);
        for (int a = 0; a < 3; a++) {
            d.origin[a] = c.origin[a];
            d.n[a] = nn > 0 ? c.normal[a] / nn : 0.0; // HalfSpace ctor normalises (AnalyticLevelSet.cpp:318-323)
            d.bmin[a] = c.box_min[a];
            d.bmax[a] = c.box_max[a];
            d.b0[a] = c.b0[a];
            d.dbdt[a] = c.dbdt[a];
        }
        d.friction = c.friction;
        d.radius = c.radius;
        d.height = c.height;
        // the capped cylinder's primitive rotation and its inverse (Eigen's 3x3
        // inverse through the cofactor matrix, as R.inverse() forms it)
        quatToMatrix(c.shape_rotation, d.Rc);
        const double* R = d.Rc;
        const double cof[9] = { R[4] * R[8] - R[5] * R[7], R[5] * R[6] - R[3] * R[8], R[3] * R[7] - R[4] * R[6],
                                R[2] * R[7] - R[1] * R[8], R[0] * R[8] - R[2] * R[6], R[1] * R[6] - R[0] * R[7],
                                R[1] * R[5] - R[2] * R[4], R[2] * R[3] - R[0] * R[5], R[0] * R[4] - R[1] * R[3] };
        const double inv = 1.0 / (cof[0] * R[0] + cof[1] * R[1] + cof[2] * R[2]);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) d.Rci[3 * i + j] = cof[3 * j + i] * inv;
        const bool q_zero = c.rotation[0] == 0 && c.rotation[1] == 0 && c.rotation[2] == 0 && c.rotation[3] == 0;
        for (int a = 0; a < 4; a++) d.q0[a] = q_zero ? (a == 0 ? 1.0 : 0.0) : c.rotation[a];
        for (int a = 0; a < 3; a++) d.omega[a] = c.omega[a];
        d.s0 = c.scale0 == 0 ? 1.0 : c.scale0;
        d.dsdt = c.dsdt;
        // translation only: x - b(t) is already material space
        const bool identity = d.q0[0] == 1 && d.q0[1] == 0 && d.q0[2] == 0 && d.q0[3] == 0;
        d.moving = (identity && c.omega[0] == 0 && c.omega[1] == 0 && c.omega[2] == 0 && d.s0 == 1 && c.dsdt == 0) ? 0 : 1;
    }
    if (!flat.empty())
        HIPCHK(h, hipMemcpyAsync(h->cols_dev, cd.data(), sizeof(ColDev) * flat.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ZMPM_OK;
}

int gridFor(int64_t items, int per, int maxg)
{
    int64_t g = (items + per - 1) / per;
    if (g < 1) g = 1;
    return (int)std::min<int64_t>(g, maxg);
}

// P2G of the substep: the MLS form, or the standard form in the compiled
// variant the launch configuration selects (VARIANT_* bits).
template <class T, class AccT>
void launchP2G(zmpm_context* h, const Dev<T, AccT>& D, hipStream_t st)
{
    const int wgs = (int)std::min<int64_t>(D.cap_blocks, h->p2g_wgs);
    constexpr int OCC = P2GWaves<T>::value;
    const int v = h->variant;
    if (D.mls)
        k_p2g<T, AccT, P2G_MLS><<<wgs, WG, 0, st>>>(D);
    else if ((v & VARIANT_P2G_LOW_OCC) && (v & VARIANT_P2G_XCD))
        k_p2g<T, AccT, P2G_STD, OCC - 1, true><<<wgs, WG, 0, st>>>(D);
    else if (v & VARIANT_P2G_LOW_OCC)
        k_p2g<T, AccT, P2G_STD, OCC - 1, false><<<wgs, WG, 0, st>>>(D);
    else if (v & VARIANT_P2G_XCD)
        k_p2g<T, AccT, P2G_STD, OCC, true><<<wgs, WG, 0, st>>>(D);
    else
        k_p2g<T, AccT, P2G_STD><<<wgs, WG, 0, st>>>(D);
}

// The kernels of one substep stage (zmpm_stage_id); launchStep runs all five.
template <class T, class AccT>
void launchStage(zmpm_context* h, int stage)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    hipStream_t st = h->stream;
    switch (stage) {
    case ZMPM_STAGE_SORT: {
        const int gp = gridFor(D.n, WG, 1 << 30);
        // the touched super pages in order, then the scan of their pages
        k_super_scan<T, AccT><<<(int)((D.num_super + PSCAN_TILE - 1) / PSCAN_TILE), WG, 0, st>>>(D);
        k_page_scan<T, AccT><<<(int)std::min<int64_t>(D.num_super, kPageScanWgs), WG, 0, st>>>(D);
        const int prep = (int)std::min<int64_t>(D.cap_blocks, h->prep_wgs);
        k_scatter<T, AccT><<<gp + prep + 1, WG, 0, st>>>(D, gp, prep);
        break;
    }
    case ZMPM_STAGE_P2G:
        launchP2G<T, AccT>(h, D, st);
        break;
    case ZMPM_STAGE_GRID_UPDATE:
        if (h->variant & VARIANT_GRID_SPLIT) {
            k_grid<T, AccT, false><<<gridFor(D.cap_nodes, WG, h->grid_wgs), WG, 0, st>>>(D);
            if (D.num_cols > 0) k_collide<T, AccT><<<gridFor(D.cap_blocks, WG / MAX_COLLIDERS, h->grid_wgs), WG, 0, st>>>(D);
        }
        else
            k_grid<T, AccT, true><<<gridFor(D.cap_nodes, WG, h->grid_wgs), WG, 0, st>>>(D);
        break;
    case ZMPM_STAGE_DAMAGE_SOLVE:
        if (D.implicit && D.any_damage) {
            // AnisoFractureSimulation::solveDamageSystem (AnisoFractureSimulation.h:435-612)
            k_dtilde<T, AccT><<<(int)std::min<int64_t>(D.cap_blocks, h->p2g_wgs), WG, 0, st>>>(D);
            // a cooperative launch: the runtime starts it only with every workgroup
            // resident, which its grid barriers need (work of other streams or
            // processes cannot leave some of them waiting to start)
            void* args[] = { &D };
            const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_cg<T, AccT>),
                                                            dim3(D.cg_wgs), dim3(CGW), args, 0, st);
            if (e != hipSuccess) h->launch_err = e;
        }
        break;
    case ZMPM_STAGE_G2P: {
        if (D.flip)
            k_g2p<T, AccT, true><<<D.max_slots, WG, 0, st>>>(D);
        else if (h->variant & VARIANT_G2P_RR)
            k_g2p<T, AccT, false, true><<<D.max_slots, WG, 0, st>>>(D);
        else
            k_g2p<T, AccT, false><<<D.max_slots, WG, 0, st>>>(D);
        k_rare<T, AccT><<<std::min(D.max_slots, h->rare_wgs), WG, 0, st>>>(D);
        std::swap(D.P, D.Q);
        std::swap(D.pid, D.qid);
        std::swap(D.pmat, D.qmat);
        // RPIC / PIC damping passes (MpmSimulationBase.cpp:318-348) on the moved
        // particles, through this substep's binning
        const int damping = D.flip ? h->desc.pic_damping_iteration : h->desc.rpic_damping_iteration;
        constexpr bool FIXED = std::is_same<AccT, long long>::value;
        for (int it = 0; it < damping; it++) {
            k_damp_prep<T, AccT><<<gridFor(D.cap_nodes, WG, h->prep_wgs) + (FIXED ? 1 : 0), WG, 0, st>>>(D);
            k_p2g<T, AccT, P2G_DAMP><<<(int)std::min<int64_t>(D.cap_blocks, h->p2g_wgs), WG, 0, st>>>(D);
            k_damp_grid<T, AccT><<<gridFor(D.cap_nodes, WG, h->grid_wgs), WG, 0, st>>>(D);
            k_damp_g2p<T, AccT><<<D.max_slots, WG, 0, st>>>(D);
        }
        break;
    }
    default:
        break;
    }
}

// One substep: the five stages, with the timing events between them when given
// (events 0..5 bracket sort, P2G, grid update, damage solve, G2P).
template <class T, class AccT>
void launchStep(zmpm_context* h, hipEvent_t* ev)
{
    for (int stage = ZMPM_STAGE_SORT; stage <= ZMPM_STAGE_G2P; stage++) {
        if (ev) (void)hipEventRecord(ev[stage], h->stream);
        launchStage<T, AccT>(h, stage);
    }
    if (ev) (void)hipEventRecord(ev[5], h->stream);
}

zmpm_status deviceStatus(zmpm_context* h, bool sync)
{
    if (sync) HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipGetLastError());
    int* counters = h->fp32 ? h->Df.counters : h->Dd.counters;
    if (!counters) return ZMPM_OK;
    int c[2] = { 0, 0 };
    HIPCHK(h, hipMemcpy(c, counters, sizeof(c), hipMemcpyDeviceToHost));
    int bits = c[1];
    if (bits & ST_NAN) return fail(h, ZMPM_E_NAN, "non-finite particle position");
    if (bits & ST_DOMAIN) return fail(h, ZMPM_E_DOMAIN, "particle base node outside [0, domain_nodes-5]");
    if (bits & ST_OVERFLOW) return fail(h, ZMPM_E_OVERFLOW, "fixed-point P2G contribution out of range");
    if (bits & ST_CAPACITY) return fail(h, ZMPM_E_NOMEM, "grid block capacity exceeded (set ZMPM_BLOCK_CAPACITY)");
    if (bits & ST_SCAN) return fail(h, ZMPM_E_HIP, "binning page scan did not complete");
    if (bits & ST_BARRIER) return fail(h, ZMPM_E_HIP, "damage solve grid barrier did not complete");
    if (bits & ST_CFL) return fail(h, ZMPM_E_CFL, "a particle moved more than dx in one substep");
    return ZMPM_OK;
}

// FP32-form position in fixed point, X = round(x / dx * 2^pb) in the scene's
// frame: POS_NAN for a non-finite x, clamped to [POS_MIN, POS_MAX] (outside the
// domain either way, which the first substep reports).
int encodePos(double x, double dx, int pb)
{
    if (!std::isfinite(x)) return POS_NAN;
    const double u = std::nearbyint(x / dx * (double)(1 << pb));
    if (!(u >= (double)POS_MIN)) return POS_MIN;
    if (!(u <= (double)POS_MAX)) return POS_MAX;
    return (int)u;
}

double decodePos(int X, double dx, int pb)
{
    if (X == POS_NAN) return std::numeric_limits<double>::quiet_NaN();
    return (double)X * dx / (double)(1 << pb);
}

template <class T, class AccT>
zmpm_status uploadT(zmpm_context* h, const zmpm_particles* p, const int32_t* scene)
{
    const int n = (int)p->n;
    zmpm_status s;
    if (n > h->cap || devOf<T, AccT>(h).P == nullptr)
        if ((s = allocate<T, AccT>(h, n)) != ZMPM_OK) return s;
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    D.n = n;
    h->n = n;
    // unless configured: collision objects in a launch of their own on grids
    // large enough for the grid update's doubled occupancy to outweigh the
    // second launch (batches; single scenes from 4M particles)
    if (!h->launch_set) {
        h->variant = (h->variant & ~VARIANT_GRID_SPLIT) | (h->scenes.size() > 1 || n >= (1 << 22) ? VARIANT_GRID_SPLIT : 0);
        // P2G workgroups: on large grids several times the resident ones, so
        // the blocks' uneven work evens out across the device (one per ~1k
        // particles, 2048 to 16384)
        if (!h->p2g_wgs_env) {
            int w = 2048;
            while (w < 16384 && (int64_t)w * 2 * 512 <= (int64_t)n) w *= 2;
            h->p2g_wgs = w;
        }
    }
    refreshTables<T, AccT>(h);
    const int64_t cap = D.cap;
    constexpr bool F32 = std::is_same<T, float>::value;
    std::vector<T> buf((size_t)NPF * cap, T(0));
    h->pscene.assign((size_t)n, 0);
    for (int i = 0; i < n; i++) {
        const int sc = scene ? scene[i] : 0;
        h->pscene[(size_t)i] = sc;
        const zmpm_scene_desc& sd = h->scenes[(size_t)sc];
        for (int a = 0; a < 3; a++) {
            if constexpr (F32) {
                const int X = encodePos(p->x[3 * i + a], sd.dx, posBits(sd.domain_nodes));
                std::memcpy(&buf[(PX + a) * cap + i], &X, sizeof(int));
            }
            else
                buf[(PX + a) * cap + i] = (T)p->x[3 * i + a]; // metres in the scene's frame
            buf[(PV + a) * cap + i] = (T)p->v[3 * i + a];
        }
        for (int k = 0; k < 9; k++) {
            buf[(PC + k) * cap + i] = (T)p->C[9 * i + k];
            // FP32 form: H = F - I, formed in double
            buf[(PF + k) * cap + i] = F32 ? (T)(p->F[9 * i + k] - (k % 4 == 0 ? 1.0 : 0.0)) : (T)p->F[9 * i + k];
        }
        buf[PM * cap + i] = (T)p->m[i];
        buf[PVOL * cap + i] = (T)p->V0[i];
        buf[PD * cap + i] = p->d ? (T)p->d[i] : T(0);
        buf[PG * cap + i] = p->g ? (T)p->g[i] : T(1);
        const int mi = p->material[i];
        const int pl = mi >= 0 && mi < (int)h->mats.size() ? h->mats[(size_t)mi].plasticity : ZMPM_PLASTICITY_NONE;
        const T ys0 = pl == ZMPM_PLASTICITY_VON_MISES_STVK_HENCKY ? (T)h->mats[(size_t)mi].yield_stress
            : pl == ZMPM_PLASTICITY_SNOW                          ? T(1)
                                                                  : T(0);
        buf[PYS * cap + i] = p->plastic_state ? (T)p->plastic_state[i] : ys0;
        buf[PDT * cap + i] = p->max_dtilde ? (T)p->max_dtilde[i] : T(0);
        // J (addFJMixedMpmForce starts it at 1); the FP32 form keeps J - 1
        const double J = p->J ? p->J[i] : 1.0;
        buf[PJ * cap + i] = F32 ? (T)(J - 1.0) : (T)J;
    }
    std::vector<int> ids(cap, 0), mat(cap, 0);
    for (int i = 0; i < n; i++) {
        ids[i] = i;
        mat[i] = p->material[i] | (h->pscene[(size_t)i] << PMAT_SCENE_SHIFT);
    }
    HIPCHK(h, hipMemcpyAsync(D.P, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(D.pid, ids.data(), cap * sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(D.pmat, mat.data(), cap * sizeof(int), hipMemcpyHostToDevice, h->stream));
    // block count and status; the scan epoch and ticket (slots 2, 3) keep counting
    HIPCHK(h, hipMemsetAsync(D.counters, 0, 2 * sizeof(int), h->stream));
    // G2P workgroups (WG particles each); the bound slots start cleared, and so do
    // the page flags and counts the last G2P may have left for a substep that
    // did not run
    D.max_slots = std::max(1, (n + WG - 1) / WG);
    HIPCHK(h, hipMemsetAsync(D.maxbits, 0, h->scenes.size() * NSC * BOUND_SLOTS * sizeof(unsigned int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.touched, 0, D.num_pages * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.page_count, 0, D.num_pages * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(D.super_touched, 0, D.num_super * sizeof(int), h->stream));
    if (n > 0) {
        // stress of the uploaded state and the first substep's fixed-point bounds
        k_stress_init<T, AccT><<<D.max_slots, WG, 0, h->stream>>>(D, 1, 1);
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipGetLastError());
    h->stepped = false;
    return ZMPM_OK;
}

template <class T, class AccT>
zmpm_status restress(zmpm_context* h)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    HIPCHK(h, hipMemsetAsync(D.maxbits, 0, h->scenes.size() * NSC * BOUND_SLOTS * sizeof(unsigned int), h->stream));
    k_stress_init<T, AccT><<<D.max_slots, WG, 0, h->stream>>>(D, 0, h->stepped ? 0 : 1);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipGetLastError());
    return ZMPM_OK;
}

template <class T, class AccT>
zmpm_status downloadT(zmpm_context* h, zmpm_particles* p)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    const int n = h->n;
    const int64_t cap = D.cap;
    std::vector<T> buf((size_t)NPF * cap);
    std::vector<int> ids(cap), mat(cap);
    HIPCHK(h, hipMemcpyAsync(buf.data(), D.P, buf.size() * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(ids.data(), D.pid, cap * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(mat.data(), D.pmat, cap * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int s = 0; s < n; s++) {
        const int i = ids[s];
        if (i < 0 || i >= n) return fail(h, ZMPM_E_HIP, "corrupt particle id table");
        const int sc = mat[s] >> PMAT_SCENE_SHIFT;
        if (sc < 0 || sc >= (int)h->scenes.size()) return fail(h, ZMPM_E_HIP, "corrupt particle scene");
        for (int a = 0; a < 3; a++) {
            if (p->x) {
                if constexpr (std::is_same<T, float>::value) {
                    int X;
                    std::memcpy(&X, &buf[(PX + a) * cap + s], sizeof(int));
                    const zmpm_scene_desc& sd = h->scenes[(size_t)sc];
                    p->x[3 * i + a] = decodePos(X, sd.dx, posBits(sd.domain_nodes));
                }
                else
                    p->x[3 * i + a] = (double)buf[(PX + a) * cap + s];
            }
            if (p->v) p->v[3 * i + a] = (double)buf[(PV + a) * cap + s];
        }
        for (int k = 0; k < 9; k++) {
            if (p->C) p->C[9 * i + k] = (double)buf[(PC + k) * cap + s];
            if (p->F)
                p->F[9 * i + k] = (double)buf[(PF + k) * cap + s]
                    + (std::is_same<T, float>::value && k % 4 == 0 ? 1.0 : 0.0);
        }
        if (p->m) p->m[i] = (double)buf[PM * cap + s];
        if (p->V0) p->V0[i] = (double)buf[PVOL * cap + s];
        if (p->material) p->material[i] = mat[s] & PMAT_MAT_MASK;
        if (p->d) p->d[i] = (double)buf[PD * cap + s];
        if (p->g) p->g[i] = (double)buf[PG * cap + s];
        if (p->laplacian) p->laplacian[i] = (double)buf[PLAP * cap + s];
        if (p->plastic_state) p->plastic_state[i] = (double)buf[PYS * cap + s];
        if (p->max_dtilde) p->max_dtilde[i] = h->desc.implicit_damage ? (double)buf[PDT * cap + s] : 0.0;
        if (p->J) p->J[i] = (double)buf[PJ * cap + s] + (std::is_same<T, float>::value ? 1.0 : 0.0);
    }
    return ZMPM_OK;
}

template <class T, class AccT>
zmpm_status binningT(zmpm_context* h, int64_t* n_groups, uint64_t* keys, int32_t* order, int32_t* groups,
                     uint64_t* block_offset)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    using B = Blk<T>;
    constexpr int bbits = B::xbits + B::ybits + B::zbits;
    constexpr int NB = 1 << bbits;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // the in-block order is complete after a substep's P2G
    if (!h->stepped && h->stage_next <= ZMPM_STAGE_P2G) {
        *n_groups = 0;
        return ZMPM_OK;
    }
    int nb = 0;
    HIPCHK(h, hipMemcpy(&nb, D.counters, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> bpage(nb), cstart((size_t)nb * NB), ccount((size_t)nb * NB);
    HIPCHK(h, hipMemcpy(bpage.data(), D.block_page, nb * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(cstart.data(), D.cell_start, cstart.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ccount.data(), D.cell_count, ccount.size() * sizeof(int), hipMemcpyDeviceToHost));
    int64_t G = 0;
    for (int b = 0; b < nb; b++) {
        int tot = 0;
        for (int e = 0; e < NB; e++) tot += ccount[(size_t)b * NB + e];
        if (tot == 0) continue;
        if (groups) {
            groups[2 * G] = cstart[(size_t)b * NB];
            groups[2 * G + 1] = cstart[(size_t)b * NB] + tot - 1;
        }
        if (block_offset) block_offset[G] = (uint64_t)(uint32_t)bpage[b];
        G++;
    }
    *n_groups = G;
    if (keys || order) {
        const int n = h->n;
        std::vector<int> ids(n), scell(n);
        HIPCHK(h, hipMemcpy(ids.data(), D.pid, n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(scell.data(), D.sorted_cell, n * sizeof(int), hipMemcpyDeviceToHost));
        const int index_bits = 32 - bbits;
        for (int s = 0; s < n; s++) {
            if (order) order[s] = ids[s];
            if (keys) {
                const int c = scell[s];
                const uint64_t page = (uint64_t)(uint32_t)bpage[c / NB];
                const uint64_t elem = (uint64_t)(c % NB);
                keys[s] = (((page << bbits) | elem) << index_bits) + (uint64_t)ids[s];
            }
        }
    }
    return ZMPM_OK;
}

// The grid of the substep in progress or last run, on the host: the nodes with
// mass ordered by (x, y, z) node coordinates (the order zmpm_get_grid reports).
template <class T, class AccT>
struct HostGrid {
    struct Rec {
        uint32_t c[3];
        int64_t node;
        int scene;
    };
    int64_t nn = 0;
    std::vector<AccT> acc;
    std::vector<T> nv; // [nn][4] node records: new velocity, phase field
    std::vector<float> sc;
    std::vector<Rec> recs;
    double val(int f, const Rec& r, int g) const
    {
        if (std::is_same<AccT, long long>::value) return (double)acc[f * nn + r.node] / (double)sc[(size_t)r.scene * NSC + g];
        return (double)acc[f * nn + r.node];
    }
};

template <class T, class AccT>
zmpm_status collectGrid(zmpm_context* h, HostGrid<T, AccT>& G)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    int nb = 0;
    HIPCHK(h, hipMemcpy(&nb, D.counters, sizeof(int), hipMemcpyDeviceToHost));
    const int64_t nn = (int64_t)nb * NB;
    G.nn = nn;
    std::vector<int> bpage(nb);
    G.acc.assign((size_t)7 * nn, AccT(0));
    G.nv.assign((size_t)4 * nn, T(0));
    const size_t ns = h->scenes.size();
    G.sc.assign(NSC * ns, 1.f);
    HIPCHK(h, hipMemcpy(bpage.data(), D.block_page, nb * sizeof(int), hipMemcpyDeviceToHost));
    for (int f = 0; f < 7; f++)
        HIPCHK(h, hipMemcpy(G.acc.data() + f * nn, D.acc + f * D.cap_nodes, nn * sizeof(AccT), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(G.nv.data(), D.newv, 4 * nn * sizeof(T), hipMemcpyDeviceToHost));
    // scales of the last substep's P2G (they are replaced only at the start of the next one)
    if (std::is_same<AccT, long long>::value)
        HIPCHK(h, hipMemcpy(G.sc.data(), D.scales, G.sc.size() * sizeof(float), hipMemcpyDeviceToHost));
    // scene of a node from its region
    auto sceneAt = [&](const uint32_t* c) -> int {
        const int rd = h->region_dim;
        const int r0 = (int)c[0] / h->region, r1 = (int)c[1] / h->region, r2 = (int)c[2] / h->region;
        if (r0 >= rd || r1 >= rd || r2 >= rd) return 0;
        return std::max(0, h->region_scene[(size_t)((r0 * rd + r1) * rd + r2)]);
    };
    using Rec = typename HostGrid<T, AccT>::Rec;
    G.recs.clear();
    for (int64_t node = 0; node < nn; node++) {
        const int b = (int)(node / NB), e = (int)(node % NB);
        uint64_t off = ((uint64_t)(uint32_t)bpage[b] << 12) | ((uint64_t)e << B::data_bits);
        Rec r { { hostBitPack(off, D.masks.x), hostBitPack(off, D.masks.y), hostBitPack(off, D.masks.z) }, node, 0 };
        r.scene = sceneAt(r.c);
        if (G.val(GM, r, SC_M) == 0) continue;
        G.recs.push_back(r);
    }
    std::sort(G.recs.begin(), G.recs.end(), [](const Rec& a, const Rec& b) {
        if (a.c[0] != b.c[0]) return a.c[0] < b.c[0];
        if (a.c[1] != b.c[1]) return a.c[1] < b.c[1];
        return a.c[2] < b.c[2];
    });
    return ZMPM_OK;
}

template <class T, class AccT>
zmpm_status gridT(zmpm_context* h, int64_t* n_nodes, int32_t* coords, double* m, double* v, double* new_v, double* pf)
{
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // a grid exists after a substep's P2G (a whole substep or its P2G stage)
    if (!h->stepped && h->stage_next <= ZMPM_STAGE_P2G) {
        *n_nodes = 0;
        return ZMPM_OK;
    }
    HostGrid<T, AccT> G;
    const zmpm_status st = collectGrid<T, AccT>(h, G);
    if (st != ZMPM_OK) return st;
    *n_nodes = (int64_t)G.recs.size();
    for (size_t q = 0; q < G.recs.size(); q++) {
        const int64_t node = G.recs[q].node;
        const double mm = G.val(GM, G.recs[q], SC_M);
        if (coords)
            for (int a = 0; a < 3; a++) coords[3 * q + a] = (int32_t)G.recs[q].c[a];
        if (m) m[q] = mm;
        if (v)
            for (int a = 0; a < 3; a++) v[3 * q + a] = G.val(GMV + a, G.recs[q], SC_MV) / mm;
        if (new_v)
            for (int a = 0; a < 3; a++) new_v[3 * q + a] = (double)G.nv[4 * node + a];
        if (pf) pf[q] = h->any_damage ? (double)G.nv[4 * node + 3] : 0.0;
    }
    return ZMPM_OK;
}

template <class T, class AccT>
zmpm_status setGridVelocityT(zmpm_context* h, int64_t n_nodes, const double* new_v)
{
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HostGrid<T, AccT> G;
    const zmpm_status st = collectGrid<T, AccT>(h, G);
    if (st != ZMPM_OK) return st;
    if (n_nodes != (int64_t)G.recs.size()) return fail(h, ZMPM_E_ARG, "node count differs from zmpm_get_grid's");
    const int64_t nn = G.nn;
    for (size_t q = 0; q < G.recs.size(); q++)
        for (int a = 0; a < 3; a++) G.nv[4 * G.recs[q].node + a] = (T)new_v[3 * q + a];
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    HIPCHK(h, hipMemcpy(D.newv, G.nv.data(), 4 * nn * sizeof(T), hipMemcpyHostToDevice));
    return ZMPM_OK;
}

} // namespace

// ================================================================= C ABI
namespace {
// Device state a substep reads or leaves for the next one: live particle
// buffers, the binning the last G2P (or the upload) did (keys, slots in the
// pages, page flags and counts), bound slots, scales, counters and clock.
template <class T, class AccT>
std::vector<std::pair<void*, size_t>> liveState(zmpm_context* h)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    return { { D.P, (size_t)NPF * D.cap * sizeof(T) },      { D.pid, (size_t)D.cap * sizeof(int) },
             { D.pmat, (size_t)D.cap * sizeof(int) },        { D.key, (size_t)D.cap * sizeof(uint64_t) },
             { D.rank, (size_t)D.cap * sizeof(int) },
             { D.touched, (size_t)D.num_pages * sizeof(int) }, { D.page_count, (size_t)D.num_pages * sizeof(int) },
             { D.super_touched, (size_t)D.num_super * sizeof(int) },
             { D.maxbits, h->scenes.size() * NSC * BOUND_SLOTS * sizeof(unsigned) },
             { D.scales, h->scenes.size() * NSC * sizeof(float) },
             { D.counters, 8 * sizeof(int) },                { D.clock, 2 * h->scenes.size() * sizeof(double) } };
}

template <class T, class AccT>
void clearScanState(zmpm_context* h)
{
    Dev<T, AccT>& D = devOf<T, AccT>(h);
    (void)hipMemsetAsync(D.scan_state, 0, D.num_super * sizeof(uint64_t), h->stream);
    (void)hipMemsetAsync(D.super_state, 0, (D.num_super + PSCAN_TILE - 1) / PSCAN_TILE * sizeof(uint64_t), h->stream);
}

} // namespace

extern "C" {

int32_t zmpm_abi_version(void) { return ZMPM_ABI_VERSION; }

const char* zmpm_build_info(void)
{
    return "zmpm HIP library: explicit APIC MPM substep + AnisoMPM damage; device code gfx950";
}

} // extern "C"

namespace {

// A scene description the reference's assertions accept
// (MpmSimulationBase::initialize / advanceOneTimeStep, MpmSimulationBase.cpp:108-128, 319, 335).
bool validScene(const zmpm_scene_desc* d)
{
    const int dn = d->domain_nodes;
    if (dn < 64 || dn > 4096 || (dn & (dn - 1)) || !(d->dx > 0) || !(d->cfl > 0)) return false;
    if (d->precision != ZMPM_PREC_FP32 && d->precision != ZMPM_PREC_FP64) return false;
    if (d->transfer_scheme != ZMPM_APIC_BLEND_RPIC && d->transfer_scheme != ZMPM_FLIP_BLEND_PIC) return false;
    const bool apic = d->transfer_scheme == ZMPM_APIC_BLEND_RPIC;
    if (d->mls_mpm && !apic) return false;
    if (apic && !(d->apic_rpic_ratio >= 0 && d->apic_rpic_ratio <= 1)) return false;
    if (d->rpic_damping_iteration < 0 || d->pic_damping_iteration < 0) return false;
    if ((d->rpic_damping_iteration > 0 && !apic) || (d->pic_damping_iteration > 0 && apic)) return false;
    if (d->implicit_damage != 0 && d->implicit_damage != 1) return false;
    return true;
}

// The options every scene of a batch shares (one launch sequence runs them all).
bool sameSharedOptions(const zmpm_scene_desc& a, const zmpm_scene_desc& b)
{
    return a.precision == b.precision && a.device == b.device && a.transfer_scheme == b.transfer_scheme
        && a.flip_pic_ratio == b.flip_pic_ratio && a.mls_mpm == b.mls_mpm
        && a.rpic_damping_iteration == b.rpic_damping_iteration && a.pic_damping_iteration == b.pic_damping_iteration
        && a.implicit_damage == b.implicit_damage;
}

uint64_t hostDeposit(uint32_t v, uint64_t mask)
{
    uint64_t r = 0;
    int bit = 0;
    for (int b = 0; b < 64; b++)
        if (mask & (1ull << b)) {
            if (v & (1u << bit)) r |= 1ull << b;
            bit++;
        }
    return r;
}

zmpm_status createScenes(const zmpm_scene_desc* descs, int32_t n_scenes, zmpm_handle* out)
{
    if (!descs || !out || n_scenes < 1 || n_scenes > 4096) return ZMPM_E_ARG;
    for (int s = 0; s < n_scenes; s++) {
        if (!validScene(&descs[s])) return ZMPM_E_ARG;
        if (!sameSharedOptions(descs[0], descs[s])) return ZMPM_E_ARG;
    }
    // the implicit damage solve couples every node of the handle: one scene only
    if (n_scenes > 1 && descs[0].implicit_damage) return ZMPM_E_ARG;
    // regions: cubes of the largest scene domain, region_dim per axis
    int region = 0;
    for (int s = 0; s < n_scenes; s++) region = std::max(region, descs[s].domain_nodes);
    int rd = 1;
    while (rd * rd * rd < n_scenes) rd++;
    int grid = 1;
    while (grid < rd * region) grid <<= 1;
    if (grid > 4096) return ZMPM_E_ARG; // the SPGrid range (MpmGrid.h:111-112)
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ZMPM_E_HIP;
    const int device = descs[0].device;
    if (device < 0 || device >= ndev) return ZMPM_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return ZMPM_E_HIP;
    // the allocation-free entry point's scratch address on this device (a
    // failure here only disables that entry point)
    (void)setUpNoallocScratch(device);
    auto* h = new zmpm_context();
    h->desc = descs[0];
    h->scenes.assign(descs, descs + n_scenes);
    h->scene_cols.assign((size_t)n_scenes, {});
    h->body_forces.assign((size_t)n_scenes, {});
    h->vfields.assign((size_t)n_scenes, {});
    h->step_dt.assign((size_t)n_scenes, 0.0);
    h->region = region;
    h->region_dim = rd;
    h->grid_domain = grid;
    h->fp32 = descs[0].precision == ZMPM_PREC_FP32;
    // regions in page (Morton) order of their first node, so a scene's grid
    // blocks are contiguous in the dense block order; scene s takes the s-th
    const int log2s = h->fp32 ? 6 : 7;
    const uint64_t mx = spreadMask(log2s, 0), my = spreadMask(log2s, 1), mz = spreadMask(log2s, 2);
    std::vector<std::pair<uint64_t, int>> order;
    for (int r = 0; r < rd * rd * rd; r++) {
        const int r0 = r / (rd * rd), r1 = (r / rd) % rd, r2 = r % rd;
        const uint64_t off = hostDeposit((uint32_t)(r0 * region), mx) | hostDeposit((uint32_t)(r1 * region), my)
            | hostDeposit((uint32_t)(r2 * region), mz);
        order.push_back({ off, r });
    }
    std::sort(order.begin(), order.end());
    h->region_scene.assign((size_t)(rd * rd * rd), -1);
    h->origin.assign((size_t)n_scenes, { 0, 0, 0 });
    for (int s = 0; s < n_scenes; s++) {
        const int r = order[(size_t)s].second;
        h->region_scene[(size_t)r] = s;
        h->origin[(size_t)s] = { (r / (rd * rd)) * region, ((r / rd) % rd) * region, (r % rd) * region };
    }
    if (const char* e = std::getenv("ZMPM_NO_GRAPH")) h->use_graphs = std::atoi(e) == 0;
    if (const char* e = std::getenv("ZMPM_P2G_WGS")) {
        h->p2g_wgs = std::max(1, std::atoi(e));
        h->p2g_wgs_env = true;
    }
    if (const char* e = std::getenv("ZMPM_GRID_WGS")) h->grid_wgs = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("ZMPM_PREP_WGS")) h->prep_wgs = std::max(1, std::atoi(e));
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return ZMPM_E_HIP;
    }
    const size_t msz = h->fp32 ? sizeof(MatDev<float>) : sizeof(MatDev<double>);
    const size_t ssz = h->fp32 ? sizeof(SceneDev<float>) : sizeof(SceneDev<double>);
    const size_t ncol = (size_t)kMaxColliders * (size_t)n_scenes;
    if (hipMalloc(&h->mats_dev, msz * kMaxMaterials) != hipSuccess || hipMalloc(&h->cols_dev, sizeof(ColDev) * ncol) != hipSuccess
        || hipMalloc(&h->colstate_dev, sizeof(ColState) * ncol) != hipSuccess
        || hipMalloc(&h->scenes_dev, ssz * (size_t)n_scenes) != hipSuccess
        || hipMalloc(&h->bfval_dev, sizeof(double) * 3 * MAX_BODY_FORCES * (size_t)n_scenes) != hipSuccess
        || hipMalloc(&h->regions_dev, sizeof(int) * h->region_scene.size()) != hipSuccess
        || hipMemcpy(h->regions_dev, h->region_scene.data(), sizeof(int) * h->region_scene.size(), hipMemcpyHostToDevice)
            != hipSuccess) {
        (void)hipGetLastError();
        for (void* q : { h->mats_dev, h->cols_dev, h->colstate_dev, h->scenes_dev, h->regions_dev, h->bfval_dev })
            if (q) (void)hipFree(q);
        (void)hipStreamDestroy(h->stream);
        delete h;
        return ZMPM_E_NOMEM;
    }
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    *out = h;
    return ZMPM_OK;
}

} // namespace

extern "C" {

zmpm_status zmpm_create(const zmpm_scene_desc* desc, zmpm_handle* out) { return createScenes(desc, 1, out); }

zmpm_status zmpm_create_batch(const zmpm_scene_desc* descs, int32_t n_scenes, zmpm_handle* out)
{
    return createScenes(descs, n_scenes, out);
}

zmpm_status zmpm_batch_origin(zmpm_handle h, int32_t scene, int32_t* origin)
{
    if (!h || !origin || scene < 0 || scene >= (int32_t)h->scenes.size()) return ZMPM_E_ARG;
    for (int a = 0; a < 3; a++) origin[a] = h->origin[(size_t)scene][a];
    return ZMPM_OK;
}

zmpm_status zmpm_destroy(zmpm_handle h)
{
    if (!h) return ZMPM_E_ARG;
    (void)hipSetDevice(h->desc.device);
    freeAll(h);
    for (hipEvent_t e : h->events) (void)hipEventDestroy(e);
    for (void* q : { h->mats_dev, h->cols_dev, h->colstate_dev, h->scenes_dev, h->regions_dev, h->bfval_dev })
        if (q) (void)hipFree(q);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return ZMPM_OK;
}

const char* zmpm_last_error(zmpm_handle h)
{
    if (!h) return "null handle";
    return h->err_static ? h->err_static : h->err.c_str();
}

zmpm_status zmpm_set_materials(zmpm_handle h, int32_t count, const zmpm_material* mats)
{
    if (!h) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (count <= 0 || count > kMaxMaterials || !mats) return fail(h, ZMPM_E_ARG, "bad material table");
    for (int i = 0; i < count; i++) {
        if (mats[i].model < 0 || mats[i].model > ZMPM_MODEL_NEO_HOOKEAN_BORDEN)
            return fail(h, ZMPM_E_ARG, "unknown constitutive model");
        if (mats[i].plasticity < ZMPM_PLASTICITY_NONE || mats[i].plasticity > ZMPM_PLASTICITY_SNOW)
            return fail(h, ZMPM_E_ARG, "unknown plasticity");
        if (mats[i].plasticity && mats[i].model == ZMPM_MODEL_EQUATION_OF_STATE)
            return fail(h, ZMPM_E_ARG, "plasticity needs a model with mu");
        if (mats[i].plasticity == ZMPM_PLASTICITY_SNOW && mats[i].model != ZMPM_MODEL_FIXED_COROTATED
            && mats[i].model != ZMPM_MODEL_NEO_HOOKEAN)
            return fail(h, ZMPM_E_ARG, "snow plasticity needs a stress linear in mu and lambda (fixed corotated, neo-Hookean)");
        if (mats[i].fj_mixed < ZMPM_FJ_NONE || mats[i].fj_mixed > ZMPM_FJ_FLUID)
            return fail(h, ZMPM_E_ARG, "unknown FJ-mixed phase");
        if (mats[i].fj_mixed
            && ((mats[i].model != ZMPM_MODEL_FIXED_COROTATED && mats[i].model != ZMPM_MODEL_NEO_HOOKEAN)
                || mats[i].has_damage || mats[i].plasticity == ZMPM_PLASTICITY_SNOW))
            return fail(h, ZMPM_E_ARG,
                        "an FJ-mixed material has a fixed-corotated or neo-Hookean F model, no phase field and no snow plasticity");
    }
    if (h->n > 0 && h->max_mat >= count)
        return fail(h, ZMPM_E_ARG, "material table smaller than the uploaded particles' material indices");
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    h->mats.assign(mats, mats + count);
    h->any_damage = false;
    for (auto& m : h->mats) h->any_damage = h->any_damage || m.has_damage;
    zmpm_status s = h->fp32 ? uploadMaterials<float>(h) : uploadMaterials<double>(h);
    if (s != ZMPM_OK) return s;
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    // particles already on the device: their stored stress and the next P2G's
    // fixed-point bounds follow the new table (the binning stays)
    if (h->n > 0) return h->fp32 ? restress<float, long long>(h) : restress<double, double>(h);
    return ZMPM_OK;
}

zmpm_status zmpm_add_body_force(zmpm_handle h, int32_t scene, double scale, double frequency, const double* direction)
{
    if (!h) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (scene < 0 || scene >= (int32_t)h->scenes.size()) return fail(h, ZMPM_E_ARG, "no such scene");
    if (!direction || !std::isfinite(scale) || !std::isfinite(frequency)) return fail(h, ZMPM_E_ARG, "bad body force");
    auto& list = h->body_forces[(size_t)scene];
    if ((int)list.size() >= MAX_BODY_FORCES) return fail(h, ZMPM_E_ARG, "at most 4 body forces per scene");
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    list.push_back({ scale, frequency, { direction[0], direction[1], direction[2] } });
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    return ZMPM_OK;
}

zmpm_status zmpm_set_velocity_field(zmpm_handle h, int32_t scene, const double* a, const double* B)
{
    if (!h) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (scene < 0 || scene >= (int32_t)h->scenes.size()) return fail(h, ZMPM_E_ARG, "no such scene");
    if (a && !B) return fail(h, ZMPM_E_ARG, "the velocity field needs B");
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    auto& vf = h->vfields[(size_t)scene];
    vf.on = a != nullptr;
    for (int k = 0; k < 3; k++) vf.a[k] = a ? a[k] : 0.0;
    for (int k = 0; k < 9; k++) vf.B[k] = a ? B[k] : 0.0;
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    return ZMPM_OK;
}

zmpm_status zmpm_set_fj_state(zmpm_handle h, int32_t state)
{
    if (!h) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (state != 0 && state != 1) return fail(h, ZMPM_E_ARG, "FJ state is 0 (SOLVE_F) or 1 (SOLVE_J)");
    if (state == h->fj_state) return ZMPM_OK;
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    h->fj_state = state;
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    // the stored stress of the particles on the device follows the new state
    if (h->n > 0) return h->fp32 ? restress<float, long long>(h) : restress<double, double>(h);
    return ZMPM_OK;
}

zmpm_status zmpm_add_scene_collider(zmpm_handle h, int32_t scene, const zmpm_collider* c, int32_t* id_out)
{
    if (!h || !c) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (scene < 0 || scene >= (int32_t)h->scenes.size()) return fail(h, ZMPM_E_ARG, "no such scene");
    auto& cols = h->scene_cols[(size_t)scene];
    if ((int)cols.size() >= kMaxColliders) return fail(h, ZMPM_E_ARG, "too many colliders");
    if (c->shape < ZMPM_SHAPE_HALF_SPACE || c->shape > ZMPM_SHAPE_CAPPED_CYLINDER) return fail(h, ZMPM_E_ARG, "unknown shape");
    if (c->type < ZMPM_STICKY || c->type > ZMPM_GHOST) return fail(h, ZMPM_E_ARG, "unknown collider type");
    if (!((c->scale0 == 0 ? 1.0 : c->scale0) > 0)) return fail(h, ZMPM_E_ARG, "collider scaling must be positive");
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    cols.push_back(*c);
    zmpm_status s = uploadColliders(h);
    if (s != ZMPM_OK) return s;
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    if (id_out) *id_out = (int32_t)cols.size() - 1;
    return ZMPM_OK;
}

zmpm_status zmpm_add_collider(zmpm_handle h, const zmpm_collider* c, int32_t* id_out)
{
    return zmpm_add_scene_collider(h, 0, c, id_out);
}

zmpm_status zmpm_clear_colliders(zmpm_handle h)
{
    if (!h) return ZMPM_E_ARG;
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    for (auto& v : h->scene_cols) v.clear();
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    return ZMPM_OK;
}

static zmpm_status uploadScenesParticles(zmpm_handle h, const zmpm_particles* p, const int32_t* scene)
{
    if (!h || !p || p->n < 0) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (p->n > 0 && (!p->x || !p->v || !p->C || !p->F || !p->m || !p->V0 || !p->material))
        return fail(h, ZMPM_E_ARG, "missing particle array");
    if (h->mats.empty()) return fail(h, ZMPM_E_ARG, "set materials before upload");
    const int bbits = h->fp32 ? 6 : 5;
    if (p->n >= ((int64_t)1 << (32 - bbits))) return fail(h, ZMPM_E_ARG, "too many particles for the key layout");
    int max_mat = -1;
    for (int64_t i = 0; i < p->n; i++) {
        if (p->material[i] < 0 || p->material[i] >= (int32_t)h->mats.size())
            return fail(h, ZMPM_E_ARG, "material index out of range");
        max_mat = std::max(max_mat, (int)p->material[i]);
        if (scene && (scene[i] < 0 || scene[i] >= (int32_t)h->scenes.size()))
            return fail(h, ZMPM_E_ARG, "scene index out of range");
    }
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    const zmpm_status s = h->fp32 ? uploadT<float, long long>(h, p, scene) : uploadT<double, double>(h, p, scene);
    if (s == ZMPM_OK) h->max_mat = max_mat;
    return s;
}

zmpm_status zmpm_upload(zmpm_handle h, const zmpm_particles* p) { return uploadScenesParticles(h, p, nullptr); }

zmpm_status zmpm_upload_batch(zmpm_handle h, const zmpm_particles* p, const int32_t* scene)
{
    if (!scene && p && p->n > 0) return fail(h, ZMPM_E_ARG, "missing scene array");
    return uploadScenesParticles(h, p, scene);
}

zmpm_status zmpm_download(zmpm_handle h, zmpm_particles* p)
{
    if (!h || !p) return ZMPM_E_ARG;
    if (p->n != h->n) return fail(h, ZMPM_E_ARG, "particle count mismatch");
    (void)hipSetDevice(h->desc.device);
    zmpm_status st = deviceStatus(h, true);
    if (st == ZMPM_E_HIP) return st;
    zmpm_status s = h->fp32 ? downloadT<float, long long>(h, p) : downloadT<double, double>(h, p);
    return s != ZMPM_OK ? s : st;
}

// `n` (even) substeps captured into one HIP graph (the particle buffers
// ping-pong, so an even count returns to the starting buffer); cached per live
// buffer and n.
static zmpm_status graphFor(zmpm_handle h, int n, hipGraphExec_t* out)
{
    const void* live = h->fp32 ? (const void*)h->Df.P : (const void*)h->Dd.P;
    for (auto& g : h->graphs)
        if (g.live == live && g.steps == n) {
            *out = g.exec;
            return ZMPM_OK;
        }
    if (h->graphs.size() >= 4) dropGraphs(h);
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(h, ZMPM_E_HIP, "hipStreamBeginCapture");
    for (int k = 0; k < n; k++) {
        if (h->fp32)
            launchStep<float, long long>(h, nullptr);
        else
            launchStep<double, double>(h, nullptr);
    }
    if (hipStreamEndCapture(h->stream, &graph) != hipSuccess) return fail(h, ZMPM_E_HIP, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(h, ZMPM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    h->graphs.push_back({ live, n, exec });
    *out = exec;
    return ZMPM_OK;
}

// `steps` substeps of every scene: scene s advances by dt[s] per substep from
// t0[s] (its device clock); the scene table is refreshed when a dt changes.
// Every scene's step dt[s] in the scene table and its device clock at t0[s].
static zmpm_status setClocks(zmpm_handle h, const double* dt, const double* t0)
{
    const size_t ns = h->scenes.size();
    bool dt_changed = false;
    for (size_t s = 0; s < ns; s++)
        if (h->step_dt[s] != dt[s]) {
            h->step_dt[s] = dt[s];
            dt_changed = true;
        }
    if (dt_changed) {
        const zmpm_status st = h->fp32 ? uploadScenes<float>(h) : uploadScenes<double>(h);
        if (st != ZMPM_OK) return st;
    }
    // each scene's device clock starts at its t0; every substep advances it by its dt
    h->clock_host.assign(2 * ns, 0.0);
    for (size_t s = 0; s < ns; s++) h->clock_host[2 * s] = h->clock_host[2 * s + 1] = t0[s];
    double* clock = h->fp32 ? h->Df.clock : h->Dd.clock;
    if (hipMemcpyAsync(clock, h->clock_host.data(), 2 * ns * sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess)
        return fail(h, ZMPM_E_HIP, "clock upload");
    return ZMPM_OK;
}

static zmpm_status runSteps(zmpm_handle h, int32_t steps, const double* dt, const double* t0)
{
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (h->n == 0) return ZMPM_OK;
    (void)hipSetDevice(h->desc.device);
    hipEvent_t* ev = nullptr;
    if (h->timing) {
        const size_t need = (size_t)steps * (kTimedGroups + 1);
        while (h->events.size() < need) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return fail(h, ZMPM_E_HIP, "hipEventCreate failed");
            h->events.push_back(e);
        }
    }
    {
        const zmpm_status st = setClocks(h, dt, t0);
        if (st != ZMPM_OK) return st;
    }
    int32_t s = 0;
    // substeps with the implicit damage solve run eagerly: its cooperative
    // launch is not captured into graphs
    const bool coop = h->fp32 ? (h->Df.implicit && h->Df.any_damage) : (h->Dd.implicit && h->Dd.any_damage);
    if (h->use_graphs && !h->timing && steps >= 2 && !coop) {
        // runs of kGraphSteps substeps, then pairs, replayed from captured graphs
        for (const int n : { kGraphSteps, 2 }) {
            if (s + n > steps) continue;
            hipGraphExec_t exec = nullptr;
            zmpm_status st = graphFor(h, n, &exec);
            if (st != ZMPM_OK) return st;
            for (; s + n <= steps; s += n)
                if (hipGraphLaunch(exec, h->stream) != hipSuccess) return fail(h, ZMPM_E_HIP, "hipGraphLaunch");
        }
    }
    for (; s < steps; s++) {
        if (h->timing) ev = &h->events[(size_t)s * (kTimedGroups + 1)];
        if (h->fp32)
            launchStep<float, long long>(h, ev);
        else
            launchStep<double, double>(h, ev);
    }
    h->stepped = true;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = h->launch_err;
    h->launch_err = hipSuccess;
    if (e != hipSuccess) return fail(h, ZMPM_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    if (h->timing) {
        if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(h, ZMPM_E_HIP, "stream sync");
        for (int g = 0; g <= kTimedGroups; g++) h->timed_ms[g] = 0;
        for (int32_t s = 0; s < steps; s++) {
            hipEvent_t* e5 = &h->events[(size_t)s * (kTimedGroups + 1)];
            for (int g = 0; g < kTimedGroups; g++) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e5[g], e5[g + 1]);
                h->timed_ms[g] += ms;
            }
            float tot = 0;
            (void)hipEventElapsedTime(&tot, e5[0], e5[kTimedGroups]);
            h->timed_ms[kTimedGroups] += tot;
        }
        h->timed_steps = steps;
    }
    return ZMPM_OK;
}

// Every scene with the same dt and start time.
static zmpm_status runUniform(zmpm_handle h, int32_t steps, double dt, double t0)
{
    const std::vector<double> dts(h->scenes.size(), dt), t0s(h->scenes.size(), t0);
    return runSteps(h, steps, dts.data(), t0s.data());
}

zmpm_status zmpm_substep(zmpm_handle h, double dt, double t)
{
    if (!h || !(dt > 0)) return ZMPM_E_ARG;
    return runUniform(h, 1, dt, t);
}

zmpm_status zmpm_run(zmpm_handle h, int32_t steps, double dt, double t0)
{
    if (!h || steps < 0 || !(dt > 0)) return ZMPM_E_ARG;
    return runUniform(h, steps, dt, t0);
}

zmpm_status zmpm_stage(zmpm_handle h, int32_t stage, double dt, double t)
{
    if (!h || stage < ZMPM_STAGE_SORT || stage > ZMPM_STAGE_G2P || !(dt > 0)) return ZMPM_E_ARG;
    if (h->scenes.size() != 1) return fail(h, ZMPM_E_ARG, "stages run on a single-scene handle");
    if (stage != h->stage_next)
        return fail(h, ZMPM_E_ARG, "stages run in order: SORT, P2G, GRID_UPDATE, DAMAGE_SOLVE, G2P");
    if (stage == ZMPM_STAGE_SORT) {
        h->stage_dt = dt;
        h->stage_t = t;
    }
    else if (dt != h->stage_dt || t != h->stage_t)
        return fail(h, ZMPM_E_ARG, "the stages of a substep share its dt and t");
    h->stage_next = stage == ZMPM_STAGE_G2P ? 0 : stage + 1;
    if (h->n == 0) return ZMPM_OK;
    (void)hipSetDevice(h->desc.device);
    if (stage == ZMPM_STAGE_SORT) {
        const zmpm_status st = setClocks(h, &dt, &t);
        if (st != ZMPM_OK) {
            h->stage_next = 0;
            return st;
        }
    }
    if (h->fp32)
        launchStage<float, long long>(h, stage);
    else
        launchStage<double, double>(h, stage);
    if (stage == ZMPM_STAGE_G2P) h->stepped = true;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = h->launch_err;
    h->launch_err = hipSuccess;
    if (e != hipSuccess) {
        h->stage_next = 0;
        return fail(h, ZMPM_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    }
    return ZMPM_OK;
}

zmpm_status zmpm_set_grid_velocity(zmpm_handle h, int64_t n_nodes, const double* new_v)
{
    if (!h || !new_v) return ZMPM_E_ARG;
    if (h->stage_next != ZMPM_STAGE_DAMAGE_SOLVE && h->stage_next != ZMPM_STAGE_G2P)
        return fail(h, ZMPM_E_ARG, "the grid velocities are replaced between the GRID_UPDATE and G2P stages");
    (void)hipSetDevice(h->desc.device);
    return h->fp32 ? setGridVelocityT<float, long long>(h, n_nodes, new_v) : setGridVelocityT<double, double>(h, n_nodes, new_v);
}

zmpm_status zmpm_run_batch(zmpm_handle h, int32_t steps, const double* dt, const double* t0)
{
    if (!h || steps < 0 || !dt || !t0) return ZMPM_E_ARG;
    for (size_t s = 0; s < h->scenes.size(); s++)
        if (!(dt[s] > 0)) return fail(h, ZMPM_E_ARG, "every scene's dt must be positive");
    return runSteps(h, steps, dt, t0);
}

zmpm_status zmpm_synchronize(zmpm_handle h)
{
    if (!h) return ZMPM_E_ARG;
    (void)hipSetDevice(h->desc.device);
    return deviceStatus(h, true);
}

zmpm_status zmpm_get_binning(zmpm_handle h, int64_t* n_groups, uint64_t* keys, int32_t* order, int32_t* groups,
                             uint64_t* block_offset)
{
    if (!h || !n_groups) return ZMPM_E_ARG;
    (void)hipSetDevice(h->desc.device);
    return h->fp32 ? binningT<float, long long>(h, n_groups, keys, order, groups, block_offset)
                   : binningT<double, double>(h, n_groups, keys, order, groups, block_offset);
}

zmpm_status zmpm_get_grid(zmpm_handle h, int64_t* n_nodes, int32_t* coords, double* m, double* v, double* new_v,
                          double* pf)
{
    if (!h || !n_nodes) return ZMPM_E_ARG;
    (void)hipSetDevice(h->desc.device);
    return h->fp32 ? gridT<float, long long>(h, n_nodes, coords, m, v, new_v, pf)
                   : gridT<double, double>(h, n_nodes, coords, m, v, new_v, pf);
}

zmpm_status zmpm_set_timing(zmpm_handle h, int32_t enable)
{
    if (!h) return ZMPM_E_ARG;
    h->timing = enable != 0;
    return ZMPM_OK;
}

zmpm_status zmpm_calculate_dt(zmpm_handle h, double max_dt, double* dt, double* max_speed, double* min_corner,
                              double* max_corner)
{
    if (!h || !(max_dt > 0)) return ZMPM_E_ARG;
    (void)hipSetDevice(h->desc.device);
    double e[8] = { 0, -3.4028234663852886e38, -3.4028234663852886e38, -3.4028234663852886e38,
                    -3.4028234663852886e38, -3.4028234663852886e38, -3.4028234663852886e38, 0 };
    double* ext = h->fp32 ? h->Df.extent : h->Dd.extent;
    if (h->n > 0 && ext) {
        if (h->fp32) {
            k_extent<float, long long><<<EXTENT_WGS, WG, 0, h->stream>>>(h->Df, 0, EXTENT_WGS, nullptr, 0.0, 0.0);
            k_extent<float, long long><<<1, WG, 0, h->stream>>>(h->Df, 1, EXTENT_WGS, nullptr, 0.0, 0.0);
        }
        else {
            k_extent<double, double><<<EXTENT_WGS, WG, 0, h->stream>>>(h->Dd, 0, EXTENT_WGS, nullptr, 0.0, 0.0);
            k_extent<double, double><<<1, WG, 0, h->stream>>>(h->Dd, 1, EXTENT_WGS, nullptr, 0.0, 0.0);
        }
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipMemcpyAsync(e, ext + EXTENT_WGS * 8, 8 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    if (max_speed) *max_speed = e[0];
    for (int a = 0; a < 3; a++) {
        if (max_corner) max_corner[a] = e[1 + a];
        if (min_corner) min_corner[a] = -e[4 + a];
    }
    // one scene: cfl dx / max speed (MpmSimulationBase.cpp:1127-1151); a batch:
    // the smallest scene step, 1 / max(speed / (cfl dx))
    if (dt) {
        if (h->scenes.size() == 1)
            *dt = e[0] != 0 ? h->desc.cfl * h->desc.dx / e[0] : max_dt;
        else
            *dt = e[7] != 0 ? 1.0 / e[7] : max_dt;
    }
    return ZMPM_OK;
}

zmpm_status zmpm_calculate_dt_noalloc(zmpm_handle h, double max_dt, int32_t reduction_wgs, double* out)
{
    if (!h) return ZMPM_E_ARG;
    if (!(max_dt > 0) || !out) return failStatic(h, ZMPM_E_ARG, "max_dt must be positive and out a device buffer");
    if (reduction_wgs < 1 || reduction_wgs > NOALLOC_MAX_WGS)
        return failStatic(h, ZMPM_E_ARG, "reduction_wgs must be in [1, 1024]: the static scratch holds 1024 partial rows");
    if (hipSetDevice(h->desc.device) != hipSuccess) return failStatic(h, ZMPM_E_HIP, "hipSetDevice failed");
    double* scr = (h->desc.device >= 0 && h->desc.device < kMaxDevices) ? g_scratch[h->desc.device] : nullptr;
    if (!scr) return failStatic(h, ZMPM_E_HIP, "static scratch not set up on this device");
    const double cfl_dx = h->desc.cfl * h->desc.dx;
    // the handle's kernel state with the partial maxima in the static scratch
    if (h->fp32) {
        Dev<float, long long> D = h->Df;
        D.extent = scr;
        if (h->n == 0) D.n = 0;
        k_extent<float, long long><<<reduction_wgs, WG, 0, h->stream>>>(D, 0, reduction_wgs, nullptr, max_dt, cfl_dx);
        k_extent<float, long long><<<1, WG, 0, h->stream>>>(D, 1, reduction_wgs, out, max_dt, cfl_dx);
    }
    else {
        Dev<double, double> D = h->Dd;
        D.extent = scr;
        if (h->n == 0) D.n = 0;
        k_extent<double, double><<<reduction_wgs, WG, 0, h->stream>>>(D, 0, reduction_wgs, nullptr, max_dt, cfl_dx);
        k_extent<double, double><<<1, WG, 0, h->stream>>>(D, 1, reduction_wgs, out, max_dt, cfl_dx);
    }
    if (hipGetLastError() != hipSuccess) return failStatic(h, ZMPM_E_HIP, "kernel launch failed");
    return ZMPM_OK;
}

zmpm_status zmpm_get_timings(zmpm_handle h, int32_t* count, const char** names, double* ms)
{
    if (!h || !count) return ZMPM_E_ARG;
    const int cap = *count;
    const int k = std::min(cap, kTimedGroups + 1);
    for (int g = 0; g < k; g++) {
        if (names) names[g] = kTimedNames[g];
        if (ms) ms[g] = h->timed_steps ? h->timed_ms[g] / h->timed_steps : 0.0;
    }
    *count = k;
    return ZMPM_OK;
}

zmpm_status zmpm_get_launch_config(zmpm_handle h, zmpm_launch_config* cfg)
{
    if (!h || !cfg) return ZMPM_E_ARG;
    *cfg = zmpm_launch_config {};
    cfg->p2g_wgs = h->p2g_wgs;
    cfg->grid_wgs = h->grid_wgs;
    cfg->prep_wgs = h->prep_wgs;
    cfg->p2g_split = h->p2g_split;
    cfg->variant = h->variant;
    return ZMPM_OK;
}

zmpm_status zmpm_set_launch_config(zmpm_handle h, const zmpm_launch_config* cfg)
{
    if (!h || !cfg) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (cfg->p2g_wgs < 1 || cfg->grid_wgs < 1 || cfg->prep_wgs < 1 || cfg->p2g_split < 1)
        return fail(h, ZMPM_E_ARG, "launch configuration entries must be >= 1");
    if (cfg->variant < 0 || cfg->variant > VARIANT_ALL) return fail(h, ZMPM_E_ARG, "unknown kernel variant bits");
    (void)hipSetDevice(h->desc.device);
    dropGraphs(h);
    h->p2g_wgs = cfg->p2g_wgs;
    h->grid_wgs = cfg->grid_wgs;
    h->prep_wgs = cfg->prep_wgs;
    h->p2g_split = cfg->p2g_split;
    h->variant = cfg->variant;
    h->launch_set = true;
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    return ZMPM_OK;
}

zmpm_status zmpm_autotune(zmpm_handle h, double dt, double t0, int32_t steps, int32_t tune_variants,
                          zmpm_launch_config* cfg)
{
    if (!h || !(dt > 0) || steps < 1) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (h->n == 0) return cfg ? zmpm_get_launch_config(h, cfg) : ZMPM_OK;
    (void)hipSetDevice(h->desc.device);
    // snapshot of the state so tuning leaves the simulation where it was
    const auto state = h->fp32 ? liveState<float, long long>(h) : liveState<double, double>(h);
    std::vector<void*> saved(state.size(), nullptr);
    for (size_t i = 0; i < state.size(); i++) {
        if (hipMalloc(&saved[i], state[i].second) != hipSuccess) {
            (void)hipGetLastError();
            for (void* q : saved)
                if (q) (void)hipFree(q);
            return fail(h, ZMPM_E_NOMEM, "autotune: no memory for the state snapshot");
        }
        HIPCHK(h, hipMemcpyAsync(saved[i], state[i].first, state[i].second, hipMemcpyDeviceToDevice, h->stream));
    }
    const bool timing = h->timing, stepped = h->stepped;
    // substeps swap the live and output particle buffers: keep which is live
    const Dev<float, long long> df = h->Df;
    const Dev<double, double> dd = h->Dd;
    double t = t0;
    // a knob: the bits `mask` of *value, the timed stage it is charged with (0
    // bin/sort, 1 p2g, 2 grid update, 4 g2p) and its candidates; none of them
    // changes a result (the P2G lane split, which does, is not tuned)
    struct Knob {
        int* value;
        int mask;
        int stage;
        std::vector<int> candidates;
    };
    std::vector<Knob> knobs = {
        { &h->p2g_wgs, ~0, 1, { 2048, 4096, 8192, 16384, 32768 } },
        { &h->prep_wgs, ~0, 0, { 512, 2048, 8192 } },
        { &h->grid_wgs, ~0, 2, { 256, 512, 1024, 2048, 4096 } },
    };
    if (tune_variants) {
        knobs.push_back({ &h->variant, VARIANT_P2G_LOW_OCC, 1, { 0, VARIANT_P2G_LOW_OCC } });
        knobs.push_back({ &h->variant, VARIANT_P2G_XCD, 1, { 0, VARIANT_P2G_XCD } });
        knobs.push_back({ &h->variant, VARIANT_G2P_RR, 4, { 0, VARIANT_G2P_RR } });
        knobs.push_back({ &h->variant, VARIANT_GRID_SPLIT, 2, { 0, VARIANT_GRID_SPLIT } });
    }
    h->timing = true;
    zmpm_status st = ZMPM_OK;
    // one untimed substep first (first-touch of the configuration's launches)
    if ((st = runUniform(h, 1, dt, t)) != ZMPM_OK) goto done;
    t += dt;
    for (Knob& k : knobs) {
        // each candidate timed in two rounds, its best round kept (one round
        // can catch a clock or cache transient)
        std::vector<double> best_of(k.candidates.size(), 1e300);
        for (int round = 0; round < 2; round++)
            for (size_t ci = 0; ci < k.candidates.size(); ci++) {
                *k.value = (*k.value & ~k.mask) | (k.candidates[ci] & k.mask);
                if (h->fp32)
                    refreshTables<float, long long>(h);
                else
                    refreshTables<double, double>(h);
                if ((st = runUniform(h, steps, dt, t)) != ZMPM_OK) goto done;
                t += steps * dt;
                best_of[ci] = std::min(best_of[ci], h->timed_ms[k.stage] / h->timed_steps);
            }
        size_t best = 0;
        for (size_t ci = 1; ci < k.candidates.size(); ci++)
            if (best_of[ci] < best_of[best]) best = ci;
        *k.value = (*k.value & ~k.mask) | (k.candidates[best] & k.mask);
    }
done:
    h->timing = timing;
    h->launch_set = true;
    if (st == ZMPM_OK) st = deviceStatus(h, true);
    // restore the state (into the buffers live before tuning); page-scan tile
    // words of the tuning substeps are cleared
    h->Df = df;
    h->Dd = dd;
    for (size_t i = 0; i < state.size(); i++)
        (void)hipMemcpyAsync(state[i].first, saved[i], state[i].second, hipMemcpyDeviceToDevice, h->stream);
    if (h->fp32)
        clearScanState<float, long long>(h);
    else
        clearScanState<double, double>(h);
    (void)hipStreamSynchronize(h->stream);
    for (void* q : saved) (void)hipFree(q);
    h->stepped = stepped;
    if (h->fp32)
        refreshTables<float, long long>(h);
    else
        refreshTables<double, double>(h);
    dropGraphs(h);
    if (st != ZMPM_OK) return st;
    if (cfg) return zmpm_get_launch_config(h, cfg);
    return ZMPM_OK;
}

// Key of the tuning cache: the device and what selects the work and the kernels.
// Raised when the kernels change enough that cached tunings should be redone.
constexpr int kTuneRevision = 4;

static std::string tuneKey(zmpm_handle h, int32_t tune_variants)
{
    hipDeviceProp_t prop {};
    (void)hipGetDeviceProperties(&prop, h->desc.device);
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    bool plastic = false, fj = false;
    for (const auto& m : h->mats) {
        plastic = plastic || m.plasticity != ZMPM_PLASTICITY_NONE;
        fj = fj || m.fj_mixed != ZMPM_FJ_NONE;
    }
    std::ostringstream k;
    k << arch << "/cu" << prop.multiProcessorCount << "/" << (h->fp32 ? "fp32" : "fp64") << "/s" << h->scenes.size()
      << "/n" << std::lround(4.0 * std::log2((double)std::max(h->n, 1))) << "/d" << (h->any_damage ? 1 : 0) << "/i"
      << h->desc.implicit_damage << "/t" << h->desc.transfer_scheme << "/m" << h->desc.mls_mpm << "/r"
      << h->desc.rpic_damping_iteration << "/p" << h->desc.pic_damping_iteration << "/j" << (fj ? 1 : 0) << "/y"
      << (plastic ? 1 : 0) << "/v" << (tune_variants ? 1 : 0) << "/k" << kTuneRevision;
    return k.str();
}

zmpm_status zmpm_autotune_cached(zmpm_handle h, double dt, double t0, int32_t steps, int32_t tune_variants,
                                 const char* cache_path, zmpm_launch_config* cfg, int32_t* from_cache)
{
    if (!h || !cache_path || !(dt > 0) || steps < 1) return ZMPM_E_ARG;
    if (midSubstep(h)) return ZMPM_E_ARG;
    if (from_cache) *from_cache = 0;
    const std::string key = tuneKey(h, tune_variants);
    {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            zmpm_launch_config c {};
            if (!(ls >> k >> c.p2g_wgs >> c.grid_wgs >> c.prep_wgs >> c.variant) || k != key) continue;
            c.p2g_split = h->p2g_split;
            const zmpm_status st = zmpm_set_launch_config(h, &c);
            if (st != ZMPM_OK) continue; // a malformed line: tune instead
            if (from_cache) *from_cache = 1;
            return cfg ? zmpm_get_launch_config(h, cfg) : ZMPM_OK;
        }
    }
    zmpm_launch_config c {};
    const zmpm_status st = zmpm_autotune(h, dt, t0, steps, tune_variants, &c);
    if (st != ZMPM_OK) return st;
    // best effort: a cache that cannot be written only costs the next run a tuning
    std::ofstream out(cache_path, std::ios::app);
    out << key << ' ' << c.p2g_wgs << ' ' << c.grid_wgs << ' ' << c.prep_wgs << ' ' << c.variant << '\n';
    if (cfg) *cfg = c;
    return ZMPM_OK;
}

} // extern "C"
