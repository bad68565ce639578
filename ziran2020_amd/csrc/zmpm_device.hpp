// zmpm_device.hpp -- gfx950 kernels of the explicit
// APIC MPM substep with AnisoMPM damage.
//
// One substep (Ziran MpmSimulationBase::advanceOneTimeStep,
// Lib/MPM/MpmSimulationBase.cpp:262-357) is this launch sequence on the
// handle's stream:
//   k_super_scan     touched super pages in order (single-pass compaction)
//   k_page_scan      dense block ids in page (Morton) order, page table, block
//                    ranges (single-pass scan)
//   k_scatter        particles to their block ranges; zero block accumulators,
//                    forward-neighbour block table, device clock and
//                    fixed-point scales (float form)
//   k_p2g            order inside each block by cell, then original index
//                    (-> reference order), then APIC scatter of m, mv, f
//                    (stress read from the particle) and of the damage phase
//                    field
//   k_grid           v = mv/m, gravity + force, body forces, phase-field
//                    normalisation; the collision objects inside it (small
//                    grids) or in k_collide after it, on the blocks they may
//                    reach
//   k_dtilde, k_cg   (implicit damage only) crack driving force history and the
//                    phase-field CG solve
//   k_g2p            APIC gather, x update, CFL flag, damage Laplacian and
//                    update, F update, stress of the new state and the next
//                    P2G's fixed-point bounds; writes particles in sorted order
//                    and bins them for the next substep (binParticle: key,
//                    page activation, per-page counts)
//   k_rare           the particles k_g2p deferred (full damage update,
//                    near-inversion stress, plasticity return mappings)
#pragma once
#include "zmpm_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

namespace zmpm {

constexpr int WG = 256;
constexpr float FIXED_TARGET_EXP = 40.0f;
constexpr int BOUND_SLOTS = 64; // atomic-max slots per fixed-point scale group
// Per-lane bound on the summed |fixed-point values| one P2G lane adds to its
// tile.  By construction one particle's contributions to a stencil column are
// below 15 x 2^40 summed over the 9 quantities and 3 nodes (each quantity's
// per-particle bound is scaled to < 2^40; the force counts 3 components x 3
// nodes), and a lane adds at most CHUNK = 128 particles, so a correct lane stays
// below 1920 x 2^40 < 2^51; checking 2^51 also keeps every value in the range
// of the exact double-to-int64 conversion used for it.
constexpr float FIXED_LANE_LIMIT = 2251799813685248.0f; // 2^51
constexpr double FIXED_MAGIC = 6755399441055744.0;      // 1.5 x 2^52
constexpr long long FIXED_MAGIC_BITS = 0x4338000000000000LL;

__device__ __forceinline__ void latch(int* counters, int bits) { atomicOr(&counters[1], bits); }

// Workgroup barrier ordering LDS only: the workgroup's global loads and
// atomics stay in flight across it (a full __syncthreads waits for them), so a
// prefetch issued before it keeps hiding its latency.  Used where threads
// exchange data through LDS alone.
__device__ __forceinline__ void ldsBarrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroup ids are dispatched round-robin over the 8 XCDs (each with its own
// L2).  The logical index below gives each XCD one contiguous run of the grid's
// work items (G2P's particle segments), so neighbours that read the same halo
// nodes run on the same XCD and meet in its L2 (G2P FP64 ~189 -> ~164 us, FP32
// ~104 -> ~101 us for 1M particles on MI355X).  A bijection on [0, nwg).
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcdLocal(int wg, int nwg)
{
    const int x = wg % NUM_XCD, q = wg / NUM_XCD;
    const int full = nwg / NUM_XCD, rem = nwg % NUM_XCD;
    return x * full + min(x, rem) + q;
}

// ------------------------------------------------------- particle field access
// Element i of a field whose base (buffer + field * capacity) is uniform: the
// byte offset is formed in 32 bits (capacity < 2^27), so the access takes the
// scalar-base + 32-bit vector-offset addressing form and needs no per-lane
// 64-bit address arithmetic.
// The base goes through readfirstlane (a no-op on a uniform value) so the
// optimiser keeps it a scalar instead of folding the field offset into the
// per-lane address; the access stays in the global address space.
template <class T>
using GlobalPtr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ GlobalPtr<char> scalarBase(const T* p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (GlobalPtr<char>)(((uint64_t)hi << 32) | lo);
}
template <class T>
__device__ __forceinline__ T ldAt(const T* base, int i)
{
    return *(GlobalPtr<const T>)(scalarBase(base) + (uint32_t)i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stAt(T* base, int i, T v)
{
    *(GlobalPtr<T>)(scalarBase(base) + (uint32_t)i * (uint32_t)sizeof(T)) = v;
}

// ---------------------------------------------------------------- positions
__device__ __forceinline__ int loadPos(const float* base, int i) { return __float_as_int(ldAt(base, i)); }
__device__ __forceinline__ double loadPos(const double* base, int i) { return ldAt(base, i); }
__device__ __forceinline__ void storePos(float* base, int i, int X) { stAt(base, i, __int_as_float(X)); }
__device__ __forceinline__ void storePos(double* base, int i, double x) { stAt(base, i, x); }

// Base node of one axis, floor(x / dx - 1/2) (baseNode<2>, BSplines.h:16-20;
// the key build of MpmSimulationBase.cpp:1545), as a node of the handle's grid
// (the scene's region starts at node `origin`), and the offset from it in
// cells, d0 = x / dx - base in [0.5, 1.5).  Returns ST_NAN / ST_DOMAIN when the
// position is not finite or the base node leaves [0, domain - 4) of the scene;
// base is then clamped into that range (and d0 set to 1) so callers index
// inside the scene's region.
// FP64 form: x (metres in the scene's frame) * (1/dx) - 0.5 in double without
// FMA, truncated (int_floor for t >= 0, MathTools.h:15-25).
__device__ __forceinline__ int cellOf(double x, double odx, int origin, int domain, int& base, double& d0, int)
{
    const double xs = mul_rn(x, odx);
    double t = sub_rn(xs, 0.5);
    int bits = 0;
    if (!isfinite(t)) {
        bits = ST_NAN;
        t = 0.0;
    }
    const double hi = double(domain - 4);
    if (!(t >= 0.0 && t < hi)) {
        bits |= ST_DOMAIN;
        t = (t < 0.0) ? 0.0 : hi - 1.0;
    }
    const int b = (int)t;
    d0 = bits ? 1.0 : xs - double(b);
    base = b + origin;
    return bits;
}

// FP32 form: the same floor, exact on the fixed-point position X (scene frame,
// pb fraction bits); d0 has at most pb + 1 significant bits, so it is exact in
// float too.
__device__ __forceinline__ int cellOf(int X, float, int origin, int domain, int& base, float& d0, int pb)
{
    const int bits0 = (X == POS_NAN) ? ST_NAN : 0;
    int bits = bits0;
    const int t = bits0 ? 0 : X - (1 << (pb - 1)); // half a cell below: floor(X / 2^pb - 1/2)
    int b = t >> pb;                               // arithmetic shift: floor
    if (!(b >= 0 && b < domain - 4)) {
        bits |= ST_DOMAIN;
        b = (b < 0) ? 0 : domain - 5;
    }
    base = b + origin;
    d0 = bits ? 1.f : (float)(X - (b << pb)) * __int_as_float((127 - pb) << 23); // times 2^-pb, exact
    return bits;
}

// The scene of a particle (material word) and of a grid node (its region).
__device__ __forceinline__ int sceneOf(int pmat) { return pmat >> PMAT_SCENE_SHIFT; }
__device__ __forceinline__ int matOf(int pmat) { return pmat & PMAT_MAT_MASK; }

template <class T, class AccT>
__device__ __forceinline__ int sceneOfNode(const Dev<T, AccT>& D, uint32_t ix, uint32_t iy, uint32_t iz)
{
    const uint32_t rd = (uint32_t)D.region_dim;
    const uint32_t rx = ix >> D.region_shift, ry = iy >> D.region_shift, rz = iz >> D.region_shift;
    if (rx >= rd || ry >= rd || rz >= rd) return -1;
    return D.region_scene[(rx * rd + ry) * rd + rz];
}

// Every third bit: b10..b0 -> bit 3*i (block coordinates < 2048, the 4096-node
// domain in the double form's 2-node x blocks).
__device__ __forceinline__ uint32_t part1by2(uint32_t v)
{
    uint64_t x = v & 0x7ffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return (uint32_t)x;
}

__device__ __forceinline__ uint32_t compact1by2(uint32_t v)
{
    uint64_t x = v & 0x49249249u;
    x = (x | (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x | (x >> 4)) & 0x100f00f00f00f00full;
    x = (x | (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x | (x >> 16)) & 0x1f00000000ffffull;
    x = (x | (x >> 32)) & 0x7ffull;
    return (uint32_t)x;
}

// SPGrid page index of block coordinates: the page bits of Linear_Offset
// (SPGrid_Mask.h:150-166) interleave the block coordinates, one bit per axis in
// the order of Blk<T>::{x,y,z}pos.  Equal to bitSpread over the full masks
// shifted right by 12 (pinned by the binning parity tests).
template <class T>
__device__ __forceinline__ uint32_t pageOfBlock(const Masks&, uint32_t Bx, uint32_t By, uint32_t Bz)
{
    using B = Blk<T>;
    return (part1by2(Bx) << B::xpos) | (part1by2(By) << B::ypos) | (part1by2(Bz) << B::zpos);
}

template <class T>
__device__ __forceinline__ void blockOfPage(uint32_t page, uint32_t& Bx, uint32_t& By, uint32_t& Bz)
{
    using B = Blk<T>;
    Bx = compact1by2(page >> B::xpos);
    By = compact1by2(page >> B::ypos);
    Bz = compact1by2(page >> B::zpos);
}

// (page, element) of a node: Linear_Offset(i, j, k) >> 12 and the element index
// inside the page (x bits above y bits above z bits).
template <class T>
__device__ __forceinline__ void pageElem(uint32_t i, uint32_t j, uint32_t k, uint32_t& page, uint32_t& elem)
{
    using B = Blk<T>;
    page = pageOfBlock<T>(Masks {}, i >> B::xbits, j >> B::ybits, k >> B::zbits);
    elem = ((i & ((1u << B::xbits) - 1)) << (B::ybits + B::zbits)) | ((j & ((1u << B::ybits) - 1)) << B::zbits)
        | (k & ((1u << B::zbits) - 1));
}

// ---------------------------------------------------------------- wave scans
__device__ __forceinline__ float waveMax(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d));
    return v;
}

__device__ __forceinline__ int waveInclusiveScan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(v, d);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ uint64_t waveInclusiveScan64(uint64_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t o = __shfl_up(v, d);
        if (lane >= d) v += o;
    }
    return v;
}

// ------------------------------------------------------------------ binning
// sortParticlesAndPolluteGrid (MpmSimulationBase.cpp:1531-1573) and
// MpmGrid::pollute (MpmGrid.h:153-193):
//   binParticle  key of every particle, page activation, per-page particle
//                count and each particle's slot in its page -- run by the
//                kernel that wrote the positions (G2P, or the upload)
//   k_page_scan  one single-pass scan over the pages of (activated, count):
//                dense block ids in page (Morton) order, the page table, and
//                each block's range of sorted positions; clears the page flags
//   k_scatter    particle -> its block's range; zeroes the blocks' grid
//                accumulators, builds the neighbour table, starts the substep
//                (device clock, fixed-point scales)
//   blockOrder   (in k_p2g) inside each block: counting sort by cell, then by
//                original index inside the cell, in LDS
// Sorting by (block in page order, cell, original index) is std::sort of the
// reference's unique 64-bit keys (MpmSimulationBase.cpp:1552).

// counters[] slots
enum : int { CNT_BLOCKS = 0, CNT_STATUS = 1, CNT_EPOCH = 2, CNT_TICKET = 3, CNT_RARE = 4, CNT_SUPER = 5, CNT_TICKET2 = 6 };

constexpr int PSCAN_ITEMS = 8;
constexpr int PSCAN_TILE = WG * PSCAN_ITEMS; // pages per scan tile
constexpr int SCAN_SPIN_LIMIT = 1 << 16;

// Page-scan tile status word: [63:55] epoch of the substep (1..511),
// [54:53] 1 = tile aggregate / 2 = inclusive prefix, [52:27] activated pages
// (< 2^26), [26:0] particles (< 2^27).  Every tile publishes every substep, so
// a word whose epoch differs is the previous substep's and is waited on.
__device__ __forceinline__ uint64_t scanWord(unsigned epoch, unsigned flag, uint64_t pair)
{
    const uint64_t pages = pair >> 32, parts = pair & 0xffffffffull;
    return ((uint64_t)epoch << 55) | ((uint64_t)flag << 53) | (pages << 27) | parts;
}

__device__ __forceinline__ uint64_t scanPair(uint64_t w)
{
    return (((w >> 27) & ((1ull << 26) - 1)) << 32) | (w & ((1ull << 27) - 1));
}

// Binning of the particle at live index p for the coming substep: base node
// (MpmSimulationBase.cpp:1545, without FMA), SPGrid key, activation of its page
// and the +1 neighbours (MpmGrid.h:153-193), and its slot in its page's count.
// Live order is the previous substep's sorted order, so equal pages come in
// runs inside a wavefront: one count atomic and one activation per run.  Every
// lane of the wavefront calls it (active or not).
template <class T, class AccT>
__device__ __forceinline__ void binParticle(const Dev<T, AccT>& D, int p, bool active, const PosT<T>* x, int scene)
{
    using B = Blk<T>;
    const int lane = threadIdx.x & 63;
    uint32_t page = 0xffffffffu - (uint32_t)lane; // inactive lanes never join a run
    int bb[3] = { 0, 0, 0 };
    if (active) {
        const SceneDev<T>& sd = D.sc[scene];
        int bits = 0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            T d0;
            bits |= cellOf(x[a], sd.odx, sd.origin[a], sd.domain, bb[a], d0, sd.pos_bits);
        }
        if (bits) latch(D.counters, bits);
        uint32_t elem;
        pageElem<T>((uint32_t)bb[0], (uint32_t)bb[1], (uint32_t)bb[2], page, elem);
        stAt(D.key, p, ((uint64_t)page << (B::xbits + B::ybits + B::zbits)) | elem);
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)page, 1);
    const bool head = (lane == 0) || (prev != page);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = heads & (~0ull >> (63 - lane));
    const int start = 63 - __clzll(upto);
    const uint64_t after = (lane == 63) ? 0ull : (heads & (~0ull << (lane + 1)));
    const int next = after ? __ffsll((unsigned long long)after) - 1 : 64;
    int slot = 0;
    if (active && head) {
        slot = atomicAdd(&D.page_count[page], next - lane);
        const uint32_t Bx = bb[0] >> B::xbits, By = bb[1] >> B::ybits, Bz = bb[2] >> B::zbits;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const uint32_t pg = pageOfBlock<T>(D.masks, Bx + a, By + b, Bz + c);
                    D.touched[pg] = 1;
                    D.super_touched[pg / PSCAN_TILE] = 1;
                }
    }
    slot = __shfl(slot, start);
    if (active) stAt(D.rank, p, slot + (lane - start));
}

// The binning of a substep is done by whatever wrote the particle positions
// (G2P of the previous substep, or the upload): it also opens the page scan of
// the coming substep.
template <class T, class AccT>
__device__ __forceinline__ void openPageScan(const Dev<T, AccT>& D)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int e = D.counters[CNT_EPOCH];
        D.counters[CNT_EPOCH] = e % 511 + 1; // the coming substep's page-scan epoch
        D.counters[CNT_TICKET] = 0;          // page-scan tile tickets
        D.counters[CNT_TICKET2] = 0;         // super-page-scan tile tickets
    }
}

// Decoupled look-back of a single-pass scan, run by wavefront 0 of the
// workgroup holding `tile`: publishes the tile's aggregate, adds its
// predecessors' published prefixes (64 per poll; lane l reads tile - 1 - l, the
// prefix is the sum up to the nearest inclusive word), publishes the inclusive
// prefix and returns the exclusive one.  A tile waits only on tiles that took
// earlier tickets, which are running; the wait is also bounded (SCAN_SPIN_LIMIT
// polls, then ST_SCAN is latched).
template <class T, class AccT>
__device__ __forceinline__ uint64_t lookBack(const Dev<T, AccT>& D, uint64_t* state, int tile, uint64_t total,
                                             unsigned epoch)
{
    const int lane = threadIdx.x & 63;
    uint64_t pre = 0;
    if (tile == 0) {
        if (lane == 0) __hip_atomic_store(&state[0], scanWord(epoch, 2, total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if (lane == 0) __hip_atomic_store(&state[tile], scanWord(epoch, 1, total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int base_j = tile - 1, spins = 0;
    while (true) {
        const int j = base_j - lane;
        uint64_t w = scanWord(epoch, 2, 0); // before tile 0: an inclusive zero
        if (j >= 0) w = __hip_atomic_load(&state[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned we = (unsigned)(w >> 55), wf = (unsigned)(w >> 53) & 3u;
        const bool valid = we == epoch && wf != 0;
        const uint64_t incl = __ballot(valid && wf == 2);
        const int first = incl ? __ffsll((unsigned long long)incl) - 1 : 63;
        const uint64_t need = (first == 63) ? ~0ull : ((2ull << first) - 1ull);
        if (__ballot(!valid) & need) { // a predecessor up to `first` is not published yet
            if (++spins > SCAN_SPIN_LIMIT) {
                if (lane == 0) latch(D.counters, ST_SCAN);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t v = lane <= first ? scanPair(w) : 0ull;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
        pre += v;
        if (incl) break;
        base_j -= 64;
    }
    if (lane == 0) __hip_atomic_store(&state[tile], scanWord(epoch, 2, pre + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return pre;
}

// Exclusive workgroup scan of one value per thread; *total gets the sum.
__device__ __forceinline__ uint64_t blockExclusiveScan(uint64_t mine, uint64_t* wsum, uint64_t* total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t inc = waveInclusiveScan64(mine);
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    uint64_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < WG / 64; w++) {
        if (w < wv) off += wsum[w];
        tot += wsum[w];
    }
    *total = tot;
    return off + inc - mine;
}

// Compaction of the touched super pages in index order (one tile of
// PSCAN_TILE super pages per workgroup, tiles in ticket order): super_list
// and their count (CNT_SUPER); the flags are cleared.  Its last tile also
// zeroes the block count, which the page scan sets when any page is active.
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_super_scan(Dev<T, AccT> D)
{
    __shared__ int s_tile;
    __shared__ uint64_t s_wsum[WG / 64];
    __shared__ uint64_t s_prefix;
    const int t = threadIdx.x, wv = t >> 6;
    const int64_t ns = D.num_super;
    const int ntiles = (int)((ns + PSCAN_TILE - 1) / PSCAN_TILE);
    const unsigned epoch = (unsigned)D.counters[CNT_EPOCH];
    if (t == 0) s_tile = atomicAdd(&D.counters[CNT_TICKET2], 1);
    __syncthreads();
    const int tile = s_tile;
    if (tile >= ntiles) return;
    const int64_t base = (int64_t)tile * PSCAN_TILE + (int64_t)t * PSCAN_ITEMS;
    uint32_t fl[PSCAN_ITEMS];
    uint64_t mine = 0;
#pragma unroll
    for (int k = 0; k < PSCAN_ITEMS; k++) {
        const int64_t sp = base + k;
        fl[k] = sp < ns && D.super_touched[sp] != 0;
        mine += fl[k];
    }
    uint64_t total = 0;
    uint64_t ex = blockExclusiveScan(mine, s_wsum, &total);
    if (wv == 0) {
        const uint64_t pre = lookBack(D, D.super_state, tile, total, epoch);
        if ((t & 63) == 0) {
            s_prefix = pre;
            if (tile == ntiles - 1) {
                D.counters[CNT_SUPER] = (int)(pre + total);
                D.counters[CNT_BLOCKS] = 0;
            }
        }
    }
    __syncthreads();
    ex += s_prefix;
#pragma unroll
    for (int k = 0; k < PSCAN_ITEMS; k++) {
        const int64_t sp = base + k;
        if (fl[k]) {
            D.super_list[ex] = (int)sp;
            D.super_touched[sp] = 0;
        }
        ex += fl[k];
    }
}

// Single-pass scan (tiles taken in ticket order, each adding its predecessors'
// published prefixes) of the pair (activated, particle count) over the pages of
// the touched super pages, one super page per tile, in page order.  Persistent:
// each workgroup takes tickets until the super pages run out.  Epilogue per
// page: dense block id or -1 in the page table, and for each block its page,
// first sorted position and particle count.  The flags and counts are cleared
// for the next substep.  Pages of untouched super pages keep stale table
// entries, which nothing reads: every page a particle or a block's neighbour
// lookup names was touched.
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_page_scan(Dev<T, AccT> D)
{
    __shared__ int s_tile;
    __shared__ uint64_t s_wsum[WG / 64];
    __shared__ uint64_t s_prefix;
    const int t = threadIdx.x, wv = t >> 6;
    const int64_t np = D.num_pages;
    const int ntiles = D.counters[CNT_SUPER];
    const unsigned epoch = (unsigned)D.counters[CNT_EPOCH];
    for (;;) {
        if (t == 0) s_tile = atomicAdd(&D.counters[CNT_TICKET], 1);
        __syncthreads();
        const int tile = s_tile;
        __syncthreads(); // s_tile read by every thread before the next ticket
        if (tile >= ntiles) return;
        const int64_t base = (int64_t)D.super_list[tile] * PSCAN_TILE + (int64_t)t * PSCAN_ITEMS;
        uint32_t fl[PSCAN_ITEMS], ct[PSCAN_ITEMS];
        uint64_t mine = 0;
#pragma unroll
        for (int k = 0; k < PSCAN_ITEMS; k++) {
            const int64_t pg = base + k;
            fl[k] = 0;
            ct[k] = 0;
            if (pg < np) {
                fl[k] = D.touched[pg] != 0;
                ct[k] = (uint32_t)D.page_count[pg];
            }
            mine += ((uint64_t)fl[k] << 32) | ct[k];
        }
        uint64_t total = 0;
        uint64_t ex = blockExclusiveScan(mine, s_wsum, &total);
        if (wv == 0) {
            const uint64_t pre = lookBack(D, D.scan_state, tile, total, epoch);
            if ((t & 63) == 0) {
                s_prefix = pre;
                if (tile == ntiles - 1) {
                    const int64_t blocks = (int64_t)((pre + total) >> 32);
                    if (blocks > D.cap_blocks) latch(D.counters, ST_CAPACITY);
                    D.counters[CNT_BLOCKS] = (int)min<int64_t>(blocks, D.cap_blocks);
                }
            }
        }
        __syncthreads();
        ex += s_prefix;
#pragma unroll
        for (int k = 0; k < PSCAN_ITEMS; k++) {
            const int64_t pg = base + k;
            if (pg < np) {
                int id = -1;
                if (fl[k]) {
                    const int bid = (int)(ex >> 32);
                    D.touched[pg] = 0;
                    if (bid < D.cap_blocks) {
                        id = bid;
                        D.block_page[bid] = (int)pg;
                        D.block_start[bid] = (int)(ex & 0xffffffffull);
                        D.block_count[bid] = (int)ct[k];
                    }
                }
                if (ct[k]) D.page_count[pg] = 0;
                D.page_table[pg] = id;
            }
            ex += ((uint64_t)fl[k] << 32) | ct[k];
        }
        __syncthreads(); // s_wsum and s_prefix free for the next tile
    }
}

// Eigen's Quaternion::toRotationMatrix of q = (w, x, y, z), row-major.
__host__ __device__ inline void quatToMatrix(const double* q, double* R)
{
    double w = q[0];
    const double x = q[1], y = q[2], z = q[3];
    if (w == 0 && x == 0 && y == 0 && z == 0) w = 1;
    const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
    const double twx = tx * w, twy = ty * w, twz = tz * w;
    const double txx = tx * x, txy = ty * x, txz = tz * x;
    const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
    R[0] = 1 - (tyy + tzz);
    R[1] = txy - twz;
    R[2] = txz + twy;
    R[3] = txy + twz;
    R[4] = 1 - (txx + tzz);
    R[5] = tyz - twx;
    R[6] = txz - twy;
    R[7] = tyz + twx;
    R[8] = 1 - (txx + tyy);
}

// The object's state at time t: b = b0 + dbdt t, s = s0 + dsdt t and the
// rotation of q(t) = (cos(|w| t/2), sin(|w| t/2) w/|w|) * q0 (constant angular
// velocity w, as the updateState callbacks of examples/boneTwist.h:144-170 set it).
__device__ inline void colliderState(const ColDev& c, double t, ColState& st)
{
#pragma unroll
    for (int a = 0; a < 3; a++) st.b[a] = c.b0[a] + c.dbdt[a] * t;
    st.s = c.s0 + c.dsdt * t;
    st.os = 1.0 / st.s;
    double q[4] = { c.q0[0], c.q0[1], c.q0[2], c.q0[3] };
    const double wn = sqrt(c.omega[0] * c.omega[0] + c.omega[1] * c.omega[1] + c.omega[2] * c.omega[2]);
    if (wn > 0) {
        const double ha = 0.5 * wn * t, cw = cos(ha), sw = sin(ha) / wn;
        const double p[4] = { cw, sw * c.omega[0], sw * c.omega[1], sw * c.omega[2] };
        const double* q0 = c.q0;
        q[0] = p[0] * q0[0] - p[1] * q0[1] - p[2] * q0[2] - p[3] * q0[3];
        q[1] = p[0] * q0[1] + p[1] * q0[0] + p[2] * q0[3] - p[3] * q0[2];
        q[2] = p[0] * q0[2] - p[1] * q0[3] + p[2] * q0[0] + p[3] * q0[1];
        q[3] = p[0] * q0[3] + p[1] * q0[2] - p[2] * q0[1] + p[3] * q0[0];
    }
    quatToMatrix(q, st.R);
}

// Start of a substep: each scene's clock and (float form) the fixed-point
// scales of this substep's P2G per scene from the bounds the previous G2P (or
// the upload) wrote, one slot per workgroup: 2^(40 - e) with max = f 2^e,
// f in [0.5, 1).
__device__ __forceinline__ float boundScale(float m, float fallback)
{
    if (!(m > 0.f && isfinite(m))) return fallback;
    int e;
    frexpf(m, &e);
    int se = (int)FIXED_TARGET_EXP - e;
    se = se < -120 ? -120 : (se > 120 ? 120 : se);
    return ldexpf(1.0f, se);
}

template <class T, class AccT>
__device__ __forceinline__ void beginStep(const Dev<T, AccT>& D)
{
    // clock[2 s] = time of the scene's next substep, clock[2 s + 1] = this
    // substep's t; the rare-path queue of this substep's G2P starts empty
    if (threadIdx.x == 0) D.counters[CNT_RARE] = 0;
    for (int sc = threadIdx.x; sc < D.nscenes; sc += WG) {
        D.clock[2 * sc + 1] = D.clock[2 * sc];
        D.clock[2 * sc] = D.clock[2 * sc] + D.sc[sc].ddt;
    }
    __syncthreads();
    // the collision objects at their scene's time t_n (the grid update sees them
    // before they move, MpmSimulationBase.cpp:291-293)
    for (int c = threadIdx.x; c < D.num_cols; c += WG) {
        const ColDev co = D.cols[c];
        colliderState(co, D.clock[2 * co.scene + 1], D.colstate[c]);
    }
    // the external body forces at the scene's time t_n
    // (addExternalBodyForce, MpmInitializationHelper.h:1436-1444)
    for (int e = threadIdx.x; e < D.nscenes * MAX_BODY_FORCES; e += WG) {
        const int sc = e / MAX_BODY_FORCES, k = e - sc * MAX_BODY_FORCES;
        const SceneDev<T>& sd = D.sc[sc];
        if (k >= sd.nbf) continue;
        const double amp = sd.bf_scale[k] * (sin(sd.bf_freq[k] * D.clock[2 * sc + 1]) + 1.0);
#pragma unroll
        for (int a = 0; a < 3; a++) D.bfval[3 * e + a] = amp * sd.bf_dir[k][a];
    }
    if constexpr (!std::is_same<AccT, long long>::value) return;
    // one wavefront per (scene, scale group) over the BOUND_SLOTS slots, which
    // are cleared for this substep's G2P
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    for (int k = g; k < D.nscenes * NSC; k += WG / 64) {
        const float v = __uint_as_float(D.maxbits[k * BOUND_SLOTS + lane]);
        D.maxbits[k * BOUND_SLOTS + lane] = 0u;
        const float m = waveMax(v);
        if (lane == 0 && m > 0.f && isfinite(m)) D.scales[k] = boundScale(m, 1.f);
    }
}

// Workgroups [0, prep_wgs): zero the accumulators of every block's nodes and
// build the forward-neighbour block table (first, so their block loops overlap
// the particle workgroups).  Workgroup prep_wgs starts the substep.  The rest,
// part_wgs of them: particle p -> block_start[block] + its slot, with its
// in-block sort key (cell, original index) beside it (the order inside a block
// is fixed by P2G's blockOrder).
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_scatter(Dev<T, AccT> D, int part_wgs, int prep_wgs)
{
    using B = Blk<T>;
    constexpr int bbits = B::xbits + B::ybits + B::zbits;
    constexpr int NB = 1 << bbits;
    const int t = threadIdx.x;
    if ((int)blockIdx.x > prep_wgs) {
        const int p = ((int)blockIdx.x - prep_wgs - 1) * WG + t;
        if (p >= D.n) return;
        const uint64_t key = D.key[p];
        const int bid = D.page_table[key >> bbits];
        if (bid < 0) return; // capacity exceeded (latched by the scan)
        const int pos = D.block_start[bid] + D.rank[p];
        if ((unsigned)pos < (unsigned)D.n) {
            D.tmp[pos] = p;
            D.tmpk[pos] = ((uint32_t)(key & (NB - 1)) << (32 - bbits)) | (uint32_t)D.pid[p];
        }
        return;
    }
    if ((int)blockIdx.x == prep_wgs) {
        beginStep(D);
        return;
    }
    const int nb = D.counters[CNT_BLOCKS];
    const int nf = D.any_damage ? NGF : GPF;
    for (int b = (int)blockIdx.x; b < nb; b += prep_wgs) {
        for (int e = t; e < nf * NB; e += WG) {
            const int f = e / NB, nd = e % NB;
            D.acc[f * D.cap_nodes + (int64_t)b * NB + nd] = AccT(0);
        }
        if (t < 8) {
            uint32_t Bx, By, Bz;
            blockOfPage<T>((uint32_t)D.block_page[b], Bx, By, Bz);
            const uint32_t nx = Bx + ((t >> 2) & 1), ny = By + ((t >> 1) & 1), nz = Bz + (t & 1);
            int id = -1;
            if ((nx << B::xbits) < (uint32_t)D.domain && (ny << B::ybits) < (uint32_t)D.domain
                && (nz << B::zbits) < (uint32_t)D.domain)
                id = D.page_table[pageOfBlock<T>(D.masks, nx, ny, nz)];
            D.nbr[b * 8 + t] = id;
        }
    }
}

// Order inside block b (run by every thread of the P2G workgroup that owns the
// block).  The block's particles are staged in LDS as key = cell << (32 -
// block_bits) | original index (the low 32 bits of the reference key, unique)
// with their live index (ska / pa); a histogram and cursors group them by cell
// (skb / pb); then every particle's position is its cell's start plus the
// number of keys in its cell below its own (all particles ranked in parallel).
// Blocks larger than the LDS stage take the same steps with the grouped arrays
// in HBM.  Writes perm, spid, sorted_cell and every cell's start and count, and
// leaves the cell counts and starts (relative to the block) in hist / cst.
template <class T, class AccT>
__device__ __forceinline__ void blockOrder(const Dev<T, AccT>& D, int b, int start, int cnt, uint32_t* ska, int* pa,
                                           uint32_t* skb, int* pb, int stage, int* hist, int* cur, int* cst)
{
    using B = Blk<T>;
    constexpr int bbits = B::xbits + B::ybits + B::zbits;
    constexpr int NB = 1 << bbits;
    constexpr int idbits = 32 - bbits;
    constexpr uint32_t idmask = (1u << idbits) - 1u;
    const int t = threadIdx.x;
    const bool staged = cnt <= stage;
    if (t < NB) hist[t] = 0;
    ldsBarrier();
    // stage (key, live index) and count cells; four rounds of loads in flight
    for (int i0 = t; i0 < cnt; i0 += 4 * WG) {
        uint32_t kk[4] = { 0u, 0u, 0u, 0u };
        int pp[4] = { 0, 0, 0, 0 };
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * WG;
            if (i < cnt) {
                kk[u] = D.tmpk[start + i];
                if (staged) pp[u] = D.tmp[start + i];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * WG;
            if (i < cnt) {
                atomicAdd(&hist[kk[u] >> idbits], 1);
                if (staged) {
                    ska[i] = kk[u];
                    pa[i] = pp[u];
                }
            }
        }
    }
    ldsBarrier();
    if (t < 64) {
        const int v = t < NB ? hist[t] : 0;
        const int ex = waveInclusiveScan(v) - v;
        if (t < NB) {
            cst[t] = ex;
            cur[t] = ex;
        }
    }
    ldsBarrier();
    if (staged) {
        // group by cell
        for (int i = t; i < cnt; i += WG) {
            const uint32_t k = ska[i];
            const int dst = atomicAdd(&cur[k >> idbits], 1);
            skb[dst] = k;
            pb[dst] = pa[i];
        }
        ldsBarrier();
        // rank inside the cell (keys are unique); write the sorted block.  The
        // cell's keys are read four at a time so the LDS reads overlap.
        for (int i = t; i < cnt; i += WG) {
            const uint32_t k = skb[i];
            const int c = (int)(k >> idbits);
            const int c0 = cst[c], c1 = c0 + hist[c];
            int pos = c0;
            int j = c0;
            for (; j + 4 <= c1; j += 4) {
                const uint32_t k0 = skb[j], k1 = skb[j + 1], k2 = skb[j + 2], k3 = skb[j + 3];
                pos += (int)(k0 < k) + (int)(k1 < k) + (int)(k2 < k) + (int)(k3 < k);
            }
            for (; j < c1; j++) pos += skb[j] < k;
            D.perm[start + pos] = pb[i];
            D.spid[start + pos] = (int)(k & idmask);
            D.sorted_cell[start + pos] = b * NB + c;
        }
    }
    else {
        // the same in HBM: group into perm/spid, then rank through tmp
        for (int i = t; i < cnt; i += WG) {
            const uint32_t k = D.tmpk[start + i];
            const int dst = atomicAdd(&cur[k >> idbits], 1);
            D.perm[start + dst] = D.tmp[start + i];
            D.spid[start + dst] = (int)(k & idmask);
        }
        __syncthreads();
        __threadfence_block();
        for (int i = t; i < cnt; i += WG) D.tmp[start + i] = D.perm[start + i];
        __syncthreads();
        for (int i = t; i < cnt; i += WG) D.sorted_cell[start + i] = D.spid[start + i];
        __syncthreads();
        for (int i = t; i < cnt; i += WG) {
            const int p = D.tmp[start + i];
            const int id = D.sorted_cell[start + i];
            int c = NB - 1;
            while (c > 0 && cst[c] > i) c--;
            const int c0 = cst[c], c1 = c0 + hist[c];
            int pos = c0;
            for (int j = c0; j < c1; j++) pos += D.sorted_cell[start + j] < id;
            D.perm[start + pos] = p;
            D.spid[start + pos] = id;
        }
        __syncthreads();
        for (int i = t; i < cnt; i += WG) {
            int c = NB - 1;
            while (c > 0 && cst[c] > i) c--;
            D.sorted_cell[start + i] = b * NB + c;
        }
    }
    if (t < NB) {
        D.cell_start[b * NB + t] = start + cst[t];
        D.cell_count[b * NB + t] = hist[t];
    }
    // LDS free for the next block; the sorted block (perm, cell starts) is read
    // back by this workgroup after the full barrier that ends its ordering loop
    ldsBarrier();
}

// ------------------------------------------------------------- P2G helpers
template <class T>
__device__ __forceinline__ Mat3<T> loadM(const T* P, int64_t cap, int f, int p)
{
    Mat3<T> M;
#pragma unroll
    for (int i = 0; i < 9; i++) M.a[i / 3][i % 3] = ldAt(P + (f + i) * cap, p);
    return M;
}

// V0 * Kirchhoff stress of one particle, symmetric part (xx xy xz yy yz zz):
// FBasedMpmForceHelper::updateState (FBasedMpmForceHelper.cpp:41-63) with
// kirchhoff = P F^T (HyperelasticConstitutiveModel.h:75-80), or
// JBasedMpmForceHelper::updateState (JBasedMpmForceHelper.cpp:28-50).  tau is
// symmetric for every model on the path; its rounding-level skew is dropped.
// Fs is the stored deformation: F (FP64 form) or H = F - I (FP32 form, whose
// EOS particles hold J - 1 in H_00; the EOS pressure is then
// -k expm1(-gamma log1p(J - 1)) J).
// FAST (G2P): returns true, with no stress written, for a corotated particle
// near inversion (see corotatedTauH).
template <bool FAST = false>
__device__ __forceinline__ bool vtauSym(const MatDev<float>& mt, const Mat3<float>& H, float V0, float g, float* S)
{
    float t6[6];
    if (mt.model == ZMPM_MODEL_EQUATION_OF_STATE) {
        const float jm1 = H.a[0][0];
        const float p = V0 * (-mt.bulk * expm1f(-mt.gamma * log1pf(jm1))) * (1.f + jm1);
        S[0] = S[3] = S[5] = p;
        S[1] = S[2] = S[4] = 0.f;
        return false;
    }
    if (mt.model == ZMPM_MODEL_FIXED_COROTATED) {
        if (corotatedTauH<FAST>(H, mt.mu, mt.lambda, t6)) return true;
    }
    else if (mt.model == ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN) {
        float J;
        qrStableTauH(H, mt.mu, mt.lambda, g, mt.kx, mt.ky, t6, &J);
    }
    else if (mt.model == ZMPM_MODEL_QR_ANISOTROPIC) {
        float J;
        qrAnisoTauH(H, mt.lambda, mt.mu, g, mt.kf[0], mt.kf[1], mt.kf[2], mt.iso != 0, t6, &J);
    }
    else if (mt.model == ZMPM_MODEL_NEO_HOOKEAN_BORDEN) {
        float J;
        bordenTauH(H, mt.mu, mt.kappa, g, t6, &J);
    }
    else
        neoHookeanTauH(H, mt.mu, mt.lambda, t6);
#pragma unroll
    for (int i = 0; i < 6; i++) S[i] = V0 * t6[i];
    return false;
}

template <bool FAST = false>
__device__ __forceinline__ bool vtauSym(const MatDev<double>& mt, const Mat3<double>& F, double V0, double g, double* S)
{
    using T = double;
    T s0, s1, s2, s3, s4, s5;
    if (mt.model == ZMPM_MODEL_EQUATION_OF_STATE) {
        const T J = F.a[0][0];
        const T p = V0 * (-mt.bulk * (dpow(J, -mt.gamma) - T(1))) * J;
        s0 = s3 = s5 = p;
        s1 = s2 = s4 = T(0);
    }
    else {
        Mat3<T> tau;
        if (mt.model == ZMPM_MODEL_FIXED_COROTATED)
            tau = corotatedTau(F, mt.mu, mt.lambda);
        else if (mt.model != ZMPM_MODEL_NEO_HOOKEAN) {
            T J, t6[6];
            if (mt.model == ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN)
                qrStableTau(F, mt.mu, mt.lambda, g, mt.kx, mt.ky, t6, &J);
            else if (mt.model == ZMPM_MODEL_QR_ANISOTROPIC)
                qrAnisoTauH(minusIdentity(F), mt.lambda, mt.mu, g, mt.kf[0], mt.kf[1], mt.kf[2], mt.iso != 0, t6, &J);
            else
                bordenTauH(minusIdentity(F), mt.mu, mt.kappa, g, t6, &J);
            S[0] = V0 * t6[0];
            S[1] = V0 * t6[1];
            S[2] = V0 * t6[2];
            S[3] = V0 * t6[3];
            S[4] = V0 * t6[4];
            S[5] = V0 * t6[5];
            return false;
        }
        else
            tau = neoHookeanTau(F, mt.mu, mt.lambda);
        s0 = V0 * tau.a[0][0];
        s1 = V0 * (T(0.5) * (tau.a[0][1] + tau.a[1][0]));
        s2 = V0 * (T(0.5) * (tau.a[0][2] + tau.a[2][0]));
        s3 = V0 * tau.a[1][1];
        s4 = V0 * (T(0.5) * (tau.a[1][2] + tau.a[2][1]));
        s5 = V0 * tau.a[2][2];
    }
    S[0] = s0;
    S[1] = s1;
    S[2] = s2;
    S[3] = s3;
    S[4] = s4;
    S[5] = s5;
    return false;
}

// Upper bounds of one particle's P2G contributions per scale group, folded
// into the running per-thread maxima b: mass w m; momentum w (m v + m C (x_i - x_p))
// with |x_i - x_p| <= 1.5 dx; force tau grad(w) with |grad w| <= 1/dx per axis;
// phase field w d and w (w <= 1).
template <class T>
__device__ __forceinline__ void momentumBounds(float* b, T m, const T* v, const T* Cm, T dx)
{
    b[SC_M] = fmaxf(b[SC_M], fabsf((float)m));
    float mv = 0.f;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        float cr = fabsf((float)Cm[3 * r]) + fabsf((float)Cm[3 * r + 1]) + fabsf((float)Cm[3 * r + 2]);
        mv = fmaxf(mv, fabsf((float)m) * (fabsf((float)v[r]) + 1.5f * (float)dx * cr));
    }
    b[SC_MV] = fmaxf(b[SC_MV], mv * 1.0001f);
}

template <class T>
__device__ __forceinline__ void forceBounds(float* b, const T* S, T d, T odx, bool dmg)
{
    const float a0 = fabsf((float)S[0]), a1 = fabsf((float)S[1]), a2 = fabsf((float)S[2]);
    const float a3 = fabsf((float)S[3]), a4 = fabsf((float)S[4]), a5 = fabsf((float)S[5]);
    const float f = (float)odx * fmaxf(a0 + a1 + a2, fmaxf(a1 + a3 + a4, a2 + a4 + a5));
    b[SC_F] = fmaxf(b[SC_F], f * 1.0001f);
    if (dmg) {
        b[SC_PF] = fmaxf(b[SC_PF], fabsf((float)d));
        b[SC_PW] = 1.f;
    }
}

// FJMixedMpmForceHelper::updateState under SOLVE_J (FJMixedMpmForceHelper.h:
// 106-121): V0 dpsi/dJ J I with QuadraticVolumePenalty's dpsi/dJ = lambda_J (J - 1)
// (QuadraticVolumePenalty.h).  lambda_J is the J model's own until the first
// evolveStrain, which sets it to the F model's lambda (:186) -- except for fluid
// particles, whose evolveStrain returns before that copy (FJ-mixed materials
// have no snow hardening, so the F model's lambda is the table's).  jm: J (FP64
// form) or J - 1 (FP32 form).
template <class T>
__device__ __forceinline__ void fjSolveJStress(const MatDev<T>& mt, bool first, T V0, T jm, T* S)
{
    const bool own = first || mt.fj == ZMPM_FJ_FLUID;
    const T lam = own ? mt.jl : mt.lambda;
    T jm1, J;
    if constexpr (sizeof(T) == 4) {
        jm1 = jm;
        J = T(1) + jm;
    }
    else {
        jm1 = jm - T(1);
        J = jm;
    }
    const T p = V0 * (lam * jm1) * J;
    S[0] = S[3] = S[5] = p;
    S[1] = S[2] = S[4] = T(0);
}

template <class T>
__device__ __forceinline__ void particleBounds(float* b, T m, const T* v, const T* Cm, const T* S, T d, T dx, T odx,
                                               bool dmg)
{
    momentumBounds(b, m, v, Cm, dx);
    forceBounds(b, S, d, odx, dmg);
}

// The per-thread bounds folded into the bound slots of each thread's scene:
// per wavefront and scene (one pass per distinct scene in the wavefront), an
// atomic max on the float bits (non-negative floats order like their bit
// patterns) into slot blockIdx % BOUND_SLOTS of each scale group.  Threads with
// no particle pass scene < 0.
__device__ __forceinline__ void writeBounds(unsigned int* maxbits, const float* b, int scene)
{
    __shared__ float wmax[NSC][WG / 64];
    __shared__ int wscene[WG / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int slot = (int)(blockIdx.x % BOUND_SLOTS);
    const uint64_t act = __ballot(scene >= 0);
    const int s0 = act ? __shfl(scene, __ffsll((unsigned long long)act) - 1) : -1;
    if (!__ballot(scene >= 0 && scene != s0)) {
        // the wavefront's particles are in one scene (or none): its maxima go to
        // the workgroup's fold below
#pragma unroll
        for (int g = 0; g < NSC; g++) {
            const float mx = waveMax(b[g]);
            if (lane == 0) wmax[g][wv] = mx;
        }
        if (lane == 0) wscene[wv] = s0;
    }
    else {
        // a wavefront across scenes' blocks: one pass per scene, its own atomics
        uint64_t left = act;
        while (left) {
            const int sc = __shfl(scene, __ffsll((unsigned long long)left) - 1);
            const bool mine = scene == sc;
            left &= ~__ballot(mine);
#pragma unroll
            for (int g = 0; g < NSC; g++) {
                const float mx = waveMax(mine ? b[g] : 0.f);
                if (lane == 0 && (mx > 0.f || !(mx == mx)))
                    atomicMax(&maxbits[(sc * NSC + g) * BOUND_SLOTS + slot], __float_as_uint(mx));
            }
        }
        if (lane == 0) wscene[wv] = -1;
    }
    __syncthreads();
    // one atomic per group for each run of wavefronts of one scene (one run in practice)
    if (threadIdx.x < NSC) {
        const int g = threadIdx.x;
        int sc = -1;
        float mx = 0.f;
#pragma unroll
        for (int w = 0; w < WG / 64; w++) {
            const int ws = wscene[w];
            if (ws < 0) continue;
            if (sc >= 0 && ws != sc) {
                if (mx > 0.f || !(mx == mx)) atomicMax(&maxbits[(sc * NSC + g) * BOUND_SLOTS + slot], __float_as_uint(mx));
                mx = 0.f;
            }
            sc = ws;
            mx = fmaxf(mx, wmax[g][w]);
        }
        if (sc >= 0 && (mx > 0.f || !(mx == mx))) atomicMax(&maxbits[(sc * NSC + g) * BOUND_SLOTS + slot], __float_as_uint(mx));
    }
    __syncthreads(); // wmax / wscene reusable by the next call
}

// The factor on a particle's stress from its material's plasticity: snow's
// hardened moduli exp(psi (1 - Jp)) (the supported models' stress is linear in
// mu and lambda); 1 otherwise.
template <class T>
__device__ __forceinline__ T snowScale(const MatDev<T>& mt, T jp)
{
    if (mt.plasticity != ZMPM_PLASTICITY_SNOW) return T(1);
    if constexpr (sizeof(T) == 4)
        return expf(mt.hardening * (1.f - jp));
    else
        return exp(mt.hardening * (1.0 - jp));
}

// Stress of the uploaded state and the P2G bounds of the first substep; with
// bin, also the binning of the first substep (grid of max_slots workgroups,
// striding over particles).  Run again without bin after a material change.
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_stress_init(Dev<T, AccT> D, int bin, int first)
{
    const int64_t cap = D.cap;
    if (bin) openPageScan(D);
    for (int base = blockIdx.x * WG; base < D.n; base += gridDim.x * WG) {
        float b[NSC] = { 0.f, 0.f, 0.f, 0.f, 0.f };
        const int p = base + threadIdx.x;
        const bool active = p < D.n;
        PosT<T> x[3] = { 0, 0, 0 };
        const int pm = active ? D.pmat[p] : 0, scene = sceneOf(pm);
        if (active) {
            const SceneDev<T>& sd = D.sc[scene];
            const MatDev<T>& mt = D.mats[matOf(pm)];
            Mat3<T> F = loadM(D.P, cap, PF, p);
            T S[6], v[3], Cm[9];
            const T jp = mt.plasticity == ZMPM_PLASTICITY_SNOW ? ldAt(D.P + PYS * cap, p) : T(1);
            if (D.fj_solve_j && mt.fj)
                fjSolveJStress(mt, first != 0, ldAt(D.P + PVOL * cap, p), ldAt(D.P + PJ * cap, p), S);
            else
                vtauSym(mt, F, ldAt(D.P + PVOL * cap, p) * snowScale(mt, jp), ldAt(D.P + PG * cap, p), S);
#pragma unroll
            for (int i = 0; i < 6; i++) stAt(D.P + (PS + i) * cap, p, S[i]);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                v[a] = ldAt(D.P + (PV + a) * cap, p);
                x[a] = loadPos(D.P + (PX + a) * cap, p);
            }
#pragma unroll
            for (int i = 0; i < 9; i++) Cm[i] = ldAt(D.P + (PC + i) * cap, p);
            const bool dmg = D.any_damage && mt.has_damage;
            particleBounds(b, ldAt(D.P + PM * cap, p), v, Cm, S, dmg ? ldAt(D.P + PD * cap, p) : T(0), sd.dx, sd.odx,
                           D.any_damage != 0);
        }
        if (bin) binParticle(D, p, active, x, scene);
        if constexpr (std::is_same<AccT, long long>::value) writeBounds(D.maxbits, b, active ? scene : -1);
    }
}

// Scene of dense block b (its region); a handle of one scene needs no lookup.
template <class T, class AccT>
__device__ __forceinline__ int blockScene(const Dev<T, AccT>& D, int b)
{
    if (D.nscenes == 1) return 0;
    using B = Blk<T>;
    uint32_t Bx, By, Bz;
    blockOfPage<T>((uint32_t)D.block_page[b], Bx, By, Bz);
    return max(0, sceneOfNode(D, Bx << B::xbits, By << B::ybits, Bz << B::zbits));
}

// Block-local LDS tile geometry: a block of BX x BY x BZ nodes plus the +2
// halo its particles' quadratic stencils reach.
template <class T>
struct Tile {
    using B = Blk<T>;
    static constexpr int BX = 1 << B::xbits, BY = 1 << B::ybits, BZ = 1 << B::zbits;
    static constexpr int NB = BX * BY * BZ;
    static constexpr int TX = BX + 2, TY = BY + 2, TZ = BZ + 2, TN = TX * TY * TZ;
    static constexpr int CHUNK = 128; // sorted particles per P2G payload stage
};

template <class T>
struct alignas(4 * sizeof(T)) V4 {
    T x, y, z, w;
};

// The four-field record of grid node `node` in a [nodes][4] array (the grid's
// new and old velocities: one aligned 16- or 32-byte access per node).
template <class T>
__device__ __forceinline__ V4<T>& nodeRec(T* a, int64_t node)
{
    return reinterpret_cast<V4<T>*>(a)[node];
}

// Adds a lane's partial sum into the LDS tile.  Float form: v * scale (exact:
// the scale is a power of two) rounded to the nearest integer as int64 -- added
// in double onto 1.5 x 2^52 it lands in the low mantissa bits (exact for
// |v * scale| < 2^51), minus the constant's bit pattern; |v * scale| is summed
// into chk, so one test of chk <= FIXED_LANE_LIMIT per lane bounds every value
// added (and catches NaN / Inf) without a branch per value.
template <class T, class AccT>
__device__ __forceinline__ void ldsAdd(AccT* a, T v, float scale, float& chk)
{
    if constexpr (std::is_same<AccT, long long>::value) {
        const float vs = (float)v * scale;
        chk += fabsf(vs);
        // the constant's low word is zero, so only the high word needs the offset
        const double q = (double)vs + FIXED_MAGIC;
        const unsigned lo = (unsigned)__double2loint(q);
        const unsigned hi = (unsigned)__double2hiint(q) - (unsigned)(FIXED_MAGIC_BITS >> 32);
        atomicAdd((unsigned long long*)a, ((unsigned long long)hi << 32) | lo);
    }
    else
        unsafeAtomicAdd(a, v);
}

// Particle fields P2G reads, loaded ahead of use.
template <class T>
struct RawP2G {
    PosT<T> x[3];
    T v[3], C[9], S[6], m, d;
};

// pos_only: the position alone (the weights' half of the P2G payload).
template <class T, class AccT, bool DAMP = false>
__device__ __forceinline__ void loadP2G(const Dev<T, AccT>& D, int s, bool dmg, RawP2G<T>& r, bool pos_only = false)
{
    int p = DAMP ? s : ldAt(D.perm, s);
    if ((unsigned)p >= (unsigned)D.n) p = 0;
    const int64_t cap = D.cap;
    const T* P = D.P;
#pragma unroll
    for (int a = 0; a < 3; a++) r.x[a] = loadPos(P + (PX + a) * cap, p);
    if (pos_only) return;
#pragma unroll
    for (int a = 0; a < 3; a++) r.v[a] = ldAt(P + (PV + a) * cap, p);
#pragma unroll
    for (int i = 0; i < 9; i++) r.C[i] = ldAt(P + (PC + i) * cap, p);
#pragma unroll
    for (int i = 0; i < 6; i++) r.S[i] = DAMP ? T(0) : ldAt(P + (PS + i) * cap, p);
    r.m = ldAt(P + PM * cap, p);
    r.d = !DAMP && dmg && D.mats[matOf(D.pmat[p])].has_damage ? ldAt(P + PD * cap, p) : T(0);
}

// particlesToGridWithForceHelper<scheme, deg 2, mls> (MpmSimulationBase.cpp:791-854).
// FLIP_blend_PIC: no affine term (C = 0).  MLS-MPM: the force weights grad(w)
// become w D^-1 (x_i - x_p) (= 4 w (i - d0) / dx per axis), set in phase A.
// DAMP: the damping passes' particlesToGridHelper (MpmSimulationBase.cpp:860-900):
// mass and momentum only, particles at their sorted positions (written there by
// G2P) binned by this substep's sort while their weights come from the moved
// positions -- the reference reads grid memory through the substep's base
// offsets (MpmGrid.h:437-488, particle_base_offset).
// reading the stress of forcesUpdateParticleState (:953-979) from the particle,
// plus the AnisoMPM phase-field transfer (AnisoFractureSimulation.h:161-176).
//
// One workgroup per grid block (grid-stride over blocks).  The block's sorted
// particles are taken CHUNK at a time, counted from the block's first particle.
// Phase A: two threads per particle (its weights; its momentum, affine and
// stress terms) write its P2G payload to LDS.  Phase B: the
// cells holding particles of the chunk are compacted; lanes are (sub-lane,
// cell, i, j) columns of the 27-node stencil; each lane walks every R-th
// particle of its cell (all share one base node), accumulates 3 nodes x 9
// quantities in registers and adds them into an LDS tile of the block + halo.
// After the last chunk the tile is flushed to HBM with one atomic per non-zero
// node value.  FP32 form: per-lane float partial sums in sorted order, then
// int64 fixed point in LDS and HBM, so the result does not depend on scheduling.
enum : int { P2G_STD = 0, P2G_MLS = 1, P2G_DAMP = 2 };

// Waves per SIMD the P2G kernel is compiled for by default: the float form fits
// 4 (at most 128 VGPRs), the double form 3.  The autotuner also times one wave
// fewer (more registers) and an XCD-contiguous block order; no variant changes
// a result (the in-block order fixes every sum).
template <class T> struct P2GWaves { static constexpr int value = sizeof(T) == 4 ? 4 : 3; };

template <class T, class AccT, int VAR = P2G_STD, int OCC = P2GWaves<T>::value, bool XCD = false>
__global__ __launch_bounds__(WG, OCC) void k_p2g(Dev<T, AccT> D)
{
    constexpr bool DAMP = VAR == P2G_DAMP, MLS = VAR == P2G_MLS;
    using B = Blk<T>;
    using TL = Tile<T>;
    constexpr int NB = TL::NB, CH = TL::CHUNK;
    constexpr bool FIXED = std::is_same<AccT, long long>::value;
    __shared__ V4<T> q[7][CH];
    __shared__ T wl[4][CH][3];
    __shared__ AccT tile[NGF][TL::TN];
    __shared__ int cs[NB], ce[NB], cl[NB];
    __shared__ int hist[NB], cur[NB], cst[NB];
    __shared__ int ncl, nbl[8], nxr[2];
    // the in-block sort stages (key, live index) in the payload area and the
    // grouped copy in the tile area, both free until the block's first chunk
    constexpr int SA = (int)(sizeof(q) / 8), SB = (int)(sizeof(tile) / 8);
    constexpr int STAGE = SA < SB ? SA : SB;
    uint32_t* const ska = reinterpret_cast<uint32_t*>(&q[0][0]);
    int* const pa = reinterpret_cast<int*>(&q[0][0]) + SA;
    uint32_t* const skb = reinterpret_cast<uint32_t*>(&tile[0][0]);
    int* const pb = reinterpret_cast<int*>(&tile[0][0]) + SB;

    const int t = threadIdx.x;
    // phase A: particle pt of the chunk.  Float form: two threads per particle,
    // its weights (t < CH) and its other terms (t >= CH); the double form, short
    // of registers, computes both in thread t < CH
    constexpr bool SPLIT = sizeof(T) == 4;
    static_assert(WG == 2 * CH, "two threads per chunk particle in phase A");
    const int pt = SPLIT ? (t & (CH - 1)) : t;
    const bool doW = !SPLIT || t < CH, doM = !SPLIT || t >= CH;
    const bool dmg = !DAMP && D.any_damage != 0;
    const int nq = DAMP ? GFO : (dmg ? NGF : GPF);
    const int64_t cap = D.cap, cn = D.cap_nodes;
    const int nblocks = D.counters[0];

    // in-block order of this workgroup's blocks first (a loop of its own, so
    // its registers are free again for the scatter)
    // blocks round-robin over the XCDs by default: neighbouring blocks flush
    // their shared halo nodes from different L2s (an XCD-contiguous mapping, as
    // in G2P, made P2G ~25 % slower on the damage slab); both loops take the
    // same blocks
    const int wgl = XCD ? xcdLocal((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    for (int b = wgl; !DAMP && b < nblocks; b += gridDim.x) {
        const int bs = D.block_start[b], cnt = D.block_count[b];
        if (cnt <= 0) { // halo-only block
            if (t < NB) {
                D.cell_start[b * NB + t] = bs;
                D.cell_count[b * NB + t] = 0;
            }
            continue;
        }
        ldsBarrier(); // LDS of the previous block's order done
        blockOrder(D, b, bs, cnt, ska, pa, skb, pb, STAGE, hist, cur, cst);
    }
    // the order just written (perm, cell starts) is read below: a full barrier;
    // the loop below exchanges data through LDS only
    __syncthreads();
    int ovf = 0;
    bool have = false; // raw holds this thread's particle of the coming chunk
    RawP2G<T> raw;
    // cell ranges and neighbour table of the workgroup's next block, loaded a
    // block ahead (thread t < NB: cell t; threads 64..71: neighbour t - 64)
    int ns = 0, ne = 0, nn = -1;
    auto fetchBlock = [&](int bb) {
        if (bb >= nblocks) return;
        if (t < NB) {
            const int c = bb * NB + t;
            ns = D.cell_start[c];
            ne = ns + D.cell_count[c];
        }
        else if (t >= 64 && t < 72)
            nn = D.nbr[bb * 8 + t - 64];
    };
    fetchBlock(wgl);
    for (int b = wgl; b < nblocks; b += gridDim.x) {
        ldsBarrier(); // previous block flushed
        if (t < NB) {
            cs[t] = ns;
            ce[t] = ne;
        }
        else if (t >= 64 && t < 72)
            nbl[t - 64] = nn;
        const int bn = b + (int)gridDim.x;
        fetchBlock(bn);
        ldsBarrier();
        const int bs = cs[0], be = ce[NB - 1];
        if (be <= bs) continue; // halo-only block
        for (int e = t; e < nq * TL::TN; e += WG) (&tile[0][0])[e] = AccT(0);
        // the block's scene: its dx and fixed-point scales
        const int bscene = blockScene(D, b);
        const SceneDev<T>& sd = D.sc[bscene];
        const float* sc = D.scales + bscene * NSC;
        const float scl[NGF] = { sc[SC_M], sc[SC_MV], sc[SC_MV], sc[SC_MV], sc[SC_F], sc[SC_F], sc[SC_F], sc[SC_PF], sc[SC_PW] };
        for (int lo = bs; lo < be; lo += CH) {
            const int hi = min(lo + CH, be);
            ldsBarrier(); // previous chunk's phase B done with the payload
            // ---- phase A: payload of particle lo + pt from the prefetched fields:
            // its B-spline weights (doW) and its momentum, affine and stress terms (doM)
            if (pt < CH && lo + pt < hi) {
                if (!have) loadP2G<T, AccT, DAMP>(D, lo + pt, dmg, raw, !doM);
                T d0s[3], xb[3];
                int base[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    cellOf(raw.x[a], sd.odx, sd.origin[a], sd.domain, base[a], d0s[a], sd.pos_bits); // errors latched by the binning
                    if constexpr (sizeof(T) == 4)
                        xb[a] = -d0s[a] * sd.dx; // x_base - x_p
                    else
                        xb[a] = T(base[a] - sd.origin[a]) * sd.dx - raw.x[a];
                }
                if (doW) {
                    T w[3][3], dw[3][3];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        bsplineFrac(d0s[a], w[a], dw[a]);
                        if constexpr (MLS) // D^-1 dx w (x_i - x_p) in place of dw
#pragma unroll
                            for (int i = 0; i < 3; i++) dw[a][i] = (sd.Dinv * sd.dx) * w[a][i] * (xb[a] + T(i) * sd.dx);
                    }
                    q[5][pt] = V4<T> { w[2][0], w[2][1], w[2][2], T(0) };
                    q[6][pt] = V4<T> { dw[2][0], dw[2][1], dw[2][2], T(0) };
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        wl[0][pt][i] = w[0][i];
                        wl[1][pt][i] = dw[0][i];
                        wl[2][pt][i] = w[1][i];
                        wl[3][pt][i] = dw[1][i];
                    }
                }
                if (doM) {
                    const T mass = raw.m;
                    T mC[9], a3[3];
#pragma unroll
                    for (int i = 0; i < 9; i++) mC[i] = D.flip ? T(0) : mass * raw.C[i];
#pragma unroll
                    for (int r = 0; r < 3; r++)
                        a3[r] = mass * raw.v[r] + mC[3 * r] * xb[0] + mC[3 * r + 1] * xb[1] + mC[3 * r + 2] * xb[2];
                    T Sv[6];
#pragma unroll
                    for (int i = 0; i < 6; i++) Sv[i] = raw.S[i] * sd.odx; // float form stores -S below
                    const T dx = sd.dx;
                    q[0][pt] = V4<T> { mass, a3[0], a3[1], a3[2] };
                    q[1][pt] = V4<T> { mC[0] * dx, mC[1] * dx, mC[2] * dx, mC[3] * dx };
                    q[2][pt] = V4<T> { mC[4] * dx, mC[5] * dx, mC[6] * dx, mC[7] * dx };
                    // the force is -V0 tau grad(w): the float form accumulates it
                    // with FMAs of the negated stress
                    const T sg = sizeof(T) == 4 ? T(-1) : T(1);
                    q[3][pt] = V4<T> { mC[8] * dx, sg * Sv[0], sg * Sv[1], sg * Sv[2] };
                    q[4][pt] = V4<T> { sg * Sv[3], sg * Sv[4], sg * Sv[5], raw.d };
                }
            }
            // ---- cells holding particles of this chunk (NB <= 64: one wavefront)
            if (t < 64) {
                bool act = false;
                if (t < NB) act = min(ce[t], hi) > max(cs[t], lo);
                const uint64_t m = __ballot(act);
                if (act) cl[__popcll(m & ((1ull << t) - 1))] = t;
                if (t == 0) ncl = __popcll(m);
            }
            // the next block's particle range, for the prefetch of its first chunk
            if (sizeof(T) == 4 && hi == be && bn < nblocks) {
                if (t == 0) nxr[0] = ns;
                if (t == NB - 1) nxr[1] = ne;
            }
            ldsBarrier();
            // float form: prefetch the next chunk of this block, or after its last
            // chunk the first chunk of the next block, while phase B runs (the
            // double form keeps its registers for phase B)
            have = false;
            if (sizeof(T) == 4) {
                if (hi + pt < be) {
                    loadP2G<T, AccT, DAMP>(D, hi + pt, dmg, raw, !doM);
                    have = true;
                }
                else if (hi == be && bn < nblocks && nxr[0] + pt < nxr[1]) {
                    loadP2G<T, AccT, DAMP>(D, nxr[0] + pt, dmg, raw, !doM);
                    have = true;
                }
            }
            // ---- phase B: lanes = (sub-lane, cell, i, j); R sub-lanes share a column
            const int per = ncl * 9;
            const int R = per > 0 ? max(1, min(D.p2g_split, WG / per)) : 1;
            for (int slot = t; slot < per * R; slot += WG) {
                const int sub = slot / per, rem = slot - per * (slot / per);
                const int cell = cl[rem / 9], col = rem - 9 * (rem / 9);
                const int ci = col / 3, cj = col - 3 * (col / 3);
                const int p0 = max(cs[cell], lo) - lo + sub, p1 = min(ce[cell], hi) - lo;
                if (p0 >= p1) continue;
                const T fi = T(ci), fj = T(cj);
                T acc[3][NGF];
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int f = 0; f < NGF; f++) acc[k][f] = T(0);
                for (int pp = p0; pp < p1; pp += R) {
                    const V4<T> q0 = q[0][pp], q1 = q[1][pp], q2 = q[2][pp], q3 = q[3][pp], q4 = q[4][pp];
                    const V4<T> q5 = q[5][pp], q6 = q[6][pp];
                    const T wx = wl[0][pp][ci], dwx = wl[1][pp][ci], wy = wl[2][pp][cj], dwy = wl[3][pp][cj];
                    const T wxy = wx * wy, gx = dwx * wy, gy = wx * dwy;
                    // momentum + m C (x_i - x_p) at the column base node (k = 0)
                    const T a0 = q0.y + q1.x * fi + q1.y * fj;
                    const T a1 = q0.z + q1.w * fi + q2.x * fj;
                    const T a2 = q0.w + q2.z * fi + q2.w * fj;
                    const T b0 = q1.z, b1 = q2.y, b2 = q3.x; // m C dx, z column
                    const T s00 = q3.y, s01 = q3.z, s02 = q3.w, s11 = q4.x, s12 = q4.y, s22 = q4.z;
                    const T wz[3] = { q5.x, q5.y, q5.z }, dwz[3] = { q6.x, q6.y, q6.z };
                    // V0 tau grad(w) at node k: float form as wz_k (S (gx, gy, 0)) +
                    // dwz_k (S e_z wxy); the double form keeps its registers for
                    // the per-node products
                    T u0, u1, u2, v0, v1, v2;
                    if constexpr (sizeof(T) == 4) {
                        u0 = s00 * gx + s01 * gy;
                        u1 = s01 * gx + s11 * gy;
                        u2 = s02 * gx + s12 * gy;
                        v0 = s02 * wxy;
                        v1 = s12 * wxy;
                        v2 = s22 * wxy;
                    }
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const T wk = wxy * wz[k];
                        const T fk = T(k);
                        acc[k][GM] += q0.x * wk;
                        acc[k][GMV + 0] += wk * (k == 0 ? a0 : a0 + b0 * fk);
                        acc[k][GMV + 1] += wk * (k == 0 ? a1 : a1 + b1 * fk);
                        acc[k][GMV + 2] += wk * (k == 0 ? a2 : a2 + b2 * fk);
                        if constexpr (DAMP) {
                        }
                        else if constexpr (sizeof(T) == 4) { // s.. hold -V0 tau / dx here
                            acc[k][GFO + 0] += wz[k] * u0;
                            acc[k][GFO + 0] += dwz[k] * v0;
                            acc[k][GFO + 1] += wz[k] * u1;
                            acc[k][GFO + 1] += dwz[k] * v1;
                            acc[k][GFO + 2] += wz[k] * u2;
                            acc[k][GFO + 2] += dwz[k] * v2;
                        }
                        else {
                            const T g0 = gx * wz[k], g1 = gy * wz[k], g2 = wxy * dwz[k];
                            acc[k][GFO + 0] -= s00 * g0 + s01 * g1 + s02 * g2;
                            acc[k][GFO + 1] -= s01 * g0 + s11 * g1 + s12 * g2;
                            acc[k][GFO + 2] -= s02 * g0 + s12 * g1 + s22 * g2;
                        }
                        if constexpr (!DAMP) {
                            acc[k][GPF] += q4.w * wk;
                            acc[k][GPW] += wk;
                        }
                    }
                }
                const int cx = cell >> (B::ybits + B::zbits), cy = (cell >> B::zbits) & (TL::BY - 1), cz = cell & (TL::BZ - 1);
                const int nbase = ((cx + ci) * TL::TY + (cy + cj)) * TL::TZ + cz;
                float chk = 0.f;
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int f = 0; f < NGF; f++)
                        if (f < nq) ldsAdd<T, AccT>(&tile[f][nbase + k], acc[k][f], scl[f], chk);
                if (FIXED && !(chk <= FIXED_LANE_LIMIT)) ovf = 1;
            }
        }
        ldsBarrier();
        // ---- flush the tile: one HBM atomic per non-zero node value
        for (int e = t; e < nq * TL::TN; e += WG) {
            const int f = e / TL::TN, nd = e - TL::TN * (e / TL::TN);
            const AccT v = tile[f][nd];
            if (v == AccT(0)) continue;
            const int tx = nd / (TL::TY * TL::TZ), ty = (nd / TL::TZ) % TL::TY, tz = nd % TL::TZ;
            const int which = ((tx >= TL::BX) << 2) | ((ty >= TL::BY) << 1) | (tz >= TL::BZ);
            const int tb = nbl[which];
            if (tb < 0) {
                ovf |= 2;
                continue;
            }
            const int elem = ((tx & (TL::BX - 1)) << (B::ybits + B::zbits)) | ((ty & (TL::BY - 1)) << B::zbits) | (tz & (TL::BZ - 1));
            AccT* dst = D.acc + f * cn + (int64_t)tb * NB + elem;
            if constexpr (FIXED)
                atomicAdd((unsigned long long*)dst, (unsigned long long)v);
            else
                unsafeAtomicAdd(dst, v);
        }
    }
    if (ovf & 1) latch(D.counters, ST_OVERFLOW);
    if (ovf & 2) latch(D.counters, ST_CAPACITY);
}

template <class T, class AccT>
__device__ __forceinline__ T accValue(const Dev<T, AccT>& D, int f, int64_t node, const float* sc, int sg)
{
    if constexpr (std::is_same<AccT, long long>::value)
        return (T)((double)D.acc[f * D.cap_nodes + node] * (1.0 / (double)sc[sg]));
    else
        return D.acc[f * D.cap_nodes + node];
}

// ------------------------------------------------------- collision objects

// AnalyticLevelSet::queryInside in material space X: inside flag and the
// material-space normal N (half space AnalyticLevelSet.cpp:331-347; the axis-
// aligned box's AutoDiff normal :562-588 -- the first largest d_a inside, sign
// +1 at 0, the clamped norm's derivative 0 at 0; Sphere::queryInside :494-511;
// CappedCylinder AnalyticLevelSet.h:290-320 in its primitive frame).
template <class T>
__device__ __forceinline__ bool levelSetInside(const ColDev& co, const T* X, T* N)
{
    N[0] = N[1] = N[2] = T(0);
    if (co.shape == ZMPM_SHAPE_HALF_SPACE) {
        const T phi = (T)co.n[0] * (X[0] - (T)co.origin[0]) + (T)co.n[1] * (X[1] - (T)co.origin[1])
            + (T)co.n[2] * (X[2] - (T)co.origin[2]);
#pragma unroll
        for (int a = 0; a < 3; a++) N[a] = (T)co.n[a];
        return phi <= T(0);
    }
    if (co.shape == ZMPM_SHAPE_BOX) {
        T dd[3], Xc[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T c0 = (T)((co.bmin[a] + co.bmax[a]) / 2), he = (T)((co.bmax[a] - co.bmin[a]) / 2);
            Xc[a] = X[a] - c0;
            dd[a] = dabs(Xc[a]) - he;
        }
        int am = 0;
        if (dd[1] > dd[am]) am = 1;
        if (dd[2] > dd[am]) am = 2;
        T q2 = T(0);
#pragma unroll
        for (int a = 0; a < 3; a++) q2 += dd[a] > T(0) ? dd[a] * dd[a] : T(0);
        const T dmax = dd[am];
        const T phi = (dmax < T(0) ? dmax : T(0)) + dsqrt(q2);
        if (q2 > T(0)) {
            const T qn = dsqrt(q2);
#pragma unroll
            for (int a = 0; a < 3; a++) N[a] = (dd[a] > T(0) ? dd[a] / qn : T(0)) * (Xc[a] < T(0) ? T(-1) : T(1));
        }
        else {
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (a == am) N[a] = Xc[a] < T(0) ? T(-1) : T(1);
        }
        return phi <= T(0);
    }
    if (co.shape == ZMPM_SHAPE_SPHERE) {
        T tc[3], d2 = T(0);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            tc[a] = X[a] - (T)co.origin[a];
            d2 += tc[a] * tc[a];
        }
        const T dist = dsqrt(d2);
        if (dist < T(1e-7))
            N[0] = T(1);
        else {
            const T id = T(1) / dist;
#pragma unroll
            for (int a = 0; a < 3; a++) N[a] = id * tc[a];
        }
        const T r = (T)co.radius;
        return d2 < r * r;
    }
    // capped cylinder, in the primitive frame Rc^-1 (X - centre)
    T Y[3], Xp[3];
#pragma unroll
    for (int a = 0; a < 3; a++) Y[a] = X[a] - (T)co.origin[a];
#pragma unroll
    for (int a = 0; a < 3; a++) Xp[a] = (T)co.Rci[3 * a] * Y[0] + (T)co.Rci[3 * a + 1] * Y[1] + (T)co.Rci[3 * a + 2] * Y[2];
    const T rxz = dsqrt(Xp[0] * Xp[0] + Xp[2] * Xp[2]);
    const T d0 = rxz - (T)co.radius, d1 = dabs(Xp[1]) - T(0.5) * (T)co.height;
    T g0[3] = { T(0), T(0), T(0) };
    const T g1y = Xp[1] < T(0) ? T(-1) : T(1);
    if (rxz != T(0)) {
        g0[0] = Xp[0] / rxz;
        g0[2] = Xp[2] / rxz;
    }
    const T q0 = d0 < T(0) ? T(0) : d0, q1 = d1 < T(0) ? T(0) : d1;
    const T mx = d0 < d1 ? d1 : d0;
    const T qn = dsqrt(q0 * q0 + q1 * q1);
    const T phi = (T(0) < mx ? T(0) : mx) + qn;
    T n[3] = { T(0), T(0), T(0) };
    if (!(T(0) < mx)) {
        if (d0 < d1)
            n[1] = g1y;
        else {
            n[0] = g0[0];
            n[2] = g0[2];
        }
    }
    if (qn != T(0)) {
        const T iq = T(1) / qn;
        n[0] += q0 * g0[0] * iq;
        n[1] += q1 * g1y * iq;
        n[2] += q0 * g0[2] * iq;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) N[a] = (T)co.Rc[3 * a] * n[0] + (T)co.Rc[3 * a + 1] * n[1] + (T)co.Rc[3 * a + 2] * n[2];
    return phi <= T(0);
}

// AnalyticCollisionObject::detectAndResolveCollision (CollisionObject.cpp:341-406)
// of grid node xi: true when the node is inside; vi resolved; n the world normal
// of a SLIP / SEPARATE object.
template <class T>
__device__ __forceinline__ bool resolveCollision(const ColDev& co, const ColState& st, const T* xi, T* vi, T* n)
{
    T xb[3], X[3], N[3];
#pragma unroll
    for (int a = 0; a < 3; a++) xb[a] = xi[a] - (T)st.b[a];
    if (co.moving) {
        const T os = (T)st.os;
#pragma unroll
        for (int a = 0; a < 3; a++)
            X[a] = ((T)st.R[a] * xb[0] + (T)st.R[3 + a] * xb[1] + (T)st.R[6 + a] * xb[2]) * os;
    }
    else {
#pragma unroll
        for (int a = 0; a < 3; a++) X[a] = xb[a];
    }
    if (!levelSetInside(co, X, N)) return false;
    // v_object = omega x (x - b) + (dsdt / s)(x - b) + dbdt
    T vo[3];
    if (co.moving) {
        const T w0 = (T)co.omega[0], w1 = (T)co.omega[1], w2 = (T)co.omega[2];
        const T sr = (T)(co.dsdt * st.os);
        vo[0] = (w1 * xb[2] - w2 * xb[1]) + sr * xb[0] + (T)co.dbdt[0];
        vo[1] = (w2 * xb[0] - w0 * xb[2]) + sr * xb[1] + (T)co.dbdt[1];
        vo[2] = (w0 * xb[1] - w1 * xb[0]) + sr * xb[2] + (T)co.dbdt[2];
#pragma unroll
        for (int a = 0; a < 3; a++) n[a] = (T)st.R[3 * a] * N[0] + (T)st.R[3 * a + 1] * N[1] + (T)st.R[3 * a + 2] * N[2];
    }
    else {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            vo[a] = (T)co.dbdt[a];
            n[a] = N[a];
        }
    }
    if (co.type == ZMPM_STICKY) {
#pragma unroll
        for (int a = 0; a < 3; a++) vi[a] = vo[a];
        return true;
    }
    T vr[3] = { vi[0] - vo[0], vi[1] - vo[1], vi[2] - vo[2] };
    const T dot = vr[0] * n[0] + vr[1] * n[1] + vr[2] * n[2];
    if (co.type == ZMPM_SLIP || dot < T(0)) {
#pragma unroll
        for (int a = 0; a < 3; a++) vr[a] -= n[a] * dot;
        const T fr = (T)co.friction;
        if (fr != T(0) && dot < T(0)) { // kinematic friction, only against approach
            const T vn = dsqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
            if (-dot * fr < vn) {
#pragma unroll
                for (int a = 0; a < 3; a++) vr[a] += (vr[a] / vn) * dot * fr;
            }
            else
                vr[0] = vr[1] = vr[2] = T(0);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) vi[a] = vr[a] + vo[a];
    return true;
}

// AnalyticCollisionObject::multiObjectCollision (CollisionObject.cpp:66-105):
// objects in insertion order; the first STICKY hit ends the loop; the normals
// of SLIP / SEPARATE hits are Gram-Schmidt orthonormalised and the loop ends
// once 3 are independent (a normal in the span of the earlier ones adds none).
// mask: bit c clear for objects that hold no node of the node's block (they
// would leave vi as it is).
template <class T>
__device__ __forceinline__ void multiObjectCollision(const ColDev* cols, const ColState* st, int ncols, const T* xi,
                                                     T* vi, uint32_t mask)
{
    int count = 0;
    T b0[3] = { T(0), T(0), T(0) }, b1[3] = { T(0), T(0), T(0) };
    for (int c = 0; c < ncols; c++) {
        if (!((mask >> c) & 1u)) continue;
        const ColDev& co = cols[c];
        if (co.type == ZMPM_GHOST) continue;
        T n[3];
        if (!resolveCollision(co, st[c], xi, vi, n)) continue;
        if (co.type == ZMPM_STICKY) break;
        if (count >= 1) {
            const T d = b0[0] * n[0] + b0[1] * n[1] + b0[2] * n[2];
#pragma unroll
            for (int a = 0; a < 3; a++) n[a] -= d * b0[a];
        }
        if (count >= 2) {
            const T d = b1[0] * n[0] + b1[1] * n[1] + b1[2] * n[2];
#pragma unroll
            for (int a = 0; a < 3; a++) n[a] -= d * b1[a];
        }
        const T len = dsqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len != T(0)) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const T u = n[a] / len;
                if (count == 0) b0[a] = u;
                if (count == 1) b1[a] = u;
            }
            if (++count == 3) break;
        }
    }
}

// Whether collision object co may hold a node of the block whose first node is
// i0 (scene frame, node units; ext nodes per axis): the block's bounding sphere,
// taken to the object's material space (a rotation, then 1 / s), against the
// level set's inside region grown by its radius.  False only when no node of
// the block is inside, so an object it rules out leaves every node as it is.
template <class T>
__device__ __forceinline__ bool colliderMayHold(const ColDev& co, const ColState& st, const int* i0, const int* ext,
                                                double h)
{
    if (co.type == ZMPM_GHOST) return false;
    double c[3], r2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double e = 0.5 * (ext[a] - 1) * h;
        c[a] = i0[a] * h + e - st.b[a];
        r2 += e * e;
    }
    // the margin covers the rounding of the per-node level-set evaluation
    double r = sqrt(r2) * 1.001 + 1e-9 * h;
    double X[3];
    if (co.moving) {
#pragma unroll
        for (int a = 0; a < 3; a++) X[a] = (st.R[a] * c[0] + st.R[3 + a] * c[1] + st.R[6 + a] * c[2]) * st.os;
        r *= fabs(st.os);
    }
    else {
#pragma unroll
        for (int a = 0; a < 3; a++) X[a] = c[a];
    }
    if (co.shape == ZMPM_SHAPE_HALF_SPACE) {
        const double phi = co.n[0] * (X[0] - co.origin[0]) + co.n[1] * (X[1] - co.origin[1]) + co.n[2] * (X[2] - co.origin[2]);
        return phi <= r * sqrt(co.n[0] * co.n[0] + co.n[1] * co.n[1] + co.n[2] * co.n[2]);
    }
    if (co.shape == ZMPM_SHAPE_BOX) {
        bool in = true;
#pragma unroll
        for (int a = 0; a < 3; a++)
            in = in && fabs(X[a] - 0.5 * (co.bmin[a] + co.bmax[a])) <= 0.5 * (co.bmax[a] - co.bmin[a]) + r;
        return in;
    }
    if (co.shape == ZMPM_SHAPE_SPHERE) {
        double d2 = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) d2 += (X[a] - co.origin[a]) * (X[a] - co.origin[a]);
        return sqrt(d2) <= co.radius + r;
    }
    double Y[3], Xp[3];
#pragma unroll
    for (int a = 0; a < 3; a++) Y[a] = X[a] - co.origin[a];
#pragma unroll
    for (int a = 0; a < 3; a++) Xp[a] = co.Rci[3 * a] * Y[0] + co.Rci[3 * a + 1] * Y[1] + co.Rci[3 * a + 2] * Y[2];
    return sqrt(Xp[0] * Xp[0] + Xp[2] * Xp[2]) <= co.radius + r && fabs(Xp[1]) <= 0.5 * co.height + r;
}

// v = mv/m (MpmSimulationBase.cpp:655-666) and gridVelocityExplicitUpdate
// (:1179-1228) of one node at its scene's time t, without the collision
// objects; phase-field normalisation (AnisoFractureSimulation.h:177-179).  A
// node without mass gets its record (the force, for the host's grid view) and
// false; otherwise v and pf are returned for the caller to store, and col tells
// whether the scene's objects apply (not ignored, no explicit velocity field).
template <class T, class AccT>
__device__ __forceinline__ bool gridNode(const Dev<T, AccT>& D, int64_t node, int b, int scene, T* vi, T& pf, bool& col)
{
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    const SceneDev<T>& sd = D.sc[scene];
    const float* sc = D.scales + scene * NSC;
    const T m = accValue(D, GM, node, sc, SC_M);
    T f[3];
#pragma unroll
    for (int a = 0; a < 3; a++) f[a] = accValue(D, GFO + a, node, sc, SC_F);
    if (m == T(0)) {
        nodeRec(D.newv, node) = V4<T> { f[0], f[1], f[2], T(0) };
        if (D.flip) nodeRec(D.oldv, node) = V4<T> { T(0), T(0), T(0), T(0) };
        return false;
    }
    const T dt = sd.dt;
    const T dt_over_mi = dt / m;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        T v = accValue(D, GMV + a, node, sc, SC_MV) / m;
        if (D.flip) D.oldv[4 * node + a] = v; // the grid's v before the update (FLIP)
        vi[a] = v + sd.gravity[a] * dt + f[a] * dt_over_mi;
    }
    // external body forces, in the order added (MpmSimulationBase.cpp:1208-1212)
    for (int k = 0; k < sd.nbf; k++)
#pragma unroll
        for (int a = 0; a < 3; a++) vi[a] += dt * (T)D.bfval[3 * (scene * MAX_BODY_FORCES + k) + a];
    col = !sd.ignore_cols && !sd.vfield; // ignoreCollisionObject (:1217-1221)
    if (sd.vfield) {
        // explicit_velocity_field (:1187-1195): new_v = a + B x_i at every node
        // with mass, in place of the update above (the objects do not apply)
        const int e = (int)(node % NB);
        uint32_t Bx, By, Bz;
        blockOfPage<T>((uint32_t)D.block_page[b], Bx, By, Bz);
        const int ix = (int)((Bx << B::xbits) | (uint32_t)(e >> (B::ybits + B::zbits))) - sd.origin[0];
        const int iy = (int)((By << B::ybits) | (uint32_t)((e >> B::zbits) & ((1 << B::ybits) - 1))) - sd.origin[1];
        const int iz = (int)((Bz << B::zbits) | (uint32_t)(e & ((1 << B::zbits) - 1))) - sd.origin[2];
        const double xi[3] = { ix * sd.ddx, iy * sd.ddx, iz * sd.ddx };
#pragma unroll
        for (int a = 0; a < 3; a++)
            vi[a] = (T)(sd.va[a] + (sd.vB[3 * a] * xi[0] + sd.vB[3 * a + 1] * xi[1] + sd.vB[3 * a + 2] * xi[2]));
    }
    pf = T(0);
    if (D.any_damage) {
        const T pw = accValue(D, GPW, node, sc, SC_PW);
        pf = pw != T(0) ? accValue(D, GPF, node, sc, SC_PF) / pw : T(0);
    }
    return true;
}

// The collision objects of a grid-update workgroup: those of its scene staged
// in LDS, read from there so the per-node loop's loads cannot alias the
// kernel's stores (with the table in global memory, stores of doubles anywhere
// in the kernel made the compiler load it per node with vector loads: ~2x
// slower on large grids); a scene's blocks are contiguous, so a workgroup's
// nodes rarely span two scenes, and those that do read the tables in global
// memory.  Per block of the iteration (WG nodes), the objects that may hold one
// of its nodes: bit c of the block's word, two words so the next iteration's
// need no second barrier.
template <class T>
struct ColLds {
    static constexpr int NB = 1 << (Blk<T>::xbits + Blk<T>::ybits + Blk<T>::zbits);
    static constexpr int BPI = WG / NB; // blocks per iteration
    ColDev cols[MAX_COLLIDERS];
    ColState cst[MAX_COLLIDERS];
    uint32_t bmask[2][BPI];
};
static_assert(sizeof(ColDev) % 8 == 0 && sizeof(ColState) % 8 == 0, "collider records are copied as 8-byte words");
static_assert(MAX_COLLIDERS == 16, "16 object lanes per block of an iteration");

// Called by every thread for the iteration of nodes [base, base + WG) (it:
// the iteration's parity): stages the objects of its scene when all its nodes
// are in one (staged / nstaged: the scene in LDS and its object count, the same
// in every thread), returns that first scene and whether all nodes are in it,
// and leaves the block words in L; it ends on a barrier.
template <class T, class AccT>
__device__ __forceinline__ void collidePrepare(const Dev<T, AccT>& D, ColLds<T>& L, int64_t base, int64_t nn, int it,
                                               int& staged, int& nstaged, int& first, bool& same)
{
    using B = Blk<T>;
    constexpr int NB = ColLds<T>::NB, BPI = ColLds<T>::BPI;
    const int nblocks = (int)(nn / NB);
    const int s_first = blockScene(D, (int)(base / NB));
    const int s_last = blockScene(D, (int)((min(base + WG, nn) - 1) / NB));
    const bool uniform = s_first == s_last;
    if (uniform && s_first != staged) {
        __syncthreads(); // the previous scene's objects are no longer read
        const SceneDev<T>& sd0 = D.sc[s_first];
        nstaged = min(sd0.ncol, MAX_COLLIDERS);
        for (int e = threadIdx.x; e < nstaged * (int)(sizeof(ColDev) / 8); e += WG)
            reinterpret_cast<uint64_t*>(L.cols)[e] = reinterpret_cast<const uint64_t*>(D.cols + sd0.col0)[e];
        for (int e = threadIdx.x; e < nstaged * (int)(sizeof(ColState) / 8); e += WG)
            reinterpret_cast<uint64_t*>(L.cst)[e] = reinterpret_cast<const uint64_t*>(D.colstate + sd0.col0)[e];
        __syncthreads();
        staged = s_first;
    }
    // thread (lb, c): block lb of the iteration, object c of its scene
    const int lb = (int)threadIdx.x / MAX_COLLIDERS, c = (int)threadIdx.x % MAX_COLLIDERS;
    const int bb = (int)(base / NB) + lb;
    bool hold = false;
    if (lb < BPI && bb < nblocks) {
        const int s = uniform ? s_first : blockScene(D, bb);
        const SceneDev<T>& sb = D.sc[s];
        if (!sb.ignore_cols && !sb.vfield && c < min(sb.ncol, MAX_COLLIDERS)) {
            uint32_t Bx, By, Bz;
            blockOfPage<T>((uint32_t)D.block_page[bb], Bx, By, Bz);
            const int i0[3] = { (int)(Bx << B::xbits) - sb.origin[0], (int)(By << B::ybits) - sb.origin[1],
                                (int)(Bz << B::zbits) - sb.origin[2] };
            const int ext[3] = { 1 << B::xbits, 1 << B::ybits, 1 << B::zbits };
            hold = uniform ? colliderMayHold<T>(L.cols[c], L.cst[c], i0, ext, sb.ddx)
                           : colliderMayHold<T>(D.cols[sb.col0 + c], D.colstate[sb.col0 + c], i0, ext, sb.ddx);
        }
    }
    const uint64_t bal = __ballot(hold);
    if (lb < BPI && c == 0) L.bmask[it][lb] = (uint32_t)(bal >> (threadIdx.x & 48)) & 0xffffu;
    first = s_first;
    same = uniform;
    ldsBarrier();
}

// The collision objects of the block words in L on node `node` (block b, of
// scene `scene`) with velocity v: multiObjectCollision / detectAndResolveCollision
// (CollisionObject.cpp:66-105, 341-406) at the scene's time t, in double in
// both forms (a node on an object's surface, a wall at a whole number of
// cells, is classified as in double).
template <class T, class AccT>
__device__ __forceinline__ void collideNode(const Dev<T, AccT>& D, const ColLds<T>& L, int nstaged, bool uniform,
                                            uint32_t cmask, int64_t node, int b, int scene, T* vi)
{
    using B = Blk<T>;
    constexpr int NB = ColLds<T>::NB;
    const int e = (int)(node % NB);
    const SceneDev<T>& sd = D.sc[scene];
    uint32_t Bx, By, Bz;
    blockOfPage<T>((uint32_t)D.block_page[b], Bx, By, Bz);
    const int ix = (int)((Bx << B::xbits) | (uint32_t)(e >> (B::ybits + B::zbits))) - sd.origin[0];
    const int iy = (int)((By << B::ybits) | (uint32_t)((e >> B::zbits) & ((1 << B::ybits) - 1))) - sd.origin[1];
    const int iz = (int)((Bz << B::zbits) | (uint32_t)(e & ((1 << B::zbits) - 1))) - sd.origin[2];
    const double xi[3] = { ix * sd.ddx, iy * sd.ddx, iz * sd.ddx };
    double vd[3] = { vi[0], vi[1], vi[2] };
    if (uniform)
        multiObjectCollision(L.cols, L.cst, nstaged, xi, vd, cmask);
    else
        multiObjectCollision(D.cols + sd.col0, D.colstate + sd.col0, min(sd.ncol, MAX_COLLIDERS), xi, vd, cmask);
#pragma unroll
    for (int a = 0; a < 3; a++) vi[a] = (T)vd[a];
}

// The grid update (gridNode); with COL also its collision objects (collideNode)
// on the nodes of the blocks they may reach, in this one launch.  Without COL
// k_collide applies them next: a second launch, but this kernel then needs no
// double-precision object code and runs at twice the occupancy.
template <class T, class AccT, bool COL>
__global__ __launch_bounds__(WG, COL ? 4 : 1) void k_grid(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    const int64_t nn = (int64_t)D.counters[0] * NB;
    const int64_t cn = D.cap_nodes;
    if (D.implicit && D.any_damage) {
        // the damage solve's accumulators, filled by k_dtilde
        for (int64_t node = (int64_t)blockIdx.x * WG + threadIdx.x; node < nn; node += (int64_t)gridDim.x * WG) {
            D.cgv[node] = 0.0;
            D.cgv[cn + node] = 0.0;
            D.cgv[4 * cn + node] = 0.0;
        }
    }
    if constexpr (!COL) {
        for (int64_t node = (int64_t)blockIdx.x * WG + threadIdx.x; node < nn; node += (int64_t)gridDim.x * WG) {
            const int b = (int)(node / NB);
            T vi[3], pf;
            bool col;
            if (gridNode(D, node, b, blockScene(D, b), vi, pf, col)) nodeRec(D.newv, node) = V4<T> { vi[0], vi[1], vi[2], pf };
        }
    }
    else {
        __shared__ ColLds<T> L;
        int staged = -1, nstaged = 0, it = 0;
        for (int64_t base = (int64_t)blockIdx.x * WG; base < nn; base += (int64_t)gridDim.x * WG, it ^= 1) {
            int s_first;
            bool uniform;
            collidePrepare(D, L, base, nn, it, staged, nstaged, s_first, uniform);
            const int64_t node = base + threadIdx.x;
            if (node >= nn) continue;
            const int b = (int)(node / NB);
            const int scene = uniform ? s_first : blockScene(D, b);
            T vi[3], pf;
            bool col = false;
            if (!gridNode(D, node, b, scene, vi, pf, col)) continue;
            const uint32_t cmask = L.bmask[it][threadIdx.x / NB];
            if (col && cmask != 0u) collideNode(D, L, nstaged, uniform, cmask, node, b, scene, vi);
            nodeRec(D.newv, node) = V4<T> { vi[0], vi[1], vi[2], pf };
        }
    }
}

// The collision objects of the grid update after k_grid<COL = false>, by
// groups of 16 blocks: lanes (block, object) test whether the object may hold
// a node of the block (reading the object tables from global memory, where 16
// lanes share each record); then, only when some block of the group is
// reached, the objects of the group's scene are staged in LDS and the nodes
// with mass of the reached blocks resolve them on the velocity k_grid stored
// (in the float form a float, widened to double as the fused update widens it,
// so the result is the one of resolving inside the update).
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_collide(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    constexpr int GB = WG / MAX_COLLIDERS; // blocks per group
    static_assert(MAX_COLLIDERS == 16 && (GB * NB) % WG == 0, "a group is whole iterations of WG nodes");
    const int nblocks = D.counters[0];
    __shared__ ColDev cols[MAX_COLLIDERS];
    __shared__ ColState cst[MAX_COLLIDERS];
    __shared__ uint32_t gmask[GB];
    __shared__ int any;
    int staged = -1, nstaged = 0; // the scene whose objects are in LDS (the same in every thread)
    for (int g0 = (int)blockIdx.x * GB; g0 < nblocks; g0 += (int)gridDim.x * GB) {
        const int s_first = blockScene(D, g0);
        const bool uniform = s_first == blockScene(D, min(g0 + GB, nblocks) - 1);
        // thread (lb, c): block g0 + lb, object c of its scene
        const int lb = (int)threadIdx.x / MAX_COLLIDERS, c = (int)threadIdx.x % MAX_COLLIDERS;
        const int bb = g0 + lb;
        bool hold = false;
        if (bb < nblocks) {
            const int sc = uniform ? s_first : blockScene(D, bb);
            const SceneDev<T>& sb = D.sc[sc];
            if (!sb.ignore_cols && !sb.vfield && c < min(sb.ncol, MAX_COLLIDERS)) {
                uint32_t Bx, By, Bz;
                blockOfPage<T>((uint32_t)D.block_page[bb], Bx, By, Bz);
                const int i0[3] = { (int)(Bx << B::xbits) - sb.origin[0], (int)(By << B::ybits) - sb.origin[1],
                                    (int)(Bz << B::zbits) - sb.origin[2] };
                const int ext[3] = { 1 << B::xbits, 1 << B::ybits, 1 << B::zbits };
                hold = colliderMayHold<T>(D.cols[sb.col0 + c], D.colstate[sb.col0 + c], i0, ext, sb.ddx);
            }
        }
        const uint64_t bal = __ballot(hold);
        __syncthreads(); // the previous group's words and flag are no longer read
        if (threadIdx.x == 0) any = 0;
        ldsBarrier();
        if ((threadIdx.x & 15) == 0) gmask[lb] = (uint32_t)(bal >> (threadIdx.x & 48)) & 0xffffu;
        if ((threadIdx.x & 63) == 0 && bal != 0) any = 1;
        ldsBarrier();
        if (!any) continue; // the same in every thread
        if (uniform && s_first != staged) {
            const SceneDev<T>& sd0 = D.sc[s_first];
            nstaged = min(sd0.ncol, MAX_COLLIDERS);
            for (int e = threadIdx.x; e < nstaged * (int)(sizeof(ColDev) / 8); e += WG)
                reinterpret_cast<uint64_t*>(cols)[e] = reinterpret_cast<const uint64_t*>(D.cols + sd0.col0)[e];
            for (int e = threadIdx.x; e < nstaged * (int)(sizeof(ColState) / 8); e += WG)
                reinterpret_cast<uint64_t*>(cst)[e] = reinterpret_cast<const uint64_t*>(D.colstate + sd0.col0)[e];
            __syncthreads();
            staged = s_first;
        }
        const int64_t gn = (int64_t)min(g0 + GB, nblocks) * NB;
        for (int64_t node = (int64_t)g0 * NB + threadIdx.x; node < gn; node += WG) {
            const int b = (int)(node / NB);
            const uint32_t cmask = gmask[b - g0];
            if (cmask == 0u) continue;
            if (D.acc[GM * D.cap_nodes + node] == AccT(0)) continue; // no mass: the record holds the force
            const int e = (int)(node % NB);
            const int scene = uniform ? s_first : blockScene(D, b);
            const SceneDev<T>& sd = D.sc[scene];
            uint32_t Bx, By, Bz;
            blockOfPage<T>((uint32_t)D.block_page[b], Bx, By, Bz);
            const int ix = (int)((Bx << B::xbits) | (uint32_t)(e >> (B::ybits + B::zbits))) - sd.origin[0];
            const int iy = (int)((By << B::ybits) | (uint32_t)((e >> B::zbits) & ((1 << B::ybits) - 1))) - sd.origin[1];
            const int iz = (int)((Bz << B::zbits) | (uint32_t)(e & ((1 << B::zbits) - 1))) - sd.origin[2];
            // in double in both forms, as the fused update
            const double xi[3] = { ix * sd.ddx, iy * sd.ddx, iz * sd.ddx };
            V4<T>& rec = nodeRec(D.newv, node);
            const V4<T> v0 = rec;
            double vd[3] = { v0.x, v0.y, v0.z };
            if (uniform)
                multiObjectCollision(cols, cst, nstaged, xi, vd, cmask);
            else
                multiObjectCollision(D.cols + sd.col0, D.colstate + sd.col0, min(sd.ncol, MAX_COLLIDERS), xi, vd, cmask);
            rec = V4<T> { (T)vd[0], (T)vd[1], (T)vd[2], v0.w };
        }
    }
}

// Damping pass (MpmSimulationBase.cpp:318-348), before its P2G: the substep's
// blocks' mass and momentum accumulators cleared; float form: the mass and
// momentum scales from the bound slots the last G2P wrote (read, not cleared:
// the next substep's scales take the maximum over every pass).
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_damp_prep(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    const int nb = D.counters[CNT_BLOCKS];
    const int t = threadIdx.x;
    if constexpr (std::is_same<AccT, long long>::value) {
        if (blockIdx.x == gridDim.x - 1) {
            // one wavefront per (scene, mass or momentum group)
            const int lane = t & 63;
            for (int k = t >> 6; k < 2 * D.nscenes; k += WG / 64) {
                const int g = (k & 1) ? SC_MV : SC_M, sc = k >> 1;
                const float m = waveMax(__uint_as_float(D.maxbits[(sc * NSC + g) * BOUND_SLOTS + lane]));
                if (lane == 0 && m > 0.f && isfinite(m)) D.scales[sc * NSC + g] = boundScale(m, 1.f);
            }
            return;
        }
    }
    const int64_t nn = (int64_t)nb * NB;
    const int nw = std::is_same<AccT, long long>::value ? (int)gridDim.x - 1 : (int)gridDim.x;
    for (int64_t e = (int64_t)blockIdx.x * WG + t; e < nn; e += (int64_t)nw * WG)
#pragma unroll
        for (int f = GM; f < GFO; f++) D.acc[f * D.cap_nodes + e] = AccT(0);
}

// Damping pass grid: v = mv/m where m != 0, else 0, and new_v = v
// (MpmSimulationBase.cpp:322-327 with the particlesToGrid tail :655-666).
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_damp_grid(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    const int64_t nn = (int64_t)D.counters[0] * NB;
    for (int64_t node = (int64_t)blockIdx.x * WG + threadIdx.x; node < nn; node += (int64_t)gridDim.x * WG) {
        const float* sc = D.scales + blockScene(D, (int)(node / NB)) * NSC;
        const T m = accValue(D, GM, node, sc, SC_M);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T v = m != T(0) ? accValue(D, GMV + a, node, sc, SC_MV) / m : T(0);
            D.newv[4 * node + a] = v;
            if (D.flip) D.oldv[4 * node + a] = v;
        }
    }
}

// The crack driving force D~ = zeta <phi - 1>_+ of one particle
// (updateDamageHelper / updateDTildeHelper, AnisoFractureSimulation.h:222-268,
// 296-332; structural tensor :365-391, undegraded Cauchy stress
// QRStableNeoHookean.h:83-93).  Fs is the stored deformation: F (FP64 form) or
// H = F - I (FP32 form).  FAST (G2P): a particle whose phi bound does not rule
// out <phi - 1>_+ > 0 is not evaluated; the function returns false and the
// caller queues the particle for k_rare, which runs the full path.  That keeps
// the eigen split and the polar rotation, rarely needed, out of G2P's registers.
template <class T, bool FAST = false>
__device__ __forceinline__ bool dTildeOf(const MatDev<T>& mt, const Mat3<T>& Fs, T& dTilde)
{
    // undegraded Cauchy stress tau / J (evaluateEffectiveCauchyStress): g = 1 and no
    // fibres (QRStableNeoHookean.h:83-93), every fibre at k_fiber_isotropic = mu
    // (QRAnisotropic.h:79-91), or g = 1 (NeoHookeanBorden.cpp:96-108); tau is
    // exactly symmetric in its closed form
    T J, t6[6];
    if (mt.model == ZMPM_MODEL_QR_ANISOTROPIC) {
        if constexpr (sizeof(T) == 4)
            qrAnisoTauH(Fs, mt.lambda, mt.mu, 1.f, mt.mu, mt.mu, mt.mu, true, t6, &J);
        else
            qrAnisoTauH(minusIdentity(Fs), mt.lambda, mt.mu, T(1), mt.mu, mt.mu, mt.mu, true, t6, &J);
    }
    else if (mt.model == ZMPM_MODEL_NEO_HOOKEAN_BORDEN) { // NeoHookeanBorden.cpp:96-108
        if constexpr (sizeof(T) == 4)
            bordenTauH(Fs, mt.mu, mt.kappa, 1.f, t6, &J);
        else
            bordenTauH(minusIdentity(Fs), mt.mu, mt.kappa, T(1), t6, &J);
    }
    else if constexpr (sizeof(T) == 4)
        qrStableTauH(Fs, mt.mu, mt.lambda, 1.f, 0.f, 0.f, t6, &J);
    else
        qrStableTau(Fs, mt.mu, mt.lambda, T(1), T(0), T(0), t6, &J);
    const T iJ = ddiv(T(1), J);
    Mat3<T> sym;
    sym.a[0][0] = t6[0] * iJ;
    sym.a[0][1] = sym.a[1][0] = t6[1] * iJ;
    sym.a[0][2] = sym.a[2][0] = t6[2] * iJ;
    sym.a[1][1] = t6[3] * iJ;
    sym.a[1][2] = sym.a[2][1] = t6[4] * iJ;
    sym.a[2][2] = t6[5] * iJ;
    // phi = tr(A s+ A s+) / sigma_c^2 <= |A|_2^2 |s+|_F^2 / sigma_c^2
    //     <= (1 + sum |alpha_k| |a0_k|^2)^2 |s|_F^2 / sigma_c^2 (phi_bound_scale):
    // when that is below 1 (with a rounding margin), <phi - 1>_+ is exactly 0 and
    // sigma^+ and the polar rotation contribute nothing
    const T sf2 = sym.a[0][0] * sym.a[0][0] + sym.a[1][1] * sym.a[1][1] + sym.a[2][2] * sym.a[2][2]
        + T(2) * (sym.a[0][1] * sym.a[0][1] + sym.a[0][2] * sym.a[0][2] + sym.a[1][2] * sym.a[1][2]);
    dTilde = T(0);
    if (!(mt.phi_bound_scale * sf2 < T(0.99))) {
        if constexpr (FAST) return false;
        const Mat3<T> sp = positivePart(sym);
        Mat3<T> R;
        if constexpr (sizeof(T) == 4)
            R = polarRotation(identityPlus(Fs));
        else
            R = polarRotation(Fs);
        Mat3<T> A;
        setIdentity(A);
        for (int k = 0; k < mt.nd; k++) {
            T a1[3];
#pragma unroll
            for (int i = 0; i < 3; i++) a1[i] = R.a[i][0] * mt.a0[k][0] + R.a[i][1] * mt.a0[k][1] + R.a[i][2] * mt.a0[k][2];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) A.a[i][j] += mt.alphas[k] * (a1[i] * a1[j]);
        }
        Mat3<T> As = mul(A, sp), sA = mul(sp, A);
        T contraction = T(0);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) contraction += As.a[i][j] * sA.a[i][j];
        const T phi = mt.inv_sigma_crit2 * contraction;
        dTilde = mt.zeta * macaulay(phi - T(1));
    }
    return true;
}

// updateDamageHelper for one particle (AnisoFractureSimulation.h:210-283) on
// F_n.  FAST (G2P): returns true, d and g unchanged, when the exact path is
// needed (the particle goes to k_rare).
template <class T, bool FAST = false>
__device__ __forceinline__ bool damageUpdate(const MatDev<T>& mt, const Mat3<T>& Fs, T lap, T dt, T& d, T& g)
{
    const T geometricResist = d - (mt.l0 * mt.l0 * lap);
    T dTilde;
    if (!dTildeOf<T, FAST>(mt, Fs, dTilde)) return true;
    const T diff = ((T(1) - d) * dTilde) - geometricResist;
    const T newD = d + (ddiv(dt, mt.eta) * macaulay(diff));
    d = newD < T(1) ? newD : T(1);
    const T k = mt.k;
    g = ((T(1) - d) * (T(1) - d) * (T(1) - k)) + k;
    return false;
}

// gridToParticlesHelper<APIC, deg 2, !MLS> (MpmSimulationBase.cpp:1412-1478),
// CFL flags (:1466-1499), evolveStrain (FBasedMpmForceHelper.cpp:93-108,
// JBasedMpmForceHelper.cpp:79-94), the AnisoMPM Laplacian gather
// (iterateKernelWithLaplacian, MpmGrid.h:552-623, with its 3D index quirk) and
// damage update on F_n, then the stress of (F_{n+1}, g_{n+1}) that the next
// substep's forcesUpdateParticleState would compute, and the next P2G's
// fixed-point bounds.  One thread per sorted particle (max_slots workgroups of
// WG particles); the grid velocities of each block segment of the workgroup's
// particles (block + halo) are staged in LDS first.  Output goes to sorted
// position s of the other buffer.
// Per-particle G2P state carried from the gather to the finish.
template <class T>
struct G2PAcc {
    T picV[3], Bp[3][3], gv[3][3], lap;
    T oldV[3]; // sum w v_old (FLIP_blend_PIC)
};

// APIC gather of one particle from the staged block tile.  The 27-node sums
// (gridToParticlesHelper, MpmSimulationBase.cpp:1412-1478) factor over the
// tensor-product weights, so they run as sums over k (z), then j (y), then i
// (x):  v = sum w v_i,  B = sum w v_i (x_i - x_p)^T,  grad v = sum v_i grad(w)^T
// with grad(w) = (dw_x w_y w_z, w_x dw_y w_z, w_x w_y dw_z) / dx, and the
// damage Laplacian sum d_i L_i with the reference's 3D weights
// L = (ddw_x(i) w_y w_z + w_x ddw_x(i) w_z + w_x w_y ddw_x(i)) / dx^2
// (MpmGrid.h:610-612 index ddw[1], ddw[2] with i as well); under the implicit
// damage solve the staged field is x - pf and the sum is sum w (x_i - pf_i)
// (damageG2PHelper, AnisoFractureSimulation.h:631-642).
// FLIP: also sum w v_old over the staged old velocities (fields 4-6).  cell >= 0:
// the stencil's tile position from that cell of the segment's block (the damping
// passes: this substep's binning, moved particle); otherwise from the particle.
template <class T, class AccT, bool FLIP = false>
__device__ __forceinline__ void g2pGather(const Dev<T, AccT>& D, const SceneDev<T>& sd, const V4<T> (*gt)[Tile<T>::TN],
                                          const PosT<T>* Xp, bool dmg, G2PAcc<T>& A, int cell = -1)
{
    using TL = Tile<T>;
    const T odx = sd.odx, dx = sd.dx;
    T w[3][3], dw[3][3], off[3][3];
    int base[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        T d0;
        cellOf(Xp[a], odx, sd.origin[a], sd.domain, base[a], d0, sd.pos_bits);
        bsplineFrac(d0, w[a], dw[a]);
#pragma unroll
        for (int i = 0; i < 3; i++) { // x_i - x_p
            if constexpr (sizeof(T) == 4)
                off[a][i] = (T(i) - d0) * dx;
            else
                off[a][i] = T(base[a] - sd.origin[a] + i) * dx - Xp[a];
        }
    }
    const T wzo[3] = { w[2][0] * off[2][0], w[2][1] * off[2][1], w[2][2] * off[2][2] };
    using B = Blk<T>;
    const int lx0 = cell >= 0 ? cell >> (B::ybits + B::zbits) : base[0] & (TL::BX - 1);
    const int ly0 = cell >= 0 ? (cell >> B::zbits) & (TL::BY - 1) : base[1] & (TL::BY - 1);
    const int lz0 = cell >= 0 ? cell & (TL::BZ - 1) : base[2] & (TL::BZ - 1);
    T lap = T(0);
#pragma unroll 1
    for (int i = 0; i < 3; i++) {
        // sums over j of the z-sums: tv = sum wy s0, ty = sum wy (y_j - y_p) s0,
        // tdy = sum dwy s0, tz = sum wy s2, tdz = sum wy s1
        T tv[3] = { T(0), T(0), T(0) }, ty[3] = { T(0), T(0), T(0) }, tdy[3] = { T(0), T(0), T(0) };
        T tz[3] = { T(0), T(0), T(0) }, tdz[3] = { T(0), T(0), T(0) };
        T to[3] = { T(0), T(0), T(0) };
        T q0 = T(0), q1 = T(0), q2 = T(0);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int nrow = ((lx0 + i) * TL::TY + (ly0 + j)) * TL::TZ + lz0;
            // s0 = sum wz v, s1 = sum dwz v, s2 = sum wz (z_k - z_p) v; p0 = sum wz d, p1 = sum d
            T s0[3] = { T(0), T(0), T(0) }, s1[3] = { T(0), T(0), T(0) }, s2[3] = { T(0), T(0), T(0) };
            T so[3] = { T(0), T(0), T(0) };
            T p0 = T(0), p1 = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const V4<T> n4 = gt[0][nrow + k];
                const T nvs[3] = { n4.x, n4.y, n4.z };
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const T nv = nvs[a];
                    s0[a] += w[2][k] * nv;
                    s1[a] += dw[2][k] * nv;
                    s2[a] += wzo[k] * nv;
                }
                if constexpr (FLIP) {
                    const V4<T> o4 = gt[1][nrow + k];
                    so[0] += w[2][k] * o4.x;
                    so[1] += w[2][k] * o4.y;
                    so[2] += w[2][k] * o4.z;
                }
                if (dmg) {
                    const T ph = n4.w;
                    p0 += w[2][k] * ph;
                    p1 += ph;
                }
            }
            const T wy = w[1][j], dwy = dw[1][j], wyo = w[1][j] * off[1][j];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                tv[a] += wy * s0[a];
                ty[a] += wyo * s0[a];
                tdy[a] += dwy * s0[a];
                tz[a] += wy * s2[a];
                tdz[a] += wy * s1[a];
                if constexpr (FLIP) to[a] += wy * so[a];
            }
            if (dmg) {
                q0 += wy * p0;
                q1 += p0;
                q2 += wy * p1;
            }
        }
        const T wx = w[0][i], dwx = dw[0][i], wxo = w[0][i] * off[0][i];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            A.picV[a] += wx * tv[a];
            A.Bp[a][0] += wxo * tv[a];
            A.Bp[a][1] += wx * ty[a];
            A.Bp[a][2] += wx * tz[a];
            A.gv[a][0] += dwx * tv[a];
            A.gv[a][1] += wx * tdy[a];
            A.gv[a][2] += wx * tdz[a];
            if constexpr (FLIP) A.oldV[a] += wx * to[a];
        }
        if (dmg) lap += D.implicit ? wx * q0 : ((i == 1) ? T(-2) : T(1)) * (q0 + wx * (q1 + q2));
    }
    // (called once per particle: the particle's stencil lies in one block tile)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) A.gv[a][c] *= odx;
    A.lap += D.implicit ? lap : lap * (odx * odx);
}

// Particle update after the gather; writes sorted position s.
// Particle fields G2P reads after the gather, loaded before it.
template <class T>
struct RawG2P {
    PosT<T> x[3];
    T F[9], m, V0, d, g;
    T v[3]; // FLIP_blend_PIC only
    T ys;   // plasticity state (when any material has plasticity)
    T J;    // FJ-mixed J (FP32 form: J - 1), when any material is FJ-mixed
    int mid, id;
};

// G2P's particle fields in two parts: what the gather needs (position and
// material word) and what only the finish needs (F, m, V0, d, g, original
// index, plastic state, J; FLIP: v).  The float form loads the second part
// after the gather, so it is not live across it: 120 instead of 134 VGPRs, 4
// wavefronts per SIMD instead of 3 (G2P ~8 % faster); the double form, at 2
// wavefronts either way, loads both up front.
template <class T, class AccT>
__device__ __forceinline__ void loadG2PGather(const Dev<T, AccT>& D, int p, RawG2P<T>& r)
{
    const int64_t cap = D.cap;
#pragma unroll
    for (int a = 0; a < 3; a++) r.x[a] = loadPos(D.P + (PX + a) * cap, p);
    r.mid = ldAt(D.pmat, p);
}

template <class T, class AccT, bool FLIP = false>
__device__ __forceinline__ void loadG2PFinish(const Dev<T, AccT>& D, int s, int p, RawG2P<T>& r)
{
    const int64_t cap = D.cap;
    const T* P = D.P;
#pragma unroll
    for (int i = 0; i < 9; i++) r.F[i] = ldAt(P + (PF + i) * cap, p);
    r.m = ldAt(P + PM * cap, p);
    r.V0 = ldAt(P + PVOL * cap, p);
    r.d = ldAt(P + PD * cap, p);
    r.g = ldAt(P + PG * cap, p);
    r.id = ldAt(D.spid, s);
    r.ys = D.any_plastic ? ldAt(P + PYS * cap, p) : T(0);
    r.J = D.any_fj ? ldAt(P + PJ * cap, p) : T(0);
    if constexpr (FLIP)
#pragma unroll
        for (int a = 0; a < 3; a++) r.v[a] = ldAt(P + (PV + a) * cap, p);
}

// Returns true for a particle deferred to k_rare (its d, g and stress are
// completed there).
template <class T, class AccT, bool FLIP = false>
__device__ __forceinline__ bool g2pFinish(const Dev<T, AccT>& D, int s, const RawG2P<T>& R, const SceneDev<T>& sd,
                                          const G2PAcc<T>& A, const MatDev<T>& mt, float* bnd, PosT<T>* xnew)
{
    bool rare = false;
    constexpr bool FIXED = std::is_same<AccT, long long>::value;
    const int64_t cap = D.cap;
    T* Q = D.Q;
    const bool dmg = D.any_damage && mt.has_damage;
    const T dt = sd.dt;
    const T dx = sd.dx;
    const T Dinv = sd.Dinv;
    const T r = sd.rpic;
    const T ca = (r + T(1)) * T(0.5), cb = (r - T(1)) * T(0.5);
    // grad v of evolveStrain: sum v_i grad(w)^T, or C under MLS-MPM (APIC)
    T gvF[3][3];
    // v, C, x first (values die early: the momentum bound is taken here)
    {
        T Cn[9], vn[3];
        if constexpr (FLIP) {
            // FLIP_blend_PIC (MpmSimulationBase.cpp:1431-1432): v <- r v, v += pic - r old;
            // no C (stored as 0)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                T va = D.flip_ratio * R.v[a];
                va += A.picV[a] - D.flip_ratio * A.oldV[a];
                vn[a] = va;
            }
#pragma unroll
            for (int i = 0; i < 9; i++) Cn[i] = T(0);
        }
        else {
#pragma unroll
            for (int a = 0; a < 3; a++) vn[a] = A.picV[a];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 3; c++) Cn[3 * a + c] = ca * (A.Bp[a][c] * Dinv) + cb * (A.Bp[c][a] * Dinv);
        }
#pragma unroll
        for (int a = 0; a < 3; a++) stAt(Q + (PV + a) * cap, s, vn[a]);
#pragma unroll
        for (int i = 0; i < 9; i++) stAt(Q + (PC + i) * cap, s, Cn[i]);
        const bool mls = !FLIP && D.mls;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++) gvF[a][c] = mls ? Cn[3 * a + c] : A.gv[a][c];
        T inc[3], inc2 = T(0);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            inc[a] = dt * A.picV[a];
            inc2 += inc[a] * inc[a];
        }
        if (inc2 > dx * dx) latch(D.counters, ST_CFL);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if constexpr (sizeof(T) == 4) {
                // fixed point: exact integer add of the rounded increment; a
                // non-finite increment marks the position (the binning latches
                // it), an out-of-range one saturates (the domain check rejects it)
                const float u = inc[a] * sd.pos_scale;
                int xn = POS_NAN;
                if (u == u && R.x[a] != POS_NAN) {
                    const int du = __float2int_rn(fminf(fmaxf(u, -1073741824.f), 1073741824.f));
                    xn = min(max(__builtin_elementwise_add_sat(R.x[a], du), POS_MIN), POS_MAX);
                }
                xnew[a] = xn;
            }
            else
                xnew[a] = R.x[a] + inc[a];
            storePos(Q + (PX + a) * cap, s, xnew[a]);
        }
        if constexpr (FIXED) momentumBounds(bnd, R.m, vn, Cn, dx);
    }
    stAt(Q + PM * cap, s, R.m);
    stAt(Q + PVOL * cap, s, R.V0);
    stAt(Q + PLAP * cap, s, dmg && !D.implicit ? A.lap : T(0));
    stAt(D.qid, s, R.id);
    stAt(D.qmat, s, R.mid);
    T d = R.d, g = R.g;
    Mat3<T> Fn;
#pragma unroll
    for (int i = 0; i < 9; i++) Fn.a[i / 3][i % 3] = R.F[i];
    // F <- (I + dt grad v) F; J <- (1 + dt tr grad v) J.  FP32 form on H = F - I:
    // H <- H + G + G H with G = dt grad v; J - 1 <- (J - 1) + dt tr(grad v) J.
    const T dtr = dt * (gvF[0][0] + gvF[1][1] + gvF[2][2]);
    // FJMixedMpmForceHelper::evolveStrain (FJMixedMpmForceHelper.h:168-187): F and
    // J = (1 + dt tr grad v) J of a solid; a fluid's are left as they are
    T jn = R.J;
    if (D.any_fj) {
        if (mt.fj == ZMPM_FJ_SOLID) {
            if constexpr (sizeof(T) == 4)
                jn = jn + dtr * (1.f + jn);
            else
                jn = (T(1) + dtr) * jn;
        }
        stAt(Q + PJ * cap, s, jn);
    }
    if (mt.model == ZMPM_MODEL_EQUATION_OF_STATE) {
        if constexpr (sizeof(T) == 4)
            Fn.a[0][0] = Fn.a[0][0] + dtr * (1.f + Fn.a[0][0]);
        else
            Fn.a[0][0] = (T(1) + dtr) * Fn.a[0][0];
    }
    else if (!D.any_fj || mt.fj != ZMPM_FJ_FLUID) {
        const Mat3<T> F = Fn;
        if (dmg && phaseFieldModel(mt.model)) {
            if (D.implicit) {
                // damageG2PHelper (AnisoFractureSimulation.h:644-678): the solve's
                // change, never healing, clamped at 1
                const T nd = d + A.lap;
                const T mx = d < nd ? nd : d; // std::max(d, new_d)
                d = T(1) < mx ? T(1) : mx;    // std::min(., 1)
                g = ((T(1) - d) * (T(1) - d) * (T(1) - mt.k)) + mt.k;
            }
            else
                rare = damageUpdate<T, true>(mt, F, A.lap, dt, d, g);
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                T v = F.a[a][c];
                if constexpr (sizeof(T) == 4) v += dt * gvF[a][c];
#pragma unroll
                for (int k = 0; k < 3; k++) v += dt * gvF[a][k] * F.a[k][c];
                Fn.a[a][c] = v;
            }
    }
    stAt(Q + PD * cap, s, d);
    stAt(Q + PG * cap, s, g);
#pragma unroll
    for (int i = 0; i < 9; i++) stAt(Q + (PF + i) * cap, s, Fn.a[i / 3][i % 3]);
    // plasticity (applyPlasticity, MpmSimulationBase.cpp:1517-1529): the return
    // mapping runs in k_rare for particles that may yield
    if (D.any_plastic) {
        stAt(Q + PYS * cap, s, R.ys);
        if (mt.plasticity) rare = rare || plasticMayProject(mt, Fn, R.ys);
    }
    // stress of the new state for the next substep's P2G
    T S[6];
    if (D.fj_solve_j && mt.fj)
        fjSolveJStress(mt, false, R.V0, jn, S);
    else
        rare = rare || vtauSym<true>(mt, Fn, R.V0 * snowScale(mt, R.ys), g, S);
    if (!rare) {
#pragma unroll
        for (int i = 0; i < 6; i++) stAt(Q + (PS + i) * cap, s, S[i]);
        if constexpr (FIXED) forceBounds(bnd, S, dmg ? d : T(0), sd.odx, D.any_damage != 0);
    }
    return rare;
}

// Grid node records of block b's tile (the block and its forward halo, through
// the neighbour table) into gt: gt[0] the new velocity and phase field, and
// with nrec = 2 gt[1] the grid's velocity before the update (FLIP).  A node
// outside the grid stages zeros.
template <class T, class AccT>
__device__ __forceinline__ void stageTile(const Dev<T, AccT>& D, int b, V4<T> (*gt)[Tile<T>::TN], int nrec)
{
    using B = Blk<T>;
    using TL = Tile<T>;
    constexpr int NB = TL::NB;
    // the neighbour table row, read once (a uniform address) and selected per node
    const int* nb = D.nbr + b * 8;
    int nbv[8];
#pragma unroll
    for (int i = 0; i < 8; i++) nbv[i] = nb[i];
    for (int e = threadIdx.x; e < nrec * TL::TN; e += WG) {
        const int r = e / TL::TN, nd = e - TL::TN * (e / TL::TN);
        const int tx = nd / (TL::TY * TL::TZ), ty = (nd / TL::TZ) % TL::TY, tz = nd % TL::TZ;
        const int which = ((tx >= TL::BX) << 2) | ((ty >= TL::BY) << 1) | (tz >= TL::BZ);
        int tb = nbv[0];
#pragma unroll
        for (int i = 1; i < 8; i++) tb = which == i ? nbv[i] : tb;
        V4<T> v { T(0), T(0), T(0), T(0) };
        if (tb >= 0) {
            const int elem = ((tx & (TL::BX - 1)) << (B::ybits + B::zbits)) | ((ty & (TL::BY - 1)) << B::zbits) | (tz & (TL::BZ - 1));
            const int64_t node = (int64_t)tb * NB + elem;
            v = nodeRec(r == 0 ? D.newv : D.oldv, node);
        }
        gt[r][nd] = v;
    }
}

// RR: workgroups take particle segments round-robin over the XCDs instead of
// XCD by XCD (an autotuner variant; results are the same).
template <class T, class AccT, bool FLIP, bool RR = false>
__global__ __launch_bounds__(WG) void k_g2p(Dev<T, AccT> D)
{
    using TL = Tile<T>;
    constexpr int NB = TL::NB;
    constexpr int NGT = FLIP ? 7 : 4;
    __shared__ V4<T> gt[(NGT + 3) / 4][TL::TN];
    __shared__ int seg_end;
    const int lo = (RR ? (int)blockIdx.x : xcdLocal((int)blockIdx.x, (int)gridDim.x)) * WG;
    const int hi = min(lo + WG, D.n);
    const int t = threadIdx.x, s = lo + t;
    const bool live = s < hi;
    RawG2P<T> R;
    // the Laplacian is gathered for every particle when any material has damage
    // (stored only for damaged ones), so no per-particle material load has to
    // complete before the gathers
    const bool dmg = D.any_damage != 0;
    constexpr bool LATE = sizeof(T) == 4; // the finish's fields loaded after the gather
    int p = 0;
    if (live) {
        p = ldAt(D.perm, s);
        if ((unsigned)p >= (unsigned)D.n) p = 0;
        loadG2PGather<T, AccT>(D, p, R);
        if (!LATE) loadG2PFinish<T, AccT, FLIP>(D, s, p, R);
    }
    G2PAcc<T> A;
#pragma unroll
    for (int a = 0; a < 3; a++) A.picV[a] = A.oldV[a] = T(0);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        A.Bp[i / 3][i % 3] = T(0);
        A.gv[i / 3][i % 3] = T(0);
    }
    A.lap = T(0);
    const int scene = live ? sceneOf(R.mid) : 0;
    int seg_lo = lo;
    while (seg_lo < hi) {
        const int b = D.sorted_cell[seg_lo] / NB;
        // LDS-only barriers: the particle loads issued above stay in flight
        // while the tile is staged
        ldsBarrier(); // previous segment's gathers done
        if (t == 0) {
            const int c = b * NB + NB - 1;
            seg_end = D.cell_start[c] + D.cell_count[c];
        }
        stageTile(D, b, gt, FLIP ? 2 : 1);
        ldsBarrier();
        const int seg_hi = max(seg_lo + 1, min(hi, seg_end));
        // the segment's particles are in its block's scene (a uniform index)
        if (live && s >= seg_lo && s < seg_hi) g2pGather<T, AccT, FLIP>(D, D.sc[blockScene(D, b)], gt, R.x, dmg, A);
        seg_lo = seg_hi;
    }
    if (LATE && live) loadG2PFinish<T, AccT, FLIP>(D, s, p, R);
    float bnd[NSC] = { 0.f, 0.f, 0.f, 0.f, 0.f };
    PosT<T> xnew[3] = { 0, 0, 0 };
    bool rare = false;
    // a wavefront of one scene (all but those across two scenes' blocks) reads
    // the scene's table with scalar loads through a wave-uniform index
    const int s_first = __builtin_amdgcn_readfirstlane(scene);
    if (!__ballot(live && scene != s_first)) {
        if (live) rare = g2pFinish<T, AccT, FLIP>(D, s, R, D.sc[s_first], A, D.mats[matOf(R.mid)], bnd, xnew);
    }
    else if (live)
        rare = g2pFinish<T, AccT, FLIP>(D, s, R, D.sc[scene], A, D.mats[matOf(R.mid)], bnd, xnew);
    // deferred particles to the rare-path queue, one atomic per wavefront
    const uint64_t rm = __ballot(rare);
    if (rm) {
        const int lane = t & 63, first = __ffsll((unsigned long long)rm) - 1;
        int at = 0;
        if (lane == first) at = atomicAdd(&D.counters[CNT_RARE], __popcll(rm));
        at = __shfl(at, first);
        if (rare) stAt(D.rare, at + __popcll(rm & ((1ull << lane) - 1)), s);
    }
    // binning of the coming substep at the new positions (live index s)
    openPageScan(D);
    binParticle(D, s, live, xnew, scene);
    if constexpr (std::is_same<AccT, long long>::value) writeBounds(D.maxbits, bnd, live ? scene : -1);
}

// VonMisesStvkHencky::projectStrain (PlasticityApplier.cpp:1480-1516) on the
// stored deformation (FP32 form: H = F - I, formed into F in double): SVD of F,
// Hencky strain eps = log(max(sigma, 1e-4)) and its deviator eps_hat; when
// |eps_hat| > yield / (2 mu) the strain returns radially, F = U exp(H) V^T, and
// the yield stress hardens by xi delta_gamma.  In double for both forms (it runs
// for the few particles G2P defers).  Returns true when F changed.
template <class T>
__device__ __noinline__ bool vonMisesProject(const MatDev<T>& mt, Mat3<T>& Fs, T& ys)
{
    constexpr double id = sizeof(T) == 4 ? 1.0 : 0.0;
    Mat3<double> F, U, V;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) F.a[i][j] = (double)Fs.a[i][j] + (i == j ? id : 0.0);
    double sg[3];
    svd3(F, U, sg, V);
    double eps[3], tr = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        eps[a] = log(fmax(sg[a], 1e-4));
        tr += eps[a];
    }
    double hat[3], n2 = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        hat[a] = eps[a] - (tr / 3.0);
        n2 += hat[a] * hat[a];
    }
    const double nrm = sqrt(n2);
    const double dgam = nrm - (double)ys / (2 * (double)mt.mu);
    if (dgam <= 0) return false;
    ys = (T)((double)ys - (double)mt.hardening * dgam);
    double e[3];
#pragma unroll
    for (int a = 0; a < 3; a++) e[a] = exp(eps[a] - (dgam / nrm) * hat[a]);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            Fs.a[r][c] = (T)((U.a[r][0] * e[0] * V.a[c][0] + U.a[r][1] * e[1] * V.a[c][1] + U.a[r][2] * e[2] * V.a[c][2])
                             - (r == c ? id : 0.0));
    return true;
}

// DruckerPragerStvkHencky::projectStrain (PlasticityApplier.cpp:833-876) on the
// stored deformation, in double as vonMisesProject; ljp is log Jp.  F is
// rebuilt from the SVD in every case (|sigma| for an inverted F), as the
// reference does.  Returns true (F and ljp written back).
template <class T>
__device__ __noinline__ bool druckerPragerProject(const MatDev<T>& mt, Mat3<T>& Fs, T& ljp)
{
    constexpr double id = sizeof(T) == 4 ? 1.0 : 0.0;
    Mat3<double> F, U, V;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) F.a[i][j] = (double)Fs.a[i][j] + (i == j ? id : 0.0);
    double sg[3];
    svd3(F, U, sg, V);
    const double coh = (double)mt.cohesion;
    double lj = (double)ljp;
    double eps[3], sum = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        eps[a] = log(fmax(fabs(sg[a]), 1e-4)) - coh;
        sum += eps[a];
    }
    const double tr = sum + lj;
    double hat[3], n2 = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        hat[a] = eps[a] - (tr / 3.0);
        n2 += hat[a] * hat[a];
    }
    double e[3];
    if (tr >= 0) { // tip
#pragma unroll
        for (int a = 0; a < 3; a++) e[a] = exp(coh);
        if (mt.volume_correction) lj = (double)mt.hardening * sum + lj;
    }
    else if ((double)mt.mu == 0)
        return false;
    else {
        lj = 0;
        const double nrm = sqrt(n2);
        const double mu = (double)mt.mu, lam = (double)mt.lambda;
        const double dgam = nrm + (3 * lam + 2 * mu) / (2 * mu) * tr * (double)mt.dp_alpha;
#pragma unroll
        for (int a = 0; a < 3; a++) e[a] = exp(dgam <= 0 ? eps[a] + coh : eps[a] - (dgam / nrm) * hat[a] + coh);
    }
    ljp = (T)lj;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            Fs.a[r][c] = (T)((U.a[r][0] * e[0] * V.a[c][0] + U.a[r][1] * e[1] * V.a[c][1] + U.a[r][2] * e[2] * V.a[c][2])
                             - (r == c ? id : 0.0));
    return true;
}

// Whether a plastic particle's return mapping may change it, so that G2P
// defers it to k_rare.  FP64 form: always (the return mapping decides there,
// on the reference's SVD).  FP32 form: the Hencky strain from the eigenvalues
// l of F F^T - I (eps = log1p(l) / 2, cancellation-free); a margin covers the
// float rounding of the test; inverted or nearly degenerate F is always
// deferred.  von Mises: outside or near |eps_hat| = yield / (2 mu).  Drucker-
// Prager: anything but strictly inside the cone with log Jp = 0 (where the
// reference's rebuild of F from its SVD returns F).
// SnowPlasticity::projectStrain (PlasticityApplier.cpp:16-50) on the stored
// deformation, in double as vonMisesProject: singular values clamped to
// [1 - theta_c, 1 + theta_s] (std::min / std::max order), F = U Sigma V^T,
// Jp <- Jp det F / det Fe clamped to [min_Jp, max_Jp].  The reference also scales
// the particle's mu and lambda by exp(psi (Jp_old - Jp_new)); snowScale gives
// the accumulated factor exp(psi (1 - Jp)) from the stored Jp.  F is rebuilt
// from the SVD in every case, as the reference does.  Returns true.
template <class T>
__device__ __noinline__ bool snowProject(const MatDev<T>& mt, Mat3<T>& Fs, T& jp)
{
    constexpr double id = sizeof(T) == 4 ? 1.0 : 0.0;
    Mat3<double> F, U, V;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) F.a[i][j] = (double)Fs.a[i][j] + (i == j ? id : 0.0);
    double sg[3];
    svd3(F, U, sg, V);
    double fe_det = 1.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double v = mt.snow_hi < sg[a] ? mt.snow_hi : sg[a];
        v = v < mt.snow_lo ? mt.snow_lo : v;
        sg[a] = v;
        fe_det *= v;
    }
    double jn = (double)jp * det3(F) / fe_det;
    if (!(jn <= mt.max_jp)) jn = mt.max_jp;
    if (!(jn >= mt.min_jp)) jn = mt.min_jp;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            Fs.a[r][c] = (T)((U.a[r][0] * sg[0] * V.a[c][0] + U.a[r][1] * sg[1] * V.a[c][1] + U.a[r][2] * sg[2] * V.a[c][2])
                             - (r == c ? id : 0.0));
    jp = (T)jn;
    return true;
}

__device__ __forceinline__ bool plasticMayProject(const MatDev<double>&, const Mat3<double>&, double) { return true; }
__device__ __forceinline__ bool plasticMayProject(const MatDev<float>& mt, const Mat3<float>& H, float state)
{
    const bool dp = mt.plasticity == ZMPM_PLASTICITY_DRUCKER_PRAGER_STVK_HENCKY;
    if (dp && state != 0.f) return true;
    if (!(jm1OfH(H) > -0.95f)) return true;
    float E[6];
    strainOfH(H, E);
    float l1, l2, l3;
    symEigvals3(E, l1, l2, l3);
    if (mt.plasticity == ZMPM_PLASTICITY_SNOW) // a singular value may leave [1 - theta_c, 1 + theta_s]
        return !(l3 > mt.snow_lam_lo + 4e-6f && l1 < mt.snow_lam_hi - 4e-6f);
    if (!(1.f + l3 > 1e-8f)) return true;
    const float c = dp ? mt.cohesion : 0.f;
    const float e0 = 0.5f * log1pf(l1) - c, e1 = 0.5f * log1pf(l2) - c, e2 = 0.5f * log1pf(l3) - c;
    const float sum = e0 + e1 + e2, tr = sum * (1.f / 3.f);
    const float h0 = e0 - tr, h1 = e1 - tr, h2 = e2 - tr;
    const float n = sqrtf(h0 * h0 + h1 * h1 + h2 * h2);
    if (dp) {
        if (!(sum < -1e-6f)) return true;
        const float dgam = n + (3.f * mt.lambda + 2.f * mt.mu) / (2.f * mt.mu) * sum * mt.dp_alpha;
        return dgam > -1e-6f;
    }
    const float lim = state / (2.f * mt.mu);
    return n - lim > -(1e-4f * lim + 1e-6f);
}

// The particles G2P deferred (queued by sorted position): the AnisoMPM damage
// update on F_n with its full path (sigma^+, structural tensor) and the stress of
// (F_{n+1}, g_{n+1}) with every branch of the models, and its force bound.  Runs
// after k_g2p of the same substep, before the buffers swap.
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_rare(Dev<T, AccT> D)
{
    const int cnt = D.counters[CNT_RARE];
    const int64_t cap = D.cap;
    for (int base = blockIdx.x * WG; base < cnt; base += gridDim.x * WG) {
        float bnd[NSC] = { 0.f, 0.f, 0.f, 0.f, 0.f };
        const int i = base + threadIdx.x;
        int s = i < cnt ? ldAt(D.rare, i) : -1;
        if ((unsigned)s >= (unsigned)D.n) s = -1;
        int scene = -1;
        if (s >= 0) {
            int p = ldAt(D.perm, s);
            if ((unsigned)p >= (unsigned)D.n) p = 0;
            const int pm = ldAt(D.qmat, s);
            scene = sceneOf(pm);
            const SceneDev<T>& sd = D.sc[scene];
            const MatDev<T>& mt = D.mats[matOf(pm)];
            const bool dmg = D.any_damage && mt.has_damage;
            T d = ldAt(D.P + PD * cap, p), g = ldAt(D.P + PG * cap, p);
            if (dmg && !D.implicit && phaseFieldModel(mt.model)) {
                const Mat3<T> F = loadM(D.P, cap, PF, p); // F_n (FP32 form: H)
                damageUpdate<T, false>(mt, F, ldAt(D.Q + PLAP * cap, s), sd.dt, d, g);
                stAt(D.Q + PD * cap, s, d);
                stAt(D.Q + PG * cap, s, g);
            }
            Mat3<T> Fn = loadM(D.Q, cap, PF, s);
            T ys = T(0);
            if (mt.plasticity) {
                ys = ldAt(D.Q + PYS * cap, s);
                const bool changed = mt.plasticity == ZMPM_PLASTICITY_VON_MISES_STVK_HENCKY ? vonMisesProject(mt, Fn, ys)
                    : mt.plasticity == ZMPM_PLASTICITY_DRUCKER_PRAGER_STVK_HENCKY             ? druckerPragerProject(mt, Fn, ys)
                                                                                             : snowProject(mt, Fn, ys);
                if (changed) {
#pragma unroll
                    for (int k = 0; k < 9; k++) stAt(D.Q + (PF + k) * cap, s, Fn.a[k / 3][k % 3]);
                    stAt(D.Q + PYS * cap, s, ys);
                }
            }
            T S[6];
            if (D.fj_solve_j && mt.fj)
                fjSolveJStress(mt, false, ldAt(D.Q + PVOL * cap, s), ldAt(D.Q + PJ * cap, s), S);
            else
                vtauSym<false>(mt, Fn, ldAt(D.Q + PVOL * cap, s) * snowScale(mt, ys), g, S);
#pragma unroll
            for (int k = 0; k < 6; k++) stAt(D.Q + (PS + k) * cap, s, S[k]);
            if constexpr (std::is_same<AccT, long long>::value) forceBounds(bnd, S, dmg ? d : T(0), sd.odx, D.any_damage != 0);
        }
        if constexpr (std::is_same<AccT, long long>::value) writeBounds(D.maxbits, bnd, scene);
    }
}

// Damping pass G2P, gridToParticles(0) (MpmSimulationBase.cpp:328, 344) with
// apic_rpic_ratio = 0 (APIC_blend_RPIC: v = sum w v_i, C the skew part of
// B D^-1) or flip_pic_ratio = 0 (FLIP_blend_PIC: v = sum w v_i); x and F are
// unchanged (dt = 0).  The grid is read through this substep's binning (each
// particle's sorted cell) with weights from its moved position, as the
// reference's iterateKernel does with particle_base_offset; v and C are
// rewritten in place (the particles are in sorted order after G2P).  Float
// form: the momentum bounds of the new v, C join the bound slots.
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_damp_g2p(Dev<T, AccT> D)
{
    using TL = Tile<T>;
    constexpr int NB = TL::NB;
    __shared__ V4<T> gt[1][TL::TN];
    __shared__ int seg_end;
    const int lo = xcdLocal((int)blockIdx.x, (int)gridDim.x) * WG;
    const int hi = min(lo + WG, D.n);
    const int t = threadIdx.x, s = lo + t;
    const bool live = s < hi;
    const int64_t cap = D.cap;
    T* P = D.P;
    PosT<T> x[3] = { 0, 0, 0 };
    T m = T(0);
    int cell = 0, scene = 0;
    if (live) {
#pragma unroll
        for (int a = 0; a < 3; a++) x[a] = loadPos(P + (PX + a) * cap, s);
        m = ldAt(P + PM * cap, s);
        cell = ldAt(D.sorted_cell, s) & (NB - 1);
        scene = sceneOf(ldAt(D.pmat, s));
    }
    const SceneDev<T>& sd = D.sc[scene];
    G2PAcc<T> A;
#pragma unroll
    for (int a = 0; a < 3; a++) A.picV[a] = A.oldV[a] = T(0);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        A.Bp[i / 3][i % 3] = T(0);
        A.gv[i / 3][i % 3] = T(0);
    }
    A.lap = T(0);
    int seg_lo = lo;
    while (seg_lo < hi) {
        const int b = D.sorted_cell[seg_lo] / NB;
        __syncthreads(); // previous segment's gathers done
        if (t == 0) {
            const int c = b * NB + NB - 1;
            seg_end = D.cell_start[c] + D.cell_count[c];
        }
        stageTile(D, b, gt, 1);
        __syncthreads();
        const int seg_hi = max(seg_lo + 1, min(hi, seg_end));
        if (live && s >= seg_lo && s < seg_hi) g2pGather<T, AccT, false>(D, D.sc[blockScene(D, b)], gt, x, false, A, cell);
        seg_lo = seg_hi;
    }
    float bnd[NSC] = { 0.f, 0.f, 0.f, 0.f, 0.f };
    if (live) {
        T Cn[9];
        const T Dinv = sd.Dinv;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                Cn[3 * a + c] = D.flip ? T(0) : T(0.5) * (A.Bp[a][c] * Dinv) + T(-0.5) * (A.Bp[c][a] * Dinv);
#pragma unroll
        for (int a = 0; a < 3; a++) stAt(P + (PV + a) * cap, s, A.picV[a]);
#pragma unroll
        for (int i = 0; i < 9; i++) stAt(P + (PC + i) * cap, s, Cn[i]);
        if constexpr (std::is_same<AccT, long long>::value) momentumBounds(bnd, m, A.picV, Cn, sd.dx);
    }
    if constexpr (std::is_same<AccT, long long>::value) writeBounds(D.maxbits, bnd, live ? scene : -1);
}

// evalMaxParticleSpeed (MpmSimulationBase.cpp:1613-1638): per workgroup the
// maxima of |v| (in the particle precision, as the reference's velocity.norm()),
// x and -x over a grid stride of the particles; the last launch folds the
// EXTENT_WGS partials into row EXTENT_WGS.  Empty: speed 0, corners at the
// reference's initial float lowest.
constexpr int EXTENT_WGS = 256;

__device__ __forceinline__ double posMetres(int X, double dx, int pb) { return (double)X * dx * (1.0 / (double)(1 << pb)); }
__device__ __forceinline__ double posMetres(double x, double, int) { return x; }

__device__ __forceinline__ double blockMax(double v)
{
    __shared__ double wm[WG / 64];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = wm[0];
#pragma unroll
    for (int w = 1; w < WG / 64; w++) m = fmax(m, wm[w]);
    return m;
}

template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_extent(Dev<T, AccT> D, int fold, int rows, double* out, double max_dt,
                                                double cfl_dx)
{
    const double lowest = -3.4028234663852886e38; // std::numeric_limits<float>::lowest()
    double r[8] = { 0.0, lowest, lowest, lowest, lowest, lowest, lowest, 0.0 };
    if (fold) {
        for (int w = threadIdx.x; w < rows; w += WG)
#pragma unroll
            for (int k = 0; k < 8; k++) r[k] = fmax(r[k], D.extent[w * 8 + k]);
    }
    else {
        const int64_t cap = D.cap;
        for (int p = blockIdx.x * WG + threadIdx.x; p < D.n; p += gridDim.x * WG) {
            const SceneDev<T>& sd = D.sc[sceneOf(ldAt(D.pmat, p))];
            const T v0 = ldAt(D.P + (PV + 0) * cap, p), v1 = ldAt(D.P + (PV + 1) * cap, p), v2 = ldAt(D.P + (PV + 2) * cap, p);
            const double speed = (double)dsqrt(v0 * v0 + v1 * v1 + v2 * v2);
            r[0] = fmax(r[0], speed);
            r[7] = fmax(r[7], speed * sd.inv_cfl_dx);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                // in the scene's own frame
                const PosT<T> X = loadPos(D.P + (PX + a) * cap, p);
                const double x = posMetres(X, sd.ddx, sd.pos_bits);
                r[1 + a] = fmax(r[1 + a], x);
                r[4 + a] = fmax(r[4 + a], -x);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) r[k] = blockMax(r[k]);
    if (threadIdx.x == 0) {
        const int row = fold ? rows : blockIdx.x;
#pragma unroll
        for (int k = 0; k < 8; k++) D.extent[row * 8 + k] = r[k];
        // the folded result as calculateDt returns it: dt, speed, min corner, max
        // corner (one scene: cfl dx / speed; a batch: the smallest scene step)
        if (fold && out) {
            if (D.nscenes == 1)
                out[0] = r[0] != 0.0 ? cfl_dx / r[0] : max_dt;
            else
                out[0] = r[7] != 0.0 ? 1.0 / r[7] : max_dt;
            out[1] = r[0];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                out[2 + a] = -r[4 + a];
                out[5 + a] = r[1 + a];
            }
        }
    }
}

// ------------------------------------------------- implicit damage solve
// AnisoFractureSimulation::solveDamageSystem (AnisoFractureSimulation.h:435-612)
// between the grid update and G2P:
//   k_dtilde  per damaged particle (block by block, sorted order): vol = V0 det F_n,
//             maxDTilde = max(D~(F_n), maxDTilde) (updateDTilde :284-337), the
//             solve's per-particle weights, and the scatter of sum vol w,
//             sum vol dt D~ / eta w (rhs :457-474) and the lumped diagonal
//             sum vol (1 + dt / eta (D~ + 1)) w (multiply / preconditioner
//             :476-544) through an LDS tile of the block and its halo
//   k_cg      ConjugateGradient::solve (ConjugateGradient.h:52-103) from x = 0,
//             Jacobi preconditioner, tolerance 1e-6, 100 iterations, as one
//             persistent launch: all workgroups resident, grid barriers between
//             the node phases and the block phase, every workgroup folding the
//             same per-workgroup partial sums in the same order (so the stop
//             decision is the same everywhere); then pf <- x - pf (:603-605),
//             which G2P gathers for damageG2P (:614-680)
// The stiffness product K s = sum_p grad w_p (vol dt / eta l0^2 sum_j s_j grad w_jp)
// runs per block: s of the block and its halo staged in LDS, one thread per
// particle gathers, then scatters into an LDS tile flushed with one HBM atomic
// per node value.  s . A s is summed as sum_i diag_i s_i^2 + sum_p kappa_p |G_p|^2
// (G_p = sum_j s_j grad w_jp), equal to the reference's dot product of s with
// the assembled q, so alpha needs no pass over q.  The solve runs in double in
// both forms; its HBM atomics make the last bits of x depend on scheduling.
constexpr int CG_MAX_WGS = 2048;
constexpr int CGW = 256;                 // k_cg workgroup size (one workgroup per CU)
constexpr int CG_ITERATIONS = 100;       // ConjugateGradient<..>(100) (AnisoFractureSimulation.h:70)
constexpr double CG_TOLERANCE = 1e-6;    // cg.setTolerance(1e-6) (AnisoFractureSimulation.h:600-601)
constexpr int CG_SPIN_LIMIT = 1 << 18;   // barrier polls before the wait is given up (ST_BARRIER)

// Barrier over every workgroup of a launch whose workgroups are all resident.
// arrivals counts up through the launch (cleared before it); barrier number k
// is passed when it reaches k * nwg.  Release before arriving and acquire after
// leaving, at agent scope, make each workgroup's writes visible to the others
// across the chip's L2 slices (one per XCD).  The fences are cache-wide
// operations, so one wave per workgroup issues them: with every wave fencing,
// a barrier of 512 workgroups took ~52 us on MI355X, ~22 us with one wave per
// workgroup, ~9 us with 256 workgroups.  A wait that exceeds the poll limit latches
// ST_BARRIER and goes on, and once that is latched no barrier waits any more,
// so the launch always ends soon (with its results flagged as invalid).
__device__ __forceinline__ void gridBarrier(unsigned* arrivals, unsigned nwg, unsigned& epoch, int* counters)
{
    // every wave's stores complete (in L2) before the workgroup barrier; then
    // one wave writes the L2 back (agent-scope release) for the whole
    // workgroup and, after the wait, invalidates it and the CU's vector cache
    // (agent-scope acquire), which the other waves share
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    epoch++;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned target = epoch * nwg;
        unsigned seen = __hip_atomic_fetch_add(arrivals, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        int polls = 0;
        while (seen < target) {
            if (++polls > CG_SPIN_LIMIT
                || (__hip_atomic_load(&counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ST_BARRIER)) {
                latch(counters, ST_BARRIER);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            seen = __hip_atomic_load(arrivals, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Sum over a workgroup of NT threads, the same value in every thread (fixed order).
template <int NT>
__device__ __forceinline__ double blockSum(double v, double* red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) r += red[w];
    return r;
}

// Sum of the nwg per-workgroup partials part[0 .. nwg), identical in every
// workgroup (each folds them in the same order).
template <int NT>
__device__ __forceinline__ double foldPartials(const double* part, int nwg, double* red)
{
    double v = 0.0;
    for (int w = threadIdx.x; w < nwg; w += NT) v += part[w];
    return blockSum<NT>(v, red);
}

// Flush an LDS tile of block b (block + forward halo) into a node array with
// one HBM atomic per non-zero value.
template <class T, class AccT, int NT = WG>
__device__ __forceinline__ void flushTile(const Dev<T, AccT>& D, int b, const double* tile, double* dst, int& ovf)
{
    using B = Blk<T>;
    using TL = Tile<T>;
    const int* nb = D.nbr + b * 8;
    for (int nd = threadIdx.x; nd < TL::TN; nd += NT) {
        const double v = tile[nd];
        if (v == 0.0) continue;
        const int tx = nd / (TL::TY * TL::TZ), ty = (nd / TL::TZ) % TL::TY, tz = nd % TL::TZ;
        const int tb = nb[((tx >= TL::BX) << 2) | ((ty >= TL::BY) << 1) | (tz >= TL::BZ)];
        if (tb < 0) {
            ovf = 1;
            continue;
        }
        const int elem = ((tx & (TL::BX - 1)) << (B::ybits + B::zbits)) | ((ty & (TL::BY - 1)) << B::zbits) | (tz & (TL::BZ - 1));
        unsafeAtomicAdd(dst + (int64_t)tb * TL::NB + elem, v);
    }
}

// Cells of block b holding sorted positions in [lo, hi): their indices into cl[],
// their count into *ncl (one wavefront; NB <= 64).
template <int NB>
__device__ __forceinline__ void activeCells(const int* cs, const int* ce, int lo, int hi, int* cl, int* ncl)
{
    const int t = threadIdx.x;
    if (t < 64) {
        bool act = false;
        if (t < NB) act = min(ce[t], hi) > max(cs[t], lo);
        const uint64_t m = __ballot(act);
        if (act) cl[__popcll(m & ((1ull << t) - 1))] = t;
        if (t == 0) *ncl = __popcll(m);
    }
}

template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_dtilde(Dev<T, AccT> D)
{
    // the implicit solve runs on one-scene handles only
    const SceneDev<T>& sd = D.sc[0];
    const T dt = sd.dt;
    using B = Blk<T>;
    using TL = Tile<T>;
    constexpr int NB = TL::NB;
    __shared__ double acc[3][TL::TN];
    __shared__ double pw[9][WG]; // per particle of the chunk: w_a(i) at [3 a + i]
    __shared__ double pc[3][WG]; // vol, vol dt D~ / eta, vol (1 + dt / eta (D~ + 1)); 0 off the solve
    __shared__ int cs[NB], ce[NB], cl[NB];
    __shared__ int ncl;
    const int t = threadIdx.x;
    if (blockIdx.x == 0 && t == 0) { // k_cg's barrier arrivals and block tickets
        D.cg_bar[0] = 0u;
        D.cg_bar[2] = 0u;
    }
    const int64_t cap = D.cap, cn = D.cap_nodes;
    const int nblocks = D.counters[CNT_BLOCKS];
    int ovf = 0;
    for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const int bs = D.block_start[b], cnt = D.block_count[b];
        if (cnt <= 0) continue;
        __syncthreads(); // previous block flushed
        for (int e = t; e < 3 * TL::TN; e += WG) (&acc[0][0])[e] = 0.0;
        if (t < NB) {
            cs[t] = D.cell_start[b * NB + t];
            ce[t] = cs[t] + D.cell_count[b * NB + t];
        }
        for (int lo = bs; lo < bs + cnt; lo += WG) {
            const int hi = min(lo + WG, bs + cnt);
            __syncthreads(); // previous chunk's columns done with pw, pc
            // ---- one thread per particle: D~ history, the solve's per-particle data
            const int s = lo + t;
            if (s < hi) {
                int p = ldAt(D.perm, s);
                if ((unsigned)p >= (unsigned)D.n) p = 0;
                const MatDev<T>& mt = D.mats[matOf(ldAt(D.pmat, p))];
                T hist = ldAt(D.P + PDT * cap, p);
                T kap = T(0);
                T d0[3] = { T(0), T(0), T(0) };
                double c3[3] = { 0.0, 0.0, 0.0 };
                if (mt.has_damage && phaseFieldModel(mt.model)) {
                    const Mat3<T> Fs = loadM(D.P, cap, PF, p);
                    T J;
                    if constexpr (sizeof(T) == 4)
                        J = 1.f + jm1OfH(Fs);
                    else
                        J = det3(Fs);
                    const T vol = ldAt(D.P + PVOL * cap, p) * J; // phase_field.vol (:425-432)
                    T dTilde;
                    dTildeOf<T, false>(mt, Fs, dTilde);
                    hist = dTilde < hist ? hist : dTilde; // std::max(dTilde, dTilde_history)
                    const T dte = dt / mt.eta;
                    kap = vol * dte * mt.l0 * mt.l0;
                    int base;
#pragma unroll
                    for (int a = 0; a < 3; a++)
                        cellOf(loadPos(D.P + (PX + a) * cap, p), sd.odx, sd.origin[a], sd.domain, base, d0[a], sd.pos_bits);
                    c3[0] = (double)vol;                                             // sum vol w
                    c3[1] = (double)((dt * hist) / mt.eta) * (double)vol;            // D~ part of rhs
                    c3[2] = (double)vol * (1.0 + (double)dte * ((double)hist + 1.0)); // diagonal
                }
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    double w[3], dw[3];
                    bsplineFrac<double>((double)d0[a], w, dw);
#pragma unroll
                    for (int i = 0; i < 3; i++) pw[3 * a + i][t] = w[i];
                }
#pragma unroll
                for (int f = 0; f < 3; f++) pc[f][t] = c3[f];
                stAt(D.Q + PDT * cap, s, hist);
#pragma unroll
                for (int a = 0; a < 3; a++) stAt(D.cgp + a * cap, s, d0[a]);
                stAt(D.cgp + 3 * cap, s, kap);
            }
            activeCells<NB>(cs, ce, lo, hi, cl, &ncl);
            __syncthreads();
            // ---- lanes = (cell, i, j) stencil columns: each walks its cell's
            // particles (one base node) and adds 3 nodes x 3 sums into the tile
            const int per = ncl * 9;
            for (int slot = t; slot < per; slot += WG) {
                const int cell = cl[slot / 9], col = slot - 9 * (slot / 9);
                const int ci = col / 3, cj = col - 3 * (col / 3);
                const int p0 = max(cs[cell], lo) - lo, p1 = min(ce[cell], hi) - lo;
                double a[3][3] = { { 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0 } };
                for (int pp = p0; pp < p1; pp++) {
                    const double wxy = pw[ci][pp] * pw[3 + cj][pp];
#pragma unroll
                    for (int f = 0; f < 3; f++) {
                        const double cf = pc[f][pp] * wxy;
#pragma unroll
                        for (int k = 0; k < 3; k++) a[f][k] += cf * pw[6 + k][pp];
                    }
                }
                const int cx = cell >> (B::ybits + B::zbits), cy = (cell >> B::zbits) & (TL::BY - 1), cz = cell & (TL::BZ - 1);
                const int nbase = ((cx + ci) * TL::TY + (cy + cj)) * TL::TZ + cz;
#pragma unroll
                for (int f = 0; f < 3; f++)
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        if (a[f][k] != 0.0) unsafeAtomicAdd(&acc[f][nbase + k], a[f][k]);
            }
        }
        __syncthreads();
        flushTile(D, b, acc[0], D.cgv, ovf);          // x slot: sum vol w
        flushTile(D, b, acc[1], D.cgv + cn, ovf);     // r slot: sum vol dt D~ / eta w
        flushTile(D, b, acc[2], D.cgv + 4 * cn, ovf); // diagonal
    }
    if (ovf) latch(D.counters, ST_CAPACITY);
}

template <class T, class AccT>
__global__ __launch_bounds__(CGW) void k_cg(Dev<T, AccT> D)
{
    using B = Blk<T>;
    using TL = Tile<T>;
    constexpr int NB = TL::NB;
    __shared__ double st[TL::TN], qt[TL::TN];
    __shared__ double pw[18][CGW]; // per particle of the segment: w_a(i) at [3 a + i], dw_a(i) at [9 + 3 a + i]
    __shared__ double pg[3][CGW];  // kappa G odx
    __shared__ int cs[NB], ce[NB], cl[NB];
    __shared__ int ncl, ticket;
    __shared__ double red[CGW / 64];
    const int t = threadIdx.x;
    const int nwg = (int)gridDim.x;
    const int nblocks = D.counters[CNT_BLOCKS];
    const int n = D.n;
    const int64_t cap = D.cap, cn = D.cap_nodes, nn = (int64_t)nblocks * NB;
    double* const X = D.cgv;
    double* const R = D.cgv + cn;
    double* const S = D.cgv + 2 * cn;
    double* const Qv = D.cgv + 3 * cn;
    const double* const DG = D.cgv + 4 * cn;
    double* const part = D.cg_part;
    const int64_t first = (int64_t)blockIdx.x * CGW + t, stride = (int64_t)nwg * CGW;
    const double odx = (double)D.sc[0].odx; // the implicit solve runs on one-scene handles only
    unsigned epoch = 0, tickets = 0;
    int ovf = 0;
    // r = b - A 0 = b with b_i = sum vol dt D~ / eta w + pf_i sum vol w; x = 0
    double prr = 0.0, prho = 0.0;
    for (int64_t node = first; node < nn; node += stride) {
        const double dg = DG[node];
        const double r = dg > 0.0 ? R[node] + (double)D.newv[4 * node + 3] * X[node] : 0.0;
        X[node] = 0.0;
        R[node] = r;
        const double mr = dg > 0.0 ? r / dg : 0.0;
        prr += r * r;
        prho += mr * r;
    }
    prr = blockSum<CGW>(prr, red);
    prho = blockSum<CGW>(prho, red);
    if (t == 0) {
        part[blockIdx.x] = prr;
        part[CG_MAX_WGS + blockIdx.x] = prho;
    }
    gridBarrier(D.cg_bar, (unsigned)nwg, epoch, D.counters);
    double rr = foldPartials<CGW>(part, nwg, red);
    double rho = foldPartials<CGW>(part + CG_MAX_WGS, nwg, red);
    const double rel = sqrt(rr) * D.cg_rel;
    double rho_old = 0.0;
    int it = 0;
    for (;; it++) {
        // search direction s = M^-1 r (+ rho / rho_old s) and q = diag s
        const double beta = it == 0 ? 0.0 : rho / rho_old;
        double psq = 0.0;
        for (int64_t node = first; node < nn; node += stride) {
            const double dg = DG[node];
            const double mr = dg > 0.0 ? R[node] / dg : 0.0;
            const double sv = it == 0 ? mr : beta * S[node] + mr;
            S[node] = sv;
            Qv[node] = dg * sv;
            psq += dg * sv * sv;
        }
        gridBarrier(D.cg_bar, (unsigned)nwg, epoch, D.counters);
        // q += K s, one grid block at a time, blocks handed out by a ticket
        // counter (the block sizes vary): s of the block and its halo is staged
        // in LDS; one thread per particle gathers G = sum_j s_j grad w_j and
        // keeps kappa G with its weights in LDS; then lanes = (cell, i, j)
        // stencil columns walk their cell's particles and add kappa G . grad w_i
        // for 3 nodes into a tile flushed with one HBM atomic per node value.
        // Every workgroup draws until it gets a ticket past the last block, so
        // each pass uses exactly nblocks + nwg tickets.
        for (;;) {
            __syncthreads(); // previous block flushed, ticket slot free
            if (t == 0) ticket = (int)(__hip_atomic_fetch_add(&D.cg_bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - tickets);
            __syncthreads();
            const int b = ticket;
            if (b >= nblocks) break;
            const int bs = D.block_start[b], cnt = D.block_count[b];
            if (cnt <= 0) continue;
            const int* nb = D.nbr + b * 8;
            for (int nd = t; nd < TL::TN; nd += CGW) {
                const int tx = nd / (TL::TY * TL::TZ), ty = (nd / TL::TZ) % TL::TY, tz = nd % TL::TZ;
                const int tb = nb[((tx >= TL::BX) << 2) | ((ty >= TL::BY) << 1) | (tz >= TL::BZ)];
                double v = 0.0;
                if (tb >= 0) {
                    const int elem = ((tx & (TL::BX - 1)) << (B::ybits + B::zbits)) | ((ty & (TL::BY - 1)) << B::zbits)
                        | (tz & (TL::BZ - 1));
                    v = S[(int64_t)tb * NB + elem];
                }
                st[nd] = v;
                qt[nd] = 0.0;
            }
            if (t < NB) {
                cs[t] = D.cell_start[b * NB + t];
                ce[t] = cs[t] + D.cell_count[b * NB + t];
            }
            __syncthreads();
            for (int lo = bs; lo < bs + cnt; lo += CGW) {
                const int hi = min(lo + CGW, bs + cnt);
                if (lo != bs) __syncthreads(); // previous chunk's columns done with pw, pg
                const int s = lo + t;
                if (s < hi) {
                    const double kap = (double)ldAt(D.cgp + 3 * cap, s);
                    double w[3][3], dw[3][3];
#pragma unroll
                    for (int a = 0; a < 3; a++) bsplineFrac<double>((double)ldAt(D.cgp + a * cap, s), w[a], dw[a]);
                    double G0 = 0.0, G1 = 0.0, G2 = 0.0;
                    if (kap != 0.0) {
                        const int e = ldAt(D.sorted_cell, s) - b * NB;
                        const int lx = e >> (B::ybits + B::zbits), ly = (e >> B::zbits) & (TL::BY - 1), lz = e & (TL::BZ - 1);
#pragma unroll
                        for (int i = 0; i < 3; i++) {
                            double ty = 0.0, tdy = 0.0, tz = 0.0;
#pragma unroll
                            for (int j = 0; j < 3; j++) {
                                const int nrow = ((lx + i) * TL::TY + (ly + j)) * TL::TZ + lz;
                                double a0 = 0.0, a1 = 0.0;
#pragma unroll
                                for (int k = 0; k < 3; k++) {
                                    const double sv = st[nrow + k];
                                    a0 += w[2][k] * sv;
                                    a1 += dw[2][k] * sv;
                                }
                                ty += w[1][j] * a0;
                                tdy += dw[1][j] * a0;
                                tz += w[1][j] * a1;
                            }
                            G0 += dw[0][i] * ty;
                            G1 += w[0][i] * tdy;
                            G2 += w[0][i] * tz;
                        }
                        G0 *= odx;
                        G1 *= odx;
                        G2 *= odx;
                        psq += kap * (G0 * G0 + G1 * G1 + G2 * G2);
                    }
                    // kappa G . grad w_ijk = kappa odx (G0 dw_i w_j w_k + G1 w_i dw_j w_k + G2 w_i w_j dw_k)
                    pg[0][t] = kap * G0 * odx;
                    pg[1][t] = kap * G1 * odx;
                    pg[2][t] = kap * G2 * odx;
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int i = 0; i < 3; i++) {
                            pw[3 * a + i][t] = w[a][i];
                            pw[9 + 3 * a + i][t] = dw[a][i];
                        }
                }
                activeCells<NB>(cs, ce, lo, hi, cl, &ncl);
                __syncthreads();
                const int per = ncl * 9;
                for (int slot = t; slot < per; slot += CGW) {
                    const int cell = cl[slot / 9], col = slot - 9 * (slot / 9);
                    const int ci = col / 3, cj = col - 3 * (col / 3);
                    const int p0 = max(cs[cell], lo) - lo, p1 = min(ce[cell], hi) - lo;
                    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                    for (int pp = p0; pp < p1; pp++) {
                        const double wx = pw[ci][pp], dwx = pw[9 + ci][pp], wy = pw[3 + cj][pp], dwy = pw[12 + cj][pp];
                        const double cxy = pg[0][pp] * dwx * wy + pg[1][pp] * wx * dwy;
                        const double cz = pg[2][pp] * wx * wy;
                        a0 += cxy * pw[6][pp] + cz * pw[15][pp];
                        a1 += cxy * pw[7][pp] + cz * pw[16][pp];
                        a2 += cxy * pw[8][pp] + cz * pw[17][pp];
                    }
                    const int cx = cell >> (B::ybits + B::zbits), cy = (cell >> B::zbits) & (TL::BY - 1), cz = cell & (TL::BZ - 1);
                    const int nbase = ((cx + ci) * TL::TY + (cy + cj)) * TL::TZ + cz;
                    if (a0 != 0.0) unsafeAtomicAdd(&qt[nbase], a0);
                    if (a1 != 0.0) unsafeAtomicAdd(&qt[nbase + 1], a1);
                    if (a2 != 0.0) unsafeAtomicAdd(&qt[nbase + 2], a2);
                }
            }
            __syncthreads();
            flushTile<T, AccT, CGW>(D, b, qt, Qv, ovf);
        }
        tickets += (unsigned)(nblocks + nwg);
        psq = blockSum<CGW>(psq, red);
        if (t == 0) part[2 * CG_MAX_WGS + blockIdx.x] = psq;
        gridBarrier(D.cg_bar, (unsigned)nwg, epoch, D.counters);
        const double sq = foldPartials<CGW>(part + 2 * CG_MAX_WGS, nwg, red);
        const double alpha = sq != 0.0 ? rho / sq : 1.7976931348623157e308;
        // x += alpha s, r -= alpha q, and the next M^-1 r . r
        prr = 0.0;
        prho = 0.0;
        for (int64_t node = first; node < nn; node += stride) {
            const double dg = DG[node];
            if (!(dg > 0.0)) continue;
            X[node] = alpha * S[node] + X[node];
            const double r = -alpha * Qv[node] + R[node];
            R[node] = r;
            prr += r * r;
            prho += (r / dg) * r;
        }
        prr = blockSum<CGW>(prr, red);
        prho = blockSum<CGW>(prho, red);
        if (t == 0) {
            part[blockIdx.x] = prr;
            part[CG_MAX_WGS + blockIdx.x] = prho;
        }
        gridBarrier(D.cg_bar, (unsigned)nwg, epoch, D.counters);
        rr = foldPartials<CGW>(part, nwg, red);
        const double rho_new = foldPartials<CGW>(part + CG_MAX_WGS, nwg, red);
        const double conv = sqrt(rr);
        if (conv < CG_TOLERANCE || conv < rel || it == CG_ITERATIONS) break;
        rho_old = rho;
        rho = rho_new;
    }
    // g.phase_field = x - phase_field: the change G2P transfers back
    for (int64_t node = first; node < nn; node += stride) {
        const double dg = DG[node];
        D.newv[4 * node + 3] = dg > 0.0 ? (T)(X[node] - (double)D.newv[4 * node + 3]) : T(0);
    }
    if (blockIdx.x == 0 && t == 0) D.cg_bar[1] = (unsigned)it;
    if (ovf) latch(D.counters, ST_CAPACITY);
}

} // namespace zmpm
