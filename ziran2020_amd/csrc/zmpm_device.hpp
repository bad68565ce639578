// zmpm_device.hpp -- gfx950 kernels of the explicit
// APIC MPM substep with AnisoMPM damage.
//
// One substep (Ziran MpmSimulationBase::advanceOneTimeStep,
// Lib/MPM/MpmSimulationBase.cpp:262-357) is this launch sequence on the
// handle's stream:
//   k_begin_step     device clock, fixed-point scales (float form)
//   k_bin            base node, SPGrid page/element key, mark touched pages
//   scan             dense block ids in page (Morton) order -> page table
//   k_prepare        zero block accumulators, forward-neighbour block table
//   k_count          per-cell occupancy (atomic slot)
//   scan             cell start offsets
//   k_scatter        counting-sort scatter
//   k_cellsort       order inside each cell by original index -> reference order
//   k_p2g            APIC scatter of m, mv, f (stress read from the particle)
//                    and of the damage phase field
//   k_grid           v = mv/m, gravity + force, collision objects, phase-field
//                    normalisation
//   k_g2p            APIC gather, x update, CFL flag, damage Laplacian and
//                    update, F update, stress of the new state (SVD / QR) and
//                    the next P2G's fixed-point bounds; writes particles in
//                    sorted order
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
constexpr int SCAN_TILE = 2048; // 256 threads x 8 elements
constexpr float FIXED_TARGET_EXP = 40.0f;
// Per-lane bound on the summed |fixed-point values| one P2G lane adds to its tile:
// by construction a lane adds at most CHUNK particles x 27 node-quantities of
// |q| < 2^40 (< 2^52); 2^56 leaves int64 tile sums far from overflow.
constexpr float FIXED_LANE_LIMIT = 72057594037927936.0f; // 2^56

__device__ __forceinline__ void latch(int* counters, int bits) { atomicOr(&counters[1], bits); }

// Every third bit: b9..b0 -> bit 3*i (block coordinates < 1024).
__device__ __forceinline__ uint32_t part1by2(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t compact1by2(uint32_t x)
{
    x &= 0x09249249u;
    x = (x | (x >> 2)) & 0x030c30c3u;
    x = (x | (x >> 4)) & 0x0300f00fu;
    x = (x | (x >> 8)) & 0x030000ffu;
    x = (x | (x >> 16)) & 0x000003ffu;
    return x;
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

// ---------------------------------------------------------------- block scan
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

// exclusive scan over a 256-thread block; returns the block total through *total
__device__ __forceinline__ int blockExclusiveScan256(int v, int* total)
{
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = waveInclusiveScan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// length of a scan: device count * mul, or the host length
__device__ __forceinline__ int64_t scanLength(const int* len_dev, int64_t mul, int64_t len_host)
{
    return len_dev ? (int64_t)(*len_dev) * mul : len_host;
}

__global__ __launch_bounds__(WG) void k_scan_reduce(const int* __restrict__ in, const int* len_dev, int64_t mul,
                                                   int64_t len_host, int* __restrict__ tile_sums)
{
    const int64_t len = scanLength(len_dev, mul, len_host);
    const int64_t ntiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int s = 0;
        const int64_t base = tile * SCAN_TILE + (int64_t)threadIdx.x * 8;
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (base + k < len) s += in[base + k];
        int tot;
        blockExclusiveScan256(s, &tot);
        if (threadIdx.x == 0) tile_sums[tile] = tot;
    }
}

// exclusive scan of the tile sums in one workgroup; total -> *total_out
// (clamped to cap if cap >= 0, with the capacity status latched)
__global__ __launch_bounds__(WG) void k_scan_top(int* __restrict__ tile_sums, const int* len_dev, int64_t mul,
                                                int64_t len_host, int* total_out, int cap, int* counters)
{
    const int64_t len = scanLength(len_dev, mul, len_host);
    const int64_t ntiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    int carry = 0;
    for (int64_t base = 0; base < ntiles; base += WG) {
        int64_t i = base + threadIdx.x;
        int v = (i < ntiles) ? tile_sums[i] : 0;
        int tot;
        int ex = blockExclusiveScan256(v, &tot);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) {
        int t = carry;
        if (cap >= 0 && t > cap) {
            t = cap;
            atomicOr(&counters[1], ST_CAPACITY);
        }
        *total_out = t;
    }
}

// Down-sweep.  ASSIGN: the input is the per-page "touched" flag and the
// epilogue writes the page table (dense block id in page order, or -1), the
// block -> page list, and clears the flag for the next substep.
template <bool ASSIGN>
__global__ __launch_bounds__(WG) void k_scan_down(int* __restrict__ in, const int* len_dev, int64_t mul,
                                                 int64_t len_host, const int* __restrict__ tile_off,
                                                 int* __restrict__ out, int* __restrict__ block_page, int cap)
{
    const int64_t len = scanLength(len_dev, mul, len_host);
    const int64_t ntiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t base = tile * SCAN_TILE + (int64_t)threadIdx.x * 8;
        int v[8], s = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[k] = (base + k < len) ? in[base + k] : 0;
            s += v[k];
        }
        int tot;
        int ex = blockExclusiveScan256(s, &tot) + tile_off[tile];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (base + k < len) {
                if (ASSIGN) {
                    int id = -1;
                    if (v[k]) {
                        in[base + k] = 0;
                        if (ex < cap) {
                            id = ex;
                            block_page[ex] = (int)(base + k);
                        }
                    }
                    out[base + k] = id;
                }
                else
                    out[base + k] = ex;
            }
            ex += v[k];
        }
    }
}

// ------------------------------------------------------------------ binning
// sortParticlesAndPolluteGrid key build (MpmSimulationBase.cpp:1545-1550) and
// MpmGrid::pollute page activation (MpmGrid.h:153-193).
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_bin(Dev<T, AccT> D)
{
    using B = Blk<T>;
    const int p = blockIdx.x * WG + threadIdx.x;
    const bool active = p < D.n;
    uint32_t page = 0xffffffffu;
    int bb[3] = { 0, 0, 0 };
    if (active) {
        int bits = 0;
        const T hi = T(D.domain - 4);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            T x = D.P[(PX + a) * D.cap + p];
            T t = sub_rn(mul_rn(x, D.odx), T(0.5));
            if (!isfinite(t)) {
                bits |= ST_NAN;
                t = T(0);
            }
            if (!(t >= T(0) && t < hi)) {
                bits |= ST_DOMAIN;
                t = (t < T(0)) ? T(0) : hi - T(1);
            }
            bb[a] = (int)t; // t >= 0: truncation is int_floor (MathTools.h:15-25)
        }
        if (bits) latch(D.counters, bits);
        uint32_t elem;
        pageElem<T>((uint32_t)bb[0], (uint32_t)bb[1], (uint32_t)bb[2], page, elem);
        D.key[p] = ((uint64_t)page << (B::xbits + B::ybits + B::zbits)) | elem;
    }
    // one activation per run of equal pages inside the wavefront
    uint32_t prev = (uint32_t)__shfl_up((int)page, 1);
    const int lane = threadIdx.x & 63;
    if (active && (lane == 0 || prev != page)) {
        uint32_t Bx = bb[0] >> B::xbits, By = bb[1] >> B::ybits, Bz = bb[2] >> B::zbits;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int c = 0; c < 2; c++) D.touched[pageOfBlock<T>(D.masks, Bx + a, By + b, Bz + c)] = 1;
    }
}

template <class T, class AccT>
__global__ __launch_bounds__(64) void k_prepare(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    const int nb = D.counters[0];
    const int nf = D.any_damage ? NGF : GPF;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int t = threadIdx.x;
        if (t < NB) {
            const int64_t node = (int64_t)b * NB + t;
            for (int f = 0; f < nf; f++) D.acc[f * D.cap_nodes + node] = AccT(0);
            D.cell_count[node] = 0;
        }
        if (t < 8) {
            uint32_t Bx, By, Bz;
            blockOfPage<T>((uint32_t)D.block_page[b], Bx, By, Bz);
            uint32_t nx = Bx + ((t >> 2) & 1), ny = By + ((t >> 1) & 1), nz = Bz + (t & 1);
            int id = -1;
            if ((nx << B::xbits) < (uint32_t)D.domain && (ny << B::ybits) < (uint32_t)D.domain
                && (nz << B::zbits) < (uint32_t)D.domain)
                id = D.page_table[pageOfBlock<T>(D.masks, nx, ny, nz)];
            D.nbr[b * 8 + t] = id;
        }
    }
}

// Per-cell occupancy and each particle's slot in its cell.  Live order is the
// previous substep's sorted order, so equal cells come in runs: one atomic per
// run inside a wavefront, slots in live order within the run.
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_count(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int bbits = B::xbits + B::ybits + B::zbits;
    const int p = blockIdx.x * WG + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int cell = -1 - lane; // distinct per inactive lane: never joins a run
    if (p < D.n) {
        const uint64_t key = D.key[p];
        const int bid = D.page_table[key >> bbits];
        if (bid >= 0)
            cell = bid * (1 << bbits) + (int)(key & ((1 << bbits) - 1));
        else
            latch(D.counters, ST_CAPACITY);
    }
    const int prev = __shfl_up(cell, 1);
    const bool head = (lane == 0) || (prev != cell);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = heads & (~0ull >> (63 - lane));      // heads at lanes <= this lane
    const int start = 63 - __clzll(upto);                       // this run's head lane
    const uint64_t after = (lane == 63) ? 0ull : (heads & (~0ull << (lane + 1)));
    const int next = after ? __ffsll((unsigned long long)after) - 1 : 64;
    int base = 0;
    if (head && cell >= 0) base = atomicAdd(&D.cell_count[cell], next - lane);
    base = __shfl(base, start);
    if (p < D.n) D.rank[p] = cell >= 0 ? base + (lane - start) : -1;
}

template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_scatter(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int bbits = B::xbits + B::ybits + B::zbits;
    const int p = blockIdx.x * WG + threadIdx.x;
    if (p >= D.n) return;
    int r = D.rank[p];
    if (r < 0) return;
    uint64_t key = D.key[p];
    int bid = D.page_table[key >> bbits];
    int64_t cell = (int64_t)bid * (1 << bbits) + (int)(key & ((1 << bbits) - 1));
    const int pos = D.cell_start[cell] + r;
    D.perm[pos] = p;
    D.spid[pos] = D.pid[p];
}

// Order inside a cell by original particle index: together with the page-ordered
// block ids and element-ordered cells this reproduces std::sort of the
// reference's unique 64-bit keys (MpmSimulationBase.cpp:1552).  A workgroup
// takes 256 consecutive cells; their particles are contiguous, so (index,
// live slot) pairs are staged in LDS with coalesced loads, each thread
// insertion-sorts its cell there, and the result and the cell ids go back
// coalesced.  Ranges larger than the LDS stage fall back to a per-cell shell
// sort in HBM.
constexpr int CELLSORT_CELLS = 256, CELLSORT_STAGE = 4096;
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_cellsort(Dev<T, AccT> D)
{
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    __shared__ int kk[CELLSORT_STAGE];
    __shared__ int vv[CELLSORT_STAGE];
    __shared__ int cc[CELLSORT_STAGE];
    const int t = threadIdx.x;
    const int64_t ncells = (int64_t)D.counters[0] * NB;
    for (int64_t c0 = (int64_t)blockIdx.x * CELLSORT_CELLS; c0 < ncells; c0 += (int64_t)gridDim.x * CELLSORT_CELLS) {
        const int64_t clast = min(c0 + CELLSORT_CELLS, ncells) - 1;
        const int S0 = D.cell_start[c0];
        const int S1 = D.cell_start[clast] + D.cell_count[clast];
        const int64_t c = c0 + t;
        const bool mine = c <= clast;
        const int cs = mine ? D.cell_start[c] : 0;
        const int cnt = mine ? D.cell_count[c] : 0;
        if (S1 - S0 <= CELLSORT_STAGE) {
            for (int i = t; i < S1 - S0; i += WG) {
                kk[i] = D.spid[S0 + i];
                vv[i] = D.perm[S0 + i];
            }
            __syncthreads();
            if (cnt > 1) {
                int* k = kk + (cs - S0);
                int* v = vv + (cs - S0);
                for (int i = 1; i < cnt; i++) {
                    const int kv = k[i], vv_i = v[i];
                    int j = i;
                    while (j > 0 && k[j - 1] > kv) {
                        k[j] = k[j - 1];
                        v[j] = v[j - 1];
                        j--;
                    }
                    k[j] = kv;
                    v[j] = vv_i;
                }
            }
            for (int i = 0; i < cnt; i++) cc[cs - S0 + i] = (int)c;
            __syncthreads();
            for (int i = t; i < S1 - S0; i += WG) {
                D.spid[S0 + i] = kk[i];
                D.perm[S0 + i] = vv[i];
                D.sorted_cell[S0 + i] = cc[i];
            }
            __syncthreads();
        }
        else {
            if (cnt > 0) {
                int* key = D.spid + cs;
                int* val = D.perm + cs;
                int gap = 1;
                while (gap < cnt / 3) gap = 3 * gap + 1;
                for (; gap >= 1; gap /= 3)
                    for (int i = gap; i < cnt; i++) {
                        const int kv = key[i], v = val[i];
                        int j = i;
                        while (j >= gap && key[j - gap] > kv) {
                            key[j] = key[j - gap];
                            val[j] = val[j - gap];
                            j -= gap;
                        }
                        key[j] = kv;
                        val[j] = v;
                    }
                for (int i = 0; i < cnt; i++) D.sorted_cell[cs + i] = (int)c;
            }
        }
    }
}

// ------------------------------------------------------------- P2G helpers
template <class T>
__device__ __forceinline__ Mat3<T> loadM(const T* P, int64_t cap, int f, int p)
{
    Mat3<T> M;
#pragma unroll
    for (int i = 0; i < 9; i++) M.a[i / 3][i % 3] = P[(f + i) * cap + p];
    return M;
}

// V0 * Kirchhoff stress of one particle, symmetric part (xx xy xz yy yz zz):
// FBasedMpmForceHelper::updateState (FBasedMpmForceHelper.cpp:41-63) with
// kirchhoff = P F^T (HyperelasticConstitutiveModel.h:75-80), or
// JBasedMpmForceHelper::updateState (JBasedMpmForceHelper.cpp:28-50).  tau is
// symmetric for every model on the path; its rounding-level skew is dropped.
template <class T>
__device__ __forceinline__ void vtauSym(const MatDev<T>& mt, const Mat3<T>& F, T V0, T g, T* S)
{
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
        else if (mt.model == ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN) {
            T J, t6[6];
            qrStableTau(F, mt.mu, mt.lambda, g, mt.kx, mt.ky, t6, &J);
            S[0] = V0 * t6[0];
            S[1] = V0 * t6[1];
            S[2] = V0 * t6[2];
            S[3] = V0 * t6[3];
            S[4] = V0 * t6[4];
            S[5] = V0 * t6[5];
            return;
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
}

__device__ __forceinline__ float waveMax(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d));
    return v;
}

// Upper bounds of one particle's P2G contributions per scale group, folded
// into the running per-thread maxima b: mass w m; momentum w (m v + m C (x_i - x_p))
// with |x_i - x_p| <= 1.5 dx; force tau grad(w) with |grad w| <= 1/dx per axis;
// phase field w d and w (w <= 1).
template <class T>
__device__ __forceinline__ void particleBounds(float* b, T m, const T* v, const T* Cm, const T* S, T d, T dx, T odx,
                                               bool dmg)
{
    b[SC_M] = fmaxf(b[SC_M], fabsf((float)m));
    float mv = 0.f;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        float cr = fabsf((float)Cm[3 * r]) + fabsf((float)Cm[3 * r + 1]) + fabsf((float)Cm[3 * r + 2]);
        mv = fmaxf(mv, fabsf((float)m) * (fabsf((float)v[r]) + 1.5f * (float)dx * cr));
    }
    const float a0 = fabsf((float)S[0]), a1 = fabsf((float)S[1]), a2 = fabsf((float)S[2]);
    const float a3 = fabsf((float)S[3]), a4 = fabsf((float)S[4]), a5 = fabsf((float)S[5]);
    const float f = (float)odx * fmaxf(a0 + a1 + a2, fmaxf(a1 + a3 + a4, a2 + a4 + a5));
    b[SC_MV] = fmaxf(b[SC_MV], mv * 1.0001f);
    b[SC_F] = fmaxf(b[SC_F], f * 1.0001f);
    if (dmg) {
        b[SC_PF] = fmaxf(b[SC_PF], fabsf((float)d));
        b[SC_PW] = 1.f;
    }
}

// Workgroup maximum of the per-thread bounds -> this workgroup's slot
// (one slot per workgroup of a fixed-size launch, no contended atomics).
__device__ __forceinline__ void writeBounds(unsigned int* maxbits, int nslots, const float* b)
{
    __shared__ float wmax[NSC][WG / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int g = 0; g < NSC; g++) {
        const float mx = waveMax(b[g]);
        if (lane == 0) wmax[g][wv] = mx;
    }
    __syncthreads();
    if (threadIdx.x < NSC) {
        float mx = 0.f;
#pragma unroll
        for (int w = 0; w < WG / 64; w++) mx = fmaxf(mx, wmax[threadIdx.x][w]);
        maxbits[threadIdx.x * nslots + blockIdx.x] = __float_as_uint(mx);
    }
}

// Stress of the uploaded state and the P2G bounds of the first substep
// (grid of exactly max_slots workgroups, striding over particles).
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_stress_init(Dev<T, AccT> D)
{
    float b[NSC] = { 0.f, 0.f, 0.f, 0.f, 0.f };
    const int64_t cap = D.cap;
    for (int p = blockIdx.x * WG + threadIdx.x; p < D.n; p += gridDim.x * WG) {
        const MatDev<T>& mt = D.mats[D.pmat[p]];
        Mat3<T> F = loadM(D.P, cap, PF, p);
        T S[6], v[3], Cm[9];
        vtauSym(mt, F, D.P[PVOL * cap + p], D.P[PG * cap + p], S);
#pragma unroll
        for (int i = 0; i < 6; i++) D.P[(PS + i) * cap + p] = S[i];
#pragma unroll
        for (int a = 0; a < 3; a++) v[a] = D.P[(PV + a) * cap + p];
#pragma unroll
        for (int i = 0; i < 9; i++) Cm[i] = D.P[(PC + i) * cap + p];
        const bool dmg = D.any_damage && mt.has_damage;
        particleBounds(b, D.P[PM * cap + p], v, Cm, S, dmg ? D.P[PD * cap + p] : T(0), D.dx, D.odx, D.any_damage != 0);
    }
    if constexpr (std::is_same<AccT, long long>::value) writeBounds(D.maxbits, D.max_slots, b);
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

// Adds a lane's partial sum into the LDS tile.  Float form: rounded to int64
// fixed point; |q| is summed into chk, so one test of chk <= FIXED_LANE_LIMIT per lane
// bounds every value added (and catches NaN / Inf) without a branch per value.
template <class T, class AccT>
__device__ __forceinline__ void ldsAdd(AccT* a, T v, float scale, float& chk)
{
    if constexpr (std::is_same<AccT, long long>::value) {
        const float q = rintf((float)v * scale);
        chk += fabsf(q);
        atomicAdd((unsigned long long*)a, (unsigned long long)(long long)q);
    }
    else
        unsafeAtomicAdd(a, v);
}

// Particle fields P2G reads, loaded ahead of use.
template <class T>
struct RawP2G {
    T x[3], v[3], C[9], S[6], m, d;
};

template <class T, class AccT>
__device__ __forceinline__ void loadP2G(const Dev<T, AccT>& D, int s, bool dmg, RawP2G<T>& r)
{
    int p = D.perm[s];
    if ((unsigned)p >= (unsigned)D.n) p = 0;
    const int64_t cap = D.cap;
    const T* P = D.P;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.x[a] = P[(PX + a) * cap + p];
        r.v[a] = P[(PV + a) * cap + p];
    }
#pragma unroll
    for (int i = 0; i < 9; i++) r.C[i] = P[(PC + i) * cap + p];
#pragma unroll
    for (int i = 0; i < 6; i++) r.S[i] = P[(PS + i) * cap + p];
    r.m = P[PM * cap + p];
    r.d = dmg && D.mats[D.pmat[p]].has_damage ? P[PD * cap + p] : T(0);
}

// particlesToGridWithForceHelper<APIC, deg 2, !MLS> (MpmSimulationBase.cpp:791-854)
// reading the stress of forcesUpdateParticleState (:953-979) from the particle,
// plus the AnisoMPM phase-field transfer (AnisoFractureSimulation.h:161-176).
//
// One workgroup per grid block (grid-stride over blocks).  The block's sorted
// particles are taken CHUNK at a time, counted from the block's first particle.
// Phase A: one thread per particle writes its P2G payload to LDS.  Phase B: the
// cells holding particles of the chunk are compacted; lanes are (sub-lane,
// cell, i, j) columns of the 27-node stencil; each lane walks every R-th
// particle of its cell (all share one base node), accumulates 3 nodes x 9
// quantities in registers and adds them into an LDS tile of the block + halo.
// After the last chunk the tile is flushed to HBM with one atomic per non-zero
// node value.  FP32 form: per-lane float partial sums in sorted order, then
// int64 fixed point in LDS and HBM, so the result does not depend on scheduling.
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_p2g(Dev<T, AccT> D)
{
    using B = Blk<T>;
    using TL = Tile<T>;
    constexpr int NB = TL::NB, CH = TL::CHUNK;
    constexpr bool FIXED = std::is_same<AccT, long long>::value;
    __shared__ V4<T> q[7][CH];
    __shared__ T wl[4][CH][3];
    __shared__ AccT tile[NGF][TL::TN];
    __shared__ int cs[NB], ce[NB], cl[NB];
    __shared__ int ncl;

    const int t = threadIdx.x;
    const bool dmg = D.any_damage != 0;
    const int nq = dmg ? NGF : GPF;
    const int64_t cap = D.cap, cn = D.cap_nodes;
    const int nblocks = D.counters[0];
    const float* sc = D.scales;
    const float scl[NGF] = { sc[SC_M], sc[SC_MV], sc[SC_MV], sc[SC_MV], sc[SC_F], sc[SC_F], sc[SC_F], sc[SC_PF], sc[SC_PW] };
    int ovf = 0;
    for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
        __syncthreads(); // previous block flushed
        if (t < NB) {
            const int c = b * NB + t;
            cs[t] = D.cell_start[c];
            ce[t] = D.cell_start[c] + D.cell_count[c];
        }
        __syncthreads();
        const int bs = cs[0], be = ce[NB - 1];
        if (be <= bs) continue; // halo-only block
        for (int e = t; e < nq * TL::TN; e += WG) (&tile[0][0])[e] = AccT(0);
        bool have = false;
        RawP2G<T> raw;
        for (int lo = bs; lo < be; lo += CH) {
            const int hi = min(lo + CH, be);
            __syncthreads(); // previous chunk's phase B done with the payload
            // ---- phase A: payload of particle lo + t from the prefetched fields
            if (t < CH && lo + t < hi) {
                if (!have) loadP2G(D, lo + t, dmg, raw);
                const T mass = raw.m;
                T w[3][3], dw[3][3], xb[3];
                int base[3];
                const T hi_t = T(D.domain - 4);
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    T xs = mul_rn(raw.x[a], D.odx);
                    T tt = sub_rn(xs, T(0.5));
                    if (!(tt >= T(0) && tt < hi_t)) tt = (tt >= hi_t) ? hi_t - T(1) : T(0); // latched by k_bin
                    base[a] = (int)tt;
                    bspline(xs, base[a], w[a], dw[a]);
                    xb[a] = T(base[a]) * D.dx - raw.x[a];
                }
                T mC[9], a3[3];
#pragma unroll
                for (int i = 0; i < 9; i++) mC[i] = mass * raw.C[i];
#pragma unroll
                for (int r = 0; r < 3; r++)
                    a3[r] = mass * raw.v[r] + mC[3 * r] * xb[0] + mC[3 * r + 1] * xb[1] + mC[3 * r + 2] * xb[2];
                T Sv[6];
#pragma unroll
                for (int i = 0; i < 6; i++) Sv[i] = raw.S[i] * D.odx;
                const T dx = D.dx;
                q[0][t] = V4<T> { mass, a3[0], a3[1], a3[2] };
                q[1][t] = V4<T> { mC[0] * dx, mC[1] * dx, mC[2] * dx, mC[3] * dx };
                q[2][t] = V4<T> { mC[4] * dx, mC[5] * dx, mC[6] * dx, mC[7] * dx };
                q[3][t] = V4<T> { mC[8] * dx, Sv[0], Sv[1], Sv[2] };
                q[4][t] = V4<T> { Sv[3], Sv[4], Sv[5], raw.d };
                q[5][t] = V4<T> { w[2][0], w[2][1], w[2][2], T(0) };
                q[6][t] = V4<T> { dw[2][0], dw[2][1], dw[2][2], T(0) };
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    wl[0][t][i] = w[0][i];
                    wl[1][t][i] = dw[0][i];
                    wl[2][t][i] = w[1][i];
                    wl[3][t][i] = dw[1][i];
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
            __syncthreads();
            // float form: prefetch the next chunk of this block while phase B runs
            // (the double form keeps its registers for phase B)
            have = false;
            if (sizeof(T) == 4 && t < CH && hi + t < be) {
                loadP2G(D, hi + t, dmg, raw);
                have = true;
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
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const T wk = wxy * wz[k];
                        const T g0 = gx * wz[k], g1 = gy * wz[k], g2 = wxy * dwz[k];
                        const T fk = T(k);
                        acc[k][GM] += q0.x * wk;
                        acc[k][GMV + 0] += wk * (a0 + b0 * fk);
                        acc[k][GMV + 1] += wk * (a1 + b1 * fk);
                        acc[k][GMV + 2] += wk * (a2 + b2 * fk);
                        acc[k][GFO + 0] -= s00 * g0 + s01 * g1 + s02 * g2;
                        acc[k][GFO + 1] -= s01 * g0 + s11 * g1 + s12 * g2;
                        acc[k][GFO + 2] -= s02 * g0 + s12 * g1 + s22 * g2;
                        acc[k][GPF] += q4.w * wk;
                        acc[k][GPW] += wk;
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
        __syncthreads();
        // ---- flush the tile: one HBM atomic per non-zero node value
        const int* nb = D.nbr + b * 8;
        for (int e = t; e < nq * TL::TN; e += WG) {
            const int f = e / TL::TN, nd = e - TL::TN * (e / TL::TN);
            const AccT v = tile[f][nd];
            if (v == AccT(0)) continue;
            const int tx = nd / (TL::TY * TL::TZ), ty = (nd / TL::TZ) % TL::TY, tz = nd % TL::TZ;
            const int which = ((tx >= TL::BX) << 2) | ((ty >= TL::BY) << 1) | (tz >= TL::BZ);
            const int tb = nb[which];
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

// Start of a substep: advance the device clock and (float form) set this
// substep's fixed-point scales from the bounds the previous G2P (or the upload)
// wrote, one slot per workgroup: 2^(40 - e) with max = f 2^e, f in [0.5,1).
constexpr int BEGIN_WG = 1024;
__global__ __launch_bounds__(BEGIN_WG) void k_begin_step(float* scales, const unsigned int* maxbits, int nslots, double* clock,
                                                  double dt)
{
    // device clock: clock[0] = time of the next substep, clock[1] = this substep's t
    if (threadIdx.x == 0) {
        clock[1] = clock[0];
        clock[0] = clock[0] + dt;
    }
    if (!scales) return;
    __shared__ float red[NSC][BEGIN_WG / 64];
    float mx[NSC] = { 0.f, 0.f, 0.f, 0.f, 0.f };
    for (int i = threadIdx.x; i < nslots; i += BEGIN_WG)
#pragma unroll
        for (int g = 0; g < NSC; g++) mx[g] = fmaxf(mx[g], __uint_as_float(maxbits[g * nslots + i]));
#pragma unroll
    for (int g = 0; g < NSC; g++) {
        const float w = waveMax(mx[g]);
        if ((threadIdx.x & 63) == 0) red[g][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < NSC) {
        float m = 0.f;
        for (int w = 0; w < BEGIN_WG / 64; w++) m = fmaxf(m, red[threadIdx.x][w]);
        if (m > 0.f && isfinite(m)) {
            int e;
            frexpf(m, &e);
            int se = (int)FIXED_TARGET_EXP - e;
            se = se < -120 ? -120 : (se > 120 ? 120 : se);
            scales[threadIdx.x] = ldexpf(1.0f, se);
        }
    }
}

template <class T, class AccT>
__device__ __forceinline__ T accValue(const Dev<T, AccT>& D, int f, int64_t node, int sg)
{
    if constexpr (std::is_same<AccT, long long>::value)
        return (T)((double)D.acc[f * D.cap_nodes + node] * (1.0 / (double)D.scales[sg]));
    else
        return D.acc[f * D.cap_nodes + node];
}

// v = mv/m (MpmSimulationBase.cpp:655-666) and gridVelocityExplicitUpdate
// (:1179-1228) with multiObjectCollision / detectAndResolveCollision
// (CollisionObject.cpp:66-105, 341-406) at time t; phase-field normalisation
// (AnisoFractureSimulation.h:177-179).
template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_grid(Dev<T, AccT> D, T dt)
{
    const double t = D.clock[1]; // collision objects at t_n (MpmSimulationBase.cpp:291-293)
    using B = Blk<T>;
    constexpr int NB = 1 << (B::xbits + B::ybits + B::zbits);
    const int64_t nn = (int64_t)D.counters[0] * NB;
    const int64_t cn = D.cap_nodes;
    for (int64_t node = (int64_t)blockIdx.x * WG + threadIdx.x; node < nn; node += (int64_t)gridDim.x * WG) {
        const T m = accValue(D, GM, node, SC_M);
        T f[3];
#pragma unroll
        for (int a = 0; a < 3; a++) f[a] = accValue(D, GFO + a, node, SC_F);
        if (m == T(0)) {
#pragma unroll
            for (int a = 0; a < 3; a++) D.newv[a * cn + node] = f[a];
            if (D.any_damage) D.pfn[node] = T(0);
            continue;
        }
        const int b = (int)(node / NB), e = (int)(node % NB);
        uint32_t Bx, By, Bz;
        blockOfPage<T>((uint32_t)D.block_page[b], Bx, By, Bz);
        const uint32_t ix = (Bx << B::xbits) | (uint32_t)(e >> (B::ybits + B::zbits));
        const uint32_t iy = (By << B::ybits) | (uint32_t)((e >> B::zbits) & ((1 << B::ybits) - 1));
        const uint32_t iz = (Bz << B::zbits) | (uint32_t)(e & ((1 << B::zbits) - 1));
        const T xi[3] = { T(ix) * D.dx, T(iy) * D.dx, T(iz) * D.dx };
        const T dt_over_mi = dt / m;
        T vi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            T v = accValue(D, GMV + a, node, SC_MV) / m;
            vi[a] = v + D.gravity[a] * dt + f[a] * dt_over_mi;
        }
        for (int c = 0; c < D.num_cols; c++) {
            const ColDev& co = D.cols[c];
            if (co.type == ZMPM_GHOST) continue;
            T X[3];
#pragma unroll
            for (int a = 0; a < 3; a++) X[a] = xi[a] - (T)(co.b0[a] + co.dbdt[a] * t);
            T phi, nrm[3] = { T(0), T(0), T(0) };
            if (co.shape == ZMPM_SHAPE_HALF_SPACE) {
                phi = (T)co.n[0] * (X[0] - (T)co.origin[0]) + (T)co.n[1] * (X[1] - (T)co.origin[1])
                    + (T)co.n[2] * (X[2] - (T)co.origin[2]);
#pragma unroll
                for (int a = 0; a < 3; a++) nrm[a] = (T)co.n[a];
            }
            else {
                T dd[3], Xc[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    T c0 = (T)((co.bmin[a] + co.bmax[a]) / 2), he = (T)((co.bmax[a] - co.bmin[a]) / 2);
                    Xc[a] = X[a] - c0;
                    dd[a] = dabs(Xc[a]) - he;
                }
                int am = 0;
                if (dd[1] > dd[am]) am = 1;
                if (dd[2] > dd[am]) am = 2;
                T q2 = T(0);
#pragma unroll
                for (int a = 0; a < 3; a++) q2 += dd[a] > T(0) ? dd[a] * dd[a] : T(0);
                T dmax = dd[am];
                phi = (dmax < T(0) ? dmax : T(0)) + dsqrt(q2);
                if (co.type != ZMPM_STICKY) {
                    if (q2 > T(0)) {
                        T qn = dsqrt(q2);
#pragma unroll
                        for (int a = 0; a < 3; a++)
                            nrm[a] = (dd[a] > T(0) ? dd[a] / qn : T(0)) * (Xc[a] >= T(0) ? T(1) : T(-1));
                    }
                    else
                        nrm[am] = Xc[am] >= T(0) ? T(1) : T(-1);
                }
            }
            if (!(phi <= T(0))) continue;
            const T vo[3] = { (T)co.dbdt[0], (T)co.dbdt[1], (T)co.dbdt[2] };
            if (co.type == ZMPM_STICKY) {
#pragma unroll
                for (int a = 0; a < 3; a++) vi[a] = vo[a];
                break;
            }
            T vr[3] = { vi[0] - vo[0], vi[1] - vo[1], vi[2] - vo[2] };
            T dot = vr[0] * nrm[0] + vr[1] * nrm[1] + vr[2] * nrm[2];
            if (co.type == ZMPM_SLIP || dot < T(0)) {
#pragma unroll
                for (int a = 0; a < 3; a++) vr[a] -= nrm[a] * dot;
                T fr = (T)co.friction;
                if (fr != T(0) && dot < T(0)) {
                    T vn = dsqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
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
        }
#pragma unroll
        for (int a = 0; a < 3; a++) D.newv[a * cn + node] = vi[a];
        if (D.any_damage) {
            const T pw = accValue(D, GPW, node, SC_PW);
            D.pfn[node] = pw != T(0) ? accValue(D, GPF, node, SC_PF) / pw : T(0);
        }
    }
}

// AnisoMPM explicit damage update of one particle (AnisoFractureSimulation.h:210-283,
// structural tensor :365-391, undegraded Cauchy stress QRStableNeoHookean.h:83-93).
template <class T>
__device__ __forceinline__ void damageUpdate(const MatDev<T>& mt, const Mat3<T>& F, T lap, T dt, T& d, T& g)
{
    const T geometricResist = d - (mt.l0 * mt.l0 * lap);
    // undegraded Cauchy stress tau(g = 1, no fibres) / J (evaluateEffectiveCauchyStress,
    // QRStableNeoHookean.h:83-93); tau is exactly symmetric in its closed form
    T J, t6[6];
    qrStableTau(F, mt.mu, mt.lambda, T(1), T(0), T(0), t6, &J);
    const T iJ = ddiv(T(1), J);
    Mat3<T> sym;
    sym.a[0][0] = t6[0] * iJ;
    sym.a[0][1] = sym.a[1][0] = t6[1] * iJ;
    sym.a[0][2] = sym.a[2][0] = t6[2] * iJ;
    sym.a[1][1] = t6[3] * iJ;
    sym.a[1][2] = sym.a[2][1] = t6[4] * iJ;
    sym.a[2][2] = t6[5] * iJ;
    const Mat3<T> sp = positivePart(sym);
    Mat3<T> R = polarRotation(F);
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
    const T dTilde = mt.zeta * macaulay(phi - T(1));
    const T diff = ((T(1) - d) * dTilde) - geometricResist;
    const T newD = d + (ddiv(dt, mt.eta) * macaulay(diff));
    d = newD < T(1) ? newD : T(1);
    const T k = mt.k;
    g = ((T(1) - d) * (T(1) - d) * (T(1) - k)) + k;
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
};

// APIC gather of one particle from the staged block tile.
template <class T, class AccT>
__device__ __forceinline__ void g2pGather(const Dev<T, AccT>& D, const T (*gt)[Tile<T>::TN], const T* Xp, bool dmg,
                                          G2PAcc<T>& A)
{
    using TL = Tile<T>;
    const T odx = D.odx, dx = D.dx, odx2 = odx * odx;
    T w[3][3], dw[3][3];
    int base[3];
    const T hi_t = T(D.domain - 4);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        T xs = mul_rn(Xp[a], odx);
        T tt = sub_rn(xs, T(0.5));
        if (!(tt >= T(0) && tt < hi_t)) tt = (tt >= hi_t) ? hi_t - T(1) : T(0);
        base[a] = (int)tt;
        bspline(xs, base[a], w[a], dw[a]);
    }
    const int lx0 = base[0] & (TL::BX - 1), ly0 = base[1] & (TL::BY - 1), lz0 = base[2] & (TL::BZ - 1);
#pragma unroll 1
    for (int i = 0; i < 3; i++) {
        const T wi = w[0][i];
        const T dwidxi = odx * dw[0][i];
        const T ddwi = (i == 1) ? T(-2) : T(1);
        const T xi = T(base[0] + i) * dx - Xp[0];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const T wj = w[1][j];
            const T wij = wi * wj;
            const T dwijdxi = dwidxi * wj;
            const T dwijdxj = wi * odx * dw[1][j];
            const T yj = T(base[1] + j) * dx - Xp[1];
            const int nrow = ((lx0 + i) * TL::TY + (ly0 + j)) * TL::TZ + lz0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const T wk = w[2][k];
                const T wijk = wij * wk;
                const T g3[3] = { dwijdxi * wk, dwijdxj * wk, wij * odx * dw[2][k] };
                const T zk = T(base[2] + k) * dx - Xp[2];
                const T nv[3] = { gt[0][nrow + k], gt[1][nrow + k], gt[2][nrow + k] };
                const T dxv[3] = { xi, yj, zk };
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    A.picV[a] += wijk * nv[a];
                    const T wv = wijk * nv[a];
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        A.Bp[a][c] += wv * dxv[c];
                        A.gv[a][c] += nv[a] * g3[c];
                    }
                }
                if (dmg) {
                    // quirk: ddw[1](i), ddw[2](i) instead of (j), (k) (MpmGrid.h:610-612)
                    const T L = odx2 * ddwi * wj * wk + odx2 * ddwi * wi * wk + odx2 * ddwi * wi * wj;
                    A.lap += gt[3][nrow + k] * L;
                }
            }
        }
    }
}

// Particle update after the gather; writes sorted position s.
// Particle fields G2P reads after the gather, loaded before it.
template <class T>
struct RawG2P {
    T x[3], F[9], m, V0, d, g;
    int mid, id;
};

template <class T, class AccT>
__device__ __forceinline__ void loadG2P(const Dev<T, AccT>& D, int s, int p, RawG2P<T>& r)
{
    const int64_t cap = D.cap;
    const T* P = D.P;
#pragma unroll
    for (int a = 0; a < 3; a++) r.x[a] = P[(PX + a) * cap + p];
#pragma unroll
    for (int i = 0; i < 9; i++) r.F[i] = P[(PF + i) * cap + p];
    r.m = P[PM * cap + p];
    r.V0 = P[PVOL * cap + p];
    r.d = P[PD * cap + p];
    r.g = P[PG * cap + p];
    r.mid = D.pmat[p];
    r.id = D.spid[s];
}

template <class T, class AccT>
__device__ __forceinline__ void g2pFinish(const Dev<T, AccT>& D, int s, const RawG2P<T>& R, T dt, const G2PAcc<T>& A,
                                          float* bnd)
{
    const int64_t cap = D.cap;
    T* Q = D.Q;
    const int mid = R.mid;
    const MatDev<T>& mt = D.mats[mid];
    const bool dmg = D.any_damage && mt.has_damage;
    const T dx = D.dx;
    const T Dinv = D.Dinv;
    const T r = D.rpic;
    const T ca = (r + T(1)) * T(0.5), cb = (r - T(1)) * T(0.5);
    T Cn[9];
#pragma unroll
    for (int a = 0; a < 3; a++) Q[(PV + a) * cap + s] = A.picV[a];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            Cn[3 * a + c] = ca * (A.Bp[a][c] * Dinv) + cb * (A.Bp[c][a] * Dinv);
            Q[(PC + 3 * a + c) * cap + s] = Cn[3 * a + c];
        }
    T inc2 = T(0);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        T inc = dt * A.picV[a];
        inc2 += inc * inc;
        Q[(PX + a) * cap + s] = R.x[a] + inc;
    }
    if (inc2 > dx * dx) latch(D.counters, ST_CFL);
    T d = R.d, g = R.g;
    const T V0 = R.V0;
    Mat3<T> Fn;
#pragma unroll
    for (int i = 0; i < 9; i++) Fn.a[i / 3][i % 3] = R.F[i];
    if (mt.model == ZMPM_MODEL_EQUATION_OF_STATE)
        Fn.a[0][0] = (T(1) + dt * (A.gv[0][0] + A.gv[1][1] + A.gv[2][2])) * Fn.a[0][0];
    else {
        const Mat3<T> F = Fn;
        if (dmg && mt.model == ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN) damageUpdate(mt, F, A.lap, dt, d, g);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                T v = F.a[a][c];
#pragma unroll
                for (int k = 0; k < 3; k++) v += dt * A.gv[a][k] * F.a[k][c];
                Fn.a[a][c] = v;
            }
    }
#pragma unroll
    for (int i = 0; i < 9; i++) Q[(PF + i) * cap + s] = Fn.a[i / 3][i % 3];
    // stress of the new state for the next substep's P2G
    T S[6];
    vtauSym(mt, Fn, V0, g, S);
#pragma unroll
    for (int i = 0; i < 6; i++) Q[(PS + i) * cap + s] = S[i];
    const T m = R.m;
    Q[PM * cap + s] = m;
    Q[PVOL * cap + s] = V0;
    Q[PD * cap + s] = d;
    Q[PG * cap + s] = g;
    Q[PLAP * cap + s] = dmg ? A.lap : T(0);
    D.qid[s] = R.id;
    D.qmat[s] = mid;
    if constexpr (std::is_same<AccT, long long>::value)
        particleBounds(bnd, m, A.picV, Cn, S, dmg ? d : T(0), dx, D.odx, D.any_damage != 0);
}

template <class T, class AccT>
__global__ __launch_bounds__(WG) void k_g2p(Dev<T, AccT> D, T dt)
{
    using B = Blk<T>;
    using TL = Tile<T>;
    constexpr int NB = TL::NB;
    __shared__ T gt[4][TL::TN];
    __shared__ int seg_end;
    const int lo = blockIdx.x * WG;
    const int hi = min(lo + WG, D.n);
    const int t = threadIdx.x, s = lo + t;
    const bool live = s < hi;
    const int64_t cn = D.cap_nodes;
    const int nfield = D.any_damage ? 4 : 3;
    RawG2P<T> R;
    bool dmg = false;
    if (live) {
        int p = D.perm[s];
        if ((unsigned)p >= (unsigned)D.n) p = 0;
        loadG2P(D, s, p, R);
        dmg = D.any_damage && D.mats[R.mid].has_damage;
    }
    G2PAcc<T> A;
#pragma unroll
    for (int a = 0; a < 3; a++) A.picV[a] = T(0);
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
        const int* nb = D.nbr + b * 8;
        for (int e = t; e < nfield * TL::TN; e += WG) {
            const int f = e / TL::TN, nd = e - TL::TN * (e / TL::TN);
            const int tx = nd / (TL::TY * TL::TZ), ty = (nd / TL::TZ) % TL::TY, tz = nd % TL::TZ;
            const int which = ((tx >= TL::BX) << 2) | ((ty >= TL::BY) << 1) | (tz >= TL::BZ);
            const int tb = nb[which];
            T v = T(0);
            if (tb >= 0) {
                const int elem = ((tx & (TL::BX - 1)) << (B::ybits + B::zbits)) | ((ty & (TL::BY - 1)) << B::zbits) | (tz & (TL::BZ - 1));
                const int64_t node = (int64_t)tb * NB + elem;
                v = f < 3 ? D.newv[f * cn + node] : D.pfn[node];
            }
            gt[f][nd] = v;
        }
        __syncthreads();
        const int seg_hi = max(seg_lo + 1, min(hi, seg_end));
        if (live && s >= seg_lo && s < seg_hi) g2pGather(D, gt, R.x, dmg, A);
        seg_lo = seg_hi;
    }
    float bnd[NSC] = { 0.f, 0.f, 0.f, 0.f, 0.f };
    if (live) g2pFinish(D, s, R, dt, A, bnd);
    if constexpr (std::is_same<AccT, long long>::value) writeBounds(D.maxbits, D.max_slots, bnd);
}

} // namespace zmpm
